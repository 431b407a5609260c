"""The fused first two transform levels against the level-by-level launches
(GPU), forced for every launch with PYROWAVE_FUSED=2: forward
(k_dwt_forward_fused), every stored coefficient and the bitstream are
identical; inverse (k_dwt_inverse_fused), every decoded plane sample is.  On
sizes that leave partial tiles at the right and bottom edges, 4:2:0 and
4:4:4, in every precision with FP32 arithmetic."""
import numpy as np
import pytest

import pyrowave_amd as pw
from oracle import oracle as O
from pyrowave_amd import synth

from helpers import decode_bytes, mask_unstored_ll

pytestmark = pytest.mark.gpu

SIZES = [(4, 4, pw.CHROMA_420), (34, 30, pw.CHROMA_420), (26, 26, pw.CHROMA_420), (33, 17, pw.CHROMA_444),
         (13, 13, pw.CHROMA_444), (200, 130, pw.CHROMA_444), (416, 240, pw.CHROMA_420),
         (1000, 750, pw.CHROMA_444), (1920, 1080, pw.CHROMA_420), (3840, 2160, pw.CHROMA_420)]


def _encode(dev, w, h, c, precision, planes, budget, force):
    dev.force_tile_variant(force)
    enc = pw.Encoder(dev, w, h, c, precision=precision)
    try:
        assert enc.encode_cpu(planes, budget) == pw.SUCCESS
        out, pk = enc.packetize(1 << 30, budget + 64)
        data = out[:sum(z for _, z in pk)].copy()
        coeffs = mask_unstored_ll(O.layout(w, h, c), enc.read_coefficients())
    finally:
        enc.close()
    return data, coeffs


@pytest.mark.parametrize("precision", [pw.PRECISION_FP32, pw.PRECISION_FP16, pw.PRECISION_FP16_LL])
@pytest.mark.parametrize("w,h,c", SIZES)
def test_fused_forward_equals_level_by_level(w, h, c, precision, monkeypatch):
    planes = synth.frame(w, h, c == pw.CHROMA_420, seed=w + h) if w * h >= 4096 else None
    if planes is None:
        from helpers import ramp_planes
        planes = ramp_planes(w, h, c)
    budget = max(2000, w * h // 4)
    monkeypatch.setenv("PYROWAVE_FUSED", "2")
    dev = pw.Device()
    try:
        fused = _encode(dev, w, h, c, precision, planes, budget, -1)
        for shape in (0, 4):
            ref = _encode(dev, w, h, c, precision, planes, budget, shape)
            mism = np.flatnonzero(fused[1].view(np.uint32) != ref[1].view(np.uint32))
            assert mism.size == 0, (shape, mism.size, mism[:8])
            assert np.array_equal(fused[0], ref[0]), shape
    finally:
        dev.close()


def _decode(dev, w, h, c, precision, data, packets, force):
    dev.force_tile_variant(force)
    dec = pw.Decoder(dev, w, h, c, precision=precision)
    try:
        return decode_bytes(dec, data, packets)
    finally:
        dec.close()


@pytest.mark.parametrize("precision", [pw.PRECISION_FP32, pw.PRECISION_FP16, pw.PRECISION_FP16_LL])
@pytest.mark.parametrize("w,h,c", SIZES)
def test_fused_inverse_equals_level_by_level(w, h, c, precision, monkeypatch):
    planes = synth.frame(w, h, c == pw.CHROMA_420, seed=w * h) if w * h >= 4096 else None
    if planes is None:
        from helpers import ramp_planes
        planes = ramp_planes(w, h, c)
    budget = max(2000, w * h // 2)
    monkeypatch.setenv("PYROWAVE_FUSED", "2")
    dev = pw.Device()
    try:
        enc = pw.Encoder(dev, w, h, c, precision=precision)
        try:
            assert enc.encode_cpu(planes, budget) == pw.SUCCESS
            out, pk = enc.packetize(1 << 30, budget + 64)
        finally:
            enc.close()
        fused = _decode(dev, w, h, c, precision, out, pk, -1)
        for shape in (0, 4):
            ref = _decode(dev, w, h, c, precision, out, pk, shape)
            for p in range(3):
                mism = np.flatnonzero(fused[p] != ref[p])
                assert mism.size == 0, (shape, p, mism.size, mism[:8])
    finally:
        dev.close()


@pytest.mark.parametrize("precision", [pw.PRECISION_FP16, pw.PRECISION_FP16_LL])
def test_fused_inverse_tall_tiles_batch(device, precision):
    """A batch decode large enough for the tall fused inverse tiles (nine
    2990x1690 4:4:4 frames: 3 x 25 x 27 = 2025 tiles of 120x64 each, 18225
    in all; partial tiles at the right and bottom) equals each frame decoded
    on its own level by level."""
    import torch
    w, h, c, n = 2990, 1690, pw.CHROMA_444, 9
    budget = 600000
    streams = []
    for i in range(n):
        enc = pw.Encoder(device, w, h, c, precision=precision)
        try:
            assert enc.encode_cpu(synth.frame(w, h, False, seed=100 + i), budget) == pw.SUCCESS
            out, pk = enc.packetize(1 << 30, budget + 64)
            streams.append((out, pk))
        finally:
            enc.close()
    decs = [pw.Decoder(device, w, h, c, precision=precision) for _ in range(n)]
    try:
        outs = [[torch.zeros((h, w), dtype=torch.uint8, device="cuda") for _ in range(3)] for _ in range(n)]
        tgs = [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in oo]) for oo in outs]
        for d, (data, pk) in zip(decs, streams):
            for o, z in pk:
                assert d.push_packet(data[o:o + z]) == pw.SUCCESS
        pw.decode_batch(decs, tgs)
        torch.cuda.synchronize()
        got = [[o.cpu().numpy() for o in oo] for oo in outs]
    finally:
        for d in decs:
            d.close()
    device.force_tile_variant(4)
    try:
        for i, (data, pk) in enumerate(streams):
            ref = _decode(device, w, h, c, precision, data, pk, 4)
            for p in range(3):
                assert np.array_equal(got[i][p], ref[p]), (i, p)
    finally:
        device.force_tile_variant(-1)
