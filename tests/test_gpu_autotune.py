"""Transform tile shapes and the autotuner (MI355X).

Every tile shape of each transform direction (forward 128x56, 64x56, and
128x48 and 128x40 without the LDS table; inverse 64x64, 32x32, 64x56 and
64x48 with 24-sample windows and once-per-sample scaling; and in both
directions the streamed 120x32 shape that lifts in registers) gives the same
coefficients bit for bit (compared directly), so a batch of uneven frames (1280x720 4:2:0, 640x480
4:4:4, 1920x1080 4:2:0, FP16 and FP32) encodes and decodes to the same
bitstreams and planes with any shape forced on every level (also with the LL
of levels 0-1 stored as fp16); pyrowave_hip_autotune_batch
returns those same results while it times the shapes, reports a valid shape
per level, and later calls (which use the tuned shapes) still agree.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from oracle import oracle as O
from pyrowave_amd import synth

from helpers import mask_unstored_ll

pytestmark = pytest.mark.gpu

SHAPES = [(1280, 720, pw.CHROMA_420), (640, 480, pw.CHROMA_444), (1920, 1080, pw.CHROMA_420)]


def _batch(torch):
    inputs = [[torch.from_numpy(p).cuda() for p in synth.frame(w, h, c == pw.CHROMA_420, seed=90 + i)]
              for i, (w, h, c) in enumerate(SHAPES)]
    frames = [pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts]) for ts in inputs]
    budgets = [int(w * h * 1.5 / 8) for w, h, _ in SHAPES]
    return inputs, frames, budgets


def _targets(torch):
    outs = [[torch.zeros(pw.plane_shape(w, h, c, p), dtype=torch.uint8, device="cuda") for p in range(3)]
            for w, h, c in SHAPES]
    return outs, [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in oo]) for oo in outs]


def _coeffs(e, w, h, c):
    # The fused forward path (no forced shape) does not store the first
    # level's LL; compare the coefficients every path stores.
    return mask_unstored_ll(O.layout(w, h, c), e.read_coefficients()).view(np.uint32)


def _collect(encs, outs, budgets):
    res = []
    for e, oo, b, (w, h, c) in zip(encs, outs, budgets, SHAPES):
        data, pk = e.packetize(1 << 30, b + 64)
        res.append((data[:sum(z for _, z in pk)].copy(), [o.cpu().numpy() for o in oo], _coeffs(e, w, h, c)))
    return res


MODES = {"shape0": 0, "shape1": 1, "shape2": 2, "shape3": 3, "shape4": 4}


@pytest.mark.parametrize("precision", [pw.PRECISION_FP16, pw.PRECISION_FP32, pw.PRECISION_FP16_LL])
def test_tile_shapes_agree_and_autotune(precision):
    import torch
    inputs, frames, budgets = _batch(torch)
    results = {}
    dev = pw.Device()
    try:
        for mode in list(MODES) + ["autotune", "tuned"]:
            dev.force_tile_variant(MODES.get(mode, -1))
            encs = [pw.Encoder(dev, w, h, c, precision=precision) for w, h, c in SHAPES]
            decs = [pw.Decoder(dev, w, h, c, precision=precision) for w, h, c in SHAPES]
            try:
                outs, tg = _targets(torch)
                if mode == "autotune":
                    fv, iv = pw.autotune_batch(encs, frames, budgets, decs, tg, repetitions=2)
                    assert all(v in (0, 1, 2, 3, 4) for v in fv + iv), (fv, iv)
                    print("autotuned shapes: forward", fv, "inverse", iv)
                else:
                    pw.encode_batch(encs, frames, budgets)
                    pw.decode_encoded_batch(decs, encs, tg)
                torch.cuda.synchronize()
                results[mode] = _collect(encs, outs, budgets)
            finally:
                for x in encs + decs:
                    x.close()
    finally:
        dev.close()
    ref = results["shape0"]
    for mode in list(MODES)[1:] + ["autotune", "tuned"]:
        for i, ((a, pa, ca), (b, pb, cb)) in enumerate(zip(ref, results[mode])):
            assert np.array_equal(ca, cb), (mode, i, int(np.count_nonzero(ca != cb)))
            assert np.array_equal(a, b), (mode, i)
            for p in range(3):
                assert np.array_equal(pa[p], pb[p]), (mode, i, p)


ODD = [(200, 130, pw.CHROMA_444), (258, 146, pw.CHROMA_420), (130, 66, pw.CHROMA_420), (1000, 562, pw.CHROMA_420)]


@pytest.mark.parametrize("precision", [pw.PRECISION_FP16, pw.PRECISION_FP32, pw.PRECISION_FP16_LL])
def test_streamed_shape_on_odd_sizes(precision):
    """The streamed shape (4) on sizes that are not multiples of the tiles, of
    32 or of the 128 minimum (edge extension, partial tiles, tiny levels):
    coefficients, bitstreams and decoded planes equal those of shape 0."""
    import torch
    inputs = [[torch.from_numpy(p).cuda() for p in synth.frame(w, h, c == pw.CHROMA_420, seed=70 + i)]
              for i, (w, h, c) in enumerate(ODD)]
    frames = [pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts]) for ts in inputs]
    budgets = [int(w * h * 2.0 / 8) for w, h, _ in ODD]
    results = {}
    dev = pw.Device()
    try:
        for shape in (0, 4):
            dev.force_tile_variant(shape)
            encs = [pw.Encoder(dev, w, h, c, precision=precision) for w, h, c in ODD]
            decs = [pw.Decoder(dev, w, h, c, precision=precision) for w, h, c in ODD]
            try:
                outs = [[torch.zeros(pw.plane_shape(w, h, c, p), dtype=torch.uint8, device="cuda")
                         for p in range(3)] for w, h, c in ODD]
                tg = [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in oo]) for oo in outs]
                pw.encode_batch(encs, frames, budgets)
                pw.decode_encoded_batch(decs, encs, tg)
                torch.cuda.synchronize()
                results[shape] = _collect(encs, outs, budgets)
            finally:
                for x in encs + decs:
                    x.close()
    finally:
        dev.close()
    for i, ((a, pa, ca), (b, pb, cb)) in enumerate(zip(results[0], results[4])):
        assert np.array_equal(ca, cb), (ODD[i], int(np.count_nonzero(ca != cb)))
        assert np.array_equal(a, b), ODD[i]
        for p in range(3):
            assert np.array_equal(pa[p], pb[p]), (ODD[i], p)


@pytest.mark.parametrize("interleaved", [False, True])
def test_streamed_shape_on_unaligned_planes(interleaved):
    """The streamed forward shape reads u8 planes through the aligned words
    that hold them: planes starting at odd byte addresses with odd pitches,
    and NV12-style interleaved chroma (Cr at Cb + 1, step 2), placed at the
    end of their buffer, encode to the same coefficients and bitstream as
    contiguous planes under shape 0."""
    import torch
    w, h = 1000, 562
    planes = synth.frame(w, h, True, seed=77)
    budget = int(w * h * 2.0 / 8)
    contiguous = [torch.from_numpy(p).cuda() for p in planes]
    cw, ch = w // 2, h // 2
    if interleaved:
        y_pitch, uv_pitch = w + 3, 2 * cw + 5
        y_off = 1
        uv_off = y_off + y_pitch * h + 2
        size = uv_off + uv_pitch * (ch - 1) + 2 * cw
        buf = torch.zeros(size, dtype=torch.uint8, device="cuda")
        buf[y_off:y_off + y_pitch * h].view(h, y_pitch)[:, :w].copy_(contiguous[0])
        uv = buf[uv_off:].as_strided((ch, cw, 2), (uv_pitch, 2, 1))
        uv[:, :, 0].copy_(contiguous[1])
        uv[:, :, 1].copy_(contiguous[2])
        base = buf.data_ptr()
        src = pw.hip_frame([(base + y_off, y_pitch, 1), (base + uv_off, uv_pitch, 2),
                            (base + uv_off + 1, uv_pitch, 2)])
    else:
        pitches = [w + 1, cw + 3, cw + 1]
        offs = [3]
        offs.append(offs[0] + pitches[0] * h + 1)
        offs.append(offs[1] + pitches[1] * ch + 2)
        size = offs[2] + pitches[2] * (ch - 1) + cw
        buf = torch.zeros(size, dtype=torch.uint8, device="cuda")
        for p, (o, pt) in enumerate(zip(offs, pitches)):
            ph, pw_ = contiguous[p].shape
            buf[o:o + pt * (ph - 1) + pw_].as_strided((ph, pw_), (pt, 1)).copy_(contiguous[p])
        base = buf.data_ptr()
        src = pw.hip_frame([(base + o, pt, 1) for o, pt in zip(offs, pitches)])
    torch.cuda.synchronize()
    results = {}
    dev = pw.Device()
    try:
        for shape, frame in ((0, pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in contiguous])), (4, src)):
            dev.force_tile_variant(shape)
            enc = pw.Encoder(dev, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP16)
            try:
                pw.encode_batch([enc], [frame], [budget])
                torch.cuda.synchronize()
                data, pk = enc.packetize(1 << 30, budget + 64)
                results[shape] = (data[:sum(z for _, z in pk)].copy(), _coeffs(enc, w, h, pw.CHROMA_420))
            finally:
                enc.close()
    finally:
        dev.close()
    assert np.array_equal(results[0][1], results[4][1]), int(np.count_nonzero(results[0][1] != results[4][1]))
    assert np.array_equal(results[0][0], results[4][0])
