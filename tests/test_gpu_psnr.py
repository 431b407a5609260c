"""PSNR-HVS-M-H on the MI355X against the oracle.

The device totals (24-bit fixed point per block, summed as integers, so
independent of summation order) equal the oracle's exactly, for every one of
the 16 height factors, on a ragged 37x29 noise pair, a 200x130 frame against
its 4:4:4 encode/decode and a 1920x1080 frame against its 4:2:0 decode at
400,000 bytes; totals accumulate over frames; bad arguments are refused.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth
from oracle import oracle as O

from helpers import decode_bytes, gpu_encode

pytestmark = pytest.mark.gpu


def _pairs(device):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 256, (29, 37), dtype=np.uint8)
    b = np.clip(a.astype(int) + rng.integers(-12, 13, a.shape), 0, 255).astype(np.uint8)
    pairs = [(a, b)]
    for w, h, chroma, budget in ((200, 130, pw.CHROMA_444, 20000), (1920, 1080, pw.CHROMA_420, 400000)):
        planes = synth.frame(w, h, chroma == pw.CHROMA_420, seed=5)
        enc = pw.Encoder(device, w, h, chroma, precision=pw.PRECISION_FP32)
        dec = pw.Decoder(device, w, h, chroma, precision=pw.PRECISION_FP32)
        try:
            data, pk = gpu_encode(enc, planes, budget)
            rec = decode_bytes(dec, data, pk)
        finally:
            enc.close()
            dec.close()
        pairs.append((planes[0], rec[0]))
    return pairs


def _device_totals(torch, device, a, b, totals=None):
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    if totals is None:
        totals = torch.zeros(16, dtype=torch.int64, device="cuda")
    h, w = a.shape
    tp = pw.psnr_hvs_m_accumulate(device, ta.data_ptr(), ta.stride(0), tb.data_ptr(), tb.stride(0), w, h,
                                  totals.data_ptr())
    torch.cuda.synchronize()
    return totals, tp


def test_psnr_hvs_m_matches_oracle(device):
    import torch
    for a, b in _pairs(device):
        totals, tp = _device_totals(torch, device, a, b)
        got = totals.cpu().numpy().view(np.uint64)
        ref, ref_tp = O.psnr_hvs_m_totals(a, b)
        assert tp == ref_tp
        assert np.array_equal(got, ref), (a.shape, got, ref)
        db = pw.psnr_hvs_m_db(got, tp)
        exp = [O.psnr_hvs_m_db(x, ref_tp) for x in ref]
        assert np.allclose(db, exp, rtol=0, atol=1e-9)
        print(a.shape, "PSNR-HVS-M-H dB (height factors 1, 2, 2.875):", np.round(db[[0, 8, 15]], 3))


def test_psnr_hvs_m_accumulates(device):
    import torch
    pairs = _pairs(device)
    totals = torch.zeros(16, dtype=torch.int64, device="cuda")
    tp = 0
    exp = np.zeros(16, np.uint64)
    exp_tp = 0
    for a, b in pairs:
        _, n = _device_totals(torch, device, a, b, totals)
        tp += n
        r, rn = O.psnr_hvs_m_totals(a, b)
        exp += r
        exp_tp += rn
    assert tp == exp_tp
    assert np.array_equal(totals.cpu().numpy().view(np.uint64), exp)


def test_psnr_hvs_m_rejects_bad_arguments(device):
    import torch
    t = torch.zeros(16, dtype=torch.int64, device="cuda")
    p = torch.zeros((8, 8), dtype=torch.uint8, device="cuda")
    with pytest.raises(pw.PyroWaveError):
        pw.psnr_hvs_m_accumulate(device, p.data_ptr(), 8, p.data_ptr(), 8, 0, 8, t.data_ptr())
    with pytest.raises(pw.PyroWaveError):
        pw.psnr_hvs_m_accumulate(device, p.data_ptr(), 4, p.data_ptr(), 8, 8, 8, t.data_ptr())
