"""Packed-FP16 lifting (PRECISION_FP16_MATH, the reference's precision 0:
shaders/dwt_common.h:29-34 with the quantizer policy of
pyrowave_encoder.cpp:144-172) at the benchmark operating points (MI355X).

- 3840x2160 4:2:0 seeded sequence (frames 0, 1, 30) at 1,600,000 B and at
  500,000 B (the reference test point): bitstreams equal the oracle's
  precision-0 run, bytes within the budget; luma PSNR of the decoded frames
  against FP32 and FP16 storage in the same run.
- 2560x1440 4:4:4 with 10 % checker patterns, eight budgets from 150 KB to
  2 MB: every bitstream equals the oracle's precision-0 run, PSNR grows with
  the budget; mean PSNR against FP32 per budget.

The precision-0 quantizer policy (6 base bits, resolution cap 512) bounds
the quality the mode reaches whatever the budget: measured, the 4K frames
decode at 56.5-56.7 dB luma at both budgets (FP32 70.1-70.2 dB at 1.6 MB,
59.4-60.9 dB at 500 KB), the 1440p sweep stops at 54.53 dB from the fifth
budget on (FP32 up to 62.84 dB) and is 0.04-0.11 dB above FP32 at the three
smallest budgets.  The loss bounds below are those figures with a margin; the
printed lines carry them.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from oracle import oracle as O
from pyrowave_amd import synth

from helpers import decode_bytes, gpu_encode

pytestmark = pytest.mark.gpu

MODES = ((pw.PRECISION_FP32, "fp32"), (pw.PRECISION_FP16, "fp16"), (pw.PRECISION_FP16_MATH, "fp16_math"))


def _oracle_streams(jobs):
    """Oracle encodes in parallel threads (the C oracle releases the GIL):
    jobs = [(L, planes, budget, sequence, precision, boundary)] -> bytes."""
    import concurrent.futures as cf

    def one(job):
        L, planes, budget, seq, prec, boundary = job
        bs, meta, _ = O.encode(L, planes, budget, sequence=seq, precision=prec)
        out, pk = O.packetize(L, meta, bs, seq, boundary)
        return out[:sum(s for _, s in pk)]

    with cf.ThreadPoolExecutor(max_workers=min(16, len(jobs))) as ex:
        return list(ex.map(one, jobs))


def _run(device, w, h, chroma, prec, frames, budgets, boundary=1 << 30):
    """Encode every (frame, budget) in order on one encoder and decode it; returns
    [(bytes, planes)] in the same order."""
    enc = pw.Encoder(device, w, h, chroma, precision=prec)
    dec = pw.Decoder(device, w, h, chroma, precision=prec)
    try:
        res = []
        for planes, budget in zip(frames, budgets):
            data, pk = gpu_encode(enc, planes, budget, boundary=boundary)
            res.append((data, decode_bytes(dec, data, pk)))
        return res
    finally:
        enc.close()
        dec.close()


@pytest.mark.parametrize("budget,max_loss_db", [(1600000, 14.0), (500000, 5.0)])
def test_fp16_math_4k_sequence(device, budget, max_loss_db):
    w, h = 3840, 2160
    gen = synth.SceneGenerator(w, h, True, seed=2024, frames=60, pan=3, redraw=0.05)
    frames = [gen.frame(t) for t in (0, 1, 30)]
    psnr = {}
    for prec, name in MODES:
        res = _run(device, w, h, pw.CHROMA_420, prec, frames, [budget] * len(frames))
        for data, _ in res:
            assert 0 < data.size <= budget, (name, data.size)
        psnr[name] = [synth.psnr(f[0], r[1][0]) for f, r in zip(frames, res)]
        if prec == pw.PRECISION_FP16_MATH:
            L = O.layout(w, h, 0, precision=0)
            refs = _oracle_streams([(L, f, budget, (i + 1) & 7, 0, 1 << 30) for i, f in enumerate(frames)])
            for i, (got, ref) in enumerate(zip(res, refs)):
                assert np.array_equal(got[0], ref), i
    delta = [a - b for a, b in zip(psnr["fp16_math"], psnr["fp32"])]
    print(f"4K {budget} B: PSNR-Y fp32 {np.round(psnr['fp32'], 4)} fp16 {np.round(psnr['fp16'], 4)} "
          f"fp16_math {np.round(psnr['fp16_math'], 4)}; fp16_math - fp32 {np.round(delta, 4)} dB")
    assert min(delta) >= -max_loss_db


BUDGETS_1440 = [int(round(150000 * (2000000 / 150000) ** (k / 7))) for k in range(8)]


def test_fp16_math_1440p_444_budget_sweep(device):
    w, h = 2560, 1440
    planes = synth.frame(w, h, False, seed=19, checker=0.10)
    curves = {}
    for prec, name in ((pw.PRECISION_FP32, "fp32"), (pw.PRECISION_FP16_MATH, "fp16_math")):
        res = _run(device, w, h, pw.CHROMA_444, prec, [planes] * 8, BUDGETS_1440, boundary=8192)
        for (data, _), budget in zip(res, BUDGETS_1440):
            assert 0 < data.size <= budget
        curves[name] = [float(np.mean([synth.psnr(planes[c], r[1][c]) for c in range(3)])) for r in res]
        if prec == pw.PRECISION_FP16_MATH:
            L = O.layout(w, h, 1, precision=0)
            refs = _oracle_streams([(L, planes, b, (i + 1) & 7, 0, 8192) for i, b in enumerate(BUDGETS_1440)])
            for b, got, ref in zip(BUDGETS_1440, res, refs):
                assert np.array_equal(got[0], ref), b
    delta = [a - b for a, b in zip(curves["fp16_math"], curves["fp32"])]
    print("budgets", BUDGETS_1440)
    print("PSNR fp32", np.round(curves["fp32"], 4), "fp16_math", np.round(curves["fp16_math"], 4),
          "delta", np.round(delta, 4))
    for vals in curves.values():
        assert all(b >= a - 1e-9 for a, b in zip(vals, vals[1:])), vals
    assert min(delta) >= -9.0
