"""The benchmark configurations as parity cases (MI355X).

- 1920x1080 4:2:0 at 400,000 B (FP32): bitstream and decode equal the oracle's,
  packets at 8 KiB decode in order, budget filled to 95-100 %.
- 3840x2160 4:2:0 seeded sequence at 1,600,000 B: FP32 and FP16 bitstreams
  of eight frames (first, panned, redrawn, last) equal the oracle run in the
  same mode; per-subband error of both transforms against the FP64 oracle;
  FP16 within 0.02 dB PSNR of FP32 on every frame.
- 2560x1440 4:4:4 with 10 % checker patterns, eight budgets from 150 KB to
  2 MB (log-spaced), FP32 and FP16: every bitstream equals the oracle's in the
  same mode, bytes never exceed the budget, PSNR grows with the budget, FP16
  within 0.02 dB of FP32 at every budget.
- A batch of 32 uneven frames (8 x 1280x720 4:2:0, 8 x 1920x1080 4:2:0,
  8 x 2560x1440 4:4:4, 8 x 3840x2160 4:2:0) at 1.6 bits/pixel, FP16, one
  launch per stage: every frame's bytes and decoded planes equal its
  single-frame encode and decode.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from oracle import oracle as O
from pyrowave_amd import synth

from helpers import decode_bytes, gpu_encode, oracle_encode

pytestmark = pytest.mark.gpu


def _bytes(out, pk):
    return out[:sum(s for _, s in pk)]


def test_config_1080p_fp32(device):
    w, h, budget = 1920, 1080, 400000
    planes = synth.frame(w, h, True, seed=101)
    L = O.layout(w, h, 0)
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    dec = pw.Decoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    try:
        data, pk = gpu_encode(enc, planes, budget, boundary=8192)
        assert len(pk) > 1
        ref, _, _ = oracle_encode(L, planes, budget, sequence=1, boundary=8192)
        assert np.array_equal(data, ref)
        assert 0.95 * budget <= data.size <= budget
        for i, (off, size) in enumerate(pk):
            assert dec.push_packet(data[off:off + size]) == pw.SUCCESS
            assert dec.is_ready() == (i == len(pk) - 1)
        got = dec.decode_cpu()
        exp = O.decode(L, [data[o:o + s].tobytes() for o, s in pk])
        for c in range(3):
            assert np.array_equal(got[c], exp[c])
    finally:
        enc.close()
        dec.close()


def _oracle_streams(jobs):
    """Oracle encodes run concurrently (the C oracle releases the GIL):
    jobs = [(L, planes, budget, sequence, precision, boundary)] -> bytes."""
    import concurrent.futures as cf

    def one(job):
        L, planes, budget, seq, prec, boundary = job
        bs, meta, _ = O.encode(L, planes, budget, sequence=seq, precision=prec)
        out, opk = O.packetize(L, meta, bs, seq, boundary)
        return _bytes(out, opk)

    with cf.ThreadPoolExecutor(max_workers=min(16, len(jobs))) as ex:
        return list(ex.map(one, jobs))


SEQ_FRAMES = (0, 1, 2, 3, 10, 20, 30, 59)


def test_config_4k_sequence_oracle_frames(device):
    """Eight frames of the 60-frame sequence (the pan moves 3 px and 5 % of the
    rectangles are redrawn per frame), FP32 and FP16: every bitstream equals
    the oracle's in the same mode, with the encoder's sequence counter."""
    w, h, budget = 3840, 2160, 1600000
    gen = synth.SceneGenerator(w, h, True, seed=2024, frames=60, pan=3, redraw=0.05)
    L = O.layout(w, h, 0)
    frames = [gen.frame(t) for t in SEQ_FRAMES]
    for prec, oracle_prec in ((pw.PRECISION_FP32, 2), (pw.PRECISION_FP16, 1)):
        enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=prec)
        try:
            got = [gpu_encode(enc, planes, budget)[0] for planes in frames]
        finally:
            enc.close()
        refs = _oracle_streams([(L, planes, budget, enc_sequence(i), oracle_prec, 1 << 30)
                                for i, planes in enumerate(frames)])
        for t, a, b in zip(SEQ_FRAMES, got, refs):
            assert 0.95 * budget <= a.size <= budget, (prec, t, a.size)
            assert np.array_equal(a, b), (prec, t)


def test_config_4k_sequence(device):
    w, h, budget = 3840, 2160, 1600000
    gen = synth.SceneGenerator(w, h, True, seed=2024, frames=60, pan=3, redraw=0.05)
    L = O.layout(w, h, 0)
    frames = [gen.frame(t) for t in (0, 1, 30)]
    psnr = {}
    for prec, oracle_prec in ((pw.PRECISION_FP32, 2), (pw.PRECISION_FP16, 1)):
        enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=prec)
        dec = pw.Decoder(device, w, h, pw.CHROMA_420, precision=prec)
        try:
            vals = []
            for i, planes in enumerate(frames):
                data, pk = gpu_encode(enc, planes, budget)
                assert 0.95 * budget <= data.size <= budget
                if i == 0:
                    # Transform error against FP64, per subband (levels 0-1 high
                    # bands carry the fp16 storage in FP16 mode).
                    gpu = enc.read_coefficients()
                    c64 = O.dwt_frame(L, planes, f64=True)
                    worst = 0.0
                    for c in range(3):
                        for l in range(5):
                            for b in range(4):
                                B = L.band[c][l][b]
                                if not B.present:
                                    continue
                                d = (O.band_view(L, gpu, c, l, b).astype(np.float64) -
                                     O.band_view(L, c64, c, l, b))
                                mx, rms = float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))
                                lim = (4e-3, 6e-4) if (prec == pw.PRECISION_FP16 and l < 2 and b) else (2e-5, 2e-6)
                                assert mx <= lim[0] and rms <= lim[1], (prec, c, l, b, mx, rms)
                                worst = max(worst, mx)
                    print(f"precision {prec}: worst per-band max-abs vs FP64 {worst:.3g}")
                    # The first encode carries sequence 1.
                    bs, meta, _ = O.encode(L, planes, budget, sequence=1, precision=oracle_prec)
                    out, opk = O.packetize(L, meta, bs, 1, 1 << 30)
                    assert np.array_equal(data, _bytes(out, opk))
                rec = decode_bytes(dec, data, pk)
                vals.append(synth.psnr(planes[0], rec[0]))
            psnr[prec] = vals
        finally:
            enc.close()
            dec.close()
    print("PSNR fp32", psnr[pw.PRECISION_FP32], "fp16", psnr[pw.PRECISION_FP16])
    for a, b in zip(psnr[pw.PRECISION_FP32], psnr[pw.PRECISION_FP16]):
        assert abs(a - b) <= 0.02


def enc_sequence(i):
    """The encoder's 3-bit sequence counter before its (i+1)-th frame."""
    return (i + 1) & 7


BUDGETS_1440 = [int(round(150000 * (2000000 / 150000) ** (k / 7))) for k in range(8)]


def test_config_1440p_444_budget_sweep(device):
    w, h = 2560, 1440
    planes = synth.frame(w, h, False, seed=19, checker=0.10)
    L = O.layout(w, h, 1)
    curves = {}
    for prec in (pw.PRECISION_FP32, pw.PRECISION_FP16):
        enc = pw.Encoder(device, w, h, pw.CHROMA_444, precision=prec)
        dec = pw.Decoder(device, w, h, pw.CHROMA_444, precision=prec)
        try:
            vals, streams = [], []
            for i, budget in enumerate(BUDGETS_1440):
                data, pk = gpu_encode(enc, planes, budget, boundary=8192)
                assert 0 < data.size <= budget
                streams.append(data)
                rec = decode_bytes(dec, data, pk)
                vals.append(np.mean([synth.psnr(planes[c], rec[c]) for c in range(3)]))
            curves[prec] = vals
            # Every budget's bitstream equals the oracle's in the same mode.
            oprec = 2 if prec == pw.PRECISION_FP32 else 1
            refs = _oracle_streams([(L, planes, budget, enc_sequence(i), oprec, 8192)
                                    for i, budget in enumerate(BUDGETS_1440)])
            for budget, a, b in zip(BUDGETS_1440, streams, refs):
                assert np.array_equal(a, b), (prec, budget)
        finally:
            enc.close()
            dec.close()
    print("budgets", BUDGETS_1440)
    print("PSNR fp32", np.round(curves[pw.PRECISION_FP32], 4), "fp16", np.round(curves[pw.PRECISION_FP16], 4))
    for prec, vals in curves.items():
        assert all(b >= a - 1e-9 for a, b in zip(vals, vals[1:])), (prec, vals)
    for a, b in zip(curves[pw.PRECISION_FP32], curves[pw.PRECISION_FP16]):
        assert abs(a - b) <= 0.02


def test_config_uneven_batch_of_32(device):
    import torch
    shapes = ([(1280, 720, pw.CHROMA_420)] * 8 + [(1920, 1080, pw.CHROMA_420)] * 8 +
              [(2560, 1440, pw.CHROMA_444)] * 8 + [(3840, 2160, pw.CHROMA_420)] * 8)
    # 1.6 bits per pixel of the luma grid.
    budgets = [int(w * h * 1.6 / 8) for w, h, _ in shapes]
    contents = [synth.frame(w, h, c == pw.CHROMA_420, seed=300 + i) for i, (w, h, c) in enumerate(shapes)]
    singles = []
    for (w, h, c), planes, budget in zip(shapes, contents, budgets):
        enc = pw.Encoder(device, w, h, c, precision=pw.PRECISION_FP16)
        dec = pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP16)
        try:
            data, pk = gpu_encode(enc, planes, budget)
            singles.append((data, decode_bytes(dec, data, pk)))
        finally:
            enc.close()
            dec.close()
    encs = [pw.Encoder(device, w, h, c, precision=pw.PRECISION_FP16) for w, h, c in shapes]
    decs = [pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP16) for w, h, c in shapes]
    try:
        tensors = [[torch.from_numpy(p).cuda() for p in planes] for planes in contents]
        outs = [[torch.empty_like(p) for p in ts] for ts in tensors]
        frames = [pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts]) for ts in tensors]
        targets = [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in out]) for out in outs]
        stream = torch.cuda.Stream()
        pw.encode_batch(encs, frames, budgets, stream=stream.cuda_stream)
        pw.decode_encoded_batch(decs, encs, targets, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        for i, (enc, (ref_bytes, ref_planes), budget) in enumerate(zip(encs, singles, budgets)):
            out, pk = enc.packetize(1 << 30, budget + 64)
            assert np.array_equal(_bytes(out, pk), ref_bytes), i
            for c in range(3):
                assert np.array_equal(outs[i][c].cpu().numpy(), ref_planes[c]), (i, c)
    finally:
        for e in encs:
            e.close()
        for d in decs:
            d.close()
