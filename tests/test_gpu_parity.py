"""Parity of the HIP path against the oracle (GPU tests).

The FP32 path is bit-exact by construction: the forward transform evaluates
every lifting step as fmaf(c, l + r, v) in the reference's order, and every
integer stage follows the reference kernels.  So the tests compare bytes:
wavelet coefficients, rate-control state, bitstreams and decoded planes.
The FP16-storage path is checked against the oracle's FP16-storage transform
(bit-exact) and against the FP64 transform with stated tolerances.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth
from oracle import oracle as O
from tests.helpers import (band_errors, decode_bytes, gpu_encode, mask_unstored_ll, oracle_encode, ramp_planes)

pytestmark = pytest.mark.gpu

SIZES = [
    (34, 30, pw.CHROMA_420),
    (34, 30, pw.CHROMA_444),
    (33, 17, pw.CHROMA_444),
    (4, 4, pw.CHROMA_420),
    (200, 130, pw.CHROMA_444),
    (1920, 1080, pw.CHROMA_420),
]


def _content(w, h, chroma, seed=1):
    if w * h < 4096:
        return ramp_planes(w, h, chroma)
    return synth.frame(w, h, chroma == pw.CHROMA_420, seed=seed)


@pytest.mark.parametrize("w,h,chroma", SIZES)
def test_dwt_fp32_bit_exact(device, w, h, chroma):
    planes = _content(w, h, chroma)
    L = O.layout(w, h, chroma)
    enc = pw.Encoder(device, w, h, chroma, precision=pw.PRECISION_FP32)
    try:
        assert enc.encode_cpu(planes, 1 << 20) == pw.SUCCESS
        gpu = enc.read_coefficients()
    finally:
        enc.close()
    ref = O.dwt_frame(L, planes, precision=2)
    assert gpu.shape == ref.shape
    gpu, ref = mask_unstored_ll(L, gpu), mask_unstored_ll(L, ref)
    mism = np.flatnonzero(gpu.view(np.uint32) != ref.view(np.uint32))
    assert mism.size == 0, f"{mism.size} coefficients differ, first at {mism[:5]}"


# Oracle precision of each storage mode.
ORACLE_PRECISION = {pw.PRECISION_FP32: 2, pw.PRECISION_FP16: 1, pw.PRECISION_FP16_LL: 3, pw.PRECISION_FP16_MATH: 0}


@pytest.mark.parametrize("precision", [pw.PRECISION_FP16, pw.PRECISION_FP16_LL, pw.PRECISION_FP16_MATH])
@pytest.mark.parametrize("w,h,chroma", [(34, 30, pw.CHROMA_420), (200, 130, pw.CHROMA_444),
                                        (1920, 1080, pw.CHROMA_420)])
def test_dwt_fp16_storage_bit_exact(device, w, h, chroma, precision):
    """Every band of every level, the LL intermediates included, equals the
    oracle in the same storage mode (FP16: high bands of levels 0-1 in fp16;
    FP16_LL: all four bands of levels 0-1, so levels 1-2 read an fp16 LL;
    FP16_MATH: packed-FP16 lifting and fp16 bands at every level)."""
    planes = _content(w, h, chroma)
    L = O.layout(w, h, chroma)
    enc = pw.Encoder(device, w, h, chroma, precision=precision)
    try:
        assert enc.encode_cpu(planes, 1 << 20) == pw.SUCCESS
        gpu = enc.read_coefficients()
    finally:
        enc.close()
    ref = O.dwt_frame(L, planes, precision=ORACLE_PRECISION[precision])
    gpu, ref = mask_unstored_ll(L, gpu), mask_unstored_ll(L, ref)
    mism = np.flatnonzero(gpu.view(np.uint32) != ref.view(np.uint32))
    assert mism.size == 0, f"{mism.size} coefficients differ, first at {mism[:5]}"


@pytest.mark.parametrize("precision,max_abs,rms", [(pw.PRECISION_FP32, 2e-5, 2e-6),
                                                   (pw.PRECISION_FP16, 4e-3, 6e-4),
                                                   (pw.PRECISION_FP16_LL, 1e-3, 1.5e-4),
                                                   (pw.PRECISION_FP16_MATH, 5e-3, 8e-4)])
def test_dwt_vs_fp64(device, precision, max_abs, rms):
    """Per-subband error against the FP64 lifting oracle.  Tolerances:
    FP32 max-abs 2e-5, RMS 2e-6; FP16 storage max-abs 4e-3, RMS 6e-4
    (coefficients are O(1); fp16 has an 11-bit significand); FP16_LL storage
    1e-3 / 1.5e-4; packed-FP16 lifting 5e-3 / 8e-4 (every step rounds to
    fp16: measured 3.1e-3 / 5.2e-4 at 1080p)."""
    w, h = 1920, 1080
    planes = synth.frame(w, h, True, seed=3)
    L = O.layout(w, h, 0)
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=precision)
    try:
        assert enc.encode_cpu(planes, 400000) == pw.SUCCESS
        gpu = enc.read_coefficients()
    finally:
        enc.close()
    ref64 = O.dwt_frame(L, planes, f64=True)
    errs = band_errors(L, gpu, ref64)
    worst_abs = max(e[0] for e in errs.values())
    worst_rms = max(e[1] for e in errs.values())
    print("per-band (max-abs, rms):", {k: ("%.2e" % v[0], "%.2e" % v[1]) for k, v in errs.items()})
    assert worst_abs <= max_abs and worst_rms <= rms


@pytest.mark.parametrize("w,h,chroma,budget", [
    (34, 30, pw.CHROMA_420, 64 * 1024),
    (34, 30, pw.CHROMA_444, 64 * 1024),
    (34, 30, pw.CHROMA_420, 600),
    (256, 256, pw.CHROMA_444, 20000),
    (1920, 1080, pw.CHROMA_420, 400000),
])
def test_encode_bit_exact_fp32(device, w, h, chroma, budget):
    planes = _content(w, h, chroma, seed=5)
    L = O.layout(w, h, chroma)
    enc = pw.Encoder(device, w, h, chroma, precision=pw.PRECISION_FP32)
    try:
        got, _ = gpu_encode(enc, planes, budget)
        rc = enc.read_rate_control(L.count32)
    finally:
        enc.close()
    ref, _, quant = oracle_encode(L, planes, budget, sequence=1)
    assert got.size <= budget
    assert got.size == ref.size and np.array_equal(got, ref)
    assert np.array_equal(rc["quant"], quant)


@pytest.mark.parametrize("w,h,chroma,budget", [
    (34, 30, pw.CHROMA_420, 64 * 1024),
    (200, 130, pw.CHROMA_444, 9000),
    (1920, 1080, pw.CHROMA_420, 400000),
])
def test_encode_bit_exact_fp16_math(device, w, h, chroma, budget):
    """The packed-FP16 encoder (reference precision 0, with its quantizer
    policy) gives the oracle's precision-0 bitstream byte for byte."""
    planes = _content(w, h, chroma, seed=37)
    L = O.layout(w, h, chroma, precision=0)
    enc = pw.Encoder(device, w, h, chroma, precision=pw.PRECISION_FP16_MATH)
    try:
        got, _ = gpu_encode(enc, planes, budget)
    finally:
        enc.close()
    ref, _, _ = oracle_encode(L, planes, budget, sequence=1, precision=0)
    assert got.size <= budget
    assert np.array_equal(got, ref)


def test_fp16_math_batch_cannot_mix_precisions(device):
    """One batch holds packed-FP16 frames or the others, not both."""
    import torch
    w, h = 256, 144
    planes = _content(w, h, pw.CHROMA_420)
    encs = [pw.Encoder(device, w, h, pw.CHROMA_420, precision=p)
            for p in (pw.PRECISION_FP16_MATH, pw.PRECISION_FP32)]
    try:
        ts = [torch.from_numpy(p).cuda() for p in planes]
        frame = pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts])
        with pytest.raises(pw.PyroWaveError):
            pw.encode_batch(encs, [frame, frame], [20000, 20000])
    finally:
        for e in encs:
            e.close()


@pytest.mark.parametrize("merged", ["1", "0"])
def test_rate_control_state_matches_oracle(merged, monkeypatch):
    """Rate-control state against the oracle, through the one-workgroup
    resolve of a lone 1080p frame (k_rc_resolve) and through the three
    launches (PYROWAVE_FUSED_RESOLVE=0)."""
    w, h = 1920, 1080
    planes = synth.frame(w, h, True, seed=7)
    L = O.layout(w, h, 0)
    monkeypatch.setenv("PYROWAVE_FUSED_RESOLVE", merged)
    device = pw.Device()
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    try:
        assert enc.encode_cpu(planes, 300000) == pw.SUCCESS
        enc.compute_num_packets(1 << 30)
        coeffs = enc.read_coefficients()
        rc = enc.read_rate_control(L.count32)
    finally:
        enc.close()
        device.close()
    qa = O.quant_analyze(L, coeffs)
    assert np.array_equal(rc["words"], qa["words"])
    assert np.array_equal(rc["bucket"], qa["bucket"])
    assert np.array_equal(rc["totals"], qa["totals"])
    assert rc["consumed"] == qa["consumed"]
    prefix = O.finalize(qa["totals"], q1=True)
    assert np.array_equal(rc["prefix"], prefix)
    assert np.array_equal(rc["quant"], O.resolve(L, qa, prefix, 300000))


def test_fp16_path_integer_stages_bit_exact(device):
    """Given the FP16 path's own coefficients, quantization, rate control and
    packing match the oracle byte for byte."""
    w, h = 1920, 1080
    planes = synth.frame(w, h, True, seed=11)
    L = O.layout(w, h, 0)
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP16)
    try:
        got, _ = gpu_encode(enc, planes, 400000)
        coeffs = enc.read_coefficients()
    finally:
        enc.close()
    qa = O.quant_analyze(L, coeffs)
    prefix = O.finalize(qa["totals"], q1=True)
    quant = O.resolve(L, qa, prefix, 400000)
    bs, meta = O.pack(L, coeffs, quant, 1)
    ref, _ = O.packetize(L, meta, bs, 1, 1 << 30)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("precision", [pw.PRECISION_FP32, pw.PRECISION_FP16, pw.PRECISION_FP16_LL,
                                       pw.PRECISION_FP16_MATH])
@pytest.mark.parametrize("w,h,chroma,budget", [
    (34, 30, pw.CHROMA_420, 64 * 1024),
    (34, 30, pw.CHROMA_444, 64 * 1024),
    (1920, 1080, pw.CHROMA_420, 400000),
])
def test_decode_bit_exact(device, w, h, chroma, budget, precision):
    """Dequantized coefficients and decoded planes equal the oracle's.  The
    FP16 decoder keeps the high bands of levels 0-1 in fp16, which the oracle's
    precision-1 decode reproduces; the FP16_LL decoder also the LL that the
    inverse of levels 1-2 writes (oracle precision 3); the FP16_MATH decoder
    every band in fp16 and the inverse in packed-FP16 lifting (precision 0)."""
    planes = _content(w, h, chroma, seed=13)
    L = O.layout(w, h, chroma)
    ref_bytes, pk, _ = oracle_encode(L, planes, budget)
    dec = pw.Decoder(device, w, h, chroma, precision=precision)
    try:
        got = decode_bytes(dec, ref_bytes, pk)
        gcoef = dec.read_coefficients()
    finally:
        dec.close()
    d = O.Decoder(L)
    for off, size in pk:
        assert d.push(ref_bytes[off:off + size].tobytes())
    rcoef = O.dequant(L, d.payload(), d.offsets())
    oprec = ORACLE_PRECISION[precision]
    for c in range(3):
        for l in range(5):
            for b in range(4):
                B = L.band[c][l][b]
                if not B.present:
                    continue
                r = O.band_view(L, rcoef, c, l, b)
                if (oprec != 2 and l < 2 and b != 0) or oprec == 0:
                    r = r.astype(np.float16).astype(np.float32)
                assert np.array_equal(O.band_view(L, gcoef, c, l, b).view(np.uint32), r.view(np.uint32)), (c, l, b)
    ref = O.idwt_frame(L, rcoef, precision=oprec)
    for c in range(3):
        assert np.array_equal(got[c], ref[c]), f"plane {c}"


def test_roundtrip_fp16_close_to_fp32(device):
    w, h = 1920, 1080
    planes = synth.frame(w, h, True, seed=17)
    out = {}
    for prec in (pw.PRECISION_FP32, pw.PRECISION_FP16):
        enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=prec)
        dec = pw.Decoder(device, w, h, pw.CHROMA_420, precision=prec)
        try:
            data, pk = gpu_encode(enc, planes, 400000)
            out[prec] = decode_bytes(dec, data, pk)
        finally:
            enc.close()
            dec.close()
    p32 = synth.psnr(planes[0], out[pw.PRECISION_FP32][0])
    p16 = synth.psnr(planes[0], out[pw.PRECISION_FP16][0])
    print("PSNR fp32 %.4f fp16 %.4f" % (p32, p16))
    assert abs(p32 - p16) <= 0.02


@pytest.mark.parametrize("budget", [150000, 455206, 2000000])
def test_budget_never_exceeded_444(device, budget):
    w, h = 2560, 1440
    planes = synth.frame(w, h, False, seed=19, checker=0.10)
    enc = pw.Encoder(device, w, h, pw.CHROMA_444, precision=pw.PRECISION_FP32)
    try:
        data, pk = gpu_encode(enc, planes, budget, boundary=8192)
    finally:
        enc.close()
    assert 0 < data.size <= budget


def test_deterministic(device):
    """Two encoders on the same frame give identical bytes, equal to the oracle's."""
    w, h = 640, 360
    planes = synth.frame(w, h, True, seed=23)
    L = O.layout(w, h, 0)
    outs = []
    for _ in range(2):
        enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
        try:
            outs.append(gpu_encode(enc, planes, 50000)[0])
        finally:
            enc.close()
    assert np.array_equal(outs[0], outs[1])
    ref, _, _ = oracle_encode(L, planes, 50000, sequence=1)
    assert np.array_equal(outs[0], ref)


def test_batch_matches_single(device):
    import torch
    sizes = [(1280, 720, pw.CHROMA_420), (640, 480, pw.CHROMA_444), (34, 30, pw.CHROMA_420),
             (1920, 1080, pw.CHROMA_420)]
    budgets = [184320, 100000, 4000, 414720]
    contents = [_content(w, h, c, seed=29 + i) for i, (w, h, c) in enumerate(sizes)]
    singles = []
    for (w, h, c), planes, budget in zip(sizes, contents, budgets):
        enc = pw.Encoder(device, w, h, c, precision=pw.PRECISION_FP32)
        try:
            singles.append(gpu_encode(enc, planes, budget)[0])
        finally:
            enc.close()
    encs = [pw.Encoder(device, w, h, c, precision=pw.PRECISION_FP32) for (w, h, c) in sizes]
    try:
        tensors = [[torch.from_numpy(p).cuda() for p in planes] for planes in contents]
        frames = [pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts]) for ts in tensors]
        stream = torch.cuda.current_stream().cuda_stream
        pw.encode_batch(encs, frames, budgets, stream=stream)
        for enc, ref, budget in zip(encs, singles, budgets):
            out, pk = enc.packetize(1 << 30, budget + 64)
            got = out[:sum(s for _, s in pk)]
            assert np.array_equal(got, ref)
    finally:
        for e in encs:
            e.close()


def test_device_roundtrip_matches_packet_decode(device):
    import torch
    w, h = 1280, 720
    planes = synth.frame(w, h, True, seed=31)
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    dec = pw.Decoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    dec2 = pw.Decoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    try:
        ts = [torch.from_numpy(p).cuda() for p in planes]
        outs = [torch.zeros_like(t) for t in ts]
        stream = torch.cuda.current_stream().cuda_stream
        pw.encode_batch([enc], [pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts])], [184320],
                        stream=stream)
        pw.decode_encoded_batch([dec], [enc], [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in outs])],
                                stream=stream)
        torch.cuda.synchronize()
        got = [o.cpu().numpy() for o in outs]
        data, pk = enc.packetize(1 << 30, 184320 + 64)
        ref = decode_bytes(dec2, data, pk)
        for c in range(3):
            assert np.array_equal(got[c], ref[c])
    finally:
        enc.close()
        dec.close()
        dec2.close()
