"""16-bit (R16_UNORM) input and output on the MI355X.

The reference encodes 10- to 16-bit y4m input, rescaled to the full 16-bit
range, through R16_UNORM images (encode.cpp:152, yuv4mpeg.cpp:221-233) and
decodes into R16_UNORM targets (decode.cpp:215).  Here an encoder created
with UNORM16_INPUT reads u16 device planes (u / 65535 - 0.5) and a decoder
with UNORM16_OUTPUT writes round(clamp(v + 0.5) * 65535).  Checked through
the C ABI against the oracle run on the same 16-bit planes:

- bitstreams byte for byte in every precision, on sizes with partial tiles,
  4:2:0 and 4:4:4, planar and interleaved (P016) chroma, padded rows;
- every stored coefficient, on a plane holding all 65536 sample values;
- 16-bit decode output equal to the oracle's 16-bit decode;
- planes of 257 * u (the rational u / 255) give the 8-bit encoder's bytes;
- a batch mixing 8- and 16-bit encoders equals the frames encoded alone;
- misaligned 16-bit planes and the 8-bit CPU-buffer entry points are refused.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from oracle import oracle as O
from pyrowave_amd import synth

from helpers import dev_planes16 as _dev16, target16 as _target16

pytestmark = pytest.mark.gpu

ORACLE_PRECISION = {pw.PRECISION_FP32: 2, pw.PRECISION_FP16: 1, pw.PRECISION_FP16_LL: 3, pw.PRECISION_FP16_MATH: 0}


def planes16(w, h, chroma, seed):
    """Full-range 16-bit content: a synthetic frame scaled by 257 plus
    +-200 of noise (16-bit detail below one 8-bit step)."""
    rng = np.random.default_rng(seed)
    base = synth.frame(w, h, chroma == pw.CHROMA_420, seed=seed)
    return [(p.astype(np.int64) * 257 + rng.integers(-200, 201, p.shape)).clip(0, 65535).astype(np.uint16)
            for p in base]


def _encode16(dev, w, h, chroma, precision, frame, budget, boundary=1 << 30):
    import torch
    enc = pw.Encoder(dev, w, h, chroma, precision=precision, flags=pw.UNORM16_INPUT)
    try:
        pw.encode_batch([enc], [frame], [budget])
        torch.cuda.synchronize()
        out, pk = enc.packetize(boundary, budget + 64)
        return out[:sum(s for _, s in pk)].copy(), pk, enc.read_coefficients()
    finally:
        enc.close()


def _decode16(dev, w, h, chroma, precision, data, pk):
    import torch
    dec = pw.Decoder(dev, w, h, chroma, precision=precision, flags=pw.UNORM16_OUTPUT)
    try:
        for off, size in pk:
            assert dec.push_packet(data[off:off + size]) == pw.SUCCESS
        outs, tgt = _target16(w, h, chroma)
        pw.decode_batch([dec], [tgt])
        torch.cuda.synchronize()
        return [o.cpu().numpy().view(np.uint16) for o in outs]
    finally:
        dec.close()


@pytest.mark.parametrize("precision", [pw.PRECISION_FP32, pw.PRECISION_FP16, pw.PRECISION_FP16_LL,
                                       pw.PRECISION_FP16_MATH])
@pytest.mark.parametrize("w,h,chroma,budget", [(34, 30, pw.CHROMA_420, 4000), (200, 130, pw.CHROMA_444, 12000),
                                               (1920, 1080, pw.CHROMA_420, 400000)])
def test_unorm16_encode_decode_vs_oracle(device, w, h, chroma, budget, precision):
    planes = planes16(w, h, chroma, seed=w + 3 * h)
    keep, frame = _dev16(planes, pad=3)  # keep: the device planes stay alive
    data, pk, _ = _encode16(device, w, h, chroma, precision, frame, budget)
    L = O.layout(w, h, chroma, precision=0 if precision == pw.PRECISION_FP16_MATH else None)
    op = ORACLE_PRECISION[precision]
    bs, meta, _ = O.encode(L, planes, budget, sequence=1, precision=op)
    ref, opk = O.packetize(L, meta, bs, 1, 1 << 30)
    ref = ref[:sum(s for _, s in opk)]
    assert 0 < data.size <= budget
    assert np.array_equal(data, ref)
    got = _decode16(device, w, h, chroma, precision, data, pk)
    exp = O.decode(L, [data[o:o + s].tobytes() for o, s in pk], precision=op, bits=16)
    for c in range(3):
        assert np.array_equal(got[c], exp[c]), c


@pytest.mark.parametrize("precision", [pw.PRECISION_FP32, pw.PRECISION_FP16_MATH])
def test_unorm16_every_sample_value(device, precision):
    """A 256x256 luma plane holding each 16-bit value once (shuffled), so the
    level-0 conversion meets all 65536 inputs: stored coefficients equal the
    oracle's bit for bit."""
    w = h = 256
    rng = np.random.default_rng(5)
    y = rng.permutation(65536).astype(np.uint16).reshape(h, w)
    cb = rng.integers(0, 65536, (h // 2, w // 2), dtype=np.uint16)
    planes = [y, cb, cb[::-1].copy()]
    keep, frame = _dev16(planes)
    _, _, coeffs = _encode16(device, w, h, pw.CHROMA_420, precision, frame, 30000)
    L = O.layout(w, h, 0, precision=0 if precision == pw.PRECISION_FP16_MATH else None)
    ref = O.dwt_frame(L, planes, precision=ORACLE_PRECISION[precision])
    if precision == pw.PRECISION_FP16_MATH:
        ref = np.float16(ref).astype(np.float32)
    mism = np.flatnonzero(coeffs.view(np.uint32) != ref.view(np.uint32))
    assert mism.size == 0, (mism.size, mism[:8])


def test_unorm16_p016_interleaved_chroma(device):
    """4:2:0 with the CbCr plane interleaved (P016): pixel stride 4 bytes, Cr at
    data + 2; equal to the planar encode."""
    import torch
    w, h, budget = 416, 240, 30000
    planes = planes16(w, h, pw.CHROMA_420, seed=8)
    keep, planar = _dev16(planes)
    planar_bytes = _encode16(device, w, h, pw.CHROMA_420, pw.PRECISION_FP16, planar, budget)[0]
    ch, cw = planes[1].shape
    y = torch.from_numpy(planes[0].view(np.int16)).cuda()
    uv = np.empty((ch, 2 * cw), np.uint16)
    uv[:, 0::2], uv[:, 1::2] = planes[1], planes[2]
    uvt = torch.from_numpy(uv.view(np.int16)).cuda()
    frame = pw.hip_frame([(y.data_ptr(), 2 * y.stride(0), 2), (uvt.data_ptr(), 2 * uvt.stride(0), 4),
                          (uvt.data_ptr() + 2, 2 * uvt.stride(0), 4)])
    data, _, _ = _encode16(device, w, h, pw.CHROMA_420, pw.PRECISION_FP16, frame, budget)
    assert np.array_equal(data, planar_bytes)


def test_unorm16_257x_equals_8bit(device):
    """16-bit planes 257 * u are the 8-bit planes u exactly (u / 255 ==
    257u / 65535): same bitstream as the 8-bit encoder, in a batch that mixes
    an 8-bit and a 16-bit encoder (each equal to its encode alone)."""
    import torch
    w, h, budget = 640, 360, 60000
    p8 = synth.frame(w, h, True, seed=12)
    p16 = [p.astype(np.uint16) * 257 for p in p8]
    enc8 = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP16)
    enc16 = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP16, flags=pw.UNORM16_INPUT)
    try:
        t8 = [torch.from_numpy(p).cuda() for p in p8]
        f8 = pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in t8])
        keep, f16 = _dev16(p16)
        pw.encode_batch([enc8, enc16], [f8, f16], [budget, budget])
        torch.cuda.synchronize()
        outs = []
        for e in (enc8, enc16):
            out, pk = e.packetize(1 << 30, budget + 64)
            outs.append(out[:sum(s for _, s in pk)].copy())
        assert np.array_equal(outs[0], outs[1])
        alone = _encode16(device, w, h, pw.CHROMA_420, pw.PRECISION_FP16, f16, budget)[0]
        # The alone encode is its encoder's first frame: sequence 1, as in the batch.
        assert np.array_equal(alone, outs[1])
    finally:
        enc8.close()
        enc16.close()


def test_unorm16_rejects_bad_planes(device):
    import torch
    w, h = 64, 48
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32, flags=pw.UNORM16_INPUT)
    dec = pw.Decoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32, flags=pw.UNORM16_OUTPUT)
    try:
        planes = planes16(w, h, pw.CHROMA_420, seed=1)
        ts, frame = _dev16(planes)
        bad = [(ts[0].data_ptr() + 1, 2 * ts[0].stride(0), 2)] + [(t.data_ptr(), 2 * t.stride(0), 2) for t in ts[1:]]
        with pytest.raises(pw.PyroWaveError):
            pw.encode_batch([enc], [pw.hip_frame(bad)], [4000])
        short = [(ts[0].data_ptr(), w, 2)] + [(t.data_ptr(), 2 * t.stride(0), 2) for t in ts[1:]]
        with pytest.raises(pw.PyroWaveError):  # row stride of 8-bit samples
            pw.encode_batch([enc], [pw.hip_frame(short)], [4000])
        # The pyrowave.h CPU-buffer formats are 8-bit.
        assert enc.encode_cpu([p.astype(np.uint8) for p in planes], 4000) == pw.ERROR_INVALID_ARGUMENT
        pw.encode_batch([enc], [frame], [4000])
        torch.cuda.synchronize()
        out, pk = enc.packetize(1 << 30, 4064)
        for off, size in pk:
            assert dec.push_packet(out[off:off + size]) == pw.SUCCESS
        outs, tgt = _target16(w, h, pw.CHROMA_420)
        odd = pw.hip_target([(outs[0].data_ptr(), 2 * w - 1)] + [(o.data_ptr(), 2 * o.stride(0)) for o in outs[1:]])
        with pytest.raises(pw.PyroWaveError):
            pw.decode_batch([dec], [odd])
        pw.decode_batch([dec], [tgt])
        torch.cuda.synchronize()
    finally:
        enc.close()
        dec.close()
