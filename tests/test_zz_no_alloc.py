"""pyrowave_hip_decode_packets_noalloc (GPU): the allocation-free decode equals
pyrowave_hip_decode_packets_batch on two calls in a row, the second with a
smaller frame, its info words equal the batch path's parse result, device
memory does not change across calls, and frames beyond its scratch are
refused."""
import ctypes as C

import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth

pytestmark = pytest.mark.gpu

# Large first, then smaller: the second call reuses the shared scratch, whose
# first entries still hold the first frame's offsets.
CASES = [(1920, 1080, pw.CHROMA_420, 400000, pw.PRECISION_FP16),
         (640, 360, pw.CHROMA_444, 60000, pw.PRECISION_FP32)]


def _targets(torch, w, h, chroma):
    shapes = [pw.plane_shape(w, h, chroma, c) for c in range(3)]
    outs = [torch.full(s, 7, dtype=torch.uint8, device="cuda") for s in shapes]
    return outs, pw.hip_target([(o.data_ptr(), o.stride(0)) for o in outs])


def _encode(device, torch, w, h, chroma, budget, precision, seed, boundary=8192):
    enc = pw.Encoder(device, w, h, chroma, precision=precision)
    planes = synth.frame(w, h, chroma == pw.CHROMA_420, seed=seed)
    assert enc.encode_cpu(planes, budget) == pw.SUCCESS
    pw.packetize_batch([enc], boundary)
    torch.cuda.synchronize()
    return enc


def test_noalloc_decode_matches_batch_twice(device):
    import torch
    encs, decs = [], []
    try:
        for i, (w, h, c, budget, prec) in enumerate(CASES):
            encs.append(_encode(device, torch, w, h, c, budget, prec, seed=70 + i))
            decs.append((pw.Decoder(device, w, h, c, precision=prec), pw.Decoder(device, w, h, c, precision=prec)))
        for rnd in range(2):
            refs, runs = [], []
            for enc, (ref_dec, na_dec), (w, h, c, budget, prec) in zip(encs, decs, CASES):
                ps = pw.device_packets(enc)[0]
                ref_out, ref_tgt = _targets(torch, w, h, c)
                pw.decode_packets_batch([ref_dec], [ps], [ref_tgt])
                torch.cuda.synchronize()
                refs.append((ref_out, pw.parse_result(ref_dec)))
                out, tgt = _targets(torch, w, h, c)
                info = torch.full((4,), -1, dtype=torch.int32, device="cuda")
                runs.append((na_dec, ps, out, tgt, info))
            torch.cuda.synchronize()
            free_before = torch.cuda.mem_get_info()[0]
            for (na_dec, ps, out, tgt, info), (w, h, c, _, _) in zip(runs, CASES):
                rc = pw.decode_packets_noalloc(na_dec, ps, tgt, C.c_void_p(info.data_ptr()))
                assert rc == pw.SUCCESS, (w, h, rc)
                torch.cuda.synchronize()
            # No device allocation happened in this round's calls.
            assert torch.cuda.mem_get_info()[0] == free_before
            for (na_dec, ps, out, tgt, info), (ref_out, (r, present, announced)), (w, h, c, _, _) in zip(
                    runs, refs, CASES):
                assert r == pw.SUCCESS
                for p in range(3):
                    assert torch.equal(out[p], ref_out[p]), (rnd, w, h, p)
                got = info.cpu().numpy().view(np.uint32)
                assert got[0] == present and got[1] == announced and got[2] == 0, (got, present, announced)
    finally:
        for e in encs:
            e.close()
        for a, b in decs:
            a.close()
            b.close()


def test_noalloc_decode_without_info_and_on_a_stream(device):
    import torch
    w, h, c, budget, prec = CASES[1]
    enc = _encode(device, torch, w, h, c, budget, prec, seed=80, boundary=1000)
    ref_dec = pw.Decoder(device, w, h, c, precision=prec)
    dec = pw.Decoder(device, w, h, c, precision=prec)
    try:
        ps = pw.device_packets(enc)[0]
        ref_out, ref_tgt = _targets(torch, w, h, c)
        pw.decode_packets_batch([ref_dec], [ps], [ref_tgt])
        s = torch.cuda.Stream()
        out, tgt = _targets(torch, w, h, c)
        torch.cuda.synchronize()
        assert pw.decode_packets_noalloc(dec, ps, tgt, None, C.c_void_p(s.cuda_stream)) == pw.SUCCESS
        torch.cuda.synchronize()
        for p in range(3):
            assert torch.equal(out[p], ref_out[p])
    finally:
        enc.close()
        ref_dec.close()
        dec.close()


def test_noalloc_decode_refuses_frames_beyond_its_scratch(device):
    import torch
    # 8192x8192 4:4:4 has about 199,000 32x32 blocks, beyond the 131072 the
    # scratch holds.
    w = h = 8192
    dec = pw.Decoder(device, w, h, pw.CHROMA_444, precision=pw.PRECISION_FP16_MATH)
    try:
        d = torch.zeros(64, dtype=torch.uint8, device="cuda")
        t = torch.zeros(2, dtype=torch.int32, device="cuda")
        cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
        ps = pw.packet_stream(d.data_ptr(), 64, t.data_ptr(), cnt.data_ptr(), 1)
        out, tgt = _targets(torch, 64, 64, pw.CHROMA_444)
        big = pw.hip_target([(out[0].data_ptr(), w)] * 3)
        assert pw.decode_packets_noalloc(dec, ps, big, None) == pw.ERROR_INVALID_ARGUMENT
    finally:
        dec.close()
