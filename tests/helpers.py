"""Shared helpers of the GPU and CPU tests."""
import numpy as np

import pyrowave_amd as pw
from oracle import oracle as O


def chroma_shape(w, h, chroma):
    return (h // 2, w // 2) if chroma == pw.CHROMA_420 else (h, w)


def ramp_planes(w, h, chroma):
    """pyrowave_c_test.cpp:288-308 content: luma 3x+5y, Cb 7x+3y, Cr 3x+5y (mod 256)."""
    ys, xs = np.mgrid[0:h, 0:w]
    y = ((3 * xs + 5 * ys) & 255).astype(np.uint8)
    ch, cw = chroma_shape(w, h, chroma)
    ys, xs = np.mgrid[0:ch, 0:cw]
    cb = ((7 * xs + 3 * ys) & 255).astype(np.uint8)
    cr = ((3 * xs + 5 * ys) & 255).astype(np.uint8)
    return [y, cb, cr]


def gpu_encode(enc, planes, budget, boundary=1 << 30):
    """Encode through the C API and packetize into one buffer; returns (bytes, packets)."""
    assert enc.encode_cpu(planes, budget) == pw.SUCCESS
    out, pk = enc.packetize(boundary, budget + 64)
    total = sum(s for _, s in pk)
    return out[:total].copy(), pk


def oracle_encode(L, planes, budget, sequence=1, q1=True, boundary=1 << 30, precision=2):
    bs, meta, quant = O.encode(L, planes, budget, sequence=sequence, q1=q1, precision=precision)
    out, pk = O.packetize(L, meta, bs, sequence, boundary)
    return out, pk, quant


def decode_bytes(dec, data, packets):
    for off, size in packets:
        assert dec.push_packet(data[off:off + size]) == pw.SUCCESS
    return dec.decode_cpu()


def band_errors(L, a, b):
    """Per-band max-abs and RMS of a - b over the transmitted bands."""
    res = {}
    for c in range(3):
        for l in range(5):
            for bd in range(4):
                B = L.band[c][l][bd]
                if not B.present:
                    continue
                x = O.band_view(L, a, c, l, bd).astype(np.float64)
                y = O.band_view(L, b, c, l, bd).astype(np.float64)
                d = x - y
                res[(c, l, bd)] = (float(np.abs(d).max()), float(np.sqrt(np.mean(d * d))))
    return res


def mask_unstored_ll(L, flat):
    """The flat coefficient array with the LL band of each component's first
    level zeroed: the fused forward launch keeps that intermediate in
    registers and never stores it (every transmitted band, and the LL of the
    later levels, is stored)."""
    out = np.array(flat, copy=True)
    for c in range(3):
        first = 1 if (c and L.chroma == 0) else 0
        B = L.band[c][first][0]
        out[B.coeff_off:B.coeff_off + B.bw * B.bh] = 0
    return out


def dev_planes16(planes, pad=0):
    """u16 device copies of 16-bit planes (row pitch padded by `pad` samples)
    and their pyrowave_hip_frame; keep the tensors alive while it is used."""
    import torch
    ts = []
    for p in planes:
        t = torch.zeros((p.shape[0], p.shape[1] + pad), dtype=torch.int16, device="cuda")
        t[:, :p.shape[1]] = torch.from_numpy(np.ascontiguousarray(p, np.uint16).view(np.int16)).cuda()
        ts.append(t)
    return ts, pw.hip_frame([(t.data_ptr(), 2 * t.stride(0), 2) for t in ts])


def target16(w, h, chroma):
    """Zeroed u16 device planes of a frame and their pyrowave_hip_decode_target."""
    import torch
    outs = [torch.zeros(chroma_shape(w, h, chroma) if c else (h, w), dtype=torch.int16, device="cuda")
            for c in range(3)]
    return outs, pw.hip_target([(o.data_ptr(), 2 * o.stride(0)) for o in outs])
