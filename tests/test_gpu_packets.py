"""Packetize and packet parsing on the device (MI355X).

- pyrowave_hip_encoder_packetize_batch writes the same bytes and the same
  packet table as the host packetize (reference pyrowave_encoder.cpp:596-623,
  1048-1115) at boundaries from 0 to 1 GiB, for a batch of frames of
  different sizes in one launch.
- pyrowave_hip_decode_packets_batch, fed the encoder's device stream with no
  host copy, decodes the same planes as the host parser
  (pyrowave_decoder_push_packet) followed by the decode; so does a reordered
  table, a table with every packet twice (the first copy of a block wins), and
  a table with packets dropped (missing blocks decode as zero on both paths).
- Parse errors: a block index past the frame, a packet past the data, a
  second sequence number, trailing bytes and a sequence header for another
  size are reported; blocks present / announced match the stream.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth

from helpers import decode_bytes, ramp_planes

pytestmark = pytest.mark.gpu

FRAMES = [(34, 30, pw.CHROMA_420, 64 * 1024), (200, 130, pw.CHROMA_444, 30000),
          (1920, 1080, pw.CHROMA_420, 400000)]
BOUNDARIES = [0, 64, 1000, 8192, 1 << 30]


def _planes(w, h, chroma, seed):
    if w * h < 4096:
        return ramp_planes(w, h, chroma)
    return synth.frame(w, h, chroma == pw.CHROMA_420, seed=seed)


def _host_stream(enc, boundary, budget):
    out, pk = enc.packetize(boundary, budget + 64)
    return out[:sum(s for _, s in pk)], pk


def _dev_targets(torch, w, h, chroma):
    shapes = [pw.plane_shape(w, h, chroma, c) for c in range(3)]
    outs = [torch.zeros(s, dtype=torch.uint8, device="cuda") for s in shapes]
    return outs, pw.hip_target([(o.data_ptr(), o.stride(0)) for o in outs])


def _upload(torch, data, table):
    """Device copies of a byte stream and a {offset, size} table."""
    d = torch.from_numpy(np.frombuffer(np.ascontiguousarray(data).tobytes() + b"\0" * 16, np.uint8).copy()).cuda()
    t = torch.tensor(np.array(table, np.int64).reshape(-1).astype(np.uint32).view(np.int32), device="cuda")
    c = torch.tensor([len(table)], dtype=torch.int32, device="cuda")
    ps = pw.packet_stream(d.data_ptr(), len(data), t.data_ptr(), c.data_ptr(), max(len(table), 1))
    return ps, (d, t, c)


@pytest.fixture(scope="module")
def encoded(device):
    """Each frame encoded once through the host path."""
    encs, res = [], []
    for i, (w, h, c, budget) in enumerate(FRAMES):
        enc = pw.Encoder(device, w, h, c, precision=pw.PRECISION_FP32)
        planes = _planes(w, h, c, seed=40 + i)
        assert enc.encode_cpu(planes, budget) == pw.SUCCESS
        encs.append(enc)
        res.append(planes)
    yield encs, res
    for e in encs:
        e.close()


@pytest.mark.parametrize("boundary", BOUNDARIES)
def test_device_packetize_matches_host(encoded, boundary):
    import torch
    encs, _ = encoded
    s = torch.cuda.Stream()
    pw.packetize_batch(encs, boundary, stream=s.cuda_stream)
    for enc, (w, h, c, budget) in zip(encs, FRAMES):
        data, pk = pw.read_packets(enc)
        ref, ref_pk = _host_stream(enc, boundary, budget)
        assert np.array_equal(data, ref), (w, h, boundary)
        assert pk == [(int(o), int(z)) for o, z in ref_pk], (w, h, boundary)


def test_device_parse_decodes_like_host(device, encoded):
    import torch
    encs, _ = encoded
    s = torch.cuda.Stream()
    pw.packetize_batch(encs, 1000, stream=s.cuda_stream)
    decs = [pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP32) for w, h, c, _ in FRAMES]
    host = [pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP32) for w, h, c, _ in FRAMES]
    try:
        targets = [_dev_targets(torch, w, h, c) for w, h, c, _ in FRAMES]
        streams = [pw.device_packets(e)[0] for e in encs]
        pw.decode_packets_batch(decs, streams, [t for _, t in targets], stream=s.cuda_stream)
        torch.cuda.synchronize()
        for enc, dec, hd, (outs, _), (w, h, c, budget) in zip(encs, decs, host, targets, FRAMES):
            data, pk = pw.read_packets(enc)
            ref = decode_bytes(hd, data, pk)
            for p in range(3):
                assert np.array_equal(outs[p].cpu().numpy(), ref[p]), (w, h, p)
            r, present, announced = pw.parse_result(dec)
            assert r == pw.SUCCESS
            nonzero = int(np.frombuffer(data[4:8].tobytes(), np.uint32)[0]) & 0xffffff
            assert present == nonzero and announced == nonzero, (present, announced, nonzero)
    finally:
        for d in decs + host:
            d.close()


@pytest.mark.parametrize("mode", ["reversed", "twice", "dropped"])
def test_device_parse_tables(device, encoded, mode):
    import torch
    encs, _ = encoded
    w, h, c, budget = FRAMES[2]
    enc = encs[2]
    s = torch.cuda.Stream()
    pw.packetize_batch([enc], 8192, stream=s.cuda_stream)
    data, pk = pw.read_packets(enc)
    assert len(pk) > 8
    if mode == "reversed":
        table = pk[::-1]
        push = pk
    elif mode == "twice":
        table = pk + pk
        push = pk + pk
    else:
        keep = [i for i in range(len(pk)) if i == 0 or i % 3]
        table = [pk[i] for i in keep]
        push = table
    dec = pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP32)
    hd = pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP32)
    try:
        ps, keep_alive = _upload(torch, data, table)
        outs, tgt = _dev_targets(torch, w, h, c)
        pw.decode_packets_batch([dec], [ps], [tgt], stream=s.cuda_stream)
        torch.cuda.synchronize()
        ref = decode_bytes(hd, data, push)
        for p in range(3):
            assert np.array_equal(outs[p].cpu().numpy(), ref[p]), (mode, p)
        r, present, announced = pw.parse_result(dec)
        assert r == pw.SUCCESS
        if mode == "dropped":
            assert present < announced
        else:
            assert present == announced
    finally:
        dec.close()
        hd.close()


def _words(b):
    return np.frombuffer(np.ascontiguousarray(b).tobytes(), np.uint32).copy()


@pytest.mark.parametrize("fault", ["block_index", "past_data", "sequence", "trailing", "header_size"])
def test_device_parse_errors(device, encoded, fault):
    import torch
    encs, _ = encoded
    w, h, c, budget = FRAMES[1]
    enc = encs[1]
    s = torch.cuda.Stream()
    pw.packetize_batch([enc], 1000, stream=s.cuda_stream)
    data, pk = pw.read_packets(enc)
    words = _words(data)
    table = list(pk)
    n32 = None
    if fault == "block_index":
        # First block after the sequence header: index field past the frame.
        words[3] = (words[3] & 0xff) | (0xffffff << 8)
    elif fault == "past_data":
        table[-1] = (table[-1][0], table[-1][1] + 64)
    elif fault == "sequence":
        words[2] ^= 1 << 28
    elif fault == "trailing":
        table[0] = (table[0][0], table[0][1] + 4) if len(table) == 1 else (table[0][0], table[0][1] - 4)
    elif fault == "header_size":
        words[0] ^= 1  # width - 1 off by one
    dec = pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP32)
    try:
        ps, keep_alive = _upload(torch, words.view(np.uint8), table)
        outs, tgt = _dev_targets(torch, w, h, c)
        pw.decode_packets_batch([dec], [ps], [tgt], stream=s.cuda_stream)
        torch.cuda.synchronize()
        r, present, announced = pw.parse_result(dec)
        assert r == pw.ERROR_INVALID_ARGUMENT, (fault, r, present, announced)
    finally:
        dec.close()


def test_device_parse_rejects_bad_arguments(device):
    import torch
    dec = pw.Decoder(device, 64, 64, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    try:
        outs, tgt = _dev_targets(torch, 64, 64, pw.CHROMA_420)
        d = torch.zeros(64, dtype=torch.uint8, device="cuda")
        t = torch.zeros(2, dtype=torch.int32, device="cuda")
        ps = pw.packet_stream(d.data_ptr() + 1, 32, t.data_ptr(), t.data_ptr(), 1)  # misaligned data
        with pytest.raises(pw.PyroWaveError):
            pw.decode_packets_batch([dec], [ps], [tgt])
        assert pw.parse_result(dec)[0] == pw.ERROR_GENERIC  # no device parse yet
    finally:
        dec.close()
