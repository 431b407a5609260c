"""Device packetize and parse against the oracle and the reference's own
parser (GPU tests).

- k_packetize (pyrowave_hip_encoder_packetize_batch) writes the stream and
  packet table that the oracle's packetize (a restatement of
  pyrowave_encoder.cpp:596-623, 1048-1115, pinned by golden fixtures of the
  reference's metal/pyrowave_bitstream.cpp) writes for the oracle's own
  bitstream, at boundaries from 0 (one block per packet) through sizes below
  the largest block (oversize blocks get packets of their own) to 1 GiB.
- k_parse (pyrowave_hip_decode_packets_batch) replays the recorded scripts of
  the reference's BitstreamParser (tests/golden/parser.npz) wherever a script
  stays within one sequence number: blocks present and announced equal the
  reference parser's after the last push, a script with a rejected push is a
  parse error, and a clean script decodes to the oracle's decode of the
  reference parser's payload and offsets.  Scripts that change sequence are
  parse errors on the device (its contract: one frame's packets).
- The opt-in exact-prefix rate control (PYROWAVE_HIP_ENCODER_EXACT_PREFIX)
  equals the oracle's finalize(q1=False) and its bitstream.
"""
import os

import numpy as np
import pytest

import pyrowave_amd as pw
from oracle import oracle as O
from pyrowave_amd import synth

from helpers import oracle_encode

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

FRAMES = [(34, 30, pw.CHROMA_420, 64 * 1024), (200, 130, pw.CHROMA_444, 30000),
          (1920, 1080, pw.CHROMA_420, 400000)]
BOUNDARIES = [0, 8, 64, 300, 1000, 4096, 8192, 65536, 1 << 30]


def _planes(w, h, chroma, seed):
    if w * h < 4096:
        from helpers import ramp_planes
        return ramp_planes(w, h, chroma)
    return synth.frame(w, h, chroma == pw.CHROMA_420, seed=seed)


@pytest.fixture(scope="module")
def encoded(device):
    encs, refs = [], []
    for i, (w, h, c, budget) in enumerate(FRAMES):
        planes = _planes(w, h, c, seed=60 + i)
        enc = pw.Encoder(device, w, h, c, precision=pw.PRECISION_FP32)
        assert enc.encode_cpu(planes, budget) == pw.SUCCESS
        L = O.layout(w, h, c)
        bs, meta, _ = O.encode(L, planes, budget, sequence=1)
        encs.append(enc)
        refs.append((L, bs, meta))
    yield encs, refs
    for e in encs:
        e.close()


@pytest.mark.parametrize("boundary", BOUNDARIES)
def test_device_packetize_matches_oracle(encoded, boundary):
    import torch
    encs, refs = encoded
    s = torch.cuda.Stream()
    pw.packetize_batch(encs, boundary, stream=s.cuda_stream)
    for enc, (L, bs, meta), (w, h, c, _) in zip(encs, refs, FRAMES):
        data, pk = pw.read_packets(enc)
        out, opk = O.packetize(L, meta, bs, 1, boundary)
        assert np.array_equal(data, out[:int(opk[:, 1].sum())]), (w, h, boundary)
        assert pk == [(int(o), int(z)) for o, z in opk], (w, h, boundary)
        assert len(pk) == O.compute_num_packets(L, meta, boundary)


_G = np.load(os.path.join(GOLDEN, "parser.npz"))
NAMES = [str(n) for n in _G["names"]]


def _script_pushes(name):
    """Packets pushed after the script's last clear (the device parse is clear
    + push of every packet), the index of the last op, and whether a mark-
    decoded op lies between them (it does not change the parse state)."""
    ops, lens = _G[f"{name}_ops"], _G[f"{name}_lens"]
    data = _G[f"{name}_data"].tobytes()
    starts = np.concatenate([[0], np.cumsum(lens)])
    first = 0
    for k, op in enumerate(ops):
        if op == 1:
            first = k + 1
    pushes = [(k, data[starts[k]:starts[k + 1]]) for k in range(first, len(ops)) if ops[k] == 0]
    return pushes


def _seqs(packet):
    """Sequence numbers of the units a packet starts with (header or block)."""
    out, u = set(), 0
    while len(packet) - u >= 8:
        w0, w1 = np.frombuffer(packet[u:u + 8], np.uint32)
        out.add(int(w0 >> 28) & 7)
        if w0 >> 31:
            u += 8
        else:
            n = int((w0 >> 16) & 0xfff) * 4
            if n < 8:
                break
            u += n
    return out


@pytest.mark.parametrize("name", NAMES)
def test_device_parse_replays_reference_parser(device, name):
    import torch
    w, h, ch = int(_G["w"]), int(_G["h"]), int(_G["chroma"])
    steps = _G[f"{name}_steps"]
    pushes = _script_pushes(name)
    seqs = set()
    for _, p in pushes:
        seqs |= _seqs(p)
    # Packets are 4-byte aligned in the device stream: pad each to a word.
    blob, table = b"", []
    for _, p in pushes:
        table.append((len(blob), len(p)))
        blob += p + b"\0" * (-len(p) % 4)
    d = torch.from_numpy(np.frombuffer(blob + b"\0" * 16, np.uint8).copy()).cuda()
    t = torch.tensor(np.array(table, np.int64).reshape(-1).astype(np.uint32).view(np.int32) if table
                     else np.zeros(2, np.int32), device="cuda")
    c = torch.tensor([len(table)], dtype=torch.int32, device="cuda")
    ps = pw.packet_stream(d.data_ptr(), len(blob), t.data_ptr(), c.data_ptr(), max(len(table), 1))
    dec = pw.Decoder(device, w, h, ch, precision=pw.PRECISION_FP32)
    try:
        shapes = [pw.plane_shape(w, h, ch, k) for k in range(3)]
        outs = [torch.zeros(s, dtype=torch.uint8, device="cuda") for s in shapes]
        pw.decode_packets_batch([dec], [ps], [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in outs])])
        torch.cuda.synchronize()
        r, present, announced = pw.parse_result(dec)
        if len(seqs) > 1:
            assert r == pw.ERROR_INVALID_ARGUMENT, name
            return
        rejected = any(steps[k][0] == 0 for k, _ in pushes)
        last = steps[-1]
        assert (r == pw.ERROR_INVALID_ARGUMENT) == rejected, (name, r)
        assert present == int(last[3]), (name, present, last)
        if pushes:
            assert announced == int(last[4]), (name, announced, last)
        if not rejected and pushes:
            # The device reads blocks where they lie in its packet buffer; the
            # reference copies first arrivals into a payload.  The same blocks
            # (the reference's present set) at their first copy's position in
            # the buffer must decode like the oracle decodes them there.
            L = O.layout(w, h, ch)
            ref_present = _G[f"{name}_offsets"] != 0xffffffff
            first = np.full(L.count32, 0xffffffff, np.uint32)
            for off, size in table:
                u = 0
                while size - u >= 8:
                    w0, w1 = np.frombuffer(blob[off + u:off + u + 8], np.uint32)
                    if w0 >> 31:
                        u += 8
                        continue
                    blk = int(w1 >> 8)
                    if first[blk] == 0xffffffff:
                        first[blk] = (off + u) // 4
                    u += int((w0 >> 16) & 0xfff) * 4
            assert np.array_equal(first != 0xffffffff, ref_present), name
            words = np.frombuffer(blob + b"\0" * (-len(blob) % 4), np.uint32)
            exp = O.idwt_frame(L, O.dequant(L, words, first))
            for k in range(3):
                assert np.array_equal(outs[k].cpu().numpy(), exp[k]), (name, k)
    finally:
        dec.close()


@pytest.mark.parametrize("merged", ["1", "0"])
def test_exact_prefix_matches_oracle(merged, monkeypatch):
    """The exact-prefix finalize on the device equals the oracle's, differs
    from the reference's scan above flat index 1024 on this content, and its
    bitstream equals the oracle's exact-prefix encode; through the
    one-workgroup resolve and the three launches (PYROWAVE_FUSED_RESOLVE=0)."""
    monkeypatch.setenv("PYROWAVE_FUSED_RESOLVE", merged)
    device = pw.Device()
    w, h, budget = 640, 360, 1000
    planes = synth.frame(w, h, False, seed=5, checker=0.3)
    rng = np.random.default_rng(1)
    planes = [np.clip(p.astype(int) + rng.integers(-40, 41, p.shape), 0, 255).astype(np.uint8) for p in planes]
    L = O.layout(w, h, 1)
    enc = pw.Encoder(device, w, h, pw.CHROMA_444, precision=pw.PRECISION_FP32, flags=pw.EXACT_PREFIX)
    try:
        assert enc.encode_cpu(planes, budget) == pw.SUCCESS
        out, pk = enc.packetize(1 << 30, budget + 64)
        got = out[:sum(s for _, s in pk)]
        rc = enc.read_rate_control(L.count32)
    finally:
        enc.close()
        device.close()
    exact = O.finalize(rc["totals"], q1=False)
    assert np.array_equal(rc["prefix"], exact)
    assert not np.array_equal(exact, O.finalize(rc["totals"], q1=True))
    ref, _, quant = oracle_encode(L, planes, budget, q1=False)
    assert np.array_equal(got, ref)
    assert np.array_equal(rc["quant"], quant)
