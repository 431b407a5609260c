"""Graph mode of the batch entry points (MI355X).

A batch of three uneven frames (1280x720 4:2:0, 640x480 4:4:4, 1920x1080
4:2:0, FP16) is encoded and decoded four times with new contents and budgets
each round, once on a device in eager mode and once on a device in graph mode
(the first round captures the launch sequence, later rounds replay it with
the per-call parameter nodes rewritten).  Between rounds the test runs other
eager work and synchronizes.  Every replay's bitstreams and decoded planes
must equal the eager ones; the same holds for decode from host-parsed packets
(pyrowave_hip_decode_batch) and from device packets
(pyrowave_hip_decode_packets_batch) in graph mode.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth

pytestmark = pytest.mark.gpu

SHAPES = [(1280, 720, pw.CHROMA_420), (640, 480, pw.CHROMA_444), (1920, 1080, pw.CHROMA_420)]


def _targets(torch):
    outs, tg = [], []
    for w, h, c in SHAPES:
        o = [torch.zeros(pw.plane_shape(w, h, c, p), dtype=torch.uint8, device="cuda") for p in range(3)]
        outs.append(o)
        tg.append(pw.hip_target([(t.data_ptr(), t.stride(0)) for t in o]))
    return outs, tg


def _run(torch, dev, rounds, inputs, budgets):
    """Encode + decode every round; returns per round [(bytes, packets, planes)] and the
    host-packet and device-packet decodes of the last round."""
    s = torch.cuda.Stream()
    encs = [pw.Encoder(dev, w, h, c, precision=pw.PRECISION_FP16) for w, h, c in SHAPES]
    decs = [pw.Decoder(dev, w, h, c, precision=pw.PRECISION_FP16) for w, h, c in SHAPES]
    decs_host = [pw.Decoder(dev, w, h, c, precision=pw.PRECISION_FP16) for w, h, c in SHAPES]
    decs_dev = [pw.Decoder(dev, w, h, c, precision=pw.PRECISION_FP16) for w, h, c in SHAPES]
    side = pw.Encoder(dev, 64, 64, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    results = []
    try:
        outs, tg = _targets(torch)
        for r in range(rounds):
            fr = [pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts]) for ts in inputs[r]]
            pw.encode_batch(encs, fr, budgets[r], stream=s.cuda_stream)
            pw.decode_encoded_batch(decs, encs, tg, stream=s.cuda_stream)
            torch.cuda.synchronize()
            res = []
            for i, e in enumerate(encs):
                out, pk = e.packetize(4096, budgets[r][i] + 64)
                res.append((out[:sum(z for _, z in pk)].copy(), pk, [o.cpu().numpy() for o in outs[i]]))
            results.append(res)
            # Other eager work between replays.
            side.encode_cpu([np.full((64, 64), 17 * r, np.uint8), np.full((32, 32), 90, np.uint8),
                             np.full((32, 32), 160, np.uint8)], 4096)
            torch.cuda.synchronize()
        # Last round's streams through the host parser and the device parser.
        last = results[-1]
        for d, (data, pk, _) in zip(decs_host, last):
            for o, z in pk:
                assert d.push_packet(data[o:o + z]) == pw.SUCCESS
        outs_h, tg_h = _targets(torch)
        pw.decode_batch(decs_host, tg_h, stream=s.cuda_stream)
        pw.packetize_batch(encs, 4096, stream=s.cuda_stream)
        outs_d, tg_d = _targets(torch)
        pw.decode_packets_batch(decs_dev, [pw.device_packets(e)[0] for e in encs], tg_d, stream=s.cuda_stream)
        torch.cuda.synchronize()
        host_planes = [[o.cpu().numpy() for o in oo] for oo in outs_h]
        dev_planes = [[o.cpu().numpy() for o in oo] for oo in outs_d]
        return results, host_planes, dev_planes
    finally:
        for o in encs + decs + decs_host + decs_dev + [side]:
            o.close()


def test_graph_replays_equal_eager():
    import torch
    rounds = 4
    rng = np.random.default_rng(5)
    inputs, budgets = [], []
    for r in range(rounds):
        inputs.append([[torch.from_numpy(p).cuda() for p in synth.frame(w, h, c == pw.CHROMA_420, seed=70 + 3 * r + i)]
                       for i, (w, h, c) in enumerate(SHAPES)])
        budgets.append([int(w * h * rng.uniform(1.0, 2.5) / 8) for w, h, _ in SHAPES])
    torch.cuda.synchronize()
    eager_dev, graph_dev = pw.Device(), pw.Device()
    graph_dev.set_graphs(True)
    try:
        eager = _run(torch, eager_dev, rounds, inputs, budgets)
        graph = _run(torch, graph_dev, rounds, inputs, budgets)
    finally:
        eager_dev.close()
        graph_dev.close()
    for r in range(rounds):
        for i in range(len(SHAPES)):
            a, b = eager[0][r][i], graph[0][r][i]
            assert np.array_equal(a[0], b[0]), (r, i)
            assert a[1] == b[1], (r, i)
            for p in range(3):
                assert np.array_equal(a[2][p], b[2][p]), (r, i, p)
    for i in range(len(SHAPES)):
        for p in range(3):
            assert np.array_equal(eager[1][i][p], graph[1][i][p]), ("host packets", i, p)
            assert np.array_equal(eager[2][i][p], graph[2][i][p]), ("device packets", i, p)
            # Host-parsed and device-parsed decodes of the same stream agree.
            assert np.array_equal(graph[1][i][p], graph[2][i][p]), ("host vs device parse", i, p)
