"""Serving loop on the MI355X (pyrowave_hip_server_* entry points).

Sixteen and sixty-four sessions of mixed resolution and chroma (1280x720
4:2:0, 1920x1080 4:2:0, 2560x1440 4:4:4, 3840x2160 4:2:0) and cadence (60 or
120 fps) are admitted against an HBM budget of exactly their summed
footprint.  Frames
arrive on a Poisson clock per session with 1 ms Gaussian jitter and budgets of
1.2 to 2.0 bits per pixel; the server batches whatever is ready over four
streams.  Every frame's packetized bitstream and decoded planes must equal the
same session's frames encoded and decoded alone, in order, by a fresh encoder
and decoder (bit-exact: batching changes nothing in the arithmetic).  The run
also checks that frames of different sessions did share launches, that the
budget refuses one more session until one closes, and prints p50/p99 latency.
"""
import time

import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth

from helpers import decode_bytes, gpu_encode

pytestmark = pytest.mark.gpu

SHAPES = [(1280, 720, pw.CHROMA_420), (1920, 1080, pw.CHROMA_420), (2560, 1440, pw.CHROMA_444),
          (3840, 2160, pw.CHROMA_420)]


def _schedule(n_sessions, n_frames, seed):
    rng = np.random.default_rng(seed)
    sessions = []
    for i in range(n_sessions):
        w, h, c = SHAPES[i % len(SHAPES)]
        fps = (60, 120)[(i // len(SHAPES)) % 2]
        bpp = rng.uniform(1.2, 2.0)
        t = rng.uniform(0, 1.0 / fps)
        times = []
        for _ in range(n_frames):
            times.append(max(0.0, t + rng.normal(0, 1e-3)))
            t += rng.exponential(1.0 / fps)
        sessions.append(dict(shape=(w, h, c), budget=int(w * h * bpp / 8), times=sorted(times), index=i))
    return sessions


@pytest.mark.parametrize("n_sessions,n_frames", [(16, 4), (64, 3)])
def test_serving_mixed_sessions_match_alone(device, n_sessions, n_frames):
    """16 sessions x 4 frames, and the 64-session serving configuration (16 of
    each shape, 60 and 120 fps, under a budget of their summed footprint,
    about 4.3 GB) x 3 frames."""
    import torch
    plan = _schedule(n_sessions, n_frames, seed=77)
    gens = {s: synth.SceneGenerator(s[0], s[1], s[2] == pw.CHROMA_420, seed=500 + k, frames=112, pan=3,
                                    redraw=0.05) for k, s in enumerate(SHAPES)}
    # A pool of frames per shape (generating 192 distinct 4K frames would
    # dominate the test); sessions take different frames of it.
    pool = {shape: [g.frame(t) for t in range(0, 112, 14)] for shape, g in gens.items()}
    contents = [[pool[s["shape"]][(k + 3 * s["index"]) % len(pool[s["shape"]])] for k in range(n_frames)]
                for s in plan]

    # Alone: one fresh encoder/decoder per session, frames in order.
    alone = []
    for s, frames in zip(plan, contents):
        w, h, c = s["shape"]
        enc = pw.Encoder(device, w, h, c, precision=pw.PRECISION_FP16)
        dec = pw.Decoder(device, w, h, c, precision=pw.PRECISION_FP16)
        try:
            res = []
            for planes in frames:
                data, pk = gpu_encode(enc, planes, s["budget"])
                res.append((data, decode_bytes(dec, data, pk)))
            alone.append(res)
        finally:
            enc.close()
            dec.close()

    # In the mix.
    probe = pw.Server(device, 1 << 40)
    footprints = [probe.footprint(*s["shape"]) for s in plan]
    probe.close()
    server = pw.Server(device, sum(footprints), streams=4, precision=pw.PRECISION_FP16, packet_boundary=8192)
    try:
        sids = [server.open_session(*s["shape"]) for s in plan]
        assert all(sid is not None for sid in sids)
        assert server.open_session(1280, 720, pw.CHROMA_420) is None  # budget exhausted

        inputs = [[[torch.from_numpy(p).cuda() for p in planes] for planes in frames] for frames in contents]
        outputs = [[[torch.zeros_like(p) for p in ts] for ts in frames] for frames in inputs]
        torch.cuda.synchronize()
        events = sorted((t, si, k) for si, s in enumerate(plan) for k, t in enumerate(s["times"]))
        ticket_of = {}
        done = {}
        t0 = time.perf_counter()
        pending = list(events)
        while pending or len(done) < len(events):
            now = time.perf_counter() - t0
            while pending and pending[0][0] <= now:
                _, si, k = pending.pop(0)
                ts, os_ = inputs[si][k], outputs[si][k]
                frame = pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in ts])
                target = pw.hip_target([(o.data_ptr(), o.stride(0)) for o in os_])
                ticket_of[server.submit(sids[si], frame, plan[si]["budget"], target)] = (si, k)
            wait = int(max(0.0, (pending[0][0] - now) * 1e6)) if pending else 100000
            for c in server.poll(64, timeout_us=min(wait, 100000)):
                assert c.status == pw.SUCCESS, (ticket_of[c.ticket], c.status)
                done[c.ticket] = (c, server.take_bitstream(c))
            assert time.perf_counter() - t0 < 90, "serving run did not finish"
        wall = time.perf_counter() - t0
        torch.cuda.synchronize()

        lat = []
        starts = {}
        for ticket, (c, (data, pk)) in done.items():
            si, k = ticket_of[ticket]
            ref_data, ref_planes = alone[si][k]
            assert np.array_equal(data, ref_data), (si, k)
            assert sum(s for _, s in pk) == data.size and all(s <= 8192 for _, s in pk[:-1]), (si, k)
            for p in range(3):
                assert np.array_equal(outputs[si][k][p].cpu().numpy(), ref_planes[p]), (si, k, p)
            lat.append((c.done_ns - c.submit_ns) / 1e3)
            starts.setdefault(c.start_ns, []).append(ticket)
        assert max(len(v) for v in starts.values()) > 1, "no launch carried frames of more than one session"
        lat = np.array(lat)
        if n_sessions == 64:
            assert sum(footprints) > 4e9
        print(f"serving {n_sessions} sessions: {len(lat)} frames in {wall:.3f} s, {len(starts)} launches, "
              f"latency p50 {np.percentile(lat, 50):.0f} us p99 {np.percentile(lat, 99):.0f} us")

        # A closed session returns its footprint to the budget.
        server.close_session(sids[0])
        again = server.open_session(*plan[0]["shape"])
        assert again is not None
        server.close_session(again)
    finally:
        server.close()


def test_serving_rejects_bad_requests(device):
    server = pw.Server(device, 1 << 34)
    try:
        sid = server.open_session(1920, 1080, pw.CHROMA_420)
        with pytest.raises(pw.PyroWaveError):
            server.submit(sid + 100, pw.hip_frame([(1, 1920, 1), (1, 960, 1), (1, 960, 1)]), 100000)
        with pytest.raises(pw.PyroWaveError):  # row stride shorter than the plane
            server.submit(sid, pw.hip_frame([(1, 1000, 1), (1, 960, 1), (1, 960, 1)]), 100000)
        with pytest.raises(pw.PyroWaveError):  # zero budget
            server.submit(sid, pw.hip_frame([(1, 1920, 1), (1, 960, 1), (1, 960, 1)]), 3)
        with pytest.raises(pw.PyroWaveError):  # odd 4:2:0 size
            server.open_session(1921, 1080, pw.CHROMA_420)
        assert server.poll(8) == []
        server.close_session(sid)
        with pytest.raises(pw.PyroWaveError):
            server.close_session(sid)
    finally:
        server.close()
