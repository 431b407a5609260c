"""A YUV4MPEG2 sequence through the codec and the PYROWAVE container (MI355X).

Three 640x480 4:2:0 frames are written to a y4m file, read back, encoded at
60,000 bytes with 8 KiB packets into a container file, read from the
container into a decoder frame by frame, and written to a y4m file with the
header decode.cpp:195-196 writes.  The decoded planes equal a direct decode
of the same bitstreams.
"""
import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import io as pio
from pyrowave_amd import synth

from helpers import decode_bytes, gpu_encode

pytestmark = pytest.mark.gpu


def test_y4m_container_roundtrip(device, tmp_path):
    w, h, budget = 640, 480, 60000
    src = str(tmp_path / "in.y4m")
    gen = synth.SceneGenerator(w, h, True, seed=9, frames=3, pan=3, redraw=0.05)
    frames = [gen.frame(t) for t in range(3)]
    with pio.Y4MFile(src, "wb", pio.y4m_params(w, h, pio.YUV420P, True, (30, 1))) as f:
        for planes in frames:
            f.write_frame(planes)

    stream = str(tmp_path / "s.pyrowave")
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    direct = pw.Decoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    expected = []
    try:
        with pio.Y4MFile(src) as f, pio.ContainerWriter(stream, f.width, f.height, f.format, pw.CHROMA_420,
                                                        f.full_range, f.frame_rate) as out:
            while True:
                planes = f.read_frame()
                if planes is None:
                    break
                data, pk = gpu_encode(enc, planes, budget, boundary=8192)
                for o, z in pk:
                    out.write_packet(data[o:o + z])
                expected.append(decode_bytes(direct, data, pk))
    finally:
        enc.close()
        direct.close()
    assert len(expected) == 3

    dst = str(tmp_path / "out.y4m")
    with pio.ContainerReader(stream) as r:
        assert (r.width, r.height, r.chroma, r.frame_rate) == (w, h, pw.CHROMA_420, (30, 1))
        dec = pw.Decoder(device, r.width, r.height, r.chroma, precision=pw.PRECISION_FP32)
        try:
            with pio.Y4MFile(dst, "wb", pio.y4m_params(r.width, r.height, r.format, r.full_range,
                                                       r.frame_rate)) as out:
                while r.read_frame(dec):
                    out.write_frame(dec.decode_cpu())
        finally:
            dec.close()
    with pio.Y4MFile(dst) as f:
        assert f.params.startswith("W640 H480 F30:1 Ip A1:1 XCOLORRANGE=FULL C420")
        for exp in expected:
            got = f.read_frame()
            assert all(np.array_equal(a, b) for a, b in zip(got, exp))
        assert f.read_frame() is None


def test_y4m_p10_container_roundtrip_16bit(device, tmp_path):
    """A 10-bit 4:2:0 y4m sequence (C420p10): read rescaled to 16-bit UNORM
    (yuv4mpeg.cpp:194-209), encoded by a 16-bit encoder (encode.cpp:152) into
    a container, read back into a 16-bit decoder (decode.cpp:215) and written
    as y4m.  The bitstreams equal the oracle's encode of the rescaled planes
    and the decoded planes its 16-bit decode."""
    import torch
    from oracle import oracle as O
    from helpers import dev_planes16, target16
    w, h, budget = 320, 192, 30000
    rng = np.random.default_rng(4)
    gen = synth.SceneGenerator(w, h, True, seed=10, frames=2, pan=3, redraw=0.05)
    frames10 = [[(p.astype(np.int64) * 4 + rng.integers(0, 4, p.shape)).astype(np.uint16) for p in gen.frame(t)]
                for t in range(2)]
    src = str(tmp_path / "in10.y4m")
    with open(src, "wb") as f:
        f.write(b"YUV4MPEG2 W%d H%d F30:1 Ip A1:1 XCOLORRANGE=FULL C420p10\n" % (w, h))
        for planes in frames10:
            f.write(b"FRAME\n")
            for p in planes:
                f.write(p.astype("<u2").tobytes())
    L = O.layout(w, h, 0)
    stream = str(tmp_path / "s16.pyrowave")
    enc = pw.Encoder(device, w, h, pw.CHROMA_420, precision=pw.PRECISION_FP16, flags=pw.UNORM16_INPUT)
    expected = []
    try:
        with pio.Y4MFile(src) as f, pio.ContainerWriter(stream, f.width, f.height, f.format, pw.CHROMA_420,
                                                        f.full_range, f.frame_rate) as out:
            assert f.format == pio.YUV420P16
            k = 0
            while True:
                planes = f.read_frame()
                if planes is None:
                    break
                assert planes[0].dtype == np.uint16 and int(planes[0].max()) > 1023
                keep, frame = dev_planes16(planes)
                pw.encode_batch([enc], [frame], [budget])
                torch.cuda.synchronize()
                out_b, pk = enc.packetize(8192, budget + 64)
                data = out_b[:sum(s for _, s in pk)].copy()
                k += 1
                bs, meta, _ = O.encode(L, planes, budget, sequence=k & 7, precision=1)
                ref, opk = O.packetize(L, meta, bs, k & 7, 8192)
                assert np.array_equal(data, ref[:sum(s for _, s in opk)])
                for o, z in pk:
                    out.write_packet(data[o:o + z])
                expected.append(O.decode(L, [data[o:o + z].tobytes() for o, z in pk], precision=1, bits=16))
    finally:
        enc.close()
    assert len(expected) == 2

    dst = str(tmp_path / "out16.y4m")
    with pio.ContainerReader(stream) as r:
        assert r.format == pio.YUV420P16
        dec = pw.Decoder(device, r.width, r.height, r.chroma, precision=pw.PRECISION_FP16, flags=pw.UNORM16_OUTPUT)
        try:
            with pio.Y4MFile(dst, "wb", pio.y4m_params(r.width, r.height, r.format, r.full_range,
                                                       r.frame_rate)) as out:
                while r.read_frame(dec):
                    outs, tgt = target16(w, h, pw.CHROMA_420)
                    pw.decode_batch([dec], [tgt])
                    torch.cuda.synchronize()
                    out.write_frame([o.cpu().numpy().view(np.uint16) for o in outs])
        finally:
            dec.close()
    with pio.Y4MFile(dst) as f:
        assert "C420p16" in f.params
        for exp in expected:
            got = f.read_frame()
            assert all(np.array_equal(a, b) for a, b in zip(got, exp))
        assert f.read_frame() is None
