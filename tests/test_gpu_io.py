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
