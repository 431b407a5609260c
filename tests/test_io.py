"""YUV4MPEG2 and PYROWAVE container files (pyrowave_amd/io.py) on the CPU.

- Headers parse as yuv4mpeg.cpp does (W/H via strtol base 0, F num:den,
  colour-space tag, XCOLORRANGE, jpeg siting) and frames round-trip.
- 10-bit 4:2:0 up-converts to 16-bit UNORM as round(v * 65535 / 1023) for
  every 10-bit value (within one step where the exact value is within float
  error of a half), and writes back to the same 10-bit value; 4:4:4
  16-bit passes through unscaled, as in the reference.
- The container round-trips its header and packet records; a frame is read
  by pushing records until the decoder reports it ready.
"""
import numpy as np

from pyrowave_amd import io as pio


def test_y4m_roundtrip_8bit(tmp_path):
    rng = np.random.default_rng(0)
    for fmt, tag in ((pio.YUV420P, "C420jpeg"), (pio.YUV444P, "C444")):
        path = str(tmp_path / f"a{fmt}.y4m")
        params = f"W64 H48 F30000:1001 Ip A1:1 {tag} XCOLORRANGE=FULL\n"
        frames = []
        with pio.Y4MFile(path, "wb", params) as f:
            assert (f.width, f.height, f.format, f.frame_rate, f.full_range) == (64, 48, fmt, (30000, 1001), True)
            for _ in range(3):
                planes = [rng.integers(0, 256, s, dtype=np.uint8) for s in f.plane_shapes()]
                f.write_frame(planes)
                frames.append(planes)
        with pio.Y4MFile(path) as f:
            assert f.params == params and f.center_chroma == (tag == "C420jpeg")
            for planes in frames:
                got = f.read_frame()
                assert all(np.array_equal(a, b) for a, b in zip(got, planes))
            assert f.read_frame() is None


def test_y4m_header_numbers(tmp_path):
    path = str(tmp_path / "h.y4m")
    with open(path, "wb") as f:
        f.write(b"YUV4MPEG2 W0x40 H060 C444 XCOLORRANGE=LIMITED\n")
    with pio.Y4MFile(path) as f:
        # strtol base 0: hex and octal.
        assert (f.width, f.height, f.frame_rate, f.full_range) == (64, 48, (60, 1), False)


def test_y4m_10bit_upconversion(tmp_path):
    path = str(tmp_path / "p10.y4m")
    v = np.arange(1024, dtype=np.uint16)
    w, h = 64, 32  # luma 2048 samples, chroma 512 each
    luma = np.concatenate([v, v[::-1]]).reshape(h, w)
    chroma = v[:512].reshape(h // 2, w // 2)
    with open(path, "wb") as f:
        f.write(b"YUV4MPEG2 W64 H32 C420p10\nFRAME\n")
        for p in (luma, chroma, chroma):
            f.write(p.astype("<u2").tobytes())
    with pio.Y4MFile(path) as f:
        assert f.format == pio.YUV420P16 and f.bits == 10
        y, cb, _ = f.read_frame()
    # Nearest of v * 65535 / 1023, except within float32 error of a half
    # (the reference computes min(1, v / 1023) * 65535 + 0.5 in float).
    got = y.reshape(-1)[:1024].astype(np.int64)
    for x in range(1024):
        q, r = divmod(x * 65535, 1023)
        exact = q + (1 if 2 * r >= 1023 else 0)
        if abs(2 * r - 1023) > 16:  # fraction more than 0.008 away from a half
            assert got[x] == exact, x
        else:
            assert abs(got[x] - exact) <= 1, x
    out = str(tmp_path / "p10_out.y4m")
    with pio.Y4MFile(out, "wb", "W64 H32 C420p10\n") as f:
        f.write_frame([y, cb, cb])
    with open(out, "rb") as f:
        raw = f.read()
    body = np.frombuffer(raw[raw.index(b"FRAME\n") + 6:], "<u2")
    assert np.array_equal(body[:2048], luma.reshape(-1))


def test_y4m_444_16bit_passthrough(tmp_path):
    path = str(tmp_path / "p16.y4m")
    v = (np.arange(16 * 16, dtype=np.uint32) * 997 % 65536).astype(np.uint16).reshape(16, 16)
    with open(path, "wb") as f:
        f.write(b"YUV4MPEG2 W16 H16 C444p12\nFRAME\n")
        for _ in range(3):
            f.write(v.astype("<u2").tobytes())
    with pio.Y4MFile(path) as f:
        assert f.format == pio.YUV444P16
        planes = f.read_frame()
    assert all(np.array_equal(p, v) for p in planes)


class _FakeDecoder:
    def __init__(self, per_frame):
        self.per_frame, self.got = per_frame, []

    def push_packet(self, data):
        self.got.append(bytes(data))
        return 0

    def is_ready(self, allow_partial):
        return len(self.got) % self.per_frame == 0


def test_container_roundtrip(tmp_path):
    path = str(tmp_path / "s.pyrowave")
    packets = [bytes([i]) * (8 + 4 * i) for i in range(6)]
    with pio.ContainerWriter(path, 1920, 1080, pio.YUV420P, 0, True, (60, 1)) as w:
        for p in packets:
            w.write_packet(p)
    raw = open(path, "rb").read()
    assert raw[:8] == b"PYROWAVE"
    assert np.frombuffer(raw[8:40], "<i4").tolist() == [1920, 1080, 0, 0, 1, 60, 1, 0]
    with pio.ContainerReader(path) as r:
        assert (r.width, r.height, r.format, r.chroma, r.full_range, r.frame_rate) == \
            (1920, 1080, pio.YUV420P, 0, True, (60, 1))
        d = _FakeDecoder(per_frame=2)
        assert r.read_frame(d) and r.read_frame(d) and r.read_frame(d)
        assert not r.read_frame(d)
        assert d.got == packets
