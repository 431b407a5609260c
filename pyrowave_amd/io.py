"""YUV4MPEG2 files and the PyroWave stream container.

Restates the reference's file formats (host code; the codec itself runs on the
device through the C ABI):

- ``Y4MFile`` follows yuv4mpeg.cpp:1-283 / yuv4mpeg.hpp: header
  ``YUV4MPEG2 <params>\\n`` with W, H, F num:den, a colour-space tag (C420,
  C444, C420p10/12/14/16, C444p10/12/14/16; anything else reads as 4:2:0
  8-bit), XCOLORRANGE=FULL/LIMITED, jpeg chroma siting; frames are
  ``FRAME\\n`` followed by the planes.  Like the reference, a 4:2:0 high-bit-
  depth file is rescaled to 16-bit UNORM on read
  (min(1, v / (2^bits - 1)) * 65535 + 0.5) and back on write; a 4:4:4 one is
  passed through unscaled (yuv4mpeg.cpp:212-259 rescales only YUV420P16).
- The container follows encode.cpp:144-165 / decode.cpp:121-207: the 8-byte
  magic ``PYROWAVE``, eight little-endian int32 parameters (width, height,
  y4m format, chroma subsampling, full range, frame rate num/den, 0), then one
  ``[u32 size][bytes]`` record per packet; a frame's packets follow each other
  and a frame ends when its blocks are all present.
"""
import struct

import numpy as np

# yuv4mpeg.hpp: enum class Format
YUV420P, YUV444P, YUV420P16, YUV444P16 = 0, 1, 2, 3

_TAGS = [("C420p10", YUV420P16, 10), ("C420p12", YUV420P16, 12), ("C420p14", YUV420P16, 14),
         ("C420p16", YUV420P16, 16), ("C444p10", YUV444P16, 10), ("C444p12", YUV444P16, 12),
         ("C444p14", YUV444P16, 14), ("C444p16", YUV444P16, 16), ("C444", YUV444P, 8)]

MAGIC = b"PYROWAVE"


def format_bytes_per_component(fmt):
    return 2 if fmt in (YUV420P16, YUV444P16) else 1


def format_has_subsampling(fmt):
    return fmt in (YUV420P, YUV420P16)


def format_tag(fmt):
    """The colour-space tag decode.cpp:142-157 writes for a format."""
    return {YUV420P: "C420", YUV420P16: "C420p16", YUV444P: "C444", YUV444P16: "C444p16"}.get(fmt, "???")


def _strtol0(s, i):
    """C strtol(s + i, &end, 0): optional space and sign, then hex (0x),
    octal (leading 0) or decimal digits.  Returns (value, end index)."""
    j = i
    while j < len(s) and s[j] in " \t\n\v\f\r":
        j += 1
    sign = 1
    if j < len(s) and s[j] in "+-":
        sign = -1 if s[j] == "-" else 1
        j += 1
    base, digits = 10, "0123456789"
    if s[j:j + 2] in ("0x", "0X") and j + 2 < len(s) and s[j + 2] in "0123456789abcdefABCDEF":
        base, digits, j = 16, "0123456789abcdefABCDEF", j + 2
    elif s[j:j + 1] == "0":
        base, digits = 8, "01234567"
    k = j
    while k < len(s) and s[k] in digits:
        k += 1
    if k == j:
        return 0, i
    return sign * int(s[j:k], base), k


class Y4MFile:
    """One YUV4MPEG2 file, read or write (yuv4mpeg.cpp)."""

    def __init__(self, path, mode="rb", params=None):
        self.f = open(path, mode)
        self.mode = mode
        self.frame_rate = (60, 1)
        if "r" in mode:
            magic = self.f.read(10)
            if magic != b"YUV4MPEG2 ":
                raise ValueError("invalid YUV4MPEG2 magic")
            line = bytearray()
            while True:
                c = self.f.read(1)
                if not c or c == b"\n":
                    break
                line += c
            self.params = line.decode("ascii", "replace") + "\n"
        else:
            if params is None:
                raise ValueError("params required for writing")
            self.params = params
            self.f.write(b"YUV4MPEG2 " + params.encode("ascii"))
        self._parse()

    def _parse(self):
        p = self.params

        def number_after(ch):
            i = p.find(ch)
            if i < 0:
                return None, None
            return _strtol0(p, i + 1)

        self.width = number_after("W")[0] or 0
        self.height = number_after("H")[0] or 0
        num, end = number_after("F")
        if num is not None:
            den = 1
            if end < len(p) and p[end] == ":":
                den = _strtol0(p, end + 1)[0]
            self.frame_rate = (num, den)
        self.format, self.bits = YUV420P, 8
        for tag, fmt, bits in _TAGS:
            if tag in p:
                self.format, self.bits = fmt, bits
                break
        self.full_range = False
        if "XCOLORRANGE=LIMITED" in p:
            self.full_range = False
        elif "XCOLORRANGE=FULL" in p:
            self.full_range = True
        self.center_chroma = "jpeg" in p or "JPEG" in p
        self.unorm_scale = float((1 << self.bits) - 1) if self.bits < 16 else float(0xFFFF)
        if self.width <= 0 or self.height <= 0:
            raise ValueError("YUV4MPEG2 header without a size")

    def plane_shapes(self):
        h, w = self.height, self.width
        ch, cw = ((h + 1) // 2, (w + 1) // 2) if format_has_subsampling(self.format) else (h, w)
        return [(h, w), (ch, cw), (ch, cw)]

    def read_frame(self):
        """[y, cb, cr] numpy planes (uint8, or uint16 UNORM for p10..p16), or
        None at the end of the file."""
        line = bytearray()
        while True:
            c = self.f.read(1)
            if not c:
                return None
            if c == b"\n":
                break
            line += c
        if bytes(line) != b"FRAME":
            raise ValueError("expected FRAME")
        wide = format_bytes_per_component(self.format) == 2
        planes = []
        for shape in self.plane_shapes():
            n = shape[0] * shape[1]
            raw = self.f.read(n * (2 if wide else 1))
            if len(raw) != n * (2 if wide else 1):
                return None
            if wide:
                v = np.frombuffer(raw, "<u2").astype(np.uint16)
                if self.format == YUV420P16:
                    f = np.minimum(np.float32(1.0), v.astype(np.float32) / np.float32(self.unorm_scale))
                    v = (f * np.float32(0xFFFF) + np.float32(0.5)).astype(np.uint16)
                planes.append(v.reshape(shape))
            else:
                planes.append(np.frombuffer(raw, np.uint8).reshape(shape).copy())
        return planes

    def write_frame(self, planes):
        self.f.write(b"FRAME\n")
        for p in planes:
            a = np.ascontiguousarray(p)
            if format_bytes_per_component(self.format) == 2:
                a = a.astype(np.uint16)
                if self.format == YUV420P16:
                    a = (np.float32(self.unorm_scale) * a.astype(np.float32) / np.float32(0xFFFF) +
                         np.float32(0.5)).astype(np.uint16)
                self.f.write(a.astype("<u2").tobytes())
            else:
                self.f.write(a.astype(np.uint8).tobytes())

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def y4m_params(width, height, fmt, full_range, frame_rate=(60, 1)):
    """Header parameters as decode.cpp:195-196 writes them."""
    return "W%d H%d F%d:%d Ip A1:1 XCOLORRANGE=%s %s\n" % (
        width, height, frame_rate[0], frame_rate[1], "FULL" if full_range else "LIMITED", format_tag(fmt))


class ContainerWriter:
    """PYROWAVE stream file (encode.cpp:144-165, 100-120)."""

    def __init__(self, path, width, height, fmt, chroma, full_range, frame_rate=(60, 1)):
        self.f = open(path, "wb")
        self.f.write(MAGIC)
        self.f.write(struct.pack("<8i", width, height, fmt, chroma, 1 if full_range else 0,
                                 frame_rate[0], frame_rate[1], 0))

    def write_packet(self, data):
        data = bytes(data)
        self.f.write(struct.pack("<I", len(data)))
        self.f.write(data)

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class ContainerReader:
    """Reads a PYROWAVE stream file (decode.cpp:121-207)."""

    def __init__(self, path):
        self.f = open(path, "rb")
        if self.f.read(8) != MAGIC:
            raise ValueError("invalid PYROWAVE magic")
        raw = self.f.read(32)
        if len(raw) != 32:
            raise ValueError("truncated PYROWAVE parameters")
        p = struct.unpack("<8i", raw)
        self.width, self.height, self.format, self.chroma = p[0], p[1], p[2], p[3]
        self.full_range = p[4] != 0
        self.frame_rate = (p[5], p[6])

    def packets(self):
        """Yields each packet's bytes in file order."""
        while True:
            h = self.f.read(4)
            if len(h) != 4:
                return
            (n,) = struct.unpack("<I", h)
            data = self.f.read(n)
            if len(data) != n:
                return
            yield data

    def read_frame(self, decoder):
        """Pushes packets into a pyrowave_amd.Decoder until a frame is ready
        (decode.cpp:121-140); returns False at the end of the file or on a
        packet the decoder rejects."""
        from . import SUCCESS
        for data in self.packets():
            if decoder.push_packet(data) != SUCCESS:
                return False
            if decoder.is_ready(False):
                return True
        return False

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
