"""C API behaviour on the GPU, case for case after the reference's own test
(reference pyrowave_c_test.cpp): create validation, empty decode, CPU-buffer
validation, small-frame round trips, the 4K stability run and the
performance-stats callback."""
import ctypes as C

import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth
from tests.helpers import ramp_planes

pytestmark = pytest.mark.gpu


def test_encoder_create_validation(device):
    # pyrowave_c_test.cpp:22-59
    with pytest.raises(pw.PyroWaveError) as e:
        pw.Encoder(None, 64, 64)
    assert e.value.code == pw.ERROR_INVALID_ARGUMENT
    enc = pw.Encoder(device, 64, 64)
    for w, h in ((0, 0), (65, 64), (64, 65)):
        with pytest.raises(pw.PyroWaveError) as e:
            pw.Encoder(device, w, h)
        assert e.value.code == pw.ERROR_INVALID_ARGUMENT
    pw.Encoder(device, 65, 64, pw.CHROMA_444).close()
    enc.close()


def test_decoder_create_validation(device):
    # pyrowave_c_test.cpp:61-103
    with pytest.raises(pw.PyroWaveError) as e:
        pw.Decoder(None, 64, 64)
    assert e.value.code == pw.ERROR_INVALID_ARGUMENT
    dec = pw.Decoder(device, 64, 64)
    for w, h in ((0, 0), (65, 64), (64, 65)):
        with pytest.raises(pw.PyroWaveError) as e:
            pw.Decoder(device, w, h)
        assert e.value.code == pw.ERROR_INVALID_ARGUMENT
    pw.Decoder(device, 65, 64, pw.CHROMA_444).close()
    pw.Decoder(device, 65, 64, pw.CHROMA_444, fragment_path=True).close()
    dec.close()


def test_decode_cpu_buffer_validation(device):
    # pyrowave_c_test.cpp:105-162: a decoder with no packets decodes mid-gray.
    dec = pw.Decoder(device, 16, 16)
    assert not dec.is_ready(False) and not dec.is_ready(True)
    luma = np.zeros((16, 16), np.uint8)
    cb = np.zeros((8, 16), np.uint8)  # strided readback
    cr = np.zeros((8, 16), np.uint8)
    buf = pw.CpuBuffer()
    buf.width = buf.height = 16
    for i, (p, stride, size) in enumerate(((luma, 16, 256), (cb, 16, 128), (cr, 16, 128))):
        buf.data[i] = p.ctypes.data
        buf.row_stride_in_bytes[i] = stride
        buf.plane_size_in_bytes[i] = size
    buf.format = pw.FORMAT_NV12
    assert dec.decode_cpu_buffer(buf) == pw.ERROR_INVALID_ARGUMENT
    buf.format = pw.FORMAT_YUV420P
    assert dec.decode_cpu_buffer(buf) == pw.SUCCESS
    assert np.isin(luma, (0x7f, 0x80)).all()
    assert np.isin(cb[:, :8], (0x7f, 0x80)).all()
    assert np.isin(cr[:, :8], (0x7f, 0x80)).all()
    dec.close()


@pytest.mark.parametrize("nv12", [False, True])
def test_encode_cpu_buffer_validation(device, nv12):
    # pyrowave_c_test.cpp:164-218
    enc = pw.Encoder(device, 16, 16)
    y = np.zeros((16, 16), np.uint8)
    cb = np.zeros((8, 8), np.uint8)
    cr = np.zeros((8, 8), np.uint8)
    buf = pw.CpuBuffer()
    buf.format = pw.FORMAT_NV12 if nv12 else pw.FORMAT_YUV420P
    buf.data[0], buf.data[1], buf.data[2] = y.ctypes.data, cb.ctypes.data, cr.ctypes.data
    buf.row_stride_in_bytes[0] = 16
    buf.row_stride_in_bytes[1] = 16 if nv12 else 8
    buf.row_stride_in_bytes[2] = 0 if nv12 else 8
    buf.plane_size_in_bytes[0] = 256
    buf.plane_size_in_bytes[1] = (2 if nv12 else 1) * 64
    buf.plane_size_in_bytes[2] = 0 if nv12 else 64
    # NV12's CbCr plane is 8 x 8 x 2 bytes; back it with enough memory.
    cbcr = np.zeros((8, 16), np.uint8)
    if nv12:
        buf.data[1] = cbcr.ctypes.data
    buf.width = buf.height = 16
    assert enc.encode_cpu_buffer(buf, 1024) == pw.SUCCESS
    buf.width, buf.height = 15, 16
    assert enc.encode_cpu_buffer(buf, 1024) == pw.ERROR_INVALID_ARGUMENT
    buf.width, buf.height = 16, 15
    assert enc.encode_cpu_buffer(buf, 1024) == pw.ERROR_INVALID_ARGUMENT
    buf.width = buf.height = 16
    buf.row_stride_in_bytes[1] = 15 if nv12 else 7
    assert enc.encode_cpu_buffer(buf, 1024) == pw.ERROR_INVALID_ARGUMENT
    buf.row_stride_in_bytes[1] = 16 if nv12 else 8
    buf.plane_size_in_bytes[1] = (2 if nv12 else 1) * 64 - 1
    assert enc.encode_cpu_buffer(buf, 1024) == pw.ERROR_INVALID_ARGUMENT
    # Budgets that are 0 after rounding down to words are rejected
    # (reference pyrowave_c.cpp:957-961).
    buf.plane_size_in_bytes[1] = (2 if nv12 else 1) * 64
    assert enc.encode_cpu_buffer(buf, 3) == pw.ERROR_INVALID_ARGUMENT
    enc.close()


@pytest.mark.parametrize("nv12", [False, True])
@pytest.mark.parametrize("chroma", [pw.CHROMA_420, pw.CHROMA_444])
def test_basic_roundtrip(device, nv12, chroma):
    # pyrowave_c_test.cpp:220-383 (the fragment-path variants are the same
    # decoder here).
    if chroma == pw.CHROMA_444 and nv12:
        pytest.skip("NV12 is 4:2:0 only")
    W, H = 34, 30
    dec = pw.Decoder(device, W, H, chroma)
    enc = pw.Encoder(device, W, H, chroma)
    ys, xs = np.mgrid[0:H, 0:W]
    luma = ((3 * xs + 5 * ys) & 255).astype(np.uint8)
    cb_full = ((7 * xs + 3 * ys) & 255).astype(np.uint8)
    cr_full = ((3 * xs + 5 * ys) & 255).astype(np.uint8)
    if nv12:
        cbcr = np.stack([cb_full, cr_full], axis=-1).reshape(H, 2 * W).copy()
        buf = pw.CpuBuffer()
        buf.format = pw.FORMAT_NV12
        buf.data[0], buf.data[1] = luma.ctypes.data, cbcr.ctypes.data
        buf.row_stride_in_bytes[0], buf.row_stride_in_bytes[1] = W, 2 * W
        buf.plane_size_in_bytes[0], buf.plane_size_in_bytes[1] = luma.nbytes, cbcr.nbytes
        buf.width, buf.height = W, H
        assert enc.encode_cpu_buffer(buf, 64 * 1024) == pw.SUCCESS
    else:
        # Full-size chroma arrays with full-width strides, as the reference test.
        fmt = pw.FORMAT_YUV444P if chroma == pw.CHROMA_444 else pw.FORMAT_YUV420P
        buf = pw.CpuBuffer()
        buf.format = fmt
        for i, p in enumerate((luma, cb_full, cr_full)):
            buf.data[i] = p.ctypes.data
            buf.row_stride_in_bytes[i] = W
            buf.plane_size_in_bytes[i] = p.nbytes
        buf.width, buf.height = W, H
        assert enc.encode_cpu_buffer(buf, 64 * 1024) == pw.SUCCESS
    assert enc.compute_num_packets(64 * 1024) == 1
    out, pk = enc.packetize(64 * 1024, 64 * 1024)
    assert len(pk) == 1 and pk[0][0] == 0 and 0 < pk[0][1] <= 64 * 1024
    data = out[:pk[0][1]].tobytes()
    assert dec.push_packet(data) == pw.SUCCESS
    assert dec.is_ready(False)
    dec.clear()
    assert not dec.is_ready(False)
    assert dec.push_packet(data) == pw.SUCCESS
    assert dec.is_ready(False)

    d_luma = np.zeros((H, W), np.uint8)
    d_cb = np.zeros((H, W), np.uint8)
    d_cr = np.zeros((H, W), np.uint8)
    buf = pw.CpuBuffer()
    buf.format = pw.FORMAT_YUV444P if chroma == pw.CHROMA_444 else pw.FORMAT_YUV420P
    for i, p in enumerate((d_luma, d_cb, d_cr)):
        buf.data[i] = p.ctypes.data
        buf.row_stride_in_bytes[i] = W
        buf.plane_size_in_bytes[i] = p.nbytes
    buf.width, buf.height = W, H
    assert dec.decode_cpu_buffer(buf) == pw.SUCCESS
    assert np.abs(d_luma.astype(int) - luma).max() <= 1
    if chroma == pw.CHROMA_444:
        assert np.abs(d_cb.astype(int) - cb_full).max() <= 1
        assert np.abs(d_cr.astype(int) - cr_full).max() <= 1
    elif nv12:
        assert np.abs(d_cb[:H // 2, :W // 2].astype(int) - cb_full[:H // 2, :W // 2]).max() <= 1
        assert np.abs(d_cr[:H // 2, :W // 2].astype(int) - cr_full[:H // 2, :W // 2]).max() <= 1
    else:
        # The planar 4:2:0 chroma was read as W/2 x H/2 with a W stride.
        src_cb = cb_full[:H // 2, :W // 2]
        src_cr = cr_full[:H // 2, :W // 2]
        assert np.abs(d_cb[:H // 2, :W // 2].astype(int) - src_cb).max() <= 1
        assert np.abs(d_cr[:H // 2, :W // 2].astype(int) - src_cr).max() <= 1
    dec.close()
    enc.close()


def test_basic_system_stability(device):
    # pyrowave_c_test.cpp:385-548, 20 of its 100 iterations.
    W, H = 3840, 2160
    dec = pw.Decoder(device, W, H)
    enc = pw.Encoder(device, W, H)
    y, cb, cr = synth.mirror_pattern(W, H)
    cbcr = np.stack([cb, cr], axis=-1).reshape(H // 2, W).copy()
    buf = pw.CpuBuffer()
    buf.format = pw.FORMAT_NV12
    buf.data[0], buf.data[1] = y.ctypes.data, cbcr.ctypes.data
    buf.row_stride_in_bytes[0] = buf.row_stride_in_bytes[1] = W
    buf.plane_size_in_bytes[0], buf.plane_size_in_bytes[1] = W * H, W * H // 2
    buf.width, buf.height = W, H
    budget = 500000
    for it in range(20):
        assert device.set_queue_type(pw.VK_QUEUE_COMPUTE_BIT if it % 2 else pw.VK_QUEUE_GRAPHICS_BIT) == pw.SUCCESS
        assert enc.encode_cpu_buffer(buf, budget) == pw.SUCCESS
        n = enc.compute_num_packets(8 * 1024)
        assert n > 1
        out, pk = enc.packetize(8 * 1024, budget)
        assert len(pk) == n
        total = 0
        for i, (o, s) in enumerate(pk):
            assert dec.push_packet(out[o:o + s]) == pw.SUCCESS
            assert dec.is_ready(False) == (i == len(pk) - 1)
            total += s
        assert 95 * budget // 100 <= total <= budget
        planes = dec.decode_cpu()
    for src, got in zip((y, cb, cr), planes):
        p = synth.psnr(src, got)
        assert 40.0 < p < 60.0, p
    msgs = device.report_performance_stats(reset=True)
    assert msgs and all("ms per frame" in m for m in msgs)
    dec.close()
    enc.close()


def test_unsupported_vulkan_paths(device):
    L = pw.load()
    assert L.pyrowave_encoder_encode_gpu_synchronous(None, None, None, None, None) == pw.ERROR_NOT_IMPLEMENTED
    assert L.pyrowave_decoder_decode_gpu_buffer(None, None, None, None) == pw.ERROR_NOT_IMPLEMENTED
    assert L.pyrowave_create_device(None, C.byref(C.c_void_p())) == pw.ERROR_NOT_IMPLEMENTED
    assert not L.pyrowave_device_confirm_interop_support(device.handle)
    assert not L.pyrowave_decoder_device_prefers_fragment_path(device.handle)
    assert device.set_queue_type(4) == pw.ERROR_INVALID_ARGUMENT
