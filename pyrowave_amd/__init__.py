"""pyrowave_amd -- Python front end of the MI355X PyroWave codec.

Thin ctypes bindings of libpyrowave-shared.so, the HIP implementation behind
the reference's pyrowave.h C API.  The classes mirror the C API one to one
(Device -> pyrowave_device, Encoder -> pyrowave_encoder, Decoder ->
pyrowave_decoder) with the same argument meaning; a non-success
pyrowave_result raises PyroWaveError carrying the code.

The compute path is the HIP library only: importing this module does not
fall back to anything else, and using it without the built library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyrowave-shared.so")

# pyrowave_result (pyrowave.h)
SUCCESS = 0
TIMEOUT = 1
ERROR_GENERIC = -1
ERROR_INVALID_ARGUMENT = -2
ERROR_OUT_OF_HOST_MEMORY = -3
ERROR_OUT_OF_DEVICE_MEMORY = -4
ERROR_NO_VULKAN = -5
ERROR_NOT_IMPLEMENTED = -6

CHROMA_420 = 0
CHROMA_444 = 1

FORMAT_NV12 = 0
FORMAT_YUV420P = 1
FORMAT_YUV444P = 2

PRECISION_FP16_MATH = 0
PRECISION_FP16 = 1
PRECISION_FP32 = 2
PRECISION_FP16_LL = 3

EXACT_PREFIX = 1
UNORM16_INPUT = 2     # encoder flag: 16-bit (R16_UNORM) input planes
UNORM16_OUTPUT = 1    # decoder flag: 16-bit (R16_UNORM) decode targets

VK_QUEUE_GRAPHICS_BIT = 1
VK_QUEUE_COMPUTE_BIT = 2

STAGES = ("DWT", "Quant", "Analyze", "Resolve", "Packing", "Dequant", "iDWT")


class PyroWaveError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with pyrowave_result {code}")
        self.code = code


class EncoderCreateInfo(C.Structure):
    _fields_ = [("device", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("chroma", C.c_int)]


class DecoderCreateInfo(C.Structure):
    _fields_ = [("device", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("chroma", C.c_int),
                ("fragment_path", C.c_bool)]


class HipEncoderCreateInfo(C.Structure):
    _fields_ = [("device", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("chroma", C.c_int),
                ("precision", C.c_int), ("flags", C.c_uint32)]


class HipDecoderCreateInfo(C.Structure):
    _fields_ = [("device", C.c_void_p), ("width", C.c_int), ("height", C.c_int), ("chroma", C.c_int),
                ("precision", C.c_int), ("flags", C.c_uint32)]


class CpuBuffer(C.Structure):
    _fields_ = [("data", C.c_void_p * 3), ("row_stride_in_bytes", C.c_size_t * 3),
                ("plane_size_in_bytes", C.c_size_t * 3), ("width", C.c_int), ("height", C.c_int),
                ("format", C.c_int)]


class RateControl(C.Structure):
    _fields_ = [("maximum_bitstream_size", C.c_size_t)]


class Packet(C.Structure):
    _fields_ = [("offset", C.c_size_t), ("size", C.c_size_t)]


class HipPlane(C.Structure):
    _fields_ = [("data", C.c_void_p), ("row_stride_in_bytes", C.c_size_t), ("pixel_stride_in_bytes", C.c_size_t)]


class HipFrame(C.Structure):
    _fields_ = [("planes", HipPlane * 3)]


class HipOutputPlane(C.Structure):
    _fields_ = [("data", C.c_void_p), ("row_stride_in_bytes", C.c_size_t)]


class HipDecodeTarget(C.Structure):
    _fields_ = [("planes", HipOutputPlane * 3)]


class HipEncodedFrame(C.Structure):
    _fields_ = [("bitstream", C.c_void_p), ("meta", C.c_void_p), ("num_blocks", C.c_uint32),
                ("sequence", C.c_uint32)]


class PacketStream(C.Structure):
    _fields_ = [("data", C.c_void_p), ("data_bytes", C.c_size_t), ("packets", C.c_void_p), ("count", C.c_void_p),
                ("max_packets", C.c_uint32)]


class ServerCreateInfo(C.Structure):
    _fields_ = [("device", C.c_void_p), ("hbm_budget_bytes", C.c_size_t), ("streams", C.c_uint32),
                ("precision", C.c_int), ("packet_boundary", C.c_size_t), ("max_batch_frames", C.c_uint32)]


class ServerCompletion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("session", C.c_uint32), ("status", C.c_int), ("bytes", C.c_size_t),
                ("packets", C.c_uint32), ("submit_ns", C.c_uint64), ("start_ns", C.c_uint64),
                ("done_ns", C.c_uint64)]


class ExternalMemoryImportInfo(C.Structure):
    _fields_ = [("device", C.c_void_p), ("fd", C.c_int), ("handle_type", C.c_int), ("size", C.c_uint64)]


class ExternalImageLayout(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("chroma", C.c_int), ("layout", C.c_int),
                ("plane_offset", C.c_uint64 * 3), ("row_pitch", C.c_uint64 * 3)]


MESSAGE_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)

_lib = None

# Every symbol pyrowave.h and pyrowave_hip.h declare.
EXPORTS = (
    "pyrowave_get_api_version", "pyrowave_create_default_device", "pyrowave_create_device_by_compat",
    "pyrowave_create_device", "pyrowave_device_set_command_buffer", "pyrowave_device_report_performance_stats",
    "pyrowave_device_get_vk_device_handles", "pyrowave_device_confirm_interop_support",
    "pyrowave_device_set_queue_type", "pyrowave_device_destroy", "pyrowave_sync_object_create",
    "pyrowave_sync_object_get_semaphore", "pyrowave_sync_object_export_handle", "pyrowave_sync_object_cpu_wait",
    "pyrowave_sync_object_cpu_signal", "pyrowave_sync_object_destroy", "pyrowave_image_create",
    "pyrowave_image_get_handle", "pyrowave_image_get_image_view", "pyrowave_image_destroy",
    "pyrowave_encoder_create", "pyrowave_encoder_encode_gpu_synchronous", "pyrowave_encoder_encode_cpu_synchronous",
    "pyrowave_encoder_compute_num_packets", "pyrowave_encoder_packetize", "pyrowave_encoder_destroy",
    "pyrowave_decoder_device_prefers_fragment_path", "pyrowave_decoder_create", "pyrowave_decoder_clear",
    "pyrowave_decoder_push_packet", "pyrowave_decoder_decode_is_ready", "pyrowave_decoder_decode_gpu_buffer",
    "pyrowave_decoder_decode_cpu_buffer_synchronous", "pyrowave_decoder_destroy",
    "pyrowave_hip_encoder_create", "pyrowave_hip_decoder_create", "pyrowave_hip_encode_batch",
    "pyrowave_hip_encoder_get_result", "pyrowave_hip_decode_batch", "pyrowave_hip_decode_encoded_batch",
    "pyrowave_hip_coefficient_count", "pyrowave_hip_encoder_read_coefficients",
    "pyrowave_hip_encoder_encode_coefficients", "pyrowave_hip_encoder_read_rate_control",
    "pyrowave_hip_decoder_read_coefficients", "pyrowave_hip_device_query_stage_times",
    "pyrowave_hip_device_set_timing", "pyrowave_hip_device_set_timing_stages",
    "pyrowave_hip_encoder_packetize_batch", "pyrowave_hip_encoder_get_packets", "pyrowave_hip_decode_packets_batch",
    "pyrowave_hip_decode_packets_noalloc",
    "pyrowave_hip_decoder_parse_result", "pyrowave_hip_encoder_read_packets", "pyrowave_hip_device_set_graphs",
    "pyrowave_hip_psnr_hvs_m_accumulate", "pyrowave_hip_psnr_hvs_m_db", "pyrowave_hip_autotune_batch",
    "pyrowave_hip_device_force_tile_variant",
    "pyrowave_hip_server_create", "pyrowave_hip_server_destroy", "pyrowave_hip_server_session_footprint",
    "pyrowave_hip_server_open_session", "pyrowave_hip_server_close_session", "pyrowave_hip_server_submit",
    "pyrowave_hip_server_poll", "pyrowave_hip_server_take_bitstream",
    "pyrowave_hip_external_memory_import", "pyrowave_hip_external_memory_allocate",
    "pyrowave_hip_external_memory_device_address", "pyrowave_hip_external_memory_size",
    "pyrowave_hip_external_memory_frame", "pyrowave_hip_external_memory_target",
    "pyrowave_hip_external_memory_destroy",
)


def load():
    """Loads the HIP library; raises if it has not been built.

    One process holds one HIP runtime: PyTorch ships its own libamdhip64, and
    a second copy loaded beside it leaves torch without a device.  So torch,
    when installed, is imported first and the library binds to its runtime
    (set PYROWAVE_NO_TORCH=1 to skip that in processes that never use torch).
    """
    global _lib
    if _lib is not None:
        return _lib
    if os.environ.get("PYROWAVE_NO_TORCH") != "1":
        try:
            import torch  # noqa: F401
        except Exception:
            pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pyrowave_amd.build`")
    L = C.CDLL(LIB_PATH)
    P, I, U32, SZ = C.c_void_p, C.c_int, C.c_uint32, C.c_size_t
    sig = {
        "pyrowave_get_api_version": (None, [C.POINTER(U32)] * 3),
        "pyrowave_create_default_device": (I, [C.POINTER(P)]),
        "pyrowave_create_device_by_compat": (I, [U32, U32, P, P, P, C.POINTER(P)]),
        "pyrowave_create_device": (I, [P, C.POINTER(P)]),
        "pyrowave_device_report_performance_stats": (None, [P, MESSAGE_CB, P, C.c_bool]),
        "pyrowave_device_set_queue_type": (I, [P, I]),
        "pyrowave_device_confirm_interop_support": (C.c_bool, [P]),
        "pyrowave_device_destroy": (None, [P]),
        "pyrowave_encoder_create": (I, [C.POINTER(EncoderCreateInfo), C.POINTER(P)]),
        "pyrowave_encoder_encode_gpu_synchronous": (I, [P, P, P, P, P]),
        "pyrowave_encoder_encode_cpu_synchronous": (I, [P, C.POINTER(CpuBuffer), C.POINTER(RateControl)]),
        "pyrowave_encoder_compute_num_packets": (I, [P, SZ, C.POINTER(SZ)]),
        "pyrowave_encoder_packetize": (I, [P, P, SZ, C.POINTER(SZ), P, SZ]),
        "pyrowave_encoder_destroy": (None, [P]),
        "pyrowave_decoder_device_prefers_fragment_path": (C.c_bool, [P]),
        "pyrowave_decoder_create": (I, [C.POINTER(DecoderCreateInfo), C.POINTER(P)]),
        "pyrowave_decoder_clear": (None, [P]),
        "pyrowave_decoder_push_packet": (I, [P, P, SZ]),
        "pyrowave_decoder_decode_is_ready": (C.c_bool, [P, C.c_bool]),
        "pyrowave_decoder_decode_gpu_buffer": (I, [P, P, P, P]),
        "pyrowave_decoder_decode_cpu_buffer_synchronous": (I, [P, C.POINTER(CpuBuffer)]),
        "pyrowave_decoder_destroy": (None, [P]),
        "pyrowave_hip_encoder_create": (I, [C.POINTER(HipEncoderCreateInfo), C.POINTER(P)]),
        "pyrowave_hip_decoder_create": (I, [C.POINTER(HipDecoderCreateInfo), C.POINTER(P)]),
        "pyrowave_hip_encode_batch": (I, [P, P, P, U32, P]),
        "pyrowave_hip_encoder_get_result": (I, [P, C.POINTER(HipEncodedFrame)]),
        "pyrowave_hip_decode_batch": (I, [P, P, U32, P]),
        "pyrowave_hip_decode_encoded_batch": (I, [P, P, P, U32, P]),
        "pyrowave_hip_coefficient_count": (SZ, [I, I, I]),
        "pyrowave_hip_encoder_read_coefficients": (I, [P, P, SZ]),
        "pyrowave_hip_encoder_encode_coefficients": (I, [P, P, SZ, C.POINTER(RateControl)]),
        "pyrowave_hip_encoder_read_rate_control": (I, [P, P, P, P, P, P, P]),
        "pyrowave_hip_decoder_read_coefficients": (I, [P, P, SZ]),
        "pyrowave_hip_device_query_stage_times": (I, [P, P, P, C.c_bool]),
        "pyrowave_hip_device_set_timing": (None, [P, C.c_bool]),
        "pyrowave_hip_device_set_timing_stages": (None, [P, C.c_uint32]),
        "pyrowave_hip_device_set_graphs": (None, [P, C.c_bool]),
        "pyrowave_hip_autotune_batch": (I, [P, P, P, P, P, U32, U32, P, P, P]),
        "pyrowave_hip_device_force_tile_variant": (None, [P, I]),
        "pyrowave_hip_psnr_hvs_m_accumulate": (I, [P, P, SZ, P, SZ, I, I, P, C.POINTER(C.c_uint64), P]),
        "pyrowave_hip_psnr_hvs_m_db": (None, [P, C.c_uint64, C.c_bool, P]),
        "pyrowave_hip_encoder_packetize_batch": (I, [P, U32, SZ, P]),
        "pyrowave_hip_encoder_get_packets": (I, [P, C.POINTER(PacketStream), C.POINTER(P)]),
        "pyrowave_hip_decode_packets_batch": (I, [P, P, P, U32, P]),
        "pyrowave_hip_decode_packets_noalloc": (I, [P, P, P, P, P]),
        "pyrowave_hip_decoder_parse_result": (I, [P, C.POINTER(U32), C.POINTER(U32)]),
        "pyrowave_hip_encoder_read_packets": (I, [P, P, SZ, C.POINTER(Packet), U32, C.POINTER(SZ),
                                                  C.POINTER(U32)]),
        "pyrowave_hip_server_create": (I, [C.POINTER(ServerCreateInfo), C.POINTER(P)]),
        "pyrowave_hip_server_destroy": (None, [P]),
        "pyrowave_hip_server_session_footprint": (I, [P, I, I, I, C.POINTER(SZ)]),
        "pyrowave_hip_server_open_session": (I, [P, I, I, I, C.POINTER(U32)]),
        "pyrowave_hip_server_close_session": (I, [P, U32]),
        "pyrowave_hip_server_submit": (I, [P, U32, C.POINTER(HipFrame), SZ, C.POINTER(HipDecodeTarget),
                                           C.POINTER(C.c_uint64)]),
        "pyrowave_hip_server_poll": (I, [P, C.POINTER(ServerCompletion), U32, C.POINTER(U32), U32]),
        "pyrowave_hip_server_take_bitstream": (I, [P, C.c_uint64, P, SZ, C.POINTER(Packet), U32]),
        "pyrowave_hip_external_memory_import": (I, [C.POINTER(ExternalMemoryImportInfo), C.POINTER(P)]),
        "pyrowave_hip_external_memory_allocate": (I, [P, C.c_uint64, C.POINTER(P), C.POINTER(I)]),
        "pyrowave_hip_external_memory_device_address": (P, [P]),
        "pyrowave_hip_external_memory_size": (C.c_uint64, [P]),
        "pyrowave_hip_external_memory_frame": (I, [P, C.POINTER(ExternalImageLayout), C.POINTER(HipFrame)]),
        "pyrowave_hip_external_memory_target": (I, [P, C.POINTER(ExternalImageLayout),
                                                    C.POINTER(HipDecodeTarget)]),
        "pyrowave_hip_external_memory_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(code, what):
    if code != SUCCESS:
        raise PyroWaveError(code, what)


def api_version():
    a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    load().pyrowave_get_api_version(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def coefficient_count(width, height, chroma):
    return int(load().pyrowave_hip_coefficient_count(width, height, chroma))


def plane_shape(width, height, chroma, c):
    if c and chroma == CHROMA_420:
        return height // 2, width // 2
    return height, width


class Device:
    def __init__(self):
        h = C.c_void_p()
        _check(load().pyrowave_create_default_device(C.byref(h)), "pyrowave_create_default_device")
        self.handle = h

    def close(self):
        if self.handle:
            load().pyrowave_device_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_queue_type(self, flags):
        return load().pyrowave_device_set_queue_type(self.handle, flags)

    def report_performance_stats(self, reset=False):
        msgs = []
        cb = MESSAGE_CB(lambda ud, m: msgs.append(m.decode()))
        load().pyrowave_device_report_performance_stats(self.handle, cb, None, reset)
        return msgs

    def stage_times(self, reset=False):
        ms = (C.c_double * 7)()
        fr = (C.c_uint64 * 7)()
        _check(load().pyrowave_hip_device_query_stage_times(self.handle, ms, fr, reset), "query_stage_times")
        return {STAGES[i]: (ms[i], fr[i]) for i in range(7)}

    def set_timing(self, enable):
        load().pyrowave_hip_device_set_timing(self.handle, enable)

    def set_graphs(self, enable):
        """Batch entry points replay a captured hipGraph per batch composition."""
        load().pyrowave_hip_device_set_graphs(self.handle, enable)

    def force_tile_variant(self, variant):
        """Every transform level uses tile shape `variant` (0-4 in both
        directions); -1 = autotuned."""
        load().pyrowave_hip_device_force_tile_variant(self.handle, variant)

    def set_timing_stages(self, stages):
        """Times only the named stages (names as in STAGES)."""
        mask = 0
        for s in stages:
            mask |= 1 << STAGES.index(s)
        load().pyrowave_hip_device_set_timing_stages(self.handle, mask)


def _cpu_buffer(planes, fmt, width, height, strides=None, sizes=None):
    buf = CpuBuffer()
    buf.width, buf.height, buf.format = width, height, fmt
    for i, p in enumerate(planes):
        if p is None:
            continue
        buf.data[i] = p.ctypes.data
        buf.row_stride_in_bytes[i] = p.strides[0] if strides is None else strides[i]
        buf.plane_size_in_bytes[i] = p.nbytes if sizes is None else sizes[i]
    return buf


class Encoder:
    def __init__(self, device, width, height, chroma=CHROMA_420, precision=None, flags=0):
        self.device, self.width, self.height, self.chroma = device, width, height, chroma
        h = C.c_void_p()
        if precision is None:
            info = EncoderCreateInfo(device.handle if device else None, width, height, chroma)
            code = load().pyrowave_encoder_create(C.byref(info), C.byref(h))
        else:
            info = HipEncoderCreateInfo(device.handle if device else None, width, height, chroma, precision, flags)
            code = load().pyrowave_hip_encoder_create(C.byref(info), C.byref(h))
        _check(code, "pyrowave_encoder_create")
        self.handle = h

    def close(self):
        if self.handle:
            load().pyrowave_encoder_destroy(self.handle)
            self.handle = None

    def encode_cpu(self, planes, budget, fmt=None, width=None, height=None):
        """planes: [y, cb, cr] (or [y, cbcr] for NV12) uint8 numpy arrays."""
        if fmt is None:
            fmt = FORMAT_YUV444P if self.chroma == CHROMA_444 else FORMAT_YUV420P
        planes = list(planes) + [None] * (3 - len(planes))
        buf = _cpu_buffer(planes, fmt, width or self.width, height or self.height)
        rc = RateControl(budget)
        return load().pyrowave_encoder_encode_cpu_synchronous(self.handle, C.byref(buf), C.byref(rc))

    def encode_cpu_buffer(self, buf, budget):
        rc = RateControl(budget)
        return load().pyrowave_encoder_encode_cpu_synchronous(self.handle, C.byref(buf), C.byref(rc))

    def compute_num_packets(self, boundary):
        n = C.c_size_t()
        _check(load().pyrowave_encoder_compute_num_packets(self.handle, boundary, C.byref(n)), "compute_num_packets")
        return n.value

    def packetize(self, boundary, size):
        n = self.compute_num_packets(boundary)
        packets = (Packet * max(n, 1))()
        out = np.zeros(size, dtype=np.uint8)
        got = C.c_size_t()
        _check(load().pyrowave_encoder_packetize(self.handle, packets, boundary, C.byref(got), out.ctypes.data, size),
               "packetize")
        return out, [(packets[i].offset, packets[i].size) for i in range(got.value)]

    def packet_bytes(self, boundary, size):
        out, pk = self.packetize(boundary, size)
        return [out[o:o + s].tobytes() for o, s in pk]

    def result(self):
        r = HipEncodedFrame()
        _check(load().pyrowave_hip_encoder_get_result(self.handle, C.byref(r)), "get_result")
        return r

    def read_coefficients(self):
        n = coefficient_count(self.width, self.height, self.chroma)
        out = np.zeros(n, dtype=np.float32)
        _check(load().pyrowave_hip_encoder_read_coefficients(self.handle, out.ctypes.data, n), "read_coefficients")
        return out

    def encode_coefficients(self, coeffs, budget):
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float32)
        rc = RateControl(budget)
        _check(load().pyrowave_hip_encoder_encode_coefficients(self.handle, coeffs.ctypes.data, coeffs.size,
                                                               C.byref(rc)), "encode_coefficients")

    def read_rate_control(self, count32):
        words = np.zeros((count32, 16), np.uint16)
        bucket = np.zeros((count32, 16), np.uint8)
        totals = np.zeros(2048, np.uint32)
        prefix = np.zeros(2048, np.uint32)
        consumed = np.zeros(1, np.uint32)
        quant = np.zeros(count32, np.int32)
        _check(load().pyrowave_hip_encoder_read_rate_control(
            self.handle, words.ctypes.data, bucket.ctypes.data, totals.ctypes.data, prefix.ctypes.data,
            consumed.ctypes.data, quant.ctypes.data), "read_rate_control")
        return dict(words=words, bucket=bucket, totals=totals, prefix=prefix, consumed=int(consumed[0]), quant=quant)


class Decoder:
    def __init__(self, device, width, height, chroma=CHROMA_420, precision=None, fragment_path=False, flags=0):
        self.device, self.width, self.height, self.chroma = device, width, height, chroma
        h = C.c_void_p()
        if precision is None:
            info = DecoderCreateInfo(device.handle if device else None, width, height, chroma, fragment_path)
            code = load().pyrowave_decoder_create(C.byref(info), C.byref(h))
        else:
            info = HipDecoderCreateInfo(device.handle if device else None, width, height, chroma, precision, flags)
            code = load().pyrowave_hip_decoder_create(C.byref(info), C.byref(h))
        _check(code, "pyrowave_decoder_create")
        self.handle = h

    def close(self):
        if self.handle:
            load().pyrowave_decoder_destroy(self.handle)
            self.handle = None

    def clear(self):
        load().pyrowave_decoder_clear(self.handle)

    def push_packet(self, data):
        data = bytes(data)
        return load().pyrowave_decoder_push_packet(self.handle, data, len(data))

    def is_ready(self, allow_partial=False):
        return bool(load().pyrowave_decoder_decode_is_ready(self.handle, allow_partial))

    def decode_cpu(self, fmt=None):
        if fmt is None:
            fmt = FORMAT_YUV444P if self.chroma == CHROMA_444 else FORMAT_YUV420P
        planes = [np.zeros(plane_shape(self.width, self.height, self.chroma, c), np.uint8) for c in range(3)]
        buf = _cpu_buffer(planes, fmt, self.width, self.height)
        _check(load().pyrowave_decoder_decode_cpu_buffer_synchronous(self.handle, C.byref(buf)), "decode_cpu")
        return planes

    def decode_cpu_buffer(self, buf):
        return load().pyrowave_decoder_decode_cpu_buffer_synchronous(self.handle, C.byref(buf))

    def read_coefficients(self):
        n = coefficient_count(self.width, self.height, self.chroma)
        out = np.zeros(n, dtype=np.float32)
        _check(load().pyrowave_hip_decoder_read_coefficients(self.handle, out.ctypes.data, n), "read_coefficients")
        return out


def _handles(objs):
    arr = (C.c_void_p * len(objs))()
    for i, o in enumerate(objs):
        arr[i] = o.handle.value if isinstance(o.handle, C.c_void_p) else o.handle
    return arr


def hip_frame(planes_ptrs):
    """planes_ptrs: [(device_ptr, row_stride, pixel_stride)] * 3."""
    f = HipFrame()
    for i, (ptr, stride, step) in enumerate(planes_ptrs):
        f.planes[i] = HipPlane(ptr, stride, step)
    return f


def hip_target(planes_ptrs):
    t = HipDecodeTarget()
    for i, (ptr, stride) in enumerate(planes_ptrs):
        t.planes[i] = HipOutputPlane(ptr, stride)
    return t


def encode_batch(encoders, frames, budgets, stream=None):
    n = len(encoders)
    fa = (HipFrame * n)(*frames)
    ra = (RateControl * n)(*[RateControl(b) for b in budgets])
    _check(load().pyrowave_hip_encode_batch(_handles(encoders), fa, ra, n, stream), "encode_batch")


def decode_batch(decoders, targets, stream=None):
    n = len(decoders)
    ta = (HipDecodeTarget * n)(*targets)
    _check(load().pyrowave_hip_decode_batch(_handles(decoders), ta, n, stream), "decode_batch")


def autotune_batch(encoders, frames, budgets, decoders=None, targets=None, repetitions=3, stream=None):
    """Encodes (and decodes) the batch while timing every transform tile shape
    per level; returns (forward variants, inverse variants) per level."""
    n = len(encoders)
    fa = (HipFrame * n)(*frames)
    ra = (RateControl * n)(*[RateControl(b) for b in budgets])
    ta = (HipDecodeTarget * n)(*targets) if targets is not None else None
    fv, iv = (C.c_int * 5)(), (C.c_int * 5)()
    _check(load().pyrowave_hip_autotune_batch(_handles(encoders), _handles(decoders) if decoders else None, fa, ra,
                                              ta, n, repetitions, stream, fv, iv), "autotune_batch")
    return list(fv), list(iv)


def decode_encoded_batch(decoders, encoders, targets, stream=None):
    n = len(decoders)
    ta = (HipDecodeTarget * n)(*targets)
    _check(load().pyrowave_hip_decode_encoded_batch(_handles(decoders), _handles(encoders), ta, n, stream),
           "decode_encoded_batch")


def packetize_batch(encoders, packet_boundary, stream=None):
    """Device packetize after the encoders' last encode on `stream`."""
    _check(load().pyrowave_hip_encoder_packetize_batch(_handles(encoders), len(encoders), packet_boundary, stream),
           "packetize_batch")


def device_packets(encoder):
    """(PacketStream, device address of the stream-length word) of the
    encoder's last device packetize."""
    ps = PacketStream()
    nbytes = C.c_void_p()
    _check(load().pyrowave_hip_encoder_get_packets(encoder.handle, C.byref(ps), C.byref(nbytes)), "get_packets")
    return ps, nbytes.value


def read_packets(encoder):
    """(stream bytes as uint8 array, [(offset, size)]) of the encoder's last
    device packetize, copied to the host."""
    import numpy as np
    nb, npk = C.c_size_t(), C.c_uint32()
    L = load()
    _check(L.pyrowave_hip_encoder_read_packets(encoder.handle, None, 0, None, 0, C.byref(nb), C.byref(npk)),
           "read_packets")
    out = np.empty(max(1, nb.value), np.uint8)
    pk = (Packet * max(1, npk.value))()
    _check(L.pyrowave_hip_encoder_read_packets(encoder.handle, out.ctypes.data, out.size, pk, npk.value,
                                               C.byref(nb), C.byref(npk)), "read_packets")
    return out[:nb.value], [(pk[i].offset, pk[i].size) for i in range(npk.value)]


def packet_stream(data_ptr, data_bytes, packets_ptr, count_ptr, max_packets):
    return PacketStream(data_ptr, data_bytes, packets_ptr, count_ptr, max_packets)


def decode_packets_batch(decoders, streams, targets, stream=None):
    n = len(decoders)
    pa = (PacketStream * n)(*streams)
    ta = (HipDecodeTarget * n)(*targets)
    _check(load().pyrowave_hip_decode_packets_batch(_handles(decoders), pa, ta, n, stream), "decode_packets_batch")


def decode_packets_noalloc(decoder, packet_stream_desc, target, device_info=None, stream=None):
    """pyrowave_hip_decode_packets_noalloc: returns the result code."""
    return load().pyrowave_hip_decode_packets_noalloc(decoder.handle, C.byref(packet_stream_desc), C.byref(target),
                                                       device_info, stream)


def parse_result(decoder):
    """(result code, blocks present, blocks announced) of the last device parse."""
    a, b = C.c_uint32(), C.c_uint32()
    r = load().pyrowave_hip_decoder_parse_result(decoder.handle, C.byref(a), C.byref(b))
    return r, a.value, b.value


def psnr_hvs_m_accumulate(device, ref_ptr, ref_stride, dist_ptr, dist_stride, width, height, totals_ptr,
                          stream=None):
    """Queues one frame's PSNR-HVS-M-H errors into 16 device uint64 totals;
    returns the frame's pixel count."""
    tp = C.c_uint64(0)
    _check(load().pyrowave_hip_psnr_hvs_m_accumulate(device.handle, ref_ptr, ref_stride, dist_ptr, dist_stride,
                                                     width, height, totals_ptr, C.byref(tp), stream),
           "psnr_hvs_m_accumulate")
    return tp.value


def psnr_hvs_m_db(totals, total_pixels, full_range=True):
    """dB per height factor from 16 host uint64 totals."""
    import numpy as np
    t = np.ascontiguousarray(totals, np.uint64)
    out = np.zeros(16, np.float64)
    load().pyrowave_hip_psnr_hvs_m_db(t.ctypes.data, total_pixels, full_range, out.ctypes.data)
    return out


EXTERNAL_HANDLE_FD = 1
PLANES_THREE, PLANES_INTERLEAVED_CHROMA = 0, 1


def external_layout(width, height, chroma, offsets, pitches, layout=PLANES_THREE):
    L = ExternalImageLayout()
    L.width, L.height, L.chroma, L.layout = width, height, chroma, layout
    for i in range(3):
        L.plane_offset[i] = offsets[i] if i < len(offsets) else 0
        L.row_pitch[i] = pitches[i] if i < len(pitches) else 0
    return L


class ExternalMemory:
    """Device memory shared with other APIs or processes through a POSIX fd
    (pyrowave_hip_external_memory_*): `allocate` makes exportable memory and
    returns its fd, `import_fd` maps an fd exported elsewhere."""

    def __init__(self, handle, fd=-1):
        self.handle = handle
        self.fd = fd  # exported fd (allocate only); the caller closes it

    @classmethod
    def allocate(cls, device, size):
        h, fd = C.c_void_p(), C.c_int(-1)
        _check(load().pyrowave_hip_external_memory_allocate(device.handle, size, C.byref(h), C.byref(fd)),
               "external_memory_allocate")
        return cls(h, fd.value)

    @classmethod
    def import_fd(cls, device, fd, size):
        info = ExternalMemoryImportInfo(device.handle, fd, EXTERNAL_HANDLE_FD, size)
        h = C.c_void_p()
        _check(load().pyrowave_hip_external_memory_import(C.byref(info), C.byref(h)), "external_memory_import")
        return cls(h)

    @property
    def address(self):
        return load().pyrowave_hip_external_memory_device_address(self.handle)

    @property
    def size(self):
        return load().pyrowave_hip_external_memory_size(self.handle)

    def frame(self, layout):
        f = HipFrame()
        _check(load().pyrowave_hip_external_memory_frame(self.handle, C.byref(layout), C.byref(f)),
               "external_memory_frame")
        return f

    def target(self, layout):
        t = HipDecodeTarget()
        _check(load().pyrowave_hip_external_memory_target(self.handle, C.byref(layout), C.byref(t)),
               "external_memory_target")
        return t

    def close(self):
        if self.handle:
            load().pyrowave_hip_external_memory_destroy(self.handle)
            self.handle = None


class Server:
    """Serving loop: sessions admitted by device footprint, frames batched
    continuously over several streams (pyrowave_hip_server_* entry points)."""

    def __init__(self, device, hbm_budget_bytes, streams=4, precision=PRECISION_FP16, packet_boundary=0,
                 max_batch_frames=16):
        info = ServerCreateInfo(device.handle, hbm_budget_bytes, streams, precision, packet_boundary,
                                max_batch_frames)
        h = C.c_void_p()
        _check(load().pyrowave_hip_server_create(C.byref(info), C.byref(h)), "pyrowave_hip_server_create")
        self.handle = h

    def close(self):
        if self.handle:
            load().pyrowave_hip_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def footprint(self, width, height, chroma):
        n = C.c_size_t()
        _check(load().pyrowave_hip_server_session_footprint(self.handle, width, height, chroma, C.byref(n)),
               "session_footprint")
        return n.value

    def open_session(self, width, height, chroma):
        """Returns the session id, or None when the footprint does not fit."""
        sid = C.c_uint32()
        r = load().pyrowave_hip_server_open_session(self.handle, width, height, chroma, C.byref(sid))
        if r == ERROR_OUT_OF_DEVICE_MEMORY:
            return None
        _check(r, "open_session")
        return sid.value

    def close_session(self, session):
        _check(load().pyrowave_hip_server_close_session(self.handle, session), "close_session")

    def submit(self, session, frame, budget, target=None):
        t = C.c_uint64()
        _check(load().pyrowave_hip_server_submit(self.handle, session, C.byref(frame), budget,
                                                 C.byref(target) if target is not None else None, C.byref(t)),
               "submit")
        return t.value

    def poll(self, max_count=64, timeout_us=0):
        arr = (ServerCompletion * max_count)()
        n = C.c_uint32()
        _check(load().pyrowave_hip_server_poll(self.handle, arr, max_count, C.byref(n), timeout_us), "poll")
        return [arr[i] for i in range(n.value)]

    def take_bitstream(self, completion):
        """(bytes as uint8 array, [(offset, size)] packets) of a finished frame."""
        import numpy as np
        out = np.empty(max(1, completion.bytes), np.uint8)
        pk = (Packet * max(1, completion.packets))()
        _check(load().pyrowave_hip_server_take_bitstream(self.handle, completion.ticket, out.ctypes.data,
                                                         out.size, pk, completion.packets), "take_bitstream")
        return out[:completion.bytes], [(pk[i].offset, pk[i].size) for i in range(completion.packets)]
