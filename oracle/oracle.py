"""ctypes/numpy front end of the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.  The product library never loads this module.
"""
import contextlib
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_bitstream.so")

LEVELS, COMPONENTS, BANDS = 5, 3, 4
FLAT = 2048


class Band(C.Structure):
    _fields_ = [("present", C.c_int), ("bw", C.c_int), ("bh", C.c_int),
                ("bx8", C.c_int), ("by8", C.c_int), ("bx32", C.c_int), ("by32", C.c_int),
                ("off8", C.c_int), ("off32", C.c_int), ("coeff_off", C.c_long)]


class Layout(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("chroma", C.c_int),
                ("aligned_w", C.c_int), ("aligned_h", C.c_int),
                ("count8", C.c_int), ("count32", C.c_int), ("coeff_count", C.c_long),
                ("per_subdivision", C.c_int),
                ("band", Band * BANDS * LEVELS * COMPONENTS),
                ("precision0", C.c_int)]

    def b(self, c, l, b):
        return self.band[c][l][b]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _lib = load_library(LIB_PATH)
    return _lib


@contextlib.contextmanager
def use_library(path):
    """Runs the block with the oracle functions bound to another build of the
    oracle sources (the mutation tests build planted mistakes this way)."""
    global _lib
    saved = _lib
    _lib = load_library(path)
    try:
        yield _lib
    finally:
        _lib = saved


def load_library(path):
    L = C.CDLL(path)
    P = C.c_void_p
    L.or_layout_init.argtypes = [C.POINTER(Layout), C.c_int, C.c_int, C.c_int]
    L.or_layout_size.restype = C.c_long
    L.or_block_descs.argtypes = [C.POINTER(Layout), P]
    L.or_quant_resolution.restype = C.c_float
    L.or_quant_resolution.argtypes = [C.c_int, C.c_int, C.c_int]
    L.or_quant_resolution_policy.restype = C.c_float
    L.or_quant_resolution_policy.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.or_round_f64_to_f16.restype = C.c_float
    L.or_round_f64_to_f16.argtypes = [C.c_double]
    L.or_rdo_distortion_scale.restype = C.c_float
    L.or_rdo_distortion_scale.argtypes = [C.POINTER(Layout), C.c_int, C.c_int, C.c_int]
    L.or_encode_quant.restype = C.c_uint8
    L.or_encode_quant.argtypes = [C.c_float]
    L.or_decode_quant.restype = C.c_float
    L.or_decode_quant.argtypes = [C.c_uint8]
    L.or_f32_to_f16.restype = C.c_uint16
    L.or_f32_to_f16.argtypes = [C.c_float]
    L.or_f16_to_f32.restype = C.c_float
    L.or_f16_to_f32.argtypes = [C.c_uint16]
    L.or_round_to_f16_array.argtypes = [P, C.c_long]
    L.or_round_f16_fast.restype = C.c_float
    L.or_round_f16_fast.argtypes = [C.c_float]
    L.or_check_round_f16_fast.restype = C.c_long
    L.or_fwd_ext.argtypes = [C.c_int, C.c_int, C.c_int]
    for name in ("or_dwt_level_f64", "or_dwt_level_f32"):
        getattr(L, name).argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, C.c_int]
    L.or_dwt_level_h.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, C.c_int]
    L.or_idwt_level_h.argtypes = [P, P, P, P, C.c_int, C.c_int, C.c_int, P, C.c_int]
    for name in ("or_idwt_level_f64", "or_idwt_level_f32"):
        getattr(L, name).argtypes = [P, P, P, P, C.c_int, C.c_int, C.c_int, P, C.c_int]
    L.or_dwt_frame.argtypes = [C.POINTER(Layout), P, P, C.c_int, P, P]
    L.or_unorm_dc.restype = C.c_float
    L.or_unorm_dc.argtypes = [C.c_uint32, C.c_int, C.c_int]
    L.or_dwt_frame_ex.argtypes = [C.POINTER(Layout), P, P, C.c_int, C.c_int, P, P]
    L.or_encode_ex.restype = C.c_long
    L.or_encode_ex.argtypes = [C.POINTER(Layout), P, P, C.c_int, C.c_int, C.c_size_t, C.c_int, C.c_int,
                               P, C.c_long, P, P]
    L.or_idwt_frame_ex.argtypes = [C.POINTER(Layout), P, C.c_int, P, P, C.c_int]
    L.or_quant_analyze.argtypes = [C.POINTER(Layout), P, P, P, P, P, P, P]
    L.or_quant_analyze_ex.argtypes = [C.POINTER(Layout), P, P, P, P, P, P, P, P]
    L.or_block8_stats.argtypes = [C.POINTER(Layout), P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  P, P, P, P]
    L.or_bucket_of_ratio.restype = C.c_uint32
    L.or_bucket_of_ratio.argtypes = [C.c_float]
    L.or_bucket_scan.argtypes = [P, P]
    L.or_finalize.argtypes = [P, P, C.c_int]
    L.or_resolve.argtypes = [C.POINTER(Layout), P, P, P, C.c_uint32, C.c_size_t, P]
    L.or_pack.restype = C.c_long
    L.or_pack.argtypes = [C.POINTER(Layout), P, P, C.c_int, P, C.c_long, P]
    L.or_encode.restype = C.c_long
    L.or_encode.argtypes = [C.POINTER(Layout), P, P, C.c_int, C.c_size_t, C.c_int, C.c_int,
                            P, C.c_long, P, P]
    L.or_compute_num_packets.restype = C.c_size_t
    L.or_compute_num_packets.argtypes = [C.POINTER(Layout), P, C.c_size_t]
    L.or_packetize.restype = C.c_size_t
    L.or_packetize.argtypes = [C.POINTER(Layout), P, P, C.c_int, C.c_size_t, P, C.c_size_t, P]
    L.or_validate_block.argtypes = [C.POINTER(Layout), P, C.c_uint32, C.c_uint32]
    L.or_decoder_create.restype = P
    L.or_decoder_create.argtypes = [C.POINTER(Layout)]
    L.or_decoder_destroy.argtypes = [P]
    L.or_decoder_clear.argtypes = [P]
    L.or_decoder_push_packet.argtypes = [P, P, C.c_size_t]
    L.or_decoder_is_ready.argtypes = [P, C.c_int]
    L.or_decoder_mark_decoded.argtypes = [P]
    L.or_decoder_payload_words.restype = C.c_long
    L.or_decoder_payload_words.argtypes = [P]
    L.or_decoder_payload.restype = P
    L.or_decoder_payload.argtypes = [P]
    L.or_decoder_offsets.restype = P
    L.or_decoder_offsets.argtypes = [P]
    L.or_decoder_decoded_blocks.argtypes = [P]
    L.or_decoder_total_blocks.argtypes = [P]
    L.or_dequant.argtypes = [C.POINTER(Layout), P, C.c_long, P, P]
    L.or_idwt_frame.argtypes = [C.POINTER(Layout), P, C.c_int, P, P]
    L.or_idwt_frame_f64.argtypes = [C.POINTER(Layout), P, P, P]
    L.or_dct8x8.argtypes = [P, P]
    L.or_maskeff.restype = C.c_float
    L.or_maskeff.argtypes = [P, P, P]
    L.or_psnr_hvs_m_tables.argtypes = [C.c_int, C.c_float, P, P, P]
    L.or_psnr_height_factor.restype = C.c_float
    L.or_psnr_height_factor.argtypes = [C.c_int]
    L.or_psnr_hvs_m_totals.restype = C.c_uint64
    L.or_psnr_hvs_m_totals.argtypes = [P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_int, P]
    L.or_psnr_hvs_m_db.restype = C.c_double
    L.or_psnr_hvs_m_db.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
    return L


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def layout(width, height, chroma, precision=None):
    """precision 0 selects the precision-0 quantizer policy for the functions
    that take the layout (quant_analyze, pack)."""
    assert lib().or_layout_size() == C.sizeof(Layout)
    L = Layout()
    if lib().or_layout_init(C.byref(L), width, height, chroma) != 0:
        raise ValueError("invalid dimensions")
    L.precision0 = 1 if precision == 0 else 0
    return L


def plane_shape(L, c):
    if c != 0 and L.chroma == 0:
        return L.height // 2, L.width // 2
    return L.height, L.width


def _plane_ptrs(planes):
    """uint16 planes are R16_UNORM samples, anything else UNORM8; returns
    (planes, pointers, pitches in samples, bits)."""
    bits = 16 if all(np.asarray(p).dtype == np.uint16 for p in planes) else 8
    dt = np.uint16 if bits == 16 else np.uint8
    planes = [np.ascontiguousarray(p, dtype=dt) for p in planes]
    arr = (C.c_void_p * 3)(*[p.ctypes.data for p in planes])
    pitches = (C.c_int * 3)(*[p.shape[1] for p in planes])
    return planes, arr, pitches, bits


BLOCK_DESC = np.dtype([("component", np.int16), ("level", np.int16), ("band", np.int16), ("pad", np.int16),
                       ("bx", np.int32), ("by", np.int32)])


def block_descs(L):
    raw = np.zeros(L.count32, dtype=BLOCK_DESC)
    lib().or_block_descs(C.byref(L), ptr(raw))
    return raw


def band_view(L, coeffs, c, l, b):
    B = L.band[c][l][b]
    return coeffs[B.coeff_off:B.coeff_off + B.bw * B.bh].reshape(B.bh, B.bw)


def unorm_dc(u, bits, precision=2):
    """The oracle's level-0 input value of sample u (float)."""
    return float(lib().or_unorm_dc(int(u), bits, precision))


def dwt_frame(L, planes, precision=2, f64=False):
    """Forward transform of u8 (UNORM8) or uint16 (R16_UNORM) planes."""
    planes, arr, pitches, bits = _plane_ptrs(planes)
    if f64:
        c64 = np.zeros(L.coeff_count, dtype=np.float64)
        lib().or_dwt_frame_ex(C.byref(L), arr, pitches, bits, precision, None, ptr(c64))
        return c64
    c32 = np.zeros(L.coeff_count, dtype=np.float32)
    lib().or_dwt_frame_ex(C.byref(L), arr, pitches, bits, precision, ptr(c32), None)
    return c32


def quant_analyze(L, coeffs, with_dist=False):
    """Quantizer + analyze; with_dist also returns dist32[n32][16], each 32x32
    block's summed distortion per quant level."""
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float32)
    words = np.zeros((L.count32, 16), dtype=np.uint16)
    bucket = np.zeros((L.count32, 16), dtype=np.uint8)
    code8 = np.zeros(L.count8, dtype=np.uint32)
    np8 = np.zeros(L.count8, dtype=np.uint8)
    totals = np.zeros(FLAT, dtype=np.uint32)
    consumed = np.zeros(1, dtype=np.uint32)
    dist32 = np.zeros((L.count32, 16), dtype=np.float32) if with_dist else None
    lib().or_quant_analyze_ex(C.byref(L), ptr(coeffs), ptr(words), ptr(bucket), ptr(code8), ptr(np8),
                              ptr(totals), ptr(consumed), ptr(dist32))
    r = dict(words=words, bucket=bucket, code8=code8, np8=np8, totals=totals, consumed=int(consumed[0]))
    if with_dist:
        r["dist32"] = dist32
    return r


def block8_stats(L, coeffs, c, l, b, x8, y8):
    """BlockStats of one 8x8 block: (cost[16] bits, dist[16], code word, msb)."""
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float32)
    cost = np.zeros(16, np.uint16)
    dist = np.zeros(16, np.float32)
    cw = C.c_uint32()
    msb = C.c_int()
    if lib().or_block8_stats(C.byref(L), ptr(coeffs), c, l, b, x8, y8, ptr(cost), ptr(dist),
                             C.byref(cw), C.byref(msb)) != 0:
        raise ValueError("block outside the band")
    return cost, dist, int(cw.value), int(msb.value)


def bucket_of_ratio(r):
    return int(lib().or_bucket_of_ratio(float(r)))


def bucket_scan(raw):
    raw = np.ascontiguousarray(raw, np.uint32)
    out = np.zeros(16, np.uint32)
    lib().or_bucket_scan(ptr(raw), ptr(out))
    return [int(x) for x in out]


def finalize(totals, q1=True):
    totals = np.ascontiguousarray(totals, dtype=np.uint32)
    prefix = np.zeros(FLAT, dtype=np.uint32)
    lib().or_finalize(ptr(totals), ptr(prefix), 1 if q1 else 0)
    return prefix


def resolve(L, qa, prefix, budget):
    quant = np.zeros(L.count32, dtype=np.int32)
    lib().or_resolve(C.byref(L), ptr(qa["words"]), ptr(qa["bucket"]), ptr(np.ascontiguousarray(prefix, np.uint32)),
                     qa["consumed"], budget, ptr(quant))
    return quant


def pack(L, coeffs, quant, sequence, capacity_words=None):
    coeffs = np.ascontiguousarray(coeffs, dtype=np.float32)
    quant = np.ascontiguousarray(quant, dtype=np.int32)
    cap = capacity_words or (L.count32 * 600 + 16)
    bs = np.zeros(cap, dtype=np.uint32)
    meta = np.zeros((L.count32, 2), dtype=np.uint32)
    n = lib().or_pack(C.byref(L), ptr(coeffs), ptr(quant), sequence, ptr(bs), cap, ptr(meta))
    if n < 0:
        raise RuntimeError("bitstream capacity exceeded")
    return bs[:n].copy(), meta


def encode(L, planes, budget, sequence=1, precision=2, q1=True):
    """Encode u8 (UNORM8) or uint16 (R16_UNORM) planes."""
    planes, arr, pitches, bits = _plane_ptrs(planes)
    cap = L.count32 * 600 + 16
    bs = np.zeros(cap, dtype=np.uint32)
    meta = np.zeros((L.count32, 2), dtype=np.uint32)
    quant = np.zeros(L.count32, dtype=np.int32)
    n = lib().or_encode_ex(C.byref(L), arr, pitches, bits, precision, budget, sequence, 1 if q1 else 0,
                           ptr(bs), cap, ptr(meta), ptr(quant))
    if n < 0:
        raise RuntimeError("bitstream capacity exceeded")
    return bs[:n].copy(), meta, quant


def compute_num_packets(L, meta, boundary):
    return int(lib().or_compute_num_packets(C.byref(L), ptr(np.ascontiguousarray(meta, np.uint32)), boundary))


def packetize(L, meta, bitstream, sequence, boundary, out_size=None):
    meta = np.ascontiguousarray(meta, np.uint32)
    bitstream = np.ascontiguousarray(bitstream, np.uint32)
    total = 8 + int(meta[:, 1].sum()) * 4
    out = np.zeros(out_size or total, dtype=np.uint8)
    npk = compute_num_packets(L, meta, boundary) + 1
    packets = np.zeros((npk, 2), dtype=np.uint64)
    n = lib().or_packetize(C.byref(L), ptr(meta), ptr(bitstream), sequence, boundary, ptr(out), out.size,
                           ptr(packets))
    return out, packets[:n].astype(np.int64)


def validate_block(L, block_words, index, meta_words):
    w = np.ascontiguousarray(block_words, np.uint32)
    return bool(lib().or_validate_block(C.byref(L), ptr(w), index, meta_words))


class Decoder:
    def __init__(self, L):
        self.L = L
        self.h = lib().or_decoder_create(C.byref(L))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_decoder_destroy(self.h)
            self.h = None

    def clear(self):
        lib().or_decoder_clear(self.h)

    def push(self, data):
        data = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8))
        return bool(lib().or_decoder_push_packet(self.h, ptr(data), data.size))

    def ready(self, partial=False):
        return bool(lib().or_decoder_is_ready(self.h, 1 if partial else 0))

    def mark_decoded(self):
        lib().or_decoder_mark_decoded(self.h)

    def decoded_blocks(self):
        return lib().or_decoder_decoded_blocks(self.h)

    def total_blocks(self):
        return lib().or_decoder_total_blocks(self.h)

    def payload(self):
        n = lib().or_decoder_payload_words(self.h)
        p = lib().or_decoder_payload(self.h)
        if n == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(n,)).copy()

    def offsets(self):
        p = lib().or_decoder_offsets(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(self.L.count32,)).copy()


def dequant(L, payload, offsets):
    payload = np.ascontiguousarray(payload, np.uint32)
    coeffs = np.zeros(L.coeff_count, dtype=np.float32)
    lib().or_dequant(C.byref(L), ptr(payload), payload.size, ptr(np.ascontiguousarray(offsets, np.uint32)),
                     ptr(coeffs))
    return coeffs


def idwt_frame(L, coeffs, precision=2, bits=8):
    """Inverse transform into u8 planes, or uint16 (R16_UNORM) with bits=16."""
    coeffs = np.ascontiguousarray(coeffs, np.float32)
    planes = [np.zeros(plane_shape(L, c), dtype=np.uint16 if bits == 16 else np.uint8) for c in range(3)]
    arr = (C.c_void_p * 3)(*[p.ctypes.data for p in planes])
    pitches = (C.c_int * 3)(*[p.shape[1] for p in planes])
    lib().or_idwt_frame_ex(C.byref(L), ptr(coeffs), precision, arr, pitches, bits)
    return planes


def idwt_frame_f64(L, coeffs64):
    coeffs64 = np.ascontiguousarray(coeffs64, np.float64)
    planes = [np.zeros(plane_shape(L, c), dtype=np.float64) for c in range(3)]
    arr = (C.c_void_p * 3)(*[p.ctypes.data for p in planes])
    pitches = (C.c_int * 3)(*[p.shape[1] for p in planes])
    lib().or_idwt_frame_f64(C.byref(L), ptr(coeffs64), arr, pitches)
    return planes


def decode(L, packets_bytes, precision=2, bits=8):
    """Push every packet, dequantize and inverse transform: u8 planes (uint16
    with bits=16)."""
    d = Decoder(L)
    for p in packets_bytes:
        if not d.push(p):
            raise RuntimeError("push_packet failed")
    coeffs = dequant(L, d.payload(), d.offsets())
    return idwt_frame(L, coeffs, precision, bits)


def dwt_level(in_arr, R_x, R_y, A_x, A_y, f64=True):
    dt = np.float64 if f64 else np.float32
    in_arr = np.ascontiguousarray(in_arr, dtype=dt)
    outs = [np.zeros((A_y // 2, A_x // 2), dtype=dt) for _ in range(4)]
    fn = lib().or_dwt_level_f64 if f64 else lib().or_dwt_level_f32
    fn(ptr(in_arr), in_arr.shape[1], R_x, R_y, A_x, A_y, *[ptr(o) for o in outs], A_x // 2)
    return outs


def idwt_level(bands, f64=True):
    dt = np.float64 if f64 else np.float32
    bands = [np.ascontiguousarray(b, dtype=dt) for b in bands]
    bh, bw = bands[0].shape
    out = np.zeros((2 * bh, 2 * bw), dtype=dt)
    fn = lib().or_idwt_level_f64 if f64 else lib().or_idwt_level_f32
    fn(*[ptr(b) for b in bands], bw, bw, bh, ptr(out), 2 * bw)
    return out


# --- the reference's own bitstream layer (oracle/_ref), for golden fixtures ---

def dct8x8(sig):
    """8x8 butterfly DCT of shaders/psnr_hvs_m.comp (float32)."""
    sig = np.ascontiguousarray(sig, np.float32).reshape(64)
    out = np.zeros(64, np.float32)
    lib().or_dct8x8(ptr(sig), ptr(out))
    return out.reshape(8, 8)


def maskeff(sig, dct, mask):
    sig = np.ascontiguousarray(sig, np.float32).reshape(64)
    dct = np.ascontiguousarray(dct, np.float32).reshape(64)
    mask = np.ascontiguousarray(mask, np.float32).reshape(64)
    return float(lib().or_maskeff(ptr(sig), ptr(dct), ptr(mask)))


def psnr_hvs_m_tables(height, factor):
    """(csf, mask, inv_mask) 8x8 float32 tables of one height factor."""
    t = [np.zeros(64, np.float32) for _ in range(3)]
    lib().or_psnr_hvs_m_tables(height, factor, *[ptr(x) for x in t])
    return [x.reshape(8, 8) for x in t]


def psnr_height_factor(i):
    return float(lib().or_psnr_height_factor(i))


def psnr_hvs_m_totals(a, b, nfactors=16):
    """Per height factor, the summed 24-bit fixed-point block errors of luma
    planes a (reference) and b; and the total pixel count."""
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    assert a.shape == b.shape
    h, w = a.shape
    totals = np.zeros(16, np.uint64)
    tp = lib().or_psnr_hvs_m_totals(ptr(a), w, ptr(b), w, w, h, nfactors, ptr(totals))
    return totals[:nfactors], int(tp)


def psnr_hvs_m_db(total, total_pixels, full_range=True):
    return float(lib().or_psnr_hvs_m_db(int(total), int(total_pixels), 1 if full_range else 0))


_ref = None


def ref():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            raise RuntimeError("oracle/_ref/libref_bitstream.so not built")
        R = C.CDLL(REF_PATH)
        P = C.c_void_p
        R.ref_layout.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, C.c_int]
        R.ref_block_count_per_subdivision.argtypes = [C.c_int]
        R.ref_compute_num_packets.restype = C.c_size_t
        R.ref_compute_num_packets.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_size_t]
        R.ref_packetize.restype = C.c_size_t
        R.ref_packetize.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_size_t, P, C.c_size_t, P, P]
        R.ref_parser_create.restype = P
        R.ref_parser_create.argtypes = [C.c_int, C.c_int, C.c_int]
        for n in ("ref_parser_destroy", "ref_parser_clear", "ref_parser_mark_decoded"):
            getattr(R, n).argtypes = [P]
        R.ref_parser_push.argtypes = [P, P, C.c_size_t]
        R.ref_parser_ready.argtypes = [P, C.c_int]
        R.ref_parser_decoded_blocks.argtypes = [P]
        R.ref_parser_total_blocks.argtypes = [P]
        R.ref_parser_payload_words.restype = C.c_long
        R.ref_parser_payload_words.argtypes = [P]
        R.ref_parser_get.argtypes = [P, P, P]
        _ref = R
    return _ref
