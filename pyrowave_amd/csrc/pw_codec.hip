// pw_codec.hip -- per-32x32-block kernels of the MI355X PyroWave codec.
//
//   k_quant_analyze : quantizer + rate analysis fused; one wave (workgroup)
//                     per coded 32x32 block (reference shaders/wavelet_quant.comp
//                     and shaders/analyze_rate_control.comp).  The per-8x8
//                     statistics stay in LDS; only 48 B per block leave.
//   k_set_params    : per-call fields; clears the rate-control accumulators.
//   k_rc_finalize   : savings scan + per-(bucket,subdivision) classification
//                     (shaders/analyze_rate_control_finalize.comp).
//   k_rc_partial    : ordered greedy walk of the partially applied entries
//                     (shaders/resolve_rate_control.comp).
//   k_rc_prepare    : final truncation per block and the block-order word
//                     scan that places blocks contiguously.
//   k_pack          : block bitstream assembled in LDS, written with
//                     coalesced stores (shaders/block_packing.comp).
//   k_dequant       : block parse + reconstruction (shaders/wavelet_dequant.comp).
//
// Lane layout inside a 32x32 block (quantizer, packer, dequantizer: one
// 64-lane wave per block): lane = 4 * k8 + j, where k8 = 4 * y8 + x8 is the
// 8x8 block in raster order and the lane holds its 4x2 subblocks 2j and
// 2j + 1 (a 4x4 region); subblock s covers x = 4*(s>>2) .. +3,
// y = 2*(s&3) .. +1 and its pixel p sits at (p >> 1, p & 1) (reference
// bitstream/bitstream.md:350-383).  Lane order is the bitstream's subblock
// order, so every per-subblock stream position is a scan over the lanes.
#include "pw_common.hpp"

#include <hip/hip_fp16.h>

#pragma clang fp contract(off)

namespace pw
{
namespace
{
// Quantizer and packer: one wave per 32x32 block, lane = 4 * k8 + j holding
// subblocks 2j and 2j+1 of 8x8 block k8.
constexpr int kBlockLanes = 64;

// Smallest floats >= 2^0.25 and 2^0.75: floor(60.5 + 2*log2(r)) for
// r = m * 2^E, m in [1,2), is 60 + 2E + [m >= 2^0.25] + [m >= 2^0.75].
constexpr float kBucketT1 = 0x1.306fe2p+0f;
constexpr float kBucketT2 = 0x1.ae89fap+0f;

__device__ __forceinline__ int find_msb(uint32_t v)
{
	return v ? 31 - __clz(v) : -1;
}

// float -> fp16 bits, round to nearest even, subnormals kept.
__device__ __forceinline__ uint32_t f32_to_f16_bits(float f)
{
	uint32_t x = __float_as_uint(f);
	uint32_t sign = (x >> 16) & 0x8000u;
	uint32_t ax = x & 0x7fffffffu;
	if (ax >= 0x7f800000u)
		return sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u);
	if (ax == 0)
		return sign;
	int e = (int)(ax >> 23) - 127;
	uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
	if (e >= -14)
	{
		uint32_t hm = mant >> 13;
		uint32_t rem = mant & 0x1fffu;
		if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u)))
			hm++;
		int he = e + 15;
		if (hm & 0x800u)
		{
			hm >>= 1;
			he++;
		}
		if (he >= 31)
			return sign | 0x7c00u;
		return sign | ((uint32_t)he << 10) | (hm & 0x3ffu);
	}
	int shift = -e - 1;
	if (shift > 24)
		return sign;
	uint32_t hm = mant >> shift;
	uint32_t rem = mant & ((1u << shift) - 1u);
	uint32_t halfway = 1u << (shift - 1);
	if (rem > halfway || (rem == halfway && (hm & 1u)))
		hm++;
	return sign | hm;
}

__device__ __forceinline__ float f16_bits_to_f32(uint32_t h)
{
	return __half2float(__ushort_as_half((unsigned short)h));
}

// Cross-lane moves with DPP: quad_perm [1,0,3,2] and [2,3,0,1] swap
// neighbours at distance 1 and 2 (after both, all four lanes of a quad hold
// the butterfly ((v0+v1)+(v2+v3))), row_half_mirror pairs lane i with lane
// 7-i of each 8-lane half-row, row_mirror lane i with 15-i of each row.
constexpr int kDppXor1 = 0xB1;
constexpr int kDppXor2 = 0x4E;
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140; // lane i <-> 15 - i within each 16-lane row
constexpr int kDppRowRor4 = 0x124;   // lane i reads lane (i + 4) mod 16 of its 16-lane row
constexpr int kDppRowRor8 = 0x128;   // lane i reads lane (i + 8) mod 16 of its 16-lane row

template <int Ctrl>
__device__ __forceinline__ int dpp_i(int v)
{
	return __builtin_amdgcn_update_dpp(0, v, Ctrl, 0xF, 0xF, false);
}

template <int Ctrl>
__device__ __forceinline__ float dpp_f(float v)
{
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xF, 0xF, false));
}

// Sum over the 64 lanes of a wave (exact for integers): row butterfly with
// DPP, then the four row sums through readlane.
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
	int x = (int)v;
	x += dpp_i<kDppXor1>(x);
	x += dpp_i<kDppXor2>(x);
	x += dpp_i<kDppHalfMirror>(x);
	x += dpp_i<kDppRowMirror>(x);
	return (uint32_t)(__builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16) +
	                  __builtin_amdgcn_readlane(x, 32) + __builtin_amdgcn_readlane(x, 48));
}

// fp16 rounding (nearest even, subnormals kept) of 0 <= x <= 60000, returned
// as float: the value float(half(x)), two hardware conversions.  The empty asm
// keeps the backend from merging the conversion with the arithmetic that
// produced x into one mixed-precision rounding.
__device__ __forceinline__ float round_f16_nonneg(float x)
{
	asm volatile("" : "+v"(x));
	return __half2float(__float2half_rn(x));
}

// Two 16-bit lanes of one register: the quantized magnitudes (at most 16383)
// of two texels, so masks, minima and sums run on both at once.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Inclusive scan over the 64 lanes of a wave with DPP (all lanes active):
// row_shr 1, 2, 4, 8 scan each 16-lane row (lanes without a source read 0),
// then row_bcast 15 carries row 0 into row 1 and row 2 into row 3, and
// row_bcast 31 carries the first two rows into rows 2 and 3.
__device__ __forceinline__ int wave_incl_scan(int v)
{
	v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
	v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);
	v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);
	v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true);
	v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);
	v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);
	return v;
}

// Quantizer state of one lane: subblocks 2j and 2j+1 of its 8x8 block (a
// 4x4 region; lane = 4 * k8 + j).  Texel h * 8 + p is pixel p of subblock
// 2j + h, at column p >> 1 and row 2h + (p & 1) of the region.
struct LaneQuant
{
	float t[16];          // scaled texels
	int n[16];            // magnitudes
	int b4max[2];         // subblock max magnitudes
	int msb;              // 8x8 msb, -1 when the 8x8 quantizes to zero
	uint32_t scale_code;
	float iq2;            // (1 / (resolution * scale))^2
};

// Sums and maxima over the four lanes of one 8x8 block (a quad), in the
// order ((v0 + v1) + (v2 + v3)) in every lane.
__device__ __forceinline__ float quad_sum_f(float v)
{
	v = v + dpp_f<kDppXor1>(v);
	v = v + dpp_f<kDppXor2>(v);
	return v;
}

__device__ __forceinline__ int quad_sum_i(int v)
{
	v += dpp_i<kDppXor1>(v);
	v += dpp_i<kDppXor2>(v);
	return v;
}

__device__ __forceinline__ uint32_t quad_or(uint32_t v)
{
	v |= (uint32_t)dpp_i<kDppXor1>((int)v);
	v |= (uint32_t)dpp_i<kDppXor2>((int)v);
	return v;
}

// Maximum over a quad of non-negative, non-NaN floats: their bit patterns
// order as unsigned integers, so the integer maximum (one DPP instruction a
// step, no float canonicalization) returns the same float.
__device__ __forceinline__ float quad_max_nonneg_f(float v)
{
	uint32_t b = __float_as_uint(v);
	b = max(b, (uint32_t)dpp_i<kDppXor1>((int)b));
	b = max(b, (uint32_t)dpp_i<kDppXor2>((int)b));
	return __uint_as_float(b);
}

// reference shaders/wavelet_quant.comp:172-205, 292-316.
__device__ __forceinline__ void quantize_lane(const BandJob &J, int bx, int by, int k8, int j, LaneQuant &Q)
{
	const int x = bx * 32 + 8 * (k8 & 3) + 4 * (j >> 1);
	const int y = by * 32 + 8 * (k8 >> 2) + 4 * (j & 1);
	// Band storage is padded to whole 32x32 blocks (pitch and rows multiples
	// of 32, 256-byte aligned), so the four rows of the region load as one
	// aligned vector each.  The padding is zero: encoder bands are cleared
	// when the slot is made and every forward transform stores only inside
	// its band (x < bw, y < bh), so samples outside the band read as 0.
	float row[4][4];
	const long i0 = (long)y * J.pitch + x;
	if (J.kind == kStoreF16)
	{
		const __half *b16 = static_cast<const __half *>(J.data);
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			const uint2 w = *reinterpret_cast<const uint2 *>(b16 + i0 + r * J.pitch);
			row[r][0] = __half2float(__ushort_as_half((unsigned short)(w.x & 0xffffu)));
			row[r][1] = __half2float(__ushort_as_half((unsigned short)(w.x >> 16)));
			row[r][2] = __half2float(__ushort_as_half((unsigned short)(w.y & 0xffffu)));
			row[r][3] = __half2float(__ushort_as_half((unsigned short)(w.y >> 16)));
		}
	}
	else
	{
		const float *b32 = static_cast<const float *>(J.data);
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			const float4 w = *reinterpret_cast<const float4 *>(b32 + i0 + r * J.pitch);
			row[r][0] = w.x;
			row[r][1] = w.y;
			row[r][2] = w.z;
			row[r][3] = w.w;
		}
	}
	float ms[2] = {0.0f, 0.0f};
#pragma unroll
	for (int h = 0; h < 2; h++)
#pragma unroll
		for (int p = 0; p < 8; p++)
		{
			// Magnitudes beyond 16383.75 would overflow the 14 planes a block
			// can describe; they are clamped (reachable only far outside the
			// nominal coefficient range).
			const float v = row[2 * h + (p & 1)][p >> 1];
			const float t = fminf(fmaxf(v * J.qres, -16383.75f), 16383.75f);
			Q.t[8 * h + p] = t;
			ms[h] = fmaxf(ms[h], fabsf(t));
		}
	// ms >= 0 and never NaN (t is clamped with fminf / fmaxf).
	float M = quad_max_nonneg_f(fmaxf(ms[0], ms[1]));

	float quant_scale;
	if (M < 1.0f)
	{
		Q.scale_code = 6;
		quant_scale = 1.0f;
	}
	else
	{
		int e;
		frexpf(M - 0.25f, &e);
		float target_max = (float)(1 << e) - 0.25f;
		float inv_scale = M / target_max;
		Q.scale_code = (uint32_t)ceilf((inv_scale - 0.25f) * 8.0f);
		quant_scale = 1.0f / ((float)Q.scale_code / 8.0f + 0.25f);
	}
#pragma unroll
	for (int p = 0; p < 16; p++)
		Q.t[p] = Q.t[p] * quant_scale;
	M = M * quant_scale;
	float inv_quant = 1.0f / (J.qres * quant_scale);
	Q.iq2 = inv_quant * inv_quant;
#pragma unroll
	for (int p = 0; p < 16; p++)
		Q.n[p] = abs((int)Q.t[p]);
	// max(|t|) * scale = max(|t * scale|): rounding is monotone.
	Q.b4max[0] = (int)(ms[0] * quant_scale);
	Q.b4max[1] = (int)(ms[1] * quant_scale);
	int maxabs = (int)M;
	Q.msb = find_msb((uint32_t)maxabs);
}

__device__ __forceinline__ float dot4_sq(float a, float b, float c, float d)
{
	return ((a * a + b * b) + c * c) + d * d;
}

// Squared magnitude of one subblock's eight values, in the reference's order.
__device__ __forceinline__ float sub_sq(const float *v)
{
	return dot4_sq(v[0], v[1], v[2], v[3]) + dot4_sq(v[4], v[5], v[6], v[7]);
}

// Code word of an 8x8 block at quant 0 (reference shaders/wavelet_quant.comp:219-229).
__device__ __forceinline__ uint32_t code_word_q0(const LaneQuant &Q, int j)
{
	if (Q.msb < 0)
		return 0;
	int qp0 = Q.msb >= 3 ? Q.msb - 2 : 0;
	const uint32_t code0 = (uint32_t)(find_msb((uint32_t)(Q.b4max[0] >> qp0)) + 1);
	const uint32_t code1 = (uint32_t)(find_msb((uint32_t)(Q.b4max[1] >> qp0)) + 1);
	uint32_t cw = quad_or((code0 | (code1 << 2)) << (4 * j));
	return cw | ((uint32_t)qp0 << 16) | ((Q.scale_code & 0xfu) << 20);
}

__device__ __forceinline__ uint32_t bucket_of_ratio(float r)
{
	if (!(r > 0.0f))
		return 0;
	int e;
	float f = frexpf(r, &e);
	float m = 2.0f * f;
	int k = 60 + 2 * (e - 1) + (m >= kBucketT1 ? 1 : 0) + (m >= kBucketT2 ? 1 : 0);
	return k < 0 ? 0u : (uint32_t)k;
}

struct BlockCoord
{
	BandJob J;
	int b;   // frame-local block index
	int bx, by;
};

// blk_band holds two words per block of the launch: the band job index and
// the block position bx | by << 16 within its band.
__device__ __forceinline__ BlockCoord locate_block(const BandJob *bands, const uint32_t *blk_band, XcdMap xm)
{
	BlockCoord C;
	const int g = xcd_swizzle((int)blockIdx.x, xm);
	const uint2 d = reinterpret_cast<const uint2 *>(blk_band)[g];
	C.J = bands[__builtin_amdgcn_readfirstlane(d.x)];
	const uint32_t pos = __builtin_amdgcn_readfirstlane(d.y);
	C.bx = (int)(pos & 0xffffu);
	C.by = (int)(pos >> 16);
	C.b = C.J.off32 + C.by * C.J.bx32 + C.bx;
	return C;
}

__global__ __launch_bounds__(kBlockLanes) void k_quant_analyze(const BandJob *__restrict__ bands,
                                                               const uint32_t *__restrict__ blk_band,
                                                               const FrameState *__restrict__ frames,
                                                               XcdMap xm)
{
	__shared__ int s_cost[16][16];
	__shared__ float s_dist[16][16];
	__shared__ int s_msb[16];
	__shared__ float s_disc[16];
	__shared__ uint32_t s_words[16];
	__shared__ float s_D[16];
	__shared__ uint32_t s_v[16];

	const BlockCoord C = locate_block(bands, blk_band, xm);
	const BandJob &J = C.J;
	const int tid = threadIdx.x, j = tid & 3, k8 = tid >> 2;
	const int x8 = C.bx * 4 + (k8 & 3), y8 = C.by * 4 + (k8 >> 2);
	const bool in_range = x8 < J.bx8 && y8 < J.by8;

	LaneQuant Q;
	quantize_lane(J, C.bx, C.by, k8, j, Q);
	const int msb = in_range ? Q.msb : -1;

	// Distortion of discarding the whole 8x8 block; the reference leaves out
	// rdo_distortion_scale here (reference shaders/wavelet_quant.comp:254).
	// Per subblock, then over the eight in the order ((s0 + s1) + (s2 + s3))
	// + ((s4 + s5) + (s6 + s7)): the lane adds its pair, the quad the rest.
	const float dd = quad_sum_f(sub_sq(Q.t) * Q.iq2 + sub_sq(Q.t + 8) * Q.iq2);
	if (j == 0)
	{
		s_msb[k8] = msb;
		s_disc[k8] = msb >= 0 ? round_f16_nonneg(fminf(dd, 60000.0f)) : 0.0f;
	}
	// The wave runs the levels of its sixteen 8x8 blocks up to their largest
	// msb (the analyze below reads level q of a block only for q <= its msb):
	// row_ror 4 and 8 reduce each 16-lane row, the four row maxima are read
	// out as scalars.
	int mrow = max(msb, dpp_i<kDppRowRor4>(msb));
	mrow = max(mrow, dpp_i<kDppRowRor8>(mrow));
	const int qmax = max(max(__builtin_amdgcn_readlane(mrow, 0), __builtin_amdgcn_readlane(mrow, 16)),
	                     max(__builtin_amdgcn_readlane(mrow, 32), __builtin_amdgcn_readlane(mrow, 48)));

	float a[16];
#pragma unroll
	for (int p = 0; p < 16; p++)
		a[p] = fabsf(Q.t[p]);
	u16x2 npair[8];
#pragma unroll
	for (int i = 0; i < 8; i++)
		npair[i] = u16x2{(unsigned short)Q.n[2 * i], (unsigned short)Q.n[2 * i + 1]};
	// Late plane count of a subblock at quant level q, max(early - 1, 0) of
	// reference shaders/wavelet_quant.comp:139-163.  With m = msb - q >= 3
	// the early count is (m - 2) quality planes + a sign plane + the planes
	// of (b4max >> q) >> (m - 2) = b4max >> (msb - 2), which does not depend
	// on q; otherwise it is 1 + the planes of b4max >> q when that is
	// nonzero, else 0.  So both cases are a subtraction from a per-subblock
	// constant (checked for every msb <= 13, b4max < 2^(msb+1), q <= msb).
	int m4[2], fq[2];
#pragma unroll
	for (int h = 0; h < 2; h++)
	{
		m4[h] = find_msb((uint32_t)Q.b4max[h]);
		fq[h] = msb >= 2 ? find_msb((uint32_t)Q.b4max[h] >> (msb - 2)) : 0;
	}
	auto late_planes = [&](int q) {
		int sum = 0;
#pragma unroll
		for (int h = 0; h < 2; h++)
			sum += q <= msb - 3 ? msb - q + fq[h] - 1 : max(m4[h] - q + 1, 0);
		return sum;
	};
	// Level 0: cost only (its distortion is stored as 0, as the reference does).
	{
		int packed = 0;
		if (msb >= 0)
		{
			int nsig = 0;
#pragma unroll
			for (int p = 0; p < 16; p++)
				nsig += Q.n[p] != 0 ? 1 : 0;
			packed = late_planes(0) | (nsig << 16);
		}
		const int sums = quad_sum_i(packed);
		if (j == 0 && msb >= 0)
		{
			s_cost[0][k8] = 8 * (sums & 0xffff) + (sums >> 16);
			s_dist[0][k8] = 0.0f;
		}
	}
	for (int q = 1; q <= qmax; q++)
	{
		// compute_quant_stats (reference shaders/wavelet_quant.comp:113-165).
		float se_lane = 0.0f;
		int packed = 0; // late planes | significant count << 16
		if (q <= msb)
		{
			// A magnitude n is significant at level q when n >= 2^q; its
			// reconstruction (floor(n / 2^q) + 1/2) * 2^q is the integer
			// r + 2^(q-1), r = n with the low q bits cleared, exact in float.
			// r is 0 or at least 2^q, so min(r, 2^(q-1)) is 2^(q-1) exactly
			// for the significant magnitudes and 0 for the others: r plus it
			// is the reconstruction (0 when not significant), and its sum
			// over the sixteen is 2^(q-1) times their significant count.
			const unsigned short keep = (unsigned short)(0xffffu << q), half_bit = (unsigned short)(1u << (q - 1));
			const u16x2 keep2 = {keep, keep}, half2 = {half_bit, half_bit};
			u16x2 msum = {0, 0};
			float err[16];
#pragma unroll
			for (int i = 0; i < 8; i++)
			{
				const u16x2 r = npair[i] & keep2;
				const u16x2 m = __builtin_elementwise_min(r, half2);
				const u16x2 recon = r + m;
				msum += m;
				err[2 * i] = a[2 * i] - (float)recon.x;
				err[2 * i + 1] = a[2 * i + 1] - (float)recon.y;
			}
			const int nsig = (int)(((uint32_t)msum.x + (uint32_t)msum.y) >> (q - 1));
			se_lane = (sub_sq(err) * J.rdo_scale) * Q.iq2 + (sub_sq(err + 8) * J.rdo_scale) * Q.iq2;
			packed = late_planes(q) | (nsig << 16);
		}
		const float se = quad_sum_f(se_lane);
		const int sums = quad_sum_i(packed);
		if (j == 0 && q <= msb)
		{
			s_cost[q][k8] = 8 * (sums & 0xffff) + (sums >> 16);
			s_dist[q][k8] = round_f16_nonneg(fminf(se, 60000.0f));
		}
	}
	__syncthreads();

	const FrameState &F = frames[J.frame];
	{
		// reference shaders/analyze_rate_control.comp:169-215.  Lanes 4q ..
		// 4q+3 sum quant level q over the 16 8x8 blocks, four blocks per lane
		// as ((d0 + d1) + (d2 + d3)); the quad then adds the lane sums, so the
		// total is (((d0+d1)+(d2+d3))+((d4+d5)+(d6+d7))) + (same for d8..d15).
		const int q = tid >> 2;
		int cost4 = 0;
		float d[4];
#pragma unroll
		for (int i = 0; i < 4; i++)
		{
			const int k = 4 * j + i;
			// errors[min(q, num_planes)]: past the last plane the entry is the
			// cost-free "discard" one; zero 8x8 blocks are (0, 0) throughout.
			const bool live = q <= s_msb[k];
			const int c = live ? s_cost[q][k] : 0;
			d[i] = live ? s_dist[q][k] : s_disc[k];
			cost4 += c ? c + 24 : 0;
		}
		uint32_t total = (uint32_t)quad_sum_i(cost4);
		const float D = quad_sum_f((d[0] + d[1]) + (d[2] + d[3]));
		if (j == 0)
		{
			if (total)
				total += 64;
			s_words[q] = (total + 31) >> 5;
			s_D[q] = D;
		}
	}
	__syncthreads();
	if (tid < 16)
	{
		// emit_rdo_operations (reference shaders/analyze_rate_control.comp:101-147).
		const int q = tid;
		uint32_t bi = 0;
		if (q != 0)
		{
			float cost_f = (float)s_words[q], base_f = (float)s_words[0];
			if (cost_f != base_f)
			{
				float dd2 = s_D[q] - s_D[0];
				dd2 = dd2 > 0.0f ? dd2 : 0.0f;
				bi = bucket_of_ratio(dd2 / (base_f - cost_f));
			}
		}
		s_v[q] = bi < 112u + (uint32_t)q ? bi : 112u + (uint32_t)q;
	}
	__syncthreads();
	{
		// Bucket of level q: the maximum over k <= q of s_v[k] + (q - k)
		// (inclusive_max_clustered16, reference shaders/analyze_rate_control.comp:85-99), that is q plus
		// the prefix maximum of s_v[k] - k over lanes 0..q: a row-shift DPP
		// max scan on s_v[k] - k + 16 (positive, so the zero the first lanes
		// of the row read in never wins).
		const int q = tid;
		int m = tid < 16 ? (int)s_v[q] - q + 16 : 0;
		m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x111, 0xF, 0xF, true));
		m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x112, 0xF, 0xF, true));
		m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x114, 0xF, 0xF, true));
		m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x118, 0xF, 0xF, true));
		if (tid >= 16)
			return;
		const uint32_t Bq = (uint32_t)(m - 16 + q);
		F.words[(long)C.b * 16 + q] = (uint16_t)s_words[q];
		F.bucket[(long)C.b * 16 + q] = (uint8_t)Bq;
		if (q != 0)
		{
			uint32_t saving = s_words[q - 1] - s_words[q];
			if (saving)
				atomicAdd(&F.totals[(C.b & (F.totals_copies - 1)) * kFlat + Bq * kSubdiv + ((uint32_t)C.b >> F.shamt)],
				          saving);
		}
		else
			// Unquantized words, also into a contiguous per-block array that
			// the finalize sums (the reference adds them with one atomic per
			// block into one word, shaders/analyze_rate_control.comp:127-131; a
			// frame's thousands of same-address atomics serialize here).
			F.words0[C.b] = (uint16_t)s_words[0];
	}
}

// A frame's per-call fields, written by one thread.
__device__ __forceinline__ void write_call_fields(FrameState &F, const CallParams &P)
{
	for (int c = 0; c < 3; c++)
	{
		F.src[c] = P.src[c];
		F.src_pitch[c] = P.src_pitch[c];
		F.src_step[c] = P.src_step[c];
		F.dst[c] = P.dst[c];
		F.dst_pitch[c] = P.dst_pitch[c];
	}
	F.sequence = P.sequence;
	F.target_words = P.target_words;
	F.flags = P.flags;
	F.dec_offsets = P.dec_offsets;
	F.dec_payload = P.dec_payload;
	F.dec_offset_stride = P.dec_offset_stride;
	F.dec_payload_bytes = P.dec_payload_bytes;
}

// Grid (clear blocks, frames of the parameter block).  Block (0, i) writes
// frame i's per-call fields; every block clears its share of the frame's
// rate-control accumulators (encoder frames only: a decoder's FrameState has
// no rate-control arrays).  A kernel, not memsets: it is captured into graphs.
__global__ void k_set_params(FrameState *__restrict__ frames, ParamBlock block, int first, int count)
{
	const int i = blockIdx.y;
	if (i >= count)
		return;
	FrameState &F = frames[first + i];
	if (F.totals)
	{
		const int kTotals = F.totals_copies * kFlat;
		const int n = F.n32 > kTotals ? F.n32 : kTotals;
		for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
		{
			if (k < kTotals)
				F.totals[k] = 0;
			if (k < F.n32)
				F.partial_quant[k] = 0;
			if (k < kRcCount)
				F.rc[k] = 0;
		}
	}
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	write_call_fields(F, block.p[i]);
}

// The finalize scan of reference shaders/analyze_rate_control_finalize.comp:
// 14-38 on 512 threads holding four entries each (v: the thread's inclusive
// in-thread prefix; w = v[3] its total).  The reference's Hillis-Steele loop
// runs steps 1 .. 128 (quirk Q1), after which thread t holds its own prefix
// plus the totals of threads t - 255 .. t - 1; the exact variant (steps up to
// 256) adds every earlier thread's.  Both are differences of one exclusive
// prefix P of the totals: P[t] - P[max(0, t - 255)], or P[t] (exact modulo
// 2^32, like the sums themselves).  Threads 0-511 hold the entries (`scanner`);
// every thread of the workgroup reaches the barriers.  On return s_scan[t]
// holds v[3] for t < 512.
__device__ __forceinline__ void window_scan512(uint32_t (&v)[4], bool exact, uint32_t *s_scan, bool scanner)
{
	__shared__ uint32_t s_wave[8];
	const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
	const uint32_t w = v[3];
	const uint32_t incl = (uint32_t)wave_incl_scan((int)w);
	if (scanner && lane == 63)
		s_wave[wave] = incl;
	__syncthreads();
	uint32_t excl = 0;
	if (scanner)
	{
		uint32_t base = 0;
#pragma unroll
		for (int k = 0; k < 8; k++)
			base += k < wave ? s_wave[k] : 0u;
		excl = base + incl - w;
		s_scan[t] = excl;
	}
	__syncthreads();
	const uint32_t add = exact ? excl : excl - (scanner && t >= 255 ? s_scan[t - 255] : 0u);
	__syncthreads();
	if (scanner)
	{
#pragma unroll
		for (int i = 0; i < 4; i++)
			v[i] += add;
		s_scan[t] = v[3];
	}
	__syncthreads();
}

// Totals 4t .. 4t + 3, summed over the copies.
__device__ __forceinline__ uint4 sum_total_copies(const uint32_t *totals, int copies, int t)
{
	uint4 s = make_uint4(0u, 0u, 0u, 0u);
	for (int r = 0; r < copies; r++)
	{
		const uint4 c = reinterpret_cast<const uint4 *>(totals + r * kFlat)[t];
		s.x += c.x;
		s.y += c.y;
		s.z += c.z;
		s.w += c.w;
	}
	return s;
}

// reference shaders/analyze_rate_control_finalize.comp:14-38 (512 threads x 4
// entries) and the skip / required logic of shaders/resolve_rate_control.comp:38-56.
__global__ __launch_bounds__(512) void k_rc_finalize(FrameState *__restrict__ frames)
{
	__shared__ uint32_t s_scan[512];
	__shared__ uint32_t s_consumed;
	const FrameState &F = frames[blockIdx.x];
	const int t = threadIdx.x;
	{
		// Words of every block at quant level 0 (consumed payload), from the
		// contiguous per-block array: eight blocks per 16-byte load (the array
		// is padded to whole groups of eight, with zeros past the frame).
		const uint4 *w = reinterpret_cast<const uint4 *>(F.words0);
		const int groups = (F.n32 + 7) >> 3;
		uint32_t acc = 0;
		for (int g = t; g < groups; g += 512)
		{
			const uint4 q = w[g];
			acc += (q.x & 0xffffu) + (q.x >> 16) + (q.y & 0xffffu) + (q.y >> 16) + (q.z & 0xffffu) + (q.z >> 16) +
			       (q.w & 0xffffu) + (q.w >> 16);
		}
		for (int d = 32; d >= 1; d >>= 1)
			acc += __shfl_xor(acc, d, 64);
		if (t == 0)
			s_consumed = 0;
		__syncthreads();
		if ((t & 63) == 0)
			atomicAdd(&s_consumed, acc);
		__syncthreads();
		if (t == 0)
			F.rc[kRcConsumed] = s_consumed;
	}
	const uint4 tv = sum_total_copies(F.totals, F.totals_copies, t);
	const uint32_t total4[4] = {tv.x, tv.y, tv.z, tv.w};
	uint32_t v[4] = {tv.x, tv.y, tv.z, tv.w};
	v[1] += v[0];
	v[2] += v[1];
	v[3] += v[2];
	window_scan512(v, (F.flags & 1u) != 0, s_scan, true);
	const int required_base = (int)s_consumed - F.target_words;
	uint32_t prev = t ? s_scan[t - 1] : 0u;
#pragma unroll
	for (int i = 0; i < 4; i++)
	{
		const int x = 4 * t + i;
		F.prefix[x] = v[i];
		bool empty;
		int required;
		if (x == 0)
		{
			empty = v[0] == 0;
			required = required_base;
		}
		else
		{
			empty = v[i] == prev;
			required = required_base - (int)prev;
		}
		uint8_t cls = kRcSkip;
		if (!empty && required > 0)
			cls = (int)total4[i] <= required ? kRcFull : kRcPartial;
		F.cls[x] = cls;
		F.required[x] = required;
		prev = v[i];
	}
}

__device__ __forceinline__ int block_incl_scan256(int v, int *s_tmp, int &total)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int incl = wave_incl_scan(v);
	if (lane == 63)
		s_tmp[wave] = incl;
	__syncthreads();
	int base = 0;
	for (int w = 0; w < wave; w++)
		base += s_tmp[w];
	total = s_tmp[0] + s_tmp[1] + s_tmp[2] + s_tmp[3];
	__syncthreads();
	return base + incl;
}

// One workgroup per (bucket, subdivision) entry; only partially applied
// entries walk their operations (reference shaders/resolve_rate_control.comp:57-79).
__global__ __launch_bounds__(256) void k_rc_partial(FrameState *__restrict__ frames)
{
	__shared__ int s_tmp[4];
	const FrameState &F = frames[blockIdx.y];
	const int x = blockIdx.x;
	if (F.cls[x] != kRcPartial)
		return;
	const int required = F.required[x];
	const uint32_t beta = (uint32_t)x >> 4;
	const int sd = x & 15;
	const int base = sd * F.per_sub;
	int end = base + F.per_sub;
	end = end < F.n32 ? end : F.n32;
	int total_saved = 0;
	for (int chunk = base; chunk < end && total_saved < required; chunk += 256)
	{
		const int b = chunk + (int)threadIdx.x;
		int saving = 0, q = 0;
		if (b < end)
		{
			for (int qq = 1; qq < 16; qq++)
			{
				if (F.bucket[(long)b * 16 + qq] == beta)
				{
					saving = (int)F.words[(long)b * 16 + qq - 1] - (int)F.words[(long)b * 16 + qq];
					q = qq;
					break;
				}
			}
		}
		int sum;
		int incl = block_incl_scan256(saving, s_tmp, sum);
		if (saving != 0 && total_saved + incl - saving < required)
			atomicMax(&F.partial_quant[b], q);
		total_saved += sum;
	}
}

// Final truncation per block and word offsets within 256-block groups.
__global__ __launch_bounds__(256) void k_rc_prepare(FrameState *__restrict__ frames, const uint32_t *__restrict__ grp_frame,
                                                    const int32_t *__restrict__ frame_group_base)
{
	__shared__ int s_tmp[4];
	const int f = (int)grp_frame[blockIdx.x];
	const FrameState &F = frames[f];
	const int g = (int)blockIdx.x - frame_group_base[f];
	const int b = g * 256 + (int)threadIdx.x;
	int wf = 0;
	if (b < F.n32)
	{
		int qf = F.partial_quant[b];
		const uint32_t sd = (uint32_t)b >> F.shamt;
		for (int q = 1; q < 16; q++)
		{
			if (F.words[(long)b * 16 + q - 1] != F.words[(long)b * 16 + q])
			{
				uint32_t x = (uint32_t)F.bucket[(long)b * 16 + q] * kSubdiv + sd;
				if (F.cls[x] == kRcFull && q > qf)
					qf = q;
			}
		}
		F.quant[b] = (uint8_t)qf;
		wf = F.words[(long)b * 16 + qf];
	}
	int total;
	int incl = block_incl_scan256(wf, s_tmp, total);
	if (b < F.n32)
		F.excl[b] = (uint32_t)(incl - wf);
	if (threadIdx.x == 0)
		F.group_sum[g] = (uint32_t)total;
}

// Inclusive scan over a 1024-thread workgroup (16 waves); total: the sum.
__device__ __forceinline__ int block_incl_scan1024(int v, int *s_wave, int &total)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int incl = wave_incl_scan(v);
	if (lane == 63)
		s_wave[wave] = incl;
	__syncthreads();
	int base = 0;
	total = 0;
#pragma unroll
	for (int w = 0; w < 16; w++)
	{
		const int x = s_wave[w];
		base += w < wave ? x : 0;
		total += x;
	}
	__syncthreads();
	return base + incl;
}

// The three rate-control steps of one frame in one 1024-thread workgroup:
// the finalize of k_rc_finalize (its 512-thread scan reproduced by threads
// 0-511), the partial entries of k_rc_partial walked one after another by the
// whole workgroup, and the per-block truncation and group offsets of
// k_rc_prepare, with the classes in LDS: two launch boundaries fewer for a
// lone frame of up to kResolveMaxBlocks blocks.
__global__ __launch_bounds__(1024) void k_rc_resolve(FrameState *__restrict__ frames)
{
	__shared__ uint32_t s_scan[512];
	__shared__ uint8_t s_cls[kFlat];
	__shared__ uint16_t s_plist[kFlat];
	__shared__ int s_preq[kFlat];
	__shared__ uint32_t s_pq[kResolveMaxBlocks / 4]; // partial truncation per block, one byte each

	__shared__ int s_wave[16];
	__shared__ uint32_t s_consumed;
	__shared__ int s_npartial;
	FrameState &F = frames[blockIdx.x];
	const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
	const int n32 = F.n32;

	// Every load that does not depend on the scan goes out first: the
	// totals, the level-0 words and the first chunk's block records.
	const bool scanner = t < 512;
	uint32_t v[4] = {0u, 0u, 0u, 0u};
	if (scanner)
	{
		const uint4 tv = sum_total_copies(F.totals, F.totals_copies, t);
		v[0] = tv.x;
		v[1] = tv.y;
		v[2] = tv.z;
		v[3] = tv.w;
	}
	// Thread t owns blocks t + 1024 i; their records (16 word counts, 16
	// buckets) are held in registers from here on.
	constexpr int kPre = 4;
	static_assert(kResolveMaxBlocks <= kPre * 1024, "every block record is held in registers");
	uint4 rw0[kPre], rw1[kPre], rbk[kPre];
#pragma unroll
	for (int i = 0; i < kPre; i++)
	{
		const int b = t + 1024 * i;
		rw0[i] = rw1[i] = rbk[i] = make_uint4(0u, 0u, 0u, 0u);
		if (b < n32)
		{
			const uint4 *wp = reinterpret_cast<const uint4 *>(F.words + (long)b * 16);
			rw0[i] = wp[0];
			rw1[i] = wp[1];
			rbk[i] = *reinterpret_cast<const uint4 *>(F.bucket + (long)b * 16);
		}
	}
	uint32_t acc = 0;
	{
		// Words at quant level 0, eight blocks per 16-byte load (the array is
		// padded to whole groups of eight with zeros).
		const uint4 *w = reinterpret_cast<const uint4 *>(F.words0);
		const int groups = (n32 + 7) >> 3;
		for (int g = t; g < groups; g += 1024)
		{
			const uint4 q = w[g];
			acc += (q.x & 0xffffu) + (q.x >> 16) + (q.y & 0xffffu) + (q.y >> 16) + (q.z & 0xffffu) + (q.z >> 16) +
			       (q.w & 0xffffu) + (q.w >> 16);
		}
	}
	for (int i = t; i < (n32 + 3) / 4; i += 1024)
		s_pq[i] = 0;
	if (t == 0)
	{
		s_consumed = 0;
		s_npartial = 0;
	}
	__syncthreads();
	for (int d = 32; d >= 1; d >>= 1)
		acc += __shfl_xor(acc, d, 64);
	if (lane == 0)
		atomicAdd(&s_consumed, acc);

	// Finalize scan: threads 0-511 hold four entries each
	// (reference shaders/analyze_rate_control_finalize.comp:14-38).
	const uint32_t total4[4] = {v[0], v[1], v[2], v[3]};
	if (scanner)
	{
		v[1] += v[0];
		v[2] += v[1];
		v[3] += v[2];
	}
	window_scan512(v, (F.flags & 1u) != 0, s_scan, scanner);
	if (scanner)
	{
		const int required_base = (int)s_consumed - F.target_words;
		uint32_t prev = t ? s_scan[t - 1] : 0u;
		uint8_t cls4[4];
		int req4[4];
#pragma unroll
		for (int i = 0; i < 4; i++)
		{
			const int x = 4 * t + i;
			const bool empty = x == 0 ? v[0] == 0 : v[i] == prev;
			const int required = x == 0 ? required_base : required_base - (int)prev;
			uint8_t cls = kRcSkip;
			if (!empty && required > 0)
				cls = (int)total4[i] <= required ? kRcFull : kRcPartial;
			cls4[i] = cls;
			req4[i] = required;
			s_cls[x] = cls;
			if (cls == kRcPartial)
			{
				const int k = atomicAdd(&s_npartial, 1);
				s_plist[k] = (uint16_t)x;
				s_preq[k] = required;
			}
			prev = v[i];
		}
		reinterpret_cast<uint4 *>(F.prefix)[t] = make_uint4(v[0], v[1], v[2], v[3]);
		reinterpret_cast<int4 *>(F.required)[t] = make_int4(req4[0], req4[1], req4[2], req4[3]);
		reinterpret_cast<uint32_t *>(F.cls)[t] = (uint32_t)cls4[0] | ((uint32_t)cls4[1] << 8) |
		                                         ((uint32_t)cls4[2] << 16) | ((uint32_t)cls4[3] << 24);
	}
	if (t == 0)
		F.rc[kRcConsumed] = s_consumed;
	__syncthreads();

	// Partially applied entries (reference shaders/resolve_rate_control.comp:57-79):
	// the blocks of the entry's subdivision in order; a block whose step into
	// the entry's bucket keeps the running saving before it under the
	// requirement takes that step.
	uint8_t *const pq = reinterpret_cast<uint8_t *>(s_pq);
	const int npartial = s_npartial;
	for (int k = 0; k < npartial; k++)
	{
		const int x = s_plist[k];
		const int required = s_preq[k];
		const uint32_t beta = (uint32_t)x >> 4;
		const int base = (x & 15) * F.per_sub;
		const int end = base + F.per_sub < n32 ? base + F.per_sub : n32;
		int total_saved = 0;
		// Chunks aligned to 1024 blocks, so that block b is the one its
		// thread holds; blocks outside [base, end) save nothing.
		auto step = [&](int ci, const uint4 &w0, const uint4 &w1, const uint4 &bk) {
			const int b = (ci << 10) + t;
			int saving = 0, q = 0;
			if (b >= base && b < end)
			{
				const uint32_t bw[4] = {bk.x, bk.y, bk.z, bk.w};
				const uint32_t wv[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
				for (int qq = 1; qq < 16; qq++)
				{
					if (((bw[qq >> 2] >> (8 * (qq & 3))) & 0xffu) == beta)
					{
						const int a = (int)((wv[(qq - 1) >> 1] >> (16 * ((qq - 1) & 1))) & 0xffffu);
						const int c = (int)((wv[qq >> 1] >> (16 * (qq & 1))) & 0xffffu);
						saving = a - c;
						q = qq;
						break;
					}
				}
			}
			int sum;
			const int incl = block_incl_scan1024(saving, s_wave, sum);
			// Blocks of one subdivision are distinct, and entries run one
			// after another, so the byte's read-modify-write is race-free.
			if (saving != 0 && total_saved + incl - saving < required && q > pq[b])
				pq[b] = (uint8_t)q;
			total_saved += sum;
		};
		const int c0 = base >> 10, c1 = (end + 1023) >> 10;
#pragma unroll
		for (int ci = 0; ci < kPre; ci++)
			if (ci >= c0 && ci < c1 && total_saved < required)
				step(ci, rw0[ci], rw1[ci], rbk[ci]);
		__syncthreads();
	}

	// Final truncation per block and word offsets within 256-block groups
	// (four groups per 1024-block chunk, four waves each).
	auto truncate = [&](int ci, const uint4 &w0, const uint4 &w1, const uint4 &bk) {
		const int b = (ci << 10) + t;
		int wf = 0;
		if (b < n32)
		{
			const uint32_t wv[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
			const uint32_t bw[4] = {bk.x, bk.y, bk.z, bk.w};
			auto words = [&](int q) { return (wv[q >> 1] >> (16 * (q & 1))) & 0xffffu; };
			int qf = pq[b];
			const uint32_t sd = (uint32_t)b >> F.shamt;
#pragma unroll
			for (int q = 1; q < 16; q++)
			{
				if (words(q - 1) != words(q))
				{
					const uint32_t bq = (bw[q >> 2] >> (8 * (q & 3))) & 0xffu;
					if (s_cls[bq * kSubdiv + sd] == kRcFull && q > qf)
						qf = q;
				}
			}
			F.quant[b] = (uint8_t)qf;
			wf = (int)words(qf);
		}
		const int incl = wave_incl_scan(wf);
		if (lane == 63)
			s_wave[wave] = incl;
		__syncthreads();
		const int grp = wave >> 2;
		int gbase = 0, gtotal = 0;
#pragma unroll
		for (int w = 0; w < 4; w++)
		{
			const int x = s_wave[4 * grp + w];
			gbase += 4 * grp + w < wave ? x : 0;
			gtotal += x;
		}
		if (b < n32)
			F.excl[b] = (uint32_t)(gbase + incl - wf);
		const int g = b >> 8;
		if ((t & 255) == 0 && g < F.ngroups)
			F.group_sum[g] = (uint32_t)gtotal;
		__syncthreads();
	};
	const int chunks = (n32 + 1023) >> 10;
#pragma unroll
	for (int ci = 0; ci < kPre; ci++)
		if (ci < chunks)
			truncate(ci, rw0[ci], rw1[ci], rbk[ci]);
}

// reference shaders/block_packing.comp:107-142: the block's 8x8 code word
// truncated by `quant` levels.  The q bits absorb the first levels; each
// level beyond them lowers every 4x2 subblock's 2-bit plane code c to
// max(c - 1, 0), which the reference does by shifting planes lo / hi / both
// (c >= 1, c >= 2, c == 3) down; written here without its loop: one level
// left gives codes (hi & ~lo) | both << 1, two give both, three or more 0
// (checked against the reference loop on 3.2 million code words).
__device__ __forceinline__ uint32_t quantize_code_word(uint32_t cw, int quant)
{
	if (cw == 0)
		return 0;
	const int q_bits = (int)((cw >> 16) & 0xf);
	const int sub = q_bits < quant ? q_bits : quant;
	const int rest = quant - sub;
	const uint32_t lo = cw & 0x5555u, hi = (cw >> 1) & 0x5555u, both = lo & hi;
	uint32_t planes = cw & 0xffffu;
	planes = rest == 1 ? (hi & ~lo) | (both << 1) : planes;
	planes = rest == 2 ? both : planes;
	planes = rest >= 3 ? 0u : planes;
	return (cw & ~0xfffffu) | ((uint32_t)(q_bits - sub) << 16) | planes;
}

__device__ __forceinline__ void lds_store_byte(uint32_t *buf, uint32_t byte_pos, uint32_t value)
{
	reinterpret_cast<uint8_t *>(buf)[byte_pos] = (uint8_t)value;
}

__global__ __launch_bounds__(kBlockLanes) void k_pack(const BandJob *__restrict__ bands,
                                                      const uint32_t *__restrict__ blk_band,
                                                      FrameState *__restrict__ frames,
                                                      XcdMap xm)
{
	__shared__ __align__(16) uint32_t s_buf[kMaxBlockWords];

	const BlockCoord C = locate_block(bands, blk_band, xm);
	const BandJob &J = C.J;
	FrameState &F = frames[J.frame];
	const int tid = threadIdx.x, j = tid & 3, k8 = tid >> 2;
	// The block's 16 word counts load with its quant level (one round trip
	// instead of two dependent ones).
	const uint4 *const w16 = reinterpret_cast<const uint4 *>(F.words + (long)C.b * 16);
	const uint4 wlo = w16[0], whi = w16[1];
	const int quant = __builtin_amdgcn_readfirstlane(F.quant[C.b]);
	const uint32_t wpairs[8] = {wlo.x, wlo.y, wlo.z, wlo.w, whi.x, whi.y, whi.z, whi.w};
	const uint32_t words_expected = (wpairs[quant >> 1] >> (16 * (quant & 1))) & 0xffffu;

	if (words_expected == 0)
	{
		if (tid == 0)
		{
			F.meta[2 * C.b] = 0;
			F.meta[2 * C.b + 1] = 0;
		}
		return;
	}

	// Two 16-byte stores per lane clear the block buffer.
	static_assert(kMaxBlockWords == 8 * kBlockLanes, "two uint4 of the block buffer per lane");
	reinterpret_cast<uint4 *>(s_buf)[tid] = make_uint4(0u, 0u, 0u, 0u);
	reinterpret_cast<uint4 *>(s_buf)[tid + kBlockLanes] = make_uint4(0u, 0u, 0u, 0u);
	// Start of this block's 256-block group: sum of the earlier groups.
	uint32_t gacc = 0;
	for (int i = tid; i < (C.b >> 8); i += kBlockLanes)
		gacc += F.group_sum[i];
	const uint32_t group_base = wave_sum_u32(gacc);

	const int x8 = C.bx * 4 + (k8 & 3), y8 = C.by * 4 + (k8 >> 2);
	const bool in_range = x8 < J.bx8 && y8 < J.by8;
	LaneQuant Q;
	quantize_lane(J, C.bx, C.by, k8, j, Q);
	const uint32_t cw0 = in_range ? code_word_q0(Q, j) : 0u;
	const uint32_t cw = quantize_code_word(cw0, quant);
	const bool active = (cw & 0xffffu) != 0;

	// Per subblock h (subblock 2j + h, code field at bit 2 (2j + h)): the
	// planes it had at quant 0 and the ones kept.  Bit planes go MSB first:
	// plane byte bit p is bit `bit` of magnitude p.  Magnitudes p and p+4
	// share a register (low / high half), so one plane of all eight is four
	// shift-and-mask steps and a fold.  The planes kept are bits
	// in_planes - 1 down to in_planes - out_planes, so a magnitude is
	// significant (its sign is sent) when it has a bit at or above
	// in_planes - out_planes, which min(n >> drop, 1) tests on both halves
	// at once.
	const uint32_t in_q = (cw0 >> 16) & 0xf, out_q = (cw >> 16) & 0xf;
	uint32_t in_planes[2], out_planes[2], sig[2], pair[2][4];
#pragma unroll
	for (int h = 0; h < 2; h++)
	{
		const int sb = 2 * j + h;
		in_planes[h] = ((cw0 >> (2 * sb)) & 3u) + in_q;
		out_planes[h] = active ? ((cw >> (2 * sb)) & 3u) + out_q : 0u;
		u16x2 np[4];
#pragma unroll
		for (int i = 0; i < 4; i++)
		{
			np[i] = u16x2{(unsigned short)Q.n[8 * h + i], (unsigned short)Q.n[8 * h + i + 4]};
			pair[h][i] = __builtin_bit_cast(uint32_t, np[i]);
		}
		sig[h] = 0;
		if (out_planes[h])
		{
			const unsigned short drop = (unsigned short)(in_planes[h] - out_planes[h]);
			const u16x2 drop2 = {drop, drop}, one2 = {1, 1};
			uint32_t t = 0;
#pragma unroll
			for (int i = 0; i < 4; i++)
				t |= __builtin_bit_cast(uint32_t, __builtin_elementwise_min(np[i] >> drop2, one2)) << i;
			sig[h] = (t & 0xfu) | ((t >> 12) & 0xf0u);
		}
	}
	const int nsig0 = __popc(sig[0]), nsig = nsig0 + __popc(sig[1]);
	const int lane_planes = (int)(out_planes[0] + out_planes[1]);
	const int cost8 = quad_sum_i(8 * lane_planes + nsig);

	// The coded-block mask: every lane of an 8x8 block has the same code
	// word, so the ballot holds a nibble 0xf or 0 per block; folding the
	// nibbles gives one bit per block.  The coded bits (cost + 24 per coded
	// 8x8) and one scan of (planes << 16 | significant count) over the lanes
	// (both sums stay below 2^16: 64 lanes x 28 planes, 1024 coefficients)
	// are wave-wide: the block is one wave.
	uint64_t act = __ballot(active) & 0x1111111111111111ull;
	act = (act | (act >> 3)) & 0x0303030303030303ull;
	act = (act | (act >> 6)) & 0x000f000f000f000full;
	act = (act | (act >> 12)) & 0x000000ff000000ffull;
	const uint32_t ballot = (uint32_t)((act | (act >> 24)) & 0xffffu);
	const uint32_t N = __popc(ballot);
	const uint32_t bits = wave_sum_u32(j == 0 && cost8 ? (uint32_t)cost8 + 24u : 0u);
	uint32_t wcount = (bits + 31) / 32;
	if (wcount)
		wcount += 2;
	const int scan_v = (lane_planes << 16) | nsig;
	const int scan_incl = wave_incl_scan(scan_v);
	const int packed_excl = scan_incl - scan_v;
	const int packed_total = __builtin_amdgcn_readlane(scan_incl, 63);
	const int plane_excl = packed_excl >> 16, sig_excl = packed_excl & 0xffff;
	const int planes_total = packed_total >> 16;
	__syncthreads();

	// Lanes own disjoint byte ranges of the block, so control words, scale
	// bytes and plane bytes are plain byte stores; sign bits share words and
	// go in with at most two atomic ORs per lane.
	const uint32_t plane_base = 8 + 3 * N;
	if (active && j == 0)
	{
		uint32_t jb = __popc(ballot & ((1u << k8) - 1u));
		lds_store_byte(s_buf, 8 + 2 * jb, cw & 0xffu);
		lds_store_byte(s_buf, 9 + 2 * jb, (cw >> 8) & 0xffu);
		lds_store_byte(s_buf, 8 + 2 * N + jb, (cw >> 16) & 0xffu);
	}
	uint32_t pos = plane_base + (uint32_t)plane_excl;
#pragma unroll
	for (int h = 0; h < 2; h++)
		for (uint32_t pi = 0; pi < out_planes[h]; pi++)
		{
			const int bit = (int)in_planes[h] - 1 - (int)pi;
			uint32_t c = 0;
#pragma unroll
			for (int i = 0; i < 4; i++)
				c |= ((pair[h][i] >> bit) & 0x00010001u) << i;
			lds_store_byte(s_buf, pos++, (c & 0xfu) | ((c >> 12) & 0xf0u));
		}
	if (nsig)
	{
		// The signs of the significant magnitudes, subblock 2j first, in
		// magnitude order: at most 16 consecutive bits.  Without branches:
		// the sign bit of every texel (a significant one is nonzero, so its
		// sign bit is t < 0), kept where significant and moved down to the
		// count of significant texels before it.
		const uint32_t sig16 = sig[0] | (sig[1] << 8);
		uint32_t neg = 0;
#pragma unroll
		for (int p = 0; p < 16; p++)
			neg |= (__float_as_uint(Q.t[p]) >> 31) << p;
		neg &= sig16;
		uint32_t sbits = 0;
#pragma unroll
		for (int p = 0; p < 16; p++)
			sbits |= ((neg >> p) & 1u) << __popc(sig16 & ((1u << p) - 1u));
		const uint32_t sbit = 8u * (plane_base + (uint32_t)planes_total) + (uint32_t)sig_excl;
		const uint64_t wide = (uint64_t)sbits << (sbit & 31);
		if ((uint32_t)wide)
			atomicOr(&s_buf[sbit >> 5], (uint32_t)wide);
		if ((uint32_t)(wide >> 32))
			atomicOr(&s_buf[(sbit >> 5) + 1], (uint32_t)(wide >> 32));
	}
	if (tid == 0)
	{
		uint32_t w0 = ballot | (wcount << 16) | ((F.sequence & 7u) << 28);
		int e = (int)((J.qcode >> 3) & 31) - quant;
		e = e > 0 ? e : 0;
		uint32_t qc = (J.qcode & ~(31u << 3)) | ((uint32_t)e << 3);
		uint32_t w1 = qc | ((uint32_t)C.b << 8);
		s_buf[0] = w0;
		s_buf[1] = w1;
	}
	__syncthreads();

	const uint32_t offset = group_base + F.excl[C.b];
	const bool fits = offset + wcount <= F.capacity_words && wcount <= (uint32_t)kMaxBlockWords;
	if (fits)
		for (uint32_t i = tid; i < wcount; i += kBlockLanes)
			F.bitstream[offset + i] = s_buf[i];
	if (tid == 0)
	{
		F.meta[2 * C.b] = offset;
		F.meta[2 * C.b + 1] = wcount;
		if (wcount != words_expected)
			atomicOr(&F.rc[kRcError], kErrWordMismatch);
		if (!fits)
			atomicOr(&F.rc[kRcError], kErrCapacity);
	}
}

// reference shaders/wavelet_dequant.comp:100-160 and 220-391.  The coded
// block (at most kMaxBlockWords words for any block an encoder can produce) is
// first copied to LDS with coalesced loads and parsed from there.  A malformed
// block that reads past its own length reads the following payload bytes, as
// the reference does; bytes past the payload read as zero.
//
// With nparams > 0 the call's fields of frames 0 .. nparams - 1 come in
// `block` (the k_set_params launch of a small decode folded into this one):
// the frame's decode inputs are read from there, and the wave of each
// frame's block 0 writes the fields into its FrameState for the transform
// launches that follow.
__global__ __launch_bounds__(kBlockLanes) void k_dequant(const BandJob *__restrict__ bands,
                                                         const uint32_t *__restrict__ blk_band,
                                                         FrameState *__restrict__ frames,
                                                         XcdMap xm, ParamBlock block, int nparams)
{
	__shared__ uint32_t s_pay[kMaxBlockWords];
	__shared__ uint32_t s_poff[16];
	__shared__ uint32_t s_ctrl[16];
	__shared__ uint32_t s_qb[16];
	__shared__ uint32_t s_sign_base;

	const BlockCoord C = locate_block(bands, blk_band, xm);
	const BandJob &J = C.J;
	const FrameState &F = frames[J.frame];
	const int tid = threadIdx.x, j = tid & 3, k8 = tid >> 2;
	const int x = C.bx * 32 + 8 * (k8 & 3) + 4 * (j >> 1);
	const int y = C.by * 32 + 8 * (k8 >> 2) + 4 * (j & 1);

	// The frame's decode inputs are read together, and a meta pair (offset,
	// words) as one 8-byte load: each dependent lookup costs a round trip.
	const bool folded = J.frame < nparams; // wave-uniform
	const CallParams &CP = block.p[folded ? J.frame : 0];
	if (folded && C.b == 0 && tid == 0)
		write_call_fields(frames[J.frame], CP);
	const uint32_t *const offsets = folded ? CP.dec_offsets : F.dec_offsets;
	const int offset_stride = folded ? CP.dec_offset_stride : F.dec_offset_stride;
	const uint32_t payload_words = (folded ? CP.dec_payload_bytes : F.dec_payload_bytes) >> 2;
	const uint32_t *const dec_payload = folded ? CP.dec_payload : F.dec_payload;
	uint32_t off;
	bool missing;
	if (offset_stride == 2)
	{
		const uint2 e = reinterpret_cast<const uint2 *>(offsets)[C.b];
		off = e.x;
		missing = off == 0xffffffffu || e.y == 0;
	}
	else
	{
		off = offsets[C.b];
		missing = off == 0xffffffffu;
	}
	if (!missing && off >= payload_words)
		missing = true;

	// Texel h * 8 + p: pixel p of subblock 2j + h (column p >> 1, row 2h + (p & 1)).
	float v[16];
#pragma unroll
	for (int p = 0; p < 16; p++)
		v[p] = 0.0f;
	if (!missing)
	{
		const uint32_t *payload = dec_payload + off;
		// The first kBlockLanes words (one per lane; most blocks are not much
		// longer) load together with the header word that gives the length,
		// so a typical block arrives in one or two round trips.
		const uint32_t first = (uint32_t)tid < payload_words - off ? payload[tid] : 0u;
		const uint32_t w0 = payload[0];
		uint32_t nwords = (w0 >> 16) & 0xfffu;
		nwords = nwords < (uint32_t)kMaxBlockWords ? nwords : (uint32_t)kMaxBlockWords;
		nwords = nwords < payload_words - off ? nwords : payload_words - off;
		if ((uint32_t)tid < nwords)
			s_pay[tid] = first;
		for (uint32_t i = tid + kBlockLanes; i < nwords; i += kBlockLanes)
			s_pay[i] = payload[i];
		__syncthreads();
		// Byte `pos` of the coded block.  A well-formed block reads only its
		// own nwords words, all in LDS; past them (only a malformed block
		// does) come the payload's following bytes, and zero past its end.
		// The two sources are separate loads, never one load through a
		// pointer chosen per lane (which would be a generic FLAT access).
		const uint32_t nbytes = nwords * 4, avail = (payload_words - off) * 4;
		auto byte_at = [&](uint32_t pos) -> uint32_t {
			const uint32_t w = s_pay[min(pos >> 2, (uint32_t)kMaxBlockWords - 1u)];
			uint32_t b = (w >> (8 * (pos & 3))) & 0xffu;
			if (pos >= nbytes)
				b = pos < avail ? (payload[pos >> 2] >> (8 * (pos & 3))) & 0xffu : 0u;
			return b;
		};
		const uint32_t ballot = w0 & 0xffffu;
		const uint32_t q_code = byte_at(4);
		const uint32_t N = __popc(ballot);
		{
			// Lanes 0-15 read 8x8 block `lane`'s control word and q / scale
			// byte and scan the plane bytes before it (row_shr DPP steps: an
			// inclusive scan within the 16-lane row; every lane takes part, the
			// other rows see no coded blocks and write nothing).
			uint32_t ctrl = 0, qb = 0;
			if (tid < 16 && ((ballot >> tid) & 1u))
			{
				uint32_t jb = __popc(ballot & ((1u << tid) - 1u));
				ctrl = byte_at(8 + 2 * jb) | (byte_at(9 + 2 * jb) << 8);
				qb = byte_at(8 + 2 * N + jb);
			}
			uint32_t lsbs = ctrl & 0x5555u, msbs = ctrl & 0xaaaau;
			msbs |= msbs >> 1;
			const int cost = __popc(lsbs) + __popc(msbs) + (int)(qb & 0xfu) * 8;
			int incl = cost;
			incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xF, 0xF, true);
			incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xF, 0xF, true);
			incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xF, 0xF, true);
			incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xF, 0xF, true);
			if (tid < 16)
			{
				s_poff[tid] = 8 + 3 * N + (uint32_t)(incl - cost);
				s_ctrl[tid] = ctrl;
				s_qb[tid] = qb;
			}
			if (tid == 15)
				s_sign_base = 8u * (8 + 3 * N + (uint32_t)incl);
		}
		__syncthreads();
		// Whether the plane bytes (everything before the sign bits) lie inside
		// the block's own words: true for every block an encoder writes, and
		// then they are read from LDS without a bounds check.
		const uint8_t *const pay8 = reinterpret_cast<const uint8_t *>(s_pay);
		const bool planes_inside = s_sign_base <= 8u * nbytes;

		int sig_count = 0;
		if ((ballot >> k8) & 1u)
		{
			const uint32_t ctrl = s_ctrl[k8], qb = s_qb[k8];
			const uint32_t q_bits = qb & 0xfu;
			const int bo = 4 * j; // code field of subblock 2j
			const uint32_t lmask = (1u << bo) - 1u;
			uint32_t lsbs = ctrl & 0x5555u, msbs = ctrl & 0xaaaau;
			msbs |= msbs >> 1;
			// Subblock 2j's planes start after those of the 8x8's earlier
			// subblocks; subblock 2j + 1's follow them.
			uint32_t byte_off = s_poff[k8] + (uint32_t)__popc(lsbs & lmask) + (uint32_t)__popc(msbs & lmask) +
			                    q_bits * (uint32_t)(2 * j);
			// reference shaders/wavelet_dequant.comp:168-175, 312-315.
			int ef = (int)(q_code >> 3);
			ef = ef > 24 ? 24 : ef;
			const int e = 4 - ef;
			const int m = (int)(q_code & 7u);
			const float qstep = (1.0f / (8.0f * 1024.0f * 1024.0f)) * (float)((8 + m) * (1 << (20 + e)));
			const float inv_scale = qstep * ((float)(qb >> 4) / 8.0f + 0.25f);
#pragma unroll
			for (int h = 0; h < 2; h++)
			{
				const int planes = (int)(q_bits + ((ctrl >> (bo + 2 * h)) & 3u));
				int mag[8] = {0, 0, 0, 0, 0, 0, 0, 0};
				if (planes <= 16)
				{
					// Magnitudes p and p+4 accumulate in the low / high half of
					// one register: a plane byte is spread once, then four
					// shifts.
					uint32_t pair[4] = {0, 0, 0, 0};
					auto spread = [&](uint32_t byte, int pl) {
						const uint32_t c = (byte & 0xfu) | ((byte & 0xf0u) << 12);
#pragma unroll
						for (int i = 0; i < 4; i++)
							pair[i] |= ((c >> i) & 0x00010001u) << pl;
					};
					if (planes_inside)
						for (int pl = planes - 1; pl >= 0; pl--)
							spread(pay8[byte_off++], pl);
					else
						for (int pl = planes - 1; pl >= 0; pl--)
							spread(byte_at(byte_off++), pl);
#pragma unroll
					for (int i = 0; i < 4; i++)
					{
						mag[i] = (int)(pair[i] & 0xffffu);
						mag[i + 4] = (int)(pair[i] >> 16);
					}
				}
				else
				{
					// Up to 18 planes (only in malformed blocks).
					for (int pl = planes - 1; pl >= 0; pl--)
					{
						uint32_t byte = byte_at(byte_off++);
#pragma unroll
						for (int p = 0; p < 8; p++)
							mag[p] |= (int)((byte >> p) & 1u) << pl;
					}
				}
#pragma unroll
				for (int p = 0; p < 8; p++)
				{
					float f = (float)mag[p];
					if (f != 0.0f)
					{
						f += 0.5f;
						sig_count++;
					}
					v[8 * h + p] = f * inv_scale;
				}
			}
		}
		const int sig_incl = wave_incl_scan(sig_count);
		// This lane's sign bits are consecutive (at most 16, subblock 2j's
		// first).  When all of the block's sign bits lie inside its own words
		// (every block an encoder writes), the two LDS words around them hold
		// them (bits read past the block's end are not used); else three
		// bytes through the bounds-checked reader.
		const uint32_t sbit = s_sign_base + (uint32_t)(sig_incl - sig_count);
		const uint32_t sign_end = s_sign_base + (uint32_t)__builtin_amdgcn_readlane(sig_incl, 63);
		uint32_t signs;
		if (sign_end <= 8u * nbytes)
		{
			const uint32_t w = sbit >> 5;
			const uint64_t two = (uint64_t)s_pay[w] | ((uint64_t)s_pay[min(w + 1, (uint32_t)kMaxBlockWords - 1u)] << 32);
			signs = (uint32_t)(two >> (sbit & 31));
		}
		else
		{
			const uint32_t sb = sbit >> 3;
			signs = (byte_at(sb) | (byte_at(sb + 1) << 8) | (byte_at(sb + 2) << 16)) >> (sbit & 7);
		}
		int n = 0;
#pragma unroll
		for (int p = 0; p < 16; p++)
		{
			if (v[p] != 0.0f)
			{
				if ((signs >> n) & 1u)
					v[p] = -v[p];
				n++;
			}
		}
	}
	// Each row of the 4x4 region goes out as one vector store (band storage
	// is padded to whole 32x32 blocks, so they are in bounds and aligned):
	// row r holds subblock r >> 1's pixels p with p & 1 == r & 1.
	const long i0 = (long)y * J.pitch + x;
	if (J.kind == kStoreF16)
	{
		__half *b16 = static_cast<__half *>(J.data);
		// Each value rounds to fp16 from its float (no merge with the
		// product that made it into one mixed-precision rounding).
#pragma unroll
		for (int p = 0; p < 16; p++)
			asm volatile("" : "+v"(v[p]));
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			const float *w = v + 8 * (r >> 1) + (r & 1);
			const uint32_t h0 = __half_as_ushort(__float2half_rn(w[0]));
			const uint32_t h1 = __half_as_ushort(__float2half_rn(w[2]));
			const uint32_t h2 = __half_as_ushort(__float2half_rn(w[4]));
			const uint32_t h3 = __half_as_ushort(__float2half_rn(w[6]));
			*reinterpret_cast<uint2 *>(b16 + i0 + r * J.pitch) = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
		}
	}
	else
	{
		float *b32 = static_cast<float *>(J.data);
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			const float *w = v + 8 * (r >> 1) + (r & 1);
			*reinterpret_cast<float4 *>(b32 + i0 + r * J.pitch) = make_float4(w[0], w[2], w[4], w[6]);
		}
	}
}
} // namespace

const void *set_params_kernel()
{
	return reinterpret_cast<const void *>(&k_set_params);
}

void launch_set_params(FrameState *frames, const CallParams *params, int nframes, int max_n32, hipStream_t stream)
{
	int blocks = 1;
	if (max_n32 > 0)
	{
		const int totals = totals_copies(max_n32) * kFlat;
		const int n = max_n32 > totals ? max_n32 : totals;
		blocks = (n + 255) / 256;
		blocks = blocks > 64 ? 64 : blocks;
	}
	for (int first = 0; first < nframes; first += kParamsPerLaunch)
	{
		ParamBlock block;
		const int count = nframes - first < kParamsPerLaunch ? nframes - first : kParamsPerLaunch;
		for (int i = 0; i < count; i++)
			block.p[i] = params[first + i];
		hipLaunchKernelGGL(k_set_params, dim3(blocks, count), dim3(256), 0, stream, frames, block, first, count);
	}
}

void launch_quant_analyze(const BandJob *bands, const uint32_t *blk_band, int nblocks, const FrameState *frames,
                          hipStream_t stream)
{
	if (nblocks > 0)
		hipLaunchKernelGGL(k_quant_analyze, dim3(nblocks), dim3(kBlockLanes), 0, stream, bands, blk_band, frames,
		                   xcd_map(dim3(nblocks)));
}


void launch_rc_finalize(FrameState *frames, int nframes, hipStream_t stream)
{
	hipLaunchKernelGGL(k_rc_finalize, dim3(nframes), dim3(512), 0, stream, frames);
}

void launch_rc_resolve(FrameState *frames, int nframes, hipStream_t stream)
{
	hipLaunchKernelGGL(k_rc_resolve, dim3(nframes), dim3(1024), 0, stream, frames);
}

void launch_rc_partial(FrameState *frames, int nframes, hipStream_t stream)
{
	hipLaunchKernelGGL(k_rc_partial, dim3(kFlat, nframes), dim3(256), 0, stream, frames);
}

void launch_rc_prepare(FrameState *frames, const uint32_t *grp_frame, int ngroups_total,
                       const int32_t *frame_group_base, hipStream_t stream)
{
	if (ngroups_total > 0)
		hipLaunchKernelGGL(k_rc_prepare, dim3(ngroups_total), dim3(256), 0, stream, frames, grp_frame,
		                   frame_group_base);
}

void launch_pack(const BandJob *bands, const uint32_t *blk_band, int nblocks, FrameState *frames,
                 hipStream_t stream)
{
	if (nblocks > 0)
		hipLaunchKernelGGL(k_pack, dim3(nblocks), dim3(kBlockLanes), 0, stream, bands, blk_band, frames,
		                   xcd_map(dim3(nblocks)));
}

void launch_dequant(const BandJob *bands, const uint32_t *blk_band, int nblocks, FrameState *frames,
                    hipStream_t stream, const CallParams *params, int nparams)
{
	if (nblocks <= 0)
		return;
	ParamBlock block;
	if (nparams < 0 || nparams > kParamsPerLaunch)
		nparams = 0;
	for (int i = 0; i < nparams; i++)
		block.p[i] = params[i];
	hipLaunchKernelGGL(k_dequant, dim3(nblocks), dim3(kBlockLanes), 0, stream, bands, blk_band, frames,
	                   xcd_map(dim3(nblocks)), block, nparams);
}
} // namespace pw
