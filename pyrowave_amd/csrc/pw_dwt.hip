// pw_dwt.hip -- forward and inverse CDF 9/7 DWT levels on gfx950.
//
// One 256-thread workgroup transforms one tile of a level; the tile and its
// 4-sample apron are staged in LDS, the two 1-D passes run over LDS rows and
// columns, and every lifting step is fmaf(c, l + r, v), the order of the
// reference shaders (reference shaders/dwt.comp:108-115,
// shaders/idwt.comp:95-108).  A window of n samples yields exact outputs at
// positions 4 .. n-5, so tiles reproduce the whole-line transform bit for bit.
//
// Forward: W x 56 domain samples per tile ((W + 8) x 64 staged), W = 128
// (256 threads) or 64 (128 threads).  Coalesced 4-sample vector loads, then
// each thread lifts one long window in registers (horizontal: 64 rows x W/32
// windows of 40 -> 32 outputs; vertical: W columns x 2 windows of 36 -> 28
// outputs), so window overlap costs 25-29 % extra arithmetic; band stores
// are coalesced rows.
//
// Inverse: T x T output samples per tile ((T + 8)^2 staged), T = 64 (256
// threads) or 32 (128 threads), 16-sample windows.  Small tiles keep LDS
// (21.6 KB at T = 64) and registers low, for seven workgroups per CU: the
// inverse stages four bands per tile and needs the extra waves in flight to
// cover its load latency.  Both shapes of either direction give the same
// coefficients bit for bit; the autotuner picks one per level.
#include "pw_common.hpp"

#include <hip/hip_fp16.h>

#include <type_traits>

#pragma clang fp contract(off)

namespace pw
{
namespace
{
// reference shaders/dwt_common.h:37-42
constexpr float kAlpha = float(-1.586134342059924);
constexpr float kBeta = float(-0.052980118572961);
constexpr float kGamma = float(0.882911075530934);
constexpr float kDelta = float(0.443506852043971);
constexpr float kK = float(1.230174104914001);
constexpr float kInvK = float(1.0 / 1.230174104914001);

constexpr int kFwdHWin = 40, kFwdHOut = 32; // horizontal window / outputs

// Forward tile of W x H outputs, (W + 8) x (H + 8) staged, 2W threads;
// vertical windows of H/2 + 8 samples (H/2 exact outputs).  LUT: convert u8
// samples through an LDS copy of the u/255 - 0.5 table (else arithmetically).
template <int W, int H, bool LUT>
struct FwdTile
{
	static constexpr int kW = W, kH = H;
	static constexpr int kThreads = 2 * W;
	static constexpr int kInW = W + 2 * kApron, kInH = H + 2 * kApron;
	static constexpr int kPitch = kInW + 1; // odd: row-per-lane reads are conflict-free
	static constexpr int kGroups = kInW / 4; // groups of 4 samples per row
	static constexpr int kSlots = kInH * kGroups;
	static constexpr int kSlotIters = (kSlots + kThreads - 1) / kThreads;
	static constexpr int kVOut = H / 2, kVWin = kVOut + 2 * kApron;
	static constexpr bool kLut = LUT;
	static_assert(kInH <= 64, "one staged row per lane in the horizontal pass");
	static_assert(kInH * (W / kFwdHOut) <= kThreads, "one horizontal window per thread");
	static_assert(W * 2 == kThreads && kVOut % 2 == 0, "two vertical windows per column");
};

// Inverse tile of TW x TH outputs, (TW + 8) x (TH + 8) staged, WIN-sample
// windows (WIN - 8 exact outputs each).  PRESCALE: each 1-D pass's K / 1/K
// scaling is applied once per sample (vertical at staging, horizontal on the
// vertical outputs) instead of once per window sample; the products are the
// same, so the results are too.
template <int TW, int TH, int THREADS, int WIN, bool PRESCALE>
struct InvTile
{
	static constexpr int kTW = TW, kTH = TH, kThreads = THREADS;
	static constexpr int kWin = WIN, kOut = WIN - 2 * kApron;
	static constexpr bool kPrescale = PRESCALE;
	static constexpr int kInW = TW + 2 * kApron, kInH = TH + 2 * kApron;
	static constexpr int kPitch = kInW + 1;
	static constexpr int kVSegs = TH / kOut;                 // vertical windows per column
	static constexpr int kHSegs = TW / kOut;                 // horizontal windows per row
	static constexpr int kPass1Tasks = kInW * kVSegs;
	static constexpr int kPass1Rounds = (kPass1Tasks + THREADS - 1) / THREADS;
	static constexpr int kPass2Tasks = TH * kHSegs;
	static constexpr int kPass2Rounds = (kPass2Tasks + THREADS - 1) / THREADS;
	static_assert(TW % kOut == 0 && TH % kOut == 0 && kOut % 8 == 0, "whole windows, 8-sample stores");
	static_assert(kInW + kInH <= THREADS, "one thread per extension-map entry");
};

__constant__ float c_u8_dc[256]; // u8 / 255 - 0.5 (UNORM8 texel with the level-0 DC shift)

__device__ __forceinline__ int mirror_repeat(int t, int R)
{
	if ((unsigned)t < (unsigned)R)
		return t;
	int period = 2 * R;
	int m = t % period;
	if (m < 0)
		m += period;
	return m < R ? m : period - 1 - m;
}

// Encoder extension: clamp-to-edge from R to A folded with whole-sample
// mirroring about 0 and A-1 (reference shaders/dwt.comp:27-36 with a 2x2
// gather footprint and a MIRRORED_REPEAT sampler).  The identity on [0, R).
__device__ __forceinline__ int fwd_ext(int p, int R, int A)
{
	// Inside the valid input (0 <= p <= R - 2; A >= R) the extension is the
	// identity: the common case, without the mirror arithmetic.
	if ((unsigned)p < (unsigned)(R - 1))
		return p;
	int c = p & ~1;
	int coord = c < 0 ? c : c + 1;
	if (coord >= 2 * A - R)
		coord = coord + 2 * (R - A) + 1;
	else if (coord > R)
		coord = R;
	return mirror_repeat(coord - 1 + (p & 1), R);
}

// Decoder extension: whole-sample symmetric about 0 and N-1 of the
// interleaved signal (reference shaders/idwt.comp:25-32).  Preserves parity.
__device__ __forceinline__ int inv_ext(int p, int N)
{
	while (p < 0 || p >= N)
	{
		if (p < 0)
			p = -p;
		if (p >= N)
			p = 2 * (N - 1) - p;
	}
	return p;
}

__device__ __forceinline__ void store_kind(void *base, int kind, long idx, float v)
{
	if (kind == kStoreF16)
	{
		// Rounded to fp16 from the float value (no merge with the product
		// that made it into one mixed-precision rounding).
		asm volatile("" : "+v"(v));
		static_cast<__half *>(base)[idx] = __float2half_rn(v);
	}
	else
		static_cast<float *>(base)[idx] = v;
}

// Forward lifting on an N-sample window starting at an even position;
// outputs at window positions 4 .. N-5 (low at even, high at odd).
template <int N>
__device__ __forceinline__ void lift_fwd(float (&v)[N])
{
#pragma unroll
	for (int i = 1; i <= N - 3; i += 2)
		v[i] = fmaf(kAlpha, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 2; i <= N - 4; i += 2)
		v[i] = fmaf(kBeta, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 3; i <= N - 5; i += 2)
		v[i] = fmaf(kGamma, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 4; i <= N - 6; i += 2)
		v[i] = fmaf(kDelta, v[i - 1] + v[i + 1], v[i]);
}

// Inverse lifting on an N-sample window starting at an even position, already
// scaled by K (even) / 1/K (odd); outputs at window positions 4 .. N-5.
template <int N>
__device__ __forceinline__ void lift_inv(float (&v)[N])
{
#pragma unroll
	for (int i = 2; i <= N - 2; i += 2)
		v[i] = fmaf(-kDelta, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 3; i <= N - 3; i += 2)
		v[i] = fmaf(-kGamma, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 4; i <= N - 4; i += 2)
		v[i] = fmaf(-kBeta, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 5; i <= N - 5; i += 2)
		v[i] = fmaf(-kAlpha, v[i - 1] + v[i + 1], v[i]);
}

// --- forward ----------------------------------------------------------------

// Four consecutive source samples as raw storage: u8 -> one dword, f16 ->
// two dwords, f32 -> four.
template <typename T>
struct Raw4;
template <>
struct Raw4<uint8_t>
{
	uint32_t w;
};
template <>
struct Raw4<__half>
{
	uint2 w;
};
template <>
struct Raw4<uint16_t>
{
	uint2 w;
};
template <>
struct Raw4<float>
{
	float4 w;
};

template <typename T>
__device__ __forceinline__ Raw4<T> load4_direct(const T *p)
{
	Raw4<T> r;
	if constexpr (sizeof(T) == 1)
		r.w = *reinterpret_cast<const uint32_t *>(p);
	else if constexpr (sizeof(T) == 2)
		r.w = *reinterpret_cast<const uint2 *>(p);
	else
		r.w = *reinterpret_cast<const float4 *>(p);
	return r;
}

template <typename T>
__device__ __forceinline__ Raw4<T> load4_gather(const T *row, const int *xmap, int c, uint32_t step)
{
	Raw4<T> r;
	if constexpr (sizeof(T) == 1)
	{
		r.w = 0;
#pragma unroll
		for (int i = 0; i < 4; i++)
			r.w |= (uint32_t)row[__umul24((uint32_t)xmap[c + i], step)] << (8 * i);
	}
	else if constexpr (std::is_same<T, uint16_t>::value)
	{
		uint32_t u[4];
#pragma unroll
		for (int i = 0; i < 4; i++)
			u[i] = row[__umul24((uint32_t)xmap[c + i], step)];
		r.w.x = u[0] | (u[1] << 16);
		r.w.y = u[2] | (u[3] << 16);
	}
	else if constexpr (sizeof(T) == 2)
	{
		__half h[4];
#pragma unroll
		for (int i = 0; i < 4; i++)
			h[i] = row[xmap[c + i]];
		r.w.x = (uint32_t)__half_as_ushort(h[0]) | ((uint32_t)__half_as_ushort(h[1]) << 16);
		r.w.y = (uint32_t)__half_as_ushort(h[2]) | ((uint32_t)__half_as_ushort(h[3]) << 16);
	}
	else
	{
		r.w.x = row[xmap[c + 0]];
		r.w.y = row[xmap[c + 1]];
		r.w.z = row[xmap[c + 2]];
		r.w.w = row[xmap[c + 3]];
	}
	return r;
}

// u / 255 - 0.5 without a table: fmaf(u, hi, u * lo) with hi + lo = 1/255
// split into two floats rounds to float(u / 255) for every u in 0..255
// (checked exhaustively against the correctly rounded quotient), then the
// DC shift; equal to the table entry.
__device__ __forceinline__ float u8_dc(uint32_t u)
{
	const float fu = (float)u;
	const float hi = __uint_as_float(0x3b808081u), lo = __uint_as_float(0xaf7efeffu);
	return fmaf(fu, hi, fu * lo) - 0.5f;
}

// u / 65535 - 0.5 the same way: hi + lo = 1/65535 split into two floats,
// fmaf(u, hi, u * lo) rounds to float(u / 65535) for every u in 0..65535
// (checked exhaustively with exact rationals), then the DC shift.
__device__ __forceinline__ float u16_unorm(uint32_t u)
{
	const float fu = (float)u;
	const float hi = __uint_as_float(0x37800080u), lo = __uint_as_float(0x27800080u);
	return fmaf(fu, hi, fu * lo);
}

__device__ __forceinline__ float u16_dc(uint32_t u)
{
	return u16_unorm(u) - 0.5f;
}

template <typename T>
__device__ __forceinline__ void unpack4(const Raw4<T> &r, const float *lut, float (&f)[4])
{
	if constexpr (sizeof(T) == 1)
	{
#pragma unroll
		for (int i = 0; i < 4; i++)
			f[i] = lut ? lut[(r.w >> (8 * i)) & 0xffu] : u8_dc((r.w >> (8 * i)) & 0xffu);
	}
	else if constexpr (std::is_same<T, uint16_t>::value)
	{
		f[0] = u16_dc(r.w.x & 0xffffu);
		f[1] = u16_dc(r.w.x >> 16);
		f[2] = u16_dc(r.w.y & 0xffffu);
		f[3] = u16_dc(r.w.y >> 16);
	}
	else if constexpr (sizeof(T) == 2)
	{
		f[0] = __half2float(__ushort_as_half((unsigned short)(r.w.x & 0xffffu)));
		f[1] = __half2float(__ushort_as_half((unsigned short)(r.w.x >> 16)));
		f[2] = __half2float(__ushort_as_half((unsigned short)(r.w.y & 0xffffu)));
		f[3] = __half2float(__ushort_as_half((unsigned short)(r.w.y >> 16)));
	}
	else
	{
		f[0] = r.w.x;
		f[1] = r.w.y;
		f[2] = r.w.z;
		f[3] = r.w.w;
	}
}

// Stages the (W + 8) x 64 source window in LDS.  Slot s covers row
// s / groups, samples 4 * (s % groups) .. +3.  A group whose four samples lie
// inside the valid source needs no extension and is one vector load; others
// gather through the extension maps.  Offsets are 32-bit element offsets from
// the uniform base (planes and bands stay below 4 GiB).  All loads are issued
// before the first LDS store.
template <typename S, typename T>
__device__ __forceinline__ void fwd_stage(const T *src, uint32_t pitch, uint32_t step, bool vec_ok, int x0, int R_x,
                                          const int *ymap, const int *xmap, const float *lut, int tid, float *sA)
{
	Raw4<T> raw[S::kSlotIters];
#pragma unroll
	for (int it = 0; it < S::kSlotIters; it++)
	{
		const int slot = tid + it * S::kThreads;
		if (it < S::kSlotIters - 1 || slot < S::kSlots)
		{
			const int r = slot / S::kGroups, k = slot - r * S::kGroups;
			const int gx = x0 - kApron + 4 * k;
			const uint32_t roff = __umul24((uint32_t)ymap[r], pitch);
			if (vec_ok && gx >= 0 && gx + 3 < R_x)
				raw[it] = load4_direct(src + (roff + (uint32_t)gx));
			else
				raw[it] = load4_gather(src + roff, xmap, 4 * k, step);
		}
	}
#pragma unroll
	for (int it = 0; it < S::kSlotIters; it++)
	{
		const int slot = tid + it * S::kThreads;
		if (it < S::kSlotIters - 1 || slot < S::kSlots)
		{
			const int r = slot / S::kGroups, k = slot - r * S::kGroups;
			float f[4];
			unpack4(raw[it], lut, f);
			float *d = sA + r * S::kPitch + 4 * k;
#pragma unroll
			for (int i = 0; i < 4; i++)
				d[i] = f[i];
		}
	}
}

template <int W, int H, bool LUT>
__global__ __launch_bounds__(2 * W) void k_dwt_forward(const DwtJob *__restrict__ jobs,
                                                       const uint16_t *__restrict__ tile_job,
                                                       const FrameState *__restrict__ frames, XcdMap xm)
{
	using S = FwdTile<W, H, LUT>;
	__shared__ float sA[S::kInH * S::kPitch];
	__shared__ int s_xmap[S::kInW];
	__shared__ int s_ymap[S::kInH];
	__shared__ float s_lut_mem[LUT ? 256 : 1];
	float *s_lut = LUT ? s_lut_mem : nullptr;

	const int tid = threadIdx.x;
	const int tile = xcd_swizzle((int)blockIdx.x, xm);
	// The job index is uniform: readfirstlane keeps J in scalar registers.
	const DwtJob J = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const int local = tile - J.tile_base;
	const int x0 = (local % J.tiles_x) * S::kW;
	const int y0 = (local / J.tiles_x) * S::kH;

	for (int i = tid; i < S::kInW + S::kInH; i += S::kThreads)
	{
		if (i < S::kInW)
			s_xmap[i] = fwd_ext(x0 - kApron + i, J.R_x, J.A_x);
		else
			s_ymap[i - S::kInW] = fwd_ext(y0 - kApron + i - S::kInW, J.R_y, J.A_y);
	}
	if constexpr (LUT)
		for (int i = tid; i < 256; i += S::kThreads)
			s_lut_mem[i] = c_u8_dc[i];
	__syncthreads();

	if (J.src_kind == kStoreU8)
	{
		const FrameState &F = frames[J.frame];
		const uint8_t *src = F.src[J.component];
		const uint32_t pitch = (uint32_t)F.src_pitch[J.component], step = (uint32_t)F.src_step[J.component];
		const bool vec_ok = step == 1 && ((reinterpret_cast<uintptr_t>(src) | (uintptr_t)pitch) & 3) == 0;
		fwd_stage<S>(src, pitch, step, vec_ok, x0, J.R_x, s_ymap, s_xmap, s_lut, tid, sA);
	}
	else if (J.src_kind == kStoreU16)
	{
		const FrameState &F = frames[J.frame];
		const uint16_t *src = reinterpret_cast<const uint16_t *>(F.src[J.component]);
		const uint32_t pitch = (uint32_t)F.src_pitch[J.component], step = (uint32_t)F.src_step[J.component];
		const bool vec_ok = step == 1 && ((reinterpret_cast<uintptr_t>(src) | (2u * (uintptr_t)pitch)) & 7) == 0;
		fwd_stage<S>(src, pitch, step, vec_ok, x0, J.R_x, s_ymap, s_xmap, s_lut, tid, sA);
	}
	else if (J.src_kind == kStoreF16)
		fwd_stage<S>(static_cast<const __half *>(J.src), (uint32_t)J.src_pitch, 1u, true, x0, J.R_x, s_ymap, s_xmap,
		             s_lut, tid, sA);
	else
		fwd_stage<S>(static_cast<const float *>(J.src), (uint32_t)J.src_pitch, 1u, true, x0, J.R_x, s_ymap, s_xmap,
		             s_lut, tid, sA);
	__syncthreads();

	// Horizontal pass: wave w takes window w (outputs 32w .. 32w+31) of the
	// staged rows, one row per lane (lanes past the last row idle).
	{
		const int row = tid & 63, seg = tid >> 6;
		const bool live = row < S::kInH;
		float v[kFwdHWin];
		if (live)
		{
#pragma unroll
			for (int j = 0; j < kFwdHWin; j++)
				v[j] = sA[row * S::kPitch + seg * kFwdHOut + j];
			lift_fwd(v);
		}
		__syncthreads();
		// Low outputs to columns 0..W/2-1, high outputs to W/2..W-1.
		if (live)
		{
#pragma unroll
			for (int k = 0; k < kFwdHOut / 2; k++)
			{
				sA[row * S::kPitch + seg * (kFwdHOut / 2) + k] = v[4 + 2 * k] * kInvK;
				sA[row * S::kPitch + W / 2 + seg * (kFwdHOut / 2) + k] = v[5 + 2 * k] * kK;
			}
		}
	}
	__syncthreads();

	// Vertical pass: lanes are consecutive columns; threads 0..W-1 take
	// output rows 0..H/2-1, threads W..2W-1 the rest; each store instruction
	// writes one band row.
	{
		const int col = tid & (W - 1), seg = tid / W;
		float v[S::kVWin];
#pragma unroll
		for (int j = 0; j < S::kVWin; j++)
			v[j] = sA[(seg * S::kVOut + j) * S::kPitch + col];
		lift_fwd(v);
		const bool low = col < W / 2;
		const int ox = (x0 >> 1) + (col & (W / 2 - 1));
		if (ox < J.bw)
		{
			void *lo = low ? J.band[0] : J.band[1];
			void *hi = low ? J.band[2] : J.band[3];
			const int lo_kind = low ? J.ll_kind : J.band_kind;
			const int oy0 = (y0 >> 1) + seg * (S::kVOut / 2);
#pragma unroll
			for (int k = 0; k < S::kVOut / 2; k++)
			{
				const int oy = oy0 + k;
				if (oy < J.bh)
				{
					const long idx = (long)oy * J.band_pitch + ox;
					store_kind(lo, lo_kind, idx, v[4 + 2 * k] * kInvK);
					store_kind(hi, J.band_kind, idx, v[5 + 2 * k] * kK);
				}
			}
		}
	}
}

// --- inverse ----------------------------------------------------------------

__device__ __forceinline__ uint8_t unorm8_store(float v)
{
	// +0.5 DC shift, clamp to [0,1], round to nearest (reference
	// shaders/idwt.comp:207-211 with an R8_UNORM store).
	v = __builtin_amdgcn_fmed3f(v + 0.5f, 0.0f, 1.0f);
	return (uint8_t)(int)fmaf(v, 255.0f, 0.5f);
}

__device__ __forceinline__ uint16_t unorm16_store(float v)
{
	// The same into an R16_UNORM target (reference decode.cpp:215).
	v = __builtin_amdgcn_fmed3f(v + 0.5f, 0.0f, 1.0f);
	return (uint16_t)(int)fmaf(v, 65535.0f, 0.5f);
}

// Output planes of the last inverse level: u8 (R8_UNORM) or u16 (R16_UNORM).
__device__ __forceinline__ bool plane_kind(int kind)
{
	return kind == kStoreU8 || kind == kStoreU16;
}

// Raw buffer access for the streamed kernels: a buffer resource over a
// band or plane base (scalar registers), a wave-uniform byte offset (the
// row) and a fixed per-lane byte offset (the column), so a load or store
// needs no address arithmetic on the vector unit.  The resource spans the
// whole address range (every offset is in bounds by construction; a load
// of the aligned word holding a sample may reach up to three bytes past a
// plane's last sample, inside the same page).
typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));
constexpr int kBufWord3 = 0x00020000; // raw buffer, 32-bit data format (gfx9 family)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_res(const void *p)
{
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, -1, kBufWord3);
}

// Two horizontally adjacent plane samples (px, px + 1) of the row that
// starts `row` samples into the plane.
template <int kind>
__device__ __forceinline__ void store_plane_pair_row(uint8_t *plane, uint32_t row, int px, int R_x, float e, float o)
{
	constexpr uint32_t sb = kind == kStoreU16 ? 2u : 1u;
	const __amdgpu_buffer_rsrc_t r = buf_res(plane);
	const uint32_t voff = (uint32_t)px * sb, soff = row * sb;
	if constexpr (kind == kStoreU16)
	{
		if (px < R_x)
			__builtin_amdgcn_raw_buffer_store_b16(unorm16_store(e), r, voff, soff, 0);
		if (px + 1 < R_x)
			__builtin_amdgcn_raw_buffer_store_b16(unorm16_store(o), r, voff, soff + 2u, 0);
	}
	else
	{
		if (px < R_x)
			__builtin_amdgcn_raw_buffer_store_b8(unorm8_store(e), r, voff, soff, 0);
		if (px + 1 < R_x)
			__builtin_amdgcn_raw_buffer_store_b8(unorm8_store(o), r, voff, soff + 1u, 0);
	}
}

// Two horizontally adjacent plane samples (px, px + 1) of row py.
__device__ __forceinline__ void store_plane_pair(uint8_t *plane, int kind, long dpitch, int py, int px, int R_x,
                                                 float e, float o)
{
	if (kind == kStoreU16)
	{
		uint16_t *d = reinterpret_cast<uint16_t *>(plane) + (long)py * dpitch + px;
		if (px < R_x)
			d[0] = unorm16_store(e);
		if (px + 1 < R_x)
			d[1] = unorm16_store(o);
	}
	else
	{
		uint8_t *d = plane + (long)py * dpitch + px;
		if (px < R_x)
			d[0] = unorm8_store(e);
		if (px + 1 < R_x)
			d[1] = unorm8_store(o);
	}
}

// Loads the four bands' (TW/2 + 4) x (TH/2 + 4) shares of the interleaved
// (TW + 8) x (TH + 8) tile
// (band b holds the positions whose (row, column) parity is (b >> 1, b & 1))
// into LDS; every global load is issued before the first LDS store.  TL / TH
// are the storage types of the LL and of the high bands.
template <typename S, typename TL, typename TH>
__device__ __forceinline__ void load_bands(const DwtJob &J, const int *ymap, const int *xmap, int tid, float *sA)
{
	constexpr int kQX = S::kInW / 2, kQY = S::kInH / 2;
	constexpr int kIters = (kQX * kQY + S::kThreads - 1) / S::kThreads;
	// 32-bit element offsets from the uniform band bases (bands stay below
	// 4 GiB), so loads use a scalar base and a 32-bit vector offset.
	const uint32_t pitch = (uint32_t)J.band_pitch;
	const TL *ll = static_cast<const TL *>(J.band[0]);
	const TH *hb[3] = {static_cast<const TH *>(J.band[1]), static_cast<const TH *>(J.band[2]),
	                   static_cast<const TH *>(J.band[3])};
	float vals[4][kIters];
#pragma unroll
	for (int k = 0; k < kIters; k++)
	{
		const int i = tid + k * S::kThreads;
#pragma unroll
		for (int b = 0; b < 4; b++)
			vals[b][k] = 0.0f;
		if (i < kQX * kQY)
		{
			const int rr = i / kQX, cc = i - rr * kQX;
			const uint32_t row0 = __umul24((uint32_t)(ymap[2 * rr] >> 1), pitch);
			const uint32_t row1 = __umul24((uint32_t)(ymap[2 * rr + 1] >> 1), pitch);
			const uint32_t col0 = (uint32_t)(xmap[2 * cc] >> 1), col1 = (uint32_t)(xmap[2 * cc + 1] >> 1);
			if constexpr (sizeof(TL) == 2)
				vals[0][k] = __half2float(ll[row0 + col0]);
			else
				vals[0][k] = ll[row0 + col0];
			if constexpr (sizeof(TH) == 2)
			{
				vals[1][k] = __half2float(hb[0][row0 + col1]);
				vals[2][k] = __half2float(hb[1][row1 + col0]);
				vals[3][k] = __half2float(hb[2][row1 + col1]);
			}
			else
			{
				vals[1][k] = hb[0][row0 + col1];
				vals[2][k] = hb[1][row1 + col0];
				vals[3][k] = hb[2][row1 + col1];
			}
		}
	}
#pragma unroll
	for (int k = 0; k < kIters; k++)
	{
		const int i = tid + k * S::kThreads;
		if (i < kQX * kQY)
		{
			const int rr = i / kQX, cc = i - rr * kQX;
#pragma unroll
			for (int b = 0; b < 4; b++)
			{
				// With PRESCALE, the vertical pass's scaling once per staged
				// sample: even rows (LL, HL) by K, odd rows (LH, HH) by 1/K.
				const float x = S::kPrescale ? vals[b][k] * ((b >> 1) ? kInvK : kK) : vals[b][k];
				sA[(2 * rr + (b >> 1)) * S::kPitch + 2 * cc + (b & 1)] = x;
			}
		}
	}
}

template <typename S>
__device__ __forceinline__ void store_u16_run(uint16_t *dst, int n_valid, const float *v)
{
	// v[0 .. kOut) to dst, 8-byte stores where aligned and whole.
#pragma unroll
	for (int c = 0; c < S::kOut / 4; c++)
	{
		uint16_t *d = dst + 4 * c;
		if (4 * c + 4 <= n_valid && ((uintptr_t)d & 7) == 0)
		{
			const uint32_t lo = (uint32_t)unorm16_store(v[4 * c]) | ((uint32_t)unorm16_store(v[4 * c + 1]) << 16);
			const uint32_t hi = (uint32_t)unorm16_store(v[4 * c + 2]) | ((uint32_t)unorm16_store(v[4 * c + 3]) << 16);
			*reinterpret_cast<uint2 *>(d) = make_uint2(lo, hi);
		}
		else
		{
#pragma unroll
			for (int j = 0; j < 4; j++)
				if (4 * c + j < n_valid)
					d[j] = unorm16_store(v[4 * c + j]);
		}
	}
}

template <typename S>
__device__ __forceinline__ void store_u8_run(uint8_t *dst, int n_valid, const float *v)
{
	// v[0 .. kOut) to dst, 8-byte stores where aligned and whole.
#pragma unroll
	for (int c = 0; c < S::kOut / 8; c++)
	{
		uint8_t *d = dst + 8 * c;
		if (8 * c + 8 <= n_valid && ((uintptr_t)d & 7) == 0)
		{
			uint32_t lo = 0, hi = 0;
#pragma unroll
			for (int j = 0; j < 4; j++)
			{
				lo |= (uint32_t)unorm8_store(v[8 * c + j]) << (8 * j);
				hi |= (uint32_t)unorm8_store(v[8 * c + 4 + j]) << (8 * j);
			}
			*reinterpret_cast<uint2 *>(d) = make_uint2(lo, hi);
		}
		else
		{
#pragma unroll
			for (int j = 0; j < 8; j++)
				if (8 * c + j < n_valid)
					d[j] = unorm8_store(v[8 * c + j]);
		}
	}
}

template <int TW, int TH, int THREADS, int WIN, bool PRESCALE>
__global__ __launch_bounds__(THREADS) void k_dwt_inverse(const DwtJob *__restrict__ jobs,
                                                         const uint16_t *__restrict__ tile_job,
                                                         const FrameState *__restrict__ frames, XcdMap xm)
{
	using S = InvTile<TW, TH, THREADS, WIN, PRESCALE>;
	constexpr int kOut = S::kOut;
	__shared__ float sA[S::kInH * S::kPitch];
	__shared__ int s_xmap[S::kInW];
	__shared__ int s_ymap[S::kInH];

	const int tid = threadIdx.x;
	const int tile = xcd_swizzle((int)blockIdx.x, xm);
	// The job index is uniform: readfirstlane keeps J in scalar registers.
	const DwtJob J = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const int local = tile - J.tile_base;
	const int ox0 = (local % J.tiles_x) * TW;
	const int oy0 = (local / J.tiles_x) * TH;

	if (tid < S::kInW)
		s_xmap[tid] = inv_ext(ox0 - kApron + tid, J.A_x);
	else if (tid < S::kInW + S::kInH)
		s_ymap[tid - S::kInW] = inv_ext(oy0 - kApron + tid - S::kInW, J.A_y);
	__syncthreads();

	// Interleaved position parity selects the band: LL (even, even), HL (x
	// odd), LH (y odd), HH.  The extension preserves parity.
	if (J.ll_kind == kStoreF16)
		load_bands<S, __half, __half>(J, s_ymap, s_xmap, tid, sA);
	else if (J.band_kind == kStoreF16)
		load_bands<S, float, __half>(J, s_ymap, s_xmap, tid, sA);
	else
		load_bands<S, float, float>(J, s_ymap, s_xmap, tid, sA);
	__syncthreads();

	// Vertical pass over all TW + 8 columns; window seg covers staged rows
	// seg * kOut .. + WIN - 1, its outputs are tile rows seg * kOut .. + kOut - 1.
	float vout[S::kPass1Rounds][kOut];
#pragma unroll
	for (int t = 0; t < S::kPass1Rounds; t++)
	{
		const int task = tid + t * THREADS;
		if (t < S::kPass1Rounds - 1 || task < S::kPass1Tasks)
		{
			const int c = task % S::kInW, seg = task / S::kInW;
			float v[WIN];
#pragma unroll
			for (int j = 0; j < WIN; j++)
			{
				const float x = sA[(seg * kOut + j) * S::kPitch + c];
				v[j] = PRESCALE ? x : x * ((j & 1) ? kInvK : kK);
			}
			lift_inv(v);
			// With PRESCALE, the horizontal pass's scaling: even columns K,
			// odd 1/K (tile columns start at an even position).
			const float sc = (c & 1) ? kInvK : kK;
#pragma unroll
			for (int j = 0; j < kOut; j++)
				vout[t][j] = PRESCALE ? v[4 + j] * sc : v[4 + j];
		}
	}
	__syncthreads();
#pragma unroll
	for (int t = 0; t < S::kPass1Rounds; t++)
	{
		const int task = tid + t * THREADS;
		if (t < S::kPass1Rounds - 1 || task < S::kPass1Tasks)
		{
			const int c = task % S::kInW, seg = task / S::kInW;
#pragma unroll
			for (int j = 0; j < kOut; j++)
				sA[(seg * kOut + j) * S::kPitch + c] = vout[t][j];
		}
	}
	__syncthreads();

	// Horizontal pass: TW / kOut lanes per row write TW consecutive outputs.
	uint8_t *plane = nullptr;
	long dpitch = J.dst_pitch;
	if (plane_kind(J.dst_kind))
	{
		plane = frames[J.frame].dst[J.component];
		dpitch = frames[J.frame].dst_pitch[J.component];
	}
#pragma unroll
	for (int t = 0; t < S::kPass2Rounds; t++)
	{
		const int task = tid + t * THREADS;
		if (t == S::kPass2Rounds - 1 && task >= S::kPass2Tasks)
			break;
		const int seg = task % S::kHSegs, row = task / S::kHSegs;
		float v[WIN];
#pragma unroll
		for (int j = 0; j < WIN; j++)
		{
			const float x = sA[row * S::kPitch + seg * kOut + j];
			v[j] = PRESCALE ? x : x * ((j & 1) ? kInvK : kK);
		}
		lift_inv(v);
		const int py = oy0 + row;
		const int px0 = ox0 + seg * kOut;
		if (J.dst_kind == kStoreU8)
		{
			if (py < J.R_y)
				store_u8_run<S>(plane + (long)py * dpitch + px0, J.R_x - px0, v + 4);
		}
		else if (J.dst_kind == kStoreU16)
		{
			if (py < J.R_y)
				store_u16_run<S>(reinterpret_cast<uint16_t *>(plane) + (long)py * dpitch + px0, J.R_x - px0, v + 4);
		}
		else if (py < J.A_y)
		{
			const long idx = (long)py * dpitch + px0;
			if (J.dst_kind == kStoreF32 && px0 + kOut <= J.A_x)
			{
				// 32-byte aligned: px0 is a multiple of 8, the pitch of 32.
				float4 *d = reinterpret_cast<float4 *>(static_cast<float *>(J.dst) + idx);
#pragma unroll
				for (int q = 0; q < kOut / 4; q++)
					d[q] = make_float4(v[4 + 4 * q], v[5 + 4 * q], v[6 + 4 * q], v[7 + 4 * q]);
			}
			else
			{
#pragma unroll
				for (int j = 0; j < kOut; j++)
				{
					if (px0 + j < J.A_x)
						store_kind(J.dst, J.dst_kind, idx + j, v[4 + j]);
				}
			}
		}
	}
}

// --- forward, streamed (shape 4) ------------------------------------------
//
// One wave per tile of kFwdVariantTileW x kFwdVariantTileH[kFwdStreamVariant]
// (120 x 32) input samples and no LDS.
// Lane l holds the even / odd input pair at columns x0 - 4 + 2l and
// x0 - 3 + 2l; lanes 2 .. 61 produce the tile's outputs, lanes 0-1 and 62-63
// are the apron.  The wave walks down the 32 + 8 rows of the tile's window
// two rows at a time: each row is lifted horizontally across the lanes
// (the neighbouring lane's sample comes from a wave-wide DPP shift), and the
// two columns each lane then holds (horizontal low and high) are lifted
// vertically by a delay line in registers, which completes row pair p when
// pair p + 2 arrives.  Every sample sees the operations of lift_fwd in the
// same order, so the coefficients equal those of the LDS-tiled shapes.
constexpr int kStreamWaves = 4; // independent waves (tiles) per workgroup
static_assert(kFwdVariantTileW[kFwdStreamVariant] == 2 * 64 - 2 * kApron, "one input pair per lane");
static_assert(kFwdVariantTileH[kFwdStreamVariant] % 2 == 0, "whole row pairs");

constexpr int kDppWaveShl1 = 0x130; // lane i reads lane i + 1
constexpr int kDppWaveShr1 = 0x138; // lane i reads lane i - 1

template <int Ctrl>
__device__ __forceinline__ float dpp_wave(float v)
{
	return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xF, 0xF, false));
}

// Horizontal lifting of one row (e, o: this lane's even and odd sample), then
// the low / high scaling.
__device__ __forceinline__ void hlift_row(float &e, float &o)
{
	o = fmaf(kAlpha, e + dpp_wave<kDppWaveShl1>(e), o);
	e = fmaf(kBeta, dpp_wave<kDppWaveShr1>(o) + o, e);
	o = fmaf(kGamma, e + dpp_wave<kDppWaveShl1>(e), o);
	e = fmaf(kDelta, dpp_wave<kDppWaveShr1>(o) + o, e);
	e = e * kInvK;
	o = o * kK;
}

// Vertical lifting of one column as a delay line.  Before pair n arrives, re
// and ro hold pair n-1 as input, po and pe pair n-2 after the first two
// lifting steps (o'(n-2), e'(n-2)), and qo the odd sample of pair n-3 after
// the third (o''(n-3)).
struct VLine
{
	float re = 0.0f, ro = 0.0f, po = 0.0f, pe = 0.0f, qo = 0.0f;

	// Takes pair n (e, o); gives the finished pair n-2 scaled.
	__device__ __forceinline__ void push(float e, float o, float &low, float &high)
	{
		const float o1 = fmaf(kAlpha, re + e, ro); // o'(n-1)
		const float e1 = fmaf(kBeta, po + o1, re); // e'(n-1)
		const float o2 = fmaf(kGamma, pe + e1, po); // o''(n-2)
		const float e2 = fmaf(kDelta, qo + o2, pe); // e''(n-2)
		low = e2 * kInvK;
		high = o2 * kK;
		re = e;
		ro = o;
		po = o1;
		pe = e1;
		qo = o2;
	}
};

// Raw loads through global (not generic) pointers: a generic load also
// counts on the LDS counter, and any wait for the oldest prefetched row would
// then wait for every later one too.
typedef const __attribute__((address_space(1))) uint32_t *GU32;

// A streamed source: its address split into the aligned word address and the
// byte offset within that word (u8 planes come from the caller and may start
// at any byte, an NV12 Cr plane at data + 1; fp16 / fp32 sources are band
// storage, 256-byte aligned).
struct StreamSrc
{
	GU32 words;
	uint32_t lo;
};

__device__ __forceinline__ StreamSrc stream_src(const void *p)
{
	const uintptr_t a = reinterpret_cast<uintptr_t>(p);
	return StreamSrc{reinterpret_cast<GU32>(a & ~uintptr_t(3)), uint32_t(a & 3u)};
}

template <int KIND>
constexpr uint32_t stream_esize()
{
	return KIND == kStoreU8 ? 1u : (KIND == kStoreF16 || KIND == kStoreU16) ? 2u : 4u;
}

// Word `w` of the row that starts `row_w` words into a word-aligned source.
__device__ __forceinline__ uint32_t load_row_word(GU32 base, uint32_t row_w, uint32_t w)
{
	return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(buf_res((const void *)base), w << 2, row_w << 2, 0);
}

// Every format is fetched as the aligned 32-bit word holding the sample (a
// byte or half is extracted when it is converted): a narrower load's
// zero-extension would be placed right after it by the compiler and wait for
// the data.  An aligned word never crosses a page, so it stays inside the
// allocation that holds the sample.
template <int KIND>
__device__ __forceinline__ uint32_t stream_raw(const StreamSrc &S, uint32_t off)
{
	return S.words[(S.lo + off * stream_esize<KIND>()) >> 2];
}

// The sample at element offset `off` out of its word, as a float (u8 with
// the DC shift).
template <int KIND>
__device__ __forceinline__ float stream_value(uint32_t raw, const StreamSrc &S, uint32_t off)
{
	const uint32_t a = S.lo + off * stream_esize<KIND>();
	if constexpr (KIND == kStoreU8)
		return u8_dc((raw >> (8 * (a & 3))) & 0xffu);
	else if constexpr (KIND == kStoreU16)
		return u16_dc((raw >> (8 * (a & 2))) & 0xffffu);
	else if constexpr (KIND == kStoreF16)
		return __half2float(__ushort_as_half((unsigned short)(raw >> (8 * (a & 2)))));
	else
		return __uint_as_float(raw);
}

// The same with the sample's bit offset within its word known.
template <int KIND>
__device__ __forceinline__ float stream_value_at(uint32_t raw, uint32_t shift)
{
	if constexpr (KIND == kStoreU8)
		return u8_dc((raw >> shift) & 0xffu);
	else if constexpr (KIND == kStoreU16)
		return u16_dc((raw >> shift) & 0xffffu);
	else if constexpr (KIND == kStoreF16)
		return __half2float(__ushort_as_half((unsigned short)(raw >> shift)));
	else
		return __uint_as_float(raw);
}

// Element `col` of the row that starts `row` elements into a band: the row
// address is formed once per wave (scalar registers) and the lane's byte
// offset is a fixed 32-bit vector offset, so a store costs no vector
// address arithmetic.
template <int kind>
__device__ __forceinline__ void stream_store_row(void *base, uint32_t row, uint32_t col, float v)
{
	constexpr uint32_t es = kind == kStoreF16 ? 2u : 4u;
	const __amdgpu_buffer_rsrc_t r = buf_res(base);
	if constexpr (kind == kStoreF16)
	{
		// The barrier keeps the compiler from merging the product that made
		// v with this conversion (v_fma_mix rounds once, straight to fp16).
		asm volatile("" : "+v"(v));
		__builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half_rn(v)), r, col * es, row * es, 0);
	}
	else
		__builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, col * es, row * es, 0);
}

// Elements col and col + 1 (col even) of a band row as one store: fp32 pairs
// are 8-byte and fp16 pairs 4-byte aligned (band pitches are multiples of 32).
template <int kind>
__device__ __forceinline__ void stream_store_pair_row(void *base, uint32_t row, uint32_t col, float e, float o)
{
	constexpr uint32_t es = kind == kStoreF16 ? 2u : 4u;
	const __amdgpu_buffer_rsrc_t r = buf_res(base);
	if constexpr (kind == kStoreF16)
	{
		asm volatile("" : "+v"(e), "+v"(o)); // two roundings, as in stream_store_row
		const uint32_t packed =
		    (uint32_t)__half_as_ushort(__float2half_rn(e)) | ((uint32_t)__half_as_ushort(__float2half_rn(o)) << 16);
		__builtin_amdgcn_raw_buffer_store_b32(packed, r, col * es, row * es, 0);
	}
	else
	{
		const v2u32 pair = {__float_as_uint(e), __float_as_uint(o)};
		__builtin_amdgcn_raw_buffer_store_b64(pair, r, col * es, row * es, 0);
	}
}

__device__ __forceinline__ void stream_store(void *base, int kind, uint32_t idx, float v)
{
	if (kind == kStoreF16)
	{
		asm volatile("" : "+v"(v)); // as in store_kind
		((__attribute__((address_space(1))) uint16_t *)base)[idx] = __half_as_ushort(__float2half_rn(v));
	}
	else
		((__attribute__((address_space(1))) float *)base)[idx] = v;
}

template <int KIND, bool WORD_ROWS, int LLK, int HK>
__device__ __forceinline__ void forward_stream_tile(const DwtJob &J, const void *src_ptr, uint32_t pitch,
                                                    uint32_t step, int x0, int y0, int lane);

// The tile with the job's output storage kinds as template arguments (LL
// FP32 with fp16 or FP32 high bands, or all four fp16), so no store in the
// loop branches on them.
template <int KIND, bool WORD_ROWS>
__device__ __forceinline__ void forward_stream_kinds(const DwtJob &J, const void *src_ptr, uint32_t pitch,
                                                     uint32_t step, int x0, int y0, int lane)
{
	if (J.ll_kind == kStoreF16)
		forward_stream_tile<KIND, WORD_ROWS, kStoreF16, kStoreF16>(J, src_ptr, pitch, step, x0, y0, lane);
	else if (J.band_kind == kStoreF16)
		forward_stream_tile<KIND, WORD_ROWS, kStoreF32, kStoreF16>(J, src_ptr, pitch, step, x0, y0, lane);
	else
		forward_stream_tile<KIND, WORD_ROWS, kStoreF32, kStoreF32>(J, src_ptr, pitch, step, x0, y0, lane);
}

// One tile of the streamed forward transform with source samples of format
// KIND (fixed per wave, so the loads need no branches).  WORD_ROWS: every
// source row starts on a 32-bit word (row pitch in bytes a multiple of 4:
// band storage always, planes of such pitches), so a lane's word offset and
// bit shift within a row are the same in every row, and a row's loads are a
// scalar row base plus that fixed per-lane offset (no address arithmetic per
// row on the vector unit).
template <int KIND, bool WORD_ROWS, int LLK, int HK>
__device__ __forceinline__ void forward_stream_tile(const DwtJob &J, const void *src_ptr, uint32_t pitch,
                                                    uint32_t step, int x0, int y0, int lane)
{
	const StreamSrc src = stream_src(src_ptr);
	constexpr int TH = kFwdVariantTileH[kFwdStreamVariant];
	constexpr int kPairs = TH / 2 + kApron;
	static_assert(kPairs % 2 == 0, "the row-pair loop is unrolled by two");
	const uint32_t xe = __umul24((uint32_t)fwd_ext(x0 - kApron + 2 * lane, J.R_x, J.A_x), step);
	const uint32_t xo = __umul24((uint32_t)fwd_ext(x0 - kApron + 2 * lane + 1, J.R_x, J.A_x), step);
	const int ox = (x0 >> 1) - kApron / 2 + lane;
	const bool store_x = lane >= kApron / 2 && lane < 64 - kApron / 2 && ox < J.bw;

	// A row pair is addressed by its two row bases (uniform: scalar registers)
	// plus this lane's even / odd column offsets; its raw words are, in order,
	// even and odd column of the even row, then of the odd row.
	// Row offsets of the window's 2 * kPairs rows, one row per lane, computed
	// once (the extension's modulo would otherwise run per row pair); a row
	// pair reads its two entries with readlane.
	static_assert(2 * kPairs <= 64, "one window row per lane");
	constexpr uint32_t kEs = stream_esize<KIND>();
	// Row offsets: elements, or words for WORD_ROWS (exact: the row's byte
	// offset is a multiple of 4).
	const uint32_t row_el =
	    lane < 2 * kPairs ? __umul24((uint32_t)fwd_ext(y0 - kApron + lane, J.R_y, J.A_y), pitch) : 0u;
	const uint32_t row_off = WORD_ROWS ? (row_el * kEs) >> 2 : row_el;
	// WORD_ROWS: the even / odd sample's word within a row and bit shift.
	const uint32_t be = src.lo + xe * kEs, bo = src.lo + xo * kEs;
	const uint32_t we = be >> 2, wo = bo >> 2, se = 8u * (be & 3u), so = 8u * (bo & 3u);
	auto rows = [&](int n, uint32_t &r0, uint32_t &r1) {
		r0 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n);
		r1 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n + 1);
	};
	auto load_pair = [&](uint32_t r0, uint32_t r1, uint32_t (&r)[4]) {
		if constexpr (WORD_ROWS)
		{
			r[0] = load_row_word(src.words, r0, we);
			r[1] = load_row_word(src.words, r0, wo);
			r[2] = load_row_word(src.words, r1, we);
			r[3] = load_row_word(src.words, r1, wo);
		}
		else
		{
			r[0] = stream_raw<KIND>(src, r0 + xe);
			r[1] = stream_raw<KIND>(src, r0 + xo);
			r[2] = stream_raw<KIND>(src, r1 + xe);
			r[3] = stream_raw<KIND>(src, r1 + xo);
		}
	};
	auto values = [&](uint32_t r0, uint32_t r1, const uint32_t (&r)[4], float (&v)[4]) {
		if constexpr (WORD_ROWS)
		{
			v[0] = stream_value_at<KIND>(r[0], se);
			v[1] = stream_value_at<KIND>(r[1], so);
			v[2] = stream_value_at<KIND>(r[2], se);
			v[3] = stream_value_at<KIND>(r[3], so);
		}
		else
		{
			v[0] = stream_value<KIND>(r[0], src, r0 + xe);
			v[1] = stream_value<KIND>(r[1], src, r0 + xo);
			v[2] = stream_value<KIND>(r[2], src, r1 + xe);
			v[3] = stream_value<KIND>(r[3], src, r1 + xo);
		}
	};
	VLine lo, hi; // columns of horizontal low / high outputs
	auto lift_pair = [&](int n, float e0, float o0, float e1, float o1) {
		hlift_row(e0, o0);
		hlift_row(e1, o1);
		float ll, lh, hl, hh;
		lo.push(e0, e1, ll, lh);
		hi.push(o0, o1, hl, hh);
		const int oy = (y0 >> 1) + n - kApron;
		if (n >= kApron && store_x && oy < J.bh)
		{
			const uint32_t row = (uint32_t)oy * (uint32_t)J.band_pitch;
			stream_store_row<LLK>(J.band[0], row, (uint32_t)ox, ll);
			stream_store_row<HK>(J.band[1], row, (uint32_t)ox, hl);
			stream_store_row<HK>(J.band[2], row, (uint32_t)ox, lh);
			stream_store_row<HK>(J.band[3], row, (uint32_t)ox, hh);
		}
	};
	// Both buffers are converted (their loads were issued an iteration ago),
	// refilled with the next two pairs, then lifted and stored: the loads of
	// one iteration are in flight during the whole of the previous one's
	// lifting, and the only wait is for loads that old.
	uint32_t a0, a1, b0, b1, ra[4], rb[4];
	rows(0, a0, a1);
	rows(1, b0, b1);
	load_pair(a0, a1, ra);
	load_pair(b0, b1, rb);
#pragma unroll 2
	for (int n = 0; n < kPairs; n += 2)
	{
		float a[4], b[4];
		values(a0, a1, ra, a);
		values(b0, b1, rb, b);
		if (n + 2 < kPairs)
		{
			rows(n + 2, a0, a1);
			rows(n + 3, b0, b1);
			load_pair(a0, a1, ra);
			load_pair(b0, b1, rb);
		}
		// The scheduler would otherwise sink the loads towards their use.
		__builtin_amdgcn_sched_barrier(0);
		lift_pair(n, a[0], a[1], a[2], a[3]);
		lift_pair(n + 1, b[0], b[1], b[2], b[3]);
	}
}

__global__ __launch_bounds__(64 * kStreamWaves) void k_dwt_forward_stream(const DwtJob *__restrict__ jobs,
                                                                          const uint16_t *__restrict__ tile_job,
                                                                          const FrameState *__restrict__ frames,
                                                                          int tiles, XcdMap xm)
{
	constexpr int TW = kFwdVariantTileW[kFwdStreamVariant], TH = kFwdVariantTileH[kFwdStreamVariant];
	const int lane = threadIdx.x & 63;
	// Wave-uniform (readfirstlane tells the compiler so: the tile origin and
	// everything derived from it stay in scalar registers).
	const int tile = __builtin_amdgcn_readfirstlane(xcd_swizzle((int)blockIdx.x, xm) *
	                                                    kStreamWaves +
	                                                (int)(threadIdx.x >> 6));
	if (tile >= tiles)
		return; // whole waves: no wave-level or workgroup barrier follows
	const DwtJob J = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const int local = tile - J.tile_base;
	const int x0 = (local % J.tiles_x) * TW;
	const int y0 = (local / J.tiles_x) * TH;
	// Band storage pitches are multiples of 32 elements: word rows always.
	if (J.src_kind == kStoreU8)
	{
		const FrameState &F = frames[J.frame];
		const uint32_t pitch = (uint32_t)F.src_pitch[J.component], step = (uint32_t)F.src_step[J.component];
		if ((pitch & 3u) == 0)
			forward_stream_kinds<kStoreU8, true>(J, F.src[J.component], pitch, step, x0, y0, lane);
		else
			forward_stream_kinds<kStoreU8, false>(J, F.src[J.component], pitch, step, x0, y0, lane);
	}
	else if (J.src_kind == kStoreU16)
	{
		const FrameState &F = frames[J.frame];
		const uint32_t pitch = (uint32_t)F.src_pitch[J.component], step = (uint32_t)F.src_step[J.component];
		if ((pitch & 1u) == 0)
			forward_stream_kinds<kStoreU16, true>(J, F.src[J.component], pitch, step, x0, y0, lane);
		else
			forward_stream_kinds<kStoreU16, false>(J, F.src[J.component], pitch, step, x0, y0, lane);
	}
	else if (J.src_kind == kStoreF16)
		forward_stream_kinds<kStoreF16, true>(J, J.src, (uint32_t)J.src_pitch, 1u, x0, y0, lane);
	else
		forward_stream_kinds<kStoreF32, true>(J, J.src, (uint32_t)J.src_pitch, 1u, x0, y0, lane);
}

// --- forward, levels l and l + 1 fused -------------------------------------
//
// One wave per tile of kFusedTileW1 x kFusedTileH1 level-(l+1) coefficients
// (per band).  The wave streams level l exactly as k_dwt_forward_stream does,
// over the level-l window that the tile's level-(l+1) outputs need: LL_l
// columns 2*ox1 - 4 .. 2*ox1 + 55 (lanes 2 .. 61) and rows 2*oy1 - 4 ..
// 2*oy1 + 2*H1 + 3, one LL_l row per row-pair step.  It stores the level-l
// high bands of the part it owns (lanes 6 .. 57, LL_l rows 2*oy1 .. +2*H1);
// LL_l goes on in registers: every two LL_l rows, the pairs of LL_l columns
// are gathered to lanes 0 .. 29 (ds_bpermute), lifted horizontally across
// those lanes and vertically by a second delay line, and lanes 2 .. 27 store
// level l + 1.  So LL_l is never written nor read back.
//
// Extension of level l + 1's input (whole-sample mirror about 0 and
// A - 1 of LL_l, as the level-(l+1) kernels read it): above and left of the
// frame the LL_l values the wave computes are already the mirror images
// (level l's input is mirrored about 0, and the lifting is symmetric, down to
// the order of every float addition); past the right edge the gather takes
// the mirrored column, past the bottom a short history of LL_l rows supplies
// the mirrored rows.  LL_l values are rounded to the storage the separate
// launches would keep them in (fp16 for FP16_LL), so the coefficients equal
// those of the level-by-level launches bit for bit.
//
// As in k_dwt_forward_stream: WORD_ROWS loads are a scalar row base plus a
// fixed per-lane word offset, and the storage kinds (JH: level l's high
// bands, KL / KH: level l + 1's LL and high bands) are template arguments.
template <int KIND, bool WORD_ROWS, bool LL16, int JH, int KL, int KH>
__device__ __forceinline__ void forward_fused_tile(const DwtJob &J, const DwtJob &K, const void *src_ptr,
                                                   uint32_t pitch, uint32_t step, int ox1, int oy1, int lane)
{
	constexpr int H1 = kFusedTileH1;
	constexpr int kSteps = 2 * H1 + 12;
	static_assert(kSteps % 2 == 0, "the row-pair loop is unrolled by two");
	const StreamSrc src = stream_src(src_ptr);
	const int x0 = 4 * ox1 - 8, y0 = 4 * oy1 - 8;
	const uint32_t xe = __umul24((uint32_t)fwd_ext(x0 - kApron + 2 * lane, J.R_x, J.A_x), step);
	const uint32_t xo = __umul24((uint32_t)fwd_ext(x0 - kApron + 2 * lane + 1, J.R_x, J.A_x), step);
	const int ox = 2 * ox1 - 6 + lane; // level-l band column (LL_l column) of this lane
	const bool store_x = lane >= 6 && lane < 58 && ox < J.bw;
	// Level l + 1: lane k < 30 lifts the LL_l column pair 2*ox1 - 4 + 2k, +1,
	// gathered from the lanes that hold them (mirrored past the right edge).
	const int A1 = K.A_x; // LL_l columns (= J.bw)
	auto ext1 = [&](int c) __attribute__((always_inline)) { return c < A1 ? c : 2 * A1 - 2 - c; };
	const int ce = 2 * ox1 - 4 + 2 * lane;
	int se = ext1(ce) - (2 * ox1 - 6), so = ext1(ce + 1) - (2 * ox1 - 6);
	se = lane < 30 && se >= 0 && se < 64 ? se : lane;
	so = lane < 30 && so >= 0 && so < 64 ? so : lane;
	const int addr_e = se << 2, addr_o = so << 2;
	const int oc = ox1 + lane - 2; // level-(l+1) band column of lane k
	const bool store_x1 = lane >= 2 && lane < 2 + kFusedTileW1 && oc < K.bw;
	const int B1 = K.A_y; // LL_l rows (= J.bh)

	constexpr uint32_t kEs = stream_esize<KIND>();
	const uint32_t be = src.lo + xe * kEs, bo = src.lo + xo * kEs;
	const uint32_t we = be >> 2, wo = bo >> 2, se_b = 8u * (be & 3u), so_b = 8u * (bo & 3u);
	// Row offsets: elements, or words for WORD_ROWS (exact: rows start on words).
	auto row_base = [&](int r) __attribute__((always_inline)) {
		const uint32_t el = __umul24((uint32_t)fwd_ext(r, J.R_y, J.A_y), pitch);
		return WORD_ROWS ? (el * kEs) >> 2 : el;
	};
	auto load_pair = [&](int n, uint32_t (&r)[4], uint32_t &r0, uint32_t &r1) __attribute__((always_inline)) {
		const int row = y0 - kApron + 2 * n;
		r0 = row_base(row);
		r1 = row_base(row + 1);
		if constexpr (WORD_ROWS)
		{
			r[0] = load_row_word(src.words, r0, we);
			r[1] = load_row_word(src.words, r0, wo);
			r[2] = load_row_word(src.words, r1, we);
			r[3] = load_row_word(src.words, r1, wo);
		}
		else
		{
			r[0] = stream_raw<KIND>(src, r0 + xe);
			r[1] = stream_raw<KIND>(src, r0 + xo);
			r[2] = stream_raw<KIND>(src, r1 + xe);
			r[3] = stream_raw<KIND>(src, r1 + xo);
		}
	};
	VLine lo, hi;   // level l
	VLine lo1, hi1; // level l + 1
	float pend = 0.0f;
	float h0 = 0.0f, h1 = 0.0f, h2 = 0.0f, h3 = 0.0f, h4 = 0.0f; // LL_l rows j-1 .. j-5 (bottom mirror)
	auto feed = [&](int n, float ll) __attribute__((always_inline)) {
		// LL_l row j = 2*oy1 - 8 + n, for n >= 4.
		const int j = 2 * oy1 - 8 + n;
		if (LL16)
		{
			// The FP32 product rounded to fp16, as the store of the separate
			// launch does (the empty asm keeps the compiler from folding the
			// product and the conversion into one v_fma_mix, which rounds the
			// exact product once).
			asm volatile("" : "+v"(ll));
			ll = __half2float(__float2half_rn(ll));
		}
		if (j >= B1)
		{
			// Row B1 + t mirrors to B1 - 2 - t: the frozen history, consumed
			// one row per step.
			ll = h1;
			h1 = h2;
			h2 = h3;
			h3 = h4;
		}
		else if (j >= B1 - 5)
		{
			h4 = h3;
			h3 = h2;
			h2 = h1;
			h1 = h0;
			h0 = ll;
		}
		if (((n - 4) & 1) == 0)
		{
			pend = ll;
			return;
		}
		// Pair m of level l + 1: LL_l rows 2*oy1 - 4 + 2m (pend) and + 1 (ll).
		const int m = (n - 5) >> 1;
		float e0 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr_e, __float_as_int(pend)));
		float o0 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr_o, __float_as_int(pend)));
		float e1 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr_e, __float_as_int(ll)));
		float o1 = __int_as_float(__builtin_amdgcn_ds_bpermute(addr_o, __float_as_int(ll)));
		hlift_row(e0, o0);
		hlift_row(e1, o1);
		float a_ll, a_lh, a_hl, a_hh;
		lo1.push(e0, e1, a_ll, a_lh);
		hi1.push(o0, o1, a_hl, a_hh);
		const int oy2 = oy1 + m - kApron;
		if (m >= kApron && store_x1 && oy2 < K.bh)
		{
			const uint32_t row = (uint32_t)oy2 * (uint32_t)K.band_pitch;
			stream_store_row<KL>(K.band[0], row, (uint32_t)oc, a_ll);
			stream_store_row<KH>(K.band[1], row, (uint32_t)oc, a_hl);
			stream_store_row<KH>(K.band[2], row, (uint32_t)oc, a_lh);
			stream_store_row<KH>(K.band[3], row, (uint32_t)oc, a_hh);
		}
	};
	auto lift_pair = [&](int n, uint32_t r0, uint32_t r1, const uint32_t (&r)[4]) __attribute__((always_inline)) {
		float e0, o0, e1, o1;
		if constexpr (WORD_ROWS)
		{
			e0 = stream_value_at<KIND>(r[0], se_b);
			o0 = stream_value_at<KIND>(r[1], so_b);
			e1 = stream_value_at<KIND>(r[2], se_b);
			o1 = stream_value_at<KIND>(r[3], so_b);
		}
		else
		{
			e0 = stream_value<KIND>(r[0], src, r0 + xe);
			o0 = stream_value<KIND>(r[1], src, r0 + xo);
			e1 = stream_value<KIND>(r[2], src, r1 + xe);
			o1 = stream_value<KIND>(r[3], src, r1 + xo);
		}
		hlift_row(e0, o0);
		hlift_row(e1, o1);
		float ll, lh, hl, hh;
		lo.push(e0, e1, ll, lh);
		hi.push(o0, o1, hl, hh);
		if (n < kApron)
			return;
		const int j = 2 * oy1 - 8 + n;
		if (n >= 8 && n < 8 + 2 * H1 && store_x && j < J.bh)
		{
			const uint32_t row = (uint32_t)j * (uint32_t)J.band_pitch;
			stream_store_row<JH>(J.band[1], row, (uint32_t)ox, hl);
			stream_store_row<JH>(J.band[2], row, (uint32_t)ox, lh);
			stream_store_row<JH>(J.band[3], row, (uint32_t)ox, hh);
		}
		feed(n, ll);
	};
	uint32_t ra[4], rb[4], a0, a1, b0, b1;
	load_pair(0, ra, a0, a1);
	load_pair(1, rb, b0, b1);
#pragma unroll 2
	for (int n = 0; n < kSteps; n += 2)
	{
		const uint32_t pa0 = a0, pa1 = a1, pb0 = b0, pb1 = b1;
		const uint32_t ca[4] = {ra[0], ra[1], ra[2], ra[3]}, cb[4] = {rb[0], rb[1], rb[2], rb[3]};
		if (n + 2 < kSteps)
		{
			load_pair(n + 2, ra, a0, a1);
			load_pair(n + 3, rb, b0, b1);
		}
		// The scheduler would otherwise sink the loads towards their use.
		__builtin_amdgcn_sched_barrier(0);
		lift_pair(n, pa0, pa1, ca);
		lift_pair(n + 1, pb0, pb1, cb);
	}
}

__global__ __launch_bounds__(64 * kStreamWaves) void k_dwt_forward_fused(const FusedJob *__restrict__ jobs,
                                                                         const uint16_t *__restrict__ tile_job,
                                                                         const FrameState *__restrict__ frames,
                                                                         int tiles, XcdMap xm)
{
	const int lane = threadIdx.x & 63;
	// Wave-uniform (readfirstlane tells the compiler so: the tile origin and
	// everything derived from it stay in scalar registers).
	const int tile = __builtin_amdgcn_readfirstlane(xcd_swizzle((int)blockIdx.x, xm) *
	                                                    kStreamWaves +
	                                                (int)(threadIdx.x >> 6));
	if (tile >= tiles)
		return; // whole waves: no wave-level or workgroup barrier follows
	const FusedJob &FJ = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const DwtJob J = FJ.a, K = FJ.b;
	const int local = tile - FJ.tile_base;
	const int ox1 = (local % FJ.tiles_x) * kFusedTileW1;
	const int oy1 = (local / FJ.tiles_x) * kFusedTileH1;
	// The fused levels are a component's first two: level l reads the u8 plane.
	const FrameState &F = frames[J.frame];
	const void *p = F.src[J.component];
	const uint32_t pitch = (uint32_t)F.src_pitch[J.component], step = (uint32_t)F.src_step[J.component];
	// The storage-kind combinations that occur: FP32 throughout; FP16 (fp16
	// high bands of levels 0-1, FP32 LL) fusing levels 0+1 or 1+2 (4:2:0
	// chroma: level 2 all FP32); FP16_LL (all four bands of levels 0-1 fp16,
	// so the carried LL rounds to fp16) fusing 0+1 or 1+2.
	auto go = [&](auto wr) {
		constexpr bool WR = decltype(wr)::value;
		if (J.band_kind == kStoreF32)
			forward_fused_tile<kStoreU8, WR, false, kStoreF32, kStoreF32, kStoreF32>(J, K, p, pitch, step, ox1, oy1,
			                                                                          lane);
		else if (J.ll_kind == kStoreF16)
		{
			if (K.ll_kind == kStoreF16)
				forward_fused_tile<kStoreU8, WR, true, kStoreF16, kStoreF16, kStoreF16>(J, K, p, pitch, step, ox1,
				                                                                         oy1, lane);
			else
				forward_fused_tile<kStoreU8, WR, true, kStoreF16, kStoreF32, kStoreF32>(J, K, p, pitch, step, ox1,
				                                                                         oy1, lane);
		}
		else if (K.band_kind == kStoreF16)
			forward_fused_tile<kStoreU8, WR, false, kStoreF16, kStoreF32, kStoreF16>(J, K, p, pitch, step, ox1, oy1,
			                                                                          lane);
		else
			forward_fused_tile<kStoreU8, WR, false, kStoreF16, kStoreF32, kStoreF32>(J, K, p, pitch, step, ox1, oy1,
			                                                                          lane);
	};
	if ((pitch & 3u) == 0)
		go(std::true_type());
	else
		go(std::false_type());
}

// --- inverse, streamed (shape 4) ------------------------------------------
//
// The inverse counterpart of k_dwt_forward_stream: one wave per tile of
// kInvVariantTileW x kInvVariantTileH output samples, no LDS.  Lane l holds
// the interleaved column pair ox0 - 4 + 2l, +1 (lanes 2 .. 61 produce the
// tile).  The wave walks down the interleaved rows two at a time: each row
// pair brings LL and HL samples (even row) and LH and HH samples (odd row),
// found through the decoder's symmetric extension; the two columns each lane
// holds go through the vertical inverse lifting as a delay line (row pair p
// completes when pair p + 2 arrives), and each completed row through the
// horizontal inverse lifting across the lanes.  The operations and their
// order are those of the LDS-tiled shapes (lift_inv, K / 1/K scaling before
// each pass), so the output is the same.

// Vertical inverse lifting of one column as a delay line.  Before pair n
// arrives: ro = odd sample of pair n-1 as input, e1 = even of pair n-1 after
// the first step, o2 = odd of pair n-2 after the second, e3 = even of pair
// n-2 after the third.
struct VLineInv
{
	float ro = 0.0f, e1 = 0.0f, o2 = 0.0f, e3 = 0.0f;

	// Takes pair n (scaled input); gives the finished pair n-2.
	__device__ __forceinline__ void push(float e, float o, float &even, float &odd)
	{
		const float ne1 = fmaf(-kDelta, ro + o, e);   // e1(n)
		const float no2 = fmaf(-kGamma, e1 + ne1, ro); // o2(n-1)
		const float ne3 = fmaf(-kBeta, o2 + no2, e1);  // e3(n-1)
		odd = fmaf(-kAlpha, e3 + ne3, o2);             // o4(n-2)
		even = e3;                                     // e3(n-2)
		ro = o;
		e1 = ne1;
		o2 = no2;
		e3 = ne3;
	}
};

// Horizontal inverse lifting of one row (e, o: this lane's even and odd
// column) after the column scaling.
__device__ __forceinline__ void hlift_inv_row(float &e, float &o)
{
	e = e * kK;
	o = o * kInvK;
	e = fmaf(-kDelta, dpp_wave<kDppWaveShr1>(o) + o, e);
	o = fmaf(-kGamma, e + dpp_wave<kDppWaveShl1>(e), o);
	e = fmaf(-kBeta, dpp_wave<kDppWaveShr1>(o) + o, e);
	o = fmaf(-kAlpha, e + dpp_wave<kDppWaveShl1>(e), o);
}

// One tile with the LL band stored as LKIND and the high bands as HKIND.
template <int LKIND, int HKIND, int DKIND>
__device__ __forceinline__ void inverse_stream_tile(const DwtJob &J, const FrameState *frames, int ox0, int oy0,
                                                    int lane)
{
	constexpr int TH = kInvVariantTileH[kInvStreamVariant];
	constexpr int kPairs = TH / 2 + kApron;
	static_assert(kPairs % 2 == 0, "the row-pair loop is unrolled by two");
	const uint32_t pitch = (uint32_t)J.band_pitch;
	const int xe = inv_ext(ox0 - kApron + 2 * lane, J.A_x), xo = inv_ext(ox0 - kApron + 2 * lane + 1, J.A_x);
	// Parity is preserved: xe is even (LL / LH column), xo odd (HL / HH).
	const uint32_t ce = (uint32_t)(xe >> 1), co = (uint32_t)(xo >> 1);
	const int px = ox0 - kApron + 2 * lane;
	const bool lane_out = lane >= kApron / 2 && lane < 64 - kApron / 2;

	uint8_t *plane = nullptr;
	long dpitch = J.dst_pitch;
	if (plane_kind(J.dst_kind))
	{
		plane = frames[J.frame].dst[J.component];
		dpitch = frames[J.frame].dst_pitch[J.component];
	}

	// Band-row offsets of the window's 2 * kPairs interleaved rows, one row
	// per lane, computed once; a row pair reads its two entries with readlane.
	static_assert(2 * kPairs <= 64, "one window row per lane");
	const uint32_t row_off =
	    lane < 2 * kPairs ? __umul24((uint32_t)(inv_ext(oy0 - kApron + lane, J.A_y) >> 1), pitch) : 0u;
	auto rows = [&](int n, uint32_t &r0, uint32_t &r1) {
		r0 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n);
		r1 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n + 1);
	};
	// Raw words: LL and HL of the even row, LH and HH of the odd row.  Band
	// storage is 256-byte aligned with pitches of 32 elements, so every row
	// starts on a word: a row's loads are a scalar row base (in words) plus
	// the lane's fixed word offset, and the bit shift within the word is the
	// lane's too.
	constexpr uint32_t kEl = stream_esize<LKIND>(), kEh = stream_esize<HKIND>();
	const GU32 ll = (GU32)J.band[0], hl = (GU32)J.band[1], lh = (GU32)J.band[2], hh = (GU32)J.band[3];
	const uint32_t wle = (ce * kEl) >> 2, whe = (ce * kEh) >> 2, who = (co * kEh) >> 2;
	const uint32_t sle = 8u * ((ce * kEl) & 3u), she = 8u * ((ce * kEh) & 3u), sho = 8u * ((co * kEh) & 3u);
	auto load_pair = [&](uint32_t r0, uint32_t r1, uint32_t (&r)[4]) {
		r[0] = load_row_word(ll, (r0 * kEl) >> 2, wle);
		r[1] = load_row_word(hl, (r0 * kEh) >> 2, who);
		r[2] = load_row_word(lh, (r1 * kEh) >> 2, whe);
		r[3] = load_row_word(hh, (r1 * kEh) >> 2, who);
	};
	// Scaled for the vertical pass: even rows by K, odd rows by 1/K.
	auto values = [&](uint32_t r0, uint32_t r1, const uint32_t (&r)[4], float (&v)[4]) {
		v[0] = stream_value_at<LKIND>(r[0], sle) * kK;
		v[1] = stream_value_at<HKIND>(r[1], sho) * kK;
		v[2] = stream_value_at<HKIND>(r[2], she) * kInvK;
		v[3] = stream_value_at<HKIND>(r[3], sho) * kInvK;
	};
	auto store_row = [&](int py, float e, float o) {
		if (!lane_out)
			return;
		if constexpr (DKIND == kStoreU8 || DKIND == kStoreU16)
		{
			if (py < J.R_y)
				store_plane_pair_row<DKIND>(plane, (uint32_t)py * (uint32_t)dpitch, px, J.R_x, e, o);
		}
		else if (py < J.A_y && px < J.A_x)
			// px is even and A_x = 2 * bw: px + 1 < A_x too.
			stream_store_pair_row<DKIND>(J.dst, (uint32_t)py * (uint32_t)dpitch, (uint32_t)px, e, o);
	};
	VLineInv ce_line, co_line; // the even and the odd column
	auto lift_pair = [&](int n, const float (&v)[4]) {
		float e0, e1, o0, o1; // rows 2p, 2p+1 of the finished pair p = n - 2; e: even column, o: odd
		ce_line.push(v[0], v[2], e0, e1);
		co_line.push(v[1], v[3], o0, o1);
		if (n < kApron)
			return;
		hlift_inv_row(e0, o0);
		hlift_inv_row(e1, o1);
		const int py = oy0 - kApron + 2 * (n - 2);
		store_row(py, e0, o0);
		store_row(py + 1, e1, o1);
	};
	uint32_t a0, a1, b0, b1, ra[4], rb[4];
	rows(0, a0, a1);
	rows(1, b0, b1);
	load_pair(a0, a1, ra);
	load_pair(b0, b1, rb);
#pragma unroll 2
	for (int n = 0; n < kPairs; n += 2)
	{
		float a[4], b[4];
		values(a0, a1, ra, a);
		values(b0, b1, rb, b);
		if (n + 2 < kPairs)
		{
			rows(n + 2, a0, a1);
			rows(n + 3, b0, b1);
			load_pair(a0, a1, ra);
			load_pair(b0, b1, rb);
		}
		// The scheduler would otherwise sink the loads towards their use.
		__builtin_amdgcn_sched_barrier(0);
		lift_pair(n, a);
		lift_pair(n + 1, b);
	}
}

__global__ __launch_bounds__(64 * kStreamWaves) void k_dwt_inverse_stream(const DwtJob *__restrict__ jobs,
                                                                          const uint16_t *__restrict__ tile_job,
                                                                          const FrameState *__restrict__ frames,
                                                                          int tiles, XcdMap xm)
{
	constexpr int TW = kInvVariantTileW[kInvStreamVariant], TH = kInvVariantTileH[kInvStreamVariant];
	static_assert(TW == 2 * 64 - 2 * kApron, "one interleaved column pair per lane");
	const int lane = threadIdx.x & 63;
	// Wave-uniform (readfirstlane tells the compiler so: the tile origin and
	// everything derived from it stay in scalar registers).
	const int tile = __builtin_amdgcn_readfirstlane(xcd_swizzle((int)blockIdx.x, xm) *
	                                                    kStreamWaves +
	                                                (int)(threadIdx.x >> 6));
	if (tile >= tiles)
		return; // whole waves: no wave-level or workgroup barrier follows
	const DwtJob J = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const int local = tile - J.tile_base;
	const int ox0 = (local % J.tiles_x) * TW;
	const int oy0 = (local / J.tiles_x) * TH;
	// Band and output storage kinds as template arguments; the pairs that
	// occur: fp16 LL and bands (FP16_LL levels 0-1) into a plane or an fp16
	// LL; FP32 LL with fp16 bands (FP16 levels 0-1) into a plane or an FP32
	// LL; FP32 throughout into a plane or an FP32 or fp16 LL (level 2 of
	// FP16_LL writes the fp16 LL of level 1).
	auto with_dst = [&](auto lk, auto hk) {
		constexpr int L = decltype(lk)::value, H = decltype(hk)::value;
		switch (J.dst_kind)
		{
		case kStoreU8:
			inverse_stream_tile<L, H, kStoreU8>(J, frames, ox0, oy0, lane);
			break;
		case kStoreU16:
			inverse_stream_tile<L, H, kStoreU16>(J, frames, ox0, oy0, lane);
			break;
		case kStoreF16:
			inverse_stream_tile<L, H, kStoreF16>(J, frames, ox0, oy0, lane);
			break;
		default:
			inverse_stream_tile<L, H, kStoreF32>(J, frames, ox0, oy0, lane);
			break;
		}
	};
	if (J.ll_kind == kStoreF16)
		with_dst(std::integral_constant<int, kStoreF16>(), std::integral_constant<int, kStoreF16>());
	else if (J.band_kind == kStoreF16)
		with_dst(std::integral_constant<int, kStoreF32>(), std::integral_constant<int, kStoreF16>());
	else
		with_dst(std::integral_constant<int, kStoreF32>(), std::integral_constant<int, kStoreF32>());
}

// --- inverse, levels l + 1 and l fused ---------------------------------------
//
// The inverse counterpart of k_dwt_forward_fused: one wave per tile of
// kInvFusedTileW x TH samples of a component's plane (level l,
// the component's first level), with two delay lines.  The level-(l+1) line
// walks the level-(l+1) bands under the tile's LL_l columns (lanes 0 .. 35;
// lanes 36 .. 63 repeat lane 35's loads) and finishes two LL_l rows per row
// pair; each finished LL_l row is gathered (ds_bpermute) from the lane that
// holds each column to the level-l lane that reads it, and enters level l's
// line with that level's HL / LH / HH rows, as the LL_l rows the separate
// launch would read back.  So LL_l is never written nor read.
//
// Extension of LL_l as level l reads it (the interleaved signal mirrored
// about 0 and A - 1: LL_l rows and columns mirror about 0 and about
// bw - 1/2): left and right, the gather takes the mirrored column; above the
// plane, the LL_l rows the level-(l+1) line computes from its own mirrored
// input are already the mirror images (whole-sample symmetry about 0
// survives every lifting step, down to the order of each float addition);
// past the bottom, where the level-(l+1) signal mirrors about bh - 1
// instead, the last LL_l row pair stands in for the next one in reverse
// order (rows bh and bh + 1 are the only ones past the bottom that reach a
// plane sample).  LL_l values are rounded to the storage the separate
// launches keep them in (fp16 for FP16_LL), so the planes equal those of the
// level-by-level launches bit for bit.
__device__ __forceinline__ int inv_ext_fast(int p, int N)
{
	return (unsigned)p < (unsigned)N ? p : inv_ext(p, N);
}

template <int TH, int KL, int KH, bool LL16, int JH>
__device__ __forceinline__ void inverse_fused_tile(const DwtJob &J, const DwtJob &K, const FrameState *frames,
                                                   int ox0, int oy0, int lane)
{
	constexpr int kPairs0 = TH / 2 + kApron; // level-l row pairs, as a streamed tile's
	static_assert(TH % 4 == 0, "LL_l row pairs start on even rows");
	static_assert(kInvFusedTileW == 2 * 64 - 2 * kApron, "one interleaved column pair per lane");
	static_assert((kPairs0 / 2) % 2 == 0, "the main loop is unrolled by two");
	const uint32_t pitch0 = (uint32_t)J.band_pitch, pitch1 = (uint32_t)K.band_pitch;
	const int c0 = ox0 >> 1; // the tile's first LL_l column (even)
	// Level l: lane k holds the interleaved column pair ox0 - 4 + 2k, +1.
	const int xe0 = inv_ext(ox0 - kApron + 2 * lane, J.A_x), xo0 = inv_ext(ox0 - kApron + 2 * lane + 1, J.A_x);
	const uint32_t ce0 = (uint32_t)(xe0 >> 1), co0 = (uint32_t)(xo0 >> 1);
	const int px = ox0 - kApron + 2 * lane;
	const bool lane_out = lane >= kApron / 2 && lane < 64 - kApron / 2;
	// Level l + 1: lane j holds the interleaved column pair c0 - 6 + 2j, +1,
	// whose horizontal inverse is LL_l columns c0 - 6 + 2j, +1; lanes 2 .. 33
	// are complete, LL_l columns c0 - 2 .. c0 + 61.
	const int j1 = lane < 36 ? lane : 35;
	const int xe1 = inv_ext(c0 - 6 + 2 * j1, K.A_x), xo1 = inv_ext(c0 - 5 + 2 * j1, K.A_x);
	const uint32_t ce1 = (uint32_t)(xe1 >> 1), co1 = (uint32_t)(xo1 >> 1);
	// Level-l lane k reads LL_l column ce0 (c0 - 2 + k extended), which lane
	// (ce0 - c0 + 6) / 2 holds as its even (ce0 even) or odd column.  For
	// every lane whose values reach a stored sample that lane is in 2 .. 33;
	// the clamp only keeps the others' addresses in the wave.
	const int gsrc = ((int)ce0 - c0 + 6) >> 1;
	const int gaddr = (gsrc < 0 ? 0 : gsrc > 63 ? 63 : gsrc) << 2;
	const bool godd = (ce0 & 1u) != 0;

	uint8_t *plane = frames[J.frame].dst[J.component];
	const uint32_t dpitch = (uint32_t)frames[J.frame].dst_pitch[J.component];
	constexpr uint32_t kEl1 = stream_esize<KL>(), kEh1 = stream_esize<KH>(), kEh0 = stream_esize<JH>();
	const GU32 ll1 = (GU32)K.band[0], hl1 = (GU32)K.band[1], lh1 = (GU32)K.band[2], hh1 = (GU32)K.band[3];
	const GU32 hl0 = (GU32)J.band[1], lh0 = (GU32)J.band[2], hh0 = (GU32)J.band[3];
	const uint32_t wle1 = (ce1 * kEl1) >> 2, whe1 = (ce1 * kEh1) >> 2, who1 = (co1 * kEh1) >> 2;
	const uint32_t sle1 = 8u * ((ce1 * kEl1) & 3u), she1 = 8u * ((ce1 * kEh1) & 3u),
	               sho1 = 8u * ((co1 * kEh1) & 3u);
	const uint32_t whe0 = (ce0 * kEh0) >> 2, who0 = (co0 * kEh0) >> 2;
	const uint32_t she0 = 8u * ((ce0 * kEh0) & 3u), sho0 = 8u * ((co0 * kEh0) & 3u);

	// Band-row offsets (elements) of the interleaved rows y (even) and y + 1;
	// wave-uniform, so scalar.
	auto rows = [&](int y, int A, uint32_t pitch, uint32_t &r0, uint32_t &r1) __attribute__((always_inline)) {
		r0 = (uint32_t)(inv_ext_fast(y, A) >> 1) * pitch;
		r1 = (uint32_t)(inv_ext_fast(y + 1, A) >> 1) * pitch;
	};
	// Level-(l+1) row pair m: interleaved rows oy0 / 2 - 6 + 2m, +1 (LL, HL;
	// LH, HH).  Level-l row pair n: rows oy0 - 4 + 2n, +1 (HL; LH, HH; its LL
	// comes from the gather).
	auto load1 = [&](int m, uint32_t (&r)[4]) __attribute__((always_inline)) {
		uint32_t r0, r1;
		rows((oy0 >> 1) - 6 + 2 * m, K.A_y, pitch1, r0, r1);
		r[0] = load_row_word(ll1, (r0 * kEl1) >> 2, wle1);
		r[1] = load_row_word(hl1, (r0 * kEh1) >> 2, who1);
		r[2] = load_row_word(lh1, (r1 * kEh1) >> 2, whe1);
		r[3] = load_row_word(hh1, (r1 * kEh1) >> 2, who1);
	};
	auto load0 = [&](int n, uint32_t (&r)[3]) __attribute__((always_inline)) {
		uint32_t r0, r1;
		rows(oy0 - kApron + 2 * n, J.A_y, pitch0, r0, r1);
		r[0] = load_row_word(hl0, (r0 * kEh0) >> 2, who0);
		r[1] = load_row_word(lh0, (r1 * kEh0) >> 2, whe0);
		r[2] = load_row_word(hh0, (r1 * kEh0) >> 2, who0);
	};

	VLineInv e1_line, o1_line; // level l + 1: the even and the odd column
	VLineInv e0_line, o0_line; // level l
	// Level-(l+1) row pair: the vertical delay line, and once it yields
	// finished rows (OUT), their horizontal inverse and the gather.
	auto step1 = [&](auto out, const uint32_t (&r)[4], float &g0, float &g1) __attribute__((always_inline)) {
		const float v0 = stream_value_at<KL>(r[0], sle1) * kK;
		const float v1 = stream_value_at<KH>(r[1], sho1) * kK;
		const float v2 = stream_value_at<KH>(r[2], she1) * kInvK;
		const float v3 = stream_value_at<KH>(r[3], sho1) * kInvK;
		float a0, a1, b0, b1;
		e1_line.push(v0, v2, a0, a1);
		o1_line.push(v1, v3, b0, b1);
		if constexpr (decltype(out)::value)
		{
			hlift_inv_row(a0, b0);
			hlift_inv_row(a1, b1);
			const float ge0 = __int_as_float(__builtin_amdgcn_ds_bpermute(gaddr, __float_as_int(a0)));
			const float go0 = __int_as_float(__builtin_amdgcn_ds_bpermute(gaddr, __float_as_int(b0)));
			const float ge1 = __int_as_float(__builtin_amdgcn_ds_bpermute(gaddr, __float_as_int(a1)));
			const float go1 = __int_as_float(__builtin_amdgcn_ds_bpermute(gaddr, __float_as_int(b1)));
			g0 = godd ? go0 : ge0;
			g1 = godd ? go1 : ge1;
			if constexpr (LL16)
			{
				g0 = __half2float(__float2half_rn(g0));
				g1 = __half2float(__float2half_rn(g1));
			}
		}
	};
	auto store_row = [&](int py, float e, float o) __attribute__((always_inline)) {
		if (lane_out && py < J.R_y)
			store_plane_pair_row<kStoreU8>(plane, (uint32_t)py * dpitch, px, J.R_x, e, o);
	};
	// Level-l row pair n with its LL_l row ll; from n = kApron on it
	// finishes plane rows.
	auto step0 = [&](auto st, int n, float ll, const uint32_t (&r)[3]) __attribute__((always_inline)) {
		const float v0 = ll * kK;
		const float v1 = stream_value_at<JH>(r[0], sho0) * kK;
		const float v2 = stream_value_at<JH>(r[1], she0) * kInvK;
		const float v3 = stream_value_at<JH>(r[2], sho0) * kInvK;
		float e0, e1, o0, o1;
		e0_line.push(v0, v2, e0, e1);
		o0_line.push(v1, v3, o0, o1);
		if constexpr (decltype(st)::value)
		{
			hlift_inv_row(e0, o0);
			hlift_inv_row(e1, o1);
			const int py = oy0 - kApron + 2 * (n - 2);
			store_row(py, e0, o0);
			store_row(py + 1, e1, o1);
		}
	};

	// Warm-up: level-(l+1) pairs 0 .. 3 (their outputs precede the first
	// LL_l row the tile needs); the last one also fetches level l's first two
	// pairs.
	uint32_t q1[4], qa[3], qb[3];
	float g0 = 0.0f, g1 = 0.0f;
	load1(0, q1);
#pragma unroll
	for (int m = 0; m < kApron; m++)
	{
		const uint32_t c1[4] = {q1[0], q1[1], q1[2], q1[3]};
		load1(m + 1, q1);
		if (m == kApron - 1)
		{
			load0(0, qa);
			load0(1, qb);
		}
		__builtin_amdgcn_sched_barrier(0);
		step1(std::false_type(), c1, g0, g1);
	}
	// Main loop: iteration t runs level-(l+1) pair t + 4, which finishes LL_l
	// rows oy0 / 2 - 2 + 2t, +1, then level-l pairs 2t, 2t + 1, which take
	// them.
	const int bh0 = J.bh;
	float p0 = 0.0f, p1 = 0.0f; // the previous LL_l row pair (bottom mirror)
	auto iteration = [&](auto st, int t) __attribute__((always_inline)) {
		const uint32_t c1[4] = {q1[0], q1[1], q1[2], q1[3]};
		const uint32_t ca[3] = {qa[0], qa[1], qa[2]}, cb[3] = {qb[0], qb[1], qb[2]};
		if (t + 1 < kPairs0 / 2)
		{
			load1(t + kApron + 1, q1);
			load0(2 * t + 2, qa);
			load0(2 * t + 3, qb);
		}
		// The scheduler would otherwise sink the loads towards their use.
		__builtin_amdgcn_sched_barrier(0);
		step1(std::true_type(), c1, g0, g1);
		if ((oy0 >> 1) - 2 + 2 * t == bh0)
		{
			// LL_l rows bh, bh + 1 mirror to bh - 1, bh - 2.
			g0 = p1;
			g1 = p0;
		}
		p0 = g0;
		p1 = g1;
		step0(st, 2 * t, g0, ca);
		step0(st, 2 * t + 1, g1, cb);
	};
	static_assert(kApron == 4, "the first two iterations finish no plane row");
	iteration(std::false_type(), 0);
	iteration(std::false_type(), 1);
#pragma unroll 2
	for (int t = 2; t < kPairs0 / 2; t++)
		iteration(std::true_type(), t);
}

template <int TH>
__global__ __launch_bounds__(64 * kStreamWaves) void k_dwt_inverse_fused(const FusedJob *__restrict__ jobs,
                                                                         const uint16_t *__restrict__ tile_job,
                                                                         const FrameState *__restrict__ frames,
                                                                         int tiles, XcdMap xm)
{
	const int lane = threadIdx.x & 63;
	// Wave-uniform (readfirstlane tells the compiler so).
	const int tile = __builtin_amdgcn_readfirstlane(xcd_swizzle((int)blockIdx.x, xm) *
	                                                    kStreamWaves +
	                                                (int)(threadIdx.x >> 6));
	if (tile >= tiles)
		return; // whole waves: no wave-level or workgroup barrier follows
	const FusedJob &FJ = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const DwtJob J = FJ.a, K = FJ.b;
	const int local = tile - FJ.tile_base;
	const int ox0 = (local % FJ.tiles_x) * kInvFusedTileW;
	const int oy0 = (local / FJ.tiles_x) * TH;
	// The storage-kind combinations that occur (level l + 1's LL and high
	// bands, whether LL_l rounds to fp16, level l's high bands): FP32
	// throughout; FP16 (fp16 high bands of levels 0-1, FP32 LL) fusing 1+0 or
	// 2+1 (4:2:0 chroma: level 2 all FP32); FP16_LL (all four bands of levels
	// 0-1 fp16) fusing 1+0 or 2+1.
	if (J.band_kind == kStoreF32)
		inverse_fused_tile<TH, kStoreF32, kStoreF32, false, kStoreF32>(J, K, frames, ox0, oy0, lane);
	else if (J.ll_kind == kStoreF16)
	{
		if (K.ll_kind == kStoreF16)
			inverse_fused_tile<TH, kStoreF16, kStoreF16, true, kStoreF16>(J, K, frames, ox0, oy0, lane);
		else
			inverse_fused_tile<TH, kStoreF32, kStoreF32, true, kStoreF16>(J, K, frames, ox0, oy0, lane);
	}
	else if (K.band_kind == kStoreF16)
		inverse_fused_tile<TH, kStoreF32, kStoreF16, false, kStoreF16>(J, K, frames, ox0, oy0, lane);
	else
		inverse_fused_tile<TH, kStoreF32, kStoreF32, false, kStoreF16>(J, K, frames, ox0, oy0, lane);
}

// --- packed FP16 arithmetic, streamed (shape 5) ------------------------------
//
// The reference's precision 0 (shaders/dwt_common.h:29-34: FLOAT =
// float16_t, f16vec2 lines): the streamed shape above with every lifting
// step in packed FP16.  The horizontal pass packs the two rows of a row pair
// (e0, e1) / (o0, o1), so one DPP shift moves both rows' neighbours; the
// vertical delay line packs the two columns a lane holds (horizontal low and
// high), so one v_pk_fma_f16 lifts both.  A step is an fp16 add and an fp16
// fused multiply-add with the fp16-rounded constant.  The K / 1/K scaling
// between the passes (and after the vertical one) multiplies by the FP32
// constant and rounds the exact product once to fp16; it also transposes the
// packing.  All sources and bands are
// fp16; level 0 reads u8 through the table fp16(fp16(u / 255) - 0.5).
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr _Float16 kAlphaH = (_Float16)kAlpha;
constexpr _Float16 kBetaH = (_Float16)kBeta;
constexpr _Float16 kGammaH = (_Float16)kGamma;
constexpr _Float16 kDeltaH = (_Float16)kDelta;

__constant__ uint16_t c_u8_dc_h[256]; // fp16(fp16(u / 255) - 0.5)

__device__ __forceinline__ h2 splat(_Float16 c) { return h2{c, c}; }
__device__ __forceinline__ h2 fma2(h2 a, h2 b, h2 c) { return __builtin_elementwise_fma(a, b, c); }

template <int Ctrl>
__device__ __forceinline__ h2 dpp_wave_h2(h2 v)
{
	return __builtin_bit_cast(h2, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), Ctrl, 0xF, 0xF, false));
}

// The K / 1/K scaling: fp16 value times an FP32 constant, the exact product
// rounded once to fp16 (v_fma_mix with S0 an fp16 half, S1 / S2 FP32, S2 =
// -0 so that zeros keep their sign).  Written out because the compiler folds
// only some of the equivalent cvt / mul / cvt sequences into this
// instruction, and a folded and an unfolded product can round differently.
// SEL: the half of `src` multiplied; mixlo writes the low half of the
// result, mixhi the high half (the low half kept).
template <int SEL>
__device__ __forceinline__ h2 mixlo(h2 src, float k)
{
	h2 d;
	const float mz = -0.0f;
	if constexpr (SEL == 0)
		asm("v_fma_mixlo_f16 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(d) : "v"(src), "v"(k), "v"(mz));
	else
		asm("v_fma_mixlo_f16 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(d) : "v"(src), "v"(k), "v"(mz));
	return d;
}

template <int SEL>
__device__ __forceinline__ h2 mixhi(h2 d, h2 src, float k)
{
	const float mz = -0.0f;
	if constexpr (SEL == 0)
		asm("v_fma_mixhi_f16 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "+v"(d) : "v"(src), "v"(k), "v"(mz));
	else
		asm("v_fma_mixhi_f16 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(d) : "v"(src), "v"(k), "v"(mz));
	return d;
}

// (a.SA * ka, b.SB * kb), each product rounded once to fp16.
template <int SA, int SB>
__device__ __forceinline__ h2 scale_pair(h2 a, float ka, h2 b, float kb)
{
	return mixhi<SB>(mixlo<SA>(a, ka), b, kb);
}

// The sample at element offset `off` of a u8 plane (with the DC shift) or an
// fp16 band, out of its aligned word.
template <int KIND>
__device__ __forceinline__ _Float16 stream_half(uint32_t raw, const StreamSrc &S, uint32_t off)
{
	const uint32_t a = S.lo + off * stream_esize<KIND>();
	if constexpr (KIND == kStoreU8)
		return __builtin_bit_cast(_Float16, c_u8_dc_h[(raw >> (8 * (a & 3))) & 0xffu]);
	else if constexpr (KIND == kStoreU16)
	{
		// fp16(fp16(float(u / 65535)) - 0.5): three roundings.  The barriers
		// keep the compiler from merging the float quotient's fma with its
		// conversion (v_fma_mix rounds once, straight to fp16), or the
		// conversion with the fp16 subtraction.
		float tf = u16_unorm((raw >> (8 * (a & 2))) & 0xffffu);
		asm volatile("" : "+v"(tf));
		_Float16 t = (_Float16)tf;
		asm volatile("" : "+v"(t));
		return t - (_Float16)0.5f;
	}
	else
		return __builtin_bit_cast(_Float16, (uint16_t)(raw >> (8 * (a & 2))));
}

__device__ __forceinline__ void store_half(void *base, uint32_t idx, _Float16 v)
{
	((__attribute__((address_space(1))) _Float16 *)base)[idx] = v;
}

// Horizontal lifting of a row pair across the lanes: e = (even sample of row
// 0, of row 1), o likewise.
__device__ __forceinline__ void hlift_rows_h(h2 &e, h2 &o)
{
	o = fma2(splat(kAlphaH), e + dpp_wave_h2<kDppWaveShl1>(e), o);
	e = fma2(splat(kBetaH), dpp_wave_h2<kDppWaveShr1>(o) + o, e);
	o = fma2(splat(kGammaH), e + dpp_wave_h2<kDppWaveShl1>(e), o);
	e = fma2(splat(kDeltaH), dpp_wave_h2<kDppWaveShr1>(o) + o, e);
}

// VLine on the lane's two columns at once (x: horizontal low, y: high).
struct VLineH
{
	h2 re = {}, ro = {}, po = {}, pe = {}, qo = {};

	// Takes pair n (e: even row, o: odd row); gives pair n-2 scaled:
	// low = (LL, HL), high = (LH, HH).
	__device__ __forceinline__ void push(h2 e, h2 o, h2 &low, h2 &high)
	{
		const h2 o1 = fma2(splat(kAlphaH), re + e, ro);
		const h2 e1 = fma2(splat(kBetaH), po + o1, re);
		const h2 o2 = fma2(splat(kGammaH), pe + e1, po);
		const h2 e2 = fma2(splat(kDeltaH), qo + o2, pe);
		low = scale_pair<0, 1>(e2, kInvK, e2, kInvK);
		high = scale_pair<0, 1>(o2, kK, o2, kK);
		re = e;
		ro = o;
		po = o1;
		pe = e1;
		qo = o2;
	}
};

template <int KIND>
__device__ __forceinline__ void forward_stream_tile_h(const DwtJob &J, const void *src_ptr, uint32_t pitch,
                                                      uint32_t step, int x0, int y0, int lane)
{
	const StreamSrc src = stream_src(src_ptr);
	constexpr int TH = kFwdVariantTileH[kHalfStreamVariant];
	constexpr int kPairs = TH / 2 + kApron;
	static_assert(kPairs % 2 == 0, "the row-pair loop is unrolled by two");
	const uint32_t xe = __umul24((uint32_t)fwd_ext(x0 - kApron + 2 * lane, J.R_x, J.A_x), step);
	const uint32_t xo = __umul24((uint32_t)fwd_ext(x0 - kApron + 2 * lane + 1, J.R_x, J.A_x), step);
	const int ox = (x0 >> 1) - kApron / 2 + lane;
	const bool store_x = lane >= kApron / 2 && lane < 64 - kApron / 2 && ox < J.bw;
	static_assert(2 * kPairs <= 64, "one window row per lane");
	const uint32_t row_off =
	    lane < 2 * kPairs ? __umul24((uint32_t)fwd_ext(y0 - kApron + lane, J.R_y, J.A_y), pitch) : 0u;
	auto rows = [&](int n, uint32_t &r0, uint32_t &r1) {
		r0 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n);
		r1 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n + 1);
	};
	auto load_pair = [&](uint32_t r0, uint32_t r1, uint32_t (&r)[4]) {
		r[0] = stream_raw<KIND>(src, r0 + xe);
		r[1] = stream_raw<KIND>(src, r0 + xo);
		r[2] = stream_raw<KIND>(src, r1 + xe);
		r[3] = stream_raw<KIND>(src, r1 + xo);
	};
	VLineH line;
	auto lift_pair = [&](int n, uint32_t r0, uint32_t r1, const uint32_t (&r)[4]) {
		h2 e = h2{stream_half<KIND>(r[0], src, r0 + xe), stream_half<KIND>(r[2], src, r1 + xe)};
		h2 o = h2{stream_half<KIND>(r[1], src, r0 + xo), stream_half<KIND>(r[3], src, r1 + xo)};
		hlift_rows_h(e, o);
		// Rows of the pair as (horizontal low, high), scaled.
		const h2 v0 = scale_pair<0, 0>(e, kInvK, o, kK);
		const h2 v1 = scale_pair<1, 1>(e, kInvK, o, kK);
		h2 low, high;
		line.push(v0, v1, low, high);
		const int oy = (y0 >> 1) + n - kApron;
		if (n >= kApron && store_x && oy < J.bh)
		{
			const uint32_t idx = __umul24((uint32_t)oy, (uint32_t)J.band_pitch) + (uint32_t)ox;
			store_half(J.band[0], idx, low.x);
			store_half(J.band[1], idx, low.y);
			store_half(J.band[2], idx, high.x);
			store_half(J.band[3], idx, high.y);
		}
	};
	uint32_t a0, a1, b0, b1, ra[4], rb[4];
	rows(0, a0, a1);
	rows(1, b0, b1);
	load_pair(a0, a1, ra);
	load_pair(b0, b1, rb);
#pragma unroll 1
	for (int n = 0; n < kPairs; n += 2)
	{
		const uint32_t pa0 = a0, pa1 = a1, pb0 = b0, pb1 = b1;
		uint32_t ca[4] = {ra[0], ra[1], ra[2], ra[3]}, cb[4] = {rb[0], rb[1], rb[2], rb[3]};
		if (n + 2 < kPairs)
		{
			rows(n + 2, a0, a1);
			rows(n + 3, b0, b1);
			load_pair(a0, a1, ra);
			load_pair(b0, b1, rb);
		}
		__builtin_amdgcn_sched_barrier(0);
		lift_pair(n, pa0, pa1, ca);
		lift_pair(n + 1, pb0, pb1, cb);
	}
}

__global__ __launch_bounds__(64 * kStreamWaves) void k_dwt_forward_stream_h(const DwtJob *__restrict__ jobs,
                                                                            const uint16_t *__restrict__ tile_job,
                                                                            const FrameState *__restrict__ frames,
                                                                            int tiles, XcdMap xm)
{
	constexpr int TW = kFwdVariantTileW[kHalfStreamVariant], TH = kFwdVariantTileH[kHalfStreamVariant];
	const int lane = threadIdx.x & 63;
	// Wave-uniform (readfirstlane tells the compiler so: the tile origin and
	// everything derived from it stay in scalar registers).
	const int tile = __builtin_amdgcn_readfirstlane(xcd_swizzle((int)blockIdx.x, xm) *
	                                                    kStreamWaves +
	                                                (int)(threadIdx.x >> 6));
	if (tile >= tiles)
		return; // whole waves: no wave-level or workgroup barrier follows
	const DwtJob J = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const int local = tile - J.tile_base;
	const int x0 = (local % J.tiles_x) * TW;
	const int y0 = (local / J.tiles_x) * TH;
	if (J.src_kind == kStoreU8)
	{
		const FrameState &F = frames[J.frame];
		forward_stream_tile_h<kStoreU8>(J, F.src[J.component], (uint32_t)F.src_pitch[J.component],
		                                (uint32_t)F.src_step[J.component], x0, y0, lane);
	}
	else if (J.src_kind == kStoreU16)
	{
		const FrameState &F = frames[J.frame];
		forward_stream_tile_h<kStoreU16>(J, F.src[J.component], (uint32_t)F.src_pitch[J.component],
		                                 (uint32_t)F.src_step[J.component], x0, y0, lane);
	}
	else
		forward_stream_tile_h<kStoreF16>(J, J.src, (uint32_t)J.src_pitch, 1u, x0, y0, lane);
}

// VLineInv on the lane's two columns at once (x: even column, y: odd).
struct VLineInvH
{
	h2 ro = {}, e1 = {}, o2 = {}, e3 = {};

	// Takes pair n (e: even row, o: odd row, scaled); gives pair n-2's rows.
	__device__ __forceinline__ void push(h2 e, h2 o, h2 &even, h2 &odd)
	{
		const h2 ne1 = fma2(splat(-kDeltaH), ro + o, e);
		const h2 no2 = fma2(splat(-kGammaH), e1 + ne1, ro);
		const h2 ne3 = fma2(splat(-kBetaH), o2 + no2, e1);
		odd = fma2(splat(-kAlphaH), e3 + ne3, o2);
		even = e3;
		ro = o;
		e1 = ne1;
		o2 = no2;
		e3 = ne3;
	}
};

// Horizontal inverse lifting of a row pair (e: even column of row 0, of row
// 1; o likewise), after the scaling.
__device__ __forceinline__ void hlift_inv_rows_h(h2 &e, h2 &o)
{
	e = fma2(splat(-kDeltaH), dpp_wave_h2<kDppWaveShr1>(o) + o, e);
	o = fma2(splat(-kGammaH), e + dpp_wave_h2<kDppWaveShl1>(e), o);
	e = fma2(splat(-kBetaH), dpp_wave_h2<kDppWaveShr1>(o) + o, e);
	o = fma2(splat(-kAlphaH), e + dpp_wave_h2<kDppWaveShl1>(e), o);
}

__device__ __forceinline__ void inverse_stream_tile_h(const DwtJob &J, const FrameState *frames, int ox0, int oy0,
                                                      int lane)
{
	constexpr int TH = kInvVariantTileH[kHalfStreamVariant];
	constexpr int kPairs = TH / 2 + kApron;
	static_assert(kPairs % 2 == 0, "the row-pair loop is unrolled by two");
	const uint32_t pitch = (uint32_t)J.band_pitch;
	const int xe = inv_ext(ox0 - kApron + 2 * lane, J.A_x), xo = inv_ext(ox0 - kApron + 2 * lane + 1, J.A_x);
	const uint32_t ce = (uint32_t)(xe >> 1), co = (uint32_t)(xo >> 1);
	const int px = ox0 - kApron + 2 * lane;
	const bool lane_out = lane >= kApron / 2 && lane < 64 - kApron / 2;

	uint8_t *plane = nullptr;
	long dpitch = J.dst_pitch;
	if (plane_kind(J.dst_kind))
	{
		plane = frames[J.frame].dst[J.component];
		dpitch = frames[J.frame].dst_pitch[J.component];
	}
	static_assert(2 * kPairs <= 64, "one window row per lane");
	const uint32_t row_off =
	    lane < 2 * kPairs ? __umul24((uint32_t)(inv_ext(oy0 - kApron + lane, J.A_y) >> 1), pitch) : 0u;
	auto rows = [&](int n, uint32_t &r0, uint32_t &r1) {
		r0 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n);
		r1 = (uint32_t)__builtin_amdgcn_readlane((int)row_off, 2 * n + 1);
	};
	const StreamSrc ll = stream_src(J.band[0]), hl = stream_src(J.band[1]), lh = stream_src(J.band[2]),
	                hh = stream_src(J.band[3]);
	auto load_pair = [&](uint32_t r0, uint32_t r1, uint32_t (&r)[4]) {
		r[0] = stream_raw<kStoreF16>(ll, r0 + ce);
		r[1] = stream_raw<kStoreF16>(hl, r0 + co);
		r[2] = stream_raw<kStoreF16>(lh, r1 + ce);
		r[3] = stream_raw<kStoreF16>(hh, r1 + co);
	};
	auto store_row = [&](int py, _Float16 e, _Float16 o) {
		if (!lane_out)
			return;
		if (plane_kind(J.dst_kind))
		{
			if (py < J.R_y)
				store_plane_pair(plane, J.dst_kind, dpitch, py, px, J.R_x, (float)e, (float)o);
		}
		else if (py < J.A_y)
		{
			const uint32_t idx = __umul24((uint32_t)py, (uint32_t)dpitch) + (uint32_t)px;
			if (px < J.A_x)
				store_half(J.dst, idx, e);
			if (px + 1 < J.A_x)
				store_half(J.dst, idx + 1, o);
		}
	};
	VLineInvH line;
	auto lift_pair = [&](int n, uint32_t r0, uint32_t r1, const uint32_t (&r)[4]) {
		// Scaled for the vertical pass: even rows by K, odd rows by 1/K.
		const h2 re = h2{stream_half<kStoreF16>(r[0], ll, r0 + ce), stream_half<kStoreF16>(r[1], hl, r0 + co)};
		const h2 ro = h2{stream_half<kStoreF16>(r[2], lh, r1 + ce), stream_half<kStoreF16>(r[3], hh, r1 + co)};
		const h2 e = scale_pair<0, 1>(re, kK, re, kK);
		const h2 o = scale_pair<0, 1>(ro, kInvK, ro, kInvK);
		h2 even, odd; // rows 2p, 2p+1 of the finished pair p = n - 2: (even column, odd column)
		line.push(e, o, even, odd);
		if (n < kApron)
			return;
		h2 he = scale_pair<0, 0>(even, kK, odd, kK);
		h2 ho = scale_pair<1, 1>(even, kInvK, odd, kInvK);
		hlift_inv_rows_h(he, ho);
		const int py = oy0 - kApron + 2 * (n - 2);
		store_row(py, he.x, ho.x);
		store_row(py + 1, he.y, ho.y);
	};
	uint32_t a0, a1, b0, b1, ra[4], rb[4];
	rows(0, a0, a1);
	rows(1, b0, b1);
	load_pair(a0, a1, ra);
	load_pair(b0, b1, rb);
#pragma unroll 1
	for (int n = 0; n < kPairs; n += 2)
	{
		const uint32_t pa0 = a0, pa1 = a1, pb0 = b0, pb1 = b1;
		uint32_t ca[4] = {ra[0], ra[1], ra[2], ra[3]}, cb[4] = {rb[0], rb[1], rb[2], rb[3]};
		if (n + 2 < kPairs)
		{
			rows(n + 2, a0, a1);
			rows(n + 3, b0, b1);
			load_pair(a0, a1, ra);
			load_pair(b0, b1, rb);
		}
		__builtin_amdgcn_sched_barrier(0);
		lift_pair(n, pa0, pa1, ca);
		lift_pair(n + 1, pb0, pb1, cb);
	}
}

__global__ __launch_bounds__(64 * kStreamWaves) void k_dwt_inverse_stream_h(const DwtJob *__restrict__ jobs,
                                                                            const uint16_t *__restrict__ tile_job,
                                                                            const FrameState *__restrict__ frames,
                                                                            int tiles, XcdMap xm)
{
	constexpr int TW = kInvVariantTileW[kHalfStreamVariant], TH = kInvVariantTileH[kHalfStreamVariant];
	static_assert(TW == 2 * 64 - 2 * kApron, "one interleaved column pair per lane");
	const int lane = threadIdx.x & 63;
	// Wave-uniform (readfirstlane tells the compiler so: the tile origin and
	// everything derived from it stay in scalar registers).
	const int tile = __builtin_amdgcn_readfirstlane(xcd_swizzle((int)blockIdx.x, xm) *
	                                                    kStreamWaves +
	                                                (int)(threadIdx.x >> 6));
	if (tile >= tiles)
		return; // whole waves: no wave-level or workgroup barrier follows
	const DwtJob J = jobs[__builtin_amdgcn_readfirstlane(tile_job[tile])];
	const int local = tile - J.tile_base;
	const int ox0 = (local % J.tiles_x) * TW;
	const int oy0 = (local / J.tiles_x) * TH;
	inverse_stream_tile_h(J, frames, ox0, oy0, lane);
}
} // namespace

void launch_dwt_forward(const DwtJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                        int variant, hipStream_t stream)
{
	if (tiles <= 0)
		return;
	// Streamed shapes: kStreamWaves tiles per workgroup; tiled shapes: one.
	const dim3 sgrid((tiles + kStreamWaves - 1) / kStreamWaves), tgrid(tiles), sblock(64 * kStreamWaves);
	if (variant == kHalfStreamVariant)
		hipLaunchKernelGGL(k_dwt_forward_stream_h, sgrid, sblock, 0, stream, jobs, tile_job, frames, tiles,
		                   xcd_map(sgrid));
	else if (variant == kFwdStreamVariant)
		hipLaunchKernelGGL(k_dwt_forward_stream, sgrid, sblock, 0, stream, jobs, tile_job, frames, tiles,
		                   xcd_map(sgrid));
	else if (variant == 1)
		hipLaunchKernelGGL((k_dwt_forward<kFwdVariantTileW[1], kFwdVariantTileH[1], false>), tgrid,
		                   dim3(2 * kFwdVariantTileW[1]), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
	else if (variant == 2)
		hipLaunchKernelGGL((k_dwt_forward<kFwdVariantTileW[2], kFwdVariantTileH[2], false>), tgrid,
		                   dim3(2 * kFwdVariantTileW[2]), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
	else if (variant == 3)
		hipLaunchKernelGGL((k_dwt_forward<kFwdVariantTileW[3], kFwdVariantTileH[3], false>), tgrid,
		                   dim3(2 * kFwdVariantTileW[3]), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
	else
		hipLaunchKernelGGL((k_dwt_forward<kFwdVariantTileW[0], kFwdVariantTileH[0], true>), tgrid,
		                   dim3(2 * kFwdVariantTileW[0]), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
}

void launch_dwt_forward_fused(const FusedJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                              hipStream_t stream)
{
	if (tiles <= 0)
		return;
	const dim3 grid((tiles + kStreamWaves - 1) / kStreamWaves), block(64 * kStreamWaves);
	hipLaunchKernelGGL(k_dwt_forward_fused, grid, block, 0, stream, jobs, tile_job, frames, tiles, xcd_map(grid));
}

void launch_dwt_inverse_fused(const FusedJob *jobs, const uint16_t *tile_job, int tiles, int tile_h,
                              const FrameState *frames, hipStream_t stream)
{
	if (tiles <= 0)
		return;
	const dim3 grid((tiles + kStreamWaves - 1) / kStreamWaves), block(64 * kStreamWaves);
	if (tile_h == kInvFusedTallH)
		hipLaunchKernelGGL(k_dwt_inverse_fused<kInvFusedTallH>, grid, block, 0, stream, jobs, tile_job, frames,
		                   tiles, xcd_map(grid));
	else if (tile_h == kInvFusedShortH)
		hipLaunchKernelGGL(k_dwt_inverse_fused<kInvFusedShortH>, grid, block, 0, stream, jobs, tile_job, frames,
		                   tiles, xcd_map(grid));
	else
		hipLaunchKernelGGL(k_dwt_inverse_fused<kInvFusedLoneH>, grid, block, 0, stream, jobs, tile_job, frames,
		                   tiles, xcd_map(grid));
}

void launch_dwt_inverse(const DwtJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                        int variant, hipStream_t stream)
{
	if (tiles <= 0)
		return;
	const dim3 sgrid((tiles + kStreamWaves - 1) / kStreamWaves), tgrid(tiles), sblock(64 * kStreamWaves);
	if (variant == kHalfStreamVariant)
		hipLaunchKernelGGL(k_dwt_inverse_stream_h, sgrid, sblock, 0, stream, jobs, tile_job, frames, tiles,
		                   xcd_map(sgrid));
	else if (variant == kInvStreamVariant)
		hipLaunchKernelGGL(k_dwt_inverse_stream, sgrid, sblock, 0, stream, jobs, tile_job, frames, tiles,
		                   xcd_map(sgrid));
	else if (variant == 1)
		hipLaunchKernelGGL((k_dwt_inverse<kInvVariantTileW[1], kInvVariantTileH[1], 128, 16, false>), tgrid,
		                   dim3(128), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
	else if (variant == 2)
		hipLaunchKernelGGL((k_dwt_inverse<kInvVariantTileW[2], kInvVariantTileH[2], 64, 16, false>), tgrid,
		                   dim3(64), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
	else if (variant == 3)
		hipLaunchKernelGGL((k_dwt_inverse<kInvVariantTileW[3], kInvVariantTileH[3], 256, 24, true>), tgrid,
		                   dim3(256), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
	else
		hipLaunchKernelGGL((k_dwt_inverse<kInvVariantTileW[0], kInvVariantTileH[0], 256, 16, false>), tgrid,
		                   dim3(256), 0, stream, jobs, tile_job, frames, xcd_map(tgrid));
}

int xcd_count()
{
	static int cache[64] = {};
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0)
		return 1;
	if (dev < 64 && cache[dev])
		return cache[dev];
	int n = 1;
	if (hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || n < 1)
		n = 1;
	if (dev < 64)
		cache[dev] = n;
	return n;
}

XcdMap xcd_map(const dim3 &grid)
{
	return make_xcd_map(int(grid.x), xcd_count());
}

void upload_constants()
{
	float lut[256];
	for (int u = 0; u < 256; u++)
		lut[u] = (float)u / 255.0f - 0.5f;
	(void)hipMemcpyToSymbol(HIP_SYMBOL(c_u8_dc), lut, sizeof(lut));
	// fp16(fp16(u / 255) - 0.5): the fp16 subtraction rounds the exact
	// difference once (a float holds it exactly, so the float difference
	// rounded to fp16 is the fp16 result).
	uint16_t lut_h[256];
	for (int u = 0; u < 256; u++)
	{
		const _Float16 t = (_Float16)((float)u / 255.0f);
		const _Float16 d = (_Float16)((float)t - 0.5f);
		lut_h[u] = __builtin_bit_cast(uint16_t, d);
	}
	(void)hipMemcpyToSymbol(HIP_SYMBOL(c_u8_dc_h), lut_h, sizeof(lut_h));
}
} // namespace pw
