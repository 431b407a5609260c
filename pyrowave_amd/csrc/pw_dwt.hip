// pw_dwt.hip -- forward and inverse CDF 9/7 DWT levels on gfx950.
//
// One 256-thread workgroup transforms a 64x64 tile of a level's domain.  The
// tile plus a 4-sample apron (72x72) is staged in LDS once; the first 1-D pass
// runs over LDS rows (forward: horizontal; inverse: vertical columns), writes a
// second LDS buffer, and the second pass writes coalesced rows of the output
// bands straight to HBM.  Every lifting step is fmaf(c, l + r, v), the order
// of the reference shaders (reference shaders/dwt.comp:108-115,
// shaders/idwt.comp:95-108), so a tile reproduces the whole-line transform
// bit for bit.
#include "pw_common.hpp"

#include <hip/hip_fp16.h>

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

constexpr int kLoadIters = (kTileIn * kTileIn + kDwtThreads - 1) / kDwtThreads;
constexpr int kPitchA = kTileIn + 1; // odd pitch: row-per-lane reads are conflict-free

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
// gather footprint and a MIRRORED_REPEAT sampler).
__device__ __forceinline__ int fwd_ext(int p, int R, int A)
{
	int c = p & ~1;
	int coord = c < 0 ? c : c + 1;
	if (coord >= 2 * A - R)
		coord = coord + 2 * (R - A) + 1;
	else if (coord > R)
		coord = R;
	return mirror_repeat(coord - 1 + (p & 1), R);
}

// Decoder extension: whole-sample symmetric about 0 and N-1 of the
// interleaved signal (reference shaders/idwt.comp:25-32).
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

__device__ __forceinline__ float load_kind(const void *base, int kind, long idx)
{
	if (kind == kStoreF16)
		return __half2float(static_cast<const __half *>(base)[idx]);
	return static_cast<const float *>(base)[idx];
}

__device__ __forceinline__ void store_kind(void *base, int kind, long idx, float v)
{
	if (kind == kStoreF16)
		static_cast<__half *>(base)[idx] = __float2half_rn(v);
	else
		static_cast<float *>(base)[idx] = v;
}

// Forward lifting on a 16-sample window whose first sample is at an even
// position; outputs are windows 4..11 (low at even, high at odd).
__device__ __forceinline__ void lift_fwd16(float (&v)[16])
{
#pragma unroll
	for (int i = 1; i <= 13; i += 2)
		v[i] = fmaf(kAlpha, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 2; i <= 12; i += 2)
		v[i] = fmaf(kBeta, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 3; i <= 11; i += 2)
		v[i] = fmaf(kGamma, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 4; i <= 10; i += 2)
		v[i] = fmaf(kDelta, v[i - 1] + v[i + 1], v[i]);
}

// Inverse lifting on a 16-sample window (already scaled by K / 1/K).
__device__ __forceinline__ void lift_inv16(float (&v)[16])
{
#pragma unroll
	for (int i = 2; i <= 14; i += 2)
		v[i] = fmaf(-kDelta, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 3; i <= 13; i += 2)
		v[i] = fmaf(-kGamma, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 4; i <= 12; i += 2)
		v[i] = fmaf(-kBeta, v[i - 1] + v[i + 1], v[i]);
#pragma unroll
	for (int i = 5; i <= 11; i += 2)
		v[i] = fmaf(-kAlpha, v[i - 1] + v[i + 1], v[i]);
}

// Loads the 72x72 source tile (rows / columns through the extension maps)
// into registers: element i = tid + 256 * k; the last round (k = 20) is
// done by the first wave only.  Every load is issued before any is consumed.
template <typename T>
__device__ __forceinline__ void load_tile(const T *src, long pitch, long step, const int *ymap, const int *xmap,
                                          const float *lut, int tid, float (&vals)[kLoadIters])
{
#pragma unroll
	for (int k = 0; k < kLoadIters; k++)
	{
		const int i = tid + k * kDwtThreads;
		vals[k] = 0.0f;
		if (k < kLoadIters - 1 || tid < kTileIn * kTileIn - (kLoadIters - 1) * kDwtThreads)
		{
			const int r = i / kTileIn, c = i - r * kTileIn;
			const T v = src[(long)ymap[r] * pitch + (long)xmap[c] * step];
			if constexpr (sizeof(T) == 1)
				vals[k] = lut[v];
			else if constexpr (sizeof(T) == 2)
				vals[k] = __half2float(v);
			else
				vals[k] = v;
		}
	}
}

__device__ __forceinline__ void store_tile(float *sA, int tid, const float (&vals)[kLoadIters])
{
#pragma unroll
	for (int k = 0; k < kLoadIters; k++)
	{
		const int i = tid + k * kDwtThreads;
		if (k < kLoadIters - 1 || tid < kTileIn * kTileIn - (kLoadIters - 1) * kDwtThreads)
		{
			const int r = i / kTileIn, c = i - r * kTileIn;
			sA[r * kPitchA + c] = vals[k];
		}
	}
}

// First-pass work: 576 windows (72 lines x 8 segments); thread tid takes
// tasks tid, tid + 256 and (first wave only) tid + 512.
constexpr int kPass1Tasks = kTileIn * 8;
constexpr int kPass1Rounds = (kPass1Tasks + kDwtThreads - 1) / kDwtThreads;

__global__ __launch_bounds__(kDwtThreads) void k_dwt_forward(const DwtJob *__restrict__ jobs,
                                                             const uint16_t *__restrict__ tile_job,
                                                             const FrameState *__restrict__ frames, int nxcd)
{
	__shared__ float sA[kTileIn * kPitchA];
	__shared__ int s_xmap[kTileIn];
	__shared__ int s_ymap[kTileIn];
	__shared__ float s_lut[256];

	const int tid = threadIdx.x;
	const int tile = xcd_swizzle((int)blockIdx.x, (int)gridDim.x, nxcd);
	const DwtJob J = jobs[tile_job[tile]];
	const int local = tile - J.tile_base;
	const int x0 = (local % J.tiles_x) * kTile;
	const int y0 = (local / J.tiles_x) * kTile;

	if (tid < kTileIn)
		s_xmap[tid] = fwd_ext(x0 - kApron + tid, J.R_x, J.A_x);
	else if (tid < 2 * kTileIn)
		s_ymap[tid - kTileIn] = fwd_ext(y0 - kApron + tid - kTileIn, J.R_y, J.A_y);
	s_lut[tid] = c_u8_dc[tid];
	__syncthreads();

	float vals[kLoadIters];
	if (J.src_kind == kStoreU8)
	{
		const FrameState &F = frames[J.frame];
		load_tile(F.src[J.component], F.src_pitch[J.component], F.src_step[J.component], s_ymap, s_xmap, s_lut, tid,
		          vals);
	}
	else if (J.src_kind == kStoreF16)
		load_tile(static_cast<const __half *>(J.src), J.src_pitch, 1, s_ymap, s_xmap, s_lut, tid, vals);
	else
		load_tile(static_cast<const float *>(J.src), J.src_pitch, 1, s_ymap, s_xmap, s_lut, tid, vals);
	store_tile(sA, tid, vals);
	__syncthreads();

	// Horizontal pass: results stay in registers until every thread has read
	// its windows, then overwrite the tile in place (low half | high half).
	float h[kPass1Rounds][8];
#pragma unroll
	for (int t = 0; t < kPass1Rounds; t++)
	{
		const int task = tid + t * kDwtThreads;
		if (t < kPass1Rounds - 1 || task < kPass1Tasks)
		{
			const int r = task % kTileIn, seg = task / kTileIn;
			float v[16];
#pragma unroll
			for (int j = 0; j < 16; j++)
				v[j] = sA[r * kPitchA + seg * 8 + j];
			lift_fwd16(v);
#pragma unroll
			for (int k = 0; k < 4; k++)
			{
				h[t][k] = v[4 + 2 * k] * kInvK;
				h[t][4 + k] = v[5 + 2 * k] * kK;
			}
		}
	}
	__syncthreads();
#pragma unroll
	for (int t = 0; t < kPass1Rounds; t++)
	{
		const int task = tid + t * kDwtThreads;
		if (t < kPass1Rounds - 1 || task < kPass1Tasks)
		{
			const int r = task % kTileIn, seg = task / kTileIn;
#pragma unroll
			for (int k = 0; k < 4; k++)
			{
				sA[r * kPitchA + seg * 4 + k] = h[t][k];
				sA[r * kPitchA + 32 + seg * 4 + k] = h[t][4 + k];
			}
		}
	}
	__syncthreads();

	// Vertical pass: 64 columns x 8 segments; lanes are consecutive columns.
#pragma unroll
	for (int t = 0; t < 2; t++)
	{
		const int task = tid + t * kDwtThreads;
		const int c = task & 63, seg = task >> 6;
		float v[16];
#pragma unroll
		for (int j = 0; j < 16; j++)
			v[j] = sA[(seg * 8 + j) * kPitchA + c];
		lift_fwd16(v);
		const int ox = (x0 >> 1) + (c & 31);
		if (ox < J.bw)
		{
			void *lo = c < 32 ? J.band[0] : J.band[1];
			void *hi = c < 32 ? J.band[2] : J.band[3];
			const int lo_kind = c < 32 ? J.ll_kind : J.band_kind;
#pragma unroll
			for (int k = 0; k < 4; k++)
			{
				const int oy = (y0 >> 1) + seg * 4 + k;
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

__device__ __forceinline__ uint8_t unorm8_store(float v)
{
	// +0.5 DC shift, clamp to [0,1], round to nearest (reference
	// shaders/idwt.comp:207-211 with an R8_UNORM store).
	v = v + 0.5f;
	v = v > 0.0f ? v : 0.0f;
	v = v < 1.0f ? v : 1.0f;
	return (uint8_t)(int)fmaf(v, 255.0f, 0.5f);
}

// Loads the four bands' 36x36 shares of the interleaved 72x72 tile (band b
// holds the positions whose (row, column) parity is (b >> 1, b & 1)) into
// LDS; every global load is issued before the first LDS store.
template <typename TH>
__device__ __forceinline__ void load_bands(const DwtJob &J, const int *ymap, const int *xmap, int tid, float *sA)
{
	constexpr int kQ = kTileIn / 2;
	constexpr int kIters = (kQ * kQ + kDwtThreads - 1) / kDwtThreads;
	const long pitch = J.band_pitch;
	const float *ll = static_cast<const float *>(J.band[0]);
	const TH *hb[3] = {static_cast<const TH *>(J.band[1]), static_cast<const TH *>(J.band[2]),
	                   static_cast<const TH *>(J.band[3])};
	float vals[4][kIters];
#pragma unroll
	for (int k = 0; k < kIters; k++)
	{
		const int i = tid + k * kDwtThreads;
#pragma unroll
		for (int b = 0; b < 4; b++)
			vals[b][k] = 0.0f;
		if (i < kQ * kQ)
		{
			const int rr = i / kQ, cc = i - rr * kQ;
			const long row0 = (long)(ymap[2 * rr] >> 1) * pitch, row1 = (long)(ymap[2 * rr + 1] >> 1) * pitch;
			const int col0 = xmap[2 * cc] >> 1, col1 = xmap[2 * cc + 1] >> 1;
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
		const int i = tid + k * kDwtThreads;
		if (i < kQ * kQ)
		{
			const int rr = i / kQ, cc = i - rr * kQ;
#pragma unroll
			for (int b = 0; b < 4; b++)
				sA[(2 * rr + (b >> 1)) * kPitchA + 2 * cc + (b & 1)] = vals[b][k];
		}
	}
}

__global__ __launch_bounds__(kDwtThreads) void k_dwt_inverse(const DwtJob *__restrict__ jobs,
                                                             const uint16_t *__restrict__ tile_job,
                                                             const FrameState *__restrict__ frames, int nxcd)
{
	__shared__ float sA[kTileIn * kPitchA];
	__shared__ int s_xmap[kTileIn];
	__shared__ int s_ymap[kTileIn];

	const int tid = threadIdx.x;
	const int tile = xcd_swizzle((int)blockIdx.x, (int)gridDim.x, nxcd);
	const DwtJob J = jobs[tile_job[tile]];
	const int local = tile - J.tile_base;
	const int ox0 = (local % J.tiles_x) * kTile;
	const int oy0 = (local / J.tiles_x) * kTile;

	if (tid < kTileIn)
		s_xmap[tid] = inv_ext(ox0 - kApron + tid, J.A_x);
	else if (tid < 2 * kTileIn)
		s_ymap[tid - kTileIn] = inv_ext(oy0 - kApron + tid - kTileIn, J.A_y);
	__syncthreads();

	// Interleaved position parity selects the band: LL (even, even), HL (x
	// odd), LH (y odd), HH.  The extension preserves parity.
	if (J.band_kind == kStoreF16)
		load_bands<__half>(J, s_ymap, s_xmap, tid, sA);
	else
		load_bands<float>(J, s_ymap, s_xmap, tid, sA);
	__syncthreads();

	// Vertical pass over all 72 columns; rows of the tile are windows 4..11.
	float vout[kPass1Rounds][8];
#pragma unroll
	for (int t = 0; t < kPass1Rounds; t++)
	{
		const int task = tid + t * kDwtThreads;
		if (t < kPass1Rounds - 1 || task < kPass1Tasks)
		{
			const int c = task % kTileIn, seg = task / kTileIn;
			float v[16];
#pragma unroll
			for (int j = 0; j < 16; j++)
				v[j] = sA[(seg * 8 + j) * kPitchA + c] * ((j & 1) ? kInvK : kK);
			lift_inv16(v);
#pragma unroll
			for (int j = 0; j < 8; j++)
				vout[t][j] = v[4 + j];
		}
	}
	__syncthreads();
#pragma unroll
	for (int t = 0; t < kPass1Rounds; t++)
	{
		const int task = tid + t * kDwtThreads;
		if (t < kPass1Rounds - 1 || task < kPass1Tasks)
		{
			const int c = task % kTileIn, seg = task / kTileIn;
#pragma unroll
			for (int j = 0; j < 8; j++)
				sA[(seg * 8 + j) * kPitchA + c] = vout[t][j];
		}
	}
	__syncthreads();

	// Horizontal pass: 8 lanes per row write 64 consecutive outputs.
	uint8_t *plane = nullptr;
	long dpitch = J.dst_pitch;
	if (J.dst_kind == kStoreU8)
	{
		plane = frames[J.frame].dst[J.component];
		dpitch = frames[J.frame].dst_pitch[J.component];
	}
#pragma unroll
	for (int t = 0; t < 2; t++)
	{
		const int task = tid + t * kDwtThreads;
		const int seg = task & 7, row = task >> 3;
		float v[16];
#pragma unroll
		for (int j = 0; j < 16; j++)
			v[j] = sA[row * kPitchA + seg * 8 + j] * ((j & 1) ? kInvK : kK);
		lift_inv16(v);
		const int py = oy0 + row;
		if (J.dst_kind == kStoreU8)
		{
			if (py < J.R_y)
			{
				const int px0 = ox0 + seg * 8;
				uint8_t *dst = plane + (long)py * dpitch + px0;
				if (px0 + 8 <= J.R_x && ((uintptr_t)dst & 7) == 0)
				{
					uint32_t lo = 0, hi = 0;
#pragma unroll
					for (int j = 0; j < 4; j++)
					{
						lo |= (uint32_t)unorm8_store(v[4 + j]) << (8 * j);
						hi |= (uint32_t)unorm8_store(v[8 + j]) << (8 * j);
					}
					*reinterpret_cast<uint2 *>(dst) = make_uint2(lo, hi);
				}
				else
				{
#pragma unroll
					for (int j = 4; j < 12; j++)
					{
						if (px0 + j - 4 < J.R_x)
							dst[j - 4] = unorm8_store(v[j]);
					}
				}
			}
		}
		else if (py < J.A_y)
		{
#pragma unroll
			for (int j = 4; j < 12; j++)
			{
				const int px = ox0 + seg * 8 + j - 4;
				if (px < J.A_x)
					store_kind(J.dst, J.dst_kind, (long)py * dpitch + px, v[j]);
			}
		}
	}
}
} // namespace

void launch_dwt_forward(const DwtJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                        bool fp16, hipStream_t stream)
{
	(void)fp16;
	if (tiles > 0)
		hipLaunchKernelGGL(k_dwt_forward, dim3(tiles), dim3(kDwtThreads), 0, stream, jobs, tile_job, frames,
		                   xcd_count());
}

void launch_dwt_inverse(const DwtJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                        bool fp16, hipStream_t stream)
{
	(void)fp16;
	if (tiles > 0)
		hipLaunchKernelGGL(k_dwt_inverse, dim3(tiles), dim3(kDwtThreads), 0, stream, jobs, tile_job, frames,
		                   xcd_count());
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

void upload_constants()
{
	float lut[256];
	for (int u = 0; u < 256; u++)
		lut[u] = (float)u / 255.0f - 0.5f;
	(void)hipMemcpyToSymbol(HIP_SYMBOL(c_u8_dc), lut, sizeof(lut));
}
} // namespace pw
