// pw_psnr.hip -- PSNR-HVS-M-H on the device.
//
// Restates reference shaders/psnr_hvs_m.comp (whose dispatch per height
// factor is psnr.cpp:126-212): every 8x8 block of the luma plane at a stride
// of 4 is transformed with an 8-point butterfly DCT (rows, then columns) in
// the reference image and the distorted image; the masking effect of each
// block weights the contrast-sensitivity tables; the CSF-weighted squared DCT
// error above the mask is summed per block and accumulated in 24-bit fixed
// point, so the totals do not depend on the order of the sums.
//
// One thread per block.  The DCTs and the block variances do not depend on
// the height factor, so each thread computes them once and then loops over
// the 16 height factors (the reference recomputes them per factor).  The
// tables of all factors sit in LDS; per-factor totals are reduced across the
// workgroup before one 64-bit atomic per factor.
//
// Arithmetic fixed here where the reference leaves it to the driver: texels
// outside the image read as 0, texel = u8 / 255 correctly rounded, no FMA
// contraction, correctly rounded sqrt (sqrtf: this HIP maps __fsqrt_rn to the
// approximate native v_sqrt_f32).
#include "pw_common.hpp"

#pragma clang fp contract(off)

namespace pw
{
namespace
{
constexpr int kPsnrThreads = 256;
constexpr int kPsnrStride = 4;

// shaders/psnr_hvs_m.comp:51-56
constexpr float kC7 = 0.19509032201612833f, kS7 = 0.9807852804032304f;
constexpr float kC6 = 0.38268343236508984f, kS6 = 0.9238795325112867f;
constexpr float kC5 = 0.5555702330196023f, kS5 = 0.8314696123025452f;
constexpr float kR18 = float(0.35355339059327373); // sqrt(1/8)
constexpr float kR28 = 0.5f;                        // sqrt(2/8)

__device__ __forceinline__ void bfly(float &a, float &b)
{
	const float pos = a + b, neg = a - b;
	a = pos;
	b = neg;
}

__device__ __forceinline__ void bfly_rot(float &a, float &b, float c, float s)
{
	const float pos = a * c + b * s;
	const float neg = b * c - a * s;
	a = pos;
	b = neg;
}

// shaders/psnr_hvs_m.comp:58-97: 8-point DCT of v[o], v[o + st], ... in place.
__device__ __forceinline__ void dct8(float *v, int o, int st)
{
	float s0 = v[o], s1 = v[o + st], s2 = v[o + 2 * st], s3 = v[o + 3 * st];
	float s4 = v[o + 4 * st], s5 = v[o + 5 * st], s6 = v[o + 6 * st], s7 = v[o + 7 * st];
	bfly(s0, s7);
	bfly(s1, s6);
	bfly(s2, s5);
	bfly(s3, s4);
	bfly(s0, s3);
	bfly(s1, s2);
	bfly(s0, s1);
	bfly_rot(s2, s3, kC6, kS6);
	bfly_rot(s4, s7, kC7, kS7);
	bfly_rot(s5, s6, kC5, kS5);
	bfly(s4, s5);
	bfly(s6, s7);
	bfly(s5, s6);
	v[o] = s0 * kR18;
	v[o + st] = s4 * kR28;
	v[o + 2 * st] = s2 * kR28;
	v[o + 3 * st] = s5 * kR18;
	v[o + 4 * st] = s1 * kR18;
	v[o + 5 * st] = s6 * kR18;
	v[o + 6 * st] = s3 * kR28;
	v[o + 7 * st] = s7 * -kR28;
}

__device__ __forceinline__ float sqr(float x)
{
	return x * x;
}

// shaders/psnr_hvs_m.comp:104-107
__device__ __forceinline__ float vari(float E, float E2, int count)
{
	const float v = E2 * (float)count - sqr(E);
	return (v > 0.0f ? v : 0.0f) / (float)(count - 1);
}

// Loads one 8x8 block (u8 / 255, zero outside the image), transforms it in
// place and returns the variance ratio of the masking effect
// (shaders/psnr_hvs_m.comp:118-153), which is independent of the tables.
__device__ __forceinline__ float load_dct(const uint8_t *p, uint32_t pitch, int w, int h, int x0, int y0,
                                          float (&v)[64])
{
	float se[2][2], se2[2][2];
#pragma unroll
	for (int y = 0; y < 8; y++)
#pragma unroll
		for (int x = 0; x < 8; x++)
		{
			const int px = x0 + x, py = y0 + y;
			v[8 * y + x] = (px < w && py < h) ? (float)p[(size_t)py * pitch + px] / 255.0f : 0.0f;
		}
	float se_total = 0.0f, se2_total = 0.0f;
#pragma unroll
	for (int qy = 0; qy < 2; qy++)
#pragma unroll
		for (int qx = 0; qx < 2; qx++)
		{
			float e = 0.0f, e2 = 0.0f;
#pragma unroll
			for (int sy = 0; sy < 4; sy++)
#pragma unroll
				for (int sx = 0; sx < 4; sx++)
				{
					const float t = v[8 * (4 * qy + sy) + 4 * qx + sx];
					e += t;
					e2 += sqr(t);
				}
			se[qy][qx] = e;
			se2[qy][qx] = e2;
			se_total += e;
			se2_total += e2;
		}
	float pop = vari(se_total, se2_total, 64);
	if (pop != 0.0f)
		pop = (vari(se[0][0], se2[0][0], 16) + vari(se[1][0], se2[1][0], 16) + vari(se[0][1], se2[0][1], 16) +
		       vari(se[1][1], se2[1][1], 16)) /
		      pop;
	// shaders/psnr_hvs_m.comp:158-175: rows, then columns.
#pragma unroll
	for (int r = 0; r < 8; r++)
		dct8(v, 8 * r, 1);
#pragma unroll
	for (int c = 0; c < 8; c++)
		dct8(v, c, 8);
	return pop;
}

__global__ __launch_bounds__(kPsnrThreads) void k_psnr_hvs_m(const uint8_t *__restrict__ a, uint32_t pitch_a,
                                                             const uint8_t *__restrict__ b, uint32_t pitch_b, int w,
                                                             int h, int nbx, int nblocks,
                                                             const float *__restrict__ tables,
                                                             unsigned long long *__restrict__ totals, XcdMap xm)
{
	// [factor][csf, mask, inv_mask][64]
	__shared__ float s_tab[kPsnrFactors * 3 * 64];
	__shared__ unsigned long long s_sum[kPsnrThreads / 64][kPsnrFactors];
	const int tid = threadIdx.x;
	for (int i = tid; i < kPsnrFactors * 3 * 64; i += kPsnrThreads)
		s_tab[i] = tables[i];
	__syncthreads();

	const int g = xcd_swizzle((int)blockIdx.x, xm) * kPsnrThreads + tid;
	float da[64], db[64];
	float pop_a = 0.0f, pop_b = 0.0f;
	const bool active = g < nblocks;
	if (active)
	{
		const int by = g / nbx, bx = g - by * nbx;
		pop_a = load_dct(a, pitch_a, w, h, bx * kPsnrStride, by * kPsnrStride, da);
		pop_b = load_dct(b, pitch_b, w, h, bx * kPsnrStride, by * kPsnrStride, db);
	}
	const int lane = tid & 63, wave = tid >> 6;
	for (int f = 0; f < kPsnrFactors; f++)
	{
		unsigned long long fixed = 0;
		if (active)
		{
			const float *csf = s_tab + f * 192, *mask = csf + 64, *inv = csf + 128;
			// shaders/psnr_hvs_m.comp:111-114, 155
			float ma = 0.0f, mb = 0.0f;
#pragma unroll
			for (int i = 0; i < 64; i++)
			{
				ma += sqr(da[i]) * mask[i];
				mb += sqr(db[i]) * mask[i];
			}
			ma = sqrtf(ma * pop_a) / 32.0f;
			mb = sqrtf(mb * pop_b) / 32.0f;
			const float m = ma > mb ? ma : mb;
			// shaders/psnr_hvs_m.comp:177-199
			float err = 0.0f;
#pragma unroll
			for (int i = 0; i < 64; i++)
			{
				float e = fabsf(da[i] - db[i]);
				const float masking = m * inv[i];
				e = e - masking;
				e = e > 0.0f ? e : 0.0f;
				err += sqr(e * csf[i]);
			}
			fixed = (unsigned long long)ldexpf(err, 24);
		}
		// Integer sums: the order of the reduction does not matter.
#pragma unroll
		for (int off = 32; off > 0; off >>= 1)
			fixed += __shfl_xor(fixed, off);
		if (lane == 0)
			s_sum[wave][f] = fixed;
	}
	__syncthreads();
	if (tid < kPsnrFactors)
	{
		unsigned long long t = 0;
		for (int wv = 0; wv < kPsnrThreads / 64; wv++)
			t += s_sum[wv][tid];
		if (t)
			atomicAdd(&totals[tid], t);
	}
}
} // namespace

void launch_psnr_hvs_m(const uint8_t *a, uint32_t pitch_a, const uint8_t *b, uint32_t pitch_b, int w, int h,
                       const float *tables, uint64_t *totals, hipStream_t stream)
{
	const int nbx = (w - 1) / kPsnrStride + 1, nby = (h - 1) / kPsnrStride + 1;
	const int nblocks = nbx * nby;
	const int groups = (nblocks + kPsnrThreads - 1) / kPsnrThreads;
	hipLaunchKernelGGL(k_psnr_hvs_m, dim3(groups), dim3(kPsnrThreads), 0, stream, a, pitch_a, b, pitch_b, w, h, nbx,
	                   nblocks, tables, reinterpret_cast<unsigned long long *>(totals), xcd_map(dim3(groups)));
}
} // namespace pw
