/*
 * psnr_hvs_m.c -- CPU restatement of the reference's PSNR-HVS-M-H metric.
 *
 * TEST INFRASTRUCTURE ONLY (part of the oracle; see pyrowave_oracle.h).
 *
 * Follows reference shaders/psnr_hvs_m.comp:1-225 (per 8x8 block at a stride
 * of 4: 8-point butterfly DCT rows then columns, masking effect, CSF-weighted
 * squared error above the mask, 24-bit fixed point per block) and
 * psnr.cpp:25-28, 126-212, 247-264, 321-324 (CSF and mask tables per height
 * factor 1 + i/8, block grid, total pixels, dB conversion).
 *
 * Conventions this restatement fixes where the reference leaves them to the
 * Vulkan driver (parity unpinned at the byte level against the reference GPU
 * code; pinned instead by the checks in tests/test_oracle_psnr.py):
 *   - texels outside the image read as 0 (robust image access);
 *   - texel = u8 / 255 correctly rounded (UNORM8);
 *   - no fused multiply-add; sqrt correctly rounded;
 *   - sqrt(1/8) folded to the nearest float.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "pyrowave_oracle.h"

/* shaders/psnr_hvs_m.comp:51-56 */
static const float C7 = 0.19509032201612833f, S7 = 0.9807852804032304f;
static const float C6 = 0.38268343236508984f, S6 = 0.9238795325112867f;
static const float C5 = 0.5555702330196023f, S5 = 0.8314696123025452f;

static void bfly(float *a, float *b)
{
	float pos = *a + *b, neg = *a - *b;
	*a = pos;
	*b = neg;
}

static void bfly_rot(float *a, float *b, float c, float s)
{
	float pos = *a * c + *b * s;
	float neg = *b * c - *a * s;
	*a = pos;
	*b = neg;
}

/* shaders/psnr_hvs_m.comp:58-97 */
static void dct8(float out[8], const float in[8])
{
	const float r18 = (float)0.35355339059327373; /* sqrt(1/8) */
	const float r28 = 0.5f;                       /* sqrt(2/8) */
	float s[8];
	memcpy(s, in, sizeof(s));
	bfly(&s[0], &s[7]);
	bfly(&s[1], &s[6]);
	bfly(&s[2], &s[5]);
	bfly(&s[3], &s[4]);
	bfly(&s[0], &s[3]);
	bfly(&s[1], &s[2]);
	bfly(&s[0], &s[1]);
	bfly_rot(&s[2], &s[3], C6, S6);
	bfly_rot(&s[4], &s[7], C7, S7);
	bfly_rot(&s[5], &s[6], C5, S5);
	bfly(&s[4], &s[5]);
	bfly(&s[6], &s[7]);
	bfly(&s[5], &s[6]);
	s[0] *= r18;
	s[1] *= r18;
	s[2] *= r28;
	s[3] *= r28;
	s[4] *= r28;
	s[5] *= r18;
	s[6] *= r18;
	s[7] *= -r28;
	out[0] = s[0];
	out[1] = s[4];
	out[2] = s[2];
	out[3] = s[5];
	out[4] = s[1];
	out[5] = s[6];
	out[6] = s[3];
	out[7] = s[7];
}

/* shaders/psnr_hvs_m.comp:158-175 */
void or_dct8x8(const float sig[64], float dct[64])
{
	for (int r = 0; r < 8; r++)
		dct8(dct + 8 * r, sig + 8 * r);
	for (int c = 0; c < 8; c++)
	{
		float t[8], o[8];
		for (int i = 0; i < 8; i++)
			t[i] = dct[8 * i + c];
		dct8(o, t);
		for (int i = 0; i < 8; i++)
			dct[8 * i + c] = o[i];
	}
}

static float sqr(float x)
{
	return x * x;
}

/* shaders/psnr_hvs_m.comp:104-107 */
static float vari(float E, float E2, int count)
{
	float v = E2 * (float)count - sqr(E);
	return (v > 0.0f ? v : 0.0f) / (float)(count - 1);
}

/* shaders/psnr_hvs_m.comp:109-156 */
float or_maskeff(const float sig[64], const float dct[64], const float mask_coeffs[64])
{
	float m = 0.0f;
	for (int i = 0; i < 64; i++)
		m += sqr(dct[i]) * mask_coeffs[i];
	float se2[2][2], se[2][2], se2_total = 0.0f, se_total = 0.0f;
	for (int y = 0; y < 2; y++)
		for (int x = 0; x < 2; x++)
		{
			se2[y][x] = 0.0f;
			se[y][x] = 0.0f;
			for (int sy = 0; sy < 4; sy++)
				for (int sx = 0; sx < 4; sx++)
				{
					float v = sig[8 * (4 * y + sy) + 4 * x + sx];
					se[y][x] += v;
					se2[y][x] += sqr(v);
				}
			se_total += se[y][x];
			se2_total += se2[y][x];
		}
	float pop = vari(se_total, se2_total, 64);
	if (pop != 0.0f)
		pop = (vari(se[0][0], se2[0][0], 16) + vari(se[1][0], se2[1][0], 16) + vari(se[0][1], se2[0][1], 16) +
		       vari(se[1][1], se2[1][1], 16)) /
		      pop;
	return sqrtf(m * pop) / 32.0f;
}

/* psnr.cpp:25-28 */
static float csf_fn(float cpd)
{
	return 2.6f * (0.0192f + 0.114f * cpd) * expf(-powf(0.114f * cpd, 1.1f));
}

/* psnr.cpp:156-204: tables of one height factor for an image `height` rows. */
void or_psnr_hvs_m_tables(int height, float height_factor, float csf[64], float mask[64], float inv_mask[64])
{
	const float pi = 3.14159265358979323846f;
	float nyquist_cpd = (float)height * height_factor * pi / 360.0f;
	for (int y = 0; y < 8; y++)
		for (int x = 0; x < 8; x++)
		{
			float H = ((float)x + 0.5f) / 8.0f, V = ((float)y + 0.5f) / 8.0f;
			float cpd = sqrtf(H * H + V * V);
			csf[8 * y + x] = 2.6f * csf_fn(cpd * nyquist_cpd);
		}
	float max_csf = 0.0f;
	for (int i = 0; i < 64; i++)
		max_csf = max_csf > csf[i] ? max_csf : csf[i];
	float norm = 1.0f / max_csf;
	norm *= norm;
	for (int i = 0; i < 64; i++)
		mask[i] = csf[i] * csf[i] * norm;
	mask[0] = 0.0f;
	for (int i = 0; i < 64; i++)
		inv_mask[i] = i ? 1.0f / mask[i] : 0.0f;
}

/* psnr.cpp:321-324 */
float or_psnr_height_factor(int index)
{
	return 1.0f + (float)index / 8.0f;
}

/* shaders/psnr_hvs_m.comp:177-199 */
static float dct_error(const float a[64], const float b[64], float mask, const float csf[64], const float inv[64])
{
	float total = 0.0f;
	for (int i = 0; i < 64; i++)
	{
		float e = fabsf(a[i] - b[i]);
		float masking = mask * inv[i];
		e = e - masking;
		e = e > 0.0f ? e : 0.0f;
		total += sqr(e * csf[i]);
	}
	return total;
}

static void load_block(const uint8_t *p, int pitch, int w, int h, int x0, int y0, float sig[64])
{
	for (int y = 0; y < 8; y++)
		for (int x = 0; x < 8; x++)
		{
			int px = x0 + x, py = y0 + y;
			sig[8 * y + x] = (px < w && py < h) ? (float)p[(long)py * pitch + px] / 255.0f : 0.0f;
		}
}

/* Per-block fixed-point errors summed over the image, per height factor
 * (shaders/psnr_hvs_m.comp:201-224, psnr.cpp:140-150).  Returns the total
 * pixel count the dB conversion uses (blocks x 64). */
uint64_t or_psnr_hvs_m_totals(const uint8_t *a, int pitch_a, const uint8_t *b, int pitch_b, int w, int h,
                              int nfactors, uint64_t *totals)
{
	enum
	{
		kStride = 4
	};
	float csf[16][64], mask[16][64], inv[16][64];
	if (nfactors > 16)
		nfactors = 16;
	for (int f = 0; f < nfactors; f++)
	{
		or_psnr_hvs_m_tables(h, or_psnr_height_factor(f), csf[f], mask[f], inv[f]);
		totals[f] = 0;
	}
	const int nbx = (w - 1) / kStride + 1, nby = (h - 1) / kStride + 1;
	float sa[64], sb[64], da[64], db[64];
	for (int by = 0; by < nby; by++)
		for (int bx = 0; bx < nbx; bx++)
		{
			load_block(a, pitch_a, w, h, bx * kStride, by * kStride, sa);
			load_block(b, pitch_b, w, h, bx * kStride, by * kStride, sb);
			or_dct8x8(sa, da);
			or_dct8x8(sb, db);
			for (int f = 0; f < nfactors; f++)
			{
				float ma = or_maskeff(sa, da, mask[f]);
				float mb = or_maskeff(sb, db, mask[f]);
				float m = ma > mb ? ma : mb;
				float err = dct_error(da, db, m, csf[f], inv[f]);
				totals[f] += (uint64_t)ldexpf(err, 24);
			}
		}
	return (uint64_t)nbx * (uint64_t)nby * 64u;
}

/* psnr.cpp:247-264 */
double or_psnr_hvs_m_db(uint64_t total, uint64_t total_pixels, int full_range)
{
	const double peak = full_range ? 1.0f * 1.0f : (223.0f * 223.0f) / (255.0f * 255.0f);
	return 10.0 * log10((double)total_pixels * peak / ldexp((double)total, -24));
}
