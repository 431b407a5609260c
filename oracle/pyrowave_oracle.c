/*
 * pyrowave_oracle.c -- CPU restatement of the PyroWave encode/decode hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pyrowave_oracle.h).  Built with
 * -ffp-contract=off so that every float operation rounds exactly where it is
 * written; fmaf() is used where a fused multiply-add is meant.
 *
 * Reference files cited below are relative to the Themaister/pyrowave tree.
 */
#include "pyrowave_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Constants                                                                */
/* ------------------------------------------------------------------------ */

/* CDF 9/7 lifting constants: shaders/dwt_common.h:37-42. */
#define OR_ALPHA ((float)-1.586134342059924)
#define OR_BETA ((float)-0.052980118572961)
#define OR_GAMMA ((float)0.882911075530934)
#define OR_DELTA ((float)0.443506852043971)
#define OR_K ((float)1.230174104914001)
#define OR_INV_K ((float)(1.0 / 1.230174104914001))

static const double D_ALPHA = -1.586134342059924;
static const double D_BETA = -0.052980118572961;
static const double D_GAMMA = 0.882911075530934;
static const double D_DELTA = 0.443506852043971;
static const double D_K = 1.230174104914001;

/* Rate control: pyrowave_encoder.cpp:17-19. */
enum { SUBDIV = 16, NUM_BUCKETS = 128, FLAT = NUM_BUCKETS * SUBDIV };

/* Smallest floats >= 2^0.25 and 2^0.75.  floor(60.5 + 2*log2(r)) for r = m*2^E,
 * m in [1,2), equals 60 + 2E + [m >= 2^0.25] + [m >= 2^0.75] exactly. */
static const float BUCKET_T1 = 0x1.306fe2p+0f;
static const float BUCKET_T2 = 0x1.ae89fap+0f;

static int align_int(int v, int a) { return (v + a - 1) & ~(a - 1); }
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
static int find_msb(uint32_t v) { return v ? 31 - __builtin_clz(v) : -1; }
static int popcount(uint32_t v) { return __builtin_popcount(v); }

/* ------------------------------------------------------------------------ */
/* Layout: pyrowave_common.cpp:191-254, bitstream/bitstream.md:99-152,487-511 */
/* ------------------------------------------------------------------------ */

long or_layout_size(void) { return (long)sizeof(or_layout); }

static int next_pow2(int v)
{
	int p = 1;
	while (p < v)
		p <<= 1;
	return p;
}

int or_layout_init(or_layout *L, int width, int height, int chroma)
{
	if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
		return -1;
	if (chroma != 0 && chroma != 1)
		return -1;
	if (chroma == 0 && ((width & 1) || (height & 1)))
		return -1;

	memset(L, 0, sizeof(*L));
	L->width = width;
	L->height = height;
	L->chroma = chroma;
	L->aligned_w = imax(align_int(width, 32), 128);
	L->aligned_h = imax(align_int(height, 32), 128);

	/* Flat coefficient storage for every band the transform produces. */
	long off = 0;
	for (int c = 0; c < OR_COMPONENTS; c++)
	{
		for (int l = 0; l < OR_LEVELS; l++)
		{
			if (l == 0 && c != 0 && chroma == 0)
				continue;
			for (int b = 0; b < OR_BANDS; b++)
			{
				or_band *B = &L->band[c][l][b];
				B->bw = (L->aligned_w >> 1) >> l;
				B->bh = (L->aligned_h >> 1) >> l;
				B->bx8 = (B->bw + 7) / 8;
				B->by8 = (B->bh + 7) / 8;
				B->bx32 = (B->bw + 31) / 32;
				B->by32 = (B->bh + 31) / 32;
				B->coeff_off = off;
				off += (long)B->bw * B->bh;
			}
		}
	}
	L->coeff_count = off;

	/* Block index order: level 4 -> 0, component, band (LL only at level 4). */
	for (int l = OR_LEVELS - 1; l >= 0; l--)
	{
		for (int c = 0; c < OR_COMPONENTS; c++)
		{
			if (l == 0 && c != 0 && chroma == 0)
				continue;
			for (int b = (l == OR_LEVELS - 1 ? 0 : 1); b < OR_BANDS; b++)
			{
				or_band *B = &L->band[c][l][b];
				B->present = 1;
				B->off8 = L->count8;
				B->off32 = L->count32;
				L->count8 += B->bx8 * B->by8;
				L->count32 += B->bx32 * B->by32;
			}
		}
	}

	/* pyrowave_encoder.cpp:21-26. */
	L->per_subdivision = next_pow2(align_int(L->count32, SUBDIV) / SUBDIV);
	return 0;
}

int or_block_descs(const or_layout *L, or_block_desc *out)
{
	for (int c = 0; c < OR_COMPONENTS; c++)
		for (int l = 0; l < OR_LEVELS; l++)
			for (int b = 0; b < OR_BANDS; b++)
			{
				const or_band *B = &L->band[c][l][b];
				if (!B->present)
					continue;
				for (int y = 0; y < B->by32; y++)
					for (int x = 0; x < B->bx32; x++)
					{
						or_block_desc *d = &out[B->off32 + y * B->bx32 + x];
						d->component = (int16_t)c;
						d->level = (int16_t)l;
						d->band = (int16_t)b;
						d->pad = 0;
						d->bx = x;
						d->by = y;
					}
			}
	return L->count32;
}

/* ------------------------------------------------------------------------ */
/* Quantizer constants: pyrowave_encoder.cpp:107-172, pyrowave_common.hpp:124-145 */
/* ------------------------------------------------------------------------ */

static float noise_power_normalized_quant_resolution(int level, int component, int band, int precision0)
{
	/* pyrowave_encoder.cpp:152-172: 8 base bits for precision >= 1, 6 for
	 * precision 0. */
	int bits = precision0 ? 6 : 8;
	if (band == 0)
		bits += 2;
	else if (band < 3)
		bits += 1;
	bits += level;
	if (component != 0)
		bits--;
	return (float)(1 << bits);
}

float or_quant_resolution_policy(int level, int component, int band, int precision0)
{
	/* pyrowave_encoder.cpp:144-150: cap 4096 for precision >= 1, 512 for
	 * precision 0 ("FP16 range is limited"). */
	float r = noise_power_normalized_quant_resolution(level, component, band, precision0);
	float cap = precision0 ? 512.0f : 4096.0f;
	return r < cap ? r : cap;
}

float or_quant_resolution(int level, int component, int band)
{
	return or_quant_resolution_policy(level, component, band, 0);
}

float or_rdo_distortion_scale(const or_layout *L, int level, int component, int band)
{
	/* pyrowave_encoder.cpp:107-142, then * (1/256) at pyrowave_encoder.cpp:459. */
	float horiz_midpoint = (band & 1) ? 0.75f : 0.25f;
	float vert_midpoint = (band & 2) ? 0.75f : 0.25f;
	const float dpi = 96.0f;
	const float viewing_distance = 1.0f;
	const float cpd_nyquist = 0.34f * viewing_distance * dpi;
	float cpd = sqrtf(horiz_midpoint * horiz_midpoint + vert_midpoint * vert_midpoint) *
	            cpd_nyquist * exp2f(-(float)level);
	if (cpd < 8.0f)
		cpd = 8.0f;
	float csf = 2.6f * (0.0192f + 0.114f * cpd) * expf(-powf(0.114f * cpd, 1.1f));
	if (component != 0 && level != OR_LEVELS - 1)
	{
		if (L->chroma == 0)
			csf *= 0.6f;
	}
	float resolution = noise_power_normalized_quant_resolution(level, component, band, L->precision0);
	float weighted = csf * resolution;
	return (weighted * weighted) * (1.0f / 256.0f);
}

uint8_t or_encode_quant(float decoder_q_scale)
{
	uint32_t v;
	memcpy(&v, &decoder_q_scale, 4);
	int e = (int)((v >> 23) & 0xff) - 127 - 4;
	int m = (int)((v >> 20) & 0x7);
	e = -e;
	return (uint8_t)((e << 3) | m);
}

float or_decode_quant(uint8_t code)
{
	/* pyrowave_common.hpp:126-133 and wavelet_dequant.comp:168-175.  Codes with
	 * an exponent field above 24 are clamped so the shift stays defined. */
	int ef = code >> 3;
	if (ef > 24)
		ef = 24;
	int e = 4 - ef;
	int m = code & 0x7;
	return (1.0f / (8.0f * 1024.0f * 1024.0f)) * (float)((8 + m) * (1 << (20 + e)));
}

/* ------------------------------------------------------------------------ */
/* fp16                                                                     */
/* ------------------------------------------------------------------------ */

uint16_t or_f32_to_f16(float f)
{
	uint32_t x;
	memcpy(&x, &f, 4);
	uint32_t sign = (x >> 16) & 0x8000u;
	uint32_t ax = x & 0x7fffffffu;
	if (ax >= 0x7f800000u)
		return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
	int e = (int)(ax >> 23) - 127;
	uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
	if (ax == 0)
		return (uint16_t)sign;
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
			return (uint16_t)(sign | 0x7c00u);
		return (uint16_t)(sign | ((uint32_t)he << 10) | (hm & 0x3ffu));
	}
	else
	{
		int shift = -e - 1;
		if (shift > 24)
			return (uint16_t)sign;
		uint32_t hm = mant >> shift;
		uint32_t rem = mant & ((1u << shift) - 1u);
		uint32_t halfway = 1u << (shift - 1);
		if (rem > halfway || (rem == halfway && (hm & 1u)))
			hm++;
		return (uint16_t)(sign | hm);
	}
}

float or_f16_to_f32(uint16_t h)
{
	uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
	uint32_t e = (h >> 10) & 0x1f;
	uint32_t m = h & 0x3ff;
	uint32_t x;
	if (e == 0)
	{
		if (m == 0)
			x = sign;
		else
		{
			/* subnormal: normalize */
			int s = 0;
			while (!(m & 0x400u))
			{
				m <<= 1;
				s++;
			}
			m &= 0x3ffu;
			x = sign | ((uint32_t)(127 - 15 + 1 - s) << 23) | (m << 13);
		}
	}
	else if (e == 31)
		x = sign | 0x7f800000u | (m << 13);
	else
		x = sign | ((e + 127 - 15) << 23) | (m << 13);
	float f;
	memcpy(&f, &x, 4);
	return f;
}

/* The kernels' shortcut for float(half(x)) on 0 <= x <= 60000: round the
 * significand to 10 bits in the normal range, to a multiple of 2^-24 below.
 * or_check_round_f16_fast() compares it with the general conversion for every
 * float in that range. */
float or_round_f16_fast(float x)
{
	if (x < 0x1p-14f)
		return rintf(x * 0x1p24f) * 0x1p-24f;
	uint32_t b;
	memcpy(&b, &x, 4);
	b = (b + 0xfffu + ((b >> 13) & 1u)) & ~0x1fffu;
	float r;
	memcpy(&r, &b, 4);
	return r;
}

long or_check_round_f16_fast(void)
{
	long mismatches = 0;
	float lim = 60000.0f;
	uint32_t end;
	memcpy(&end, &lim, 4);
	for (uint32_t b = 0; b <= end; b++)
	{
		float x;
		memcpy(&x, &b, 4);
		float a = or_round_f16_fast(x);
		float r = or_f16_to_f32(or_f32_to_f16(x));
		uint32_t ab, rb;
		memcpy(&ab, &a, 4);
		memcpy(&rb, &r, 4);
		if (ab != rb)
			mismatches++;
	}
	return mismatches;
}

void or_round_to_f16_array(float *v, long n)
{
	for (long i = 0; i < n; i++)
		v[i] = or_f16_to_f32(or_f32_to_f16(v[i]));
}

float or_round_f64_to_f16(double d)
{
	/* Round to nearest even, subnormals kept (the magnitudes here stay far
	 * from the overflow threshold). */
	if (d == 0.0 || d != d)
		return (float)d;
	double a = fabs(d);
	int e;
	frexp(a, &e); /* a = m * 2^e, m in [0.5, 1) */
	/* fp16 spacing: 2^(e - 11) in the normal range, 2^-24 below 2^-14. */
	int qe = e - 11;
	if (qe < -24)
		qe = -24;
	double q = ldexp(a, -qe);
	double r = nearbyint(q); /* default rounding mode: to nearest even */
	double out = ldexp(r, qe);
	if (out > 65504.0)
		out = INFINITY;
	return (float)(d < 0 ? -out : out);
}

/* ------------------------------------------------------------------------ */
/* Forward DWT: shaders/dwt.comp                                            */
/* ------------------------------------------------------------------------ */

static int mirror_repeat(int t, int R)
{
	/* VK_SAMPLER_ADDRESS_MODE_MIRRORED_REPEAT on texel index t, size R. */
	int period = 2 * R;
	int m = t % period;
	if (m < 0)
		m += period;
	return m < R ? m : period - 1 - m;
}

int or_fwd_ext(int p, int R, int A)
{
	/* dwt.comp:27-36 plus the 2x2 textureGather footprint: the gather for the
	 * even coordinate c covers samples c and c+1. */
	int c = p & ~1;
	int coord = c < 0 ? c : c + 1;
	if (coord >= 2 * A - R)
		coord = coord + 2 * (R - A) + 1;
	else if (coord > R)
		coord = R;
	return mirror_repeat(coord - 1 + (p & 1), R);
}

/* Extended line of A+8 samples for positions [-4, A+4). */
#define APRON 4

static void lift_fwd_f32(float *e, int A)
{
	/* dwt.comp:108-115; e[i + APRON] holds position i. */
	float *x = e + APRON;
	for (int p = -3; p <= A + 1; p += 2)
		x[p] = fmaf(OR_ALPHA, x[p - 1] + x[p + 1], x[p]);
	for (int p = -2; p <= A; p += 2)
		x[p] = fmaf(OR_BETA, x[p - 1] + x[p + 1], x[p]);
	for (int p = -1; p <= A - 1; p += 2)
		x[p] = fmaf(OR_GAMMA, x[p - 1] + x[p + 1], x[p]);
	for (int p = 0; p <= A - 2; p += 2)
		x[p] = fmaf(OR_DELTA, x[p - 1] + x[p + 1], x[p]);
}

static void lift_fwd_f64(double *e, int A)
{
	double *x = e + APRON;
	for (int p = -3; p <= A + 1; p += 2)
		x[p] += D_ALPHA * (x[p - 1] + x[p + 1]);
	for (int p = -2; p <= A; p += 2)
		x[p] += D_BETA * (x[p - 1] + x[p + 1]);
	for (int p = -1; p <= A - 1; p += 2)
		x[p] += D_GAMMA * (x[p - 1] + x[p + 1]);
	for (int p = 0; p <= A - 2; p += 2)
		x[p] += D_DELTA * (x[p - 1] + x[p + 1]);
}

/* FP16 arithmetic (precision 0, dwt_common.h:29-34 FLOAT = float16_t): every
 * value is an fp16 number held in a float.  A lifting step is an fp16 add
 * followed by a fused fp16 multiply-add with the fp16-rounded constant (both
 * evaluated exactly in double, then rounded once: the products of two fp16
 * numbers and the sums here fit in a double's significand).  The K / 1/K
 * scaling multiplies by the FP32 constant and rounds the exact product once
 * to fp16 (an fp16 x fp32 product has 35 significant bits). */
static float h_add(float a, float b) { return or_round_f64_to_f16((double)a + (double)b); }
static float h_fma(float c, float s, float v) { return or_round_f64_to_f16((double)c * (double)s + (double)v); }
static float h_scale(float v, float k) { return or_round_f64_to_f16((double)v * (double)k); }
static float h_const(float c) { return or_f16_to_f32(or_f32_to_f16(c)); }

static void lift_fwd_h(float *e, int A)
{
	const float a = h_const(OR_ALPHA), b = h_const(OR_BETA), g = h_const(OR_GAMMA), d = h_const(OR_DELTA);
	float *x = e + APRON;
	for (int p = -3; p <= A + 1; p += 2)
		x[p] = h_fma(a, h_add(x[p - 1], x[p + 1]), x[p]);
	for (int p = -2; p <= A; p += 2)
		x[p] = h_fma(b, h_add(x[p - 1], x[p + 1]), x[p]);
	for (int p = -1; p <= A - 1; p += 2)
		x[p] = h_fma(g, h_add(x[p - 1], x[p + 1]), x[p]);
	for (int p = 0; p <= A - 2; p += 2)
		x[p] = h_fma(d, h_add(x[p - 1], x[p + 1]), x[p]);
}

void or_dwt_level_h(const float *in, int in_pitch, int R_x, int R_y, int A_x, int A_y,
                    float *ll, float *hl, float *lh, float *hh, int out_pitch)
{
	/* or_dwt_level_f32 in fp16 arithmetic; `in` holds fp16 values. */
	int W2 = A_x / 2, H2 = A_y / 2;
	float *hrows = (float *)malloc(sizeof(float) * (size_t)A_x * R_y);
	float *line = (float *)malloc(sizeof(float) * (size_t)(imax(A_x, A_y) + 2 * APRON));
	for (int y = 0; y < R_y; y++)
	{
		const float *row = in + (size_t)y * in_pitch;
		for (int p = -APRON; p < A_x + APRON; p++)
			line[p + APRON] = row[or_fwd_ext(p, R_x, A_x)];
		lift_fwd_h(line, A_x);
		float *o = hrows + (size_t)y * A_x;
		for (int k = 0; k < W2; k++)
		{
			o[k] = h_scale(line[APRON + 2 * k], OR_INV_K);
			o[W2 + k] = h_scale(line[APRON + 2 * k + 1], OR_K);
		}
	}
	for (int x = 0; x < A_x; x++)
	{
		for (int p = -APRON; p < A_y + APRON; p++)
			line[p + APRON] = hrows[(size_t)or_fwd_ext(p, R_y, A_y) * A_x + x];
		lift_fwd_h(line, A_y);
		float *lo = x < W2 ? ll : hl;
		float *hi = x < W2 ? lh : hh;
		int ox = x < W2 ? x : x - W2;
		for (int k = 0; k < H2; k++)
		{
			lo[(size_t)k * out_pitch + ox] = h_scale(line[APRON + 2 * k], OR_INV_K);
			hi[(size_t)k * out_pitch + ox] = h_scale(line[APRON + 2 * k + 1], OR_K);
		}
	}
	free(hrows);
	free(line);
}

void or_dwt_level_f32(const float *in, int in_pitch, int R_x, int R_y, int A_x, int A_y,
                      float *ll, float *hl, float *lh, float *hh, int out_pitch)
{
	/* Horizontal pass over every (extended) row, then vertical pass over
	 * columns (dwt.comp:195-234).  Rows outside [0, A_y) are the mapped rows, so
	 * the horizontal result of row r is that of row ext(r). */
	int W2 = A_x / 2, H2 = A_y / 2;
	float *hrows = (float *)malloc(sizeof(float) * (size_t)A_x * R_y);
	float *line = (float *)malloc(sizeof(float) * (size_t)(imax(A_x, A_y) + 2 * APRON));
	int *xmap = (int *)malloc(sizeof(int) * (size_t)(A_x + 2 * APRON));
	for (int p = -APRON; p < A_x + APRON; p++)
		xmap[p + APRON] = or_fwd_ext(p, R_x, A_x);

	for (int y = 0; y < R_y; y++)
	{
		const float *row = in + (size_t)y * in_pitch;
		for (int p = -APRON; p < A_x + APRON; p++)
			line[p + APRON] = row[xmap[p + APRON]];
		lift_fwd_f32(line, A_x);
		float *o = hrows + (size_t)y * A_x;
		for (int k = 0; k < W2; k++)
		{
			o[k] = line[APRON + 2 * k] * OR_INV_K;
			o[W2 + k] = line[APRON + 2 * k + 1] * OR_K;
		}
	}

	for (int x = 0; x < A_x; x++)
	{
		for (int p = -APRON; p < A_y + APRON; p++)
			line[p + APRON] = hrows[(size_t)or_fwd_ext(p, R_y, A_y) * A_x + x];
		lift_fwd_f32(line, A_y);
		float *lo = x < W2 ? ll : hl;
		float *hi = x < W2 ? lh : hh;
		int ox = x < W2 ? x : x - W2;
		for (int k = 0; k < H2; k++)
		{
			lo[(size_t)k * out_pitch + ox] = line[APRON + 2 * k] * OR_INV_K;
			hi[(size_t)k * out_pitch + ox] = line[APRON + 2 * k + 1] * OR_K;
		}
	}
	free(hrows);
	free(line);
	free(xmap);
}

void or_dwt_level_f64(const double *in, int in_pitch, int R_x, int R_y, int A_x, int A_y,
                      double *ll, double *hl, double *lh, double *hh, int out_pitch)
{
	int W2 = A_x / 2, H2 = A_y / 2;
	double *hrows = (double *)malloc(sizeof(double) * (size_t)A_x * R_y);
	double *line = (double *)malloc(sizeof(double) * (size_t)(imax(A_x, A_y) + 2 * APRON));
	const double inv_k = 1.0 / D_K;
	for (int y = 0; y < R_y; y++)
	{
		const double *row = in + (size_t)y * in_pitch;
		for (int p = -APRON; p < A_x + APRON; p++)
			line[p + APRON] = row[or_fwd_ext(p, R_x, A_x)];
		lift_fwd_f64(line, A_x);
		double *o = hrows + (size_t)y * A_x;
		for (int k = 0; k < W2; k++)
		{
			o[k] = line[APRON + 2 * k] * inv_k;
			o[W2 + k] = line[APRON + 2 * k + 1] * D_K;
		}
	}
	for (int x = 0; x < A_x; x++)
	{
		for (int p = -APRON; p < A_y + APRON; p++)
			line[p + APRON] = hrows[(size_t)or_fwd_ext(p, R_y, A_y) * A_x + x];
		lift_fwd_f64(line, A_y);
		double *lo = x < W2 ? ll : hl;
		double *hi = x < W2 ? lh : hh;
		int ox = x < W2 ? x : x - W2;
		for (int k = 0; k < H2; k++)
		{
			lo[(size_t)k * out_pitch + ox] = line[APRON + 2 * k] * inv_k;
			hi[(size_t)k * out_pitch + ox] = line[APRON + 2 * k + 1] * D_K;
		}
	}
	free(hrows);
	free(line);
}

/* ------------------------------------------------------------------------ */
/* Inverse DWT: shaders/idwt.comp                                           */
/* ------------------------------------------------------------------------ */

static int isym(int p, int N)
{
	/* Whole-sample symmetric extension of the interleaved signal about 0 and
	 * N-1 (idwt.comp:25-32: low band whole-sample left / half-sample right,
	 * high band the opposite). */
	while (p < 0 || p >= N)
	{
		if (p < 0)
			p = -p;
		if (p >= N)
			p = 2 * (N - 1) - p;
	}
	return p;
}

static void lift_inv_f32(float *e, int N)
{
	/* idwt.comp:91-108.  Scaling applied by the caller. */
	float *x = e + APRON;
	for (int p = -2; p <= N + 2; p += 2)
		x[p] = fmaf(-OR_DELTA, x[p - 1] + x[p + 1], x[p]);
	for (int p = -1; p <= N + 1; p += 2)
		x[p] = fmaf(-OR_GAMMA, x[p - 1] + x[p + 1], x[p]);
	for (int p = 0; p <= N; p += 2)
		x[p] = fmaf(-OR_BETA, x[p - 1] + x[p + 1], x[p]);
	for (int p = 1; p <= N - 1; p += 2)
		x[p] = fmaf(-OR_ALPHA, x[p - 1] + x[p + 1], x[p]);
}

static void lift_inv_f64(double *e, int N)
{
	double *x = e + APRON;
	for (int p = -2; p <= N + 2; p += 2)
		x[p] -= D_DELTA * (x[p - 1] + x[p + 1]);
	for (int p = -1; p <= N + 1; p += 2)
		x[p] -= D_GAMMA * (x[p - 1] + x[p + 1]);
	for (int p = 0; p <= N; p += 2)
		x[p] -= D_BETA * (x[p - 1] + x[p + 1]);
	for (int p = 1; p <= N - 1; p += 2)
		x[p] -= D_ALPHA * (x[p - 1] + x[p + 1]);
}

static void lift_inv_h(float *e, int N)
{
	const float a = -h_const(OR_ALPHA), b = -h_const(OR_BETA), g = -h_const(OR_GAMMA), d = -h_const(OR_DELTA);
	float *x = e + APRON;
	for (int p = -2; p <= N + 2; p += 2)
		x[p] = h_fma(d, h_add(x[p - 1], x[p + 1]), x[p]);
	for (int p = -1; p <= N + 1; p += 2)
		x[p] = h_fma(g, h_add(x[p - 1], x[p + 1]), x[p]);
	for (int p = 0; p <= N; p += 2)
		x[p] = h_fma(b, h_add(x[p - 1], x[p + 1]), x[p]);
	for (int p = 1; p <= N - 1; p += 2)
		x[p] = h_fma(a, h_add(x[p - 1], x[p + 1]), x[p]);
}

void or_idwt_level_h(const float *ll, const float *hl, const float *lh, const float *hh,
                     int in_pitch, int bw, int bh, float *out, int out_pitch)
{
	/* or_idwt_level_f32 in fp16 arithmetic; the bands hold fp16 values. */
	int W = 2 * bw, H = 2 * bh;
	float *cols = (float *)malloc(sizeof(float) * (size_t)W * H);
	float *line = (float *)malloc(sizeof(float) * (size_t)(imax(W, H) + 2 * APRON + 2));
	for (int x = 0; x < W; x++)
	{
		const float *lo = (x & 1) ? hl : ll;
		const float *hi = (x & 1) ? hh : lh;
		int bx = x >> 1;
		for (int p = -APRON; p < H + APRON; p++)
		{
			int q = isym(p, H);
			float v = (q & 1) ? hi[(size_t)(q >> 1) * in_pitch + bx] : lo[(size_t)(q >> 1) * in_pitch + bx];
			line[p + APRON] = h_scale(v, (p & 1) ? OR_INV_K : OR_K);
		}
		lift_inv_h(line, H);
		for (int y = 0; y < H; y++)
			cols[(size_t)y * W + x] = line[APRON + y];
	}
	for (int y = 0; y < H; y++)
	{
		for (int p = -APRON; p < W + APRON; p++)
		{
			float v = cols[(size_t)y * W + isym(p, W)];
			line[p + APRON] = h_scale(v, (p & 1) ? OR_INV_K : OR_K);
		}
		lift_inv_h(line, W);
		for (int x = 0; x < W; x++)
			out[(size_t)y * out_pitch + x] = line[APRON + x];
	}
	free(cols);
	free(line);
}

void or_idwt_level_f32(const float *ll, const float *hl, const float *lh, const float *hh,
                       int in_pitch, int bw, int bh, float *out, int out_pitch)
{
	/* Vertical inverse first, then horizontal (idwt.comp:182-214). */
	int W = 2 * bw, H = 2 * bh;
	float *cols = (float *)malloc(sizeof(float) * (size_t)W * H);
	float *line = (float *)malloc(sizeof(float) * (size_t)(imax(W, H) + 2 * APRON + 2));
	for (int x = 0; x < W; x++)
	{
		const float *lo = (x & 1) ? hl : ll;
		const float *hi = (x & 1) ? hh : lh;
		int bx = x >> 1;
		for (int p = -APRON; p < H + APRON; p++)
		{
			int q = isym(p, H);
			float v = (q & 1) ? hi[(size_t)(q >> 1) * in_pitch + bx] : lo[(size_t)(q >> 1) * in_pitch + bx];
			line[p + APRON] = (p & 1) ? v * OR_INV_K : v * OR_K;
		}
		lift_inv_f32(line, H);
		for (int y = 0; y < H; y++)
			cols[(size_t)y * W + x] = line[APRON + y];
	}
	for (int y = 0; y < H; y++)
	{
		for (int p = -APRON; p < W + APRON; p++)
		{
			float v = cols[(size_t)y * W + isym(p, W)];
			line[p + APRON] = (p & 1) ? v * OR_INV_K : v * OR_K;
		}
		lift_inv_f32(line, W);
		for (int x = 0; x < W; x++)
			out[(size_t)y * out_pitch + x] = line[APRON + x];
	}
	free(cols);
	free(line);
}

void or_idwt_level_f64(const double *ll, const double *hl, const double *lh, const double *hh,
                       int in_pitch, int bw, int bh, double *out, int out_pitch)
{
	int W = 2 * bw, H = 2 * bh;
	const double inv_k = 1.0 / D_K;
	double *cols = (double *)malloc(sizeof(double) * (size_t)W * H);
	double *line = (double *)malloc(sizeof(double) * (size_t)(imax(W, H) + 2 * APRON + 2));
	for (int x = 0; x < W; x++)
	{
		const double *lo = (x & 1) ? hl : ll;
		const double *hi = (x & 1) ? hh : lh;
		int bx = x >> 1;
		for (int p = -APRON; p < H + APRON; p++)
		{
			int q = isym(p, H);
			double v = (q & 1) ? hi[(size_t)(q >> 1) * in_pitch + bx] : lo[(size_t)(q >> 1) * in_pitch + bx];
			line[p + APRON] = (p & 1) ? v * inv_k : v * D_K;
		}
		lift_inv_f64(line, H);
		for (int y = 0; y < H; y++)
			cols[(size_t)y * W + x] = line[APRON + y];
	}
	for (int y = 0; y < H; y++)
	{
		for (int p = -APRON; p < W + APRON; p++)
		{
			double v = cols[(size_t)y * W + isym(p, W)];
			line[p + APRON] = (p & 1) ? v * inv_k : v * D_K;
		}
		lift_inv_f64(line, W);
		for (int x = 0; x < W; x++)
			out[(size_t)y * out_pitch + x] = line[APRON + x];
	}
	free(cols);
	free(line);
}

/* ------------------------------------------------------------------------ */
/* Whole-frame forward transform: pyrowave_encoder.cpp:490-594              */
/* ------------------------------------------------------------------------ */

static float unorm8_dc(uint8_t u)
{
	/* UNORM8 texel (u / 255) with the level-0 DC shift (dwt.comp:44-48). */
	return (float)u / 255.0f - 0.5f;
}

static float unorm8_dc_h(uint8_t u)
{
	/* Precision 0: the texel converted to fp16 (VEC4 = f16vec4), then the DC
	 * shift in fp16 (dwt.comp:44-48). */
	return h_add(or_f16_to_f32(or_f32_to_f16((float)u / 255.0f)), -0.5f);
}

static float unorm16_dc(uint16_t u)
{
	/* R16_UNORM texel (u / 65535, the quotient rounded to float once) with
	 * the level-0 DC shift: encode.cpp:152 samples 16-bit y4m input (rescaled
	 * to the full 16-bit range, yuv4mpeg.cpp:221-233) through an R16_UNORM
	 * image with the same dwt.comp:44-48 shift as UNORM8. */
	return (float)u / 65535.0f - 0.5f;
}

static float unorm16_dc_h(uint16_t u)
{
	return h_add(or_f16_to_f32(or_f32_to_f16((float)u / 65535.0f)), -0.5f);
}

float or_unorm_dc(uint32_t u, int bits, int precision)
{
	if (bits == 16)
		return precision == 0 ? unorm16_dc_h((uint16_t)u) : unorm16_dc((uint16_t)u);
	return precision == 0 ? unorm8_dc_h((uint8_t)u) : unorm8_dc((uint8_t)u);
}

static void plane_dims(const or_layout *L, int c, int *w, int *h)
{
	*w = L->width;
	*h = L->height;
	if (c != 0 && L->chroma == 0)
	{
		*w /= 2;
		*h /= 2;
	}
}

int or_dwt_frame(const or_layout *L, const uint8_t *const planes[3], const int pitches[3],
                 int precision, float *coeffs, double *coeffs64)
{
	const void *p[3] = {planes[0], planes[1], planes[2]};
	return or_dwt_frame_ex(L, p, pitches, 8, precision, coeffs, coeffs64);
}

int or_dwt_frame_ex(const or_layout *L, const void *const planes[3], const int pitches[3], int bits,
                    int precision, float *coeffs, double *coeffs64)
{
	int use_f64 = coeffs64 != NULL;
	for (int c = 0; c < OR_COMPONENTS; c++)
	{
		int first_level = (c != 0 && L->chroma == 0) ? 1 : 0;
		int pw, ph;
		plane_dims(L, c, &pw, &ph);
		for (int l = first_level; l < OR_LEVELS; l++)
		{
			int A_x = L->aligned_w >> l, A_y = L->aligned_h >> l;
			int R_x, R_y, in_pitch;
			float *in32 = NULL;
			double *in64 = NULL;
			const or_band *B = L->band[c][l];
			if (l == first_level)
			{
				R_x = pw;
				R_y = ph;
				in_pitch = pw;
				if (use_f64)
					in64 = (double *)malloc(sizeof(double) * (size_t)pw * ph);
				else
					in32 = (float *)malloc(sizeof(float) * (size_t)pw * ph);
				for (int y = 0; y < ph; y++)
					for (int x = 0; x < pw; x++)
					{
						const size_t at = (size_t)y * pitches[c] + x;
						const size_t to = (size_t)y * pw + x;
						const uint32_t u = bits == 16 ? ((const uint16_t *)planes[c])[at]
						                              : ((const uint8_t *)planes[c])[at];
						if (use_f64)
							in64[to] = (double)u / (bits == 16 ? 65535.0 : 255.0) - 0.5;
						else
							in32[to] = or_unorm_dc(u, bits, precision);
					}
			}
			else
			{
				const or_band *P = &L->band[c][l - 1][0];
				R_x = A_x;
				R_y = A_y;
				in_pitch = P->bw;
				if (use_f64)
					in64 = coeffs64 + P->coeff_off;
				else
					in32 = coeffs + P->coeff_off;
			}
			if (use_f64)
				or_dwt_level_f64(in64, in_pitch, R_x, R_y, A_x, A_y,
				                 coeffs64 + B[0].coeff_off, coeffs64 + B[1].coeff_off,
				                 coeffs64 + B[2].coeff_off, coeffs64 + B[3].coeff_off, B[0].bw);
			else if (precision == 0)
				or_dwt_level_h(in32, in_pitch, R_x, R_y, A_x, A_y,
				               coeffs + B[0].coeff_off, coeffs + B[1].coeff_off,
				               coeffs + B[2].coeff_off, coeffs + B[3].coeff_off, B[0].bw);
			else
			{
				or_dwt_level_f32(in32, in_pitch, R_x, R_y, A_x, A_y,
				                 coeffs + B[0].coeff_off, coeffs + B[1].coeff_off,
				                 coeffs + B[2].coeff_off, coeffs + B[3].coeff_off, B[0].bw);
				/* precision 3 rounds the LL too: it is the next level's input. */
				if ((precision == 1 || precision == 3) && l < 2)
					for (int b = precision == 3 ? 0 : 1; b < 4; b++)
						or_round_to_f16_array(coeffs + B[b].coeff_off, (long)B[b].bw * B[b].bh);
			}
			if (l == first_level)
			{
				free(in32);
				free(in64);
			}
		}
	}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Quantizer: shaders/wavelet_quant.comp                                    */
/* ------------------------------------------------------------------------ */

typedef struct q8x8
{
	float t[8][8];          /* scaled texels, [subblock][pixel] */
	int n[8][8];            /* magnitudes */
	int b4max[8];           /* per-subblock max magnitude */
	uint32_t scale_code;
	int maxabs, msb;        /* msb = -1 when all zero */
	uint32_t code_word;     /* code word at q = 0 */
	uint16_t cost[16];      /* errors[min(q, num_planes)].payload_cost */
	uint16_t dist[16];      /* errors[min(q, num_planes)].square_error (fp16) */
} q8x8;

static float dot4_sq(const float *v)
{
	/* dot(v, v) evaluated left to right without contraction. */
	return ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3];
}

static float butterfly8(const float *v)
{
	/* Value every lane of an 8-lane cluster holds after xor-butterfly adds with
	 * offsets 1, 2, 4. */
	return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

static void quantize_8x8(const float *band, int pitch, int bw, int bh, int x0, int y0,
                         float quant_resolution, float rdo_scale, q8x8 *Q)
{
	/* wavelet_quant.comp:292-316 (gather layout) and 172-290 (encode_payload). */
	float max_sub[8];
	for (int s = 0; s < 8; s++)
	{
		int sx = x0 + 4 * (s >> 2), sy = y0 + 2 * (s & 3);
		max_sub[s] = 0.0f;
		for (int p = 0; p < 8; p++)
		{
			int x = sx + (p >> 1), y = sy + (p & 1);
			float v = (x < bw && y < bh) ? band[(size_t)y * pitch + x] : 0.0f;
			/* Magnitudes beyond 16383.75 would need more than the 14 planes a
			 * block can describe; they are clamped (only reachable far outside
			 * the nominal coefficient range; the reference has no guard). */
			float t = fminf(fmaxf(v * quant_resolution, -16383.75f), 16383.75f);
			Q->t[s][p] = t;
			float a = fabsf(t);
			if (a > max_sub[s])
				max_sub[s] = a;
		}
	}
	float M = 0.0f;
	for (int s = 0; s < 8; s++)
		if (max_sub[s] > M)
			M = max_sub[s];

	/* compute_quant_scale (wavelet_quant.comp:95-111). */
	float quant_scale;
	if (M < 1.0f)
	{
		Q->scale_code = 6;
		quant_scale = 1.0f;
	}
	else
	{
		int e;
		frexpf(M - 0.25f, &e);
		float target_max = (float)(1 << e) - 0.25f;
		float inv_scale = M / target_max;
		Q->scale_code = (uint32_t)ceilf((inv_scale - 0.25f) * 8.0f);
		quant_scale = 1.0f / ((float)Q->scale_code / 8.0f + 0.25f);
	}

	for (int s = 0; s < 8; s++)
		for (int p = 0; p < 8; p++)
			Q->t[s][p] *= quant_scale;
	M *= quant_scale;
	for (int s = 0; s < 8; s++)
		max_sub[s] *= quant_scale;

	float overall = quant_resolution * quant_scale;
	float inv_quant = 1.0f / overall;
	float inv_quant_squared = inv_quant * inv_quant;

	for (int s = 0; s < 8; s++)
	{
		for (int p = 0; p < 8; p++)
			Q->n[s][p] = abs((int)Q->t[s][p]);
		Q->b4max[s] = (int)max_sub[s];
	}
	Q->maxabs = (int)M;

	if (Q->maxabs == 0)
	{
		Q->msb = -1;
		Q->code_word = 0;
		for (int q = 0; q < 16; q++)
		{
			Q->cost[q] = 0;
			Q->dist[q] = 0;
		}
		return;
	}

	/* With |t| <= 16383.75 the scale code keeps maxabs <= 16383, so
	 * num_planes = msb + 1 <= 14 fits the 15 BlockStats entries. */
	int msb = find_msb((uint32_t)Q->maxabs);
	Q->msb = msb;

	uint16_t cost_tab[15];
	uint16_t dist_tab[15];

	for (int q = 0; q <= msb; q++)
	{
		/* compute_quant_stats / compute_square_error (wavelet_quant.comp:113-165). */
		float se_lane[8];
		int late = 0, nsig = 0;
		for (int s = 0; s < 8; s++)
		{
			float err[8];
			for (int p = 0; p < 8; p++)
			{
				int iv = Q->n[s][p] >> q;
				float a = fabsf(Q->t[s][p]);
				float recon;
				if (iv == 0)
					recon = 0.0f;
				else
				{
					nsig++;
					recon = q == 0 ? (float)iv : ((float)iv + 0.5f) * (float)(1 << q);
				}
				err[p] = a - recon;
			}
			se_lane[s] = ((dot4_sq(err) + dot4_sq(err + 4)) * rdo_scale) * inv_quant_squared;

			int b = Q->b4max[s] >> q;
			int early = b > 0 ? 1 : 0;
			int m = msb - q;
			if (m >= 3)
			{
				int qp = m - 2;
				early = qp + 1;
				b >>= qp;
			}
			early += find_msb((uint32_t)b) + 1;
			late += imax(early - 1, 0);
		}
		cost_tab[q] = (uint16_t)(8 * late + nsig);
		if (q == 0)
			dist_tab[0] = 0;
		else
		{
			float se = butterfly8(se_lane);
			dist_tab[q] = or_f32_to_f16(se < 60000.0f ? se : 60000.0f);
		}
	}

	/* Distortion of discarding the block (wavelet_quant.comp:253-256).  The
	 * reference omits rdo_distortion_scale here. */
	{
		float lane[8];
		for (int s = 0; s < 8; s++)
			lane[s] = (dot4_sq(Q->t[s]) + dot4_sq(Q->t[s] + 4)) * inv_quant_squared;
		float se = butterfly8(lane);
		dist_tab[msb + 1] = or_f32_to_f16(se < 60000.0f ? se : 60000.0f);
		cost_tab[msb + 1] = 0;
	}

	for (int q = 0; q < 16; q++)
	{
		int idx = imin(q, msb + 1);
		Q->cost[q] = cost_tab[idx];
		Q->dist[q] = dist_tab[idx];
	}

	/* Code word at q = 0 (wavelet_quant.comp:219-229). */
	int qp0 = msb >= 3 ? msb - 2 : 0;
	uint32_t cw = (uint32_t)qp0 << 16;
	cw |= (Q->scale_code & 0xfu) << 20;
	for (int s = 0; s < 8; s++)
	{
		uint32_t plane_code = (uint32_t)(find_msb((uint32_t)(Q->b4max[s] >> qp0)) + 1);
		cw |= plane_code << (2 * s);
	}
	Q->code_word = cw;
}

static uint32_t bucket_of_ratio(float r)
{
	/* uint(max(60 + 2*log2(r) + 0.5, 0)) evaluated exactly. */
	if (!(r > 0.0f))
		return 0;
	int e;
	float f = frexpf(r, &e);
	float m = 2.0f * f;
	int E = e - 1;
	int k = 60 + 2 * E + (m >= BUCKET_T1) + (m >= BUCKET_T2);
	return k < 0 ? 0u : (uint32_t)k;
}

static void band_constants(const or_layout *L, int c, int l, int b, float *qres, float *rdo)
{
	/* pyrowave_encoder.cpp:450-459. */
	float quant_res = or_quant_resolution_policy(l, c, b, L->precision0);
	*qres = 1.0f / or_decode_quant(or_encode_quant(1.0f / quant_res));
	*rdo = or_rdo_distortion_scale(L, l, c, b);
}

uint32_t or_bucket_of_ratio(float r) { return bucket_of_ratio(r); }

void or_bucket_scan(const uint32_t raw[16], uint32_t out[16])
{
	/* inclusive_max_clustered16 (analyze_rate_control.comp:85-99): clamp to
	 * 128 - 16 + q, then the inclusive max of v[k] + (q - k) over k <= q. */
	uint32_t v[16];
	for (int q = 0; q < 16; q++)
		v[q] = raw[q] < 112u + (uint32_t)q ? raw[q] : 112u + (uint32_t)q;
	for (int q = 0; q < 16; q++)
	{
		uint32_t best = v[q];
		for (int k = 0; k < q; k++)
		{
			uint32_t cand = v[k] + (uint32_t)(q - k);
			if (cand > best)
				best = cand;
		}
		out[q] = best;
	}
}

int or_block8_stats(const or_layout *L, const float *coeffs, int c, int l, int b, int x8, int y8,
                    uint16_t cost[16], float dist[16], uint32_t *code_word, int *msb)
{
	const or_band *B = &L->band[c][l][b];
	if (!B->present || x8 < 0 || y8 < 0 || x8 >= B->bx8 || y8 >= B->by8)
		return -1;
	float qres, rdo;
	band_constants(L, c, l, b, &qres, &rdo);
	q8x8 *Q = (q8x8 *)malloc(sizeof(q8x8));
	quantize_8x8(coeffs + B->coeff_off, B->bw, B->bw, B->bh, 8 * x8, 8 * y8, qres, rdo, Q);
	for (int q = 0; q < 16; q++)
	{
		cost[q] = Q->cost[q];
		dist[q] = or_f16_to_f32(Q->dist[q]);
	}
	*code_word = Q->code_word;
	*msb = Q->msb;
	free(Q);
	return 0;
}

int or_quant_analyze(const or_layout *L, const float *coeffs,
                     uint16_t *words, uint8_t *bucket, uint32_t *code8, uint8_t *np8,
                     uint32_t *totals, uint32_t *consumed)
{
	return or_quant_analyze_ex(L, coeffs, words, bucket, code8, np8, totals, consumed, NULL);
}

int or_quant_analyze_ex(const or_layout *L, const float *coeffs,
                        uint16_t *words, uint8_t *bucket, uint32_t *code8, uint8_t *np8,
                        uint32_t *totals, uint32_t *consumed, float *dist32)
{
	memset(totals, 0, sizeof(uint32_t) * FLAT);
	*consumed = 0;
	int shamt = find_msb((uint32_t)L->per_subdivision);
	q8x8 *Q = (q8x8 *)malloc(sizeof(q8x8));

	for (int c = 0; c < OR_COMPONENTS; c++)
		for (int l = 0; l < OR_LEVELS; l++)
			for (int b = 0; b < OR_BANDS; b++)
			{
				const or_band *B = &L->band[c][l][b];
				if (!B->present)
					continue;
				float qres, rdo;
				band_constants(L, c, l, b, &qres, &rdo);
				const float *band = coeffs + B->coeff_off;

				for (int by = 0; by < B->by32; by++)
					for (int bx = 0; bx < B->bx32; bx++)
					{
						uint32_t block = (uint32_t)(B->off32 + by * B->bx32 + bx);
						uint16_t cost[16][16];
						float dist[16][16];
						for (int k = 0; k < 16; k++)
						{
							int x8 = bx * 4 + (k & 3), y8 = by * 4 + (k >> 2);
							int in_range = x8 < B->bx8 && y8 < B->by8;
							if (in_range)
							{
								quantize_8x8(band, B->bw, B->bw, B->bh, 8 * x8, 8 * y8, qres, rdo, Q);
								size_t i8 = (size_t)B->off8 + (size_t)y8 * B->bx8 + x8;
								code8[i8] = Q->code_word;
								np8[i8] = (uint8_t)(Q->msb + 1);
							}
							for (int q = 0; q < 16; q++)
							{
								cost[q][k] = in_range ? Q->cost[q] : 0;
								dist[q][k] = in_range ? or_f16_to_f32(Q->dist[q]) : 0.0f;
							}
						}

						/* analyze_rate_control.comp:169-215. */
						uint32_t w[16];
						float D[16];
						for (int q = 0; q < 16; q++)
						{
							uint32_t total = 0;
							for (int k = 0; k < 16; k++)
								total += cost[q][k] ? cost[q][k] + 24u : 0u;
							if (total)
								total += 64;
							w[q] = (total + 31) >> 5;
							const float *d = dist[q];
							D[q] = (((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7]))) +
							       (((d[8] + d[9]) + (d[10] + d[11])) + ((d[12] + d[13]) + (d[14] + d[15])));
							if (dist32)
								dist32[block * 16 + q] = D[q];
						}

						/* emit_rdo_operations (analyze_rate_control.comp:101-147). */
						uint32_t v[16];
						for (int q = 0; q < 16; q++)
						{
							uint32_t bi = 0;
							if (q != 0)
							{
								float cost_f = (float)w[q], base_f = (float)w[0];
								if (cost_f != base_f)
								{
									float dd = D[q] - D[0];
									if (dd < 0.0f)
										dd = 0.0f;
									bi = bucket_of_ratio(dd / (base_f - cost_f));
								}
							}
							v[q] = bi;
						}
						uint32_t B_q[16];
						or_bucket_scan(v, B_q);

						*consumed += w[0];
						for (int q = 0; q < 16; q++)
						{
							words[block * 16 + q] = (uint16_t)w[q];
							bucket[block * 16 + q] = (uint8_t)B_q[q];
						}
						for (int q = 1; q < 16; q++)
						{
							uint32_t saving = w[q - 1] - w[q];
							if (saving)
							{
								uint32_t subdiv = block >> shamt;
								totals[B_q[q] * SUBDIV + subdiv] += saving;
							}
						}
					}
			}
	free(Q);
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Rate control finalize / resolve                                          */
/* ------------------------------------------------------------------------ */

void or_finalize(const uint32_t *totals, uint32_t *prefix, int q1)
{
	/* analyze_rate_control_finalize.comp:14-38: 512 threads x uvec4, local
	 * inclusive scan, then Hillis-Steele over thread totals with steps
	 * 1..128 (window of 256 threads) -- quirk Q1.  q1 == 0: full prefix. */
	uint32_t tsum[512];
	for (int t = 0; t < 512; t++)
		tsum[t] = totals[4 * t] + totals[4 * t + 1] + totals[4 * t + 2] + totals[4 * t + 3];
	for (int t = 0; t < 512; t++)
	{
		int lo = q1 ? imax(0, t - 255) : 0;
		uint32_t before = 0;
		for (int u = lo; u < t; u++)
			before += tsum[u];
		uint32_t run = before;
		for (int i = 0; i < 4; i++)
		{
			run += totals[4 * t + i];
			prefix[4 * t + i] = run;
		}
	}
}

int or_resolve(const or_layout *L, const uint16_t *words, const uint8_t *bucket,
               const uint32_t *prefix, uint32_t consumed, size_t budget, int32_t *quant)
{
	/* pyrowave_encoder.cpp:283-338 and resolve_rate_control.comp:38-79. */
	size_t target = budget & ~(size_t)3;
	if (target >= 8)
		target -= 8;
	int target_words = (int)(target / 4);
	int required_base = (int)consumed - target_words;
	int per_sub = L->per_subdivision;

	memset(quant, 0, sizeof(int32_t) * (size_t)L->count32);
	for (int x = 0; x < FLAT; x++)
	{
		int required;
		if (x != 0)
		{
			int prev = (int)prefix[x - 1];
			if ((int)prefix[x] == prev)
				continue;
			required = required_base - prev;
		}
		else
		{
			if (prefix[0] == 0)
				continue;
			required = required_base;
		}
		if (required <= 0)
			continue;

		int sd = x & 15;
		uint32_t beta = (uint32_t)(x >> 4);
		uint32_t total_saved = 0;
		for (int i = 0; i < per_sub && (int)total_saved < required; i++)
		{
			int blk = sd * per_sub + i;
			if (blk >= L->count32)
				continue;
			for (int q = 1; q < 16; q++)
			{
				uint32_t saving = (uint32_t)words[blk * 16 + q - 1] - words[blk * 16 + q];
				if (saving && bucket[blk * 16 + q] == beta)
				{
					if ((int)total_saved < required && q > quant[blk])
						quant[blk] = q;
					total_saved += saving;
					break;
				}
			}
		}
	}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Block packing: shaders/block_packing.comp                                */
/* ------------------------------------------------------------------------ */

static uint32_t quantize_code_word(uint32_t cw, int quant)
{
	/* block_packing.comp:107-142. */
	if (quant != 0 && cw != 0)
	{
		int q_bits = (int)((cw >> 16) & 0xf);
		int sub = imin(q_bits, quant);
		q_bits -= sub;
		quant -= sub;
		if (quant != 0)
		{
			quant = imin(quant, 3);
			uint32_t plane0 = cw & 0x5555u;
			uint32_t plane1 = (cw & 0xaaaau) >> 1;
			uint32_t plane2 = plane0 & plane1;
			do
			{
				plane0 = plane1;
				plane1 = plane2;
				plane2 = 0;
				quant--;
			} while (quant != 0);
			plane0 &= ~plane1;
			uint32_t ncw = plane0 | (plane1 << 1);
			cw = (cw & ~0xffffu) | (ncw & 0xffffu);
		}
		cw = (cw & ~(0xfu << 16)) | ((uint32_t)q_bits << 16);
	}
	return cw;
}

static uint32_t required_8x8_size(uint32_t cw)
{
	/* block_packing.comp:97-105. */
	uint32_t q_bits = (cw >> 16) & 0xf;
	uint32_t lsbs = cw & 0x5555u;
	uint32_t msbs = cw & 0xaaaau;
	msbs |= msbs >> 1;
	return (uint32_t)popcount(lsbs) + (uint32_t)popcount(msbs) + q_bits * 8;
}

static uint32_t modify_quant_code(uint32_t code, int quant)
{
	/* block_packing.comp:161-167. */
	int e = (int)((code >> 3) & 31);
	e = imax(e - quant, 0);
	return (code & ~(31u << 3)) | ((uint32_t)e << 3);
}

long or_pack(const or_layout *L, const float *coeffs, const int32_t *quant, int sequence,
             uint32_t *bitstream, long capacity_words, uint32_t *meta)
{
	q8x8 *Q = (q8x8 *)malloc(sizeof(q8x8) * 16);
	uint8_t *buf = (uint8_t *)malloc(4 * 4096);
	long out_words = 0;

	/* Visit bands in block-index order so the output is index-ordered. */
	for (int l = OR_LEVELS - 1; l >= 0; l--)
		for (int c = 0; c < OR_COMPONENTS; c++)
			for (int b = 0; b < OR_BANDS; b++)
			{
				const or_band *B = &L->band[c][l][b];
				if (!B->present)
					continue;
				float qres, rdo;
				band_constants(L, c, l, b, &qres, &rdo);
				uint32_t qcode_band = or_encode_quant(1.0f / or_quant_resolution_policy(l, c, b, L->precision0));
				const float *band = coeffs + B->coeff_off;

				for (int by = 0; by < B->by32; by++)
					for (int bx = 0; bx < B->bx32; bx++)
					{
						uint32_t block = (uint32_t)(B->off32 + by * B->bx32 + bx);
						int qn = quant[block];
						uint32_t cw[16], ballot = 0, plane_bytes[16], sign_bits[16], total_bits = 0;
						int in_range[16];
						for (int k = 0; k < 16; k++)
						{
							int x8 = bx * 4 + (k & 3), y8 = by * 4 + (k >> 2);
							in_range[k] = x8 < B->bx8 && y8 < B->by8;
							uint32_t cost = 0;
							cw[k] = 0;
							if (in_range[k])
							{
								quantize_8x8(band, B->bw, B->bw, B->bh, 8 * x8, 8 * y8, qres, rdo, &Q[k]);
								cost = Q[k].cost[qn < 15 ? qn : 15];
								cw[k] = Q[k].code_word;
							}
							cw[k] = quantize_code_word(cw[k], qn);
							if (cw[k] & 0xffffu)
								ballot |= 1u << k;
							plane_bytes[k] = required_8x8_size(cw[k]);
							sign_bits[k] = cost - 8 * plane_bytes[k];
							total_bits += cost ? cost + 24 : 0;
						}
						uint32_t wcount = (total_bits + 31) / 32;
						if (wcount)
							wcount += 2;
						if (!wcount)
						{
							meta[2 * block] = 0;
							meta[2 * block + 1] = 0;
							continue;
						}
						if (out_words + wcount > capacity_words)
						{
							free(Q);
							free(buf);
							return -1;
						}

						memset(buf, 0, wcount * 4);
						uint32_t w0 = ballot | (wcount << 16) | ((uint32_t)(sequence & 7) << 28);
						uint32_t w1 = modify_quant_code(qcode_band, qn) | (block << 8);
						memcpy(buf, &w0, 4);
						memcpy(buf + 4, &w1, 4);
						int N = popcount(ballot);
						uint32_t planes_total = 0;
						for (int k = 0; k < 16; k++)
							planes_total += (ballot >> k) & 1 ? plane_bytes[k] : 0;
						uint32_t plane_off = 8 + 3 * (uint32_t)N;
						uint32_t sign_base = plane_off + planes_total;
						uint32_t sign_pos = 0;
						int j = 0;
						for (int k = 0; k < 16; k++)
						{
							if (!((ballot >> k) & 1))
								continue;
							uint16_t ctrl = (uint16_t)(cw[k] & 0xffffu);
							memcpy(buf + 8 + 2 * j, &ctrl, 2);
							buf[8 + 2 * N + j] = (uint8_t)(cw[k] >> 16);
							j++;

							uint32_t in_q = (Q[k].code_word >> 16) & 0xf;
							uint32_t out_q = (cw[k] >> 16) & 0xf;
							for (int s = 0; s < 8; s++)
							{
								uint32_t out_planes = ((cw[k] >> (2 * s)) & 3) + out_q;
								uint32_t in_planes = ((Q[k].code_word >> (2 * s)) & 3) + in_q;
								if (!out_planes)
									continue;
								uint32_t sig = 0;
								for (uint32_t pi = 0; pi < out_planes; pi++)
								{
									int bit = (int)(in_planes - 1 - pi);
									uint32_t byte = 0;
									for (int p = 0; p < 8; p++)
										byte |= (uint32_t)((Q[k].n[s][p] >> bit) & 1) << p;
									sig |= byte;
									buf[plane_off++] = (uint8_t)byte;
								}
								for (int p = 0; p < 8; p++)
								{
									if ((sig >> p) & 1)
									{
										uint32_t sbit = Q[k].t[s][p] < 0.0f ? 1u : 0u;
										buf[sign_base + (sign_pos >> 3)] |= (uint8_t)(sbit << (sign_pos & 7));
										sign_pos++;
									}
								}
							}
						}
						(void)sign_bits;
						memcpy(bitstream + out_words, buf, wcount * 4);
						meta[2 * block] = (uint32_t)out_words;
						meta[2 * block + 1] = wcount;
						out_words += wcount;
					}
			}
	free(Q);
	free(buf);
	return out_words;
}

long or_encode(const or_layout *L, const uint8_t *const planes[3], const int pitches[3],
               int precision, size_t budget, int sequence, int q1,
               uint32_t *bitstream, long capacity_words, uint32_t *meta, int32_t *quant_out)
{
	const void *p[3] = {planes[0], planes[1], planes[2]};
	return or_encode_ex(L, p, pitches, 8, precision, budget, sequence, q1, bitstream, capacity_words, meta,
	                    quant_out);
}

long or_encode_ex(const or_layout *L, const void *const planes[3], const int pitches[3], int bits,
                  int precision, size_t budget, int sequence, int q1,
                  uint32_t *bitstream, long capacity_words, uint32_t *meta, int32_t *quant_out)
{
	float *coeffs = (float *)calloc((size_t)L->coeff_count, sizeof(float));
	uint16_t *words = (uint16_t *)malloc(sizeof(uint16_t) * 16 * (size_t)L->count32);
	uint8_t *bucket = (uint8_t *)malloc(16 * (size_t)L->count32);
	uint32_t *code8 = (uint32_t *)calloc((size_t)L->count8, sizeof(uint32_t));
	uint8_t *np8 = (uint8_t *)calloc((size_t)L->count8, 1);
	uint32_t totals[FLAT], prefix[FLAT], consumed;
	int32_t *quant = quant_out ? quant_out : (int32_t *)malloc(sizeof(int32_t) * (size_t)L->count32);
	/* Precision 0 selects its own quantizer policy (pyrowave_encoder.cpp:144-172). */
	or_layout L0 = *L;
	if (precision == 0)
		L0.precision0 = 1;
	L = &L0;

	or_dwt_frame_ex(L, planes, pitches, bits, precision, coeffs, NULL);
	or_quant_analyze(L, coeffs, words, bucket, code8, np8, totals, &consumed);
	or_finalize(totals, prefix, q1);
	or_resolve(L, words, bucket, prefix, consumed, budget, quant);
	long n = or_pack(L, coeffs, quant, sequence, bitstream, capacity_words, meta);

	free(coeffs);
	free(words);
	free(bucket);
	free(code8);
	free(np8);
	if (!quant_out)
		free(quant);
	return n;
}

/* ------------------------------------------------------------------------ */
/* Packetize: pyrowave_encoder.cpp:596-623, 1048-1115                       */
/* ------------------------------------------------------------------------ */

size_t or_compute_num_packets(const or_layout *L, const uint32_t *meta, size_t boundary)
{
	size_t num_packets = 0, size_in_packet = 8;
	for (int i = 0; i < L->count32; i++)
	{
		size_t sz = (size_t)meta[2 * i + 1] * 4;
		if (!sz)
			continue;
		if (size_in_packet + sz > boundary)
		{
			size_in_packet = 0;
			num_packets++;
		}
		size_in_packet += sz;
	}
	if (size_in_packet)
		num_packets++;
	return num_packets;
}

size_t or_packetize(const or_layout *L, const uint32_t *meta, const uint32_t *bitstream, int sequence,
                    size_t boundary, uint8_t *out, size_t out_size, size_t *packets)
{
	size_t num_packets = 0, size_in_packet = 0, packet_offset = 0, out_off = 0;
	uint32_t nonzero = 0;
	for (int i = 0; i < L->count32; i++)
		if (meta[2 * i + 1])
			nonzero++;
	uint32_t h0 = (uint32_t)(L->width - 1) | ((uint32_t)(L->height - 1) << 14) |
	              ((uint32_t)(sequence & 7) << 28) | (1u << 31);
	uint32_t h1 = (nonzero & 0xffffffu) | (0u << 24) | ((uint32_t)(L->chroma == 1) << 26);
	if (out_size < 8)
		return 0;
	memcpy(out, &h0, 4);
	memcpy(out + 4, &h1, 4);
	out_off = 8;
	size_in_packet = 8;
	for (int i = 0; i < L->count32; i++)
	{
		size_t sz = (size_t)meta[2 * i + 1] * 4;
		if (!sz)
			continue;
		if (size_in_packet + sz > boundary)
		{
			packets[2 * num_packets] = packet_offset;
			packets[2 * num_packets + 1] = size_in_packet;
			num_packets++;
			size_in_packet = 0;
			packet_offset = out_off;
		}
		if (out_off + sz > out_size)
			return num_packets;
		memcpy(out + out_off, bitstream + meta[2 * i], sz);
		out_off += sz;
		size_in_packet += sz;
	}
	if (size_in_packet)
	{
		packets[2 * num_packets] = packet_offset;
		packets[2 * num_packets + 1] = size_in_packet;
		num_packets++;
	}
	return num_packets;
}

/* ------------------------------------------------------------------------ */
/* Structural validator: pyrowave_encoder.cpp:960-1046                      */
/* ------------------------------------------------------------------------ */

static int block_geometry(const or_layout *L, uint32_t block_index, int *w8, int *h8)
{
	for (int c = 0; c < OR_COMPONENTS; c++)
		for (int l = 0; l < OR_LEVELS; l++)
			for (int b = 0; b < OR_BANDS; b++)
			{
				const or_band *B = &L->band[c][l][b];
				if (!B->present)
					continue;
				int n = B->bx32 * B->by32;
				if ((int)block_index >= B->off32 && (int)block_index < B->off32 + n)
				{
					int r = (int)block_index - B->off32;
					int x = r % B->bx32, y = r / B->bx32;
					*w8 = imin(4, B->bx8 - 4 * x);
					*h8 = imin(4, B->by8 - 4 * y);
					return 1;
				}
			}
	return 0;
}

int or_validate_block(const or_layout *L, const uint32_t *block, uint32_t block_index, uint32_t meta_words)
{
	if (meta_words == 0)
		return 1;
	uint32_t w0 = block[0], w1 = block[1];
	if ((w1 >> 8) != block_index)
		return 0;
	if (((w0 >> 16) & 0xfff) != meta_words)
		return 0;
	int w8, h8;
	if (!block_geometry(L, block_index, &w8, &h8))
		return 0;
	uint32_t ballot = w0 & 0xffff;
	int N = popcount(ballot);
	const uint8_t *u8 = (const uint8_t *)block;
	uint32_t offset = 8 + 3 * (uint32_t)N;
	if (offset > meta_words * 4)
		return 0;
	int nsig = 0, j = 0;
	for (int bit = 0; bit < 16; bit++)
	{
		if (!((ballot >> bit) & 1))
			continue;
		int x = bit & 3, y = bit >> 2;
		if (x >= w8 || y >= h8)
			return 0;
		uint16_t ctrl;
		memcpy(&ctrl, u8 + 8 + 2 * j, 2);
		uint32_t q_bits = u8[8 + 2 * N + j] & 0xf;
		for (int s = 0; s < 8; s++)
		{
			uint32_t np = q_bits + ((ctrl >> (2 * s)) & 3);
			uint32_t sig = 0;
			for (uint32_t p = 0; p < np; p++)
			{
				if (offset >= meta_words * 4)
					return 0;
				sig |= u8[offset++];
			}
			nsig += popcount(sig);
		}
		j++;
	}
	offset += (uint32_t)(nsig + 7) / 8;
	return (offset + 3) / 4 == meta_words;
}

/* ------------------------------------------------------------------------ */
/* Decoder packet parser: pyrowave_decoder.cpp:126-262, 780-794, 832-840    */
/* ------------------------------------------------------------------------ */

struct or_decoder
{
	const or_layout *L;
	uint32_t *offsets;
	uint32_t *payload;
	long payload_size, payload_cap;
	int decoded_blocks;
	int total_blocks;
	uint32_t last_seq;
	int decoded_frame;
};

or_decoder *or_decoder_create(const or_layout *L)
{
	or_decoder *d = (or_decoder *)calloc(1, sizeof(or_decoder));
	d->L = L;
	d->offsets = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)L->count32);
	d->payload_cap = 1024;
	d->payload = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)d->payload_cap);
	or_decoder_clear(d);
	return d;
}

void or_decoder_destroy(or_decoder *d)
{
	if (!d)
		return;
	free(d->offsets);
	free(d->payload);
	free(d);
}

void or_decoder_clear(or_decoder *d)
{
	for (int i = 0; i < d->L->count32; i++)
		d->offsets[i] = UINT32_MAX;
	d->decoded_blocks = 0;
	d->last_seq = UINT32_MAX;
	d->decoded_frame = 0;
	d->total_blocks = d->L->count32;
	d->payload_size = 0;
}

/* pyrowave_decoder.cpp:126-150.  The block is counted and its offset set
 * before the length check, as in the reference. */
static int decode_packet(or_decoder *d, const uint32_t *words)
{
	uint32_t block = words[1] >> 8;
	uint32_t pw = (words[0] >> 16) & 0xfff;
	if (d->offsets[block] != UINT32_MAX)
		return 1;
	d->decoded_blocks++;
	d->offsets[block] = (uint32_t)d->payload_size;
	if (2 > pw)
		return 0;
	if (d->payload_size + pw > d->payload_cap)
	{
		while (d->payload_size + pw > d->payload_cap)
			d->payload_cap *= 2;
		d->payload = (uint32_t *)realloc(d->payload, sizeof(uint32_t) * (size_t)d->payload_cap);
	}
	memcpy(d->payload + d->payload_size, words, pw * 4);
	d->payload_size += pw;
	return 1;
}

int or_decoder_push_packet(or_decoder *d, const void *data_, size_t size)
{
	const uint8_t *data = (const uint8_t *)data_;
	while (size >= 8)
	{
		uint32_t w[2];
		memcpy(w, data, 8);
		uint32_t seq = (w[0] >> 28) & 7;
		if (w[0] >> 31)
		{
			uint32_t chroma_res = (w[1] >> 26) & 1;
			if ((int)chroma_res != d->L->chroma)
				return 0;
			uint32_t diff = (seq - d->last_seq) & 7;
			if (d->last_seq != UINT32_MAX && diff > 3)
				return 1;
			if (d->last_seq == UINT32_MAX || diff != 0)
			{
				or_decoder_clear(d);
				d->last_seq = seq;
			}
			uint32_t code = (w[1] >> 24) & 3;
			if (code == 0)
			{
				int wd = (int)(w[0] & 0x3fff) + 1, ht = (int)((w[0] >> 14) & 0x3fff) + 1;
				if (wd != d->L->width || ht != d->L->height)
					return 0;
				d->total_blocks = (int)(w[1] & 0xffffff);
			}
			else
				return 0;
			data += 8;
			size -= 8;
			continue;
		}

		size_t packet_size = (size_t)((w[0] >> 16) & 0xfff) * 4;
		if (packet_size > size)
			return 0;
		int restart;
		if (d->last_seq == UINT32_MAX)
			restart = 1;
		else
		{
			uint32_t diff = (seq - d->last_seq) & 7;
			if (diff > 3)
				return 1;
			restart = diff != 0;
		}
		if (restart)
		{
			or_decoder_clear(d);
			d->last_seq = seq;
		}
		if ((w[1] >> 8) >= (uint32_t)d->L->count32)
			return 0;
		/* A zero-length block already present: the reference's loop would
		 * never advance (pyrowave_decoder.cpp:236-246); reported as a parse
		 * error instead. */
		if (packet_size == 0 && d->offsets[w[1] >> 8] != UINT32_MAX)
			return 0;
		uint32_t *tmp = (uint32_t *)malloc(packet_size < 8 ? 8 : packet_size);
		memcpy(tmp, data, packet_size < 8 ? 8 : packet_size); /* size >= 8 here */
		int ok = decode_packet(d, tmp);
		free(tmp);
		if (!ok)
			return 0;
		data += packet_size;
		size -= packet_size;
	}
	return size == 0;
}

int or_decoder_is_ready(const or_decoder *d, int allow_partial)
{
	if (d->decoded_frame)
		return 0;
	if (d->last_seq == UINT32_MAX)
		return 0;
	if (d->decoded_blocks < d->total_blocks)
		if (!allow_partial || d->decoded_blocks <= d->total_blocks / 2)
			return 0;
	return 1;
}

void or_decoder_mark_decoded(or_decoder *d) { d->decoded_frame = 1; }
long or_decoder_payload_words(const or_decoder *d) { return d->payload_size; }
const uint32_t *or_decoder_payload(const or_decoder *d) { return d->payload; }
const uint32_t *or_decoder_offsets(const or_decoder *d) { return d->offsets; }
int or_decoder_decoded_blocks(const or_decoder *d) { return d->decoded_blocks; }
int or_decoder_total_blocks(const or_decoder *d) { return d->total_blocks; }

/* ------------------------------------------------------------------------ */
/* Dequant: shaders/wavelet_dequant.comp                                    */
/* ------------------------------------------------------------------------ */

static uint32_t payload_byte(const uint32_t *payload, long payload_words, long byte_index)
{
	/* Reads past the payload return 0 (the reference relies on 16 B of padding). */
	if (byte_index < 0 || byte_index >= payload_words * 4)
		return 0;
	return (payload[byte_index >> 2] >> (8 * (byte_index & 3))) & 0xffu;
}

int or_dequant(const or_layout *L, const uint32_t *payload, long payload_words,
               const uint32_t *offsets, float *coeffs)
{
	for (int c = 0; c < OR_COMPONENTS; c++)
		for (int l = 0; l < OR_LEVELS; l++)
			for (int b = 0; b < OR_BANDS; b++)
			{
				const or_band *B = &L->band[c][l][b];
				if (!B->present)
					continue;
				float *band = coeffs + B->coeff_off;
				for (int by = 0; by < B->by32; by++)
					for (int bx = 0; bx < B->bx32; bx++)
					{
						uint32_t block = (uint32_t)(B->off32 + by * B->bx32 + bx);
						float out[32][32];
						memset(out, 0, sizeof(out));
						uint32_t off = offsets[block];
						if (off != UINT32_MAX)
						{
							long base = (long)off * 4;
							uint32_t ballot = payload_byte(payload, payload_words, base) |
							                  (payload_byte(payload, payload_words, base + 1) << 8);
							uint32_t q_code = payload_byte(payload, payload_words, base + 4);
							int N = popcount(ballot);
							long plane_off = base + 8 + 3 * N;
							/* Per-8x8 plane byte offsets (wavelet_dequant.comp:258-285). */
							long p_off[16];
							uint32_t ctrl[16], qb[16];
							int j = 0;
							for (int k = 0; k < 16; k++)
							{
								ctrl[k] = qb[k] = 0;
								if ((ballot >> k) & 1)
								{
									ctrl[k] = payload_byte(payload, payload_words, base + 8 + 2 * j) |
									          (payload_byte(payload, payload_words, base + 9 + 2 * j) << 8);
									qb[k] = payload_byte(payload, payload_words, base + 8 + 2 * N + j);
									j++;
								}
								p_off[k] = plane_off;
								uint32_t lsbs = ctrl[k] & 0x5555u, msbs = ctrl[k] & 0xaaaau;
								msbs |= msbs >> 1;
								plane_off += popcount(lsbs) + popcount(msbs) + (qb[k] & 0xf) * 8;
							}
							long sign_bit = plane_off * 8;
							float q = or_decode_quant((uint8_t)q_code);
							for (int k = 0; k < 16; k++)
							{
								if (!((ballot >> k) & 1))
									continue;
								uint32_t q_bits = qb[k] & 0xf;
								float inv_scale = q * ((float)(qb[k] >> 4) / 8.0f + 0.25f);
								for (int s = 0; s < 8; s++)
								{
									int bo = 2 * s;
									uint32_t lmask = (1u << bo) - 1u;
									uint32_t lsbs = ctrl[k] & 0x5555u, msbs = ctrl[k] & 0xaaaau;
									msbs |= msbs >> 1;
									long byte_off = p_off[k] + popcount(lsbs & lmask) + popcount(msbs & lmask) +
									                (long)q_bits * s;
									int planes = (int)(q_bits + ((ctrl[k] >> bo) & 3));
									int mag[8] = {0, 0, 0, 0, 0, 0, 0, 0};
									for (int pl = planes - 1; pl >= 0; pl--)
									{
										uint32_t byte = payload_byte(payload, payload_words, byte_off++);
										for (int p = 0; p < 8; p++)
											mag[p] |= (int)((byte >> p) & 1) << pl;
									}
									int x0 = 8 * (k & 3) + 4 * (s >> 2), y0 = 8 * (k >> 2) + 2 * (s & 3);
									for (int p = 0; p < 8; p++)
									{
										float v = (float)mag[p];
										if (v != 0.0f)
										{
											v += 0.5f;
											v *= inv_scale;
											uint32_t sb = payload_byte(payload, payload_words, sign_bit >> 3);
											if ((sb >> (sign_bit & 7)) & 1)
												v = -v;
											sign_bit++;
										}
										else
											v *= inv_scale;
										out[y0 + (p & 1)][x0 + (p >> 1)] = v;
									}
								}
							}
						}
						for (int y = 0; y < 32; y++)
							for (int x = 0; x < 32; x++)
							{
								int gx = bx * 32 + x, gy = by * 32 + y;
								if (gx < B->bw && gy < B->bh)
									band[(size_t)gy * B->bw + gx] = out[y][x];
							}
					}
			}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Inverse transform chain: pyrowave_decoder.cpp:699-778, idwt.comp:201-213 */
/* ------------------------------------------------------------------------ */

static uint8_t unorm8_store(float v)
{
	/* +0.5 DC shift, clamp to [0,1], round to nearest (ties up). */
	v = v + 0.5f;
	if (!(v > 0.0f))
		v = 0.0f;
	if (v > 1.0f)
		v = 1.0f;
	return (uint8_t)(int)fmaf(v, 255.0f, 0.5f);
}

static uint16_t unorm16_store(float v)
{
	/* R16_UNORM image store (idwt.comp:201-213 into an R16_UNORM target,
	 * decode.cpp:215): +0.5 DC shift, clamp to [0,1], round to nearest (ties
	 * up), as the UNORM8 store. */
	v = v + 0.5f;
	if (!(v > 0.0f))
		v = 0.0f;
	if (v > 1.0f)
		v = 1.0f;
	return (uint16_t)(int)fmaf(v, 65535.0f, 0.5f);
}

int or_idwt_frame(const or_layout *L, const float *coeffs_in, int precision,
                  uint8_t *const planes[3], const int pitches[3])
{
	void *p[3] = {planes[0], planes[1], planes[2]};
	return or_idwt_frame_ex(L, coeffs_in, precision, p, pitches, 8);
}

int or_idwt_frame_ex(const or_layout *L, const float *coeffs_in, int precision,
                     void *const planes[3], const int pitches[3], int bits)
{
	float *coeffs = (float *)malloc(sizeof(float) * (size_t)L->coeff_count);
	memcpy(coeffs, coeffs_in, sizeof(float) * (size_t)L->coeff_count);
	for (int l = OR_LEVELS - 1; l >= 0; l--)
	{
		for (int c = 0; c < OR_COMPONENTS; c++)
		{
			int first_level = (c != 0 && L->chroma == 0) ? 1 : 0;
			if (l < first_level)
				continue;
			const or_band *B = L->band[c][l];
			int bw = B[0].bw, bh = B[0].bh;
			/* precision 1: the decoder holds the high bands of levels 0-1 in
			 * fp16 storage; LL intermediates stay FP32.  precision 3: the LL
			 * of levels 0-1 (written by the level above) in fp16 as well. */
			if ((precision == 1 || precision == 3) && l < 2)
				for (int b = precision == 3 ? 0 : 1; b < 4; b++)
					or_round_to_f16_array(coeffs + B[b].coeff_off, (long)B[b].bw * B[b].bh);
			/* precision 0: every band in fp16 storage, fp16 arithmetic. */
			if (precision == 0)
				for (int b = l == OR_LEVELS - 1 ? 0 : 1; b < 4; b++)
					or_round_to_f16_array(coeffs + B[b].coeff_off, (long)B[b].bw * B[b].bh);
			void (*level)(const float *, const float *, const float *, const float *, int, int, int, float *,
			              int) = precision == 0 ? or_idwt_level_h : or_idwt_level_f32;
			if (l == first_level)
			{
				int pw, ph;
				plane_dims(L, c, &pw, &ph);
				float *tmp = (float *)malloc(sizeof(float) * (size_t)4 * bw * bh);
				level(coeffs + B[0].coeff_off, coeffs + B[1].coeff_off, coeffs + B[2].coeff_off,
				      coeffs + B[3].coeff_off, bw, bw, bh, tmp, 2 * bw);
				for (int y = 0; y < ph && y < 2 * bh; y++)
					for (int x = 0; x < pw && x < 2 * bw; x++)
					{
						const float v = tmp[(size_t)y * 2 * bw + x];
						const size_t at = (size_t)y * pitches[c] + x;
						if (bits == 16)
							((uint16_t *)planes[c])[at] = unorm16_store(v);
						else
							((uint8_t *)planes[c])[at] = unorm8_store(v);
					}
				free(tmp);
			}
			else
			{
				const or_band *P = &L->band[c][l - 1][0];
				level(coeffs + B[0].coeff_off, coeffs + B[1].coeff_off, coeffs + B[2].coeff_off,
				      coeffs + B[3].coeff_off, bw, bw, bh, coeffs + P->coeff_off, P->bw);
			}
		}
	}
	free(coeffs);
	return 0;
}

int or_idwt_frame_f64(const or_layout *L, const double *coeffs_in, double *const planes[3], const int pitches[3])
{
	double *coeffs = (double *)malloc(sizeof(double) * (size_t)L->coeff_count);
	memcpy(coeffs, coeffs_in, sizeof(double) * (size_t)L->coeff_count);
	for (int l = OR_LEVELS - 1; l >= 0; l--)
	{
		for (int c = 0; c < OR_COMPONENTS; c++)
		{
			int first_level = (c != 0 && L->chroma == 0) ? 1 : 0;
			if (l < first_level)
				continue;
			const or_band *B = L->band[c][l];
			int bw = B[0].bw, bh = B[0].bh;
			if (l == first_level)
			{
				int pw, ph;
				plane_dims(L, c, &pw, &ph);
				double *tmp = (double *)malloc(sizeof(double) * (size_t)4 * bw * bh);
				or_idwt_level_f64(coeffs + B[0].coeff_off, coeffs + B[1].coeff_off, coeffs + B[2].coeff_off,
				                  coeffs + B[3].coeff_off, bw, bw, bh, tmp, 2 * bw);
				for (int y = 0; y < ph && y < 2 * bh; y++)
					for (int x = 0; x < pw && x < 2 * bw; x++)
						planes[c][(size_t)y * pitches[c] + x] = tmp[(size_t)y * 2 * bw + x] + 0.5;
				free(tmp);
			}
			else
			{
				const or_band *P = &L->band[c][l - 1][0];
				or_idwt_level_f64(coeffs + B[0].coeff_off, coeffs + B[1].coeff_off, coeffs + B[2].coeff_off,
				                  coeffs + B[3].coeff_off, bw, bw, bh, coeffs + P->coeff_off, P->bw);
			}
		}
	}
	free(coeffs);
	return 0;
}
