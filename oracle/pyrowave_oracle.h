/*
 * pyrowave_oracle.h -- CPU restatement of the PyroWave encode/decode hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The shipped library
 * (pyrowave_amd/libpyrowave-shared.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference files
 * are named relative to the Themaister/pyrowave source tree.
 *
 * Pinning status of each part (what checks it besides the restated formula):
 *   layout / block index order ... golden fixtures from the reference's own
 *                                  metal/pyrowave_bitstream.cpp (oracle/_ref)
 *   packetize / num_packets ...... same fixtures
 *   decoder packet parser ........ same fixtures (push/ready/clear sequences)
 *   FP64 DWT / iDWT .............. perfect reconstruction, constant input ->
 *                                  zero high bands, cubic vanishing moments,
 *                                  band orientation, brute-force convolution
 *                                  with the textbook 9/7 analysis filters
 *   FP32 DWT / iDWT .............. error bound against the FP64 transform
 *   16-bit (R16_UNORM) input ..... planes of 257 * u give the 8-bit transform
 *                                  and bitstream exactly; the level-0 sample
 *                                  against exact rationals
 *   16-bit output ................ round(clamp(v) * 65535) of the FP64
 *                                  inverse within one unit
 *                                  (tests/test_oracle_unorm16.py)
 *   precision 0 (fp16 math) ...... numpy float16 arithmetic step by step
 *                                  and the FP64 transform within fp16 bounds
 *                                  (tests/test_oracle_fp16.py)
 *   fp16 conversion .............. exhaustive comparison with numpy float16
 *   quant scale codes ............ exact-rational minimal code keeping the
 *                                  block peak under 2^e - 1/4
 *   magnitudes / quant levels .... closed-form deadzone bound at step 2^q*D
 *                                  through pack -> parse -> dequant
 *   distortion tables ............ squared reconstruction error of the
 *                                  dequantized block at every quant level,
 *                                  times the distortion weight, fp16-stored
 *   cost tables / words .......... structural size of every block packed at
 *                                  every quant level (validate_bitstream)
 *   buckets ...................... float64 log2 away from ties; literal run of
 *                                  the shader's max-scan, clamp included
 *   savings totals / consumed .... recomputed from the word/bucket tables
 *   distortion weight ............ published Mannos-Sakrison CSF at the band
 *                                  midpoints
 *   quant resolution, band term .. measured analysis gains of the FP64 9/7
 *   quant resolution .............. noise-power normalization: 2^(9 + log2 of
 *                                  the band's FP64 synthesis-basis L2 norm) at
 *                                  its stage count in its own plane (band,
 *                                  level and 4:2:0 chroma terms, base bits)
 *   (tests/test_oracle_pinning.py; tests/test_oracle_mutants.py plants a
 *   mistake in each of these and requires a check to fail)
 *   transcribed constants (pyrowave_encoder.cpp:144-172), fixed by no
 *   bitstream rule or mathematical property: one bit less for 4:4:4 chroma
 *   and the caps 4096 / 512; the check above asserts the capped values.
 *   finalize (quirk Q1) .......... literal simulation of the reference's
 *                                  512-thread Hillis-Steele loop
 *   resolve ...................... literal simulation of the reference's
 *                                  per-(bucket,subdivision) greedy loop
 *   pack ......................... structural validator (restates
 *                                  validate_bitstream), decode round trip,
 *                                  analyzed words == packed words
 *   dequant / iDWT output ........ round trip |decoded - input| <= 1 at a
 *                                  64 KiB budget (pyrowave_c_test.cpp:347-379)
 *   parity unpinned (byte level vs the reference GPU codec): GLSL log2
 *   rounding at bucket boundaries, GLSL division rounding in the scale code,
 *   float summation order of subgroup reductions, UNORM rounding on store.
 */
#ifndef PYROWAVE_ORACLE_H
#define PYROWAVE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_LEVELS = 5, OR_COMPONENTS = 3, OR_BANDS = 4 };

typedef struct or_band
{
	int present;       /* transmitted band (has block indices) */
	int bw, bh;        /* band size in coefficients */
	int bx8, by8;      /* 8x8 block grid */
	int bx32, by32;    /* 32x32 block grid */
	int off8, off32;   /* first 8x8 / 32x32 block index */
	long coeff_off;    /* offset of this band in a flat coefficient array */
} or_band;

typedef struct or_layout
{
	int width, height, chroma; /* chroma: 0 = 4:2:0, 1 = 4:4:4 */
	int aligned_w, aligned_h;
	int count8, count32;
	long coeff_count;          /* total floats of the flat coefficient array */
	int per_subdivision;       /* next_pow2(ceil(count32 / 16)) */
	or_band band[OR_COMPONENTS][OR_LEVELS][OR_BANDS];
	int precision0;            /* 1: quantizer policy of precision 0 (set by the caller) */
} or_layout;

/* 32x32 block -> band lookup, in block-index order. */
typedef struct or_block_desc
{
	int16_t component, level, band, pad;
	int32_t bx, by;            /* 32x32 position within the band */
} or_block_desc;

int or_layout_init(or_layout *L, int width, int height, int chroma);
long or_layout_size(void);
int or_block_descs(const or_layout *L, or_block_desc *out);

/* Encoder-side quantizer constants (pyrowave_encoder.cpp:107-172, 458-459). */
float or_quant_resolution(int level, int component, int band);
/* precision0 != 0: the policy of the reference's precision 0 (6 base bits,
 * cap 512); else precision >= 1 (8 base bits, cap 4096).
 * or_rdo_distortion_scale and the quantizer follow L->precision0. */
float or_quant_resolution_policy(int level, int component, int band, int precision0);
float or_rdo_distortion_scale(const or_layout *L, int level, int component, int band);
uint8_t or_encode_quant(float decoder_q_scale);
float or_decode_quant(uint8_t code);

/* fp16 helpers (round to nearest even, subnormals kept). */
float or_round_f64_to_f16(double d);
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);
void or_round_to_f16_array(float *v, long n);
float or_round_f16_fast(float x);
long or_check_round_f16_fast(void);

/* Forward extension index for the encoder's sampler trick (dwt.comp:27-36). */
int or_fwd_ext(int p, int R, int A);

/* One 2-D DWT level, FP64 and FP32.  `in` has R_y rows of R_x valid samples
 * (pitch in elements); the transform domain is A_x x A_y; the four outputs are
 * (A_x/2) x (A_y/2) with pitch out_pitch.  dc_shift subtracts 0.5 from inputs
 * in the u8 variants. */
void or_dwt_level_f64(const double *in, int in_pitch, int R_x, int R_y, int A_x, int A_y,
                      double *ll, double *hl, double *lh, double *hh, int out_pitch);
void or_dwt_level_f32(const float *in, int in_pitch, int R_x, int R_y, int A_x, int A_y,
                      float *ll, float *hl, float *lh, float *hh, int out_pitch);
void or_idwt_level_f64(const double *ll, const double *hl, const double *lh, const double *hh,
                       int in_pitch, int bw, int bh, double *out, int out_pitch);
void or_idwt_level_f32(const float *ll, const float *hl, const float *lh, const float *hh,
                       int in_pitch, int bw, int bh, float *out, int out_pitch);
/* The same levels in fp16 arithmetic (precision 0): inputs hold fp16 values;
 * lifting steps are an fp16 add then an fp16 fused multiply-add with the
 * fp16-rounded constant; the K / 1/K scaling is the exact product of the
 * fp16 value and the FP32 constant, rounded once to fp16. */
void or_dwt_level_h(const float *in, int in_pitch, int R_x, int R_y, int A_x, int A_y,
                    float *ll, float *hl, float *lh, float *hh, int out_pitch);
void or_idwt_level_h(const float *ll, const float *hl, const float *lh, const float *hh,
                     int in_pitch, int bw, int bh, float *out, int out_pitch);

/* Whole-frame forward transform into the flat coefficient array.
 * precision 2: FP32 everywhere.  precision 1: FP32 math; the high bands
 * (HL, LH, HH) of levels 0-1 rounded to fp16 storage, LL kept FP32.
 * precision 3: FP32 math; all four bands of levels 0-1 rounded to fp16, so
 * levels 1 and 2 transform the fp16-rounded LL (the reference's precision 1
 * storage, pyrowave_common.cpp:133-189: one R16F image holds every band of
 * levels 0-1, and component_ll_views feed the next level from it).
 * precision 0: fp16 arithmetic and fp16 values at every level (the
 * reference's precision 0: FLOAT = float16_t, R16F storage at all levels),
 * level-0 input fp16(fp16(u / 255) - 0.5).
 * use_f64 != 0: FP64 math; result written to coeffs64 (may be NULL otherwise). */
int or_dwt_frame(const or_layout *L, const uint8_t *const planes[3], const int pitches[3],
                 int precision, float *coeffs, double *coeffs64);
/* As or_dwt_frame with input planes of `bits` 8 (UNORM8, u / 255) or 16
 * (R16_UNORM, u / 65535: encode.cpp:152); pitches in samples.  A 16-bit
 * sample 257 * u is the same rational as the 8-bit u, so such planes give
 * the 8-bit transform exactly. */
/* The level-0 input sample: UNORM8 u / 255 or R16_UNORM u / 65535 (bits
 * 8 / 16), rounded to float, minus 0.5 in float; precision 0 rounds the
 * texel to fp16 and subtracts in fp16. */
float or_unorm_dc(uint32_t u, int bits, int precision);
int or_dwt_frame_ex(const or_layout *L, const void *const planes[3], const int pitches[3], int bits,
                    int precision, float *coeffs, double *coeffs64);

/* Quantizer + analyze (wavelet_quant.comp, analyze_rate_control.comp).
 * Outputs, per 32x32 block b: words[b*16+q], bucket[b*16+q] (inclusive bucket,
 * 0 for q == 0); per 8x8: code8 (code word at q=0), np8 (num_planes).
 * totals[2048] receives the per-(bucket,subdivision) savings, *consumed the
 * sum of words at q=0. */
int or_quant_analyze(const or_layout *L, const float *coeffs,
                     uint16_t *words, uint8_t *bucket, uint32_t *code8, uint8_t *np8,
                     uint32_t *totals, uint32_t *consumed);
/* The same, also writing dist32[b*16+q]: the 32x32 block's summed distortion
 * at quant level q (analyze_rate_control.comp:169-215), when not NULL. */
int or_quant_analyze_ex(const or_layout *L, const float *coeffs,
                        uint16_t *words, uint8_t *bucket, uint32_t *code8, uint8_t *np8,
                        uint32_t *totals, uint32_t *consumed, float *dist32);
/* One 8x8 block of band (c, l, b) at 8x8 position (x8, y8): the BlockStats
 * entries errors[min(q, num_planes)] for q = 0..15 (payload cost in bits and
 * the fp16-stored distortion), the code word at q = 0 and msb (num_planes - 1,
 * -1 when the block quantizes to zero).  Returns -1 outside the band. */
int or_block8_stats(const or_layout *L, const float *coeffs, int c, int l, int b, int x8, int y8,
                    uint16_t cost[16], float dist[16], uint32_t *code_word, int *msb);
/* uint(max(60 + 2 * log2(r) + 0.5, 0)), evaluated exactly
 * (analyze_rate_control.comp:73-83). */
uint32_t or_bucket_of_ratio(float r);
/* The monotone bucket assignment over the 16 quant levels of a 32x32 block
 * (analyze_rate_control.comp:85-99): raw[0] must be 0. */
void or_bucket_scan(const uint32_t raw[16], uint32_t out[16]);

/* analyze_rate_control_finalize.comp.  q1 != 0 reproduces the reference's
 * 8-step scan (windowed sums for flat index >= 1024); q1 == 0 is a true
 * inclusive prefix sum. */
void or_finalize(const uint32_t *totals, uint32_t *prefix, int q1);

/* resolve_rate_control.comp: greedy truncation; writes quant[count32]. */
int or_resolve(const or_layout *L, const uint16_t *words, const uint8_t *bucket,
               const uint32_t *prefix, uint32_t consumed, size_t budget, int32_t *quant);

/* block_packing.comp: writes blocks in block-index order.  meta[2*b] = offset
 * in words (0 for empty blocks), meta[2*b+1] = words.  Returns total words or
 * -1 if bitstream_capacity_words is too small. */
long or_pack(const or_layout *L, const float *coeffs, const int32_t *quant, int sequence,
             uint32_t *bitstream, long bitstream_capacity_words, uint32_t *meta);

/* Full encode = dwt + quant_analyze + finalize + resolve + pack.  precision 0
 * also selects the precision-0 quantizer policy. */
long or_encode(const or_layout *L, const uint8_t *const planes[3], const int pitches[3],
               int precision, size_t budget, int sequence, int q1,
               uint32_t *bitstream, long capacity_words, uint32_t *meta, int32_t *quant_out);
long or_encode_ex(const or_layout *L, const void *const planes[3], const int pitches[3], int bits,
                  int precision, size_t budget, int sequence, int q1,
                  uint32_t *bitstream, long capacity_words, uint32_t *meta, int32_t *quant_out);

/* Encoder::Impl::compute_num_packets / packetize (pyrowave_encoder.cpp:596-623, 1048-1115). */
size_t or_compute_num_packets(const or_layout *L, const uint32_t *meta, size_t boundary);
size_t or_packetize(const or_layout *L, const uint32_t *meta, const uint32_t *bitstream, int sequence,
                    size_t boundary, uint8_t *out, size_t out_size, size_t *packets /* 2 per packet */);

/* Structural check of one coded block (Encoder::Impl::validate_bitstream,
 * pyrowave_encoder.cpp:960-1046).  Returns 1 if valid. */
int or_validate_block(const or_layout *L, const uint32_t *block_words, uint32_t block_index, uint32_t meta_words);

/* Decoder packet parser (pyrowave_decoder.cpp:126-262, 780-794, 832-840). */
typedef struct or_decoder or_decoder;
or_decoder *or_decoder_create(const or_layout *L);
void or_decoder_destroy(or_decoder *d);
void or_decoder_clear(or_decoder *d);
int or_decoder_push_packet(or_decoder *d, const void *data, size_t size);
int or_decoder_is_ready(const or_decoder *d, int allow_partial);
void or_decoder_mark_decoded(or_decoder *d);
long or_decoder_payload_words(const or_decoder *d);
const uint32_t *or_decoder_payload(const or_decoder *d);
const uint32_t *or_decoder_offsets(const or_decoder *d);
int or_decoder_decoded_blocks(const or_decoder *d);
int or_decoder_total_blocks(const or_decoder *d);

/* wavelet_dequant.comp: writes every transmitted band of `coeffs`. */
int or_dequant(const or_layout *L, const uint32_t *payload, long payload_words,
               const uint32_t *offsets, float *coeffs);

/* idwt.comp chain: levels 4 -> 0, u8 output planes with DC shift and clamp.
 * precision 1 rounds the input high bands of levels 0-1 to fp16 first (the
 * decoder's fp16 band storage); precision 3 also rounds the LL of levels 0-1,
 * which the inverse of the level above wrote (R16F in the reference);
 * precision 0 rounds every input band to fp16 and runs or_idwt_level_h. */
int or_idwt_frame(const or_layout *L, const float *coeffs, int precision,
                  uint8_t *const planes[3], const int pitches[3]);
/* bits 16: R16_UNORM output planes (decode.cpp:215), round(clamp(v + 0.5) * 65535). */
int or_idwt_frame_ex(const or_layout *L, const float *coeffs, int precision,
                     void *const planes[3], const int pitches[3], int bits);
int or_idwt_frame_f64(const or_layout *L, const double *coeffs, double *const planes[3], const int pitches[3]);

/* PSNR-HVS-M-H (psnr_hvs_m.c; reference shaders/psnr_hvs_m.comp, psnr.cpp). */
void or_dct8x8(const float sig[64], float dct[64]);
float or_maskeff(const float sig[64], const float dct[64], const float mask_coeffs[64]);
void or_psnr_hvs_m_tables(int height, float height_factor, float csf[64], float mask[64], float inv_mask[64]);
float or_psnr_height_factor(int index);
uint64_t or_psnr_hvs_m_totals(const uint8_t *a, int pitch_a, const uint8_t *b, int pitch_b, int w, int h,
                              int nfactors, uint64_t *totals);
double or_psnr_hvs_m_db(uint64_t total, uint64_t total_pixels, int full_range);

#ifdef __cplusplus
}
#endif
#endif
