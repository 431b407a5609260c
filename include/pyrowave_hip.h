/*
 * pyrowave_hip.h -- HIP-native entry points of the MI355X PyroWave codec.
 *
 * These replace the reference's Vulkan-image paths
 * (pyrowave_encoder_encode_gpu_synchronous / pyrowave_decoder_decode_gpu_buffer,
 * reference pyrowave.h:457-462 and 528-532) with device pointers and a
 * hipStream_t, in the shape of the reference's Vulkan-free Metal API
 * (reference metal/pyrowave_metal.h:175-198, 266-291): result codes keep the
 * meanings of pyrowave_result.  They add batching: one call encodes or decodes
 * frames of different sizes and chroma formats with one launch per kernel
 * stage.
 *
 * Streams are passed as `void *` holding a hipStream_t (NULL = the device's
 * own stream).  All calls are asynchronous with respect to the host unless
 * stated otherwise.
 */
#ifndef PYROWAVE_HIP_H_
#define PYROWAVE_HIP_H_

#include "pyrowave.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pyrowave_hip_precision
{
	/* The reference's precision 0 (dwt_common.h:29-34, FLOAT = float16_t;
	 * pyrowave_common.cpp:133-189, R16F at every level): the lifting steps in
	 * packed FP16 (v_pk_add_f16 + v_pk_fma_f16 on two rows or two columns at
	 * once), the K / 1/K scaling of each pass in FP32 rounded to FP16, every
	 * band of every level stored as FP16, and the precision-0 quantizer
	 * policy (6 base bits, resolution capped at 512;
	 * pyrowave_encoder.cpp:144-172).  One batch must not mix it with the other
	 * precisions. */
	PYROWAVE_HIP_PRECISION_FP16_MATH = 0,
	/* FP32 lifting; the high bands (HL, LH, HH) of levels 0-1 are stored as
	 * FP16, everything else as FP32.  Unlike the reference's precision 1
	 * (pyrowave_common.cpp:133-157) the LL intermediates of levels 0-1 stay
	 * FP32: storing them as FP16 lowers PSNR(Y) by 0.11 dB on a 4K 4:2:0
	 * sequence at 1.6 MB per frame, the high bands by nothing measurable.
	 * Quantizer policy as the reference's precision >= 1. */
	PYROWAVE_HIP_PRECISION_FP16 = 1,
	/* FP32 lifting and storage everywhere (reference precision 2). */
	PYROWAVE_HIP_PRECISION_FP32 = 2,
	/* The reference's precision 1 storage exactly: FP32 lifting, all four
	 * bands of levels 0-1 (the LL intermediates included) stored as FP16
	 * (pyrowave_common.cpp:133-157, one R16F image for levels 0-1), levels
	 * 2-4 as FP32.  Less memory traffic than PYROWAVE_HIP_PRECISION_FP16. */
	PYROWAVE_HIP_PRECISION_FP16_LL = 3,
	PYROWAVE_HIP_PRECISION_INT_MAX = 0x7fffffff
} pyrowave_hip_precision;

typedef enum pyrowave_hip_encoder_flag_bits
{
	/* Rate-control finalize computes a true prefix sum over all 2048
	 * (bucket, subdivision) entries.  Without the flag the reference's scan is
	 * reproduced, whose 8 Hillis-Steele steps leave entries at flat index
	 * >= 1024 holding 256-thread windowed sums
	 * (reference shaders/analyze_rate_control_finalize.comp:244). */
	PYROWAVE_HIP_ENCODER_EXACT_PREFIX_BIT = 1u << 0,
	/* The input planes hold 16-bit samples (R16_UNORM: u / 65535, the
	 * reference's format for 10- to 16-bit y4m input rescaled to the full
	 * range, encode.cpp:152 and yuv4mpeg.cpp:221-233).  Plane data must be
	 * 2-byte aligned, row and pixel strides even (pixel stride 2 for planar
	 * data, 4 for an interleaved P016 CbCr plane).  Such an encoder takes
	 * device planes only (pyrowave_hip_encode_batch): the CPU-buffer formats
	 * of pyrowave.h are 8-bit. */
	PYROWAVE_HIP_ENCODER_UNORM16_INPUT_BIT = 1u << 1
} pyrowave_hip_encoder_flag_bits;

typedef struct pyrowave_hip_encoder_create_info
{
	pyrowave_device device;
	int width;
	int height;
	pyrowave_chroma_subsampling chroma;
	pyrowave_hip_precision precision;
	uint32_t flags; /* pyrowave_hip_encoder_flag_bits */
} pyrowave_hip_encoder_create_info;

typedef enum pyrowave_hip_decoder_flag_bits
{
	/* The decode targets take 16-bit samples (R16_UNORM,
	 * round(clamp(v + 0.5) * 65535): the reference decodes 16-bit streams
	 * into R16_UNORM images, decode.cpp:215).  Target data 2-byte aligned,
	 * row strides even.  Device targets only, as for 16-bit encoders. */
	PYROWAVE_HIP_DECODER_UNORM16_OUTPUT_BIT = 1u << 0
} pyrowave_hip_decoder_flag_bits;

typedef struct pyrowave_hip_decoder_create_info
{
	pyrowave_device device;
	int width;
	int height;
	pyrowave_chroma_subsampling chroma;
	pyrowave_hip_precision precision;
	uint32_t flags; /* pyrowave_hip_decoder_flag_bits */
} pyrowave_hip_decoder_create_info;

/* One 8-bit plane in device memory.  pixel_stride_in_bytes is 1 for planar
 * data and 2 for an interleaved NV12 CbCr plane (pass data + 1 for Cr). */
typedef struct pyrowave_hip_plane
{
	const void *data;
	size_t row_stride_in_bytes;
	size_t pixel_stride_in_bytes;
} pyrowave_hip_plane;

typedef struct pyrowave_hip_frame
{
	pyrowave_hip_plane planes[3];
} pyrowave_hip_frame;

typedef struct pyrowave_hip_output_plane
{
	void *data;
	size_t row_stride_in_bytes;
} pyrowave_hip_output_plane;

typedef struct pyrowave_hip_decode_target
{
	pyrowave_hip_output_plane planes[3];
} pyrowave_hip_decode_target;

/* Device-resident result of an encoder's last encode. */
typedef struct pyrowave_hip_encoded_frame
{
	/* Coded 32x32 blocks, contiguous, in block-index order. */
	const uint32_t *bitstream;
	/* Two words per 32x32 block: offset into bitstream (words), word count
	 * (0 = block not coded). */
	const uint32_t *meta;
	uint32_t num_blocks;
	uint32_t sequence;
} pyrowave_hip_encoded_frame;

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_create(const pyrowave_hip_encoder_create_info *info, pyrowave_encoder *encoder);

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_decoder_create(const pyrowave_hip_decoder_create_info *info, pyrowave_decoder *decoder);

/* Encodes frames[i] with encoders[i] for i < count (encoders must be distinct
 * and belong to one device; sizes and chroma formats may differ).  Each frame
 * gets its own rate control with rate_controls[i].  Results stay on the device
 * (pyrowave_hip_encoder_get_result); pyrowave_encoder_compute_num_packets and
 * pyrowave_encoder_packetize work on them as after a CPU-buffer encode. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encode_batch(pyrowave_encoder *encoders, const pyrowave_hip_frame *frames,
                          const pyrowave_rate_control *rate_controls, uint32_t count, void *hip_stream);

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_get_result(pyrowave_encoder encoder, pyrowave_hip_encoded_frame *result);

/* Decodes the packets queued in decoders[i] (pyrowave_decoder_push_packet)
 * into targets[i].  The payload is uploaded on the given stream. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_decode_batch(pyrowave_decoder *decoders, const pyrowave_hip_decode_target *targets,
                          uint32_t count, void *hip_stream);

/* Decodes encoders[i]'s device-resident result with decoders[i] into
 * targets[i] without a host round trip (the serving path).  Frame sizes and
 * chroma formats of each pair must match. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_decode_encoded_batch(pyrowave_decoder *decoders, pyrowave_encoder *encoders,
                                  const pyrowave_hip_decode_target *targets, uint32_t count, void *hip_stream);

/* Graph mode for the batch entry points (pyrowave_hip_encode_batch,
 * pyrowave_hip_decode_batch, pyrowave_hip_decode_encoded_batch,
 * pyrowave_hip_decode_packets_batch).  The first call for a batch composition
 * (the same encoders or decoders in the same order) captures its launch
 * sequence into a hipGraph on the caller's stream; later calls rewrite the
 * graph's per-call parameter nodes (planes, budgets, sequence numbers) and
 * launch it: one launch per batch instead of one per stage.  Stage timing is
 * not recorded for graph launches.  Off by default.  No counterpart in the
 * reference (Vulkan command buffers are re-recorded per call). */
PYROWAVE_PUBLIC_API void
pyrowave_hip_device_set_graphs(pyrowave_device device, bool enable);

/* Autotuner for the transform tile shapes (forward 128x56, 64x56, 128x48 or
 * 128x40 samples per workgroup, inverse 64x64, 32x32, 64x56 or 64x48, and in
 * both directions 120x32 per wave streamed down its rows with the lifting in
 * registers).
 * Encodes the batch (and, with decoders, decodes it from the encoders'
 * results) like pyrowave_hip_encode_batch / pyrowave_hip_decode_encoded_batch,
 * timing every shape of every transform level over three rounds of
 * `repetitions` launches back to back (HIP events on the stream; synchronous)
 * and keeping the fastest per level (the default, the streamed shape 4 on
 * levels of at least 2048 of its tiles and shape 0 on smaller ones, unless
 * another is at least 2% faster) for this batch composition and for later batches of the
 * same level shapes on the device.
 * All shapes give the same coefficients bit for bit, so the call's results
 * are those of a plain encode / decode.  The chosen shape index per level
 * is written to forward_variants[5] / inverse_variants[5]
 * when they are not null.  No counterpart in the reference. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_autotune_batch(pyrowave_encoder *encoders, pyrowave_decoder *decoders, const pyrowave_hip_frame *frames,
                            const pyrowave_rate_control *rate_controls, const pyrowave_hip_decode_target *targets,
                            uint32_t count, uint32_t repetitions, void *hip_stream, int *forward_variants,
                            int *inverse_variants);

/* Forces one transform tile shape (forward 0-4, inverse 0-4; a shape a
 * direction does not have leaves that direction as it was) for every level;
 * -1 returns to the autotuned / default shapes.  For tests and comparisons. */
PYROWAVE_PUBLIC_API void
pyrowave_hip_device_force_tile_variant(pyrowave_device device, int variant);

/* ---- Packets on the device ------------------------------------------------
 *
 * The reference packetizes on the host after a fence wait
 * (pyrowave_encoder.cpp:596-623, 1048-1115, reached through
 * pyrowave_encoder_compute_num_packets / pyrowave_encoder_packetize) and
 * parses packets on the host (pyrowave_decoder.cpp:126-262, reached through
 * pyrowave_decoder_push_packet).  These entry points do both on the device,
 * so a serving loop never walks or copies a bitstream on the host. */

/* A frame's packets in device memory. */
typedef struct pyrowave_hip_packet_stream
{
	const void *data;        /* packets back to back, 4-byte aligned */
	size_t data_bytes;       /* readable bytes at data (below 4 GiB) */
	const uint32_t *packets; /* {offset, size} in bytes per packet, in push order */
	const uint32_t *count;   /* one device word: the number of packets */
	uint32_t max_packets;    /* entries the packet table holds */
} pyrowave_hip_packet_stream;

/* After each encoder's last encode (same stream): writes the sequence header
 * in front of the coded blocks, so header + blocks form one contiguous byte
 * stream, and the packet table for packet_boundary, with the split rule of
 * pyrowave_encoder_packetize.  Asynchronous. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_packetize_batch(pyrowave_encoder *encoders, uint32_t count, size_t packet_boundary,
                                     void *hip_stream);

/* Device pointers of the encoder's packetized stream: `packets` describes it
 * (data = the stream, count = the packet count word); *stream_bytes (if not
 * null) receives the device word holding the stream's length in bytes. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_get_packets(pyrowave_encoder encoder, pyrowave_hip_packet_stream *packets,
                                 const uint32_t **stream_bytes);

/* Waits for the encoder's last device packetize and copies the stream (if
 * `data` is not null and `size` holds it) and the packet table (if `packets`
 * is not null and max_packets holds it) to the host; *bytes and *num_packets
 * receive the sizes. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_read_packets(pyrowave_encoder encoder, void *data, size_t size, pyrowave_packet *packets,
                                  uint32_t max_packets, size_t *bytes, uint32_t *num_packets);

/* Parses each decoder's packets on the device and decodes the frame into the
 * target.  Equivalent to pyrowave_decoder_clear then
 * pyrowave_decoder_push_packet of every packet in order, for packets that all
 * carry one sequence number (one frame's packets: the reference's restart on
 * a newer sequence and its dropping of stale packets are not done here; a
 * second sequence number is a parse error).  A packet that push_packet would
 * reject is a parse error too, and its units before the rejected one count,
 * as in the reference: a block shorter than its 8-byte header is present
 * (first copy in push order wins) and ends its packet.  Errors are reported
 * by pyrowave_hip_decoder_parse_result; the frame still decodes from the
 * blocks accepted, as a reference decoder would after a rejected push.
 * Missing blocks decode as zero.  Asynchronous. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_decode_packets_batch(pyrowave_decoder *decoders, const pyrowave_hip_packet_stream *packets,
                                  const pyrowave_hip_decode_target *targets, uint32_t count, void *hip_stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (a serving loop that must not call hipMalloc, a harness that forbids
 * it).  pyrowave_hip_decode_packets_batch for one decoder that allocates
 * nothing: the parse's block offset table lives in a static device array of
 * the library (131072 entries, 512 KiB per GPU, held until the process ends),
 * so frames of more than 131072 32x32 blocks (beyond 8192x8192 4:4:4) are
 * refused with PYROWAVE_ERROR_INVALID_ARGUMENT.  The parse's info words
 * {blocks present, blocks announced, error bits, sequence} go to
 * device_info[4] (device memory) when it is not null; the call does not
 * update pyrowave_hip_decoder_parse_result.  Every call on a GPU uses that
 * GPU's one array: calls that overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller.  Never captured in a
 * graph and not timed.  Asynchronous. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_decode_packets_noalloc(pyrowave_decoder decoder, const pyrowave_hip_packet_stream *packets,
                                    const pyrowave_hip_decode_target *target, uint32_t *device_info,
                                    void *hip_stream);

/* Waits for the decoder's last device parse.  Blocks present and blocks the
 * sequence header announced; PYROWAVE_ERROR_INVALID_ARGUMENT after a parse
 * error, PYROWAVE_ERROR_GENERIC if there was no device parse. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_decoder_parse_result(pyrowave_decoder decoder, uint32_t *blocks_present, uint32_t *blocks_announced);

/* ---- PSNR-HVS-M-H quality metric ----------------------------------------
 *
 * Reference psnr.cpp:126-264 with shaders/psnr_hvs_m.comp: 8x8 DCT blocks at a
 * stride of 4 over a luma plane, masking effect, CSF tables for 16 viewing
 * height factors 1 + i/8.  One launch covers all 16 factors (the reference
 * dispatches once per factor). */

/* Adds, per height factor, the frame's 24-bit fixed-point block errors to
 * device_totals[16] (device memory, caller-cleared) and the frame's pixel
 * count (blocks x 64) to *total_pixels.  Planes are u8 in device memory.
 * Texels past the image edge read as 0.  Asynchronous. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_psnr_hvs_m_accumulate(pyrowave_device device, const void *reference, size_t reference_stride,
                                   const void *distorted, size_t distorted_stride, int width, int height,
                                   uint64_t *device_totals, uint64_t *total_pixels, void *hip_stream);

/* PSNR-HVS-M-H in dB per height factor from totals copied to the host
 * (full_range false: peak 223/255, studio swing). */
PYROWAVE_PUBLIC_API void
pyrowave_hip_psnr_hvs_m_db(const uint64_t *totals, uint64_t total_pixels, bool full_range, double *db);

/* ---- Introspection (synchronous; for parity tests and tools) ---- */

/* Number of floats of the flat coefficient layout for a frame geometry:
 * for component 0..2, level 0..4 (chroma level 0 absent under 4:2:0),
 * band 0..3 (LL, HL, LH, HH), a dense row-major band of
 * (aligned_width/2 >> level) x (aligned_height/2 >> level) values, where
 * aligned = max(align(size, 32), 128).  LL bands below level 4 are the
 * transform's intermediates. */
PYROWAVE_PUBLIC_API size_t
pyrowave_hip_coefficient_count(int width, int height, pyrowave_chroma_subsampling chroma);

/* Copies the encoder's wavelet bands of the last encode, as float. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_read_coefficients(pyrowave_encoder encoder, float *coefficients, size_t count);

/* Quantizes, rate-controls and packs caller-supplied coefficients (flat
 * layout, converted to the encoder's storage precision) instead of running
 * the forward transform: the reference's Encoder::encode_pre_transformed
 * (reference pyrowave_encoder.cpp:1138-1151) with the encoder's own quantizer
 * resolutions. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_encode_coefficients(pyrowave_encoder encoder, const float *coefficients, size_t count,
                                         const pyrowave_rate_control *rate_control);

/* Rate-control state of the last encode.  Any pointer may be NULL.
 * words/buckets: 16 entries per 32x32 block (quant level 0..15);
 * totals/prefix: 2048 entries; quant: final truncation per block. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_encoder_read_rate_control(pyrowave_encoder encoder, uint16_t *words, uint8_t *buckets,
                                       uint32_t *totals, uint32_t *prefix, uint32_t *consumed,
                                       int32_t *quant);

/* Copies the decoder's dequantized bands of the last decode, as float. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_decoder_read_coefficients(pyrowave_decoder decoder, float *coefficients, size_t count);

/* Per-stage GPU time in milliseconds, summed since the last reset, and the
 * number of frames it covers.  Stage order: DWT, Quant, Analyze, Resolve,
 * Packing, Dequant, iDWT (Quant and Analyze run as one fused kernel; its time
 * is reported under Quant and Analyze reads 0). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_device_query_stage_times(pyrowave_device device, double ms[7], uint64_t frames[7], bool reset);

/* Enables or disables per-stage event timing (on by default; off removes the
 * event records from the launch sequence). */
PYROWAVE_PUBLIC_API void
pyrowave_hip_device_set_timing(pyrowave_device device, bool enable);

/* Event timing for a subset of stages: bit s of stage_mask enables stage s
 * (order as in pyrowave_hip_device_query_stage_times).  Two event records per
 * timed stage and call. */
PYROWAVE_PUBLIC_API void
pyrowave_hip_device_set_timing_stages(pyrowave_device device, uint32_t stage_mask);

/* ---- Serving loop ------------------------------------------------------
 *
 * No counterpart in the reference: its callers (pyrofling) run one encoder
 * per stream themselves.  A server owns one encoder + decoder pair per
 * session (device-resident state kept between frames), admits sessions by
 * their HBM footprint against a budget, and batches continuously: whenever a
 * stream is free, the frames ready at that moment (the oldest queued frame of
 * each session whose previous frame is done, up to max_batch_frames) go to
 * the device as one batch, one launch per stage, so a request that arrives
 * while others are on the GPU joins the next batch and its work overlaps
 * theirs on another stream.  Each finished frame leaves its packetized
 * bitstream on the host (and its decode in the request's target). */
typedef struct pyrowave_hip_server_opaque *pyrowave_hip_server;

typedef struct pyrowave_hip_server_create_info
{
	pyrowave_device device;
	size_t hbm_budget_bytes;    /* admission limit for session state */
	uint32_t streams;           /* concurrent streams (0: 4) */
	pyrowave_hip_precision precision;
	size_t packet_boundary;     /* packet split for the host bitstreams (0: one packet) */
	uint32_t max_batch_frames;  /* frames per launch (0: 16) */
} pyrowave_hip_server_create_info;

typedef struct pyrowave_hip_server_completion
{
	uint64_t ticket;
	uint32_t session;
	pyrowave_result status;
	size_t bytes;               /* packetized bitstream size */
	uint32_t packets;
	uint64_t submit_ns, start_ns, done_ns; /* steady-clock times */
} pyrowave_hip_server_completion;

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_server_create(const pyrowave_hip_server_create_info *info, pyrowave_hip_server *server);

/* Waits for all work, closes every session. */
PYROWAVE_PUBLIC_API void
pyrowave_hip_server_destroy(pyrowave_hip_server server);

/* HBM bytes one session of this geometry holds (encoder + decoder). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_server_session_footprint(pyrowave_hip_server server, int width, int height,
                                      pyrowave_chroma_subsampling chroma, size_t *bytes);

/* Admits a session if its footprint fits the remaining budget; otherwise
 * PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY (retry once sessions close). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_server_open_session(pyrowave_hip_server server, int width, int height,
                                 pyrowave_chroma_subsampling chroma, uint32_t *session);

/* Waits for the session's queued frames, then frees its state. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_server_close_session(pyrowave_hip_server server, uint32_t session);

/* Queues one frame of a session (device planes, budget; target may be null to
 * skip the decode).  Frames of one session run in submission order.  The
 * planes must stay valid until the frame completes. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_server_submit(pyrowave_hip_server server, uint32_t session, const pyrowave_hip_frame *frame,
                           size_t maximum_bitstream_size, const pyrowave_hip_decode_target *target,
                           uint64_t *ticket);

/* Returns up to `max` finished frames (oldest first); waits up to timeout_us
 * for the first one. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_server_poll(pyrowave_hip_server server, pyrowave_hip_server_completion *completions, uint32_t max,
                         uint32_t *count, uint32_t timeout_us);

/* Copies a finished frame's packetized bitstream (and its packet table, two
 * size_t per packet, if `packets` is not null) out and releases it. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_server_take_bitstream(pyrowave_hip_server server, uint64_t ticket, void *data, size_t size,
                                   pyrowave_packet *packets, uint32_t max_packets);

/* ---- External memory: zero-copy frames shared with other APIs -----------
 *
 * The reference imports memory of other APIs as a Vulkan image from an
 * external handle and takes plane views of it (pyrowave_image_create,
 * pyrowave_image_get_image_view, pyrowave_image_destroy: pyrowave.h:269-352,
 * pyrowave_c.cpp:524-792).  Here shared memory is a linear device buffer
 * passed as a POSIX file descriptor; planes inside it are described by a
 * layout and become an encoder source (pyrowave_hip_frame) or a decoder
 * target (pyrowave_hip_decode_target) without a copy. */

typedef struct pyrowave_hip_external_memory_opaque *pyrowave_hip_external_memory;

typedef enum pyrowave_hip_external_handle_type
{
	/* A POSIX fd of an exported device allocation (HIP or Vulkan opaque fd, or
	 * a dma-buf). */
	PYROWAVE_HIP_EXTERNAL_HANDLE_FD = 1,
	PYROWAVE_HIP_EXTERNAL_HANDLE_INT_MAX = 0x7fffffff
} pyrowave_hip_external_handle_type;

typedef struct pyrowave_hip_external_memory_import_info
{
	pyrowave_device device;
	/* Not consumed: the library imports a duplicate, the caller still owns
	 * (and may close) this descriptor. */
	int fd;
	pyrowave_hip_external_handle_type handle_type;
	/* Bytes of the exported allocation. */
	uint64_t size;
} pyrowave_hip_external_memory_import_info;

typedef enum pyrowave_hip_plane_layout
{
	/* Y, Cb, Cr planes, one byte per sample (I420 / I444). */
	PYROWAVE_HIP_PLANES_THREE = 0,
	/* Y plane, then one plane of interleaved Cb, Cr bytes (NV12 / NV24);
	 * encoder input only. */
	PYROWAVE_HIP_PLANES_INTERLEAVED_CHROMA = 1,
	PYROWAVE_HIP_PLANES_INT_MAX = 0x7fffffff
} pyrowave_hip_plane_layout;

typedef struct pyrowave_hip_external_image_layout
{
	/* Luma size (even under 4:2:0). */
	uint32_t width;
	uint32_t height;
	pyrowave_chroma_subsampling chroma;
	pyrowave_hip_plane_layout layout;
	/* Byte offset of each plane from the start of the memory and its row
	 * pitch in bytes (interleaved chroma: planes 0 and 1 only). */
	uint64_t plane_offset[3];
	uint64_t row_pitch[3];
} pyrowave_hip_external_image_layout;

/* Maps memory exported elsewhere (hipImportExternalMemory).  Replaces
 * pyrowave_image_create (reference pyrowave.h:330-331) for this backend.
 * PYROWAVE_ERROR_FAILED_EXTERNAL_HANDLE when the runtime refuses the handle. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_import(const pyrowave_hip_external_memory_import_info *info,
                                    pyrowave_hip_external_memory *memory);

/* Allocates device memory that other processes or APIs can import and
 * returns a POSIX fd for it (owned by the caller; close it when it has been
 * handed over).  The size is rounded up to the allocation granularity
 * (pyrowave_hip_external_memory_size).  No counterpart in the reference, which
 * only imports; a producer that cannot allocate shareable memory itself
 * writes into this one. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_allocate(pyrowave_device device, uint64_t size, pyrowave_hip_external_memory *memory,
                                      int *exported_fd);

/* Device address and size of the mapping (for the caller's own copies). */
PYROWAVE_PUBLIC_API void *pyrowave_hip_external_memory_device_address(pyrowave_hip_external_memory memory);
PYROWAVE_PUBLIC_API uint64_t pyrowave_hip_external_memory_size(pyrowave_hip_external_memory memory);

/* Encoder source / decoder target over planes in the memory; every plane must
 * lie inside it (PYROWAVE_ERROR_INVALID_ARGUMENT otherwise).  Replaces
 * pyrowave_image_get_image_view (reference pyrowave.h:347-349). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_frame(pyrowave_hip_external_memory memory, const pyrowave_hip_external_image_layout *layout,
                                   pyrowave_hip_frame *frame);
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_target(pyrowave_hip_external_memory memory,
                                    const pyrowave_hip_external_image_layout *layout,
                                    pyrowave_hip_decode_target *target);

/* Unmaps and releases the memory (replaces pyrowave_image_destroy, reference
 * pyrowave.h:351).  Encodes and decodes that use it must have completed. */
PYROWAVE_PUBLIC_API void pyrowave_hip_external_memory_destroy(pyrowave_hip_external_memory memory);

#ifdef __cplusplus
}
#endif

#endif
