// pw_host.hpp -- host-side objects of the MI355X PyroWave codec.
#pragma once

#include "pw_common.hpp"

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "pyrowave.h"
#include "pyrowave_hip.h"

namespace pw
{
inline bool valid_precision(int precision)
{
	return precision == PYROWAVE_HIP_PRECISION_FP16 || precision == PYROWAVE_HIP_PRECISION_FP32 ||
	       precision == PYROWAVE_HIP_PRECISION_FP16_LL || precision == PYROWAVE_HIP_PRECISION_FP16_MATH;
}

enum Stage
{
	kStageDWT = 0,
	kStageQuant,
	kStageAnalyze,
	kStageResolve,
	kStagePacking,
	kStageDequant,
	kStageIDWT,
	kStageCount
};

// Band geometry of one frame (reference pyrowave_common.cpp:191-254,
// bitstream/bitstream.md:99-152, 487-511).
struct BandGeom
{
	bool present = false; // transmitted (has block indices)
	bool exists = false;  // produced by the transform (incl. intermediate LL)
	int bw = 0, bh = 0, bx8 = 0, by8 = 0, bx32 = 0, by32 = 0;
	int off8 = 0, off32 = 0;
	long flat_off = 0;      // offset in the flat float coefficient layout
	size_t store_off = 0;   // byte offset in the band storage allocation
	int pitch = 0;          // storage pitch (elements), multiple of 32
	int kind = kStoreF32;
};

struct Geometry
{
	int width = 0, height = 0, chroma = 0;
	int aligned_w = 0, aligned_h = 0;
	int count8 = 0, count32 = 0;
	int per_sub = 0, shamt = 0, ngroups = 0;
	long flat_count = 0;
	size_t store_bytes = 0;
	int precision = PYROWAVE_HIP_PRECISION_FP32;
	int sample_bytes = 1; // encoder input / decoder output planes: 1 (UNORM8) or 2 (R16_UNORM)
	BandGeom band[kComponents][kLevels][kBands];

	bool init(int w, int h, int chroma, int precision);
	int plane_w(int c) const { return (c && chroma == 0) ? width / 2 : width; }
	int plane_h(int c) const { return (c && chroma == 0) ? height / 2 : height; }
	int first_level(int c) const { return (c && chroma == 0) ? 1 : 0; }
	uint32_t capacity_words() const;
};

// A caller's input plane c fits the encoder's geometry: data present, rows
// long enough for the plane's width at the pixel stride (default: one
// sample), and for 16-bit samples data and strides 2-byte aligned.
inline bool plane_fits(const Geometry &G, int c, const pyrowave_hip_plane &p)
{
	const size_t sb = size_t(G.sample_bytes);
	const size_t step = p.pixel_stride_in_bytes ? p.pixel_stride_in_bytes : sb;
	if (!p.data || step < sb || p.row_stride_in_bytes < (size_t(G.plane_w(c)) - 1) * step + sb)
		return false;
	return sb == 1 || ((reinterpret_cast<uintptr_t>(p.data) | p.row_stride_in_bytes | step) & 1u) == 0;
}

inline bool frame_fits(const Geometry &G, const pyrowave_hip_frame &f)
{
	for (int c = 0; c < 3; c++)
		if (!plane_fits(G, c, f.planes[c]))
			return false;
	return true;
}

// A decode target's planes fit the decoder's geometry (rows of plane-width
// samples; 16-bit targets 2-byte aligned).
inline bool target_fits(const Geometry &G, const pyrowave_hip_decode_target &t)
{
	const size_t sb = size_t(G.sample_bytes);
	for (int c = 0; c < 3; c++)
	{
		const pyrowave_hip_output_plane &p = t.planes[c];
		if (!p.data || p.row_stride_in_bytes < size_t(G.plane_w(c)) * sb)
			return false;
		if (sb == 2 && ((reinterpret_cast<uintptr_t>(p.data) | p.row_stride_in_bytes) & 1u))
			return false;
	}
	return true;
}

struct DeviceBuffer
{
	void *ptr = nullptr;
	size_t size = 0;
	~DeviceBuffer();
	bool alloc(size_t bytes);
	void release();
};

struct HostBuffer
{
	void *ptr = nullptr;
	size_t size = 0;
	~HostBuffer();
	bool alloc(size_t bytes);
	void release();
};

struct TimingRecord
{
	hipEvent_t start, end;
	int stage;
	int frames;
};

struct Device
{
	int hip_device = 0;
	hipStream_t stream = nullptr;
	std::recursive_mutex lock;
	uint32_t timing = (1u << kStageCount) - 1; // stages with event timing
	double stage_ms[kStageCount] = {};
	uint64_t stage_frames[kStageCount] = {};
	std::vector<TimingRecord> pending;
	std::vector<hipEvent_t> event_pool;
	std::atomic<int> live_objects{0};
	bool graphs = false; // batch entry points replay a captured graph per batch composition
	// Autotuned transform tile variants, keyed by (direction, level, tiles of
	// variant 0); new plans of the same shape start from these.
	std::map<uint64_t, int> tuned_variant;
	int forced_variant = -1; // >= 0: every transform level uses this variant
	bool fused_resolve = true; // rate control of a frame in one launch (k_rc_resolve) where it fits
	// A component's first two transform levels in one launch, forward
	// (k_dwt_forward_fused) and inverse (k_dwt_inverse_fused): 0 never, 1 the
	// forward for launches of at least kFusedMinTiles tiles and the inverse
	// always, 2 both always.
	int fused = 1;
	int tune_reps = 0;       // > 0: the next eager encode / decode autotunes
	// PSNR-HVS-M-H tables per image height (device copies).
	std::map<int, std::unique_ptr<DeviceBuffer>> psnr_tables;

	hipEvent_t get_event();
	void begin_stage(hipStream_t s, int stage, int frames, hipEvent_t *start);
	void end_stage(hipStream_t s, int stage, int frames, hipEvent_t start);
	void resolve_timings(bool wait);
	~Device();
};

struct Plan;

// Byte offsets of a slot's parts inside its one device allocation.
struct SlotLayout
{
	size_t bands, words, words0, bucket, partial_quant, quant, excl, group_sum, meta, totals, prefix, cls, required, rc,
	    bitstream, packets, total;
};
SlotLayout slot_layout(const Geometry &G, bool encoder);

// One frame slot: everything one encode or decode of a frame needs on the
// device, plus the host mirrors of the CPU-buffer paths.
struct Slot
{
	Device *dev = nullptr;
	Geometry geom;
	bool encoder = true;
	bool exact_prefix = false;

	DeviceBuffer mem;        // bands + rate-control arrays + bitstream/payload
	FrameState fs = {};      // static fields (device pointers into mem)
	uint8_t *bands = nullptr;

	// CPU-buffer path staging.
	DeviceBuffer dplanes;    // u8 input (encoder) or output (decoder) planes
	HostBuffer hplanes;      // pinned staging of the planes
	hipEvent_t staging_done = nullptr;

	// Encoder results mirrored to the host.
	HostBuffer hmeta;        // n32 * 2 words + rc words
	HostBuffer hbits;        // bitstream words
	uint32_t hbits_words = 0;
	// Device packetize output: the sequence header sits in the two words
	// before fs.bitstream; table [n32 + 1][2] {offset, size}, then 2 info
	// words {stream bytes, packets}.
	uint32_t *dtable = nullptr;
	uint32_t *dpinfo = nullptr;
	bool device_packets = false; // a device packetize has been queued
	hipEvent_t result_ready = nullptr;
	bool have_result = false;
	bool result_on_host = false;
	uint32_t sequence = 0;   // 3-bit frame counter (reference pyrowave_encoder.cpp:1155)
	size_t last_budget = 0;

	// Decoder packet state (reference pyrowave_decoder.cpp:34-39, 126-262).
	std::vector<uint32_t> offsets;
	std::vector<uint32_t> payload;
	int decoded_blocks = 0;
	int total_blocks = 0;
	uint32_t last_seq = UINT32_MAX;
	bool decoded_frame = false;
	DeviceBuffer dparse;     // device parse: offsets [n32], then 4 info words
	bool parsed_on_device = false;
	DeviceBuffer dpayload;   // device payload (+16 B padding)
	DeviceBuffer doffsets;
	HostBuffer hpayload;     // pinned staging for payload + offsets

	std::unique_ptr<Plan> single_plan;

	bool init(Device *dev, int w, int h, int chroma, int precision, bool encoder);
	~Slot();

	// decoder parser
	void clear_packets();
	bool push_packet(const void *data, size_t size);
	bool decode_is_ready(bool allow_partial) const;
};

// Reusable memory for launch tables: a caller that builds a plan per batch
// (a serving loop) keeps one of these per stream instead of allocating.  The
// previous plan built into it must have finished on the device.
struct PlanStorage
{
	DeviceBuffer dev;
	HostBuffer host;
};

// Launch tables for a list of slots of one kind; cached per composition.
struct Plan
{
	std::vector<Slot *> slots;
	bool encoder = true;
	bool half_math = false; // FP16-arithmetic frames (transform variant kHalfStreamVariant only)
	DeviceBuffer mem;
	FrameState *d_frames = nullptr;

	// Transform launch tables per level and tile variant; variant[l] is the
	// one the launches use (the autotuner's choice, 0 by default).
	DwtJob *d_dwt_jobs[kLevels][kDwtVariants] = {};
	uint16_t *d_tile_job[kLevels][kDwtVariants] = {};
	int tiles[kLevels][kDwtVariants] = {};
	int variant[kLevels] = {};
	// Plans of FP32-arithmetic frames: the fused first two levels of every
	// component, and the per-level tables of the levels that remain.
	bool has_fused = false;
	FusedJob *d_fused_jobs = nullptr;
	uint16_t *d_fused_tile_job = nullptr;
	int fused_tiles = 0;
	int inv_fused_tile_h = kInvFusedShortH; // decoder plans: kInvFusedTallH, kInvFusedShortH or kInvFusedLoneH
	int inv_fused_short_tiles = 0;          // decoder plans: fused tiles at kInvFusedShortH rows
	DwtJob *d_rest_jobs[kLevels][kDwtVariants] = {};
	uint16_t *d_rest_tile_job[kLevels][kDwtVariants] = {};
	int rest_tiles[kLevels][kDwtVariants] = {};
	BandJob *d_bands = nullptr;
	uint32_t *d_blk_band = nullptr;
	int32_t *d_frame_block_base = nullptr;
	uint32_t *d_grp_frame = nullptr;
	int32_t *d_frame_group_base = nullptr;
	int nblocks = 0;
	int ngroups = 0;
	int max_n32 = 0;

	// Without storage the plan owns its tables (synchronous upload); with
	// storage they go into it, uploaded on `stream`.
	bool build(const std::vector<Slot *> &slots, bool encoder, PlanStorage *storage = nullptr,
	           hipStream_t stream = nullptr);

	// Graph mode: the plan's launch sequence, captured once and replayed with
	// the per-call parameter nodes updated.
	hipGraph_t graph = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	std::vector<hipGraphNode_t> param_nodes;
	std::vector<hipKernelNodeParams> param_node_base;
	std::vector<int> param_first, param_count;
	~Plan();
	// Writes the per-call fields (call(i)) into the device FrameState array.
	FrameState *stage_frames(hipStream_t stream);
	CallParams &call(int i) { return params[i]; }
	std::vector<CallParams> params;
};

// Shared helpers.  precision0: the quantizer policy of the reference's
// precision 0 (PYROWAVE_HIP_PRECISION_FP16_MATH).
float quant_resolution(int level, int component, int band, bool precision0);
float rdo_distortion_scale(int chroma, int level, int component, int band, bool precision0);
uint8_t encode_quant(float decoder_q_scale);
float decode_quant(uint8_t code);

// `timed`: record the device's per-stage timing events.
void run_encode(Plan &plan, hipStream_t stream, bool timed = true);
void run_decode(Plan &plan, hipStream_t stream, bool timed = true, bool capture = false);
// Replays the plan's captured launch sequence (capturing it on first use).
bool run_graph(Plan &plan, hipStream_t stream);
// Autotuner: while Device::tune_reps > 0, run_encode / run_decode time every
// tile variant of each transform level (three rounds of `tune_reps` launches
// back to back, HIP events on the launch stream, synchronous) and keep the
// fastest per level
// (the default shape unless another is 2% faster) in the plan and in
// Device::tuned_variant.  A level's launch does not modify
// its own inputs, so repeating it leaves the call's results valid.
// Key of Device::tuned_variant.
uint64_t tune_key(bool encoder, int level, int tiles0);

// Budget in bytes -> words of coded blocks (the sequence header comes off).
int target_words_of(size_t target_bytes);
// Per-call fields of an encode (advances the slot's 3-bit sequence counter)
// and of a decode that reads an encoder slot's device result.
void fill_encode_call(CallParams &F, Slot &S, const pyrowave_hip_frame &frame, size_t budget);
void fill_decode_call(CallParams &F, Slot &S, const Slot *encoder, const pyrowave_hip_decode_target &target);
// Copies an encoder slot's meta, rc words and first `words` bitstream words to
// its pinned host mirrors on `stream`.
bool queue_result_download(Slot &S, hipStream_t stream, uint32_t words);
} // namespace pw

struct pyrowave_device_opaque
{
	pw::Device dev;
};

struct pyrowave_encoder_opaque
{
	pw::Slot slot;
	pyrowave_device device = nullptr;
};

struct pyrowave_decoder_opaque
{
	pw::Slot slot;
	pyrowave_device device = nullptr;
};
