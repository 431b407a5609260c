// pw_common.hpp -- definitions shared by the host code and the HIP kernels.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace pw
{
// Geometry (reference pyrowave_common.hpp:112-117, bitstream/bitstream.md:103-109).
constexpr int kLevels = 5;
constexpr int kComponents = 3;
constexpr int kBands = 4;
constexpr int kAlignment = 1 << kLevels;
constexpr int kMinimumImageSize = 4 << kLevels;

// Rate control (reference pyrowave_encoder.cpp:17-19).
constexpr int kSubdiv = 16;
constexpr int kBuckets = 128;
constexpr int kFlat = kBuckets * kSubdiv;
// The quantizer adds each block's savings into one of several copies of the
// 2048 totals (block b into copy b % copies), and the finalize sums the
// copies: a frame's blocks meet on few (bucket, subdivision) entries, and
// device-scope atomics on one address serialize (a lone 4K frame's quantizer
// took 39 us with one copy, 20 with sixteen).  One copy per
// kBlocksPerTotalsCopy blocks, a power of two up to kTotalsCopies (a lone
// 1080p frame, 3261 blocks, keeps one).  Integer sums in any order, so the
// totals are the same.
constexpr int kTotalsCopies = 16;
constexpr int kBlocksPerTotalsCopy = 4096;
constexpr int totals_copies(int n32)
{
	int c = 1;
	while (c < kTotalsCopies && c * kBlocksPerTotalsCopy < n32)
		c *= 2;
	return c;
}
static_assert(totals_copies(3261) == 1 && totals_copies(12429) == 4 && totals_copies(1 << 20) == kTotalsCopies,
              "copies by frame size");

// DWT tiles, one workgroup each, with a 4-sample apron on every side.  Two
// tile shapes per direction; the autotuner picks one per level (both give
// the same coefficients, bit for bit).
// Forward: W x H samples of the level domain ((W + 8) x (H + 8) staged); the
// horizontal pass is H + 8 rows x W/32 windows of 40 and the vertical pass W
// columns x 2 windows of H/2 + 8: 2W threads.  Variant 0: 128 x 56 (256
// threads), 1: 64 x 56 (128 threads); 2: 128 x 48 and 3: 128 x 40 (256
// threads, u8 conversion without an LDS table: 31 / 27 KB of LDS, five / six
// workgroups per CU instead of four).
// Inverse: W x H output samples ((W + 8) x (H + 8) staged).  Variant 0:
// 64 x 64 (256 threads), 1: 32 x 32 (128 threads), 2: 64 x 56 (256 threads,
// 19 KB of LDS: eight workgroups per CU, not seven), all with 16-sample
// windows; 3: 64 x 48 (256 threads) with 24-sample windows and each pass's
// scaling applied once per sample (a third fewer lifting instructions).
// Variant 4 of each direction streams instead: one wave per 120 x 32 tile
// walks down its rows with the lifting in registers (no LDS).  Variant 5 is
// the same streamed shape in packed FP16 arithmetic; it is the only shape of
// an FP16-arithmetic plan, and never one of another plan.
constexpr int kApron = 4;
constexpr int kFwdVariants = 5;  // FP32-arithmetic shapes
constexpr int kInvVariants = 5;
constexpr int kDwtVariants = 6;  // table slots per level
constexpr int kFwdStreamVariant = 4;
constexpr int kInvStreamVariant = 4;
constexpr int kHalfStreamVariant = 5;
// Shape of a level that has not been autotuned: the streamed one when the
// level has at least kStreamMinTiles of its tiles (six waves per CU; it is
// the fastest on the large levels of a batch and on level 1 of a lone 4K
// 4:2:0 frame, 1632 tiles), else a tiled shape (a
// streamed wave walks its rows one after another, which a level with few
// tiles cannot hide: one 1080p frame took 176 instead of 129 us): the small
// tiles the autotuner keeps on the levels of lone 720p, 1080p and 4K frames,
// forward 32x32 (shape 1) from 512 streamed tiles and 32x16 (shape 2) below,
// inverse 32x32 (shape 1) from 128 and 16x16 (shape 2) below (a lone 1080p
// FP16 frame 118 -> 101 us against shape 0 everywhere).  The autotuner keeps
// the default unless another shape is 2% faster.
constexpr int kStreamMinTiles = 1536;
static_assert(kFwdStreamVariant == kInvStreamVariant, "one streamed shape index");
constexpr int default_dwt_variant(bool forward, int stream_tiles)
{
	return stream_tiles >= kStreamMinTiles ? kFwdStreamVariant
	       : forward                       ? (stream_tiles >= 512 ? 1 : 2)
	                                       : (stream_tiles >= 128 ? 1 : 2);
}
constexpr int kFwdVariantTileW[kDwtVariants] = {128, 32, 32, 128, 120, 120};
constexpr int kFwdVariantTileH[kDwtVariants] = {56, 32, 16, 40, 32, 32};
constexpr int kInvVariantTileW[kDwtVariants] = {64, 32, 16, 64, 120, 120};
constexpr int kInvVariantTileH[kDwtVariants] = {64, 32, 16, 48, 32, 32};

// Largest coded 32x32 block: 16 8x8 blocks x 8 subblocks x 14 planes, 1024
// sign bits, 8 B header and 3 B per 8x8 block, rounded up to words.
constexpr int kMaxBlockWords = 512;

enum StorageKind : int32_t
{
	kStoreF32 = 0,
	kStoreF16 = 1,
	kStoreU8 = 2,
	kStoreU16 = 3 // R16_UNORM planes (16-bit input / output)
};

// Per-(bucket, subdivision) resolve class.
enum RcClass : uint8_t
{
	kRcSkip = 0,
	kRcFull = 1,
	kRcPartial = 2
};

// Error bits reported by the device in FrameState::rc[kRcError].
enum DeviceError : uint32_t
{
	kErrWordMismatch = 1u << 0,   // packed words != analyzed words
	kErrCapacity = 1u << 1        // bitstream buffer too small
};

// Indices into FrameState::rc.
enum RcWord
{
	kRcConsumed = 0,
	kRcError = 1,
	kRcCount = 4
};

// One frame slot of a batch: every device array the encode or decode of
// that frame touches.  Lives in device memory; kernels index it by slot.
struct FrameState
{
	// Encoder rate-control arrays.
	uint16_t *words;        // [n32][16] words at each quant level
	uint16_t *words0;       // [n32 rounded up to 8] words at quant level 0
	uint8_t *bucket;        // [n32][16] inclusive R-D bucket
	int32_t *partial_quant; // [n32] truncation chosen by partially applied buckets
	uint8_t *quant;         // [n32] final truncation
	uint32_t *excl;         // [n32] exclusive word prefix within a 256-block group
	uint32_t *group_sum;    // [ngroups] words per 256-block group
	uint32_t *meta;         // [n32][2] offset (words), words
	uint32_t *bitstream;    // coded blocks in block-index order
	uint32_t *totals;       // [totals_copies][2048] savings per (bucket, subdivision)
	uint32_t *prefix;       // [2048] finalized scan
	uint8_t *cls;           // [2048] RcClass
	int32_t *required;      // [2048] savings still required at each entry
	uint32_t *rc;           // [kRcCount] RcWord
	// Per-call encoder inputs: u8 or u16 planes (device pointers; pitch and
	// step in samples), sequence counter and word budget for the coded blocks.
	const uint8_t *src[3];
	int32_t src_pitch[3];
	int32_t src_step[3];
	uint32_t sequence;
	int32_t target_words;
	// Per-call decoder outputs: u8 or u16 planes (device pointers; pitch in
	// samples).
	uint8_t *dst[3];
	int32_t dst_pitch[3];
	// Decoder input: offsets (u32 per block, or meta pairs) and payload.
	const uint32_t *dec_offsets;
	const uint32_t *dec_payload;
	int32_t dec_offset_stride;   // 1: offset table (UINT32_MAX = missing); 2: meta pairs
	uint32_t dec_payload_bytes;
	int32_t n32;
	int32_t per_sub;
	int32_t shamt;
	int32_t ngroups;
	int32_t totals_copies; // totals_copies(n32)
	uint32_t capacity_words;
	int32_t width, height;
	uint32_t flags;   // bit 0: exact prefix in rate-control finalize
};

// Per-call fields of one frame, passed to k_set_params as kernel arguments
// (no host-to-device copy on the launch stream).
struct CallParams
{
	const uint8_t *src[3];
	uint8_t *dst[3];
	const uint32_t *dec_offsets;
	const uint32_t *dec_payload;
	int32_t src_pitch[3];
	int32_t src_step[3];
	int32_t dst_pitch[3];
	uint32_t sequence;
	int32_t target_words;
	uint32_t flags;
	int32_t dec_offset_stride;
	uint32_t dec_payload_bytes;
};

constexpr int kParamsPerLaunch = 24;

struct ParamBlock
{
	CallParams p[kParamsPerLaunch];
};

// One component of one frame at one DWT level (forward or inverse).
struct DwtJob
{
	const void *src;    // forward: LL of the previous level (null: u8 plane from FrameState)
	void *dst;          // inverse: LL of the next level up, or u8 plane (null: from the decode target)
	void *band[4];      // forward: outputs, inverse: inputs
	int32_t src_kind;   // forward input storage (kStoreU8 for planes)
	int32_t dst_kind;   // inverse output storage
	int32_t band_kind;  // storage of the high bands (HL, LH, HH)
	int32_t ll_kind;    // storage of the LL band
	int32_t src_pitch;  // elements
	int32_t dst_pitch;  // elements (bytes for u8)
	int32_t band_pitch; // elements
	int32_t R_x, R_y;   // forward: valid input size; inverse: valid output size
	int32_t A_x, A_y;   // forward: transform domain; inverse: 2 * band size
	int32_t bw, bh;     // band size
	int32_t tiles_x;
	int32_t tile_base;
	int32_t frame;      // frame slot
	int32_t component;
};

// The first two levels (l, l + 1) of one component of one frame in one
// forward launch: a wave streams a column strip of level l and feeds the LL
// rows it produces straight into level l + 1, so LL_l never leaves the
// registers (it is not written).  A tile is kFusedTileW1 x kFusedTileH1
// level-(l+1) coefficients per band.
struct FusedJob
{
	DwtJob a; // level l (its source is the component's input; its LL is not written)
	DwtJob b; // level l + 1 (b.src unused)
	int32_t tiles_x;
	int32_t tile_base;
};
constexpr int kFusedTileW1 = 26;
constexpr int kFusedTileH1 = 32;
// A fused wave walks 76 row pairs (twice a streamed tile's) at 80 VGPRs: it
// pays only when there are waves enough to hide that, i.e. batches of about
// seventeen 4K frames or more; smaller launches run level by level.
constexpr int kFusedMinTiles = 16384;
// The inverse counterpart (k_dwt_inverse_fused): a FusedJob's a is level l
// (writes the component's plane), b level l + 1 (reads its four bands); a
// tile is kInvFusedTileW plane samples wide and kInvFusedTallH high in
// launches of at least kInvFusedTallTiles such tiles, kInvFusedShortH high
// otherwise (measured, 4K 4:2:0 iDWT per frame: batches of 32, 15.1 against
// 15.5 us; one frame, 39.4 against 41.7 us), kInvFusedLoneH high in launches
// of fewer than kInvFusedLoneTiles short tiles, where twice the waves with
// half a short tile's walk each outweigh the longer apron (lone frames,
// encode + decode: 1080p FP16 101.4 -> 98.9 us, 4K FP32 173.5 -> 172.5 us).
// With these tiles every launch size is faster fused than level by level
// (lone 256x144 to 1280x720 frames 0.7 to 2.4 us), so the inverse always
// fuses.
constexpr int kInvFusedTileW = 120;
constexpr int kInvFusedTallH = 64;
constexpr int kInvFusedShortH = 32;
constexpr int kInvFusedLoneH = 16;
constexpr int kInvFusedTallTiles = 16384;
constexpr int kInvFusedLoneTiles = 4096;

// One transmitted band of one frame, for the per-32x32-block kernels.
struct BandJob
{
	void *data;         // padded band storage: pitch >= bx32 * 32, rows >= by32 * 32
	int32_t kind;       // kStoreF32 / kStoreF16
	int32_t pitch;      // elements
	int32_t bw, bh;
	int32_t bx8, by8;
	int32_t bx32, by32;
	int32_t off32;      // first frame-local 32x32 block index
	int32_t frame;      // frame slot
	float qres;         // 1 / decode_quant(encode_quant(1 / resolution))
	float rdo_scale;
	uint32_t qcode;     // encode_quant(1 / resolution)
	int32_t pad;
};

// Workgroups are dispatched round-robin over the XCDs (workgroup i runs on
// XCD i % nxcd), and each XCD has its own L2.  xcd_swizzle maps the dispatch
// order to a contiguous range of work items per XCD, so that neighbouring
// tiles and blocks, which share cache lines and filter aprons, meet in one L2:
// a bijection on [0, n) for any n.  The quotient and remainder of n / nxcd
// come with the launch (XcdMap, made on the host for the launch's grid), and
// a power-of-two XCD count divides by shifting, so a workgroup spends a few
// scalar instructions on it instead of two integer divisions.
struct XcdMap
{
	int nxcd;  // 1: identity (one XCD, or fewer workgroups than XCDs)
	int shift; // log2(nxcd) when nxcd is a power of two, else -1
	int q, r;  // n / nxcd, n % nxcd
};

__host__ __device__ inline int xcd_swizzle(int b, const XcdMap &m)
{
	if (m.nxcd <= 1)
		return b;
	int x, i;
	if (m.shift >= 0)
	{
		x = b & (m.nxcd - 1);
		i = b >> m.shift;
	}
	else
	{
		x = b % m.nxcd;
		i = b / m.nxcd;
	}
	return x < m.r ? x * (m.q + 1) + i : m.r * (m.q + 1) + (x - m.r) * m.q + i;
}

__host__ __device__ inline XcdMap make_xcd_map(int n, int nxcd)
{
	XcdMap m = {1, 0, n, 0};
	if (nxcd <= 1 || n < nxcd)
		return m;
	m.nxcd = nxcd;
	m.shift = -1;
	for (int k = 0; k < 31; k++)
		if ((1 << k) == nxcd)
			m.shift = k;
	m.q = n / nxcd;
	m.r = n % nxcd;
	return m;
}

// Number of XCDs of the current device (1 where the runtime cannot tell).
int xcd_count();

// The XcdMap of a launch of `grid` workgroups on the current device.
XcdMap xcd_map(const dim3 &grid);

// One frame of a device packetize launch (reference
// pyrowave_encoder.cpp:596-623, 1048-1115): the sequence header goes in the
// two words before the coded blocks, so header + blocks form one contiguous
// byte stream; the packet table splits it at the boundary.
struct PacketizeJob
{
	const uint32_t *meta; // [n32][2] {offset words, words}, block-index order
	uint32_t *stream;     // header (2 words), then the coded blocks
	uint32_t *table;      // [n32 + 1][2] {offset, size} in bytes
	uint32_t *info;       // [0] stream bytes, [1] packets
	uint32_t n32;
	uint32_t h0;          // header word 0 (size, sequence, header bit)
	uint32_t h1;          // header word 1 without the coded-block count
	uint32_t boundary;    // packet boundary in bytes
};
constexpr int kPacketizeJobs = 16;
struct PacketizeBlock
{
	PacketizeJob j[kPacketizeJobs];
};

// One frame of a device parse launch: the packets of one frame, in push order,
// parsed into the block offset table the dequant kernel reads (the state
// machine of reference pyrowave_decoder.cpp:126-262 for packets that all carry
// one sequence number).
struct ParseJob
{
	const uint32_t *data;    // packets back to back (4-byte aligned)
	const uint32_t *packets; // [count][2] {offset, size} in bytes
	const uint32_t *count;   // device word: number of packets
	uint32_t *offsets;       // [n32] word offset of each block in data (UINT32_MAX: missing)
	uint32_t *info;          // [0] blocks present, [1] blocks announced, [2] error bits, [3] sequence
	uint32_t data_bytes;
	uint32_t max_packets;
	uint32_t n32;
	uint32_t h0_size;        // expected header word 0, bits 0..27 (width - 1, height - 1)
	uint32_t chroma;         // 1 for 4:4:4
};
constexpr int kParseJobs = 16;
struct ParseBlock
{
	ParseJob j[kParseJobs];
};

// Error bits of ParseJob::info[2].
enum ParseError : uint32_t
{
	kParseBadTable = 1u << 0,    // packet outside data, misaligned, or too many packets
	kParseBadHeader = 1u << 1,   // sequence header for another size or chroma, reserved bits set
	kParseBadBlock = 1u << 2,    // block index out of range, block shorter than 8 bytes or past its packet
	kParseTrailing = 1u << 3,    // bytes left in a packet that hold no whole unit
	kParseSequence = 1u << 4     // units of more than one sequence number
};

// PSNR-HVS-M-H: height factors 1 + i / 8, i < kPsnrFactors (reference
// psnr.cpp:247, 321-324); tables per factor: csf, mask, inv_mask (64 each).
constexpr int kPsnrFactors = 16;

// Launch wrappers (defined in the .hip files).
void launch_dwt_forward(const DwtJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                        int variant, hipStream_t stream);
void launch_dwt_inverse(const DwtJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                        int variant, hipStream_t stream);
void launch_dwt_forward_fused(const FusedJob *jobs, const uint16_t *tile_job, int tiles, const FrameState *frames,
                              hipStream_t stream);
void launch_dwt_inverse_fused(const FusedJob *jobs, const uint16_t *tile_job, int tiles, int tile_h,
                              const FrameState *frames, hipStream_t stream);
// blk_band: two words per block (band job index, bx | by << 16).
void launch_quant_analyze(const BandJob *bands, const uint32_t *blk_band, int nblocks, const FrameState *frames,
                          hipStream_t stream);
void launch_rc_finalize(FrameState *frames, int nframes, hipStream_t stream);
// Finalize, partial resolve and truncation of each frame in one workgroup:
// launches of at most kResolveMaxFrames frames of at most kResolveMaxBlocks
// blocks (1080p: 3328), whose block records stay in registers.  Measured, it
// saves 3 us of a lone 1080p frame's encode; a lone 4K frame (12,400 blocks)
// resolves faster over the three launches above, and in batches the 1024-
// thread workgroups wait for whole CUs behind other streams' kernels
// (batched 4K: 87.5 against 83.8 us per frame on two streams).
constexpr int kResolveMaxBlocks = 4096;
constexpr int kResolveMaxFrames = 2;
void launch_rc_resolve(FrameState *frames, int nframes, hipStream_t stream);
void launch_rc_partial(FrameState *frames, int nframes, hipStream_t stream);
void launch_rc_prepare(FrameState *frames, const uint32_t *grp_frame, int ngroups_total,
                       const int32_t *frame_group_base, hipStream_t stream);
void launch_pack(const BandJob *bands, const uint32_t *blk_band, int nblocks, FrameState *frames,
                 hipStream_t stream);
// With nparams in 1 .. kParamsPerLaunch, also writes those frames' per-call
// fields (params[0 .. nparams - 1]; no k_set_params launch before it).
void launch_dequant(const BandJob *bands, const uint32_t *blk_band, int nblocks, FrameState *frames,
                    hipStream_t stream, const CallParams *params = nullptr, int nparams = 0);
// Writes the per-call fields and, for encoder frames, clears the rate-control
// accumulators (max_n32: the largest frame's 32x32 block count, 0 to skip).
void launch_set_params(FrameState *frames, const CallParams *params, int nframes, int max_n32, hipStream_t stream);
// The per-call parameter kernel (identifies its nodes in a captured graph).
const void *set_params_kernel();
void launch_packetize(const PacketizeJob *jobs, int count, hipStream_t stream);
void launch_parse(const ParseJob *jobs, int count, hipStream_t stream);
// The allocation-free decode's parse scratch: kNoAllocMaxBlocks offsets and
// four info words per GPU.  noalloc_scratch_init (with hip_device current)
// looks up its device address once; noalloc_scratch returns it (null before).
constexpr uint32_t kNoAllocMaxBlocks = 1u << 17;
constexpr int kMaxDevices = 64;
uint32_t *noalloc_scratch_init(int hip_device);
uint32_t *noalloc_scratch(int hip_device);
// Adds each height factor's fixed-point block errors to totals[kPsnrFactors].
void launch_psnr_hvs_m(const uint8_t *a, uint32_t pitch_a, const uint8_t *b, uint32_t pitch_b, int w, int h,
                       const float *tables, uint64_t *totals, hipStream_t stream);
void upload_constants();
} // namespace pw
