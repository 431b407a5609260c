// pw_host.cpp -- geometry, device memory, launch plans and the packet parser.
#include "pw_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>

namespace pw
{
namespace
{
int align_int(int v, int a) { return (v + a - 1) & ~(a - 1); }
size_t align_size(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

int next_pow2(int v)
{
	int p = 1;
	while (p < v)
		p <<= 1;
	return p;
}

int floor_log2(int v)
{
	int r = 0;
	while ((1 << (r + 1)) <= v)
		r++;
	return r;
}
} // namespace

// ---------------------------------------------------------------------------
// Quantizer constants (reference pyrowave_encoder.cpp:107-172,
// pyrowave_common.hpp:124-145).
// ---------------------------------------------------------------------------

static float noise_power_normalized_quant_resolution(int level, int component, int band, bool precision0)
{
	// 8 base bits for precision >= 1, 6 for precision 0.
	int bits = precision0 ? 6 : 8;
	if (band == 0)
		bits += 2;
	else if (band < 3)
		bits += 1;
	bits += level;
	if (component != 0)
		bits--;
	return float(1 << bits);
}

float quant_resolution(int level, int component, int band, bool precision0)
{
	// "FP16 range is limited": cap 512 for precision 0, 4096 otherwise.
	return std::min<float>(precision0 ? 512.0f : 4096.0f,
	                       noise_power_normalized_quant_resolution(level, component, band, precision0));
}

float rdo_distortion_scale(int chroma, int level, int component, int band, bool precision0)
{
	float horiz_midpoint = (band & 1) ? 0.75f : 0.25f;
	float vert_midpoint = (band & 2) ? 0.75f : 0.25f;
	constexpr float dpi = 96.0f;
	constexpr float viewing_distance = 1.0f;
	constexpr float cpd_nyquist = 0.34f * viewing_distance * dpi;
	float cpd = std::sqrt(horiz_midpoint * horiz_midpoint + vert_midpoint * vert_midpoint) * cpd_nyquist *
	            std::exp2(-float(level));
	cpd = std::max(cpd, 8.0f);
	float csf = 2.6f * (0.0192f + 0.114f * cpd) * std::exp(-std::pow(0.114f * cpd, 1.1f));
	if (component != 0 && level != kLevels - 1 && chroma == PYROWAVE_CHROMA_SUBSAMPLING_420)
		csf *= 0.6f;
	float resolution = noise_power_normalized_quant_resolution(level, component, band, precision0);
	float weighted = csf * resolution;
	return (weighted * weighted) * (1.0f / 256.0f);
}

uint8_t encode_quant(float decoder_q_scale)
{
	uint32_t v;
	memcpy(&v, &decoder_q_scale, sizeof(v));
	int e = int((v >> 23) & 0xff) - 127 - 4;
	int m = int((v >> 20) & 0x7);
	e = -e;
	return uint8_t((e << 3) | m);
}

float decode_quant(uint8_t code)
{
	int e = 4 - (code >> 3);
	int m = code & 0x7;
	return (1.0f / (8.0f * 1024.0f * 1024.0f)) * float((8 + m) * (1 << (20 + e)));
}

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------

bool Geometry::init(int w, int h, int chroma_, int precision_)
{
	if (w <= 0 || h <= 0 || w > 16384 || h > 16384)
		return false;
	*this = Geometry();
	width = w;
	height = h;
	chroma = chroma_;
	precision = precision_;
	aligned_w = std::max(align_int(w, kAlignment), kMinimumImageSize);
	aligned_h = std::max(align_int(h, kAlignment), kMinimumImageSize);

	long flat = 0;
	size_t store = 0;
	for (int c = 0; c < kComponents; c++)
		for (int l = 0; l < kLevels; l++)
		{
			if (l == 0 && c != 0 && chroma == PYROWAVE_CHROMA_SUBSAMPLING_420)
				continue;
			for (int b = 0; b < kBands; b++)
			{
				BandGeom &B = band[c][l][b];
				B.exists = true;
				B.bw = (aligned_w >> 1) >> l;
				B.bh = (aligned_h >> 1) >> l;
				B.bx8 = (B.bw + 7) / 8;
				B.by8 = (B.bh + 7) / 8;
				B.bx32 = (B.bw + 31) / 32;
				B.by32 = (B.bh + 31) / 32;
				B.flat_off = flat;
				flat += long(B.bw) * B.bh;
				// FP16: the transmitted high bands of levels 0-1 are stored as
				// fp16 and the LL intermediates stay FP32 (fp16 LL storage costs
				// 0.11 dB PSNR at high rates, the high bands ~0); FP16_LL: all
				// four bands of levels 0-1 as fp16, as the reference stores them.
				// FP16_MATH: every band of every level as fp16.
				const bool f16 = (l < 2 && ((precision == PYROWAVE_HIP_PRECISION_FP16 && b != 0) ||
				                            precision == PYROWAVE_HIP_PRECISION_FP16_LL)) ||
				                 precision == PYROWAVE_HIP_PRECISION_FP16_MATH;
				B.kind = f16 ? kStoreF16 : kStoreF32;
				B.pitch = B.bx32 * 32;
				size_t bytes = size_t(B.pitch) * size_t(B.by32 * 32) * (B.kind == kStoreF16 ? 2 : 4);
				B.store_off = store;
				store += align_size(bytes, 256);
			}
		}
	flat_count = flat;
	store_bytes = store;

	for (int l = kLevels - 1; l >= 0; l--)
		for (int c = 0; c < kComponents; c++)
		{
			if (l == 0 && c != 0 && chroma == PYROWAVE_CHROMA_SUBSAMPLING_420)
				continue;
			for (int b = (l == kLevels - 1 ? 0 : 1); b < kBands; b++)
			{
				BandGeom &B = band[c][l][b];
				B.present = true;
				B.off8 = count8;
				B.off32 = count32;
				count8 += B.bx8 * B.by8;
				count32 += B.bx32 * B.by32;
			}
		}
	per_sub = next_pow2(align_int(count32, kSubdiv) / kSubdiv);
	shamt = floor_log2(per_sub);
	ngroups = (count32 + 255) / 256;
	return true;
}

uint32_t Geometry::capacity_words() const
{
	return uint32_t(count32) * uint32_t(kMaxBlockWords);
}

// ---------------------------------------------------------------------------
// Buffers
// ---------------------------------------------------------------------------

DeviceBuffer::~DeviceBuffer() { release(); }

bool DeviceBuffer::alloc(size_t bytes)
{
	release();
	if (!bytes)
		return true;
	if (hipMalloc(&ptr, bytes) != hipSuccess)
	{
		ptr = nullptr;
		return false;
	}
	size = bytes;
	return true;
}

void DeviceBuffer::release()
{
	if (ptr)
		(void)hipFree(ptr);
	ptr = nullptr;
	size = 0;
}

HostBuffer::~HostBuffer() { release(); }

bool HostBuffer::alloc(size_t bytes)
{
	release();
	if (!bytes)
		return true;
	if (hipHostMalloc(&ptr, bytes, hipHostMallocDefault) != hipSuccess)
	{
		ptr = nullptr;
		return false;
	}
	size = bytes;
	return true;
}

void HostBuffer::release()
{
	if (ptr)
		(void)hipHostFree(ptr);
	ptr = nullptr;
	size = 0;
}

// ---------------------------------------------------------------------------
// Device: stage timing with HIP events
// ---------------------------------------------------------------------------

hipEvent_t Device::get_event()
{
	if (!event_pool.empty())
	{
		hipEvent_t e = event_pool.back();
		event_pool.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	(void)hipEventCreate(&e);
	return e;
}

void Device::begin_stage(hipStream_t s, int stage, int frames, hipEvent_t *start)
{
	(void)frames;
	*start = nullptr;
	if (!((timing >> stage) & 1u))
		return;
	*start = get_event();
	(void)hipEventRecord(*start, s);
}

void Device::end_stage(hipStream_t s, int stage, int frames, hipEvent_t start)
{
	if (!start)
		return;
	hipEvent_t end = get_event();
	(void)hipEventRecord(end, s);
	pending.push_back({start, end, stage, frames});
	if (pending.size() > 4096)
		resolve_timings(true);
}

void Device::resolve_timings(bool wait)
{
	size_t keep = 0;
	for (size_t i = 0; i < pending.size(); i++)
	{
		TimingRecord &r = pending[i];
		if (!wait && hipEventQuery(r.end) != hipSuccess)
		{
			pending[keep++] = r;
			continue;
		}
		(void)hipEventSynchronize(r.end);
		float ms = 0.0f;
		(void)hipEventElapsedTime(&ms, r.start, r.end);
		stage_ms[r.stage] += ms;
		stage_frames[r.stage] += uint64_t(r.frames);
		event_pool.push_back(r.start);
		event_pool.push_back(r.end);
	}
	pending.resize(keep);
}

Device::~Device()
{
	resolve_timings(true);
	for (hipEvent_t e : event_pool)
		(void)hipEventDestroy(e);
	if (stream)
		(void)hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// Slot
// ---------------------------------------------------------------------------

// Device memory of one slot: band storage, then (encoder only) the
// rate-control arrays and the bitstream, each 256-byte aligned.
SlotLayout slot_layout(const Geometry &G, bool encoder)
{
	SlotLayout L = {};
	const size_t n32 = size_t(G.count32);
	auto take = [&](size_t bytes) {
		size_t off = L.total;
		L.total += align_size(bytes, 256);
		return off;
	};
	L.bands = take(G.store_bytes);
	if (encoder)
	{
		L.words = take(n32 * 16 * sizeof(uint16_t));
		// Padded to whole groups of eight (zeros past the frame: the slot is
		// cleared once and the quantizer writes only real blocks).
		L.words0 = take(((n32 + 7) & ~size_t(7)) * sizeof(uint16_t));
		L.bucket = take(n32 * 16);
		L.partial_quant = take(n32 * sizeof(int32_t));
		L.quant = take(n32);
		L.excl = take(n32 * sizeof(uint32_t));
		L.group_sum = take(size_t(G.ngroups) * sizeof(uint32_t));
		L.meta = take(n32 * 2 * sizeof(uint32_t));
		L.totals = take(size_t(totals_copies(n32)) * kFlat * sizeof(uint32_t));
		L.prefix = take(kFlat * sizeof(uint32_t));
		L.cls = take(kFlat);
		L.required = take(kFlat * sizeof(int32_t));
		L.rc = take(kRcCount * sizeof(uint32_t));
		// Two words in front of the blocks hold the sequence header of a
		// device packetize.
		L.bitstream = take((size_t(G.capacity_words()) + 2) * sizeof(uint32_t));
		L.packets = take((2 * (n32 + 1) + 2) * sizeof(uint32_t));
	}
	return L;
}

bool Slot::init(Device *dev_, int w, int h, int chroma, int precision, bool encoder_)
{
	dev = dev_;
	encoder = encoder_;
	if (!geom.init(w, h, chroma, precision))
		return false;
	const Geometry &G = geom;
	const SlotLayout L = slot_layout(G, encoder);
	if (!mem.alloc(L.total))
		return false;
	if (hipMemset(mem.ptr, 0, L.total) != hipSuccess)
		return false;
	uint8_t *base = static_cast<uint8_t *>(mem.ptr);
	bands = base + L.bands;

	memset(&fs, 0, sizeof(fs));
	if (encoder)
	{
		fs.words = reinterpret_cast<uint16_t *>(base + L.words);
		fs.words0 = reinterpret_cast<uint16_t *>(base + L.words0);
		fs.bucket = base + L.bucket;
		fs.partial_quant = reinterpret_cast<int32_t *>(base + L.partial_quant);
		fs.quant = base + L.quant;
		fs.excl = reinterpret_cast<uint32_t *>(base + L.excl);
		fs.group_sum = reinterpret_cast<uint32_t *>(base + L.group_sum);
		fs.meta = reinterpret_cast<uint32_t *>(base + L.meta);
		fs.totals = reinterpret_cast<uint32_t *>(base + L.totals);
		fs.prefix = reinterpret_cast<uint32_t *>(base + L.prefix);
		fs.cls = base + L.cls;
		fs.required = reinterpret_cast<int32_t *>(base + L.required);
		fs.rc = reinterpret_cast<uint32_t *>(base + L.rc);
		fs.bitstream = reinterpret_cast<uint32_t *>(base + L.bitstream) + 2;
		dtable = reinterpret_cast<uint32_t *>(base + L.packets);
		dpinfo = dtable + 2 * (size_t(G.count32) + 1);
		fs.capacity_words = G.capacity_words();
	}
	fs.n32 = G.count32;
	fs.per_sub = G.per_sub;
	fs.shamt = G.shamt;
	fs.ngroups = G.ngroups;
	fs.totals_copies = totals_copies(G.count32);
	fs.width = G.width;
	fs.height = G.height;
	fs.flags = exact_prefix ? 1u : 0u;

	if (hipEventCreateWithFlags(&staging_done, hipEventDisableTiming) != hipSuccess)
		return false;
	if (hipEventCreateWithFlags(&result_ready, hipEventDisableTiming) != hipSuccess)
		return false;

	offsets.assign(size_t(G.count32), UINT32_MAX);
	clear_packets();
	return true;
}

Slot::~Slot()
{
	if (dev && dev->stream)
		(void)hipStreamSynchronize(dev->stream);
	single_plan.reset();
	if (staging_done)
		(void)hipEventDestroy(staging_done);
	if (result_ready)
		(void)hipEventDestroy(result_ready);
}

// Packet parser: reference pyrowave_decoder.cpp:126-262, 780-794, 832-840.
void Slot::clear_packets()
{
	std::fill(offsets.begin(), offsets.end(), UINT32_MAX);
	decoded_blocks = 0;
	last_seq = UINT32_MAX;
	decoded_frame = false;
	total_blocks = geom.count32;
	payload.clear();
}

bool Slot::push_packet(const void *data_, size_t size)
{
	const uint8_t *data = static_cast<const uint8_t *>(data_);
	while (size >= 8)
	{
		uint32_t w[2];
		memcpy(w, data, 8);
		const uint32_t seq = (w[0] >> 28) & 7u;
		if (w[0] >> 31)
		{
			// Sequence header (bitstream/bitstream.md:204-315).
			if (int((w[1] >> 26) & 1u) != geom.chroma)
				return false;
			uint32_t diff = (seq - last_seq) & 7u;
			if (last_seq != UINT32_MAX && diff > 3)
				return true;
			if (last_seq == UINT32_MAX || diff != 0)
			{
				clear_packets();
				last_seq = seq;
			}
			if (((w[1] >> 24) & 3u) != 0)
				return false;
			int wd = int(w[0] & 0x3fffu) + 1, ht = int((w[0] >> 14) & 0x3fffu) + 1;
			if (wd != geom.width || ht != geom.height)
				return false;
			total_blocks = int(w[1] & 0xffffffu);
			data += 8;
			size -= 8;
			continue;
		}

		size_t packet_size = size_t((w[0] >> 16) & 0xfffu) * 4;
		if (packet_size > size)
			return false;
		bool restart;
		if (last_seq == UINT32_MAX)
			restart = true;
		else
		{
			uint32_t diff = (seq - last_seq) & 7u;
			if (diff > 3)
				return true;
			restart = diff != 0;
		}
		if (restart)
		{
			clear_packets();
			last_seq = seq;
		}
		uint32_t block = w[1] >> 8;
		if (block >= uint32_t(geom.count32))
			return false;
		if (offsets[block] == UINT32_MAX)
		{
			// Counted and placed before the length check
			// (pyrowave_decoder.cpp:126-150).
			decoded_blocks++;
			offsets[block] = uint32_t(payload.size());
			if (packet_size < 8)
				return false;
			size_t n = payload.size();
			payload.resize(n + packet_size / 4);
			memcpy(payload.data() + n, data, packet_size);
		}
		else if (packet_size == 0)
		{
			// The reference's loop would never advance past a zero-length
			// duplicate (pyrowave_decoder.cpp:236-246); a parse error here.
			return false;
		}
		data += packet_size;
		size -= packet_size;
	}
	return size == 0;
}

bool Slot::decode_is_ready(bool allow_partial) const
{
	if (decoded_frame)
		return false;
	if (last_seq == UINT32_MAX)
		return false;
	if (decoded_blocks < total_blocks)
		if (!allow_partial || decoded_blocks <= total_blocks / 2)
			return false;
	return true;
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------

static bool fusable(const Geometry &G, int c)
{
	// The fused launches read and write 8-bit planes only.
	return G.plane_w(c) >= 13 && G.plane_h(c) >= 13 && G.sample_bytes == 1;
}

bool Plan::build(const std::vector<Slot *> &slots_, bool encoder_, PlanStorage *storage, hipStream_t stream)
{
	slots = slots_;
	encoder = encoder_;
	const int n = int(slots.size());
	std::vector<FrameState> host_frames(n);
	params.assign(n, CallParams{});
	std::vector<DwtJob> jobs[kLevels][kDwtVariants];
	std::vector<uint16_t> tile_map[kLevels][kDwtVariants];
	std::vector<DwtJob> rjobs[kLevels][kDwtVariants];
	std::vector<uint16_t> rtile_map[kLevels][kDwtVariants];
	std::vector<FusedJob> fjobs;
	std::vector<uint16_t> ftile_map;
	std::vector<BandJob> band_jobs;
	std::vector<uint32_t> blk_band;
	std::vector<int32_t> frame_block_base(n), frame_group_base(n);
	std::vector<uint32_t> grp_frame;
	max_n32 = 0;
	// FP16-arithmetic frames have their own transform kernels (one shape);
	// a batch holds either those or the others.
	half_math = slots[0]->geom.precision == PYROWAVE_HIP_PRECISION_FP16_MATH;
	for (Slot *S : slots)
		if ((S->geom.precision == PYROWAVE_HIP_PRECISION_FP16_MATH) != half_math)
			return false;
	// Both directions fuse a component's first two levels (FP32-arithmetic
	// frames with 8-bit planes).
	has_fused = !half_math;
	// Inverse tile height: tall tiles where the launch has waves enough.
	inv_fused_tile_h = kInvFusedShortH;
	inv_fused_short_tiles = 0;
	if (has_fused && !encoder)
	{
		long tall = 0, short_tiles = 0;
		for (Slot *S : slots)
			for (int c = 0; c < kComponents; c++)
				if (fusable(S->geom, c))
				{
					const long tx = (S->geom.plane_w(c) + kInvFusedTileW - 1) / kInvFusedTileW;
					tall += tx * ((S->geom.plane_h(c) + kInvFusedTallH - 1) / kInvFusedTallH);
					short_tiles += tx * ((S->geom.plane_h(c) + kInvFusedShortH - 1) / kInvFusedShortH);
				}
		inv_fused_short_tiles = int(std::min(short_tiles, long(INT32_MAX)));
		if (tall >= kInvFusedTallTiles)
			inv_fused_tile_h = kInvFusedTallH;
		else if (short_tiles < kInvFusedLoneTiles)
			inv_fused_tile_h = kInvFusedLoneH;
	}

	for (int f = 0; f < n; f++)
	{
		Slot *S = slots[f];
		const Geometry &G = S->geom;
		host_frames[f] = S->fs;
		max_n32 = std::max(max_n32, G.count32);
		auto band_ptr = [&](int c, int l, int b) { return static_cast<void *>(S->bands + G.band[c][l][b].store_off); };

		DwtJob first_two[kComponents][2] = {};
		for (int l = 0; l < kLevels; l++)
		{
			for (int c = 0; c < kComponents; c++)
			{
				const int first = G.first_level(c);
				if (l < first)
					continue;
				// The fused forward launch takes the LL values it computes above
				// and left of the frame as the mirror images of those inside,
				// which holds when the plane's extension mirrors the 12 rows and
				// columns its window reaches past the edge: planes of at least 13
				// samples.  (The fused inverse needs no such bound: its band
				// dimensions are even and at least 16.)
				const bool in_fused = has_fused && l <= first + 1 && fusable(G, c);
				const BandGeom &B0 = G.band[c][l][0];
				DwtJob J = {};
				J.frame = f;
				J.component = c;
				J.ll_kind = B0.kind;
				J.band_kind = G.band[c][l][1].kind;
				J.band_pitch = B0.pitch;
				J.bw = B0.bw;
				J.bh = B0.bh;
				for (int b = 0; b < kBands; b++)
					J.band[b] = band_ptr(c, l, b);
				if (encoder)
				{
					J.A_x = G.aligned_w >> l;
					J.A_y = G.aligned_h >> l;
					if (l == first)
					{
						J.src = nullptr;
						J.src_kind = G.sample_bytes == 2 ? kStoreU16 : kStoreU8;
						J.R_x = G.plane_w(c);
						J.R_y = G.plane_h(c);
					}
					else
					{
						const BandGeom &P = G.band[c][l - 1][0];
						J.src = band_ptr(c, l - 1, 0);
						J.src_kind = P.kind;
						J.src_pitch = P.pitch;
						J.R_x = J.A_x;
						J.R_y = J.A_y;
					}
				}
				else
				{
					J.A_x = 2 * B0.bw;
					J.A_y = 2 * B0.bh;
					if (l == first)
					{
						J.dst = nullptr;
						J.dst_kind = G.sample_bytes == 2 ? kStoreU16 : kStoreU8;
						J.R_x = G.plane_w(c);
						J.R_y = G.plane_h(c);
					}
					else
					{
						const BandGeom &P = G.band[c][l - 1][0];
						J.dst = band_ptr(c, l - 1, 0);
						J.dst_kind = P.kind;
						J.dst_pitch = P.pitch;
						J.R_x = J.A_x;
						J.R_y = J.A_y;
					}
				}
				if (in_fused)
					first_two[c][l - first] = J;
				for (int v = 0; v < kDwtVariants; v++)
				{
					if ((v == kHalfStreamVariant) != half_math)
						continue;
					const int tw = encoder ? kFwdVariantTileW[v] : kInvVariantTileW[v];
					const int th = encoder ? kFwdVariantTileH[v] : kInvVariantTileH[v];
					DwtJob Jv = J;
					Jv.tiles_x = (J.A_x + tw - 1) / tw;
					const int tiles_y = (J.A_y + th - 1) / th;
					Jv.tile_base = int(tile_map[l][v].size());
					const uint16_t idx = uint16_t(jobs[l][v].size());
					jobs[l][v].push_back(Jv);
					tile_map[l][v].insert(tile_map[l][v].end(), size_t(Jv.tiles_x) * tiles_y, idx);
					if (!in_fused)
					{
						Jv.tile_base = int(rtile_map[l][v].size());
						const uint16_t ridx = uint16_t(rjobs[l][v].size());
						rjobs[l][v].push_back(Jv);
						rtile_map[l][v].insert(rtile_map[l][v].end(), size_t(Jv.tiles_x) * tiles_y, ridx);
					}
				}
			}
		}
		if (has_fused)
			for (int c = 0; c < kComponents; c++)
			{
				if (!fusable(G, c))
					continue;
				FusedJob FJ = {};
				FJ.a = first_two[c][0];
				FJ.b = first_two[c][1];
				// Forward tiles cover level l + 1's bands; inverse tiles the
				// plane (only the plane is written).
				FJ.tiles_x = encoder ? (FJ.b.bw + kFusedTileW1 - 1) / kFusedTileW1
				                     : (FJ.a.R_x + kInvFusedTileW - 1) / kInvFusedTileW;
				const int tiles_y = encoder ? (FJ.b.bh + kFusedTileH1 - 1) / kFusedTileH1
				                            : (FJ.a.R_y + inv_fused_tile_h - 1) / inv_fused_tile_h;
				FJ.tile_base = int(ftile_map.size());
				const uint16_t idx = uint16_t(fjobs.size());
				fjobs.push_back(FJ);
				ftile_map.insert(ftile_map.end(), size_t(FJ.tiles_x) * tiles_y, idx);
			}

		frame_block_base[f] = int32_t(blk_band.size() / 2);
		for (int l = kLevels - 1; l >= 0; l--)
			for (int c = 0; c < kComponents; c++)
				for (int b = 0; b < kBands; b++)
				{
					const BandGeom &B = G.band[c][l][b];
					if (!B.present)
						continue;
					BandJob J = {};
					J.data = band_ptr(c, l, b);
					J.kind = B.kind;
					J.pitch = B.pitch;
					J.bw = B.bw;
					J.bh = B.bh;
					J.bx8 = B.bx8;
					J.by8 = B.by8;
					J.bx32 = B.bx32;
					J.by32 = B.by32;
					J.off32 = B.off32;
					J.frame = f;
					const bool p0 = G.precision == PYROWAVE_HIP_PRECISION_FP16_MATH;
					float res = quant_resolution(l, c, b, p0);
					J.qres = 1.0f / decode_quant(encode_quant(1.0f / res));
					J.rdo_scale = rdo_distortion_scale(G.chroma, l, c, b, p0);
					J.qcode = encode_quant(1.0f / res);
					uint32_t idx = uint32_t(band_jobs.size());
					band_jobs.push_back(J);
					// Two words per block: band job, position bx | by << 16.
					for (int by = 0; by < B.by32; by++)
						for (int bx = 0; bx < B.bx32; bx++)
						{
							blk_band.push_back(idx);
							blk_band.push_back(uint32_t(bx) | (uint32_t(by) << 16));
						}
				}
		frame_group_base[f] = int32_t(grp_frame.size());
		grp_frame.insert(grp_frame.end(), size_t(G.ngroups), uint32_t(f));
	}
	nblocks = int(blk_band.size() / 2);
	ngroups = int(grp_frame.size());
	Device &D = *slots[0]->dev;
	fused_tiles = int(ftile_map.size());
	if (fjobs.size() > 65535)
		return false;
	for (int l = 0; l < kLevels; l++)
	{
		for (int v = 0; v < kDwtVariants; v++)
		{
			tiles[l][v] = int(tile_map[l][v].size());
			rest_tiles[l][v] = int(rtile_map[l][v].size());
			if (jobs[l][v].size() > 65535)
				return false;
		}
		if (half_math)
		{
			variant[l] = kHalfStreamVariant;
			continue;
		}
		variant[l] = default_dwt_variant(encoder, tiles[l][kFwdStreamVariant]);
		auto it = D.tuned_variant.find(tune_key(encoder, l, tiles[l][0]));
		if (it != D.tuned_variant.end())
			variant[l] = it->second;
	}

	// One device allocation for every table.
	size_t total = 0;
	auto take = [&](size_t bytes) {
		size_t off = total;
		total += align_size(bytes, 256);
		return off;
	};
	size_t o_frames = take(sizeof(FrameState) * n);
	size_t o_jobs[kLevels][kDwtVariants], o_tiles[kLevels][kDwtVariants];
	size_t o_rjobs[kLevels][kDwtVariants], o_rtiles[kLevels][kDwtVariants];
	for (int l = 0; l < kLevels; l++)
		for (int v = 0; v < kDwtVariants; v++)
		{
			o_jobs[l][v] = take(sizeof(DwtJob) * jobs[l][v].size());
			o_tiles[l][v] = take(sizeof(uint16_t) * tile_map[l][v].size());
			o_rjobs[l][v] = take(sizeof(DwtJob) * rjobs[l][v].size());
			o_rtiles[l][v] = take(sizeof(uint16_t) * rtile_map[l][v].size());
		}
	size_t o_fjobs = take(sizeof(FusedJob) * fjobs.size());
	size_t o_ftiles = take(sizeof(uint16_t) * ftile_map.size());
	size_t o_bands = take(sizeof(BandJob) * band_jobs.size());
	size_t o_blk = take(sizeof(uint32_t) * blk_band.size());
	size_t o_fbb = take(sizeof(int32_t) * n);
	size_t o_grp = take(sizeof(uint32_t) * grp_frame.size());
	size_t o_fgb = take(sizeof(int32_t) * n);
	uint8_t *base = nullptr, *host = nullptr;
	std::vector<uint8_t> host_vec;
	if (storage)
	{
		// Grow by half again so a stream of slightly larger batches does not
		// reallocate every time.
		if (storage->dev.size < total && !storage->dev.alloc(total + total / 2))
			return false;
		if (storage->host.size < total && !storage->host.alloc(total + total / 2))
			return false;
		base = static_cast<uint8_t *>(storage->dev.ptr);
		host = static_cast<uint8_t *>(storage->host.ptr);
	}
	else
	{
		if (!mem.alloc(total))
			return false;
		base = static_cast<uint8_t *>(mem.ptr);
		host_vec.assign(total, 0);
		host = host_vec.data();
	}
	auto put = [&](size_t off, const void *src, size_t bytes) {
		if (bytes)
			memcpy(host + off, src, bytes);
	};
	for (int l = 0; l < kLevels; l++)
		for (int v = 0; v < kDwtVariants; v++)
		{
			put(o_jobs[l][v], jobs[l][v].data(), sizeof(DwtJob) * jobs[l][v].size());
			put(o_tiles[l][v], tile_map[l][v].data(), sizeof(uint16_t) * tile_map[l][v].size());
			d_dwt_jobs[l][v] = reinterpret_cast<DwtJob *>(base + o_jobs[l][v]);
			d_tile_job[l][v] = reinterpret_cast<uint16_t *>(base + o_tiles[l][v]);
			put(o_rjobs[l][v], rjobs[l][v].data(), sizeof(DwtJob) * rjobs[l][v].size());
			put(o_rtiles[l][v], rtile_map[l][v].data(), sizeof(uint16_t) * rtile_map[l][v].size());
			d_rest_jobs[l][v] = reinterpret_cast<DwtJob *>(base + o_rjobs[l][v]);
			d_rest_tile_job[l][v] = reinterpret_cast<uint16_t *>(base + o_rtiles[l][v]);
		}
	put(o_fjobs, fjobs.data(), sizeof(FusedJob) * fjobs.size());
	put(o_ftiles, ftile_map.data(), sizeof(uint16_t) * ftile_map.size());
	d_fused_jobs = reinterpret_cast<FusedJob *>(base + o_fjobs);
	d_fused_tile_job = reinterpret_cast<uint16_t *>(base + o_ftiles);
	put(o_bands, band_jobs.data(), sizeof(BandJob) * band_jobs.size());
	put(o_blk, blk_band.data(), sizeof(uint32_t) * blk_band.size());
	put(o_fbb, frame_block_base.data(), sizeof(int32_t) * n);
	put(o_grp, grp_frame.data(), sizeof(uint32_t) * grp_frame.size());
	put(o_fgb, frame_group_base.data(), sizeof(int32_t) * n);
	put(o_frames, host_frames.data(), sizeof(FrameState) * n);
	if (storage)
	{
		if (hipMemcpyAsync(base, host, total, hipMemcpyHostToDevice, stream) != hipSuccess)
			return false;
	}
	else if (hipMemcpy(base, host, total, hipMemcpyHostToDevice) != hipSuccess)
		return false;
	d_frames = reinterpret_cast<FrameState *>(base + o_frames);
	d_bands = reinterpret_cast<BandJob *>(base + o_bands);
	d_blk_band = reinterpret_cast<uint32_t *>(base + o_blk);
	d_frame_block_base = reinterpret_cast<int32_t *>(base + o_fbb);
	d_grp_frame = reinterpret_cast<uint32_t *>(base + o_grp);
	d_frame_group_base = reinterpret_cast<int32_t *>(base + o_fgb);

	return true;
}

FrameState *Plan::stage_frames(hipStream_t stream)
{
	launch_set_params(d_frames, params.data(), int(params.size()), encoder ? max_n32 : 0, stream);
	return d_frames;
}

// ---------------------------------------------------------------------------
// Launch sequences
// ---------------------------------------------------------------------------

uint64_t tune_key(bool encoder, int level, int tiles0)
{
	return (uint64_t(encoder) << 40) | (uint64_t(level) << 32) | uint32_t(tiles0);
}

int level_variant(const Plan &P, int l)
{
	// A forced shape this direction does not have (no tiles) falls back.
	const int f = P.slots[0]->dev->forced_variant;
	return f >= 0 && f < kDwtVariants && P.tiles[l][f] > 0 ? f : P.variant[l];
}

// A candidate shape replaces the level's default one (default_dwt_variant)
// only when it is this much faster: shapes closer than that are within launch-to-launch
// noise, and keeping the default there makes the choice the same from run to
// run.
constexpr float kTuneMargin = 0.02f;

// Rounds of the autotuner's measurement per level (each round times every
// variant once, in turn, so clock and cache drift fall on all of them alike).
constexpr int kTuneRounds = 3;

// One transform level, autotuned: in each round, every variant launches
// `reps` times back to back between two HIP events (a level's launch reads
// inputs it does not write, so repeating it is harmless); a variant's time is
// its best round divided by `reps`, which prices launches as the pipeline
// issues them rather than one isolated dispatch.  The fastest is kept if it
// beats the default by kTuneMargin.  The last launch leaves the level's
// output in place for the next level.
template <typename Launch>
void tune_level(Plan &P, bool encoder, int l, int reps, hipStream_t stream, Launch launch)
{
	Device &D = *P.slots[0]->dev;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
	{
		launch(level_variant(P, l));
		if (e0)
			(void)hipEventDestroy(e0);
		return;
	}
	float v_ms[kDwtVariants];
	for (int v = 0; v < kDwtVariants; v++)
		v_ms[v] = std::numeric_limits<float>::infinity();
	for (int r = 0; r < kTuneRounds; r++)
		for (int v = 0; v < kDwtVariants; v++)
		{
			if (P.tiles[l][v] == 0)
				continue;
			(void)hipEventRecord(e0, stream);
			for (int k = 0; k < reps; k++)
				launch(v);
			(void)hipEventRecord(e1, stream);
			float ms = 0.0f;
			if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
				v_ms[v] = std::min(v_ms[v], ms / float(reps));
		}
	int best = P.variant[l];
	float best_ms = std::numeric_limits<float>::infinity();
	for (int v = 0; v < kDwtVariants; v++)
		if (P.tiles[l][v] > 0 && v_ms[v] < best_ms)
		{
			best = v;
			best_ms = v_ms[v];
		}
	const int dflt = default_dwt_variant(encoder, P.tiles[l][kFwdStreamVariant]);
	const float default_ms = P.tiles[l][dflt] > 0 ? v_ms[dflt] : -1.0f;
	if (best != dflt && default_ms >= 0.0f && best_ms > default_ms * (1.0f - kTuneMargin))
		best = dflt;
	// The chosen variant runs last, so its output is the one the next level
	// reads (all variants write the same values).
	launch(best);
	if (P.variant[l] != best && P.graph_exec)
	{
		// A captured graph holds the old launches: capture again on next use.
		(void)hipGraphExecDestroy(P.graph_exec);
		(void)hipGraphDestroy(P.graph);
		P.graph_exec = nullptr;
		P.graph = nullptr;
	}
	P.variant[l] = best;
	D.tuned_variant[tune_key(encoder, l, P.tiles[l][0])] = best;
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
}

void run_encode(Plan &P, hipStream_t stream, bool timed)
{
	Device &D = *P.slots[0]->dev;
	const int n = int(P.slots.size());
	FrameState *frames = P.stage_frames(stream);
	hipEvent_t ev;
	const uint32_t saved_timing = D.timing;
	if (!timed)
		D.timing = 0;

	const int reps = timed ? D.tune_reps : 0;
	D.begin_stage(stream, kStageDWT, n, &ev);
	if (P.has_fused && D.forced_variant < 0 && reps == 0 &&
	    (D.fused >= 2 || (D.fused == 1 && P.fused_tiles >= kFusedMinTiles)))
	{
		// The first two levels of every component in one launch, then the
		// remaining levels with their tuned shapes.
		launch_dwt_forward_fused(P.d_fused_jobs, P.d_fused_tile_job, P.fused_tiles, frames, stream);
		for (int l = 0; l < kLevels; l++)
		{
			const int v = P.variant[l];
			launch_dwt_forward(P.d_rest_jobs[l][v], P.d_rest_tile_job[l][v], P.rest_tiles[l][v], frames, v, stream);
		}
	}
	else
		for (int l = 0; l < kLevels; l++)
		{
			auto launch = [&](int v) {
				launch_dwt_forward(P.d_dwt_jobs[l][v], P.d_tile_job[l][v], P.tiles[l][v], frames, v, stream);
			};
			if (reps > 0 && !P.half_math)
				tune_level(P, true, l, reps, stream, launch);
			else
				launch(level_variant(P, l));
		}
	D.end_stage(stream, kStageDWT, n, ev);

	D.begin_stage(stream, kStageQuant, n, &ev);
	launch_quant_analyze(P.d_bands, P.d_blk_band, P.nblocks, frames, stream);
	D.end_stage(stream, kStageQuant, n, ev);

	D.begin_stage(stream, kStageResolve, n, &ev);
	if (D.fused_resolve && n <= kResolveMaxFrames && P.max_n32 <= kResolveMaxBlocks)
		launch_rc_resolve(frames, n, stream);
	else
	{
		launch_rc_finalize(frames, n, stream);
		launch_rc_partial(frames, n, stream);
		launch_rc_prepare(frames, P.d_grp_frame, P.ngroups, P.d_frame_group_base, stream);
	}
	D.end_stage(stream, kStageResolve, n, ev);

	D.begin_stage(stream, kStagePacking, n, &ev);
	launch_pack(P.d_bands, P.d_blk_band, P.nblocks, frames, stream);
	D.end_stage(stream, kStagePacking, n, ev);
	D.timing = saved_timing;
}

void run_decode(Plan &P, hipStream_t stream, bool timed, bool capture)
{
	Device &D = *P.slots[0]->dev;
	const int n = int(P.slots.size());
	// Eager decodes of up to kParamsPerLaunch frames pass the call's fields
	// to the dequantizer (one launch fewer: about 3 us of a lone frame); a
	// captured graph keeps the k_set_params launch, whose node each replay
	// rewrites.
	const bool fold = !capture && P.nblocks > 0 && int(P.params.size()) <= kParamsPerLaunch;
	FrameState *frames = fold ? P.d_frames : P.stage_frames(stream);
	hipEvent_t ev;
	const uint32_t saved_timing = D.timing;
	if (!timed)
		D.timing = 0;

	D.begin_stage(stream, kStageDequant, n, &ev);
	launch_dequant(P.d_bands, P.d_blk_band, P.nblocks, frames, stream, fold ? P.params.data() : nullptr,
	               fold ? int(P.params.size()) : 0);
	D.end_stage(stream, kStageDequant, n, ev);

	const int reps = timed ? D.tune_reps : 0;
	D.begin_stage(stream, kStageIDWT, n, &ev);
	if (P.has_fused && D.forced_variant < 0 && reps == 0 &&
	    D.fused >= 1)
	{
		// The levels above each component's first two with their tuned
		// shapes, then those two in one launch.
		for (int l = kLevels - 1; l >= 0; l--)
		{
			const int v = P.variant[l];
			launch_dwt_inverse(P.d_rest_jobs[l][v], P.d_rest_tile_job[l][v], P.rest_tiles[l][v], frames, v, stream);
		}
		launch_dwt_inverse_fused(P.d_fused_jobs, P.d_fused_tile_job, P.fused_tiles, P.inv_fused_tile_h, frames,
		                         stream);
	}
	else
		for (int l = kLevels - 1; l >= 0; l--)
		{
			auto launch = [&](int v) {
				launch_dwt_inverse(P.d_dwt_jobs[l][v], P.d_tile_job[l][v], P.tiles[l][v], frames, v, stream);
			};
			if (reps > 0 && !P.half_math)
				tune_level(P, false, l, reps, stream, launch);
			else
				launch(level_variant(P, l));
		}
	D.end_stage(stream, kStageIDWT, n, ev);
	D.timing = saved_timing;
}

Plan::~Plan()
{
	if (graph_exec)
		(void)hipGraphExecDestroy(graph_exec);
	if (graph)
		(void)hipGraphDestroy(graph);
}

// Graph mode.  The first call captures the plan's launch sequence on the
// caller's stream (no stage timing events, no memset nodes: every buffer the
// sequence clears is cleared by a kernel) and keeps the k_set_params nodes;
// each call then rewrites those nodes' arguments with the call's parameters
// and launches the graph.
bool run_graph(Plan &P, hipStream_t stream)
{
	if (!P.graph_exec)
	{
		if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
			return false;
		if (P.encoder)
			run_encode(P, stream, false);
		else
			run_decode(P, stream, false, true);
		hipGraph_t g = nullptr;
		if (hipStreamEndCapture(stream, &g) != hipSuccess || !g)
			return false;
		hipGraphExec_t e = nullptr;
		if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess)
		{
			(void)hipGraphDestroy(g);
			return false;
		}
		size_t n = 0;
		(void)hipGraphGetNodes(g, nullptr, &n);
		std::vector<hipGraphNode_t> nodes(n);
		(void)hipGraphGetNodes(g, nodes.data(), &n);
		P.param_nodes.clear();
		P.param_node_base.clear();
		P.param_first.clear();
		P.param_count.clear();
		for (hipGraphNode_t node : nodes)
		{
			hipGraphNodeType t;
			if (hipGraphNodeGetType(node, &t) != hipSuccess || t != hipGraphNodeTypeKernel)
				continue;
			hipKernelNodeParams kp = {};
			if (hipGraphKernelNodeGetParams(node, &kp) != hipSuccess || kp.func != set_params_kernel())
				continue;
			if (!kp.kernelParams)
			{
				(void)hipGraphExecDestroy(e);
				(void)hipGraphDestroy(g);
				return false;
			}
			P.param_nodes.push_back(node);
			P.param_node_base.push_back(kp);
			P.param_first.push_back(*static_cast<int *>(kp.kernelParams[2]));
			P.param_count.push_back(*static_cast<int *>(kp.kernelParams[3]));
		}
		int covered = 0;
		for (int c : P.param_count)
			covered += c;
		if (covered != int(P.params.size()))
		{
			(void)hipGraphExecDestroy(e);
			(void)hipGraphDestroy(g);
			return false;
		}
		P.graph = g;
		P.graph_exec = e;
	}
	for (size_t k = 0; k < P.param_nodes.size(); k++)
	{
		ParamBlock block;
		const int first = P.param_first[k], count = P.param_count[k];
		for (int i = 0; i < count; i++)
			block.p[i] = P.params[first + i];
		FrameState *frames = P.d_frames;
		int a_first = first, a_count = count;
		void *args[4] = {&frames, &block, &a_first, &a_count};
		hipKernelNodeParams kp = P.param_node_base[k];
		kp.kernelParams = args;
		kp.extra = nullptr;
		if (hipGraphExecKernelNodeSetParams(P.graph_exec, P.param_nodes[k], &kp) != hipSuccess)
			return false;
	}
	return hipGraphLaunch(P.graph_exec, stream) == hipSuccess;
}
} // namespace pw
