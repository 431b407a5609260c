// pw_capi.cpp -- the pyrowave.h C ABI and the pyrowave_hip.h extension.
//
// Behaviour follows the reference C layer (reference pyrowave_c.cpp): the
// same validation rules and result codes, the same asynchronous encode /
// blocking packet queries, the same packet parser.  Vulkan-object entry points
// report PYROWAVE_ERROR_NOT_IMPLEMENTED.
#include "pw_host.hpp"

#include <cmath>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

using namespace pw;

namespace pw
{
int target_words_of(size_t target_bytes)
{
	// reference pyrowave_encoder.cpp:290-291, 326: the sequence header is
	// taken off the budget, the rest is counted in words.
	if (target_bytes >= 8)
		target_bytes -= 8;
	return int(target_bytes / 4);
}

// Copies the encoder's meta (and rc words) and the first `words` bitstream
// words to its pinned host mirrors on `stream`.
bool queue_result_download(Slot &S, hipStream_t stream, uint32_t words)
{
	const size_t meta_bytes = (size_t(S.geom.count32) * 2 + kRcCount) * sizeof(uint32_t);
	if (!S.hmeta.ptr && !S.hmeta.alloc(meta_bytes))
		return false;
	words = std::min(words, S.fs.capacity_words);
	if (S.hbits.size < size_t(words) * 4)
		if (!S.hbits.alloc(std::max<size_t>(size_t(words) * 4, 4096)))
			return false;
	uint32_t *hm = static_cast<uint32_t *>(S.hmeta.ptr);
	(void)hipMemcpyAsync(hm, S.fs.meta, size_t(S.geom.count32) * 8, hipMemcpyDeviceToHost, stream);
	(void)hipMemcpyAsync(hm + 2 * S.geom.count32, S.fs.rc, kRcCount * sizeof(uint32_t), hipMemcpyDeviceToHost,
	                     stream);
	if (words)
		(void)hipMemcpyAsync(S.hbits.ptr, S.fs.bitstream, size_t(words) * 4, hipMemcpyDeviceToHost, stream);
	S.hbits_words = words;
	return true;
}

void set_encode_params(CallParams &F, Slot &S, size_t budget)
{
	S.sequence = (S.sequence + 1) & 7u;
	S.last_budget = budget;
	F.sequence = S.sequence;
	F.target_words = target_words_of(budget & ~size_t(3));
	F.flags = S.exact_prefix ? 1u : 0u;
}

void fill_encode_call(CallParams &F, Slot &S, const pyrowave_hip_frame &frame, size_t budget)
{
	for (int c = 0; c < 3; c++)
	{
		// Pitch and step in samples.
		const pyrowave_hip_plane &p = frame.planes[c];
		const size_t sb = size_t(S.geom.sample_bytes);
		F.src[c] = static_cast<const uint8_t *>(p.data);
		F.src_pitch[c] = int32_t(p.row_stride_in_bytes / sb);
		F.src_step[c] = int32_t(p.pixel_stride_in_bytes ? p.pixel_stride_in_bytes / sb : 1);
	}
	set_encode_params(F, S, budget);
}

void fill_decode_call(CallParams &F, Slot &S, const Slot *E, const pyrowave_hip_decode_target &target)
{
	if (E)
	{
		F.dec_offsets = E->fs.meta;
		F.dec_offset_stride = 2;
		F.dec_payload = E->fs.bitstream;
		F.dec_payload_bytes = E->fs.capacity_words * 4u;
	}
	for (int c = 0; c < 3; c++)
	{
		F.dst[c] = static_cast<uint8_t *>(target.planes[c].data);
		F.dst_pitch[c] = int32_t(target.planes[c].row_stride_in_bytes / size_t(S.geom.sample_bytes));
	}
	S.decoded_frame = true;
}

} // namespace pw

namespace
{
std::mutex g_device_creation_lock;

// The precision of the reference entry points (pyrowave_encoder_create,
// pyrowave_decoder_create).  Reference pyrowave_common.cpp:15-28: the build's
// PYROWAVE_PRECISION, 1 on desktop builds (CMakeLists.txt:44-49), unless the
// PYROWAVE_PRECISION environment variable names another in [0, 2]; other
// values are reported on stderr and ignored.  0 is FP16 arithmetic and
// storage, 1 FP32 arithmetic with levels 0-1 stored as FP16, 2 FP32.
constexpr int kReferenceDefaultPrecision = 1;

int default_precision()
{
	int precision = kReferenceDefaultPrecision;
	if (const char *env = getenv("PYROWAVE_PRECISION"))
	{
		const long v = strtol(env, nullptr, 0);
		if (v < 0 || v > 2)
			fprintf(stderr, "pyrowave: precision must be in range [0, 2].\n");
		else
			precision = int(v);
	}
	switch (precision)
	{
	case 0:
		return PYROWAVE_HIP_PRECISION_FP16_MATH;
	case 2:
		return PYROWAVE_HIP_PRECISION_FP32;
	default:
		return PYROWAVE_HIP_PRECISION_FP16_LL;
	}
}

struct PlanCache
{
	std::vector<std::unique_ptr<Plan>> plans;

	Plan *get(const std::vector<Slot *> &slots, bool encoder)
	{
		for (auto &p : plans)
			if (p->encoder == encoder && p->slots == slots)
				return p.get();
		std::unique_ptr<Plan> p(new Plan);
		if (!p->build(slots, encoder))
			return nullptr;
		if (plans.size() >= 32)
			plans.erase(plans.begin());
		plans.push_back(std::move(p));
		return plans.back().get();
	}

	void forget(Slot *slot)
	{
		plans.erase(std::remove_if(plans.begin(), plans.end(),
		                           [&](const std::unique_ptr<Plan> &p) {
			                           return std::find(p->slots.begin(), p->slots.end(), slot) != p->slots.end();
		                           }),
		            plans.end());
	}
};

std::mutex g_cache_lock;
std::map<Device *, PlanCache> g_caches;

// A batch holds packed-FP16 frames or frames of the other precisions, not
// both (their transforms are different kernels).
bool mixes_half_math(const std::vector<Slot *> &slots)
{
	for (Slot *s : slots)
		if ((s->geom.precision == PYROWAVE_HIP_PRECISION_FP16_MATH) !=
		    (slots[0]->geom.precision == PYROWAVE_HIP_PRECISION_FP16_MATH))
			return true;
	return false;
}

Plan *get_plan(Device *dev, const std::vector<Slot *> &slots, bool encoder)
{
	if (slots.size() == 1)
	{
		Slot *s = slots[0];
		if (!s->single_plan)
		{
			std::unique_ptr<Plan> p(new Plan);
			if (!p->build(slots, encoder))
				return nullptr;
			s->single_plan = std::move(p);
		}
		return s->single_plan.get();
	}
	std::lock_guard<std::mutex> g(g_cache_lock);
	return g_caches[dev].get(slots, encoder);
}

void forget_slot(Device *dev, Slot *slot)
{
	std::lock_guard<std::mutex> g(g_cache_lock);
	auto it = g_caches.find(dev);
	if (it != g_caches.end())
		it->second.forget(slot);
}

bool valid_budget(const pyrowave_rate_control *rc)
{
	if (!rc)
		return false;
	size_t target = rc->maximum_bitstream_size & ~size_t(3);
	return target != 0 && target <= UINT32_MAX;
}


// Waits for the last encode and makes sure the host mirrors hold the whole
// result.  Returns false if there is none or the device reported an error.
bool finish_result(Slot &S)
{
	if (!S.have_result)
		return false;
	(void)hipSetDevice(S.dev->hip_device);
	(void)hipEventSynchronize(S.result_ready);
	if (!S.result_on_host)
	{
		uint32_t guess = uint32_t(target_words_of(S.last_budget & ~size_t(3))) + 64;
		if (!queue_result_download(S, S.dev->stream, guess))
			return false;
		(void)hipStreamSynchronize(S.dev->stream);
		S.result_on_host = true;
	}
	const uint32_t *hm = static_cast<const uint32_t *>(S.hmeta.ptr);
	if (hm[2 * S.geom.count32 + kRcError] != 0)
	{
		if (getenv("PYROWAVE_DEBUG"))
			fprintf(stderr, "pyrowave: device reported error bits 0x%x\n", hm[2 * S.geom.count32 + kRcError]);
		return false;
	}
	uint32_t total = 0;
	for (int i = 0; i < S.geom.count32; i++)
		total = std::max(total, hm[2 * i] + hm[2 * i + 1]);
	if (total > S.hbits_words)
	{
		if (S.hbits.size < size_t(total) * 4 && !S.hbits.alloc(size_t(total) * 4))
			return false;
		if (hipMemcpy(S.hbits.ptr, S.fs.bitstream, size_t(total) * 4, hipMemcpyDeviceToHost) != hipSuccess)
			return false;
		S.hbits_words = total;
	}
	return true;
}

bool ensure_plane_buffers(Slot &S)
{
	const Geometry &G = S.geom;
	size_t bytes = size_t(G.width) * G.height + 2 * size_t(G.plane_w(1)) * G.plane_h(1);
	if (!S.dplanes.ptr && !S.dplanes.alloc(bytes))
		return false;
	if (!S.hplanes.ptr && !S.hplanes.alloc(bytes))
		return false;
	return true;
}

float half_to_float(uint16_t h)
{
	uint32_t sign = uint32_t(h & 0x8000u) << 16;
	uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
	if (e == 0)
	{
		if (m == 0)
			x = sign;
		else
		{
			int s = 0;
			while (!(m & 0x400u))
			{
				m <<= 1;
				s++;
			}
			x = sign | (uint32_t(127 - 15 + 1 - s) << 23) | ((m & 0x3ffu) << 13);
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

void copy_coefficients_out(const Slot &S, float *out)
{
	const Geometry &G = S.geom;
	std::vector<uint8_t> tmp;
	for (int c = 0; c < kComponents; c++)
		for (int l = 0; l < kLevels; l++)
			for (int b = 0; b < kBands; b++)
			{
				const BandGeom &B = G.band[c][l][b];
				if (!B.exists)
					continue;
				size_t es = B.kind == kStoreF16 ? 2 : 4;
				tmp.resize(size_t(B.pitch) * B.bh * es);
				(void)hipMemcpy(tmp.data(), S.bands + B.store_off, tmp.size(), hipMemcpyDeviceToHost);
				for (int y = 0; y < B.bh; y++)
					for (int x = 0; x < B.bw; x++)
					{
						size_t i = size_t(y) * B.pitch + x;
						float v;
						if (es == 2)
						{
							uint16_t h;
							memcpy(&h, tmp.data() + 2 * i, 2);
							v = half_to_float(h);
						}
						else
							memcpy(&v, tmp.data() + 4 * i, 4);
						out[B.flat_off + long(y) * B.bw + x] = v;
					}
			}
}

uint16_t float_to_half_rne(float f)
{
	uint32_t x;
	memcpy(&x, &f, 4);
	uint32_t sign = (x >> 16) & 0x8000u;
	uint32_t ax = x & 0x7fffffffu;
	if (ax >= 0x7f800000u)
		return uint16_t(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
	if (ax == 0)
		return uint16_t(sign);
	int e = int(ax >> 23) - 127;
	uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
	if (e >= -14)
	{
		uint32_t hm = mant >> 13, rem = mant & 0x1fffu;
		if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u)))
			hm++;
		int he = e + 15;
		if (hm & 0x800u)
		{
			hm >>= 1;
			he++;
		}
		if (he >= 31)
			return uint16_t(sign | 0x7c00u);
		return uint16_t(sign | (uint32_t(he) << 10) | (hm & 0x3ffu));
	}
	int shift = -e - 1;
	if (shift > 24)
		return uint16_t(sign);
	uint32_t hm = mant >> shift, rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
	if (rem > halfway || (rem == halfway && (hm & 1u)))
		hm++;
	return uint16_t(sign | hm);
}

pyrowave_result validate_cpu_buffer(const Geometry &G, int width, int height, const pyrowave_cpu_buffer *buffers,
                                    bool decode)
{
	// reference pyrowave_c.cpp:1062-1088 (encode) and 1345-1373 (decode).
	if (!buffers)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (buffers->width != width || buffers->height != height)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (decode && buffers->format == PYROWAVE_CPU_BUFFER_FORMAT_NV12)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	// The CPU-buffer formats are 8-bit.
	if (G.sample_bytes != 1)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (G.chroma == PYROWAVE_CHROMA_SUBSAMPLING_420 && buffers->format == PYROWAVE_CPU_BUFFER_FORMAT_YUV444P)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (G.chroma == PYROWAVE_CHROMA_SUBSAMPLING_444 && buffers->format != PYROWAVE_CPU_BUFFER_FORMAT_YUV444P)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (buffers->format != PYROWAVE_CPU_BUFFER_FORMAT_NV12 && buffers->format != PYROWAVE_CPU_BUFFER_FORMAT_YUV420P &&
	    buffers->format != PYROWAVE_CPU_BUFFER_FORMAT_YUV444P)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	int num_planes = buffers->format == PYROWAVE_CPU_BUFFER_FORMAT_NV12 ? 2 : 3;
	for (int p = 0; p < num_planes; p++)
	{
		size_t pw_ = size_t(G.plane_w(p)), ph_ = size_t(G.plane_h(p));
		size_t bpp = num_planes == 2 && p == 1 ? 2 : 1;
		if (buffers->row_stride_in_bytes[p] < pw_ * bpp)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		if (buffers->row_stride_in_bytes[p] * ph_ > buffers->plane_size_in_bytes[p])
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		if (!buffers->data[p])
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
	}
	return PYROWAVE_SUCCESS;
}

void fill_src_from_internal_planes(CallParams &F, Slot &S, bool nv12)
{
	const Geometry &G = S.geom;
	uint8_t *base = static_cast<uint8_t *>(S.dplanes.ptr);
	size_t y_bytes = size_t(G.width) * G.height;
	size_t c_bytes = size_t(G.plane_w(1)) * G.plane_h(1);
	F.src[0] = base;
	F.src_pitch[0] = G.width;
	F.src_step[0] = 1;
	if (nv12)
	{
		F.src[1] = base + y_bytes;
		F.src[2] = base + y_bytes + 1;
		F.src_pitch[1] = F.src_pitch[2] = 2 * G.plane_w(1);
		F.src_step[1] = F.src_step[2] = 2;
	}
	else
	{
		F.src[1] = base + y_bytes;
		F.src[2] = base + y_bytes + c_bytes;
		F.src_pitch[1] = F.src_pitch[2] = G.plane_w(1);
		F.src_step[1] = F.src_step[2] = 1;
	}
}

void fill_dst_from_internal_planes(CallParams &F, Slot &S)
{
	const Geometry &G = S.geom;
	uint8_t *base = static_cast<uint8_t *>(S.dplanes.ptr);
	size_t y_bytes = size_t(G.width) * G.height;
	size_t c_bytes = size_t(G.plane_w(1)) * G.plane_h(1);
	F.dst[0] = base;
	F.dst[1] = base + y_bytes;
	F.dst[2] = base + y_bytes + c_bytes;
	F.dst_pitch[0] = G.width;
	F.dst_pitch[1] = F.dst_pitch[2] = G.plane_w(1);
}

// Uploads the decoder's queued payload (+16 B of padding) and offset table.
bool upload_packets(Slot &S, hipStream_t stream, CallParams &F)
{
	const size_t pay_bytes = S.payload.size() * 4;
	const size_t off_bytes = S.offsets.size() * 4;
	const size_t need = pay_bytes + 16;
	if (S.dpayload.size < need && !S.dpayload.alloc(std::max<size_t>(need * 2, 64 * 1024)))
		return false;
	if (!S.doffsets.ptr && !S.doffsets.alloc(off_bytes))
		return false;
	(void)hipEventSynchronize(S.staging_done);
	if (S.hpayload.size < need + off_bytes && !S.hpayload.alloc(std::max<size_t>((need + off_bytes) * 2, 64 * 1024)))
		return false;
	uint8_t *h = static_cast<uint8_t *>(S.hpayload.ptr);
	memcpy(h, S.offsets.data(), off_bytes);
	if (pay_bytes)
		memcpy(h + off_bytes, S.payload.data(), pay_bytes);
	memset(h + off_bytes + pay_bytes, 0, 16);
	(void)hipMemcpyAsync(S.doffsets.ptr, h, off_bytes, hipMemcpyHostToDevice, stream);
	(void)hipMemcpyAsync(S.dpayload.ptr, h + off_bytes, pay_bytes + 16, hipMemcpyHostToDevice, stream);
	(void)hipEventRecord(S.staging_done, stream);
	F.dec_offsets = static_cast<const uint32_t *>(S.doffsets.ptr);
	F.dec_offset_stride = 1;
	F.dec_payload = static_cast<const uint32_t *>(S.dpayload.ptr);
	F.dec_payload_bytes = uint32_t(pay_bytes);
	return true;
}

hipStream_t pick_stream(Device &D, void *s)
{
	return s ? static_cast<hipStream_t>(s) : D.stream;
}
} // namespace

extern "C" {

PYROWAVE_PUBLIC_API void pyrowave_get_api_version(uint32_t *major, uint32_t *minor, uint32_t *patch)
{
	if (major)
		*major = PYROWAVE_API_VERSION_MAJOR;
	if (minor)
		*minor = PYROWAVE_API_VERSION_MINOR;
	if (patch)
		*patch = PYROWAVE_API_VERSION_PATCH;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_create_device_by_compat(uint32_t vid, uint32_t pid,
                                                                     const pyrowave_uuid *device_uuid,
                                                                     const pyrowave_uuid *driver_uuid,
                                                                     const pyrowave_luid *device_luid,
                                                                     pyrowave_device *device)
{
	(void)pid;
	if (!device)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::lock_guard<std::mutex> g(g_device_creation_lock);
	if (device_uuid || driver_uuid || device_luid)
		return PYROWAVE_ERROR_NO_VULKAN;
	if (vid && vid != 0x1002)
		return PYROWAVE_ERROR_NO_VULKAN;
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
		return PYROWAVE_ERROR_NO_VULKAN;
	std::unique_ptr<pyrowave_device_opaque> d(new pyrowave_device_opaque);
	d->dev.hip_device = 0;
	if (hipSetDevice(0) != hipSuccess)
		return PYROWAVE_ERROR_NO_VULKAN;
	if (hipStreamCreateWithFlags(&d->dev.stream, hipStreamNonBlocking) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	upload_constants();
	// The allocation-free decode's scratch address on this GPU (current).
	if (!noalloc_scratch_init(d->dev.hip_device))
		return PYROWAVE_ERROR_GENERIC;
	if (const char *env = getenv("PYROWAVE_TIMING"))
		d->dev.timing = strtol(env, nullptr, 0) != 0 ? (1u << pw::kStageCount) - 1 : 0u;
	if (const char *env = getenv("PYROWAVE_FUSED"))
		d->dev.fused = (int)strtol(env, nullptr, 0);
	if (const char *env = getenv("PYROWAVE_FUSED_RESOLVE"))
		d->dev.fused_resolve = strtol(env, nullptr, 0) != 0;
	*device = d.release();
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_create_default_device(pyrowave_device *device)
{
	return pyrowave_create_device_by_compat(0, 0, nullptr, nullptr, nullptr, device);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_create_device(const pyrowave_device_create_info *info,
                                                           pyrowave_device *device)
{
	(void)info;
	(void)device;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API void pyrowave_device_report_performance_stats(pyrowave_device device, pyrowave_message_cb cb,
                                                                  void *userdata, bool reset)
{
	if (!device || !cb)
		return;
	static const char *names[kStageCount] = {"DWT", "Quant", "Analyze", "Resolve", "Packing", "Dequant", "iDWT"};
	Device &D = device->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	D.resolve_timings(true);
	for (int s = 0; s < kStageCount; s++)
	{
		if (!D.stage_frames[s])
			continue;
		char msg[256];
		snprintf(msg, sizeof(msg), "%s: %.3f ms per frame\n", names[s], D.stage_ms[s] / double(D.stage_frames[s]));
		cb(userdata, msg);
	}
	if (reset)
		for (int s = 0; s < kStageCount; s++)
		{
			D.stage_ms[s] = 0.0;
			D.stage_frames[s] = 0;
		}
}

PYROWAVE_PUBLIC_API void pyrowave_device_get_vk_device_handles(pyrowave_device device, VkInstance *vk_instance,
                                                               VkPhysicalDevice *vk_physical_device, VkDevice *vk_device)
{
	(void)device;
	if (vk_instance)
		*vk_instance = nullptr;
	if (vk_physical_device)
		*vk_physical_device = nullptr;
	if (vk_device)
		*vk_device = nullptr;
}

PYROWAVE_PUBLIC_API void pyrowave_device_set_command_buffer(pyrowave_device device, VkCommandBuffer cmd)
{
	(void)device;
	(void)cmd;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_device_set_queue_type(pyrowave_device device, VkQueueFlagBits queue_flags)
{
	// reference pyrowave_c.cpp:386-395.
	if (!device)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (queue_flags != VK_QUEUE_GRAPHICS_BIT && queue_flags != VK_QUEUE_COMPUTE_BIT)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API bool pyrowave_device_confirm_interop_support(pyrowave_device device)
{
	(void)device;
	return false;
}

PYROWAVE_PUBLIC_API void pyrowave_device_destroy(pyrowave_device device)
{
	if (!device)
		return;
	(void)hipSetDevice(device->dev.hip_device);
	(void)hipStreamSynchronize(device->dev.stream);
	{
		std::lock_guard<std::mutex> g(g_cache_lock);
		g_caches.erase(&device->dev);
	}
	delete device;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_sync_object_create(const pyrowave_sync_object_create_info *info,
                                                                pyrowave_sync_object *sync)
{
	(void)info;
	if (sync)
		*sync = nullptr;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API VkSemaphore pyrowave_sync_object_get_semaphore(pyrowave_sync_object sync)
{
	(void)sync;
	return nullptr;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_sync_object_export_handle(pyrowave_sync_object sync,
                                                                       pyrowave_os_handle *handle)
{
	(void)sync;
	(void)handle;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_sync_object_cpu_wait(pyrowave_sync_object sync, uint64_t value,
                                                                  uint64_t timeout)
{
	(void)sync;
	(void)value;
	(void)timeout;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_sync_object_cpu_signal(pyrowave_sync_object sync, uint64_t value)
{
	(void)sync;
	(void)value;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API void pyrowave_sync_object_destroy(pyrowave_sync_object sync)
{
	(void)sync;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_image_create(const pyrowave_image_create_info *info, pyrowave_image *image)
{
	(void)info;
	if (image)
		*image = nullptr;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API VkImage pyrowave_image_get_handle(pyrowave_image image)
{
	(void)image;
	return nullptr;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_image_get_image_view(pyrowave_image image, VkImageAspectFlagBits aspect,
                                                                  VkImageUsageFlagBits usage, pyrowave_image_view *view)
{
	(void)image;
	(void)aspect;
	(void)usage;
	(void)view;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API void pyrowave_image_destroy(pyrowave_image image)
{
	(void)image;
}

static pyrowave_result create_encoder(pyrowave_device device, int width, int height, pyrowave_chroma_subsampling chroma,
                                      int precision, uint32_t flags, pyrowave_encoder *encoder)
{
	// reference pyrowave_c.cpp:807-836.
	if (!device || !encoder)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (width <= 0 || height <= 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (chroma != PYROWAVE_CHROMA_SUBSAMPLING_420 && chroma != PYROWAVE_CHROMA_SUBSAMPLING_444)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (chroma == PYROWAVE_CHROMA_SUBSAMPLING_420 && (width % 2 || height % 2))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (!pw::valid_precision(precision))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	(void)hipSetDevice(device->dev.hip_device);
	std::unique_ptr<pyrowave_encoder_opaque> e(new pyrowave_encoder_opaque);
	e->device = device;
	e->slot.exact_prefix = (flags & PYROWAVE_HIP_ENCODER_EXACT_PREFIX_BIT) != 0;
	if (!e->slot.init(&device->dev, width, height, int(chroma), precision, true))
		return PYROWAVE_ERROR_GENERIC;
	e->slot.geom.sample_bytes = (flags & PYROWAVE_HIP_ENCODER_UNORM16_INPUT_BIT) ? 2 : 1;
	device->dev.live_objects++;
	*encoder = e.release();
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_encoder_create(const pyrowave_encoder_create_info *info,
                                                            pyrowave_encoder *encoder)
{
	if (!info)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	return create_encoder(info->device, info->width, info->height, info->chroma, default_precision(), 0, encoder);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_create(const pyrowave_hip_encoder_create_info *info,
                                                                pyrowave_encoder *encoder)
{
	if (!info)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	return create_encoder(info->device, info->width, info->height, info->chroma, int(info->precision), info->flags,
	                      encoder);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_encoder_encode_gpu_synchronous(pyrowave_encoder encoder,
                                                                            const pyrowave_gpu_sync_operation *acquire,
                                                                            const pyrowave_gpu_sync_operation *release,
                                                                            const pyrowave_gpu_buffers *buffers,
                                                                            const pyrowave_rate_control *rate_control)
{
	(void)encoder;
	(void)acquire;
	(void)release;
	(void)buffers;
	(void)rate_control;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_encoder_encode_cpu_synchronous(pyrowave_encoder encoder,
                                                                            const pyrowave_cpu_buffer *buffers,
                                                                            const pyrowave_rate_control *rate_control)
{
	if (!encoder)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	Device &D = *S.dev;
	const Geometry &G = S.geom;
	pyrowave_result r = validate_cpu_buffer(G, G.width, G.height, buffers, false);
	if (r != PYROWAVE_SUCCESS)
		return r;
	if (!valid_budget(rate_control))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	if (!ensure_plane_buffers(S))
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	Plan *P = get_plan(&D, {&S}, true);
	if (!P)
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;

	// Stage the planes in pinned memory, then one asynchronous upload.
	const bool nv12 = buffers->format == PYROWAVE_CPU_BUFFER_FORMAT_NV12;
	(void)hipEventSynchronize(S.staging_done);
	uint8_t *h = static_cast<uint8_t *>(S.hplanes.ptr);
	size_t off = 0;
	const int num_planes = nv12 ? 2 : 3;
	for (int p = 0; p < num_planes; p++)
	{
		size_t row = size_t(G.plane_w(p)) * (nv12 && p == 1 ? 2 : 1);
		const uint8_t *src = static_cast<const uint8_t *>(buffers->data[p]);
		for (int y = 0; y < G.plane_h(p); y++)
			memcpy(h + off + y * row, src + y * buffers->row_stride_in_bytes[p], row);
		off += row * G.plane_h(p);
	}
	(void)hipMemcpyAsync(S.dplanes.ptr, h, off, hipMemcpyHostToDevice, D.stream);
	(void)hipEventRecord(S.staging_done, D.stream);

	CallParams &F = P->call(0);
	fill_src_from_internal_planes(F, S, nv12);
	set_encode_params(F, S, rate_control->maximum_bitstream_size);
	run_encode(*P, D.stream);
	if (!queue_result_download(S, D.stream, uint32_t(F.target_words) + 64))
		return PYROWAVE_ERROR_OUT_OF_HOST_MEMORY;
	(void)hipEventRecord(S.result_ready, D.stream);
	S.have_result = true;
	S.result_on_host = true;
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_encoder_compute_num_packets(pyrowave_encoder encoder,
                                                                         size_t packet_boundary, size_t *num_packets)
{
	if (!encoder || !num_packets)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	std::lock_guard<std::recursive_mutex> g(S.dev->lock);
	if (!finish_result(S))
		return PYROWAVE_ERROR_GENERIC;
	// reference pyrowave_encoder.cpp:596-623.
	const uint32_t *meta = static_cast<const uint32_t *>(S.hmeta.ptr);
	size_t n = 0, in_packet = 8;
	for (int i = 0; i < S.geom.count32; i++)
	{
		size_t sz = size_t(meta[2 * i + 1]) * 4;
		if (!sz)
			continue;
		if (in_packet + sz > packet_boundary)
		{
			in_packet = 0;
			n++;
		}
		in_packet += sz;
	}
	if (in_packet)
		n++;
	*num_packets = n;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_encoder_packetize(pyrowave_encoder encoder, pyrowave_packet *packets,
                                                               size_t packet_boundary, size_t *out_packets,
                                                               void *bitstream, size_t size)
{
	if (!encoder || !out_packets || !bitstream || !packets)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	std::lock_guard<std::recursive_mutex> g(S.dev->lock);
	if (!finish_result(S))
		return PYROWAVE_ERROR_GENERIC;
	// reference pyrowave_encoder.cpp:1048-1115.  Blocks are already stored
	// contiguously in block-index order, so every packet is one memcpy span.
	const Geometry &G = S.geom;
	const uint32_t *meta = static_cast<const uint32_t *>(S.hmeta.ptr);
	const uint8_t *bits = static_cast<const uint8_t *>(S.hbits.ptr);
	uint8_t *out = static_cast<uint8_t *>(bitstream);
	uint32_t nonzero = 0;
	for (int i = 0; i < G.count32; i++)
		nonzero += meta[2 * i + 1] ? 1 : 0;
	if (size < 8)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	uint32_t h0 = uint32_t(G.width - 1) | (uint32_t(G.height - 1) << 14) | ((S.sequence & 7u) << 28) | (1u << 31);
	uint32_t h1 = (nonzero & 0xffffffu) | (uint32_t(G.chroma == PYROWAVE_CHROMA_SUBSAMPLING_444) << 26);
	memcpy(out, &h0, 4);
	memcpy(out + 4, &h1, 4);
	size_t n = 0, in_packet = 8, packet_offset = 0, out_off = 8;
	size_t span_src = 0, span_len = 0; // pending contiguous copy
	for (int i = 0; i < G.count32; i++)
	{
		size_t sz = size_t(meta[2 * i + 1]) * 4;
		if (!sz)
			continue;
		if (in_packet + sz > packet_boundary)
		{
			packets[n].offset = packet_offset;
			packets[n].size = in_packet;
			n++;
			in_packet = 0;
			packet_offset = out_off;
		}
		if (out_off + sz > size)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		size_t src = size_t(meta[2 * i]) * 4;
		if (span_len && span_src + span_len == src)
			span_len += sz;
		else
		{
			if (span_len)
				memcpy(out + out_off - span_len, bits + span_src, span_len);
			span_src = src;
			span_len = sz;
		}
		out_off += sz;
		in_packet += sz;
	}
	if (span_len)
		memcpy(out + out_off - span_len, bits + span_src, span_len);
	if (in_packet)
	{
		packets[n].offset = packet_offset;
		packets[n].size = in_packet;
		n++;
	}
	*out_packets = n;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API void pyrowave_encoder_destroy(pyrowave_encoder encoder)
{
	if (!encoder)
		return;
	Device *D = encoder->slot.dev;
	if (D)
	{
		std::lock_guard<std::recursive_mutex> g(D->lock);
		(void)hipSetDevice(D->hip_device);
		(void)hipStreamSynchronize(D->stream);
		D->resolve_timings(true);
		forget_slot(D, &encoder->slot);
		D->live_objects--;
	}
	delete encoder;
}

static pyrowave_result create_decoder(pyrowave_device device, int width, int height, pyrowave_chroma_subsampling chroma,
                                      int precision, uint32_t flags, pyrowave_decoder *decoder)
{
	// reference pyrowave_c.cpp:1198-1226 (init failure is INVALID_ARGUMENT).
	if (!device || !decoder)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (width <= 0 || height <= 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (chroma != PYROWAVE_CHROMA_SUBSAMPLING_420 && chroma != PYROWAVE_CHROMA_SUBSAMPLING_444)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (chroma == PYROWAVE_CHROMA_SUBSAMPLING_420 && (width % 2 || height % 2))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (!pw::valid_precision(precision))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	(void)hipSetDevice(device->dev.hip_device);
	std::unique_ptr<pyrowave_decoder_opaque> d(new pyrowave_decoder_opaque);
	d->device = device;
	if (!d->slot.init(&device->dev, width, height, int(chroma), precision, false))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	d->slot.geom.sample_bytes = (flags & PYROWAVE_HIP_DECODER_UNORM16_OUTPUT_BIT) ? 2 : 1;
	// The single-frame launch tables up front, so that decoding one frame
	// (pyrowave_hip_decode_packets_noalloc in particular) allocates nothing.
	if (!get_plan(&device->dev, {&d->slot}, false))
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	device->dev.live_objects++;
	*decoder = d.release();
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API bool pyrowave_decoder_device_prefers_fragment_path(pyrowave_device device)
{
	(void)device;
	return false;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_decoder_create(const pyrowave_decoder_create_info *info,
                                                            pyrowave_decoder *decoder)
{
	if (!info)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	return create_decoder(info->device, info->width, info->height, info->chroma, default_precision(), 0, decoder);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_decoder_create(const pyrowave_hip_decoder_create_info *info,
                                                                pyrowave_decoder *decoder)
{
	if (!info)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	return create_decoder(info->device, info->width, info->height, info->chroma, int(info->precision), info->flags,
	                      decoder);
}

PYROWAVE_PUBLIC_API void pyrowave_decoder_clear(pyrowave_decoder decoder)
{
	if (decoder)
		decoder->slot.clear_packets();
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_decoder_push_packet(pyrowave_decoder decoder, const void *data, size_t size)
{
	if (!decoder || (!data && size))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	return decoder->slot.push_packet(data, size) ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_INVALID_ARGUMENT;
}

PYROWAVE_PUBLIC_API bool pyrowave_decoder_decode_is_ready(pyrowave_decoder decoder, bool allow_partial_frame)
{
	return decoder && decoder->slot.decode_is_ready(allow_partial_frame);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_decoder_decode_gpu_buffer(pyrowave_decoder decoder,
                                                                       const pyrowave_gpu_sync_operation *acquire,
                                                                       const pyrowave_gpu_sync_operation *release,
                                                                       const pyrowave_gpu_buffers *buffers)
{
	(void)decoder;
	(void)acquire;
	(void)release;
	(void)buffers;
	return PYROWAVE_ERROR_NOT_IMPLEMENTED;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_decoder_decode_cpu_buffer_synchronous(pyrowave_decoder decoder,
                                                                                   const pyrowave_cpu_buffer *buffers)
{
	if (!decoder)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = decoder->slot;
	Device &D = *S.dev;
	const Geometry &G = S.geom;
	pyrowave_result r = validate_cpu_buffer(G, G.width, G.height, buffers, true);
	if (r != PYROWAVE_SUCCESS)
		return r;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	if (!ensure_plane_buffers(S))
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	Plan *P = get_plan(&D, {&S}, false);
	if (!P)
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	CallParams &F = P->call(0);
	if (!upload_packets(S, D.stream, F))
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	fill_dst_from_internal_planes(F, S);
	run_decode(*P, D.stream);
	S.decoded_frame = true;

	size_t total = size_t(G.width) * G.height + 2 * size_t(G.plane_w(1)) * G.plane_h(1);
	(void)hipMemcpyAsync(S.hplanes.ptr, S.dplanes.ptr, total, hipMemcpyDeviceToHost, D.stream);
	if (hipStreamSynchronize(D.stream) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	const uint8_t *h = static_cast<const uint8_t *>(S.hplanes.ptr);
	size_t off = 0;
	for (int p = 0; p < 3; p++)
	{
		size_t row = size_t(G.plane_w(p));
		uint8_t *dst = static_cast<uint8_t *>(buffers->data[p]);
		for (int y = 0; y < G.plane_h(p); y++)
			memcpy(dst + y * buffers->row_stride_in_bytes[p], h + off + y * row, row);
		off += row * G.plane_h(p);
	}
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API void pyrowave_decoder_destroy(pyrowave_decoder decoder)
{
	if (!decoder)
		return;
	Device *D = decoder->slot.dev;
	if (D)
	{
		std::lock_guard<std::recursive_mutex> g(D->lock);
		(void)hipSetDevice(D->hip_device);
		(void)hipStreamSynchronize(D->stream);
		D->resolve_timings(true);
		forget_slot(D, &decoder->slot);
		D->live_objects--;
	}
	delete decoder;
}

// ---------------------------------------------------------------------------
// HIP-native entry points
// ---------------------------------------------------------------------------

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encode_batch(pyrowave_encoder *encoders, const pyrowave_hip_frame *frames,
                                                              const pyrowave_rate_control *rate_controls, uint32_t count,
                                                              void *hip_stream)
{
	if (!encoders || !frames || !rate_controls || count == 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::vector<Slot *> slots(count);
	for (uint32_t i = 0; i < count; i++)
	{
		if (!encoders[i] || encoders[i]->device != encoders[0]->device)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		if (!valid_budget(&rate_controls[i]))
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		slots[i] = &encoders[i]->slot;
		for (uint32_t j = 0; j < i; j++)
			if (slots[j] == slots[i])
				return PYROWAVE_ERROR_INVALID_ARGUMENT;
		if (!frame_fits(slots[i]->geom, frames[i]))
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
	}
	Device &D = *slots[0]->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	hipStream_t stream = pick_stream(D, hip_stream);
	if (mixes_half_math(slots))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Plan *P = get_plan(&D, slots, true);
	if (!P)
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	for (uint32_t i = 0; i < count; i++)
		fill_encode_call(P->call(int(i)), *slots[i], frames[i], rate_controls[i].maximum_bitstream_size);
	if (D.graphs)
	{
		if (!run_graph(*P, stream))
			return PYROWAVE_ERROR_GENERIC;
	}
	else
		run_encode(*P, stream);
	for (uint32_t i = 0; i < count; i++)
	{
		(void)hipEventRecord(slots[i]->result_ready, stream);
		slots[i]->have_result = true;
		slots[i]->result_on_host = false;
	}
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_get_result(pyrowave_encoder encoder,
                                                                    pyrowave_hip_encoded_frame *result)
{
	if (!encoder || !result)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	if (!S.have_result)
		return PYROWAVE_ERROR_GENERIC;
	result->bitstream = S.fs.bitstream;
	result->meta = S.fs.meta;
	result->num_blocks = uint32_t(S.geom.count32);
	result->sequence = S.sequence;
	return PYROWAVE_SUCCESS;
}

static pyrowave_result decode_batch_common(pyrowave_decoder *decoders, pyrowave_encoder *encoders,
                                           const pyrowave_hip_decode_target *targets, uint32_t count, void *hip_stream)
{
	if (!decoders || !targets || count == 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::vector<Slot *> slots(count);
	for (uint32_t i = 0; i < count; i++)
	{
		if (!decoders[i] || decoders[i]->device != decoders[0]->device)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		slots[i] = &decoders[i]->slot;
		for (uint32_t j = 0; j < i; j++)
			if (slots[j] == slots[i])
				return PYROWAVE_ERROR_INVALID_ARGUMENT;
		const Geometry &G = slots[i]->geom;
		if (!target_fits(G, targets[i]))
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		if (encoders)
		{
			if (!encoders[i] || !encoders[i]->slot.have_result)
				return PYROWAVE_ERROR_INVALID_ARGUMENT;
			const Geometry &E = encoders[i]->slot.geom;
			if (E.width != G.width || E.height != G.height || E.chroma != G.chroma)
				return PYROWAVE_ERROR_INVALID_ARGUMENT;
		}
	}
	Device &D = *slots[0]->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	hipStream_t stream = pick_stream(D, hip_stream);
	if (mixes_half_math(slots))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Plan *P = get_plan(&D, slots, false);
	if (!P)
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	for (uint32_t i = 0; i < count; i++)
	{
		CallParams &F = P->call(int(i));
		fill_decode_call(F, *slots[i], encoders ? &encoders[i]->slot : nullptr, targets[i]);
		if (!encoders && !upload_packets(*slots[i], stream, F))
			return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	}
	if (D.graphs)
	{
		if (!run_graph(*P, stream))
			return PYROWAVE_ERROR_GENERIC;
	}
	else
		run_decode(*P, stream);
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_decode_batch(pyrowave_decoder *decoders,
                                                              const pyrowave_hip_decode_target *targets, uint32_t count,
                                                              void *hip_stream)
{
	return decode_batch_common(decoders, nullptr, targets, count, hip_stream);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_decode_encoded_batch(pyrowave_decoder *decoders,
                                                                      pyrowave_encoder *encoders,
                                                                      const pyrowave_hip_decode_target *targets,
                                                                      uint32_t count, void *hip_stream)
{
	if (!encoders)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	return decode_batch_common(decoders, encoders, targets, count, hip_stream);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_packetize_batch(pyrowave_encoder *encoders, uint32_t count,
                                                                         size_t packet_boundary, void *hip_stream)
{
	if (!encoders || count == 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::vector<PacketizeJob> jobs(count);
	for (uint32_t i = 0; i < count; i++)
	{
		if (!encoders[i] || encoders[i]->device != encoders[0]->device || !encoders[i]->slot.have_result)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		Slot &S = encoders[i]->slot;
		const Geometry &G = S.geom;
		PacketizeJob &J = jobs[i];
		J.meta = S.fs.meta;
		J.stream = S.fs.bitstream - 2;
		J.table = S.dtable;
		J.info = S.dpinfo;
		J.n32 = uint32_t(G.count32);
		// Sequence header (bitstream/bitstream.md:204-315), as the host
		// packetize writes it.
		J.h0 = uint32_t(G.width - 1) | (uint32_t(G.height - 1) << 14) | ((S.sequence & 7u) << 28) | (1u << 31);
		J.h1 = uint32_t(G.chroma == PYROWAVE_CHROMA_SUBSAMPLING_444) << 26;
		J.boundary = uint32_t(std::min<size_t>(packet_boundary, UINT32_MAX));
	}
	Device &D = *encoders[0]->slot.dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	hipStream_t stream = pick_stream(D, hip_stream);
	launch_packetize(jobs.data(), int(count), stream);
	for (uint32_t i = 0; i < count; i++)
	{
		Slot &S = encoders[i]->slot;
		(void)hipEventRecord(S.result_ready, stream);
		S.device_packets = true;
	}
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_read_packets(pyrowave_encoder encoder, void *data, size_t size,
                                                                      pyrowave_packet *packets, uint32_t max_packets,
                                                                      size_t *bytes, uint32_t *num_packets)
{
	if (!encoder)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	if (!S.device_packets)
		return PYROWAVE_ERROR_GENERIC;
	(void)hipSetDevice(S.dev->hip_device);
	if (hipEventSynchronize(S.result_ready) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	uint32_t info[2] = {};
	if (hipMemcpy(info, S.dpinfo, sizeof(info), hipMemcpyDeviceToHost) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	if (bytes)
		*bytes = info[0];
	if (num_packets)
		*num_packets = info[1];
	if (data && size >= info[0] &&
	    hipMemcpy(data, S.fs.bitstream - 2, info[0], hipMemcpyDeviceToHost) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	if (packets && max_packets >= info[1])
	{
		std::vector<uint32_t> t(size_t(info[1]) * 2);
		if (hipMemcpy(t.data(), S.dtable, t.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
			return PYROWAVE_ERROR_GENERIC;
		for (uint32_t i = 0; i < info[1]; i++)
			packets[i] = {t[2 * i], t[2 * i + 1]};
	}
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_get_packets(pyrowave_encoder encoder,
                                                                     pyrowave_hip_packet_stream *packets,
                                                                     const uint32_t **stream_bytes)
{
	if (!encoder || !packets)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const Slot &S = encoder->slot;
	packets->data = S.fs.bitstream - 2;
	packets->data_bytes = (size_t(S.fs.capacity_words) + 2) * 4;
	packets->packets = S.dtable;
	packets->count = S.dpinfo + 1;
	packets->max_packets = uint32_t(S.geom.count32) + 1;
	if (stream_bytes)
		*stream_bytes = S.dpinfo;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_decode_packets_batch(pyrowave_decoder *decoders,
                                                                      const pyrowave_hip_packet_stream *packets,
                                                                      const pyrowave_hip_decode_target *targets,
                                                                      uint32_t count, void *hip_stream)
{
	if (!decoders || !packets || !targets || count == 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	std::vector<Slot *> slots(count);
	for (uint32_t i = 0; i < count; i++)
	{
		if (!decoders[i] || decoders[i]->device != decoders[0]->device)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		slots[i] = &decoders[i]->slot;
		for (uint32_t j = 0; j < i; j++)
			if (slots[j] == slots[i])
				return PYROWAVE_ERROR_INVALID_ARGUMENT;
		const Geometry &G = slots[i]->geom;
		if (!target_fits(G, targets[i]))
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		const pyrowave_hip_packet_stream &P = packets[i];
		if (!P.data || (reinterpret_cast<uintptr_t>(P.data) & 3) || P.data_bytes > UINT32_MAX || !P.packets ||
		    !P.count)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
	}
	Device &D = *slots[0]->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	hipStream_t stream = pick_stream(D, hip_stream);
	std::vector<ParseJob> jobs(count);
	for (uint32_t i = 0; i < count; i++)
	{
		Slot &S = *slots[i];
		const Geometry &G = S.geom;
		const size_t bytes = (size_t(G.count32) + 4) * 4;
		if (S.dparse.size < bytes && !S.dparse.alloc(bytes))
			return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
		const pyrowave_hip_packet_stream &P = packets[i];
		ParseJob &J = jobs[i];
		J.data = static_cast<const uint32_t *>(P.data);
		J.packets = P.packets;
		J.count = P.count;
		J.offsets = static_cast<uint32_t *>(S.dparse.ptr);
		J.info = J.offsets + G.count32;
		J.data_bytes = uint32_t(P.data_bytes);
		J.max_packets = P.max_packets;
		J.n32 = uint32_t(G.count32);
		J.h0_size = uint32_t(G.width - 1) | (uint32_t(G.height - 1) << 14);
		J.chroma = G.chroma == PYROWAVE_CHROMA_SUBSAMPLING_444 ? 1u : 0u;
	}
	if (mixes_half_math(slots))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Plan *Pl = get_plan(&D, slots, false);
	if (!Pl)
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	launch_parse(jobs.data(), int(count), stream);
	for (uint32_t i = 0; i < count; i++)
	{
		Slot &S = *slots[i];
		(void)hipEventRecord(S.result_ready, stream);
		S.parsed_on_device = true;
		CallParams &F = Pl->call(int(i));
		fill_decode_call(F, S, nullptr, targets[i]);
		F.dec_offsets = jobs[i].offsets;
		F.dec_offset_stride = 1;
		F.dec_payload = jobs[i].data;
		F.dec_payload_bytes = jobs[i].data_bytes;
	}
	if (D.graphs)
	{
		if (!run_graph(*Pl, stream))
			return PYROWAVE_ERROR_GENERIC;
	}
	else
		run_decode(*Pl, stream);
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_decode_packets_noalloc(pyrowave_decoder decoder,
                                                                       const pyrowave_hip_packet_stream *packets,
                                                                       const pyrowave_hip_decode_target *target,
                                                                       uint32_t *device_info, void *hip_stream)
{
	// pyrowave_hip_decode_packets_batch for one decoder, with the parse's
	// offset table in the device's static scratch instead of decoder memory
	// allocated on first use; no allocation, no graph capture, no timing
	// events.
	if (!decoder || !packets || !target)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = decoder->slot;
	const Geometry &G = S.geom;
	if (!target_fits(G, *target))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const pyrowave_hip_packet_stream &P = *packets;
	if (!P.data || (reinterpret_cast<uintptr_t>(P.data) & 3) || P.data_bytes > UINT32_MAX || !P.packets || !P.count)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (uint32_t(G.count32) > kNoAllocMaxBlocks)
	{
		if (getenv("PYROWAVE_DEBUG"))
			fprintf(stderr, "pyrowave: %d blocks exceed the allocation-free decode's %u\n", G.count32,
			        kNoAllocMaxBlocks);
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	}
	Device &D = *S.dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	uint32_t *offsets = noalloc_scratch(D.hip_device);
	Plan *Pl = S.single_plan.get();
	if (!offsets || !Pl)
		return PYROWAVE_ERROR_GENERIC;
	hipStream_t stream = pick_stream(D, hip_stream);
	ParseJob J = {};
	J.data = static_cast<const uint32_t *>(P.data);
	J.packets = P.packets;
	J.count = P.count;
	J.offsets = offsets;
	J.info = device_info ? device_info : offsets + kNoAllocMaxBlocks;
	J.data_bytes = uint32_t(P.data_bytes);
	J.max_packets = P.max_packets;
	J.n32 = uint32_t(G.count32);
	J.h0_size = uint32_t(G.width - 1) | (uint32_t(G.height - 1) << 14);
	J.chroma = G.chroma == PYROWAVE_CHROMA_SUBSAMPLING_444 ? 1u : 0u;
	launch_parse(&J, 1, stream);
	CallParams &F = Pl->call(0);
	fill_decode_call(F, S, nullptr, *target);
	F.dec_offsets = offsets;
	F.dec_offset_stride = 1;
	F.dec_payload = J.data;
	F.dec_payload_bytes = J.data_bytes;
	run_decode(*Pl, stream, false);
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_decoder_parse_result(pyrowave_decoder decoder, uint32_t *blocks_present,
                                                                      uint32_t *blocks_announced)
{
	if (!decoder)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = decoder->slot;
	if (!S.parsed_on_device)
		return PYROWAVE_ERROR_GENERIC;
	(void)hipSetDevice(S.dev->hip_device);
	if (hipEventSynchronize(S.result_ready) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	uint32_t info[4] = {};
	const uint32_t *d = static_cast<const uint32_t *>(S.dparse.ptr) + S.geom.count32;
	if (hipMemcpy(info, d, sizeof(info), hipMemcpyDeviceToHost) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	if (blocks_present)
		*blocks_present = info[0];
	if (blocks_announced)
		*blocks_announced = info[1];
	return info[2] ? PYROWAVE_ERROR_INVALID_ARGUMENT : PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_autotune_batch(pyrowave_encoder *encoders, pyrowave_decoder *decoders,
                                                                const pyrowave_hip_frame *frames,
                                                                const pyrowave_rate_control *rate_controls,
                                                                const pyrowave_hip_decode_target *targets,
                                                                uint32_t count, uint32_t repetitions, void *hip_stream,
                                                                int *forward_variants, int *inverse_variants)
{
	if (!encoders || count == 0 || !encoders[0] || (decoders && !targets))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Device &D = encoders[0]->device->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	const bool graphs = D.graphs;
	D.graphs = false;
	D.tune_reps = int(std::max<uint32_t>(1u, std::min<uint32_t>(repetitions, 64u)));
	pyrowave_result r = pyrowave_hip_encode_batch(encoders, frames, rate_controls, count, hip_stream);
	if (r == PYROWAVE_SUCCESS && decoders)
		r = pyrowave_hip_decode_encoded_batch(decoders, encoders, targets, count, hip_stream);
	D.tune_reps = 0;
	D.graphs = graphs;
	if (r != PYROWAVE_SUCCESS)
		return r;
	std::vector<Slot *> enc_slots(count), dec_slots;
	for (uint32_t i = 0; i < count; i++)
		enc_slots[i] = &encoders[i]->slot;
	if (Plan *P = get_plan(&D, enc_slots, true))
		for (int l = 0; l < kLevels && forward_variants; l++)
			forward_variants[l] = P->variant[l];
	if (decoders)
	{
		for (uint32_t i = 0; i < count; i++)
			dec_slots.push_back(&decoders[i]->slot);
		if (Plan *P = get_plan(&D, dec_slots, false))
			for (int l = 0; l < kLevels && inverse_variants; l++)
				inverse_variants[l] = P->variant[l];
	}
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API void pyrowave_hip_device_force_tile_variant(pyrowave_device device, int variant)
{
	if (!device)
		return;
	std::lock_guard<std::recursive_mutex> g(device->dev.lock);
	device->dev.forced_variant = variant >= 0 && variant < kDwtVariants ? variant : -1;
}

PYROWAVE_PUBLIC_API void pyrowave_hip_device_set_graphs(pyrowave_device device, bool enable)
{
	if (!device)
		return;
	std::lock_guard<std::recursive_mutex> g(device->dev.lock);
	device->dev.graphs = enable;
}

} // extern "C"

namespace
{
// reference psnr.cpp:25-28
float contrast_sensitivity(float cpd)
{
	return 2.6f * (0.0192f + 0.114f * cpd) * std::exp(-std::pow(0.114f * cpd, 1.1f));
}

// reference psnr.cpp:154-204, 321-324: csf, mask and inverse mask tables of
// every height factor for an image `height` rows tall.
void psnr_tables(int height, float *out)
{
	for (int f = 0; f < kPsnrFactors; f++)
	{
		float *csf = out + f * 192, *mask = csf + 64, *inv = csf + 128;
		const float height_factor = 1.0f + float(f) / 8.0f;
		const float nyquist_cpd = float(height) * height_factor * 3.14159265358979323846f / 360.0f;
		for (int y = 0; y < 8; y++)
			for (int x = 0; x < 8; x++)
			{
				const float H = (float(x) + 0.5f) / 8.0f, V = (float(y) + 0.5f) / 8.0f;
				const float cpd = std::sqrt(H * H + V * V);
				csf[8 * y + x] = 2.6f * contrast_sensitivity(cpd * nyquist_cpd);
			}
		float max_csf = 0.0f;
		for (int i = 0; i < 64; i++)
			max_csf = std::max(max_csf, csf[i]);
		float norm = 1.0f / max_csf;
		norm *= norm;
		for (int i = 0; i < 64; i++)
			mask[i] = csf[i] * csf[i] * norm;
		mask[0] = 0.0f;
		for (int i = 0; i < 64; i++)
			inv[i] = i ? 1.0f / mask[i] : 0.0f;
	}
}
} // namespace

extern "C" {

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_psnr_hvs_m_accumulate(pyrowave_device device, const void *reference,
                                                                       size_t reference_stride, const void *distorted,
                                                                       size_t distorted_stride, int width, int height,
                                                                       uint64_t *device_totals, uint64_t *total_pixels,
                                                                       void *hip_stream)
{
	if (!device || !reference || !distorted || !device_totals || width < 1 || height < 1 || width > 16384 ||
	    height > 16384 || reference_stride < size_t(width) || distorted_stride < size_t(width) ||
	    reference_stride > UINT32_MAX || distorted_stride > UINT32_MAX)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Device &D = device->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	auto it = D.psnr_tables.find(height);
	if (it == D.psnr_tables.end())
	{
		std::vector<float> t(size_t(kPsnrFactors) * 192);
		psnr_tables(height, t.data());
		std::unique_ptr<DeviceBuffer> buf(new DeviceBuffer);
		if (!buf->alloc(t.size() * sizeof(float)))
			return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
		if (hipMemcpy(buf->ptr, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
			return PYROWAVE_ERROR_GENERIC;
		it = D.psnr_tables.emplace(height, std::move(buf)).first;
	}
	launch_psnr_hvs_m(static_cast<const uint8_t *>(reference), uint32_t(reference_stride),
	                  static_cast<const uint8_t *>(distorted), uint32_t(distorted_stride), width, height,
	                  static_cast<const float *>(it->second->ptr), device_totals, pick_stream(D, hip_stream));
	if (total_pixels)
		*total_pixels += uint64_t((width - 1) / 4 + 1) * uint64_t((height - 1) / 4 + 1) * 64u;
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API void pyrowave_hip_psnr_hvs_m_db(const uint64_t *totals, uint64_t total_pixels, bool full_range,
                                                    double *db)
{
	if (!totals || !db)
		return;
	// reference psnr.cpp:247-264
	const double peak = full_range ? 1.0f * 1.0f : (223.0f * 223.0f) / (255.0f * 255.0f);
	for (int i = 0; i < kPsnrFactors; i++)
		db[i] = 10.0 * std::log10(double(total_pixels) * peak / std::ldexp(double(totals[i]), -24));
}

PYROWAVE_PUBLIC_API size_t pyrowave_hip_coefficient_count(int width, int height, pyrowave_chroma_subsampling chroma)
{
	Geometry G;
	if (chroma == PYROWAVE_CHROMA_SUBSAMPLING_420 && (width % 2 || height % 2))
		return 0;
	if (!G.init(width, height, int(chroma), PYROWAVE_HIP_PRECISION_FP32))
		return 0;
	return size_t(G.flat_count);
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_read_coefficients(pyrowave_encoder encoder, float *coefficients,
                                                                           size_t count)
{
	if (!encoder || !coefficients || count < size_t(encoder->slot.geom.flat_count))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	std::lock_guard<std::recursive_mutex> g(S.dev->lock);
	(void)hipSetDevice(S.dev->hip_device);
	if (hipDeviceSynchronize() != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	copy_coefficients_out(S, coefficients);
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_decoder_read_coefficients(pyrowave_decoder decoder, float *coefficients,
                                                                           size_t count)
{
	if (!decoder || !coefficients || count < size_t(decoder->slot.geom.flat_count))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = decoder->slot;
	std::lock_guard<std::recursive_mutex> g(S.dev->lock);
	(void)hipSetDevice(S.dev->hip_device);
	if (hipDeviceSynchronize() != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	copy_coefficients_out(S, coefficients);
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_encode_coefficients(pyrowave_encoder encoder,
                                                                             const float *coefficients, size_t count,
                                                                             const pyrowave_rate_control *rate_control)
{
	if (!encoder || !coefficients || count < size_t(encoder->slot.geom.flat_count))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (!valid_budget(rate_control))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	Device &D = *S.dev;
	const Geometry &G = S.geom;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	if (hipStreamSynchronize(D.stream) != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	std::vector<uint8_t> tmp;
	for (int c = 0; c < kComponents; c++)
		for (int l = 0; l < kLevels; l++)
			for (int b = 0; b < kBands; b++)
			{
				const BandGeom &B = G.band[c][l][b];
				if (!B.present)
					continue;
				size_t es = B.kind == kStoreF16 ? 2 : 4;
				tmp.assign(size_t(B.pitch) * B.by32 * 32 * es, 0);
				for (int y = 0; y < B.bh; y++)
					for (int x = 0; x < B.bw; x++)
					{
						float v = coefficients[B.flat_off + long(y) * B.bw + x];
						size_t i = size_t(y) * B.pitch + x;
						if (es == 2)
						{
							uint16_t hv = float_to_half_rne(v);
							memcpy(tmp.data() + 2 * i, &hv, 2);
						}
						else
							memcpy(tmp.data() + 4 * i, &v, 4);
					}
				if (hipMemcpy(S.bands + B.store_off, tmp.data(), tmp.size(), hipMemcpyHostToDevice) != hipSuccess)
					return PYROWAVE_ERROR_GENERIC;
			}
	Plan *P = get_plan(&D, {&S}, true);
	if (!P)
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	CallParams &F = P->call(0);
	set_encode_params(F, S, rate_control->maximum_bitstream_size);
	FrameState *frames = P->stage_frames(D.stream);
	launch_quant_analyze(P->d_bands, P->d_blk_band, P->nblocks, frames, D.stream);
	launch_rc_finalize(frames, 1, D.stream);
	launch_rc_partial(frames, 1, D.stream);
	launch_rc_prepare(frames, P->d_grp_frame, P->ngroups, P->d_frame_group_base, D.stream);
	launch_pack(P->d_bands, P->d_blk_band, P->nblocks, frames, D.stream);
	if (!queue_result_download(S, D.stream, uint32_t(F.target_words) + 64))
		return PYROWAVE_ERROR_OUT_OF_HOST_MEMORY;
	(void)hipEventRecord(S.result_ready, D.stream);
	S.have_result = true;
	S.result_on_host = true;
	return hipGetLastError() == hipSuccess ? PYROWAVE_SUCCESS : PYROWAVE_ERROR_GENERIC;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_encoder_read_rate_control(pyrowave_encoder encoder, uint16_t *words,
                                                                           uint8_t *buckets, uint32_t *totals,
                                                                           uint32_t *prefix, uint32_t *consumed,
                                                                           int32_t *quant)
{
	if (!encoder)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Slot &S = encoder->slot;
	std::lock_guard<std::recursive_mutex> g(S.dev->lock);
	(void)hipSetDevice(S.dev->hip_device);
	if (hipDeviceSynchronize() != hipSuccess)
		return PYROWAVE_ERROR_GENERIC;
	const size_t n32 = size_t(S.geom.count32);
	if (words)
		(void)hipMemcpy(words, S.fs.words, n32 * 16 * 2, hipMemcpyDeviceToHost);
	if (buckets)
		(void)hipMemcpy(buckets, S.fs.bucket, n32 * 16, hipMemcpyDeviceToHost);
	if (totals)
	{
		// The quantizer's copies of the totals, summed.
		const int ncopies = totals_copies(S.geom.count32);
		std::vector<uint32_t> copies(size_t(ncopies) * kFlat);
		(void)hipMemcpy(copies.data(), S.fs.totals, copies.size() * 4, hipMemcpyDeviceToHost);
		for (int x = 0; x < kFlat; x++)
		{
			uint32_t s = 0;
			for (int r = 0; r < ncopies; r++)
				s += copies[size_t(r) * kFlat + x];
			totals[x] = s;
		}
	}
	if (prefix)
		(void)hipMemcpy(prefix, S.fs.prefix, kFlat * 4, hipMemcpyDeviceToHost);
	if (consumed)
		(void)hipMemcpy(consumed, S.fs.rc + kRcConsumed, 4, hipMemcpyDeviceToHost);
	if (quant)
	{
		std::vector<uint8_t> q(n32);
		(void)hipMemcpy(q.data(), S.fs.quant, n32, hipMemcpyDeviceToHost);
		for (size_t i = 0; i < n32; i++)
			quant[i] = q[i];
	}
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result pyrowave_hip_device_query_stage_times(pyrowave_device device, double ms[7],
                                                                          uint64_t frames[7], bool reset)
{
	if (!device || !ms || !frames)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	Device &D = device->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	D.resolve_timings(true);
	for (int s = 0; s < kStageCount; s++)
	{
		ms[s] = D.stage_ms[s];
		frames[s] = D.stage_frames[s];
		if (reset)
		{
			D.stage_ms[s] = 0.0;
			D.stage_frames[s] = 0;
		}
	}
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API void pyrowave_hip_device_set_timing(pyrowave_device device, bool enable)
{
	if (!device)
		return;
	std::lock_guard<std::recursive_mutex> g(device->dev.lock);
	device->dev.timing = enable ? (1u << pw::kStageCount) - 1 : 0u;
}

PYROWAVE_PUBLIC_API void pyrowave_hip_device_set_timing_stages(pyrowave_device device, uint32_t stage_mask)
{
	if (!device)
		return;
	std::lock_guard<std::recursive_mutex> g(device->dev.lock);
	device->dev.timing = stage_mask & ((1u << pw::kStageCount) - 1);
}
} // extern "C"
