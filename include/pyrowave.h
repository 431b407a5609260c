/*
 * pyrowave.h -- C API of the MI355X PyroWave codec.
 *
 * ABI-compatible with the reference's pyrowave.h, API version 0.4.0
 * (reference pyrowave.h:21-23): the same 34 entry points
 * (pyrowave-shared.def:3-37), enums, struct layouts and error codes.  The
 * implementation runs the encode/decode hot path as HIP kernels on an AMD
 * Instinct MI355X (gfx950); there is no Vulkan underneath.
 *
 * Differences a caller can observe, all on the Vulkan-interop side:
 *  - Entry points that take or return Vulkan objects (external images, sync
 *    objects, borrowed VkDevice / VkCommandBuffer) return
 *    PYROWAVE_ERROR_NOT_IMPLEMENTED (or false / NULL where the reference
 *    signature has no result code).  The HIP-native entry points in
 *    pyrowave_hip.h take device pointers and a hipStream_t instead.
 *  - The reference header refuses to compile unless the Vulkan headers were
 *    included first (reference pyrowave.h:7-9).  This header uses the real
 *    Vulkan types when they are present and otherwise declares ABI-identical
 *    stand-ins (dispatchable and non-dispatchable handles are pointers on
 *    64-bit targets; enums are 32-bit).
 */
#ifndef PYROWAVE_H_
#define PYROWAVE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h>
#endif

#if !defined(VULKAN_CORE_H_)
/* Minimal, layout-identical declarations of the Vulkan types this API names. */
#define VK_UUID_SIZE 16u
#define VK_LUID_SIZE 8u
typedef struct VkInstance_T *VkInstance;
typedef struct VkPhysicalDevice_T *VkPhysicalDevice;
typedef struct VkDevice_T *VkDevice;
typedef struct VkQueue_T *VkQueue;
typedef struct VkCommandBuffer_T *VkCommandBuffer;
typedef struct VkImage_T *VkImage;
typedef struct VkSemaphore_T *VkSemaphore;
typedef uint32_t VkFlags;
typedef VkFlags VkSemaphoreImportFlags;
typedef void (*PFN_vkVoidFunction)(void);
typedef PFN_vkVoidFunction (*PFN_vkGetInstanceProcAddr)(VkInstance instance, const char *pName);
typedef struct VkInstanceCreateInfo VkInstanceCreateInfo;
typedef struct VkDeviceCreateInfo VkDeviceCreateInfo;
typedef struct VkImageCreateInfo VkImageCreateInfo;
typedef enum VkQueueFlagBits
{
	VK_QUEUE_GRAPHICS_BIT = 0x00000001,
	VK_QUEUE_COMPUTE_BIT = 0x00000002,
	VK_QUEUE_TRANSFER_BIT = 0x00000004,
	VK_QUEUE_FLAG_BITS_MAX_ENUM = 0x7FFFFFFF
} VkQueueFlagBits;
typedef enum VkFormat { VK_FORMAT_UNDEFINED = 0, VK_FORMAT_R8_UNORM = 9, VK_FORMAT_MAX_ENUM = 0x7FFFFFFF } VkFormat;
typedef enum VkImageLayout { VK_IMAGE_LAYOUT_UNDEFINED = 0, VK_IMAGE_LAYOUT_GENERAL = 1, VK_IMAGE_LAYOUT_MAX_ENUM = 0x7FFFFFFF } VkImageLayout;
typedef enum VkImageAspectFlagBits { VK_IMAGE_ASPECT_COLOR_BIT = 0x00000001, VK_IMAGE_ASPECT_FLAG_BITS_MAX_ENUM = 0x7FFFFFFF } VkImageAspectFlagBits;
typedef enum VkComponentSwizzle { VK_COMPONENT_SWIZZLE_IDENTITY = 0, VK_COMPONENT_SWIZZLE_MAX_ENUM = 0x7FFFFFFF } VkComponentSwizzle;
typedef enum VkImageUsageFlagBits { VK_IMAGE_USAGE_SAMPLED_BIT = 0x00000004, VK_IMAGE_USAGE_STORAGE_BIT = 0x00000008, VK_IMAGE_USAGE_FLAG_BITS_MAX_ENUM = 0x7FFFFFFF } VkImageUsageFlagBits;
typedef enum VkExternalSemaphoreHandleTypeFlagBits { VK_EXTERNAL_SEMAPHORE_HANDLE_TYPE_FLAG_BITS_MAX_ENUM = 0x7FFFFFFF } VkExternalSemaphoreHandleTypeFlagBits;
typedef enum VkExternalMemoryHandleTypeFlagBits { VK_EXTERNAL_MEMORY_HANDLE_TYPE_FLAG_BITS_MAX_ENUM = 0x7FFFFFFF } VkExternalMemoryHandleTypeFlagBits;
typedef enum VkSemaphoreType { VK_SEMAPHORE_TYPE_BINARY = 0, VK_SEMAPHORE_TYPE_TIMELINE = 1, VK_SEMAPHORE_TYPE_MAX_ENUM = 0x7FFFFFFF } VkSemaphoreType;
#endif

/* API and ABI are not stable until the major version reaches 1. */
#define PYROWAVE_API_VERSION_MAJOR 0
#define PYROWAVE_API_VERSION_MINOR 4
#define PYROWAVE_API_VERSION_PATCH 0

#if !defined(PYROWAVE_PUBLIC_API)
#if defined(PYROWAVE_EXPORT_SYMBOLS) && defined(__GNUC__)
#define PYROWAVE_PUBLIC_API __attribute__((visibility("default")))
#else
#define PYROWAVE_PUBLIC_API
#endif
#endif

typedef void (*pyrowave_message_cb)(void *userdata, const char *msg);

typedef enum pyrowave_result
{
	PYROWAVE_SUCCESS = 0,
	PYROWAVE_TIMEOUT = 1,
	PYROWAVE_ERROR_GENERIC = -1,
	PYROWAVE_ERROR_INVALID_ARGUMENT = -2,
	PYROWAVE_ERROR_OUT_OF_HOST_MEMORY = -3,
	PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY = -4,
	PYROWAVE_ERROR_NO_VULKAN = -5,
	PYROWAVE_ERROR_NOT_IMPLEMENTED = -6,
	PYROWAVE_ERROR_UNSUPPORTED_EXTERNAL_HANDLE = -7,
	PYROWAVE_ERROR_FAILED_EXTERNAL_HANDLE = -8,
	PYROWAVE_ERROR_INT_MAX = 0x7fffffff
} pyrowave_result;

typedef enum pyrowave_chroma_subsampling
{
	PYROWAVE_CHROMA_SUBSAMPLING_420 = 0,
	PYROWAVE_CHROMA_SUBSAMPLING_444 = 1,
	PYROWAVE_CHROMA_SUBSAMPLING_INT_MAX = 0x7fffffff
} pyrowave_chroma_subsampling;

typedef struct pyrowave_encoder_opaque *pyrowave_encoder;
typedef struct pyrowave_decoder_opaque *pyrowave_decoder;
typedef struct pyrowave_device_opaque *pyrowave_device;
typedef struct pyrowave_sync_object_opaque *pyrowave_sync_object;
typedef struct pyrowave_image_opaque *pyrowave_image;

/* Stable: reports the API version the library implements. */
PYROWAVE_PUBLIC_API void pyrowave_get_api_version(uint32_t *major, uint32_t *minor, uint32_t *patch);

/* ---- Device ---- */

/* Opens the first visible HIP device.  Returns PYROWAVE_ERROR_NO_VULKAN
 * (the reference's "no usable device" code) when there is none. */
PYROWAVE_PUBLIC_API pyrowave_result pyrowave_create_default_device(pyrowave_device *device);

typedef struct pyrowave_device_create_queue_info
{
	VkQueue queue;
	uint32_t familyIndex;
	uint32_t index;
} pyrowave_device_create_queue_info;

typedef void (*pyrowave_queue_lock_cb)(void *userdata);

typedef struct pyrowave_device_create_info
{
	PFN_vkGetInstanceProcAddr GetInstanceProcAddr;
	VkInstance instance;
	VkPhysicalDevice physical_device;
	VkDevice device;
	const VkInstanceCreateInfo *instance_create_info;
	const VkDeviceCreateInfo *device_create_info;
	pyrowave_device_create_queue_info *queue_info;
	uint32_t queue_info_count;
	pyrowave_queue_lock_cb queue_lock_callback;
	pyrowave_queue_lock_cb queue_unlock_callback;
	void *userdata;
} pyrowave_device_create_info;

typedef struct pyrowave_uuid
{
	uint8_t uuid[VK_UUID_SIZE];
} pyrowave_uuid;

typedef struct pyrowave_luid
{
	uint8_t luid[VK_LUID_SIZE];
} pyrowave_luid;

/* Borrowing a caller's VkDevice: not available (PYROWAVE_ERROR_NOT_IMPLEMENTED). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_create_device(const pyrowave_device_create_info *info, pyrowave_device *device);

/* Device selection by PCI vendor/device id (vid 0x1002 for AMD; 0 = any).
 * UUID / LUID filters are Vulkan identities and are ignored when NULL;
 * non-NULL filters select nothing and yield PYROWAVE_ERROR_NO_VULKAN. */
PYROWAVE_PUBLIC_API pyrowave_result pyrowave_create_device_by_compat(
	uint32_t vid, uint32_t pid,
	const pyrowave_uuid *device_uuid,
	const pyrowave_uuid *driver_uuid,
	const pyrowave_luid *device_luid,
	pyrowave_device *device);

/* Per-stage GPU time averaged over the frames since the last reset, as
 * "<stage>: <ms> ms per frame\n" strings (stages: DWT, Quant, Analyze,
 * Resolve, Packing, Dequant, iDWT). */
PYROWAVE_PUBLIC_API void
pyrowave_device_report_performance_stats(pyrowave_device device, pyrowave_message_cb cb, void *userdata, bool reset);

/* Writes NULL handles: there is no Vulkan device. */
PYROWAVE_PUBLIC_API void pyrowave_device_get_vk_device_handles(
	pyrowave_device device,
	VkInstance *vk_instance, VkPhysicalDevice *vk_physical_device,
	VkDevice *vk_device);

/* Borrowed command buffers are a Vulkan concept; the call is accepted and has
 * no effect. */
PYROWAVE_PUBLIC_API void
pyrowave_device_set_command_buffer(pyrowave_device device, VkCommandBuffer cmd);

/* VK_QUEUE_GRAPHICS_BIT or VK_QUEUE_COMPUTE_BIT; anything else is
 * PYROWAVE_ERROR_INVALID_ARGUMENT.  Both map to the device's HIP stream. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_device_set_queue_type(pyrowave_device device, VkQueueFlagBits queue_flags);

/* Vulkan external memory / semaphore interop: always false. */
PYROWAVE_PUBLIC_API bool
pyrowave_device_confirm_interop_support(pyrowave_device device);

/* All encoders and decoders must be destroyed before the device. */
PYROWAVE_PUBLIC_API void pyrowave_device_destroy(pyrowave_device device);

/* ---- External sync objects (Vulkan semaphores): not implemented ---- */
typedef uintptr_t pyrowave_os_handle;

typedef struct pyrowave_sync_object_create_info
{
	pyrowave_device device;
	pyrowave_os_handle external_handle;
	VkExternalSemaphoreHandleTypeFlagBits handle_type;
	VkSemaphoreType semaphore_type;
	VkSemaphoreImportFlags import_flags;
} pyrowave_sync_object_create_info;

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_sync_object_create(const pyrowave_sync_object_create_info *info, pyrowave_sync_object *sync);
PYROWAVE_PUBLIC_API VkSemaphore
pyrowave_sync_object_get_semaphore(pyrowave_sync_object sync);
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_sync_object_export_handle(pyrowave_sync_object sync, pyrowave_os_handle *handle);
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_sync_object_cpu_wait(pyrowave_sync_object sync, uint64_t value, uint64_t timeout);
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_sync_object_cpu_signal(pyrowave_sync_object sync, uint64_t value);
PYROWAVE_PUBLIC_API void pyrowave_sync_object_destroy(pyrowave_sync_object sync);

/* ---- External images (Vulkan): not implemented ---- */
typedef struct pyrowave_image_view
{
	VkImage image;
	uint32_t width;
	uint32_t height;
	VkFormat image_format;
	VkFormat view_format;
	uint32_t mip_level;
	uint32_t layer;
	VkImageAspectFlagBits aspect;
	VkComponentSwizzle swizzle;
	VkImageLayout layout;
} pyrowave_image_view;

typedef struct pyrowave_image_create_info
{
	pyrowave_device device;
	pyrowave_os_handle external_handle;
	VkExternalMemoryHandleTypeFlagBits handle_type;
	const VkImageCreateInfo *image_create_info;
} pyrowave_image_create_info;

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_image_create(const pyrowave_image_create_info *info, pyrowave_image *image);
PYROWAVE_PUBLIC_API VkImage
pyrowave_image_get_handle(pyrowave_image image);
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_image_get_image_view(pyrowave_image image, VkImageAspectFlagBits aspect,
                              VkImageUsageFlagBits usage, pyrowave_image_view *view);
PYROWAVE_PUBLIC_API void pyrowave_image_destroy(pyrowave_image image);

/* ---- Encoder ---- */
typedef struct pyrowave_encoder_create_info
{
	pyrowave_device device;
	/* For 4:2:0 both must be even.  1..16384. */
	int width;
	int height;
	pyrowave_chroma_subsampling chroma;
} pyrowave_encoder_create_info;

typedef struct pyrowave_packet
{
	size_t offset;
	size_t size;
} pyrowave_packet;

typedef struct pyrowave_sync_point
{
	VkSemaphore semaphore;
	uint64_t value;
} pyrowave_sync_point;

typedef struct pyrowave_gpu_external_reference
{
	pyrowave_image image;
	uint32_t queue_family_index;
} pyrowave_gpu_external_reference;

typedef struct pyrowave_gpu_buffers
{
	pyrowave_image_view planes[3];
} pyrowave_gpu_buffers;

typedef struct pyrowave_gpu_sync_operation
{
	const pyrowave_gpu_external_reference *images;
	size_t num_images;
	pyrowave_sync_point sync;
} pyrowave_gpu_sync_operation;

typedef enum pyrowave_cpu_buffer_format
{
	PYROWAVE_CPU_BUFFER_FORMAT_NV12 = 0,    /* Y 8bpp, then interleaved CbCr 16bpp.  Encode only. */
	PYROWAVE_CPU_BUFFER_FORMAT_YUV420P = 1, /* three planes, chroma at half size */
	PYROWAVE_CPU_BUFFER_FORMAT_YUV444P = 2, /* three planes, full size */
	PYROWAVE_CPU_BUFFER_FORMAT_INT_MAX = 0x7fffffff
} pyrowave_cpu_buffer_format;

typedef struct pyrowave_cpu_buffer
{
	/* Written by the decoder, read by the encoder. */
	void *data[3];
	/* At least the plane width times the texel size. */
	size_t row_stride_in_bytes[3];
	/* At least row_stride times the plane height. */
	size_t plane_size_in_bytes[3];
	/* Luma size; must match the encoder / decoder. */
	int width;
	int height;
	pyrowave_cpu_buffer_format format;
} pyrowave_cpu_buffer;

typedef struct pyrowave_rate_control
{
	/* Hard cap on the packetized size of one frame, in bytes. */
	size_t maximum_bitstream_size;
} pyrowave_rate_control;

/* Encoder entry points are not thread-safe. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_encoder_create(const pyrowave_encoder_create_info *info, pyrowave_encoder *encoder);

/* Vulkan-image input: PYROWAVE_ERROR_NOT_IMPLEMENTED (see pyrowave_hip.h). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_encoder_encode_gpu_synchronous(pyrowave_encoder encoder,
                                        const pyrowave_gpu_sync_operation *acquire,
                                        const pyrowave_gpu_sync_operation *release,
                                        const pyrowave_gpu_buffers *buffers,
                                        const pyrowave_rate_control *rate_control);

/* Uploads the planes and queues the encode; returns without waiting for the
 * GPU.  The next encode replaces this frame's result. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_encoder_encode_cpu_synchronous(pyrowave_encoder encoder, const pyrowave_cpu_buffer *buffers,
                                        const pyrowave_rate_control *rate_control);

/* Waits for the last encode, then counts packets of at most packet_boundary
 * bytes (a single block larger than the boundary gets a packet of its own). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_encoder_compute_num_packets(pyrowave_encoder encoder, size_t packet_boundary, size_t *num_packets);

/* Waits for the last encode, then writes the sequence header and the coded
 * blocks in block-index order into bitstream and the packet list. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_encoder_packetize(pyrowave_encoder encoder, pyrowave_packet *packets, size_t packet_boundary,
                           size_t *out_packets, void *bitstream, size_t size);

PYROWAVE_PUBLIC_API void
pyrowave_encoder_destroy(pyrowave_encoder encoder);

/* ---- Decoder ---- */
typedef struct pyrowave_decoder_create_info
{
	pyrowave_device device;
	int width;
	int height;
	pyrowave_chroma_subsampling chroma;
	/* Mobile render-pass path of the reference.  Accepted and ignored: the
	 * compute path is always used on MI355X. */
	bool fragment_path;
} pyrowave_decoder_create_info;

/* Always false on MI355X. */
PYROWAVE_PUBLIC_API bool
pyrowave_decoder_device_prefers_fragment_path(pyrowave_device device);

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_decoder_create(const pyrowave_decoder_create_info *info, pyrowave_decoder *decoder);

/* Drops every queued packet. */
PYROWAVE_PUBLIC_API void pyrowave_decoder_clear(pyrowave_decoder decoder);

/* Parses one packet.  Packets of an older frame are dropped; a packet of a
 * newer frame discards the queued one and starts over. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_decoder_push_packet(pyrowave_decoder decoder, const void *data, size_t size);

/* True once every block the sequence header announced has arrived (or more
 * than half of them, with allow_partial_frame). */
PYROWAVE_PUBLIC_API bool
pyrowave_decoder_decode_is_ready(pyrowave_decoder decoder, bool allow_partial_frame);

/* Vulkan-image output: PYROWAVE_ERROR_NOT_IMPLEMENTED (see pyrowave_hip.h). */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_decoder_decode_gpu_buffer(pyrowave_decoder decoder,
                                   const pyrowave_gpu_sync_operation *acquire,
                                   const pyrowave_gpu_sync_operation *release,
                                   const pyrowave_gpu_buffers *buffers);

/* Decodes whatever has arrived (missing blocks are zero) into the planes. */
PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_decoder_decode_cpu_buffer_synchronous(pyrowave_decoder decoder, const pyrowave_cpu_buffer *buffers);

PYROWAVE_PUBLIC_API void pyrowave_decoder_destroy(pyrowave_decoder decoder);

#ifdef __cplusplus
}
#endif

#endif
