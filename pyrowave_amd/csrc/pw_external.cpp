// pw_external.cpp -- external memory: zero-copy frame input and output.
//
// The reference takes frames from other APIs by importing their memory as a
// Vulkan image (pyrowave_image_create / pyrowave_image_get_image_view,
// reference pyrowave.h:269-352, pyrowave_c.cpp:524-792) and encodes from, or
// decodes into, views of its planes.  Here the shared memory is a linear
// device buffer handed over as a POSIX file descriptor:
//   - pyrowave_hip_external_memory_import maps an fd exported by another API
//     or process (hipImportExternalMemory + hipExternalMemoryGetMappedBuffer);
//   - pyrowave_hip_external_memory_allocate makes such a buffer here and
//     exports it (HIP virtual memory: hipMemCreate with a POSIX fd handle
//     type, hipMemExportToShareableHandle), for a producer elsewhere to import;
//   - pyrowave_hip_external_memory_frame / _target describe planes inside
//     either kind as an encoder source or a decoder target, after checking
//     that every plane lies inside the buffer.
#include "pw_host.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <new>

struct pyrowave_hip_external_memory_opaque
{
	pyrowave_device device = nullptr;
	void *ptr = nullptr;
	uint64_t size = 0;
	// Imported memory.
	hipExternalMemory_t ext = nullptr;
	int fd = -1; // the duplicate handed to the import, when it is still ours to close
	// Allocated (exportable) memory.
	bool allocated = false;
	hipMemGenericAllocationHandle_t handle{};
};

namespace
{
using Mem = pyrowave_hip_external_memory_opaque;

pyrowave_result check_layout(const Mem &m, const pyrowave_hip_external_image_layout &L, bool for_output)
{
	if (L.width < 1 || L.height < 1 || L.width > 16384 || L.height > 16384)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const bool sub = L.chroma == PYROWAVE_CHROMA_SUBSAMPLING_420;
	if (!sub && L.chroma != PYROWAVE_CHROMA_SUBSAMPLING_444)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (sub && ((L.width & 1) || (L.height & 1)))
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const bool interleaved = L.layout == PYROWAVE_HIP_PLANES_INTERLEAVED_CHROMA;
	if (!interleaved && L.layout != PYROWAVE_HIP_PLANES_THREE)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	// The decoder writes one byte per sample per plane.
	if (interleaved && for_output)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const uint64_t cw = sub ? L.width / 2 : L.width, ch = sub ? L.height / 2 : L.height;
	const int planes = interleaved ? 2 : 3;
	for (int p = 0; p < planes; p++)
	{
		const uint64_t w = p == 0 ? L.width : interleaved ? 2 * cw : cw;
		const uint64_t h = p == 0 ? L.height : ch;
		const uint64_t pitch = L.row_pitch[p], off = L.plane_offset[p];
		if (pitch < w)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
		// Last byte of the plane: off + (h - 1) * pitch + w - 1 < size.
		if (off > m.size || (h - 1) > (m.size - off) / pitch || (h - 1) * pitch + w > m.size - off)
			return PYROWAVE_ERROR_INVALID_ARGUMENT;
	}
	return PYROWAVE_SUCCESS;
}

void release(Mem *m)
{
	if (!m)
		return;
	if (m->device)
		(void)hipSetDevice(m->device->dev.hip_device);
	if (m->allocated)
	{
		if (m->ptr)
		{
			(void)hipMemUnmap(m->ptr, m->size);
			(void)hipMemAddressFree(m->ptr, m->size);
		}
		(void)hipMemRelease(m->handle);
	}
	else
	{
		if (m->ptr)
			(void)hipFree(m->ptr);
		if (m->ext)
			(void)hipDestroyExternalMemory(m->ext);
		// Closed here only while still open: a runtime that owns the
		// imported descriptor closes it on destroy.
		if (m->fd >= 0 && fcntl(m->fd, F_GETFD) != -1)
			(void)close(m->fd);
	}
	if (m->device)
		m->device->dev.live_objects--;
	delete m;
}
} // namespace

extern "C" {

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_import(const pyrowave_hip_external_memory_import_info *info,
                                    pyrowave_hip_external_memory *memory)
{
	if (!info || !memory || !info->device || info->fd < 0 || info->size == 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	if (info->handle_type != PYROWAVE_HIP_EXTERNAL_HANDLE_FD)
		return PYROWAVE_ERROR_UNSUPPORTED_EXTERNAL_HANDLE;
	*memory = nullptr;
	pw::Device &D = info->device->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	// The caller keeps its descriptor; the import gets a duplicate.
	const int dupfd = dup(info->fd);
	if (dupfd < 0)
		return PYROWAVE_ERROR_FAILED_EXTERNAL_HANDLE;
	hipExternalMemoryHandleDesc desc{};
	desc.type = hipExternalMemoryHandleTypeOpaqueFd;
	desc.handle.fd = dupfd;
	desc.size = info->size;
	hipExternalMemory_t ext = nullptr;
	if (hipImportExternalMemory(&ext, &desc) != hipSuccess)
	{
		(void)hipGetLastError();
		(void)close(dupfd);
		return PYROWAVE_ERROR_FAILED_EXTERNAL_HANDLE;
	}
	Mem *m = new (std::nothrow) Mem;
	if (!m)
	{
		(void)hipDestroyExternalMemory(ext);
		return PYROWAVE_ERROR_OUT_OF_HOST_MEMORY;
	}
	m->device = info->device;
	D.live_objects++;
	m->ext = ext;
	m->size = info->size;
	// A successful import may take the descriptor over; close it on destroy
	// only while it is still open here.
	m->fd = fcntl(dupfd, F_GETFD) != -1 ? dupfd : -1;
	hipExternalMemoryBufferDesc bd{};
	bd.offset = 0;
	bd.size = info->size;
	if (hipExternalMemoryGetMappedBuffer(&m->ptr, ext, &bd) != hipSuccess)
	{
		(void)hipGetLastError();
		m->ptr = nullptr;
		release(m);
		return PYROWAVE_ERROR_FAILED_EXTERNAL_HANDLE;
	}
	*memory = m;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_allocate(pyrowave_device device, uint64_t size, pyrowave_hip_external_memory *memory,
                                      int *exported_fd)
{
	if (!device || !memory || !exported_fd || size == 0)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	*memory = nullptr;
	*exported_fd = -1;
	pw::Device &D = device->dev;
	std::lock_guard<std::recursive_mutex> g(D.lock);
	(void)hipSetDevice(D.hip_device);
	hipMemAllocationProp prop{};
	prop.type = hipMemAllocationTypePinned;
	prop.requestedHandleTypes = hipMemHandleTypePosixFileDescriptor;
	prop.location.type = hipMemLocationTypeDevice;
	prop.location.id = D.hip_device;
	size_t gran = 0;
	if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum) != hipSuccess || !gran)
	{
		(void)hipGetLastError();
		return PYROWAVE_ERROR_NOT_IMPLEMENTED;
	}
	const uint64_t padded = (size + gran - 1) / gran * gran;
	Mem *m = new (std::nothrow) Mem;
	if (!m)
		return PYROWAVE_ERROR_OUT_OF_HOST_MEMORY;
	if (hipMemCreate(&m->handle, padded, &prop, 0) != hipSuccess)
	{
		(void)hipGetLastError();
		delete m;
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	}
	m->device = device;
	D.live_objects++;
	m->allocated = true;
	m->size = padded;
	int fd = -1;
	void *va = nullptr;
	hipMemAccessDesc access{};
	access.location = prop.location;
	access.flags = hipMemAccessFlagsProtReadWrite;
	if (hipMemExportToShareableHandle(&fd, m->handle, hipMemHandleTypePosixFileDescriptor, 0) != hipSuccess ||
	    fd < 0)
	{
		(void)hipGetLastError();
		release(m);
		return PYROWAVE_ERROR_FAILED_EXTERNAL_HANDLE;
	}
	if (hipMemAddressReserve(&va, padded, 0, nullptr, 0) != hipSuccess)
	{
		(void)hipGetLastError();
		(void)close(fd);
		release(m);
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	}
	if (hipMemMap(va, padded, 0, m->handle, 0) != hipSuccess)
	{
		(void)hipGetLastError();
		(void)hipMemAddressFree(va, padded);
		(void)close(fd);
		release(m);
		return PYROWAVE_ERROR_OUT_OF_DEVICE_MEMORY;
	}
	m->ptr = va;
	if (hipMemSetAccess(va, padded, &access, 1) != hipSuccess)
	{
		(void)hipGetLastError();
		(void)close(fd);
		release(m);
		return PYROWAVE_ERROR_GENERIC;
	}
	*memory = m;
	*exported_fd = fd;
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API void *pyrowave_hip_external_memory_device_address(pyrowave_hip_external_memory memory)
{
	return memory ? memory->ptr : nullptr;
}

PYROWAVE_PUBLIC_API uint64_t pyrowave_hip_external_memory_size(pyrowave_hip_external_memory memory)
{
	return memory ? memory->size : 0;
}

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_frame(pyrowave_hip_external_memory memory, const pyrowave_hip_external_image_layout *layout,
                                   pyrowave_hip_frame *frame)
{
	if (!memory || !layout || !frame)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const pyrowave_result r = check_layout(*memory, *layout, false);
	if (r != PYROWAVE_SUCCESS)
		return r;
	const char *base = static_cast<const char *>(memory->ptr);
	frame->planes[0] = {base + layout->plane_offset[0], layout->row_pitch[0], 1};
	if (layout->layout == PYROWAVE_HIP_PLANES_INTERLEAVED_CHROMA)
	{
		frame->planes[1] = {base + layout->plane_offset[1], layout->row_pitch[1], 2};
		frame->planes[2] = {base + layout->plane_offset[1] + 1, layout->row_pitch[1], 2};
	}
	else
	{
		frame->planes[1] = {base + layout->plane_offset[1], layout->row_pitch[1], 1};
		frame->planes[2] = {base + layout->plane_offset[2], layout->row_pitch[2], 1};
	}
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API pyrowave_result
pyrowave_hip_external_memory_target(pyrowave_hip_external_memory memory,
                                    const pyrowave_hip_external_image_layout *layout,
                                    pyrowave_hip_decode_target *target)
{
	if (!memory || !layout || !target)
		return PYROWAVE_ERROR_INVALID_ARGUMENT;
	const pyrowave_result r = check_layout(*memory, *layout, true);
	if (r != PYROWAVE_SUCCESS)
		return r;
	char *base = static_cast<char *>(memory->ptr);
	for (int p = 0; p < 3; p++)
		target->planes[p] = {base + layout->plane_offset[p], layout->row_pitch[p]};
	return PYROWAVE_SUCCESS;
}

PYROWAVE_PUBLIC_API void pyrowave_hip_external_memory_destroy(pyrowave_hip_external_memory memory)
{
	if (!memory)
		return;
	std::lock_guard<std::recursive_mutex> g(memory->device->dev.lock);
	release(memory);
}

} // extern "C"
