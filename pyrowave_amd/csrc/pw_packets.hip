// pw_packets.hip -- packetize and packet parsing on the device.
//
// k_packetize: one workgroup per frame.  The coded blocks already sit in
// block-index order, contiguous, right after two reserved words, so the
// packetized stream is the sequence header followed by the blocks; what is
// left is the packet split (reference pyrowave_encoder.cpp:596-623,
// 1048-1115): a block that would take the open packet past the boundary
// starts a new packet.  The block positions are staged in LDS in tiles and
// one wave walks them, 64 blocks per step: a ballot finds the first block
// that overflows the open packet, and the walk repeats inside the step from
// there, so a step costs one ballot per packet it closes.
//
// k_parse: one workgroup per frame, one thread per packet.  Each thread walks
// its packet's units (sequence headers and blocks), validates them as
// pyrowave_decoder.cpp:126-262 does, and records each block's word offset
// with an atomic minimum, so the earliest copy of a block in push order wins,
// as in the reference.  Packets must all carry one sequence number.
#include "pw_common.hpp"

#include <mutex>

namespace pw
{
// Parse scratch of the allocation-free decode
// (pyrowave_hip_decode_packets_noalloc): one frame's block offset table and
// four info words.  Each GPU holds its own copy until the process ends; its
// device address comes from hipGetSymbolAddress (the array's name is a
// host-side stand-in, not a device address).
static __device__ __align__(256) uint32_t scratch[kNoAllocMaxBlocks + 4];

namespace
{
std::mutex g_scratch_lock;
uint32_t *g_scratch_ptr[kMaxDevices] = {};
} // namespace

uint32_t *noalloc_scratch_init(int hip_device)
{
	if (hip_device < 0 || hip_device >= kMaxDevices)
		return nullptr;
	std::lock_guard<std::mutex> g(g_scratch_lock);
	if (!g_scratch_ptr[hip_device])
	{
		int current = -1;
		if (hipGetDevice(&current) != hipSuccess || current != hip_device)
			return nullptr;
		void *ptr = nullptr;
		if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
			return nullptr;
		g_scratch_ptr[hip_device] = static_cast<uint32_t *>(ptr);
	}
	return g_scratch_ptr[hip_device];
}

uint32_t *noalloc_scratch(int hip_device)
{
	if (hip_device < 0 || hip_device >= kMaxDevices)
		return nullptr;
	std::lock_guard<std::mutex> g(g_scratch_lock);
	return g_scratch_ptr[hip_device];
}

namespace
{
constexpr int kPkThreads = 256;
constexpr int kPkTile = 2048;

__global__ __launch_bounds__(kPkThreads) void k_packetize(PacketizeBlock blk)
{
	__shared__ uint32_t s_pos[kPkTile];
	__shared__ uint32_t s_end[kPkTile];
	const PacketizeJob &J = blk.j[blockIdx.x];
	const int tid = threadIdx.x, lane = tid & 63;
	const uint32_t B = J.boundary;

	// Walk state, held by wave 0 (uniform across its lanes).
	uint32_t cur = 0, n = 0, nonzero = 0, stream_end = 8;
	for (uint32_t t0 = 0; t0 < J.n32; t0 += kPkTile)
	{
		const uint32_t cnt = J.n32 - t0 < (uint32_t)kPkTile ? J.n32 - t0 : (uint32_t)kPkTile;
		for (uint32_t i = tid; i < cnt; i += kPkThreads)
		{
			const uint32_t off = J.meta[2 * (t0 + i)], words = J.meta[2 * (t0 + i) + 1];
			s_pos[i] = 8u + 4u * off;
			s_end[i] = words ? 8u + 4u * (off + words) : 0u; // 0: block not coded
		}
		__syncthreads();
		if (tid < 64)
		{
			for (uint32_t c = 0; c < cnt; c += 64)
			{
				const uint32_t i = c + lane;
				const uint32_t e = i < cnt ? s_end[i] : 0u;
				const uint32_t p = i < cnt ? s_pos[i] : 0u;
				const bool nz = e != 0;
				const uint64_t nzmask = __ballot(nz);
				if (nzmask)
				{
					nonzero += (uint32_t)__popcll(nzmask);
					stream_end = __shfl(e, 63 - __clzll(nzmask));
				}
				uint64_t m = __ballot(nz && e - cur > B);
				while (m)
				{
					const int L = __ffsll((unsigned long long)m) - 1;
					const uint32_t start = __shfl(p, L);
					if (lane == 0)
					{
						J.table[2 * n] = cur;
						J.table[2 * n + 1] = start - cur;
					}
					n++;
					cur = start;
					m = __ballot(nz && lane > L && e - cur > B);
				}
			}
		}
		__syncthreads();
	}
	if (tid == 0)
	{
		J.table[2 * n] = cur;
		J.table[2 * n + 1] = stream_end - cur;
		J.info[0] = stream_end;
		J.info[1] = n + 1;
		J.stream[0] = J.h0;
		J.stream[1] = (nonzero & 0xffffffu) | J.h1;
	}
}

__global__ __launch_bounds__(kPkThreads) void k_parse(ParseBlock blk)
{
	__shared__ uint32_t s_err, s_seq, s_total, s_present;
	const ParseJob &J = blk.j[blockIdx.x];
	const int tid = threadIdx.x;
	if (tid == 0)
	{
		s_err = 0;
		s_seq = UINT32_MAX;
		s_total = J.n32;
		s_present = 0;
	}
	for (uint32_t i = tid; i < J.n32; i += kPkThreads)
		J.offsets[i] = UINT32_MAX;
	__threadfence();
	__syncthreads();

	uint32_t count = *J.count;
	uint32_t err = 0;
	if (count > J.max_packets)
	{
		err |= kParseBadTable;
		count = J.max_packets;
	}
	for (uint32_t pk = tid; pk < count; pk += kPkThreads)
	{
		const uint32_t off = J.packets[2 * pk], size = J.packets[2 * pk + 1];
		if ((off & 3u) || off > J.data_bytes || size > J.data_bytes - off)
		{
			err |= kParseBadTable;
			continue;
		}
		uint32_t u = 0, bad = 0;
		while (size - u >= 8u)
		{
			const uint32_t w0 = J.data[(off + u) >> 2], w1 = J.data[((off + u) >> 2) + 1];
			const uint32_t seq = (w0 >> 28) & 7u;
			const uint32_t prev = atomicCAS(&s_seq, UINT32_MAX, seq);
			if (prev != UINT32_MAX && prev != seq)
			{
				bad = kParseSequence;
				break;
			}
			if (w0 >> 31)
			{
				// Sequence header (bitstream/bitstream.md:204-315).
				if (((w1 >> 26) & 1u) != J.chroma || ((w1 >> 24) & 3u) != 0 || (w0 & 0x0fffffffu) != J.h0_size)
				{
					bad = kParseBadHeader;
					break;
				}
				atomicExch(&s_total, w1 & 0xffffffu);
				u += 8u;
				continue;
			}
			const uint32_t psize = ((w0 >> 16) & 0xfffu) * 4u;
			const uint32_t block = w1 >> 8;
			if (psize > size - u || block >= J.n32)
			{
				bad = kParseBadBlock;
				break;
			}
			// Counted and placed before the length check, as in
			// pyrowave_decoder.cpp:126-150: a block shorter than its header is
			// present (its first copy wins) and ends the packet with an error.
			atomicMin(&J.offsets[block], (off + u) >> 2);
			if (psize < 8u)
			{
				bad = kParseBadBlock;
				break;
			}
			u += psize;
		}
		if (!bad && u != size)
			bad = kParseTrailing;
		err |= bad;
	}
	if (err)
		atomicOr(&s_err, err);
	__threadfence();
	__syncthreads();

	uint32_t present = 0;
	for (uint32_t i = tid; i < J.n32; i += kPkThreads)
		present += J.offsets[i] != UINT32_MAX ? 1u : 0u;
	atomicAdd(&s_present, present);
	__syncthreads();
	if (tid == 0)
	{
		J.info[0] = s_present;
		J.info[1] = s_total;
		J.info[2] = s_err;
		J.info[3] = s_seq;
	}
}
} // namespace

void launch_packetize(const PacketizeJob *jobs, int count, hipStream_t stream)
{
	for (int first = 0; first < count; first += kPacketizeJobs)
	{
		PacketizeBlock blk;
		const int n = count - first < kPacketizeJobs ? count - first : kPacketizeJobs;
		for (int i = 0; i < n; i++)
			blk.j[i] = jobs[first + i];
		hipLaunchKernelGGL(k_packetize, dim3(n), dim3(kPkThreads), 0, stream, blk);
	}
}

void launch_parse(const ParseJob *jobs, int count, hipStream_t stream)
{
	for (int first = 0; first < count; first += kParseJobs)
	{
		ParseBlock blk;
		const int n = count - first < kParseJobs ? count - first : kParseJobs;
		for (int i = 0; i < n; i++)
			blk.j[i] = jobs[first + i];
		hipLaunchKernelGGL(k_parse, dim3(n), dim3(kPkThreads), 0, stream, blk);
	}
}
} // namespace pw
