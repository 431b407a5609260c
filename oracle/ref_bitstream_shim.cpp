// ref_bitstream_shim.cpp -- extern "C" wrapper over the reference's Vulkan-free
// bitstream layer (metal/pyrowave_bitstream.{hpp,cpp} in Themaister/pyrowave).
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference source where it lies under /root/reference into
// oracle/_ref/libref_bitstream.so.  tests/golden/make_golden.py uses it to write
// the golden fixtures that pin the oracle's layout, packetize and parser
// restatements.  Nothing here is part of the product library.
#include "pyrowave_bitstream.hpp"

#include <stdint.h>
#include <string.h>

using namespace PyroWave;

extern "C" {

// Layout: returns count32; fills count8, per-band info (c,l,b -> 4 ints:
// off8, stride8, off32, stride32) and per-32x32 mapping (4 ints each).
int ref_layout(int width, int height, int chroma, int *count8, int *band_info, int *mapping, int mapping_cap)
{
	BlockLayout L;
	if (!L.init(width, height, chroma ? ChromaSubsampling::Chroma444 : ChromaSubsampling::Chroma420))
		return -1;
	*count8 = L.block_count_8x8;
	for (int c = 0; c < NumComponents; c++)
		for (int l = 0; l < DecompositionLevels; l++)
			for (int b = 0; b < NumFrequencyBandsPerLevel; b++)
			{
				int *o = band_info + 4 * ((c * DecompositionLevels + l) * NumFrequencyBandsPerLevel + b);
				const auto &m = L.block_meta[c][l][b];
				o[0] = m.block_offset_8x8;
				o[1] = m.block_stride_8x8;
				o[2] = m.block_offset_32x32;
				o[3] = m.block_stride_32x32;
			}
	int n = int(L.block_32x32_to_8x8_mapping.size());
	for (int i = 0; i < n && i < mapping_cap; i++)
	{
		const auto &m = L.block_32x32_to_8x8_mapping[i];
		mapping[4 * i + 0] = m.block_offset_8x8;
		mapping[4 * i + 1] = m.block_stride_8x8;
		mapping[4 * i + 2] = m.block_width_8x8;
		mapping[4 * i + 3] = m.block_height_8x8;
	}
	return L.block_count_32x32;
}

int ref_block_count_per_subdivision(int num_blocks)
{
	return compute_block_count_per_subdivision(num_blocks);
}

size_t ref_compute_num_packets(int width, int height, int chroma, const uint32_t *meta, size_t boundary)
{
	BlockLayout L;
	L.init(width, height, chroma ? ChromaSubsampling::Chroma444 : ChromaSubsampling::Chroma420);
	return compute_num_packets(L, meta, boundary);
}

size_t ref_packetize(int width, int height, int chroma, size_t *packets, size_t boundary,
                     uint8_t *out, size_t size, const uint32_t *meta, const uint32_t *bitstream)
{
	BlockLayout L;
	L.init(width, height, chroma ? ChromaSubsampling::Chroma444 : ChromaSubsampling::Chroma420);
	static_assert(sizeof(Packet) == 2 * sizeof(size_t), "Packet layout");
	return packetize(L, reinterpret_cast<Packet *>(packets), boundary, out, size, meta, bitstream);
}

struct RefParser
{
	BlockLayout layout;
	BitstreamParser parser;
};

void *ref_parser_create(int width, int height, int chroma)
{
	auto *p = new RefParser;
	if (!p->layout.init(width, height, chroma ? ChromaSubsampling::Chroma444 : ChromaSubsampling::Chroma420))
	{
		delete p;
		return nullptr;
	}
	p->parser.init(&p->layout);
	return p;
}

void ref_parser_destroy(void *p) { delete static_cast<RefParser *>(p); }
int ref_parser_push(void *p, const void *data, size_t size) { return static_cast<RefParser *>(p)->parser.push_packet(data, size) ? 1 : 0; }
int ref_parser_ready(void *p, int partial) { return static_cast<RefParser *>(p)->parser.decode_is_ready(partial != 0) ? 1 : 0; }
void ref_parser_clear(void *p) { static_cast<RefParser *>(p)->parser.clear(); }
void ref_parser_mark_decoded(void *p) { static_cast<RefParser *>(p)->parser.mark_frame_decoded(); }
int ref_parser_decoded_blocks(void *p) { return static_cast<RefParser *>(p)->parser.get_decoded_blocks(); }
int ref_parser_total_blocks(void *p) { return static_cast<RefParser *>(p)->parser.get_total_blocks_in_sequence(); }
long ref_parser_payload_words(void *p) { return long(static_cast<RefParser *>(p)->parser.payload().size()); }

void ref_parser_get(void *p, uint32_t *offsets, uint32_t *payload)
{
	auto *r = static_cast<RefParser *>(p);
	const auto &o = r->parser.dequant_offsets();
	memcpy(offsets, o.data(), o.size() * sizeof(uint32_t));
	const auto &pl = r->parser.payload();
	if (!pl.empty())
		memcpy(payload, pl.data(), pl.size() * sizeof(uint32_t));
}
}
