#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own
bitstream layer (metal/pyrowave_bitstream.cpp, compiled by oracle/Makefile
into oracle/_ref/libref_bitstream.so).

The reference holds no golden vectors of its own.  Its Vulkan-free bitstream
layer does build here, so these fixtures record what that code computes for:

  layout_<w>x<h>_<chroma>.npz   block counts, per-band 8x8/32x32 offsets and
                                strides, the 32x32 -> 8x8 mapping, the block
                                count per rate-control subdivision
  packetize.npz                 compute_num_packets and packetize over
                                synthetic meta/bitstream at several boundaries
  parser.npz                    BitstreamParser state after each step of
                                scripted packet sequences (in order, grouped,
                                duplicates, missing blocks, stale and wrapped
                                sequence numbers, malformed packets)

Only the inputs and outputs are committed.  Run from the repository root
after `make -C oracle`:  python3 tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

LAYOUTS = [(16, 16, 0), (34, 30, 0), (34, 30, 1), (33, 31, 1), (128, 128, 0), (200, 100, 1),
           (1280, 720, 0), (1920, 1080, 0), (2560, 1440, 1), (3840, 2160, 0)]
PARSER_LAYOUT = (160, 96, 0)


def ref_layout(w, h, chroma):
    R = O.ref()
    count8 = C.c_int()
    band = np.zeros((3, 5, 4, 4), dtype=np.int32)
    cap = 1 << 16
    mapping = np.zeros((cap, 4), dtype=np.int32)
    n32 = R.ref_layout(w, h, chroma, C.byref(count8), O.ptr(band), O.ptr(mapping), cap)
    assert 0 < n32 <= cap
    return dict(count8=np.int32(count8.value), count32=np.int32(n32), band=band, mapping=mapping[:n32].copy(),
                per_subdivision=np.int32(R.ref_block_count_per_subdivision(n32)))


def synth_stream(count32, seed, zero_frac=0.25, max_words=40):
    """A meta table and a bitstream of synthetic blocks in block order: each
    coded block starts with a block header (sequence 5) and carries random
    words; a fraction of blocks is left uncoded (0 words)."""
    rng = np.random.default_rng(seed)
    words = rng.integers(2, max_words, size=count32).astype(np.uint32)
    words[rng.random(count32) < zero_frac] = 0
    meta = np.zeros((count32, 2), dtype=np.uint32)
    off = 0
    chunks = []
    for b in range(count32):
        n = int(words[b])
        if n == 0:
            continue
        blk = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        blk[0] = (blk[0] & 0xffff) | (n << 16) | (5 << 28)
        blk[1] = (blk[1] & 0xff) | (b << 8)
        meta[b] = (off, n)
        chunks.append(blk)
        off += n
    if meta[0, 1] == 0:
        # packetize reads the sequence from the first meta entry's offset.
        meta[0, 0] = 0
    bits = np.concatenate(chunks) if chunks else np.zeros(2, np.uint32)
    return meta, bits


def ref_packetize(w, h, chroma, meta, bits, boundary, out_size):
    R = O.ref()
    npk = R.ref_compute_num_packets(w, h, chroma, O.ptr(meta), boundary)
    packets = np.zeros((npk + 2, 2), dtype=np.uint64)
    out = np.zeros(out_size, dtype=np.uint8)
    n = R.ref_packetize(w, h, chroma, O.ptr(packets), boundary, O.ptr(out), out_size, O.ptr(meta), O.ptr(bits))
    return int(npk), packets[:n].copy(), out


def make_packetize():
    flat = {}
    cases = []
    for j, ((w, h, ch), seed) in enumerate((((160, 96, 0), 1), ((192, 128, 1), 2), ((34, 30, 0), 3))):
        L = O.layout(w, h, ch)
        meta, bits = synth_stream(L.count32, seed, max_words=24)
        flat[f"l{j}_dims"] = np.array([w, h, ch], dtype=np.int32)
        flat[f"l{j}_meta"] = meta
        flat[f"l{j}_bits"] = bits
        total = 8 + int(meta[:, 1].sum()) * 4
        for boundary in (64, 200, 1000, 1 << 20):
            for out_size in (total, total // 2):
                npk, packets, out = ref_packetize(w, h, ch, meta, bits, boundary, out_size)
                i = len(cases)
                cases.append(i)
                flat[f"c{i}_params"] = np.array([j, boundary, out_size, npk], dtype=np.int64)
                flat[f"c{i}_packets"] = packets
                flat[f"c{i}_out"] = out
    flat["count"] = np.int32(len(cases))
    np.savez_compressed(os.path.join(HERE, "packetize.npz"), **flat)
    return len(cases)


# --- parser scripts -------------------------------------------------------

def block_packet(block, words, seq, payload_words=None, seed=0):
    """`words` (>= 2) u32 words of a coded block; the header's payload_words
    field says `payload_words` (default: words)."""
    rng = np.random.default_rng(seed * 7919 + block)
    blk = rng.integers(0, 1 << 32, size=words, dtype=np.uint64).astype(np.uint32)
    pw = words if payload_words is None else payload_words
    blk[0] = (blk[0] & 0xffff) | ((pw & 0xfff) << 16) | ((seq & 7) << 28)
    blk[1] = (blk[1] & 0xff) | ((block & 0xffffff) << 8)
    return blk.tobytes()


def seq_header(w, h, chroma, seq, total, code=0):
    w0 = ((w - 1) & 0x3fff) | (((h - 1) & 0x3fff) << 14) | ((seq & 7) << 28) | (1 << 31)
    w1 = (total & 0xffffff) | ((code & 3) << 24) | ((chroma & 1) << 26)
    return np.array([w0, w1], dtype=np.uint32).tobytes()


def parser_scripts(w, h, ch, count32):
    rng = np.random.default_rng(11)
    sizes = rng.integers(2, 30, size=count32)

    def blocks(seq, order=None, salt=0):
        order = range(count32) if order is None else order
        return [block_packet(b, int(sizes[b]), seq, seed=salt) for b in order]

    scripts = {}
    full = [seq_header(w, h, ch, 1, count32)] + blocks(1)
    scripts["in_order"] = [("push", p) for p in full]
    grouped, cur = [], b""
    for p in full:
        if len(cur) + len(p) > 300 and cur:
            grouped.append(cur)
            cur = b""
        cur += p
    grouped.append(cur)
    scripts["grouped"] = [("push", p) for p in grouped]
    order = list(rng.permutation(count32)) + list(rng.choice(count32, 10))
    scripts["shuffled_duplicates"] = [("push", seq_header(w, h, ch, 2, count32))] + \
        [("push", p) for p in blocks(2, order)]
    keep = sorted(rng.choice(count32, int(count32 * 0.6), replace=False))
    scripts["sixty_percent"] = [("push", seq_header(w, h, ch, 3, count32))] + [("push", p) for p in blocks(3, keep)]
    keep = sorted(rng.choice(count32, count32 // 2, replace=False))
    scripts["half_exactly"] = [("push", p) for p in blocks(4, keep)]
    scripts["stale_then_new"] = [("push", p) for p in blocks(2, range(0, count32, 2))] + \
        [("push", p) for p in blocks(1, range(1, count32, 2), salt=1)] + \
        [("push", p) for p in blocks(3, range(count32), salt=2)]
    scripts["wrap"] = [("push", p) for p in blocks(6, range(0, 8))] + [("push", p) for p in blocks(7, range(8, 16))] + \
        [("push", p) for p in blocks(0, range(count32), salt=3)] + [("push", p) for p in blocks(7, range(3), salt=4)]
    scripts["seq_header_smaller_total"] = [("push", seq_header(w, h, ch, 5, 12))] + \
        [("push", p) for p in blocks(5, range(12))]
    scripts["no_seq_header"] = [("push", p) for p in blocks(1)]
    scripts["clear_repush"] = [("push", p) for p in full] + [("clear", None)] + [("push", p) for p in full]
    scripts["mark_decoded_then_next"] = [("push", p) for p in full] + [("mark", None)] + \
        [("push", p) for p in [seq_header(w, h, ch, 2, count32)] + blocks(2, salt=5)]
    bad = []
    bad.append(("push", seq_header(w + 1, h, ch, 1, count32)))             # dimension mismatch
    bad.append(("push", seq_header(w, h, 1 - ch, 1, count32)))             # chroma mismatch
    bad.append(("push", seq_header(w, h, ch, 1, count32, code=1)))         # unknown code
    bad.append(("push", block_packet(count32, 4, 1)))                      # block index out of range
    bad.append(("push", block_packet(3, 4, 1)[:12]))                       # truncated payload
    bad.append(("push", block_packet(4, 4, 1) + b"\x00\x00\x00"))           # trailing bytes
    bad.append(("push", b"\x01\x02\x03"))                                  # shorter than a header
    bad.append(("push", b""))                                              # empty
    bad.append(("push", block_packet(5, 2, 1, payload_words=1)))           # header longer than payload
    bad.append(("push", block_packet(5, 2, 1, payload_words=1)))           # same block again (advances 4 B)
    bad.append(("push", block_packet(6, 2, 1, payload_words=0)))           # zero-length block, first time
    bad.append(("push", seq_header(w, h, ch, 1, count32)[:4] + block_packet(7, 3, 1)))
    scripts["malformed"] = bad
    return scripts


def run_ref_parser(w, h, ch, script):
    R = O.ref()
    p = R.ref_parser_create(w, h, ch)
    L = O.layout(w, h, ch)
    steps = []
    for op, data in script:
        rc = -1
        if op == "push":
            buf = np.frombuffer(data, dtype=np.uint8).copy() if data else np.zeros(1, np.uint8)
            rc = R.ref_parser_push(p, O.ptr(buf), len(data))
        elif op == "clear":
            R.ref_parser_clear(p)
        elif op == "mark":
            R.ref_parser_mark_decoded(p)
        steps.append((rc, R.ref_parser_ready(p, 0), R.ref_parser_ready(p, 1), R.ref_parser_decoded_blocks(p),
                      R.ref_parser_total_blocks(p), R.ref_parser_payload_words(p)))
    nwords = R.ref_parser_payload_words(p)
    offsets = np.zeros(L.count32, dtype=np.uint32)
    payload = np.zeros(max(nwords, 1), dtype=np.uint32)
    R.ref_parser_get(p, O.ptr(offsets), O.ptr(payload))
    R.ref_parser_destroy(p)
    return np.array(steps, dtype=np.int64), offsets, payload[:nwords].copy()


def make_parser():
    w, h, ch = PARSER_LAYOUT
    L = O.layout(w, h, ch)
    scripts = parser_scripts(w, h, ch, L.count32)
    flat = {"w": np.int32(w), "h": np.int32(h), "chroma": np.int32(ch), "names": np.array(sorted(scripts))}
    opcode = {"push": 0, "clear": 1, "mark": 2}
    for name in sorted(scripts):
        script = scripts[name]
        steps, offsets, payload = run_ref_parser(w, h, ch, script)
        data = b"".join(d or b"" for _, d in script)
        lens = np.array([len(d or b"") for _, d in script], dtype=np.int64)
        flat[f"{name}_ops"] = np.array([opcode[o] for o, _ in script], dtype=np.int32)
        flat[f"{name}_lens"] = lens
        flat[f"{name}_data"] = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(0, np.uint8)
        flat[f"{name}_steps"] = steps
        flat[f"{name}_offsets"] = offsets
        flat[f"{name}_payload"] = payload
    np.savez_compressed(os.path.join(HERE, "parser.npz"), **flat)
    return len(scripts)


def main():
    for w, h, ch in LAYOUTS:
        r = ref_layout(w, h, ch)
        np.savez_compressed(os.path.join(HERE, f"layout_{w}x{h}_{'444' if ch else '420'}.npz"), **r)
    n_pk = make_packetize()
    n_ps = make_parser()
    print(f"wrote {len(LAYOUTS)} layouts, {n_pk} packetize cases, {n_ps} parser scripts")


if __name__ == "__main__":
    main()
