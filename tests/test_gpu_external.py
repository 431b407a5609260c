"""Zero-copy frames in external memory (MI355X).

The reference imports external memory as a Vulkan image and encodes from /
decodes into views of its planes (pyrowave.h:269-352).  Here a buffer made
exportable by pyrowave_hip_external_memory_allocate is exported as a POSIX
fd, imported back with pyrowave_hip_external_memory_import (a second mapping
of the same memory, as another process or API would see it), and:

- a frame written through the exporting mapping encodes from the imported
  mapping to the same bitstream as from ordinary device memory (planar I420
  with padded pitches, and NV12 with interleaved chroma);
- a decode into the imported mapping reads back, through the exporting
  mapping, equal to a decode into ordinary device memory;
- layouts whose planes leave the buffer are rejected.
"""
import os

import numpy as np
import pytest
import torch

import pyrowave_amd as pw
from pyrowave_amd import synth

pytestmark = pytest.mark.gpu

W, H, BUDGET = 256, 144, 12000


class _DeviceBytes:
    """A raw device range as a torch uint8 tensor (CUDA array interface)."""

    def __init__(self, ptr, nbytes):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False),
                                         "version": 2}


def _view(ptr, nbytes):
    try:
        return torch.as_tensor(_DeviceBytes(ptr, nbytes), device="cuda")
    except (TypeError, RuntimeError) as e:
        pytest.skip(f"torch cannot wrap a raw device range here: {e}")


def _shared(device, nbytes):
    """(exporting memory, imported memory) over one allocation, or skip when
    the runtime cannot export or import."""
    try:
        a = pw.ExternalMemory.allocate(device, nbytes)
    except pw.PyroWaveError as e:
        pytest.skip(f"exportable allocation not available: {e}")
    try:
        b = pw.ExternalMemory.import_fd(device, a.fd, a.size)
    except pw.PyroWaveError as e:
        a.close()
        os.close(a.fd)
        pytest.skip(f"external memory import not available: {e}")
    os.close(a.fd)  # the import holds its own duplicate
    return a, b


def _open_fds():
    return set(int(x) for x in os.listdir("/proc/self/fd"))


def test_import_descriptor_ownership(device):
    """Who closes the duplicate descriptor an import hands to the runtime:
    across import and release, the process's open descriptors come back to
    the set before the import (nothing leaks, nothing else is closed), and a
    descriptor opened while the import is alive (it may take the number the
    runtime gave up) survives the release."""
    try:
        a = pw.ExternalMemory.allocate(device, 1 << 20)
    except pw.PyroWaveError as e:
        pytest.skip(f"exportable allocation not available: {e}")
    try:
        # A first import and release: descriptors the runtime opens once and
        # keeps (a render node) are open before the measured one.
        try:
            pw.ExternalMemory.import_fd(device, a.fd, a.size).close()
        except pw.PyroWaveError as e:
            pytest.skip(f"external memory import not available: {e}")
        torch.cuda.synchronize()
        before = _open_fds()
        b = pw.ExternalMemory.import_fd(device, a.fd, a.size)
        during = _open_fds()
        sentinel = os.open("/dev/null", os.O_RDONLY)
        try:
            b.close()
            torch.cuda.synchronize()
            os.fstat(sentinel)  # raises if the release closed it
            after = _open_fds() - {sentinel}
        finally:
            os.close(sentinel)
        assert after == before, (sorted(before), sorted(during), sorted(after))
        os.fstat(a.fd)  # the caller's own descriptor stays open
    finally:
        a.close()
        os.close(a.fd)


def _planar_layout():
    pitch = [320, 192, 192]
    off = [0, 320 * H, 320 * H + 192 * (H // 2)]
    return pw.external_layout(W, H, pw.CHROMA_420, off, pitch), off, pitch


def _encode(device, frame):
    enc = pw.Encoder(device, W, H, pw.CHROMA_420, precision=pw.PRECISION_FP32)
    try:
        pw.encode_batch([enc], [frame], [BUDGET])
        torch.cuda.synchronize()
        out, pk = enc.packetize(1 << 30, BUDGET + 64)
        return out[:sum(s for _, s in pk)].copy()
    finally:
        enc.close()


def _frame_of(tensors):
    return pw.hip_frame([(t.data_ptr(), t.stride(0), 1) for t in tensors])


def test_encode_and_decode_through_imported_memory(device):
    planes = synth.SceneGenerator(W, H, True, seed=21, frames=1, pan=3, redraw=0.05).frame(0)
    layout, off, pitch = _planar_layout()
    a, b = _shared(device, off[2] + pitch[2] * (H // 2))
    try:
        va = _view(a.address, a.size)
        for p, (o, pt) in enumerate(zip(off, pitch)):
            h, w = planes[p].shape
            va[o:o + pt * h].view(h, pt)[:, :w].copy_(torch.from_numpy(planes[p]))
        torch.cuda.synchronize()

        ref_dev = [torch.from_numpy(p).cuda() for p in planes]
        expected = _encode(device, _frame_of(ref_dev))
        got = _encode(device, b.frame(layout))
        assert np.array_equal(got, expected)

        # Decode into the imported mapping; read back through the exporting one.
        enc = pw.Encoder(device, W, H, pw.CHROMA_420, precision=pw.PRECISION_FP32)
        dec = pw.Decoder(device, W, H, pw.CHROMA_420, precision=pw.PRECISION_FP32)
        try:
            pw.encode_batch([enc], [_frame_of(ref_dev)], [BUDGET])
            va.zero_()
            torch.cuda.synchronize()
            pw.decode_encoded_batch([dec], [enc], [b.target(layout)])
            outs = [torch.empty_like(t) for t in ref_dev]
            pw.decode_encoded_batch([dec], [enc], [pw.hip_target([(o.data_ptr(), o.stride(0)) for o in outs])])
            torch.cuda.synchronize()
            for p, (o, pt) in enumerate(zip(off, pitch)):
                h, w = planes[p].shape
                assert torch.equal(va[o:o + pt * h].view(h, pt)[:, :w], outs[p])
        finally:
            enc.close()
            dec.close()
    finally:
        b.close()
        a.close()


def test_nv12_input(device):
    planes = synth.SceneGenerator(W, H, True, seed=22, frames=1, pan=3, redraw=0.05).frame(0)
    pitch = [256, 256]
    off = [0, 256 * H]
    layout = pw.external_layout(W, H, pw.CHROMA_420, off, pitch, layout=pw.PLANES_INTERLEAVED_CHROMA)
    a, b = _shared(device, off[1] + pitch[1] * (H // 2))
    try:
        va = _view(a.address, a.size)
        va[0:W * H].view(H, W).copy_(torch.from_numpy(planes[0]))
        uv = np.stack([planes[1], planes[2]], axis=-1).reshape(H // 2, W)
        va[off[1]:off[1] + W * (H // 2)].view(H // 2, W).copy_(torch.from_numpy(np.ascontiguousarray(uv)))
        torch.cuda.synchronize()
        expected = _encode(device, _frame_of([torch.from_numpy(p).cuda() for p in planes]))
        assert np.array_equal(_encode(device, b.frame(layout)), expected)
        # The decoder writes planar output only.
        with pytest.raises(pw.PyroWaveError):
            b.target(layout)
    finally:
        b.close()
        a.close()


def test_layout_checks(device):
    layout, off, pitch = _planar_layout()
    a, b = _shared(device, off[2] + pitch[2] * (H // 2))
    try:
        size = b.size
        bad = [
            pw.external_layout(W, H, pw.CHROMA_420, [0, 0, size - 100], pitch),  # Cr runs past the end
            pw.external_layout(W, H, pw.CHROMA_420, off, [W - 1, 192, 192]),     # pitch below the width
            pw.external_layout(W + 1, H, pw.CHROMA_420, off, pitch),             # odd 4:2:0 width
            pw.external_layout(W, H, pw.CHROMA_420, [size, 0, 0], pitch),        # offset at the end
        ]
        for L in bad:
            with pytest.raises(pw.PyroWaveError):
                b.frame(L)
            with pytest.raises(pw.PyroWaveError):
                b.target(L)
        b.frame(layout)
        b.target(layout)
    finally:
        b.close()
        a.close()
