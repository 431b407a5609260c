"""The C-ABI library on a machine without a GPU: it loads, exports every
function include/*.h declares and nothing else, its ctypes mirrors match the
C struct layouts, and calls that need no device behave as the reference's
C API does.  No compute call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import pyrowave_amd as pw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for h in ("pyrowave.h", "pyrowave_hip.h"):
        text = open(os.path.join(INCLUDE, h)).read()
        names += re.findall(r"PYROWAVE_PUBLIC_API\s+[\w\s\*]+?\b(pyrowave_\w+)\s*\(", text)
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(pw.LIB_PATH):
        from pyrowave_amd import build
        build.build()
    return pw.load()


def test_headers_declare_the_reference_api():
    names = set(declared_functions())
    # The 34 functions of the reference's pyrowave-shared.def:3-37.
    reference = {
        "pyrowave_get_api_version", "pyrowave_create_default_device", "pyrowave_create_device_by_compat",
        "pyrowave_create_device", "pyrowave_device_set_command_buffer", "pyrowave_device_report_performance_stats",
        "pyrowave_device_get_vk_device_handles", "pyrowave_device_confirm_interop_support",
        "pyrowave_device_set_queue_type", "pyrowave_device_destroy", "pyrowave_sync_object_create",
        "pyrowave_sync_object_get_semaphore", "pyrowave_sync_object_export_handle",
        "pyrowave_sync_object_cpu_wait", "pyrowave_sync_object_cpu_signal", "pyrowave_sync_object_destroy",
        "pyrowave_image_create", "pyrowave_image_get_handle", "pyrowave_image_get_image_view",
        "pyrowave_image_destroy", "pyrowave_encoder_create", "pyrowave_encoder_encode_gpu_synchronous",
        "pyrowave_encoder_encode_cpu_synchronous", "pyrowave_encoder_compute_num_packets",
        "pyrowave_encoder_packetize", "pyrowave_encoder_destroy", "pyrowave_decoder_device_prefers_fragment_path",
        "pyrowave_decoder_create", "pyrowave_decoder_clear", "pyrowave_decoder_push_packet",
        "pyrowave_decoder_decode_is_ready", "pyrowave_decoder_decode_gpu_buffer",
        "pyrowave_decoder_decode_cpu_buffer_synchronous", "pyrowave_decoder_destroy",
    }
    assert len(reference) == 34
    assert reference <= names, sorted(reference - names)
    assert set(pw.EXPORTS) == names


def test_library_exports_exactly_the_declared_functions(lib):
    names = set(declared_functions())
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", pw.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    assert exported == names, (sorted(exported - names), sorted(names - exported))


def test_soname_and_version(lib):
    out = subprocess.run(["readelf", "-d", pw.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "libpyrowave-shared.so.0" in out
    ma, mi, pa = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib.pyrowave_get_api_version(C.byref(ma), C.byref(mi), C.byref(pa))
    assert (ma.value, mi.value, pa.value) == (0, 4, 0)


def test_library_does_not_link_the_oracle():
    out = subprocess.run(["readelf", "-d", pw.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", pw.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "or_" not in " ".join(s.split()[-1][:3] for s in syms.splitlines())


LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "pyrowave_hip.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  S(pyrowave_encoder_create_info) O(pyrowave_encoder_create_info, chroma)
  S(pyrowave_decoder_create_info) O(pyrowave_decoder_create_info, fragment_path)
  S(pyrowave_hip_encoder_create_info) O(pyrowave_hip_encoder_create_info, flags)
  S(pyrowave_hip_decoder_create_info) O(pyrowave_hip_decoder_create_info, precision)
  S(pyrowave_cpu_buffer) O(pyrowave_cpu_buffer, width) O(pyrowave_cpu_buffer, format)
  S(pyrowave_rate_control) S(pyrowave_packet) O(pyrowave_packet, size)
  S(pyrowave_hip_plane) O(pyrowave_hip_plane, pixel_stride_in_bytes) S(pyrowave_hip_frame)
  S(pyrowave_hip_output_plane) S(pyrowave_hip_decode_target)
  S(pyrowave_hip_encoded_frame) O(pyrowave_hip_encoded_frame, num_blocks) O(pyrowave_hip_encoded_frame, sequence)
  return 0;
}
"""


def test_ctypes_mirrors_match_c_layout(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(LAYOUT_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    mirrors = {
        "pyrowave_encoder_create_info": pw.EncoderCreateInfo, "pyrowave_decoder_create_info": pw.DecoderCreateInfo,
        "pyrowave_hip_encoder_create_info": pw.HipEncoderCreateInfo,
        "pyrowave_hip_decoder_create_info": pw.HipDecoderCreateInfo, "pyrowave_cpu_buffer": pw.CpuBuffer,
        "pyrowave_rate_control": pw.RateControl, "pyrowave_packet": pw.Packet, "pyrowave_hip_plane": pw.HipPlane,
        "pyrowave_hip_frame": pw.HipFrame, "pyrowave_hip_output_plane": pw.HipOutputPlane,
        "pyrowave_hip_decode_target": pw.HipDecodeTarget, "pyrowave_hip_encoded_frame": pw.HipEncodedFrame,
    }
    for key, val in got.items():
        if "." in key:
            t, f = key.split(".")
            assert getattr(mirrors[t], f).offset == int(val), key
        else:
            assert C.sizeof(mirrors[key]) == int(val), key


def test_null_handles_are_rejected(lib):
    """Null device or object arguments give INVALID_ARGUMENT and no crash
    (pyrowave_c.cpp:812-819, 1202-1209)."""
    info = pw.EncoderCreateInfo(None, 64, 64, pw.CHROMA_420)
    enc = C.c_void_p()
    assert lib.pyrowave_encoder_create(C.byref(info), C.byref(enc)) == pw.ERROR_INVALID_ARGUMENT
    dinfo = pw.DecoderCreateInfo(None, 64, 64, pw.CHROMA_420, False)
    dec = C.c_void_p()
    assert lib.pyrowave_decoder_create(C.byref(dinfo), C.byref(dec)) == pw.ERROR_INVALID_ARGUMENT
    assert lib.pyrowave_decoder_push_packet(None, b"\0" * 8, 8) == pw.ERROR_INVALID_ARGUMENT
    assert not lib.pyrowave_decoder_decode_is_ready(None, False)
    lib.pyrowave_encoder_destroy(None)
    lib.pyrowave_decoder_destroy(None)
    lib.pyrowave_device_destroy(None)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(pw, "_lib", None)
    monkeypatch.setattr(pw, "LIB_PATH", os.path.join(ROOT, "no_such_dir", "libpyrowave-shared.so"))
    with pytest.raises(Exception):
        pw.load()


def test_synthetic_scene_is_deterministic():
    from pyrowave_amd import synth
    a = synth.SceneGenerator(320, 180, True, seed=9, frames=3, pan=3, redraw=0.05)
    b = synth.SceneGenerator(320, 180, True, seed=9, frames=3, pan=3, redraw=0.05)
    for t in range(3):
        for x, y in zip(a.frame(t), b.frame(t)):
            assert x.dtype == np.uint8 and np.array_equal(x, y)
    y0, y1 = a.frame(0)[0], a.frame(1)[0]
    assert not np.array_equal(y0, y1)
