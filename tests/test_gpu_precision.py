"""Precision of the reference entry points (pyrowave_encoder_create /
pyrowave_decoder_create): the reference's desktop default, precision 1
(CMakeLists.txt:44-49), unless PYROWAVE_PRECISION names another in [0, 2];
other values are reported and ignored (pyrowave_common.cpp:15-28).  GPU tests:
the bitstream and planes of a reference-API encoder / decoder equal those of
the HIP-API object created with the expected precision."""
import os

import numpy as np
import pytest

import pyrowave_amd as pw
from pyrowave_amd import synth
from tests.helpers import decode_bytes, gpu_encode

pytestmark = pytest.mark.gpu

W, H, BUDGET = 320, 192, 30000


def _run(device, precision, planes):
    enc = pw.Encoder(device, W, H, pw.CHROMA_420, precision=precision)
    dec = pw.Decoder(device, W, H, pw.CHROMA_420, precision=precision)
    try:
        data, pk = gpu_encode(enc, planes, BUDGET)
        return data, decode_bytes(dec, data, pk)
    finally:
        enc.close()
        dec.close()


@pytest.mark.parametrize("env,expected", [(None, pw.PRECISION_FP16_LL), ("1", pw.PRECISION_FP16_LL),
                                          ("2", pw.PRECISION_FP32), ("0", pw.PRECISION_FP16_MATH),
                                          ("7", pw.PRECISION_FP16_LL), ("-1", pw.PRECISION_FP16_LL)])
def test_reference_entry_point_precision(device, capfd, env, expected):
    planes = synth.frame(W, H, True, seed=41)
    saved = os.environ.pop("PYROWAVE_PRECISION", None)
    try:
        if env is not None:
            os.environ["PYROWAVE_PRECISION"] = env
        got_bytes, got_planes = _run(device, None, planes)
    finally:
        os.environ.pop("PYROWAVE_PRECISION", None)
        if saved is not None:
            os.environ["PYROWAVE_PRECISION"] = saved
    err = capfd.readouterr().err
    if env in ("7", "-1"):
        assert "precision must be in range [0, 2]" in err
    ref_bytes, ref_planes = _run(device, expected, planes)
    assert np.array_equal(got_bytes, ref_bytes)
    for a, b in zip(got_planes, ref_planes):
        assert np.array_equal(a, b)
    # The precisions are distinguishable on this content.
    other = pw.PRECISION_FP32 if expected != pw.PRECISION_FP32 else pw.PRECISION_FP16_LL
    other_bytes, _ = _run(device, other, planes)
    assert not np.array_equal(other_bytes, ref_bytes)
