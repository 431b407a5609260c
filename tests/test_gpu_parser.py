"""The product decoder's packet parser (pyrowave_decoder_push_packet /
decode_is_ready / clear through the C ABI) replayed over the golden parser
scripts recorded from the reference's BitstreamParser, and the decode of the
resulting payload compared with the oracle's decode of the reference's
payload and offsets."""
import os

import numpy as np
import pytest

import pyrowave_amd as pw
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_G = np.load(os.path.join(GOLDEN, "parser.npz"))
NAMES = [str(n) for n in _G["names"]]


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_parser_script_matches_reference(device, name):
    g = _G
    w, h, ch = int(g["w"]), int(g["h"]), int(g["chroma"])
    dec = pw.Decoder(device, w, h, ch, precision=pw.PRECISION_FP32)
    try:
        ops, lens, data = g[f"{name}_ops"], g[f"{name}_lens"], g[f"{name}_data"].tobytes()
        steps = g[f"{name}_steps"]
        pos = 0
        for k, (op, n) in enumerate(zip(ops, lens)):
            if op == 0:
                rc = dec.push_packet(data[pos:pos + n])
                pos += int(n)
                exp_ok = bool(steps[k][0])
                assert (rc == pw.SUCCESS) == exp_ok, (name, k, rc)
                if not exp_ok:
                    assert rc == pw.ERROR_INVALID_ARGUMENT
            elif op == 1:
                dec.clear()
            else:
                # The reference's decode marks the frame decoded.
                dec.decode_cpu()
            assert dec.is_ready(False) == bool(steps[k][1]), (name, k)
            assert dec.is_ready(True) == bool(steps[k][2]), (name, k)
        if dec.is_ready(True):
            got = dec.decode_cpu()
            L = O.layout(w, h, ch)
            exp = O.idwt_frame(L, O.dequant(L, g[f"{name}_payload"], g[f"{name}_offsets"]))
            for c in range(3):
                assert np.array_equal(got[c], exp[c]), (name, c)
    finally:
        dec.close()
