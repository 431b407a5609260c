"""Mutation tests of the oracle's pins: each case plants one plausible mistake
in a copy of the oracle sources (a changed constant, a dropped term, floor for
ceil, ...), builds it on the CPU into a temporary library, and requires at
least one check of tests/test_oracle_pinning.py to fail against it.  The
shipped oracle must pass every check (test_oracle_pinning.py)."""
import os
import shutil
import subprocess

import pytest

from oracle import oracle as O

import test_oracle_pinning as P

ORACLE_DIR = os.path.dirname(O.LIB_PATH)

# (name, text in oracle/pyrowave_oracle.c, replacement)
MUTANTS = [
    ("csf_exponent", "expf(-powf(0.114f * cpd, 1.1f))", "expf(-powf(0.114f * cpd, 1.2f))"),
    ("csf_slope", "2.6f * (0.0192f + 0.114f * cpd)", "2.6f * (0.0192f + 0.120f * cpd)"),
    ("csf_floor", "if (cpd < 8.0f)\n\t\tcpd = 8.0f;", "if (cpd < 4.0f)\n\t\tcpd = 4.0f;"),
    ("chroma_weight", "csf *= 0.6f;", "csf *= 0.5f;"),
    ("rdo_normalisation", "return (weighted * weighted) * (1.0f / 256.0f);",
     "return (weighted * weighted) * (1.0f / 128.0f);"),
    ("rdo_dropped", "se_lane[s] = ((dot4_sq(err) + dot4_sq(err + 4)) * rdo_scale) * inv_quant_squared;",
     "se_lane[s] = (dot4_sq(err) + dot4_sq(err + 4)) * inv_quant_squared;"),
    ("recon_offset", "((float)iv + 0.5f) * (float)(1 << q)", "((float)iv + 0.25f) * (float)(1 << q)"),
    ("discard_weighted", "lane[s] = (dot4_sq(Q->t[s]) + dot4_sq(Q->t[s] + 4)) * inv_quant_squared;",
     "lane[s] = ((dot4_sq(Q->t[s]) + dot4_sq(Q->t[s] + 4)) * rdo_scale) * inv_quant_squared;"),
    ("dist_unclamped_order", "dist_tab[q] = or_f32_to_f16(se < 60000.0f ? se : 60000.0f);",
     "dist_tab[q] = or_f32_to_f16(se < 60000.0f ? se * 1.01f : 60000.0f);"),
    ("scale_code_floor", "Q->scale_code = (uint32_t)ceilf((inv_scale - 0.25f) * 8.0f);",
     "Q->scale_code = (uint32_t)floorf((inv_scale - 0.25f) * 8.0f);"),
    ("scale_target", "float target_max = (float)(1 << e) - 0.25f;", "float target_max = (float)(1 << e) - 0.5f;"),
    ("cost_signs_dropped", "cost_tab[q] = (uint16_t)(8 * late + nsig);", "cost_tab[q] = (uint16_t)(8 * late);"),
    ("cost_ctrl_bits", "total += cost[q][k] ? cost[q][k] + 24u : 0u;", "total += cost[q][k] ? cost[q][k] + 16u : 0u;"),
    ("cost_header_bits", "if (total)\n\t\t\t\t\t\t\t\ttotal += 64;", "if (total)\n\t\t\t\t\t\t\t\ttotal += 32;"),
    ("bucket_offset", "int k = 60 + 2 * E + (m >= BUCKET_T1) + (m >= BUCKET_T2);",
     "int k = 61 + 2 * E + (m >= BUCKET_T1) + (m >= BUCKET_T2);"),
    ("bucket_natural_log", "float m = 2.0f * f;\n\tint E = e - 1;\n\tint k",
     "float m = 2.0f * f;\n\tint E = (int)floorf(logf(r) / logf(2.0f) * 0.6931472f);\n\tint k"),
    ("bucket_cap", "raw[q] < 112u + (uint32_t)q ? raw[q] : 112u + (uint32_t)q;",
     "raw[q] < 100u + (uint32_t)q ? raw[q] : 100u + (uint32_t)q;"),
    ("bucket_scan_distance", "uint32_t cand = v[k] + (uint32_t)(q - k);", "uint32_t cand = v[k] + 1u;"),
    ("bucket_ratio_order", "bi = bucket_of_ratio(dd / (base_f - cost_f));",
     "bi = bucket_of_ratio((base_f - cost_f) / (dd > 0.0f ? dd : 1.0f));"),
    ("savings_subdivision", "uint32_t subdiv = block >> shamt;", "uint32_t subdiv = (block >> shamt) & 7u;"),
    ("band_term", "else if (band < 3)\n\t\tbits += 1;", "else if (band < 3)\n\t\tbits += 2;"),
    ("level_term", "bits += level;", "bits += level + (level > 2);"),
    ("chroma_term_dropped", "if (component != 0)\n\t\tbits--;", "if (component > 2)\n\t\tbits--;"),
    ("base_bits", "int bits = precision0 ? 6 : 8;", "int bits = precision0 ? 6 : 7;"),
    ("base_bits_precision0", "int bits = precision0 ? 6 : 8;", "int bits = precision0 ? 5 : 8;"),
]


def _build(tmpdir, name, old, new):
    src = open(os.path.join(ORACLE_DIR, "pyrowave_oracle.c")).read()
    assert src.count(old) == 1, f"mutant {name}: pattern not found exactly once"
    d = os.path.join(tmpdir, name)
    os.makedirs(d)
    for f in ("pyrowave_oracle.h", "psnr_hvs_m.c"):
        shutil.copy(os.path.join(ORACLE_DIR, f), d)
    with open(os.path.join(d, "pyrowave_oracle.c"), "w") as fh:
        fh.write(src.replace(old, new))
    lib = os.path.join(d, "liboracle_mutant.so")
    r = subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared",
                        "-o", lib, os.path.join(d, "pyrowave_oracle.c"), os.path.join(d, "psnr_hvs_m.c"), "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return lib


def _failing_checks():
    failed = []
    for check in P.CHECKS:
        try:
            check()
        except AssertionError:
            failed.append(check.__name__)
    return failed


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_caught(tmp_path, name, old, new):
    lib = _build(str(tmp_path), name, old, new)
    with O.use_library(lib):
        failed = _failing_checks()
    assert failed, f"mutant {name} passes every pinning check"
