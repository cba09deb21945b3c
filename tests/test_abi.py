"""The C-ABI library: it loads, exports every entry point include/clb.h
declares, its host-side pieces (the seeded geometry) agree with the oracle and
the reference's generator, and it reports a missing device as an error code."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import complab_amd as clb
import oracle as O
from complab_amd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "clb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    L = clb.lib()
    names = declared_entry_points()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(clb.exported_symbols()) == names
    assert L.clb_abi_version() == 1


def test_ctypes_params_match_c_layout():
    out = np.zeros(7, dtype=np.int64)
    O.lib().orc_params_layout(out.ctypes.data_as(C.c_void_p))
    assert out[0] == C.sizeof(P.CaseParams)


@pytest.mark.parametrize("seed", [1, 2, 7, 13])
def test_geometry_pack_matches_oracle(seed):
    a = clb.geom_pack(30, 16, 14, 2.5, 0.42, seed)
    b = O.geom_pack(30, 16, 14, 2.5, 0.42, seed)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_geometry_generator_matches_reference_vectors():
    d = np.load(os.path.join(GOLD, "geometry_generate.npz"))
    assert np.array_equal(clb.rng_draws(int(d["rng_seed"]), 32), d["rng"])
    assert np.array_equal(clb.geom_spheres(24, 16, 12, 3.0, 14, 77, 1), d["spheres"])


def test_footprint_counts_populations():
    p = P.default_params(258, 128, 128, 4)
    p.flow_on = 1
    f64 = clb.footprint(p)
    p.precision = P.FP32
    f32 = clb.footprint(p)
    N = 258 * 128 * 128
    assert f64 > N * (4 * 7 * 2 + 19 * 2) * 8
    assert f32 < f64


def test_no_device_is_an_error_code_not_a_crash():
    import subprocess, sys
    code = ("import sys; sys.path.insert(0, %r); import complab_amd as c\n"
            "try:\n    c.Context(0)\n    print('DEVICE')\nexcept c.ClbError as e:\n    print('ERR', e.code)\n"
            % os.path.join(ROOT, "complab3d-gems_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", "")))
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("DEVICE") or r.stdout.startswith("ERR")
