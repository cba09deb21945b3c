"""A case driven from C++ alone (include/clb_driver.hh through the example
complab3d-gems_amd/clb_run_case, built with the library) gives the same set-up
and the bit-identical fields of the same sequence driven from Python."""
import os
import subprocess

import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs, driver
from complab_amd import params as P

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(clb.LIB_PATH), "clb_run_case")


def test_cpp_driver_matches_python(ctx, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("clb_run_case is not built (run the build)")
    w = configs.config1(scale=0.25)
    w.transport.stabilize_steps = 40
    steps = 60
    p = w.params
    (tmp_path / "p.bin").write_bytes(bytes(p))
    (tmp_path / "g.bin").write_bytes(np.ascontiguousarray(w.geometry, dtype=np.int8).tobytes())
    fs, ts = w.flow, w.transport
    lines = ["pe %r" % fs.pe, "char_length %r" % fs.char_length, "delta_p %r" % fs.delta_p,
             "ns_tau %r" % fs.ns_tau, "max_iter %d" % fs.max_iter, "rule %d" % fs.rule, "eps %r" % fs.eps,
             "has_biofilm %d" % int(fs.has_biofilm), "stabilize %d" % ts.stabilize_steps, "dx %r" % ts.dx,
             "steps %d" % steps]
    lines += ["d_pore %r" % d for d in ts.d_pore] + ["d_biofilm %r" % d for d in ts.d_biofilm]
    (tmp_path / "s.txt").write_text("\n".join(lines) + "\n")
    r = subprocess.run([EXE, str(tmp_path / "p.bin"), str(tmp_path / "g.bin"), str(tmp_path / "s.txt"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rep_c = dict(line.split(None, 1) for line in r.stdout.strip().splitlines())
    fields_c = np.fromfile(str(tmp_path / "out.bin"), dtype=np.float64).reshape(p.n_species, *w.geometry.shape)

    g = clb.Case(ctx, p, w.geometry)
    rep = driver.setup(g, p, fs, ts)
    assert g.run_steps(steps) == (0, steps)
    assert [int(x) for x in rep_c["iters"].split()] == rep.iters
    assert float(rep_c["delta_p"]) == rep.delta_p and float(rep_c["dt"]) == rep.dt
    for s in range(p.n_species):
        assert np.array_equal(fields_c[s], g.read(P.FIELD_CONC, s)), s
    g.close()
