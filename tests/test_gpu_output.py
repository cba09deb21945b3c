"""A CompLaB.xml run through the runner on the device against the same run on
the oracle: every snapshot file and the summary rows."""
import os

import numpy as np
import pytest

from complab_amd import Case, inputs, output, runner
import oracle as O
from test_output import write_run

pytestmark = pytest.mark.gpu


def test_device_run_writes_the_oracle_files(ctx, tmp_path):
    print("case: runner on the 10x6x5 abiotic XML with flow, 25 steps", flush=True)
    r = inputs.load(write_run(str(tmp_path)))
    g = runner.run(r, case=Case(ctx, r.params, r.geometry), out_dir=str(tmp_path / "gpu"))
    o = runner.run(r, case=O.OracleCase(r.params, r.geometry), out_dir=str(tmp_path / "cpu"))
    assert g.steps == o.steps == 25 and g.flow_iters == o.flow_iters
    names = sorted(os.listdir(g.out_dir))
    assert names == sorted(os.listdir(o.out_dir)) and len(names) == 17
    for n in names:
        if not n.endswith(".vti"):
            continue
        ga, gd = output.read_vti(os.path.join(g.out_dir, n))
        oa, od = output.read_vti(os.path.join(o.out_dir, n))
        assert gd == od and sorted(ga) == sorted(oa)
        for k in ga:
            a, b = ga[k].astype(np.float64), oa[k].astype(np.float64)
            if n.startswith("maskL"):
                assert np.array_equal(a, b), n
            else:
                tol = 1e-10 if ga[k].dtype == np.float64 else 2e-6   # FP64 fields; Float32 files round once
                scale = max(float(np.max(np.abs(b))), 1e-300)
                assert float(np.max(np.abs(a - b))) / scale <= tol, (n, k)
    gc, grows = output.read_summary(g.summary_csv)
    oc, orows = output.read_summary(o.summary_csv)
    assert gc == oc and grows.shape == orows.shape == (5, 3 + 5 * 2)
    assert np.array_equal(grows[:, :3], orows[:, :3])
    rel = np.abs(grows - orows) / np.maximum(np.abs(orows), 1e-300)
    assert float(np.max(rel[:, 3:][np.abs(orows[:, 3:]) > 1e-14])) <= 1e-10
    assert g.verdict == o.verdict == "PASS"


def test_runner_restart_from_a_checkpoint(ctx, tmp_path):
    print("case: runner with checkpoints every 10 steps, restart from step 10", flush=True)
    r = inputs.load(write_run(str(tmp_path)))
    r.chk_interval = 10
    full = runner.run(r, case=Case(ctx, r.params, r.geometry), out_dir=str(tmp_path / "full"))
    assert [os.path.basename(p) for p in full.checkpoints] == ["checkpoint_0000010.clbchk", "checkpoint_0000020.clbchk"]
    again = runner.run(r, case=Case(ctx, r.params, r.geometry), out_dir=str(tmp_path / "again"),
                       restart=full.checkpoints[0])
    assert again.steps == 25 and again.flow_iters == []
    names = sorted(n for n in os.listdir(again.out_dir) if n.endswith(".vti"))
    assert names == sorted("%s_%07d.vti" % (n, it) for n in ("A", "B", "maskL", "nsLattice") for it in (10, 20, 25))
    for n in names:
        a, _ = output.read_vti(os.path.join(full.out_dir, n))
        b, _ = output.read_vti(os.path.join(again.out_dir, n))
        for k in a:
            assert np.array_equal(a[k], b[k]), (n, k)


def test_rate_capture_matches_oracle(ctx):
    """The captured rates of a biotic CA case with an abiotic reaction, on the
    device against the oracle (1e-10), and the steps after capture is turned
    off continuing on the fused path to the same fields."""
    from complab_amd import configs, driver
    from complab_amd import params as P
    print("case: biotic 0.12-scale pack + S2 -> S3, rates captured over steps 31-35", flush=True)
    w = configs.config2(scale=0.12)
    w.transport.stabilize_steps = 20
    w.params.ns_max_iter = 3000
    p = w.params
    p.n_rxn = 1
    p.rxn_r1[0], p.rxn_r2[0], p.rxn_k[0] = 2, 0, 0.5
    p.rxn_stoich[0][2], p.rxn_stoich[0][3] = -1.0, 1.0
    g = Case(ctx, p, w.geometry)
    o = O.OracleCase(p, w.geometry)
    rg = driver.setup(g, p, w.flow, w.transport)
    ro = driver.setup(o, p, w.flow, w.transport)
    assert rg.iters == ro.iters
    for c in (g, o):
        c.set_omegas(ro.omega_pore, ro.omega_biofilm, 1.0)
        assert c.run_steps(30) == (0, 30)
        c.rate_capture(True)
        assert c.run_steps(5) == (0, 5)
    om = o.read(P.FIELD_MASK) >= 2
    for s in range(4):
        a, b = g.read(P.FIELD_RATE, s), o.read(P.FIELD_RATE, s)
        assert np.max(np.abs(a - b)) <= 1e-10 * max(np.max(np.abs(b)), 1e-300), s
        assert np.all(a[~om] == 0.0)
    assert np.max(np.abs(o.read(P.FIELD_RATE, 3))) > 0
    for c in (g, o):
        c.rate_capture(False)
        assert c.run_steps(5) == (0, 5)
    assert np.array_equal(g.read(P.FIELD_MASK), o.read(P.FIELD_MASK))
    for s in range(4):
        a, b = g.read(P.FIELD_CONC, s)[om], o.read(P.FIELD_CONC, s)[om]
        assert np.max(np.abs(a - b)) <= 1e-10 * max(np.max(np.abs(b)), 1e-300), s
    g.close()
