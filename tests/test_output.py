"""The output contract: the VTK ImageData layout against hand-written known
answers, summary.csv, and a CompLaB.xml run driven by the runner (on the
oracle's case here; the device run is compared with it in the GPU tests)."""
import base64
import os
import struct

import numpy as np
import pytest

from complab_amd import inputs, output, runner
from complab_amd import params as P
from complab_amd.diagnostics import FieldStat, Row, row_of
import oracle as O

RUN_XML = """<?xml version="1.0" ?>
<parameters>
  <path><input_path>in</input_path><output_path>out</output_path></path>
  <simulation_mode><biotic_mode>false</biotic_mode></simulation_mode>
  <LB_numerics>
    <domain>
      <nx>10</nx><ny>6</ny><nz>5</nz><dx>1</dx><unit>mm</unit>
      <filename>g.dat</filename><characteristic_length>8</characteristic_length>
      <material_numbers><pore>2</pore><solid>0</solid><bounce_back>1</bounce_back></material_numbers>
    </domain>
    <delta_P>1e-4</delta_P><Peclet>1.0</Peclet><tau>0.8</tau>
    <iteration>
      <ns_max_iT1>3000</ns_max_iT1><ns_max_iT2>300</ns_max_iT2>
      <ns_converge_iT1>1e-8</ns_converge_iT1><ns_converge_iT2>1e-6</ns_converge_iT2>
      <ade_max_iT>25</ade_max_iT>
    </iteration>
  </LB_numerics>
  <chemistry>
    <number_of_substrates>2</number_of_substrates>
    <substrate0><name_of_substrates>A</name_of_substrates><initial_concentration>0.1</initial_concentration>
      <substrate_diffusion_coefficients><in_pore>1e-9</in_pore><in_biofilm>1e-9</in_biofilm></substrate_diffusion_coefficients>
      <left_boundary_type>Dirichlet</left_boundary_type><left_boundary_condition>1</left_boundary_condition>
      <right_boundary_type>Neumann</right_boundary_type><right_boundary_condition>0</right_boundary_condition></substrate0>
    <substrate1><name_of_substrates>B</name_of_substrates><initial_concentration>0.5</initial_concentration>
      <substrate_diffusion_coefficients><in_pore>2e-9</in_pore><in_biofilm>2e-9</in_biofilm></substrate_diffusion_coefficients>
      <left_boundary_type>closed</left_boundary_type><left_boundary_condition>0</left_boundary_condition>
      <right_boundary_type>closed</right_boundary_type><right_boundary_condition>0</right_boundary_condition></substrate1>
  </chemistry>
  <IO><save_VTK_interval>10</save_VTK_interval><mask_filename>maskL</mask_filename></IO>
  <diagnostics><enabled>true</enabled><interval>5</interval><tolerance>1e-6</tolerance><conserve>B</conserve></diagnostics>
</parameters>
"""
NX, NY, NZ = 10, 6, 5


def write_run(folder):
    """The XML above and a seeded geometry under `folder` (walls at y = 0 and
    ny-1, 15 % single-voxel grains and a 3x3x3 block, every solid voxel that
    touches water a bounce-back one, so the block keeps one NoDynamics voxel);
    returns the XML path."""
    os.makedirs(os.path.join(folder, "in"), exist_ok=True)
    rng = np.random.default_rng(4)
    g = np.full((NZ, NY, NX), 2)
    g[rng.random((NZ, NY, NX)) < 0.15] = 0
    g[1:4, 1:4, 4:7] = 0
    g[:, 0, :] = 0
    g[:, -1, :] = 0
    wet = np.zeros(g.shape, bool)
    for ax in range(3):
        for sh in (1, -1):
            nb = np.roll(g == 2, sh, axis=ax)
            edge = [slice(None)] * 3
            edge[ax] = 0 if sh == 1 else -1                   # no wrap-around
            nb[tuple(edge)] = False
            wet |= nb
    g[(g == 0) & wet] = 1
    assert (g == 0).any()
    with open(os.path.join(folder, "in", "g.dat"), "w") as f:        # x slowest, z fastest
        f.write("\n".join(str(v) for v in g.transpose(2, 1, 0).reshape(-1)) + "\n")
    path = os.path.join(folder, "CompLaB.xml")
    with open(path, "w") as f:
        f.write(RUN_XML)
    return path


def test_vti_known_answer(tmp_path):
    """A 2x1x1 grid, Density (1.5, -2.0) Float64, written out by hand."""
    p = str(tmp_path / "a.vti")
    output.write_vti(p, [("Density", np.array([[[1.5, -2.0]]]), "Float64")])
    expect = ('<?xml version="1.0"?>\n'
              '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">\n'
              '<ImageData WholeExtent="0 1 0 0 0 0" Origin="0 0 0" Spacing="1 1 1">\n'
              '<Piece Extent="0 1 0 0 0 0">\n'
              '<PointData>\n'
              '<DataArray type="Float64" Name="Density" format="binary" encoding="base64" NumberOfComponents="1">\n'
              # UInt32 16, then 1.5 = 0x3FF8000000000000 and -2.0 = 0xC000000000000000, little endian
              'EAAAAA==' 'AAAAAAAA+D8AAAAAAAAAwA=='
              '\n</DataArray>\n'
              '</PointData>\n</Piece>\n</ImageData>\n</VTKFile>\n')
    assert open(p).read() == expect


def test_vti_vectors_interleave_x_fastest(tmp_path):
    """(c, z, y, x) arrays go out as x-fastest points of interleaved
    components; a scalar in Float32 rounds like numpy's cast."""
    u = np.arange(3 * 2 * 2 * 3, dtype=np.float64).reshape(3, 2, 2, 3) * 0.5
    s = np.linspace(-1, 1, 12).reshape(2, 2, 3)
    p = str(tmp_path / "v.vti")
    output.write_vti(p, [("velocityNorm", s, "Float32"), ("velocity", u, "Float32")])
    text = open(p).read()
    assert 'WholeExtent="0 2 0 1 0 1"' in text
    body = text.split('Name="velocity" format="binary" encoding="base64" NumberOfComponents="3">\n')[1].split("\n")[0]
    n = struct.unpack("<I", base64.b64decode(body[:8]))[0]
    vals = struct.unpack("<%df" % (n // 4), base64.b64decode(body[8:]))
    assert n == 3 * 12 * 4
    for z in range(2):
        for y in range(2):
            for x in range(3):
                pt = (z * 2 + y) * 3 + x
                assert vals[3 * pt: 3 * pt + 3] == tuple(float(np.float32(u[c, z, y, x])) for c in range(3))
    arrs, dims = output.read_vti(p)
    assert dims == (3, 2, 2)
    assert np.array_equal(arrs["velocity"], u.astype(np.float32))
    assert np.array_equal(arrs["velocityNorm"], s.astype(np.float32))


def test_snapshot_names():
    assert output.snapshot_name("A_", 900) == "A_0000900.vti"
    assert output.snapshot_name("rate_A_", 1234567) == "rate_A_1234567.vti"


def test_summary_csv_layout(tmp_path):
    p = str(tmp_path / "s" / "summary.csv")
    w = output.SummaryCsv(p)
    rows = [Row(0, 0.25, 100, [FieldStat("A", 1.0 / 3, 1.0 / 300, 0.0, 0.1, 2e-17), FieldStat("M", 5.0, 0.05, 0.0, 1.0, 0.0)]),
            Row(10, 0.2, 80, [FieldStat("A", 0.1, 0.1 / 80, -1e-300, 7.0, 1e300), FieldStat("M", 5.5, 0.06875, 0.0, 1.5, 0.5)])]
    for r in rows:
        w.append(r)
    lines = open(p).read().splitlines()
    assert lines[0] == "# CompLB3D run summary, one row per diagnostic interval"
    assert sum(ln.startswith("#") for ln in lines) == 8
    assert lines[8] == ("iteration,porosity,open_voxels,A_total,A_mean,A_min,A_max,A_held,"
                        "M_total,M_mean,M_min,M_max,M_held")
    # the digits glibc's printf("%.17g") gives for these doubles
    assert lines[9] == ("0,0.25,100,0.33333333333333331,0.0033333333333333335,0,0.10000000000000001,"
                        "2.0000000000000001e-17,5,0.050000000000000003,0,1,0")
    cols, vals = output.read_summary(p)
    assert len(cols) == 13 and vals.shape == (2, 13)
    assert vals[1, 6] == 7.0 and vals[1, 7] == 1e300 and vals[1, 5] == -1e-300      # %.17g round-trips


def test_runner_writes_the_reference_files(tmp_path):
    """The runner on the oracle's case: file set, the box crop, the closed
    voxels' fictitious densities, masks, velocities and summary rows."""
    r = inputs.load(write_run(str(tmp_path)))
    assert (r.vtk_interval, r.mask_filename, r.summary_csv) == (10, "maskL", "summary.csv")
    assert r.output_path == os.path.join(str(tmp_path), "out")
    o = O.OracleCase(r.params, r.geometry)
    res = runner.run(r, case=o)
    assert res.steps == 25
    names = sorted(os.listdir(res.out_dir))
    expect = sorted(["%s_%07d.vti" % (n, it) for n in ("A", "B", "maskL", "nsLattice") for it in (0, 10, 20, 25)]
                    + ["summary.csv"])
    assert names == expect
    mask = o.read(P.FIELD_MASK)
    box = mask[:, :, 1:-1]
    for s, n in enumerate("AB"):
        arr, dims = output.read_vti(os.path.join(res.out_dir, "%s_0000025.vti" % n))
        assert dims == (NX, NY, NZ)
        d, c = arr["Density"], o.read(P.FIELD_CONC, s)[:, :, 1:-1]
        assert np.array_equal(d[box >= P.PORE], c[box >= P.PORE])
        assert np.all(d[box == P.SOLID] == 1.0) and np.all(d[box == P.WALL] == 0.0)
    m, _ = output.read_vti(os.path.join(res.out_dir, "maskL_0000025.vti"))
    assert np.array_equal(m["Density"], box.astype(np.float64))
    v, _ = output.read_vti(os.path.join(res.out_dir, "nsLattice_0000025.vti"))
    u = np.stack([o.read(P.FIELD_VEL, c)[:, :, 1:-1] for c in range(3)])
    u[:, box < P.PORE] = 0
    assert np.array_equal(v["velocity"], u.astype(np.float32))
    assert np.allclose(v["velocityNorm"], np.sqrt((u * u).sum(0)), rtol=1e-6, atol=0)
    assert np.all(u[0][box >= P.PORE].mean() > 0)                      # the flow runs along +x
    cols, rows = output.read_summary(res.summary_csv)
    assert list(rows[:, 0]) == [0, 5, 10, 15, 20]
    assert cols[3:8] == ["A_total", "A_mean", "A_min", "A_max", "A_held"]
    # the last row against a second run stopped there
    o2 = O.OracleCase(r.params, r.geometry)
    runner.run(r, case=o2, out_dir=str(tmp_path / "out2"), steps=20)
    want = row_of(o2, 20, r.species, r.microbes)
    got = rows[-1]
    assert got[1] == want.porosity and got[2] == want.open_voxels
    flat = [x for f in want.fields for x in (f.total, f.mean, f.minv, f.maxv, f.held)]
    assert list(got[3:]) == flat
    assert res.verdict == "PASS"                                        # B is closed at both planes: conserved


def test_runner_writes_rate_snapshots(tmp_path):
    """rate_<species>_<iT>.vti (writeRateVTI): the step after each
    concentration snapshot and the last step, "Rate" = the requested
    increment / dt.  For A -> B at k = 1e-3 the rate of A is -k times A at the
    start of the step and the rate of B its negative (a sign, a species or a
    step mixed up fails one of them)."""
    r = inputs.load(write_run(str(tmp_path)))
    p = r.params
    p.n_rxn = 1
    p.rxn_r1[0], p.rxn_r2[0], p.rxn_k[0] = 0, -1, 1e-3
    p.rxn_stoich[0][0], p.rxn_stoich[0][1] = -1.0, 1.0
    o = O.OracleCase(p, r.geometry)
    res = runner.run(r, case=o)
    names = sorted(os.listdir(res.out_dir))
    for n in "AB":
        for it in (0, 10, 20, 25):
            assert "rate_%s_%07d.vti" % (n, it) in names
    box = o.read(P.FIELD_MASK)[:, :, 1:-1] >= P.PORE
    # the same run stopped before step 21, then that step captured
    o2 = O.OracleCase(p, r.geometry)
    runner.run(r, case=o2, out_dir=str(tmp_path / "o2"), steps=20)
    a0 = o2.read(P.FIELD_CONC, 0)[:, :, 1:-1]
    o2.rate_capture(True)
    assert o2.run_steps(1) == (0, 1)
    ra, _ = output.read_vti(os.path.join(res.out_dir, "rate_A_0000020.vti"))
    rb, _ = output.read_vti(os.path.join(res.out_dir, "rate_B_0000020.vti"))
    assert np.array_equal(ra["Rate"], o2.read(P.FIELD_RATE, 0)[:, :, 1:-1])
    assert np.allclose(ra["Rate"][box], -1e-3 * a0[box] / p.dt, rtol=1e-12, atol=0)
    assert np.array_equal(rb["Rate"], -ra["Rate"])
    assert np.all(ra["Rate"][~box] == 0.0) and ra["Rate"][box].min() < 0
