"""CompLaB.xml and geometry.dat readers: the tag mapping on a file written
here, the geometry orders and ghost columns, the reference's shipped inputs
(read in this container only: skipped where the reference tree is absent),
and one such input run through the oracle's set-up."""
import json
import os

import numpy as np
import pytest

from complab_amd import inputs
from complab_amd import params as P
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"

XML = """<?xml version="1.0" ?>
<parameters>
  <path><input_path>in</input_path></path>
  <simulation_mode><biotic_mode>true</biotic_mode><abiotic_rate_scale>0.5</abiotic_rate_scale></simulation_mode>
  <LB_numerics>
    <domain>
      <nx>6</nx><ny>4</ny><nz>3</nz><dx>2</dx><unit>mm</unit>
      <characteristic_length>12</characteristic_length>
      <filename>g.dat</filename>
      <material_numbers><pore>5 6</pore><solid>0</solid><bounce_back>9</bounce_back><microbe0>7</microbe0></material_numbers>
    </domain>
    <delta_P>2e-5</delta_P><Peclet>1.5</Peclet><tau>0.9</tau>
    <iteration>
      <ns_max_iT1>1234</ns_max_iT1><ns_max_iT2>55</ns_max_iT2>
      <ns_converge_iT1>1e-7</ns_converge_iT1><ns_converge_iT2>1e-5</ns_converge_iT2>
      <ns_update_interval>4</ns_update_interval><ade_max_iT>321</ade_max_iT>
    </iteration>
  </LB_numerics>
  <chemistry>
    <number_of_substrates>3</number_of_substrates>
    <substrate0><name_of_substrates>A</name_of_substrates><initial_concentration>0.25</initial_concentration>
      <substrate_diffusion_coefficients><in_pore>1e-9</in_pore><in_biofilm>5e-10</in_biofilm></substrate_diffusion_coefficients>
      <left_boundary_type>Dirichlet</left_boundary_type><left_boundary_condition>1</left_boundary_condition>
      <right_boundary_type>neumann</right_boundary_type><right_boundary_condition>0</right_boundary_condition></substrate0>
    <substrate1><name_of_substrates>B</name_of_substrates><initial_concentration>0</initial_concentration>
      <substrate_diffusion_coefficients><in_pore>2e-9</in_pore><in_biofilm>1e-9</in_biofilm></substrate_diffusion_coefficients>
      <left_boundary_type>no_flux</left_boundary_type><left_boundary_condition>0</left_boundary_condition>
      <right_boundary_type>Wall</right_boundary_type><right_boundary_condition>0</right_boundary_condition></substrate1>
    <substrate2><name_of_substrates>M</name_of_substrates><initial_concentration>0</initial_concentration>
      <immobile>yes</immobile>
      <left_boundary_type>closed</left_boundary_type><left_boundary_condition>0</left_boundary_condition>
      <right_boundary_type>closed</right_boundary_type><right_boundary_condition>0</right_boundary_condition></substrate2>
  </chemistry>
  <microbiology>
    <number_of_microbes>1</number_of_microbes><maximum_biomass_density>40</maximum_biomass_density>
    <thrd_biofilm_fraction>0.1</thrd_biofilm_fraction><CA_method>half</CA_method>
    <microbe0><name_of_microbes>Bug</name_of_microbes><solver_type>CA</solver_type><reaction_type>kinetics</reaction_type>
      <initial_densities>3.5</initial_densities><decay_coefficient>1e-6</decay_coefficient>
      <viscosity_ratio_in_biofilm>10</viscosity_ratio_in_biofilm></microbe0>
  </microbiology>
  <equilibrium><enabled>true</enabled><components>A B</components>
    <stoichiometry><species0>1 0</species0><species1>0 1</species1><species2>1 1</species2></stoichiometry>
    <logK><species2>3.5</species2></logK></equilibrium>
  <precipitation><enabled>true</enabled><solid_substrate>2</solid_substrate><max_precipRho>8</max_precipRho>
    <perm_ratio>0</perm_ratio><update_interval>50</update_interval></precipitation>
  <dissolution><enabled>true</enabled><reopen_fraction>0.8</reopen_fraction><update_interval>25</update_interval>
    <phase0><material_number>0</material_number><substrate>2</substrate><full_density>8</full_density>
      <is_precipitate>true</is_precipitate></phase0></dissolution>
  <diagnostics><enabled>true</enabled><interval>10</interval><tolerance>1e-8</tolerance>
    <conserve>A+M</conserve><conserve1>B+M</conserve1></diagnostics>
</parameters>
"""


def write_case(tmp_path, values):
    (tmp_path / "in").mkdir()
    (tmp_path / "CompLaB.xml").write_text(XML)
    (tmp_path / "in" / "g.dat").write_text("\n".join(str(v) for v in values) + "\n")
    return str(tmp_path / "CompLaB.xml")


def test_tags_map_onto_the_parameter_block(tmp_path):
    nx, ny, nz = 6, 4, 3
    rng = np.random.default_rng(1)
    vals = rng.choice([0, 5, 6, 7, 9], size=nx * ny * nz)
    r = inputs.load(write_case(tmp_path, vals))
    p = r.params
    assert (p.nx, p.ny, p.nz) == (8, 4, 3)
    assert r.dx == pytest.approx(2e-3) and r.flow.char_length == pytest.approx(6.0)
    assert (r.flow.pe, r.flow.delta_p, r.flow.ns_tau, r.flow.max_iter, r.flow.eps) == (1.5, 2e-5, 0.9, 1234, 1e-7)
    assert r.flow.has_biofilm and p.flow_on == 1
    assert (p.ns_max_iter, p.ns_conv_eps, p.ns_update_interval, p.ade_update_interval) == (55, 1e-5, 4, 1)
    assert p.ns_conv_rule == P.CONV_TRACER and r.steps == 321
    assert r.species == ["A", "B", "M"] and r.microbes == ["Bug"]
    assert list(p.left_btype[:3]) == [P.BC_DIRICHLET, P.BC_CLOSED, P.BC_CLOSED]
    assert list(p.right_btype[:3]) == [P.BC_NEUMANN, P.BC_CLOSED, P.BC_CLOSED]
    assert list(p.mobile[:3]) == [1, 1, 0] and p.c0[0] == 0.25 and p.left_bc[0] == 1.0
    assert r.transport.d_pore == [1e-9, 2e-9, 0.0] and r.transport.d_biofilm == [5e-10, 1e-9, 0.0]
    assert r.transport.dx == pytest.approx(2e-3)
    assert p.n_microbes == 1 and p.b0[0] == 3.5 and p.kdecay[0] == 1e-6
    assert p.perm_ratio == pytest.approx(0.1) and p.bmax == 40 and p.thrd_bfilm_frac == 0.1
    assert p.ca_method == P.CA_HALF and p.abiotic_rate_scale == 0.5
    assert p.eq_on == 1 and p.eq_nc == 2 and p.eq_solver == P.EQ_AAPCF
    assert [list(p.eq_stoich[s][:2]) for s in range(3)] == [[1, 0], [0, 1], [1, 1]]
    assert list(p.eq_logk[:3]) == [0, 0, 3.5] and list(p.eq_comp_species[:2]) == [0, 1]
    assert (p.eq_tol, p.eq_max_iter, p.eq_anderson_depth) == (1e-10, 200, 4)
    assert (p.precip_on, p.precip_species, p.precip_max_rho, p.precip_interval, p.precip_perm_ratio) == (1, 2, 8, 50, 0)
    assert (p.dissol_on, p.dissol_n_phases, p.dissol_species[0], p.dissol_is_precipitate[0], p.precip_phase_id) == (1, 1, 2, 1, 1)
    assert (p.dissol_full[0], p.dissol_reopen_frac, p.dissol_interval, p.dissol_surface_only) == (8, 0.8, 25, 1)
    assert r.diag_enabled and r.diag_interval == 10 and r.diag_tol == 1e-8 and r.conserve == ["A+M", "B+M"]
    # materials: 5, 6 pore; 9 bounce-back; 7 the microbe's biofilm; 0 solid
    raw = inputs.read_geometry_dat(str(tmp_path / "in" / "g.dat"), nx, ny, nz)
    lut = {0: P.SOLID, 5: P.PORE, 6: P.PORE, 7: P.BIOFILM, 9: P.WALL}
    assert np.array_equal(r.geometry, np.vectorize(lut.get)(raw).astype(np.int8))


def test_geometry_orders_and_ghost_columns(tmp_path):
    nx, ny, nz = 4, 3, 2
    f = tmp_path / "g.dat"
    f.write_text("\n".join(str(v) for v in range(nx * ny * nz)))
    g = inputs.read_geometry_dat(str(f), nx, ny, nz)
    # readGeometry: x slowest, then y, z fastest; value = (x * ny + y) * nz + z
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                assert g[z, y, x + 1] == (x * ny + y) * nz + z
    assert np.array_equal(g[:, :, 0], g[:, :, 1]) and np.array_equal(g[:, :, -1], g[:, :, -2])
    h = inputs.read_geometry_dat(str(f), nx, ny, nz, order="x_fastest")
    assert h[1, 2, 3 + 1] == 3 + nx * (2 + ny * 1)
    with pytest.raises(inputs.InputError):
        inputs.read_geometry_dat(str(f), nx + 1, ny, nz)


def test_refusals(tmp_path):
    path = write_case(tmp_path, [5] * (6 * 4 * 3))
    for old, new in [("<solver_type>CA", "<solver_type>FD"), ("<reaction_type>kinetics", "<reaction_type>glpk"),
                     ("Dirichlet</left", "Robin</left"), ("<unit>mm", "<unit>km"),
                     ("<reopen_fraction>0.8", "<reopen_fraction>1.0"), ("<pore>5 6</pore>", "<pore>5</pore>")]:
        (tmp_path / "CompLaB.xml").write_text(XML.replace(old, new))
        if old.startswith("<pore>"):
            (tmp_path / "in" / "g.dat").write_text("6\n" * (6 * 4 * 3))   # a material the XML no longer declares
        with pytest.raises(inputs.InputError):
            inputs.load(path)
        (tmp_path / "in" / "g.dat").write_text("5\n" * (6 * 4 * 3))


needs_ref = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")), reason="the reference tree is not here")


@needs_ref
@pytest.mark.parametrize("ex", ["01_flow_only", "02_diffusion_only", "03_abiotic_kinetics", "04_equilibrium",
                                "05_biotic_cellular_automaton", "13_precipitation", "14_dissolution",
                                "15_precip_and_dissolution", "17_symbolic_law", "19_thermodynamic_gate",
                                "20_upscaling"])
def test_reference_examples_load(ex):
    r = inputs.load(os.path.join(REF, "examples", ex, "CompLaB.xml"))
    p = r.params
    assert r.geometry.shape == (p.nz, p.ny, p.nx)
    assert set(np.unique(r.geometry)) <= {P.SOLID, P.WALL, P.PORE, P.BIOFILM, P.BIOFILM + 1}
    assert len(r.species) == p.n_species


@needs_ref
@pytest.mark.parametrize("ex", ["06_biotic_finite_difference", "07_biotic_lattice_boltzmann", "09_fba_glpk", "11_surrogate"])
def test_reference_examples_off_the_path_are_refused(ex):
    with pytest.raises(inputs.InputError):
        inputs.load(os.path.join(REF, "examples", ex, "CompLaB.xml"))


@needs_ref
def test_abc_geometry_file_against_its_index_porosity():
    gdir = os.path.join(REF, "ABC_sweep_package", "geometries")
    dims = [int(t) for t in open(os.path.join(gdir, "geom_0000", "geometry.dims")).read().split()]
    raw = inputs.read_geometry_dat(os.path.join(gdir, "geom_0000", "geometry.dat"), *dims)
    core = raw[:, :, 1:-1]
    fix = np.load(os.path.join(HERE, "golden", "abc_geom_0000.npz"))["codes"]
    assert np.array_equal(core, fix)
    lines = open(os.path.join(gdir, "index.csv")).read().splitlines()
    header = lines[0].split(",")
    idx = {c[header.index("name")]: c for c in (l.split(",") for l in lines[1:])}
    row = idx["geom_0000"]
    nx, ny, nz = dims
    pores = float((core == 2).sum())
    # porosity_block over the whole block; porosity inside the y/z walls
    assert abs(pores / (nx * ny * nz) - float(row[header.index("porosity_block")])) < 1e-6
    assert abs(pores / (nx * (ny - 2) * (nz - 2)) - float(row[header.index("porosity")])) < 1e-6


@needs_ref
def test_reference_example_runs_through_the_oracle_setup():
    r = inputs.load(os.path.join(REF, "examples", "02_diffusion_only", "CompLaB.xml"))
    r.transport.stabilize_steps = 0
    from complab_amd import driver
    o = O.OracleCase(r.params, r.geometry)
    rep = driver.setup(o, r.params, r.flow, r.transport)
    o.set_omegas(rep.omega_pore, rep.omega_biofilm, rep.dt)
    assert o.run_steps(5) == (0, 5)
    tot = o.species_totals()
    assert np.all(np.isfinite(tot))


def test_abiotic_symbolic_file_hook(tmp_path):
    """<symbolic><abiotic_file> binds an extent law to the run's substrates
    (complab3d_integration.hh:264-300, complab.cpp:2455-2493); refused beside
    dissolution and for a biotic (growth) law."""
    nx, ny, nz = 6, 4, 3
    xml = write_case(tmp_path, [5] * (nx * ny * nz))
    dissol = XML[XML.index("  <dissolution>"):XML.index("  <diagnostics>")]
    sym = "  <symbolic><enabled>true</enabled><abiotic_file>ab.sym</abiotic_file></symbolic>\n"
    (tmp_path / "ab.sym").write_text("units per_hour\nvars A B\nreaction A -1 B 2\nrate extent = 3.6 * A * B\n")
    with open(xml, "w") as fh:
        fh.write(XML.replace("  <diagnostics>", sym + "  <diagnostics>"))
    with pytest.raises(inputs.InputError):                     # with dissolution
        inputs.load(xml)
    with open(xml, "w") as fh:
        fh.write(XML.replace(dissol, sym))
    r = inputs.load(xml)
    prog = r.symbolic[P.SYM_ABIOTIC]
    assert prog.n_vars == 2 and list(prog.var_species[:2]) == [0, 1]
    assert list(prog.rate_species[:3]) == [-1, 0, 1] and prog.unit_scale == 1.0 / 3600
    out = O.sym_evaluate(prog, [0.5, 0.25, 0.0], 0.0)
    assert out[1] == pytest.approx(-3.6 * 0.125 / 3600, rel=1e-15) and out[2] == -2 * out[1]
    (tmp_path / "ab.sym").write_text("vars A Bug\nrate growth = A\n")
    with pytest.raises(inputs.InputError):
        inputs.load(xml)
    with open(xml, "w") as fh:                                 # not read while <symbolic> is off
        fh.write(XML.replace(dissol, sym.replace("true", "false")))
    assert P.SYM_ABIOTIC not in inputs.load(xml).symbolic


THM_BLOCK = """temperature 298.15
energy_units kJ
reaction R1
  stoich A -1 B 1
  dG0 -20
  microbe {microbe}
"""


def two_symbolic_microbes(tmp_path, microbe, kinds=("symbolic", "symbolic")):
    nx, ny, nz = 6, 4, 3
    xml = write_case(tmp_path, [5] * (nx * ny * nz))
    (tmp_path / "g0.sym").write_text("vars A Bug\nrate growth = 0.1 * A\n")
    (tmp_path / "g1.sym").write_text("vars A Bug2\nrate growth = 0.2 * A\n")
    (tmp_path / "r.thm").write_text(THM_BLOCK.format(microbe=microbe))
    text = XML.replace("<number_of_microbes>1</number_of_microbes>", "<number_of_microbes>2</number_of_microbes>")
    text = text.replace("<microbe0>7</microbe0>", "<microbe0>7</microbe0><microbe1>8</microbe1>")
    m0 = text[text.index("    <microbe0><name_of_microbes>"):text.index("  </microbiology>")]
    m1 = m0.replace("microbe0>", "microbe1>").replace(">Bug<", ">Bug2<")
    m0s = m0.replace("<reaction_type>kinetics</reaction_type>",
                     "<reaction_type>%s</reaction_type><expressions_file>g0.sym</expressions_file>" % kinds[0])
    m1s = m1.replace("<reaction_type>kinetics</reaction_type>",
                     "<reaction_type>%s</reaction_type><expressions_file>g1.sym</expressions_file>" % kinds[1])
    text = text.replace(m0, m0s + m1s)
    text = text.replace("  <diagnostics>", "  <symbolic><enabled>true</enabled></symbolic>\n"
                        "  <thermodynamics><enabled>true</enabled><energetics_file>r.thm</energetics_file></thermodynamics>\n"
                        "  <diagnostics>")
    with open(xml, "w") as fh:
        fh.write(text)
    return xml, text


def test_thm_block_gates_the_organism_it_names(tmp_path):
    """A block naming one organism gates that organism alone
    (specOfMicrobe, complab3d_thermo.hh:640-648), and only it gets the
    dG_/FT_ snapshot pair (gatedMicrobes, complab3d_energyfield.hh:190-197).
    An organism other than microbe0 is accepted when no organism uses the
    compiled kinetics path (complab3d_integration.hh:838)."""
    from complab_amd import runner
    xml, _ = two_symbolic_microbes(tmp_path, "Bug")
    r = inputs.load(xml)
    assert r.params.thm_on == 1 and r.params.thm_microbe == 0
    assert runner.gated_microbes(r) == [(0, "Bug")]
    (tmp_path / "b").mkdir()
    xml, _ = two_symbolic_microbes(tmp_path / "b", "Bug2")
    r = inputs.load(xml)
    assert r.params.thm_microbe == 1 and runner.gated_microbes(r) == [(1, "Bug2")]
    (tmp_path / "c").mkdir()
    xml, _ = two_symbolic_microbes(tmp_path / "c", "all")
    r = inputs.load(xml)
    assert r.params.thm_microbe == -1 and runner.gated_microbes(r) == [(0, "Bug"), (1, "Bug2")]
    (tmp_path / "d").mkdir()
    xml, _ = two_symbolic_microbes(tmp_path / "d", "Bug2", kinds=("kinetics", "symbolic"))
    with pytest.raises(inputs.InputError, match="microbe0"):          # the compiled path gates as microbe0 only
        inputs.load(xml)


def test_symbolic_runs_refuse_fixed_substrates_and_totals(tmp_path):
    """The symbolic processors skip substrates held fixed and can read
    speciation totals (complab3d_processors_symbolic.hh:124, 167); this path
    does neither, so a symbolic run that asks for them is refused rather than
    run differently."""
    xml, text = two_symbolic_microbes(tmp_path, "all")
    with open(xml, "w") as fh:
        fh.write(text.replace("<chemistry>", "<chemistry><fix_concentration>no yes no</fix_concentration>"))
    with pytest.raises(inputs.InputError, match="fix_concentration"):
        inputs.load(xml)
    with open(xml, "w") as fh:
        fh.write(text.replace("<chemistry>", "<chemistry><fix_concentration>no no no</fix_concentration>"))
    inputs.load(xml)
    with open(xml, "w") as fh:
        fh.write(text.replace("</simulation_mode>", "<fba_concentration_basis>total</fba_concentration_basis></simulation_mode>"))
    with pytest.raises(inputs.InputError, match="total"):
        inputs.load(xml)
    with open(xml, "w") as fh:
        fh.write(text.replace("</simulation_mode>", "<fba_concentration_basis>Free</fba_concentration_basis></simulation_mode>"))
    inputs.load(xml)


def test_several_dissolution_phases(tmp_path):
    """<phase0>, <phase1>, ... take ids 1, 2, ... in order; a grain phase maps
    its material number onto the solid or bounce-back code, a precipitate
    phase takes none and becomes the phase precipitation stamps (the last one
    declared, complab.cpp:516-531); more than four phases are refused, as is a
    grain phase on a pore material."""
    two = ("<phase0><material_number>9</material_number><substrate>2</substrate><full_density>20</full_density>"
           "<initial_fill>20</initial_fill></phase0>"
           "<phase1><material_number>0</material_number><substrate>2</substrate><full_density>8</full_density>"
           "<is_precipitate>true</is_precipitate></phase1>")
    one = XML[XML.index("    <phase0>"):XML.index("</dissolution>")]
    vals = np.random.default_rng(1).choice([0, 5, 6, 7, 9], size=6 * 4 * 3)
    path = write_case(tmp_path, vals)
    with open(path, "w") as fh:
        fh.write(XML.replace(one, two))
    p = inputs.load(path).params
    assert (p.dissol_n_phases, p.precip_phase_id) == (2, 2)
    assert list(p.dissol_material[:2]) == [P.WALL, -1] and list(p.dissol_is_precipitate[:2]) == [0, 1]
    assert list(p.dissol_full[:2]) == [20, 8] and list(p.dissol_initial[:2]) == [20, 0]
    five = "".join("<phase%d><material_number>0</material_number><substrate>2</substrate><full_density>1</full_density>"
                   "</phase%d>" % (k, k) for k in range(5))
    with open(path, "w") as fh:
        fh.write(XML.replace(one, five))
    with pytest.raises(inputs.InputError):
        inputs.load(path)
    with open(path, "w") as fh:
        fh.write(XML.replace(one, two.replace("<material_number>9</material_number>", "<material_number>5</material_number>")))
    with pytest.raises(inputs.InputError):
        inputs.load(path)
