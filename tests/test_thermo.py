"""The thermodynamic gate of the biotic rates: the oracle's factor against the
checks of the reference's tests/test_thermo.cpp (the three composition
regimes by hand, the two limits a sign error breaks, the exact zero at the
threshold, the activity floor, chi), the gate's effect on the rate vector, and
the .thm reader (a file written here; the reference's example 19 where the
reference tree is present)."""
import math
import os

import numpy as np
import pytest

from complab_amd import inputs
from complab_amd import params as P
import oracle as O

REF = "/root/reference"
RGAS = 8.314462618e-3
AOM = """# energetics of anaerobic methane oxidation
provenance He et al. 2021   # a comment after the value
version 1
temperature 277.15
energy_units kJ
concentration_scale 1
reaction AOM
  microbe   ANME
  stoich    CH4 -1  SO4 -1  HS 1  HCO3 1
  dG0       -33.12
  atp       0.25
  dGatp     50.0
  chi       1.0
  cmin      1e-9
"""
SPECIES = ["CH4", "SO4", "HS", "HCO3"]


def gated(dG0=-33.12, chi=1.0, dgloss=0.0, scale=1.0):
    p = P.default_params(6, 3, 3, 4)
    p.thm_on = 1
    p.thm_rt = RGAS * 277.15
    p.thm_dg0 = dG0
    p.thm_dgloss = dgloss
    p.thm_threshold = 0.25 * 50.0
    p.thm_chi = chi
    p.thm_cmin = 1e-9
    p.thm_conc_scale = scale
    for s, v in enumerate([-1.0, -1.0, 1.0, 1.0]):
        p.thm_nu[s] = v
    return p


def by_hand(c, dG0=-33.12, chi=1.0):
    """Craig (2024) Eqs. 3.3, 3.4, 3.6 written out for CH4 + SO4 -> HS + HCO3."""
    a = [max(x, 1e-9) for x in c]
    dG = dG0 + RGAS * 277.15 * math.log((a[2] * a[3]) / (a[0] * a[1]))
    f = dG + 0.25 * 50.0
    return dG, max(0.0, 1.0 - math.exp(f / (chi * RGAS * 277.15)))


@pytest.mark.parametrize("c", [[2.0e-3, 2.0e-2, 1.0e-5, 1.0e-4],      # rim: substrate rich
                               [6.0e-5, 5.0e-4, 6.0e-3, 2.0e-2],      # interior: products built up
                               [1.0e-6, 1.0e-5, 3.0e-2, 5.0e-2]])     # dead zone: past the threshold
def test_three_regimes_by_hand(c):
    F, dG = O.thermo_factor(gated(), c)
    dGh, Fh = by_hand(c)
    assert dG == pytest.approx(dGh, rel=1e-12, abs=1e-12)
    assert F == pytest.approx(Fh, rel=1e-10, abs=1e-14)


def test_the_limits_a_sign_error_breaks():
    F, dG = O.thermo_factor(gated(), [1.0, 1.0, 1e-9, 1e-9])          # strongly exergonic
    assert dG < -50.0 and F == pytest.approx(1.0, abs=1e-9)
    F, dG = O.thermo_factor(gated(), [1e-9, 1e-9, 1.0, 1.0])          # endergonic
    assert dG > 0.0 and F == 0.0
    F, dG = O.thermo_factor(gated(dG0=-12.5), [1.0, 1.0, 1.0, 1.0])   # ln Q = 0: dG = dG0 = -m dG_ATP
    assert dG == -12.5 and F == 0.0
    F, _ = O.thermo_factor(gated(dG0=-12.5 - 1e-6), [1.0, 1.0, 1.0, 1.0])
    assert F > 0.0


def test_zero_concentration_floors_at_cmin():
    F0, dG0 = O.thermo_factor(gated(), [0.0, 0.0, 1e-3, 1e-3])
    F1, dG1 = O.thermo_factor(gated(), [1e-9, 1e-9, 1e-3, 1e-3])
    assert math.isfinite(dG0) and dG0 == dG1 and F0 == 0.0


def test_chi_stretches_without_moving_the_threshold():
    c = [1e-3, 1e-3, 1e-3, 1e-3]
    Fa, _ = O.thermo_factor(gated(chi=1.0), c)
    Fb, _ = O.thermo_factor(gated(chi=4.0), c)
    assert 0 < Fb < Fa
    for chi in (1.0, 4.0):
        assert O.thermo_factor(gated(dG0=-12.5, chi=chi), [1.0] * 4)[0] == 0.0


def test_loss_and_scale_enter_dG():
    c = [2.0e-3, 2.0e-2, 1.0e-5, 1.0e-4]
    _, dG = O.thermo_factor(gated(), c)
    _, dGl = O.thermo_factor(gated(dgloss=25.7), c)
    assert dGl - dG == pytest.approx(25.7, rel=1e-12)
    _, dGs = O.thermo_factor(gated(scale=1e-3), c)                    # mmol/L lattices: ln Q shifts by 0
    assert dGs == pytest.approx(dG, rel=1e-12)                        # (two reactants, two products)


def test_gate_scales_the_whole_biotic_rate_vector():
    p = gated()
    p.n_microbes = 1
    p.bio_law = P.BIO_DUAL_MONOD
    p.donor, p.acceptor = 0, 1
    p.mu_max[0], p.ks[0], p.ke[0], p.kdecay[0] = 1e-3, 0.1, 0.1, 1e-5
    for s, v in enumerate([-1.0, -1.0, 1.0, 1.0]):
        p.bio_stoich[0][s] = v
    c = [6.0e-5, 5.0e-4, 6.0e-3, 2.0e-2]
    sr, br, _ = O.rates(p, c, [3.0], 2)
    p.thm_on = 0
    sr0, br0, _ = O.rates(p, c, [3.0], 2)
    F, _ = O.thermo_factor(gated(), c)
    assert 0 < F < 1
    assert np.array_equal(sr, sr0 * F) and np.array_equal(br, br0 * F)


def test_thm_reader_binds_and_refuses(tmp_path):
    f = tmp_path / "aom.thm"
    f.write_text(AOM)
    M = inputs.read_thm(str(f))
    assert M.provenance == "He et al. 2021" and M.temperature == 277.15 and M.rt() == RGAS * 277.15
    r = M.reactions[0]
    assert (r.name, r.microbe, r.species, r.nu) == ("AOM", "ANME", SPECIES, [-1, -1, 1, 1])
    assert (r.dG0, r.atp, r.dGatp, r.chi, r.cmin, r.dGloss) == (-33.12, 0.25, 50.0, 1.0, 1e-9, 0.0)
    p = P.default_params(6, 3, 3, 5)
    inputs.apply_thm(p, M, SPECIES + ["X"], ["ANME"])
    assert p.thm_on == 1 and p.thm_threshold == 12.5 and list(p.thm_nu[:5]) == [-1, -1, 1, 1, 0]
    f.write_text(AOM.replace("energy_units kJ", "energy_units J").replace("-33.12", "-33120"))
    assert inputs.read_thm(str(f)).reactions[0].dG0 == pytest.approx(-33.12)
    bad = {"unknown keyword": AOM + "  dGwrong 1\n",
           "odd stoich": AOM.replace("HCO3 1", "HCO3"),
           "zero cmin": AOM.replace("cmin      1e-9", "cmin      0"),
           "negative loss": AOM + "  dGloss -1\n",
           "key before a block": "dG0 -1\n" + AOM}
    for what, text in bad.items():
        f.write_text(text)
        with pytest.raises(inputs.InputError):
            inputs.read_thm(str(f))
    f.write_text(AOM)
    with pytest.raises(inputs.InputError):                             # a species this run lacks
        inputs.apply_thm(P.default_params(6, 3, 3, 3), inputs.read_thm(str(f)), SPECIES[:3], ["ANME"])
    with pytest.raises(inputs.InputError):                             # an organism other than microbe0
        inputs.apply_thm(P.default_params(6, 3, 3, 4), inputs.read_thm(str(f)), SPECIES, ["SRB", "ANME"])
    f.write_text(AOM + AOM[AOM.index("reaction"):].replace("reaction AOM", "reaction AOM2"))
    with pytest.raises(inputs.InputError):                             # two blocks on the kinetics path
        inputs.apply_thm(P.default_params(6, 3, 3, 4), inputs.read_thm(str(f)), SPECIES, ["ANME"])


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
def test_reference_example_19_gate():
    r = inputs.load(os.path.join(REF, "examples", "19_thermodynamic_gate", "CompLaB.xml"))
    p = r.params
    assert r.species == SPECIES and r.microbes == ["ANME"]
    assert p.thm_on == 1 and p.thm_rt == pytest.approx(RGAS * 277.15, rel=1e-15)
    assert (p.thm_dg0, p.thm_dgloss, p.thm_threshold, p.thm_chi, p.thm_cmin) == (-33.12, 25.7, 12.5, 1.0, 1e-9)
    assert list(p.thm_nu[:4]) == [-1, -1, 1, 1]


def test_energy_snapshot_files(tmp_path):
    """dG_<microbe>/FT_<microbe> (computeEnergyField3D): at every open voxel of
    x = 1..nx-2 the gate's dG and F_T at the standing concentrations (negative
    ones read 0), 0 elsewhere; the files carry them under "DeltaG" and "F_T"."""
    from complab_amd import configs, output
    w = configs.config2(scale=0.08)
    p = w.params
    p.thm_on, p.thm_microbe = 1, 0
    p.thm_rt = RGAS * 298.15
    p.thm_dg0, p.thm_threshold = -20.0, 12.5
    for s, v in enumerate([-1.0, -1.0, 1.0, 1.0]):
        p.thm_nu[s] = v
    p.flow_on = 0
    o = O.OracleCase(p, w.geometry)
    o.set_omegas([1.2] * 4, [1.1] * 4, 1.0)
    o.run_steps(20)
    snaps = output.Snapshots(str(tmp_path), ["S%d" % s for s in range(4)], ["Bug"], energy=[(0, "Bug")])
    snaps.write(o, 20)
    g, _ = output.read_vti(str(tmp_path / "dG_Bug_0000020.vti"))
    f, _ = output.read_vti(str(tmp_path / "FT_Bug_0000020.vti"))
    mask = o.read(P.FIELD_MASK)[:, :, 1:-1]
    conc = [o.read(P.FIELD_CONC, s)[:, :, 1:-1] for s in range(4)]
    open_ = mask >= P.PORE
    assert np.all(g["DeltaG"][~open_] == 0.0) and np.all(f["F_T"][~open_] == 0.0)
    idx = np.argwhere(open_)
    for k in np.linspace(0, len(idx) - 1, 25).astype(int):
        z, y, x = idx[k]
        c = [max(conc[s][z, y, x], 0.0) for s in range(4)]
        F, dG = O.thermo_factor(p, c)
        lnQ = sum(v * math.log(max(ci, p.thm_cmin)) for v, ci in zip([-1, -1, 1, 1], c))
        assert dG == pytest.approx(p.thm_dg0 + p.thm_rt * lnQ, rel=1e-12)     # the equation, by hand
        assert g["DeltaG"][z, y, x] == dG and f["F_T"][z, y, x] == F
