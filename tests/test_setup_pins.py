"""The set-up formulas of complab.cpp:1220-1252 (ADE relaxation and time step)
and the tau screen of complab.cpp:1315-1378, pinned to values worked by hand.

Worked example (three species of different diffusivity):
  Pe = 2, mean pore |u| = 0.01, L = 50 lattice units, dx = 1e-5 m
  refNu = 0.01 * 50 / 2 = 0.25;  refTau = 0.25 / (1/4) + 1/2 = 1.5
  D_pore = 1e-9, 1.5e-9, 0.5e-9  ->  nu = 0.25, 0.375, 0.125
                                   ->  tau = 1.5, 2.0, 1.0
  D_biofilm = 0.2e-9, 0.3e-9, 0.1e-9 -> nu = 0.05, 0.075, 0.025
                                   ->  tau = 0.7, 0.8, 0.6
  dt = refNu dx^2 / D_pore[0] = 0.25 * 1e-10 / 1e-9 = 0.025 s
"""
import pytest

from complab_amd import driver
from complab_amd import params as P


def setup_case(d_pore, d_bio, mobile=None):
    p = P.default_params(8, 4, 4, len(d_pore))
    if mobile is not None:
        for s, m in enumerate(mobile):
            p.mobile[s] = m
    fs = driver.FlowSetup(pe=2.0, char_length=50.0, delta_p=1e-4)
    ts = driver.TransportSetup(d_pore=d_pore, d_biofilm=d_bio, dx=1e-5)
    return p, fs, ts


def test_ade_relaxation_hand_values():
    p, fs, ts = setup_case([1e-9, 1.5e-9, 0.5e-9], [0.2e-9, 0.3e-9, 0.1e-9])
    rep = driver.SetupReport()
    omp, omb, dt = driver.ade_relaxation(p, 0.01, fs, ts, rep)
    assert rep.ref_tau == pytest.approx(1.5, rel=1e-15)
    assert dt == pytest.approx(0.025, rel=1e-14)
    assert omp == pytest.approx([1 / 1.5, 0.5, 1.0], rel=1e-14)
    assert omb == pytest.approx([1 / 0.7, 1 / 0.8, 1 / 0.6], rel=1e-14)
    assert rep.tau_notes == []                 # 0.6 .. 2.0: nothing to report


def test_no_flow_takes_the_lattice_tau():
    """Pe = 0: refTau = <tau>, refNu = cs^2 (tau - 1/2) (complab.cpp:1227)."""
    p, fs, ts = setup_case([2e-9, 1e-9], [2e-9, 1e-9])
    fs.pe = 0.0
    ts.tau_if_no_flow = 0.9
    rep = driver.SetupReport()
    omp, omb, dt = driver.ade_relaxation(p, 0.0, fs, ts, rep)
    assert dt == pytest.approx(0.1 * 1e-10 / 2e-9, rel=1e-14)      # nu = 0.25 * 0.4
    assert omp == pytest.approx([1 / 0.9, 1 / 0.7], rel=1e-14)      # nu 0.05 -> tau 0.7
    assert omb == pytest.approx([1 / 0.9, 1 / 0.7], rel=1e-14)


def test_tau_screen_refuses_below_0_51():
    """A species 111 times less mobile than substrate 0: nu = 0.25 * 0.009 =
    0.00225, tau = 0.509 < 0.51, and the reference terminates."""
    p, fs, ts = setup_case([1e-9, 0.9e-11], [1e-9, 1e-9])
    with pytest.raises(driver.TauError, match="0.509"):
        driver.ade_relaxation(p, 0.01, fs, ts, driver.SetupReport())
    p, fs, ts = setup_case([1e-9, 1e-9], [1e-9, 0.9e-11])          # the biofilm tau is screened too
    with pytest.raises(driver.TauError, match="biofilm"):
        driver.ade_relaxation(p, 0.01, fs, ts, driver.SetupReport())


def test_tau_screen_skips_immobile_species():
    """complab.cpp:1329-1334: an immobile species never streams, so its tau is
    not checked (examples 13-15 declare D = 0 for the mineral)."""
    p, fs, ts = setup_case([1e-9, 0.0], [1e-9, 0.0], mobile=[1, 0])
    rep = driver.SetupReport()
    driver.ade_relaxation(p, 0.01, fs, ts, rep)
    assert rep.tau_notes == []


def test_tau_screen_warnings():
    """tau in [0.51, 0.55) rings a little, tau > 2 over-diffuses: both run,
    both are reported (complab.cpp:1342-1352)."""
    # nu = 0.25 * 0.03 = 0.0075 -> tau 0.53 ; nu = 0.25 * 1.8 = 0.45 -> tau 2.3
    p, fs, ts = setup_case([1e-9, 0.03e-9, 1.8e-9], [1e-9, 1e-9, 1e-9])
    rep = driver.SetupReport()
    driver.ade_relaxation(p, 0.01, fs, ts, rep)
    assert len(rep.tau_notes) == 2
    assert any("2.3" in n and "above 2" in n for n in rep.tau_notes)
    assert any("0.53" in n and "close to 0.5" in n for n in rep.tau_notes)
