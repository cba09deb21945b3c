"""The transport pull with its bounce-back store forced on (CLB_TUNE_ADE_BB),
against the oracle, on the cases of the parity tests whose small lattices
have too many closed links per voxel for the automatic choice to take the
store: every kernel that changes populations between launches (split
kinetics, reopening, field writes, the first launch, restarts) must keep the
store equal to the own slots it mirrors, and the fields then equal the
oracle's as in the plain runs.  Cases with equilibrium are left as they are
(the knob refuses the store there)."""
import numpy as np
import pytest

import complab_amd as clb
from complab_amd import configs
from complab_amd import params as P

import test_gpu_checkpoint as TC
import test_gpu_dissolution as TD
import test_gpu_edges as TE
import test_gpu_output as TO
import test_gpu_parity as TP
import test_gpu_symbolic as TS

pytestmark = pytest.mark.gpu


@pytest.fixture
def store_on(monkeypatch):
    """Every device case created in the test takes the store (unless it runs
    equilibrium); counts the cases that did."""
    made = []
    init = clb.Case.__init__

    def forced(self, ctx, params, geometry, biomass0=None):
        init(self, ctx, params, geometry, biomass0)
        if not params.eq_on:
            self.set_tuning(P.TUNE_ADE_BB, 1)
            made.append(self)
    monkeypatch.setattr(clb.Case, "__init__", forced)
    return made


def test_knob_rules(ctx):
    w = configs.config2(scale=0.1)
    c = clb.Case(ctx, w.params, w.geometry)
    c.set_tuning(P.TUNE_ADE_BB, 1)
    c.set_tuning(P.TUNE_ADE_BB, 0)
    with pytest.raises(clb.ClbError):
        c.set_tuning(P.TUNE_ADE_BB, 2)
    c.set_tuning(P.TUNE_ADE_BB, 1)
    c.ade_steps(1)
    with pytest.raises(clb.ClbError):                      # after the first transport launch
        c.set_tuning(P.TUNE_ADE_BB, 0)
    c.close()
    e = configs.config3(scale=1 / 16, precision=P.FP64)
    c = clb.Case(ctx, e.params, e.geometry)
    with pytest.raises(clb.ClbError):                      # equilibrium does not keep the store in step
        c.set_tuning(P.TUNE_ADE_BB, 1)
    c.set_tuning(P.TUNE_ADE_BB, 0)
    c.close()


def test_biotic_ca_with_store(ctx, store_on):
    TP.test_biotic_coupled_fp64(ctx)
    assert store_on


def test_ca_spill_with_store(ctx, store_on):
    TP.test_ca_spill_invariants_and_parity(ctx)
    assert store_on


def test_flow_and_transport_with_store(ctx, store_on):
    TP.test_flow_and_transport_fp64(ctx)
    assert store_on


def test_precipitation_reopening_with_store(ctx, store_on):
    TP.test_precipitation_parity(ctx)
    assert store_on


def test_fp32_with_store(ctx, store_on):
    TP.test_fp32_close_to_fp64(ctx)
    assert store_on


def test_thermodynamic_gate_with_store(ctx, store_on):
    TP.test_thermodynamic_gate_biotic_fp64(ctx)
    assert store_on


def test_restart_against_oracle_with_store(ctx, store_on):
    TC.test_restarted_biotic_case_matches_uninterrupted_oracle(ctx)
    assert store_on


def test_restart_bit_identical_with_store(ctx, store_on):
    TC.test_biotic_ca_case_resumes_bit_identically(ctx)
    assert store_on


def test_dissolution_with_store(ctx, store_on):
    TD.test_reopening_with_flow_resolve(ctx)
    TD.test_two_grain_phases_match_oracle(ctx)
    assert store_on


def test_symbolic_laws_with_store(ctx, store_on, tmp_path):
    TS.test_symbolic_law_matches_oracle(ctx, tmp_path)
    TS.test_abiotic_symbolic_beside_compiled_biotic_law(ctx, tmp_path)
    assert store_on


def test_rate_capture_with_store(ctx, store_on):
    TO.test_rate_capture_matches_oracle(ctx)
    assert store_on


def test_edges_with_store(ctx, store_on):
    TE.test_closed_x_planes(ctx)
    TE.test_eight_mobile_species(ctx)
    TE.test_two_microbes(ctx)
    assert store_on
