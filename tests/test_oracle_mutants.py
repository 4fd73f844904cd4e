"""Planted mistakes in the oracle are caught by the mathematical checks.

Each mutant is the oracle rebuilt with -DORC_MUTANT=k (see the switch list in
oracle/mhd_oracle.c): a flipped sign, a transposed operand, a dropped term, a
derivative taken along the wrong axis, a wrong wall-stencil plane, a wrong
BCON coefficient, a dropped factor in the CFL, a dropped or misplaced metric
term of the stretched z grid, the wrong Reynolds number in the magnetic CFL
limit.  Every one must make at least one check of test_oracle_math.py
(mutants 1-8), test_oracle_ext.py (the off-default branches, 9-15 and 22) or
test_oracle_stretched.py (the z metric and the magnetic CFL limit, 16-21)
fail, otherwise those checks would not pin what they claim to pin.  Runs on
the CPU only.
"""
import pytest

import test_oracle_ext as te
import test_oracle_math as tm
import test_oracle_relax as tr
import test_oracle_stretched as ts
from oracle import oracle as om

MUTANTS = {
    1: "Joule heating cross term with the wrong sign",
    2: "Lorentz force FU with BZ in place of BY",
    3: "induction WW3 += BY*D1y(w) subtracted",
    4: "viscous heating (D1x w)^2 term dropped",
    5: "D1x(v) taken along y",
    6: "bottom-wall one-sided D2z(BX) reading the wrong plane",
    7: "izCon=1 bottom temperature with C23 in place of C13",
    8: "CFL fast speed without GAMMA",
    9: "density diffusion without the RO ghost-plane copy",
    10: "density diffusion without the plane-mean removal",
    11: "lRem mean-profile term without DKAPA",
    12: "lShr resetting FT with the advection term",
    13: "heat loss times RO instead of 1/RO",
    14: "plume Gaussian with the heat-loss exponent",
    15: "lRem CFL limit with RKAPA in place of 1 + RKAPA",
    16: "D2z without its d2zzdz2 metric term",
    17: "top-wall one-sided D2z(BX) with dzzdz in place of d2zzdz2",
    18: "magnetic CFL limit with RE in place of RM",
    19: "D1z taking the metric of the next plane",
    20: "D2z with dzzdz not squared",
    21: "top-wall one-sided D1z(RW) with the metric of plane 2",
    22: "lShr with lRem updating TTM as without lShr",
    23: "RLAX convective flux of TT instead of TT - TTM",
    24: "RLAX ALPHA without the 1/(1+4|ALPHA|) limiter",
    25: "RLAX rank ghost planes with the real FCONM",
    26: "RLAX iTCon 1 sum subtracting each rank once",
    27: "RLAX TFAC steps of DT/TOFF",
}


def _checks():
    yield lambda: tm.test_mhd_fluxes_equal_continuum_for_polynomial_fields(1)
    yield tm.test_wall_one_sided_stencils
    yield tm.test_d2z_one_sided_is_exact_for_cubics
    yield tm.test_conservative_fluxes_telescope
    for iz, it in [(0, 0), (1, 1), (1, 3), (0, 4)]:
        yield lambda iz=iz, it=it: tm.test_bcon_variants(iz, it)
    yield lambda: tm.test_cfl_advective_limit(False)
    yield lambda: tm.test_cfl_advective_limit(True)
    for dh, dv in [(0.7, 0.0), (0.0, 0.9), (0.7, 0.9)]:
        yield lambda dh=dh, dv=dv: te.test_density_diffusion_iD1(dh, dv)
    yield te.test_density_diffusion_iD2_weights_are_gaussians
    yield te.test_lrem_conduction_splits_mean_and_fluctuation
    yield te.test_lrem_cfl_takes_one_plus_rkapa
    yield lambda: te.test_lrem_walls_keep_a_linear_profile(1)
    yield lambda: te.test_lrem_walls_keep_a_linear_profile(4)
    yield te.test_lshr_carries_ft_and_keeps_only_conduction
    yield te.test_lshr_walls_move
    yield te.test_lshr_with_lrem_conducts_the_pressure
    yield te.test_heat_loss_gaussian_and_amplitude
    yield te.test_heat_loss_without_time_factor
    yield lambda: te.test_plume_wall(0)
    yield lambda: te.test_plume_wall(1)
    yield lambda: ts.test_every_flux_equals_its_continuum_value_on_a_mapped_grid(1)
    yield ts.test_density_varying_along_the_map
    yield lambda: ts.test_wall_planes_on_a_mapped_grid(0)
    yield lambda: ts.test_wall_planes_on_a_mapped_grid(1)
    yield ts.test_z_derivatives_converge_on_the_tanh_grid
    yield ts.test_wall_d2z_converges_on_the_tanh_grid
    yield lambda: ts.test_cfl_limits_on_the_tanh_grid(25.0)
    yield lambda: ts.test_cfl_limits_on_the_tanh_grid(1e-3)
    yield lambda: tr.test_correct_is_alpha_times_height_at_rest(3)
    yield tr.test_convective_flux_mean
    yield lambda: tr.test_rank_boundaries_take_no_convective_flux(2)
    yield lambda: tr.test_itcon1_subtracts_each_rank_end_twice(2)
    yield tr.test_tfac_steps_and_cap


def test_clean_oracle_passes_every_check():
    for chk in _checks():
        chk()


@pytest.mark.parametrize("k", sorted(MUTANTS), ids=[MUTANTS[k] for k in sorted(MUTANTS)])
def test_planted_mistake_is_caught(k, tmp_path, monkeypatch):
    path = om.build_variant(str(tmp_path / f"libmut{k}.so"), [f"ORC_MUTANT={k}"])
    monkeypatch.setattr(om, "_lib", om.load(path))
    caught = 0
    for chk in _checks():
        try:
            chk()
        except AssertionError:
            caught += 1
    assert caught >= 1, MUTANTS[k]
