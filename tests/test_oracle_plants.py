"""Each pin of the oracle must catch the mistake it is there for.

Copies of the oracle are compiled with -DORC_PLANT=n, each carrying one
plausible misreading of the reference (the ORC_PLANT blocks of
oracle/cgfd_oracle.c), and the pin named for it must fail on that copy while
it passes on the oracle itself (tests/test_oracle_pins*.py).  CPU only.
"""
import pytest

import oracle
import test_oracle_pins as pins_flat
import test_oracle_pins_curvilinear as pins_curv

# plant number -> (what is planted, the pin that must catch it)
PLANTS = {
    1: ("zeta_x and zeta_y swapped in the zeta term of the Vx rate (calculate.cu:4880)",
        pins_curv.pin_affine_linear_stage),
    2: ("eta_x and eta_y swapped in the eta term of the Txy rate (calculate.cu:4915)",
        pins_curv.pin_affine_linear_stage),
    3: ("VelPDcoeff: entry (0,1) of the xi block transposed (calculate.cu:3370)",
        pins_curv.pin_closure_matrices),
    4: ("VelPDcoeff: PML ratio Bx/Bz instead of Bz/Bx (calculate.cu:3338)",
        pins_curv.pin_closure_matrices),
    5: ("traction image: even image above the surface (calculate.cu:5507-5510)",
        pins_curv.pin_free_surface_plane_waves),
    6: ("traction image: Jacobian dropped from the zeta traction vector (calculate.cu:5495)",
        pins_curv.pin_free_surface_plane_waves),
    7: ("LoadMoment: Jacobian dropped (calculate.cu:6230)",
        pins_curv.pin_source_magnitudes),
    8: ("LoadForce: no doubling on the surface row (calculate.cu:6168)",
        pins_curv.pin_source_magnitudes),
    9: ("F/B schedule: z signs of step phases 1 and 3 follow the x pattern (simulation.cc:377-413, 605-641)",
        pins_curv.pin_fb_schedule),
    10: ("surface auxiliary term without the Bx factor (calculate.cu:5021-5031)",
         pins_curv.pin_affine_linear_stage),
    11: ("the eta part of the surface closure routed into the x auxiliary (calculate.cu:5021-5031)",
         pins_curv.pin_affine_linear_stage),
    12: ("y auxiliary rate of Txy scaled by the x profile (calculate.cu:5052)",
         pins_curv.pin_affine_linear_stage),
    13: ("two-point closure row: forward difference with its sign flipped (typenew.h:86-88)",
         pins_curv.pin_affine_linear_stage),
    14: ("2-4 closure row: backward difference divided by |h| (typenew.h:82-84)",
         pins_curv.pin_fb_schedule),
    15: ("DRP: DA2 and DA3 exchanged in the backward sum (typenew.h:59)",
         lambda lib: pins_flat.pin_linear_stress_first_moment(False, lib)),
    16: ("RK4: second and fourth stage weights exchanged (calculate.cu:13)",
         pins_flat.pin_rk4_amplification),
    17: ("hybrid grid: lambda in place of lam2mu on the zz diagonal of the Cartesian branch (calculate.cu:3940)",
         pins_curv.pin_hybrid_cartesian_branch),
    18: ("SimpleCar: lam2mu in place of lambda in the eta term of Txx (calculate.cu:4381)",
         pins_curv.pin_simplecar_is_the_identity_xi_eta_metrics),
    19: ("SimpleCar: the x-row stress in CalTIMGmix's eta traction vector (calculate.cu:5735)",
         pins_curv.pin_simplecar_is_the_identity_xi_eta_metrics),
}


@pytest.fixture(scope="module")
def plant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_plants"))


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_pin_catches_planted_mistake(plant, plant_dir):
    what, pin = PLANTS[plant]
    lib = oracle.load_planted(plant, plant_dir, double=True)
    with pytest.raises(AssertionError):
        pin(lib)
