"""The reference's system tests on the oracle (CPU): the HORIZONS one-day
ephemeris, the SI circular two-body orbit, and the two-body leapfrog
convergence and closure, each with the reference's own bounds.  These pin
the oracle's Simulation loop (bbox, keys, tree, theta = 0 direct walk,
leapfrog) against the reference's golden data; the GPU tests run the same
cases through the library."""

import numpy as np
import pytest

import system_cases as S


@pytest.mark.parametrize("bits", [10, 21])
def test_oracle_horizons_one_day(orc, bits):
    fx, a, b = S.horizons()
    sim = S.OracleSimulation(orc, a, S.ephemeris_time_step(fx), fx["theta"], fx["softening"],
                             G=S.SI_G, bits=bits)
    sim.initialize()
    S.run_ephemeris(sim, fx, b)


def test_oracle_horizons_fixture_is_the_published_table():
    # spot values of the reference's table (solar_system_reference_fixture.cpp:25-44)
    fx, a, b = S.horizons()
    assert fx["names"][3] == "Earth" and fx["steps"] == 96 and fx["duration_s"] == 86400.0
    assert a["x"][3] == 133801656387.7991 and b["vz"][9] == -708.5605640010033
    assert fx["bodies"][0]["gm_km3_s2"] == 132712440041.27942
    assert a["m"][0] == 132712440041.27942 * 1.0e9 / 6.67430e-11
    # Sun's mass from GM / G (about 1.989e30 kg)
    assert abs(a["m"][0] / 1.98847e30 - 1.0) < 1e-4


def test_oracle_si_two_body_orbit(orc):
    p, period, dt = S.si_two_body()
    sim = S.OracleSimulation(orc, p, dt, 0.5, 0.0, G=S.SI_G, bits=10)
    sim.initialize()
    S.run_si_two_body(sim, p, period)


def test_oracle_leapfrog_second_order(orc):
    errs = []
    for k in S.QUARTER_STEPS:
        sim = S.OracleSimulation(orc, S.unit_two_body(), (S.ORBIT_PERIOD / 4.0) / k, 0.5, 0.0,
                                 G=1.0, bits=10)
        sim.initialize()
        sim.run(k)
        errs.append(S.quarter_orbit_error(sim.state()))
    S.check_convergence(errs)


def test_oracle_one_period_closure(orc):
    p = S.unit_two_body()
    sim = S.OracleSimulation(orc, p, S.ORBIT_PERIOD / 128, 0.5, 0.0, G=1.0, bits=10)
    sim.initialize()
    S.run_one_period(sim, p)


def test_oracle_ephemeris_catches_a_planted_sign_error(orc):
    # The ephemeris bounds are tight enough to fail a plausible mistake: the
    # oracle run with the velocities' sign flipped for one body misses.
    fx, a, b = S.horizons()
    a["vy"][1] = -a["vy"][1]
    sim = S.OracleSimulation(orc, a, S.ephemeris_time_step(fx), 0.0, 0.0, G=S.SI_G, bits=10)
    sim.initialize()
    with pytest.raises(AssertionError):
        S.run_ephemeris(sim, fx, b)
    # and a wrong time step (one step short) misses as well
    fx2, a2, b2 = S.horizons()
    sim = S.OracleSimulation(orc, a2, S.ephemeris_time_step(fx2), 0.0, 0.0, G=S.SI_G, bits=10)
    sim.initialize()
    sim.run(fx2["steps"] - 1)
    pos, _ = S.ephemeris_errors(sim.state(), b2)
    assert np.any(pos > np.array(fx2["position_tolerance_m"]))
