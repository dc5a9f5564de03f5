"""System cases of the reference's integration tests, shared by the oracle
(CPU) tests and the GPU tests.  A case drives any object with the
Simulation interface (initialize, step, run, state(), step_number(),
time()) through the reference's scenario and checks the reference's bounds.

- The HORIZONS one-day ephemeris: ten bodies (Sun, Mercury ... Pluto) from
  the DE441 state at epoch A, 96 steps of 900 s with theta 0, softening 0 and
  the SI gravitational constant, compared with the epoch-B state under
  per-body position and velocity tolerances
  (tests/integration/cuda/solar_system_ephemeris_test.cu:141-196; data in
  golden/horizons_de441_1day.json, written by golden/make_horizons_fixture.py).
- The circular two-body orbit in SI units through a Simulation loaded from a
  particle file: m = 1e10 kg each, separation 1 m, 128 steps per period,
  theta 0.5, softening 0; energy error < 5e-6, separation error < 2e-3 and
  centre-of-mass drift < 1e-12 at every step, closure within 3e-3
  (tests/integration/cuda/simulation_test.cu:195-283).
- The two-body leapfrog convergence and one-period closure at G = 1:
  m = 1, 1 at separation 1, v = 1/sqrt(2); a quarter orbit at 32, 64, 128
  and 256 steps must improve monotonically with order log2(e(dt)/e(dt/2)) in
  (1.8, 2.2) for positions and velocities; one period at 128 steps keeps the
  separation within 2e-3 and the energy within 5e-6, and returns within
  3e-3 (positions) and 4e-3 (velocities)
  (tests/integration/cuda/two_body_orbit_test.cu:320-880).
"""

import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HORIZONS = os.path.join(HERE, "golden", "horizons_de441_1day.json")
SI_G = 6.67430e-11


# ------------------------------------------------------------- ephemeris
def horizons():
    """The fixture and the epoch A and epoch B particle sets (host arrays)."""
    fx = json.load(open(HORIZONS))
    g = fx["gravitational_constant"]
    # solar_system_reference_fixture.cpp:88-91: mass = GM [km^3/s^2] * 1e9 / G
    m = np.array([b["gm_km3_s2"] * 1.0e9 / g for b in fx["bodies"]])

    def state(epoch):
        s = np.array([b[epoch] for b in fx["bodies"]])
        return {"m": m.copy(), "x": s[:, 0].copy(), "y": s[:, 1].copy(), "z": s[:, 2].copy(),
                "vx": s[:, 3].copy(), "vy": s[:, 4].copy(), "vz": s[:, 5].copy()}

    return fx, state("epoch_a"), state("epoch_b")


def ephemeris_time_step(fx):
    return fx["duration_s"] / fx["steps"]


def ephemeris_errors(st, ref):
    """Per-body 3-D position and velocity errors against the epoch-B state."""
    pos = np.sqrt((st["x"] - ref["x"]) ** 2 + (st["y"] - ref["y"]) ** 2 +
                  (st["z"] - ref["z"]) ** 2)
    vel = np.sqrt((st["vx"] - ref["vx"]) ** 2 + (st["vy"] - ref["vy"]) ** 2 +
                  (st["vz"] - ref["vz"]) ** 2)
    return pos, vel


def run_ephemeris(sim, fx, ref):
    """Runs the 96 steps on an initialized `sim`; returns the final state and
    the per-body errors, asserting the reference's bounds."""
    sim.run(fx["steps"])
    assert sim.step_number() == fx["steps"]
    assert sim.time() == fx["duration_s"]  # EXPECT_DOUBLE_EQ: 96 x 900 exactly
    st = sim.state()
    assert np.array_equal(st["m"], ref["m"])
    pos, vel = ephemeris_errors(st, ref)
    for k, name in enumerate(fx["names"]):
        print(f"{name:8s} position error {pos[k]:12.4f} m (bound {fx['position_tolerance_m'][k]:g}),"
              f" velocity error {vel[k]:.3e} m/s (bound {fx['velocity_tolerance_m_s'][k]:g})")
    assert np.all(pos <= np.array(fx["position_tolerance_m"])), pos
    assert np.all(vel <= np.array(fx["velocity_tolerance_m_s"])), vel
    return st, pos, vel


# --------------------------------------------------- SI circular two-body
SI_TWO_BODY_MASS = 1.0e10
SI_TWO_BODY_STEPS = 128


def si_two_body():
    """(particles, orbital period, time step) of simulation_test.cu:195-216."""
    m, sep = SI_TWO_BODY_MASS, 1.0
    v = math.sqrt(SI_G * m / (2.0 * sep))
    omega = 2.0 * v / sep
    period = 2.0 * math.acos(-1.0) / omega
    p = {"m": np.array([m, m]), "x": np.array([-0.5 * sep, 0.5 * sep]), "y": np.zeros(2),
         "z": np.zeros(2), "vx": np.zeros(2), "vy": np.array([-v, v]), "vz": np.zeros(2)}
    return p, period, period / SI_TWO_BODY_STEPS


def run_si_two_body(sim, p, period):
    """The per-step invariants and the closure of simulation_test.cu:218-282 on
    an initialized `sim`."""
    m = SI_TWO_BODY_MASS
    v = p["vy"][1]
    e0 = m * v * v - SI_G * m * m / 1.0
    max_e = max_sep = max_com = 0.0
    for _ in range(SI_TWO_BODY_STEPS):
        sim.step()
        s = sim.state()
        dx, dy, dz = s["x"][1] - s["x"][0], s["y"][1] - s["y"][0], s["z"][1] - s["z"][0]
        r = math.sqrt(dx * dx + dy * dy + dz * dz)
        ke = 0.5 * m * float((s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2).sum())
        pe = -SI_G * m * m / r
        max_e = max(max_e, abs((ke + pe - e0) / e0))
        max_sep = max(max_sep, abs(r - 1.0))
        cx, cy, cz = (0.5 * (s[f][0] + s[f][1]) for f in ("x", "y", "z"))
        max_com = max(max_com, math.sqrt(cx * cx + cy * cy + cz * cz))
    s = sim.state()
    print(f"SI two-body: energy {max_e:.3e}, separation {max_sep:.3e}, COM drift {max_com:.3e}")
    assert sim.step_number() == SI_TWO_BODY_STEPS
    assert abs(sim.time() - period) <= period * 1.0e-13
    assert max_e < 5.0e-6
    assert max_sep < 2.0e-3
    assert max_com < 1.0e-12
    for f in ("x", "y", "vx", "vy"):
        assert np.all(np.abs(s[f] - p[f]) <= 3.0e-3), f
    return max_e, max_sep, max_com


# ------------------------------------------------- two bodies at G = 1
ORBIT_V = 1.0 / math.sqrt(2.0)
ORBIT_PERIOD = 2.0 * math.acos(-1.0) / math.sqrt(2.0)


def unit_two_body():
    """m = 1, 1 at (-0.5, 0), (0.5, 0), counter-clockwise with v = 1/sqrt(2)."""
    return {"m": np.array([1.0, 1.0]), "x": np.array([-0.5, 0.5]), "y": np.zeros(2),
            "z": np.zeros(2), "vx": np.zeros(2), "vy": np.array([-ORBIT_V, ORBIT_V]),
            "vz": np.zeros(2)}


def quarter_orbit_error(s):
    """Distance of the state from the exact quarter-orbit state: particle 0 at
    (0, -0.5) moving (+v, 0), particle 1 at (0, 0.5) moving (-v, 0)."""
    pe = (s["x"][0] ** 2 + (s["y"][0] + 0.5) ** 2 + s["z"][0] ** 2 +
          s["x"][1] ** 2 + (s["y"][1] - 0.5) ** 2 + s["z"][1] ** 2)
    ve = ((s["vx"][0] - ORBIT_V) ** 2 + s["vy"][0] ** 2 + s["vz"][0] ** 2 +
          (s["vx"][1] + ORBIT_V) ** 2 + s["vy"][1] ** 2 + s["vz"][1] ** 2)
    return math.sqrt(pe), math.sqrt(ve)


QUARTER_STEPS = (32, 64, 128, 256)


def check_convergence(errors):
    """errors[k] = (position, velocity) at QUARTER_STEPS[k]."""
    orders = []
    for k in range(1, len(errors)):
        (p0, v0), (p1, v1) = errors[k - 1], errors[k]
        assert p1 < p0 and v1 < v0
        po, vo = math.log2(p0 / p1), math.log2(v0 / v1)
        orders.append((po, vo))
        assert 1.8 < po < 2.2, po
        assert 1.8 < vo < 2.2, vo
    print("leapfrog quarter-orbit errors", errors, "orders", orders)
    return orders


def unit_energy(s):
    ke = 0.5 * float((s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2).sum())
    r = math.sqrt((s["x"][1] - s["x"][0]) ** 2 + (s["y"][1] - s["y"][0]) ** 2 +
                  (s["z"][1] - s["z"][0]) ** 2)
    return ke - 1.0 / r, r


def run_one_period(sim, p0):
    """128 steps of one period on an initialized `sim` with the per-step
    invariants and closure bounds of two_body_orbit_test.cu:640-880."""
    e0, _ = unit_energy(p0)
    max_e = max_sep = 0.0
    for _ in range(128):
        sim.step()
        e, r = unit_energy(sim.state())
        max_e = max(max_e, abs((e - e0) / e0))
        max_sep = max(max_sep, abs(r - 1.0))
    s = sim.state()
    assert max_sep < 2.0e-3
    assert max_e < 5.0e-6
    for f, tol in (("x", 3e-3), ("y", 3e-3), ("vx", 4e-3), ("vy", 4e-3)):
        assert np.all(np.abs(s[f] - p0[f]) <= tol), f
    return max_e, max_sep


class OracleSimulation:
    """The oracle's Simulation loop behind the interface the cases drive."""

    def __init__(self, orc, p, dt, theta, softening, G=SI_G, bits=10):
        self.sim = orc.Sim(p, dt, theta, softening, G, bits)
        self.m = np.array(p["m"], dtype=np.float64)
        self.dt = dt
        self.steps = 0
        self.t = 0.0

    def initialize(self):
        self.sim.initialize()

    def step(self):
        self.sim.step(1)
        self.steps += 1
        self.t += self.dt  # Simulation::step adds timeStep (src/simulation.cpp:116)

    def run(self, k):
        for _ in range(k):
            self.step()

    def step_number(self):
        return self.steps

    def time(self):
        return self.t

    def state(self):
        s = self.sim.state()
        s["m"] = self.m.copy()
        return s
