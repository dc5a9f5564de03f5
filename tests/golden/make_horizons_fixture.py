"""Writes horizons_de441_1day.json, the solar-system ephemeris fixture of the
reference's system test, from the reference's own test sources (read as text;
nothing of the reference is compiled, imported or run).

Data taken:
- tests/support/solar_system_reference_fixture.cpp:12-45 (the table kBodies):
  for the Sun and the nine planets (Pluto included) in SPK-ID order, GM in
  km^3/s^2 and the barycentric ICRF state (x, y, z in m; vx, vy, vz in m/s)
  at JD TDB 2461278.5 (epoch A) and 2461279.5 (epoch B), NASA/JPL HORIZONS
  DE441; the particle mass is GM * 1e9 / G with G = 6.67430e-11
  (solar_system_reference_fixture.cpp:88-91).
- tests/integration/cuda/solar_system_ephemeris_test.cu:141-196: 96 steps
  over 86,400 s, Simulation config {timeStep, theta 0, softening 0}, the
  body names and the per-body position (m) and velocity (m/s) tolerances.

Run in a container that holds the reference at /root/reference:
    python tests/golden/make_horizons_fixture.py
"""

import json
import os
import re
import sys

REF = "/root/reference"
FIXTURE = os.path.join(REF, "tests/support/solar_system_reference_fixture.cpp")
TEST = os.path.join(REF, "tests/integration/cuda/solar_system_ephemeris_test.cu")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "horizons_de441_1day.json")

NUM = r"[-+]?\d[\d']*\.?\d*(?:[eE][-+]?\d+)?"


def numbers(text):
    return [float(t.replace("'", "")) for t in re.findall(NUM, text)]


def body_table(src):
    start = src.index("kBodies{{")
    end = src.index("}};", start)
    vals = numbers(src[start + len("kBodies{{"):end])
    if len(vals) != 10 * 13:
        raise SystemExit(f"expected 130 numbers in kBodies, found {len(vals)}")
    bodies = []
    for b in range(10):
        v = vals[13 * b:13 * (b + 1)]
        bodies.append({"gm_km3_s2": v[0], "epoch_a": v[1:7], "epoch_b": v[7:13]})
    return bodies


def array_after(src, name):
    i = src.index(name)
    j = src.index("{", i)
    k = src.index("}", j)
    return src[j + 1:k]


def main():
    src = open(FIXTURE).read()
    test = open(TEST).read()
    bodies = body_table(src)
    g = numbers(re.search(r"kGravitationalConstant\s*=\s*(" + NUM + ")", src).group(1))[0]
    names = re.findall(r'"([A-Za-z]+)"', array_after(test, "bodyNames"))
    pos_tol = numbers(array_after(test, "positionToleranceMeters"))
    vel_tol = numbers(array_after(test, "velocityToleranceMetersPerSecond"))
    steps = int(numbers(re.search(r"steps\s*=\s*(\d+)", test).group(1))[0])
    duration = numbers(re.search(r"durationSeconds\s*=\s*(" + NUM + ")", test).group(1))[0]
    if not (len(names) == len(pos_tol) == len(vel_tol) == len(bodies) == 10):
        raise SystemExit("inconsistent body counts")
    out = {
        "source": "NASA/JPL HORIZONS DE441 geometric barycentric ICRF vectors, SI units, "
                  "epoch A JD TDB 2461278.5, epoch B JD TDB 2461279.5 "
                  "(reference tests/support/solar_system_reference_fixture.cpp:12-45; "
                  "steps, config and tolerances from "
                  "tests/integration/cuda/solar_system_ephemeris_test.cu:141-196)",
        "gravitational_constant": g,
        "steps": steps,
        "duration_s": duration,
        "theta": 0.0,
        "softening": 0.0,
        "names": names,
        "position_tolerance_m": pos_tol,
        "velocity_tolerance_m_s": vel_tol,
        "bodies": bodies,
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    sys.exit(main())
