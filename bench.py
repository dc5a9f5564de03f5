"""Benchmark: Barnes-Hut leapfrog steps on one MI355X.

Headline workload (the measured metric): Plummer sphere, N = 4,194,304,
theta 0.5, softening 1e-3, dt 1e-4, G = 1, 63-bit Morton keys.  One step =
kick-drift + bounding box, Morton keys, radix sort, leaf groups, Karras tree,
radix->octree plan, octree, mass/COM, walk layout, traversal, closing kick --
the reference's Simulation::step (src/simulation.cpp:100-119) -- with inputs
resident in HBM.  `value` is the mixed-FP32 pipeline's Mparticle-steps/s over
exactly --steps timed steps after --warmup steps, with the per-phase timing
events off; the phase breakdown comes from separate steps that record them.

The same JSON line carries:
  - the plain FP64 pipeline on the same workload (rate, phases, roofline)
    and the mixed/FP64 ratio;
  - accuracy of the mixed forces against the FP64 pipeline (bit-exact with
    the oracle) and against an FP64 direct sum on 8192 targets, and energy /
    momentum drift over 50 steps, at theta 0.3, 0.5 and 0.8;
  - `configs`: the other workloads of the benchmark set -- 1M uniform FP64
    (20 steps, tree checked against the oracle, 4096 direct-sum targets),
    16M clumps mixed (20 steps, phase breakdown with per-phase HBM GB/s),
    the reference's own benchmark blob at 100K and 250K particles, and the
    differentiable disk + bulge workload;
  - the traversal roofline (VALU FLOP/s against the MI355X peaks), with the
    traffic figure taken from the committed rocprofv3 counter profile;
  - the CPU oracle timed on the host on bounded samples (median of 3).

Energy: kinetic + potential, the potential from an FP64 walk over a tree
built at the state (Simulation.energy of an FP64 simulation of that state).
"""

import argparse
import json
import math
import os
import platform
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "amsc-nbody_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

FP32_VALU_PEAK_TFLOPS = 157.3
FP64_VALU_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
FLOPS_PER_INTERACTION = 20
TRAFFIC_PROFILE = os.path.join("profiles", "walk_mixed_plummer4m_traffic.md")


_T0 = time.perf_counter()


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def rel_err(a, b):
    a = np.stack(a)
    b = np.stack(b)
    r = np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))
    return float(np.sqrt((r ** 2).mean())), float(r.max())


def g4(x):
    return None if x is None else float(f"{x:.4g}")


def momentum(s):
    return np.array([math.fsum(s["m"] * s[k]) for k in ("vx", "vy", "vz")])


# Simulations no longer needed are closed at the end of the run, not
# between measurements: freeing gigabytes of device memory just before a
# timed region (on a device whose memory a previous process fragmented) has
# stalled a later device synchronisation inside the timed region by 30-50 ms.
_RETIRED = []


def retire(sim):
    _RETIRED.append(sim)


def release_retired():
    while _RETIRED:
        _RETIRED.pop().close()


def time_steps(torch, sim, steps):
    """Seconds for `steps` steps with the per-phase timing events off."""
    sim.set_timing(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def phase_steps(sim, steps=2):
    """Mean per-phase times (ms) of `steps` steps that record HIP events
    between the phases (not part of the timed steps)."""
    sim.set_timing(True)
    phases = None
    for _ in range(steps):
        sim.step()
        pt = sim.phase_times()
        if phases is None:
            phases = {k: 0.0 for k in pt}
        for k, v in pt.items():
            phases[k] += v / steps
    sim.set_timing(False)
    return {k: round(v, 4) for k, v in phases.items()}


def algorithmic_bytes(n, g, m):
    """Algorithmic HBM bytes per phase of one step: N particles, G leaf
    groups, R = 2G - 1 radix nodes, M octree nodes; 63-bit keys (8 B),
    indices 4 B, FP64 state."""
    r = 2 * g - 1
    return {
        "kick_drift_bbox": 120 * n, "encode": 36 * n, "sort": 200 * n, "gather": 84 * n,
        "groups": 8 * n + 16 * g, "karras": 8 * g + 20 * (g - 1) + 4 * r, "plan": 24 * r,
        "octree": 24 * r + 56 * m, "com": 36 * n + 100 * m, "walk_layout": 92 * m,
        "kick": 72 * n,
    }


def phase_bandwidth(phases, n, g, m):
    out = {}
    for k, b in algorithmic_bytes(n, g, m).items():
        ms = phases.get(k, 0.0)
        if ms > 0:
            out[k] = {"ms": ms, "GB_per_s": round(b / (ms * 1e-3) / 1e9, 1),
                      "frac_of_hbm_peak": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def sizes(sim):
    s = sim.structure()
    return s.n_groups, s.n_octree_nodes


def traversal_traffic():
    """HBM bytes per traversal dispatch from the committed counter profile
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, calibrated against a copy of
    known size), or None when it is absent."""
    try:
        for line in open(os.path.join(ROOT, TRAFFIC_PROFILE)):
            if line.startswith("traffic_bytes_per_dispatch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def energy_of(nb, state, dt, theta, eps, G):
    """(E, K, U) of a state: an FP64 simulation of it, initialised (tree and
    forces), then its FP64 potential walk."""
    s = nb.Simulation(state, dt, theta, eps, gravitational_constant=G,
                      precision=nb.PRECISION_FP64, bits_per_axis=21)
    s.initialize()
    k, u = s.energy()
    retire(s)
    return k + u, k, u


def direct_reference(nb, torch, p, targets, eps, G):
    dev = [torch.from_numpy(p[k]).cuda() for k in ("m", "x", "y", "z")]
    out = [t.cpu().numpy() for t in
           nb.direct_sum(*dev, torch.from_numpy(targets.view(np.int32)).cuda(), eps, G)]
    del dev
    return out


def sample_targets(n, count, seed=0):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, min(count, n), replace=False)).astype(np.uint32)


def drift_run(nb, torch, p, theta, args, steps, targets=None, exact=None):
    """A mixed simulation at `theta`: force accuracy at t = 0 (vs the FP64
    pipeline and the direct sum), then `steps` steps with energy and
    momentum drift."""
    common = dict(gravitational_constant=args.g, bits_per_axis=21)
    mixed = nb.Simulation(p, args.dt, theta, args.softening, precision=nb.PRECISION_MIXED,
                          **common)
    mixed.initialize()
    a_mixed = mixed.accelerations()
    f64 = nb.Simulation(p, args.dt, theta, args.softening, precision=nb.PRECISION_FP64, **common)
    f64.initialize()
    a64 = f64.accelerations()
    retire(f64)
    out = {}
    out["rms_vs_fp64_bh"], out["max_vs_fp64_bh"] = map(g4, rel_err(a_mixed, a64))
    if exact is not None:
        out["rms_vs_direct"], out["max_vs_direct"] = map(
            g4, rel_err([a[targets] for a in a_mixed], exact))
        out["fp64_rms_vs_direct"], out["fp64_max_vs_direct"] = map(
            g4, rel_err([a[targets] for a in a64], exact))
    s0 = mixed.state()
    e0 = energy_of(nb, s0, args.dt, theta, args.softening, args.g)
    secs = time_steps(torch, mixed, steps)
    s1 = mixed.state()
    e1 = energy_of(nb, s1, args.dt, theta, args.softening, args.g)
    retire(mixed)
    out.update({"steps": steps, "ms_per_step": round(1e3 * secs / steps, 3),
                "energy_rel_drift": g4((e1[0] - e0[0]) / abs(e0[0])),
                "momentum_drift": g4(float(np.abs(momentum(s1) - momentum(s0)).max()))})
    return out


# ------------------------------------------------------------------ legs
def headline(nb, torch, args):
    n = args.n
    p = nb.generate(args.workload, n, 42)
    if args.particle_mass is not None:
        p["m"][:] = args.particle_mass
    common = dict(gravitational_constant=args.g, bits_per_axis=21)
    targets = sample_targets(n, args.direct_targets)
    exact = direct_reference(nb, torch, p, targets, args.softening, args.g)

    mixed = nb.Simulation(p, args.dt, args.theta, args.softening, precision=nb.PRECISION_MIXED,
                          **common)
    mixed.initialize()
    a_mixed0 = mixed.accelerations()
    pp, pn, visits = mixed.count_interactions(with_visits=True)
    interactions = pp + pn
    acc = {}
    s0 = mixed.state()
    e0 = None if args.no_energy else energy_of(nb, s0, args.dt, args.theta, args.softening, args.g)
    for _ in range(args.warmup):
        # untimed; each followed by a device synchronisation, so that a
        # one-off stall of the runtime lands here rather than in the timed
        # region
        mixed.step()
        torch.cuda.synchronize()
    secs = time_steps(torch, mixed, args.steps)
    phases = phase_steps(mixed)
    # further steps to the drift span (50 steps in all by default)
    extra = max(0, args.drift_steps - args.warmup - args.steps - 2)
    for _ in range(extra):
        mixed.step()
    s1 = mixed.state()
    e1 = None if args.no_energy else energy_of(nb, s1, args.dt, args.theta, args.softening, args.g)
    g, m = sizes(mixed)
    ms_per_step = 1e3 * secs / args.steps
    value = n * args.steps / secs / 1e6

    fp64 = None
    if not args.no_fp64:
        f64 = nb.Simulation(p, args.dt, args.theta, args.softening, precision=nb.PRECISION_FP64,
                            **common)
        f64.initialize()
        a64 = f64.accelerations()
        pp64, pn64, visits64 = f64.count_interactions(with_visits=True)
        acc["rms_vs_fp64_bh"], acc["max_vs_fp64_bh"] = rel_err(a_mixed0, a64)
        acc["fp64_rms_vs_direct"], acc["fp64_max_vs_direct"] = rel_err(
            [a[targets] for a in a64], exact)
        f64.step()  # warm-up
        secs64 = time_steps(torch, f64, args.fp64_steps)
        phases64 = phase_steps(f64, 1)
        retire(f64)
        fp64 = {"value": round(n * args.fp64_steps / secs64 / 1e6, 3),
                "ms_per_step": round(1e3 * secs64 / args.fp64_steps, 3),
                "steps": args.fp64_steps, "phase_ms": phases64}
        t64 = phases64["traversal"]
        fp64["interactions_per_step"] = pp64 + pn64
        fp64["packet_efficiency"] = round((pp64 + pn64) / (64.0 * visits64), 4) if visits64 else None
        a64_tf = FLOPS_PER_INTERACTION * (pp64 + pn64) / (t64 * 1e-3) / 1e12
        fp64["roofline"] = {"bound": "valu", "achieved": round(a64_tf, 3),
                            "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": round(a64_tf / FP64_VALU_PEAK_TFLOPS, 4),
                            "kernel": "walk_fp64_kernel"}
    acc["rms_vs_direct"], acc["max_vs_direct"] = rel_err([a[targets] for a in a_mixed0], exact)

    log("headline FP64 leg done")
    by_theta = {}
    if args.workload == "plummer" and args.thetas:
        for th in (float(t) for t in args.thetas.split(",")):
            log(f"theta {th} accuracy and drift")
            by_theta[str(th)] = drift_run(nb, torch, p, th, args, args.drift_steps, targets,
                                          exact)

    frames = frame_capture_timing(nb, torch, mixed, n)
    retire(mixed)

    trav_ms = phases["traversal"]
    achieved = FLOPS_PER_INTERACTION * interactions / (trav_ms * 1e-3) / 1e12
    traffic = traversal_traffic() if (args.workload == "plummer" and n == 4_194_304 and
                                      args.theta == 0.5) else None
    line = {
        "metric": "Mparticle-steps/s, N=4M theta=0.5: mixed-FP32 vs FP64 at <=1e-3 rms accel err",
        "value": round(value, 3),
        "unit": "Mparticle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 traversal / f64 build+integrate",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"{args.workload} N={n} theta={args.theta} "
                               f"eps={args.softening} dt={args.dt} G={args.g:g}"
                               + (f" m={args.particle_mass:g}" if args.particle_mass else "")
                               + " 63-bit keys",
                   "n_particles": n, "theta": args.theta, "softening": args.softening,
                   "dt": args.dt, "precision": "mixed"},
        "steps_per_s": round(1e3 / ms_per_step, 4),
        "interactions_per_step": interactions,
        "interactions_pp_pn": [pp, pn],
        "interactions_per_s": interactions * 1e3 / ms_per_step,
        "packet_efficiency": round(interactions / (64.0 * visits), 4) if visits else None,
        "groups": g, "octree_nodes": m,
        "fp64": fp64,
        "fp64_value": None if fp64 is None else fp64["value"],
        "speedup_vs_fp64": None if fp64 is None else round(value / fp64["value"], 3),
        "accuracy": {k: g4(v) for k, v in acc.items()},
        "momentum_drift": g4(float(np.abs(momentum(s1) - momentum(s0)).max())),
        "energy": None if e0 is None else {
            "E_start": e0[0], "E_end": e1[0], "kinetic_start": e0[1],
            "potential_start": e0[2], "rel_drift": g4((e1[0] - e0[0]) / abs(e0[0])),
            "steps": args.warmup + args.steps + 2 + extra},
        "accuracy_by_theta": by_theta,
        "phase_ms": phases,
        "phase_bandwidth": phase_bandwidth(phases, n, g, m),
        "frame_capture": frames,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 4),
                     "traffic": traffic,
                     "traffic_source": (f"profile-sourced: {TRAFFIC_PROFILE} (rocprofv3 "
                                        "FETCH_SIZE + WRITE_SIZE of this kernel on this "
                                        "workload, calibrated against a 1 GiB copy)")
                     if traffic else None,
                     "kernel": "walk_mixed_kernel",
                     "flops_per_interaction": FLOPS_PER_INTERACTION,
                     "frac_of_non_packed_fp32": round(achieved / 78.6, 4)},
    }
    return line, p


def config_uniform_1m(nb, torch, args):
    """1,048,576 uniform in [-1,1]^3, lognormal masses, FP64, theta 0.5,
    eps 1e-3, dt 1e-4, 20 steps; tree vs the oracle; 4096 direct-sum
    targets; the mixed pipeline on the same workload for the ratio."""
    import oracle
    n, steps = 1_048_576, 20
    p = nb.generate("uniform", n, 43)
    common = dict(gravitational_constant=1.0, bits_per_axis=21)
    targets = sample_targets(n, 4096, seed=1)
    exact = direct_reference(nb, torch, p, targets, 1e-3, 1.0)
    out = {"workload": "uniform [-1,1]^3 N=1048576 lognormal masses theta=0.5 eps=1e-3 "
                       "dt=1e-4 20 steps 63-bit keys"}
    for name, prec in (("fp64", nb.PRECISION_FP64), ("mixed", nb.PRECISION_MIXED)):
        sim = nb.Simulation(p, 1e-4, 0.5, 1e-3, precision=prec, **common)
        sim.initialize()
        r, mx = rel_err([a[targets] for a in sim.accelerations()], exact)
        leg = {"rms_vs_direct": g4(r), "max_vs_direct": g4(mx)}
        if prec == nb.PRECISION_FP64:
            ref = oracle.Sim(p, 1e-4, 0.5, 1e-3, 1.0, 21)
            ref.rebuild()
            g, o = sim.structure_host(), ref.structure()
            leg["tree_bit_exact_vs_oracle"] = bool(
                np.array_equal(g["keys"], o["keys"]) and
                np.array_equal(g["indices"], o["indices"]) and
                np.array_equal(g["children"], o["children"]) and
                np.array_equal(g["first"], o["first"]) and
                np.array_equal(g["count"], o["count"]))
            del ref
        secs = time_steps(torch, sim, steps)
        leg.update({"steps": steps, "ms_per_step": round(1e3 * secs / steps, 3),
                    "Mparticle_steps_per_s": round(n * steps / secs / 1e6, 3)})
        leg["phase_ms"] = phase_steps(sim, 1)
        retire(sim)
        out[name] = leg
    out["mixed_over_fp64"] = round(out["mixed"]["Mparticle_steps_per_s"] /
                                   out["fp64"]["Mparticle_steps_per_s"], 3)
    return out


def config_clumps_16m(nb, torch, args):
    """16,777,216 equal masses in 256 Gaussian clumps, cold start, mixed,
    theta 0.6, eps 1e-3, dt 5e-5, 20 steps: phase breakdown with per-phase
    HBM GB/s, traversal FLOP/s, accuracy vs the FP64 pipeline and the direct
    sum, energy drift; the FP64 pipeline's rate for the ratio."""
    n, steps, theta, eps, dt = 16_777_216, 20, 0.6, 1e-3, 5e-5
    p = nb.generate("clumps", n, 44)
    common = dict(gravitational_constant=1.0, bits_per_axis=21)
    targets = sample_targets(n, 8192, seed=2)
    exact = direct_reference(nb, torch, p, targets, eps, 1.0)
    mixed = nb.Simulation(p, dt, theta, eps, precision=nb.PRECISION_MIXED, **common)
    mixed.initialize()
    a_mixed = mixed.accelerations()
    pp, pn = mixed.count_interactions()
    s0 = mixed.state()
    e0 = energy_of(nb, s0, dt, theta, eps, 1.0)
    mixed.step()
    secs = time_steps(torch, mixed, steps)
    phases = phase_steps(mixed, 2)
    s1 = mixed.state()
    e1 = energy_of(nb, s1, dt, theta, eps, 1.0)
    g, m = sizes(mixed)
    retire(mixed)
    f64 = nb.Simulation(p, dt, theta, eps, precision=nb.PRECISION_FP64, **common)
    f64.initialize()
    a64 = f64.accelerations()
    secs64 = time_steps(torch, f64, 3)
    retire(f64)
    trav = phases["traversal"]
    tf = FLOPS_PER_INTERACTION * (pp + pn) / (trav * 1e-3) / 1e12
    r_bh, mx_bh = rel_err(a_mixed, a64)
    r_d, mx_d = rel_err([a[targets] for a in a_mixed], exact)
    return {
        "workload": "256 Gaussian clumps N=16777216 cold start theta=0.6 eps=1e-3 dt=5e-5 "
                    "20 steps 63-bit keys",
        "steps": steps, "ms_per_step": round(1e3 * secs / steps, 3),
        "Mparticle_steps_per_s": round(n * steps / secs / 1e6, 3),
        "fp64_Mparticle_steps_per_s": round(n * 3 / secs64 / 1e6, 3),
        "mixed_over_fp64": round((n * steps / secs) / (n * 3 / secs64), 3),
        "groups": g, "octree_nodes": m, "interactions_per_step": pp + pn,
        "phase_ms": phases, "phase_bandwidth": phase_bandwidth(phases, n, g, m),
        "traversal_TFLOP_per_s": round(tf, 3),
        "traversal_frac_of_fp32_peak": round(tf / FP32_VALU_PEAK_TFLOPS, 4),
        "rms_vs_fp64_bh": g4(r_bh), "max_vs_fp64_bh": g4(mx_bh),
        "rms_vs_direct": g4(r_d), "max_vs_direct": g4(mx_d),
        "energy_rel_drift": g4((e1[0] - e0[0]) / abs(e0[0])),
        "momentum_drift": g4(float(np.abs(momentum(s1) - momentum(s0)).max())),
    }


def config_reference_blob(nb, torch, n):
    """The reference's own benchmark (README.md:571-588): an N(0,1)^3 blob,
    velocities sigma 0.01, masses 1e30 kg, SI G, theta 0.5, eps 1e-6,
    dt 1e-3, 10 steps after initialisation; published 249 ms (100K) and
    798 ms (250K) per FP64 step on an RTX 3070."""
    p = nb.generate("blob", n, 42)
    p["m"][:] = 1e30
    out = {}
    for name, prec in (("mixed", nb.PRECISION_MIXED), ("fp64", nb.PRECISION_FP64)):
        sim = nb.Simulation(p, 1e-3, 0.5, 1e-6, gravitational_constant=6.6743e-11,
                            precision=prec, bits_per_axis=21)
        sim.initialize()
        secs = time_steps(torch, sim, 10)
        out[f"{name}_ms_per_step"] = round(1e3 * secs / 10, 3)
        if prec == nb.PRECISION_MIXED:
            out["mixed_phase_ms"] = phase_steps(sim, 2)
        retire(sim)
    ref = {100_000: 249.0, 250_000: 798.0}.get(n)
    if ref:
        out["reference_fp64_ms_per_step_rtx3070"] = ref
    return out


def frame_capture_timing(nb, torch, sim, n, reps=20):
    """Device frame capture of the simulation's current positions (after
    the timed steps): bounds + quantize at the display sample size and at
    the full particle count, and the whole DeviceFrameWriter.write path
    (capture, copy to the host, atomic file publish) at the display size.
    Algorithmic bytes per sampled particle: 24 read by the bounds pass,
    24 read and 6 written by the quantize pass."""
    import tempfile
    v = sim.particles_view()
    out = {}
    for s in (nb.DEFAULT_STREAM_PARTICLES, n):
        nb.capture_frame(v.x, v.y, v.z, s, n)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            nb.capture_frame(v.x, v.y, v.z, s, n)
        dt = (time.perf_counter() - t0) / reps
        out[f"capture_{s}"] = {"ms": round(dt * 1e3, 4),
                               "GB_per_s": round(54 * s / dt / 1e9, 1)}
    types = torch.zeros(n, dtype=torch.uint8, device="cuda")
    writer = nb.DeviceFrameWriter.from_simulation(sim, types)
    with tempfile.TemporaryDirectory() as d:
        spool = nb.FrameSpool(d, "bench")
        writer.write(spool, 0, 0, 0.0)
        t0 = time.perf_counter()
        for k in range(1, 6):
            writer.write(spool, k, 0, 0.0)
        out["write_ms"] = round((time.perf_counter() - t0) / 5 * 1e3, 3)
    out["sample"] = writer.sample_particle_count
    return out


# ----------------------------------------------------------------- CPU
def host_description():
    model = platform.processor() or ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(p, args):
    log("CPU baseline")
    return _cpu_baseline(p, args)


def _cpu_baseline(p, args):
    """The oracle (the reference algorithm in FP64 C++/OpenMP) on the host,
    median of 3 repeats each: (1) the headline workload, one full rebuild
    plus the forces of a bounded sample of Morton-ordered targets,
    extrapolated linearly to all N targets (stated in `sample`); (2) one
    full step of the 1M uniform workload (rebuild + every target), measured
    whole; (3) the plain CPU case end to end (Plummer 16,384, 10 steps)."""
    import oracle
    import nbody_amd as nb
    n = len(p["m"])
    sample = min(args.cpu_targets, n)
    runs = []
    for _ in range(3):
        sim = oracle.Sim(p, args.dt, args.theta, args.softening, args.g, 21)
        t0 = time.perf_counter()
        sim.rebuild()
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        sim.partial_forces(0, sample)
        t_pf = time.perf_counter() - t0
        t_forces = max(t_pf - t_build, 1e-9)
        runs.append((t_build + t_forces * n / sample, t_build, t_forces))
        del sim
    step, t_build, t_forces = sorted(runs)[1]
    # one whole step of the 1M uniform workload
    pu = nb.generate("uniform", 1_048_576, 43)
    whole = []
    for _ in range(3):
        sim = oracle.Sim(pu, 1e-4, 0.5, 1e-3, 1.0, 21)
        sim.initialize()
        t0 = time.perf_counter()
        sim.step(1)
        whole.append(time.perf_counter() - t0)
        del sim
    t_1m = statistics.median(whole)
    return {"value": round(n / step / 1e6, 4), "unit": "Mparticle-steps/s",
            "cores": oracle.threads(), "kind": "port",
            "host": host_description(), "omp_threads": oracle.threads(), "repeats": 3,
            "statistic": "median",
            "sample": (f"headline workload: one rebuild at N={n} ({t_build:.2f} s) + forces "
                       f"of {sample} Morton-ordered targets ({t_forces:.2f} s); the forces are "
                       f"extrapolated linearly from {sample} to {n} targets"),
            "uniform_1m_full_step": {"s_per_step": round(t_1m, 3),
                                     "Mparticle_steps_per_s": round(1_048_576 / t_1m / 1e6, 4),
                                     "measured": "whole step, every target, not extrapolated"},
            "small_plummer_end_to_end": small_plummer_end_to_end()}


def small_plummer_end_to_end():
    """The plain CPU case end to end: Plummer N=16384, G=1, theta=0.5,
    eps=0.01, dt=1e-3, 10 KDK steps, 30-bit keys, on the oracle (median of 3)
    and on the GPU pipeline (whose FP64 states must equal the oracle's bit for
    bit)."""
    import oracle
    import nbody_amd as nb
    import torch
    n, steps = 16384, 10
    p = nb.generate("plummer", n, 42)
    times = []
    for _ in range(3):
        ref = oracle.Sim(p, 1e-3, 0.5, 0.01, 1.0, 10)
        ref.initialize()
        t0 = time.perf_counter()
        ref.step(steps)
        times.append(time.perf_counter() - t0)
    out = {"n": n, "steps": steps,
           "cpu_ms_per_step": round(1e3 * statistics.median(times) / steps, 3)}
    rs = ref.state()
    for name, prec in (("fp64", nb.PRECISION_FP64), ("mixed", nb.PRECISION_MIXED)):
        sim = nb.Simulation(p, 1e-3, 0.5, 0.01, gravitational_constant=1.0,
                            precision=prec, bits_per_axis=10)
        sim.initialize()
        secs = time_steps(torch, sim, steps)
        out[f"gpu_{name}_ms_per_step"] = round(1e3 * secs / steps, 3)
        st = sim.state()
        if prec == nb.PRECISION_FP64:
            out["fp64_state_bit_exact"] = all(np.array_equal(st[f], rs[f])
                                             for f in ("x", "y", "z", "vx", "vy", "vz"))
        else:
            dx = np.sqrt(sum((st[f] - rs[f]) ** 2 for f in ("x", "y", "z")))
            out["mixed_max_position_diff"] = float(dx.max())
        retire(sim)
    return out


# ------------------------------------------------------------- adjoint
# FLOPs of one interaction of the force VJP: separation and r^2 (9), rsq
# and its powers (4), d . abar (5), the two coefficients (4), the three
# gradient components (9), the target sum (3), the mass term (1)
VJP_FLOPS_PER_INTERACTION = 35


def vjp_roofline(nb, torch, p, precision, n_reps=3):
    """Time of one force VJP on the disk + bulge state (walk, push-down and
    combination; HIP events on the default stream, on which the library
    runs) and its achieved FLOP rate over the interactions of that tree."""
    s = nb.Simulation(p, 1e-3, 0.5, 0.01, gravitational_constant=1.0, precision=precision,
                      bits_per_axis=21)
    s.initialize()
    pp, pn = s.count_interactions()
    ab = [torch.from_numpy(np.random.default_rng(5).normal(size=len(p["m"]))) for _ in range(3)]
    s.force_vjp(*ab)  # warm-up
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(n_reps):
        ev0.record()
        s.force_vjp(*ab)
        ev1.record()
        torch.cuda.synchronize()
        ms.append(ev0.elapsed_time(ev1))
    retire(s)
    t = float(np.median(ms))
    peak = FP32_VALU_PEAK_TFLOPS if precision == nb.PRECISION_MIXED else FP64_VALU_PEAK_TFLOPS
    achieved = VJP_FLOPS_PER_INTERACTION * (pp + pn) / (t * 1e-3) / 1e12
    return {"vjp_ms": round(t, 3), "interactions": pp + pn, "bound": "valu",
            "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "flops_per_interaction": VJP_FLOPS_PER_INTERACTION}


def adjoint_workload(args, probes=None, quiet=False, precision="fp64"):
    """The differentiable workload: K leapfrog steps on the disk + bulge
    galaxy, reverse mode with checkpoints, and the gradient of
    L = sum_i |r_i(T) - r*_i|^2 (r* a perturbed copy of the final positions)
    with respect to the initial velocities and the masses.  precision
    "fp64" differentiates the FP64 pipeline, "mixed" the mixed one.  Central
    differences on a few coordinates of the FP64 evolution check either
    gradient; the mixed gradient is also compared with the FP64 one."""
    import torch
    import nbody_amd as nb

    n, K, every = args.adjoint_n, args.adjoint_steps, args.adjoint_every
    nprobe = args.adjoint_probes if probes is None else probes
    prec = nb.PRECISION_MIXED if precision == "mixed" else nb.PRECISION_FP64
    p = nb.generate("disk", n, 42)
    adj = nb.Adjoint(n, 1e-3, 0.5, 0.01, K, every, gravitational_constant=1.0, precision=prec)
    ref = adj if prec == nb.PRECISION_FP64 else nb.Adjoint(
        n, 1e-3, 0.5, 0.01, K, every, gravitational_constant=1.0, precision=nb.PRECISION_FP64)
    state = [torch.from_numpy(p[k]).cuda() for k in ("x", "y", "z", "vx", "vy", "vz")]
    m = torch.from_numpy(p["m"]).cuda()

    def final(mm, st):
        return np.stack([t.cpu().numpy() for t in ref.forward(mm, *st)[:3]])

    r_end = final(m, state)
    target = r_end + 1e-3 * np.random.default_rng(1).normal(size=r_end.shape)

    def loss(mm, st):
        # L(run) - L(base run), from the small difference of the final
        # positions: sum d (d + 2 (r_end - r*)), d = r(T) - r_end
        d = final(mm, st) - r_end
        return float((d * (d + 2.0 * (r_end - target))).sum())

    def gradient(a):
        out = a.forward(m, *state)
        cot = [torch.from_numpy(2.0 * (o.cpu().numpy() - target[k])).cuda()
               for k, o in enumerate(out[:3])]
        zeros = [torch.zeros_like(state[0]) for _ in range(3)]
        g = a.backward(*cot, *zeros, with_mass=True)
        return np.stack([t.cpu().numpy() for t in g[3:6]]), g[6].cpu().numpy()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = adj.forward(m, *state)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    cot = [torch.from_numpy(2.0 * (o.cpu().numpy() - target[k])).cuda()
           for k, o in enumerate(out[:3])]
    zeros = [torch.zeros_like(state[0]) for _ in range(3)]
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    grads = adj.backward(*cot, *zeros, with_mass=True)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    gv = np.stack([g.cpu().numpy() for g in grads[3:6]])
    gm = grads[6].cpu().numpy()

    rng = np.random.default_rng(2)
    results = []
    for k in range(nprobe):
        i = int(rng.integers(n))
        on_mass = k % 2 == 1
        c = int(rng.integers(3))
        h = 1e-5 * p["m"][i] if on_mass else 1e-7

        def f(d):
            if on_mass:
                mm = m.clone()
                mm[i] += d
                return loss(mm, state)
            st = [t.clone() for t in state]
            st[3 + c][i] += d
            return loss(m, st)
        fd1 = (f(h) - f(-h)) / (2 * h)
        fd2 = (f(2 * h) - f(-2 * h)) / (4 * h)
        ad = float(gm[i] if on_mass else gv[c, i])
        results.append({"what": "mass" if on_mass else f"v{'xyz'[c]}", "index": i,
                        "adjoint": ad, "fd": fd1,
                        "smooth": bool(abs(fd1 - fd2) <= 1e-4 * abs(fd1)),
                        "rel_err": abs(ad - fd1) / max(abs(fd1), 1e-300)})

    def worst(kind):
        e = [q["rel_err"] for q in results if q["smooth"] and (q["what"] == "mass") == kind]
        return (len(e), max(e) if e else None)
    fwd_ms, bwd_ms = 1e3 * (t1 - t0), 1e3 * (t3 - t2)
    line = {
        "metric": "adjoint Mparticle-steps/s (forward + reverse), disk + bulge",
        "value": round(n * K / ((t1 - t0) + (t3 - t2)) / 1e6, 3),
        "unit": "Mparticle-steps/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32 interactions / f64 accumulation" if prec == nb.PRECISION_MIXED else "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"disk+bulge N={n} theta=0.5 eps=0.01 dt=1e-3 K={K} "
                               f"checkpoints every {every}",
                   "n_particles": n, "steps": K, "checkpoint_every": every,
                   "precision": precision},
        "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
        "backward_over_forward": round(bwd_ms / fwd_ms, 3),
        "loss": "sum |r(T) - r*|^2, r* = r(T) + 1e-3 N(0,1)",
        "fd_probes": results,
        # probes whose steps h and 2h agree to 1e-4 (no opening decision
        # flipped within 2h), and their largest relative error; the
        # differences are of the FP64 evolution
        "fd_velocity_smooth_max_rel_err": worst(False),
        "fd_mass_smooth_max_rel_err": worst(True),
        "vjp": vjp_roofline(nb, torch, p, prec),
    }
    if ref is not adj:
        gv64, gm64 = gradient(ref)
        line["vs_fp64_gradient"] = {
            "v0_rel_l2": float(np.linalg.norm(gv - gv64) / np.linalg.norm(gv64)),
            "m_rel_l2": float(np.linalg.norm(gm - gm64) / np.linalg.norm(gm64))}
        retire(ref)
    retire(adj)
    if not quiet:
        print(json.dumps(line))
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4_194_304)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--softening", type=float, default=1e-3)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--workload", default="plummer")
    ap.add_argument("--g", type=float, default=1.0, help="gravitational constant")
    ap.add_argument("--particle-mass", type=float, default=None,
                    help="give every particle this mass (the reference's benchmark "
                         "uses 1e30 with --g 6.6743e-11)")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--drift-steps", type=int, default=50,
                    help="steps over which energy and momentum drift are measured")
    ap.add_argument("--direct-targets", type=int, default=8192)
    ap.add_argument("--cpu-targets", type=int, default=131072)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first (the "
                         "reference dispatch for FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--thetas", default="0.3,0.8",
                    help="extra opening angles for the accuracy and drift table (Plummer only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-energy", action="store_true",
                    help="skip the FP64 energy check (it builds a second, FP64 simulation)")
    ap.add_argument("--no-configs", action="store_true",
                    help="measure the headline workload only (no other configs)")
    ap.add_argument("--adjoint", action="store_true",
                    help="measure only the differentiable workload (forward + reverse)")
    ap.add_argument("--adjoint-precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--config", choices=("uniform_1m", "clumps_16m", "blob_100k", "blob_250k"),
                    help="measure only this configuration and print its JSON object")
    ap.add_argument("--adjoint-n", type=int, default=262_144)
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--adjoint-every", type=int, default=10)
    ap.add_argument("--adjoint-probes", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.adjoint:
        adjoint_workload(args, precision=args.adjoint_precision)
        return

    import torch
    import nbody_amd as nb

    if args.config:
        legs = {"uniform_1m": lambda: config_uniform_1m(nb, torch, args),
                "clumps_16m": lambda: config_clumps_16m(nb, torch, args),
                "blob_100k": lambda: config_reference_blob(nb, torch, 100_000),
                "blob_250k": lambda: config_reference_blob(nb, torch, 250_000)}
        print(json.dumps({args.config: legs[args.config]()}))
        return

    if args.calibrate_copy_mib > 0:
        words = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(words, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    log("headline workload")
    line, p = headline(nb, torch, args)
    log(f"headline: {line['value']} Mparticle-steps/s, {line['ms_per_step']} ms/step")
    if not args.no_configs:
        cfg = {}
        cfg["uniform_1m_fp64"] = config_uniform_1m(nb, torch, args)
        log("uniform 1M done")
        cfg["clumps_16m_mixed"] = config_clumps_16m(nb, torch, args)
        log("clumps 16M done")
        cfg["reference_blob_100k"] = config_reference_blob(nb, torch, 100_000)
        cfg["reference_blob_250k"] = config_reference_blob(nb, torch, 250_000)
        log("reference blobs done")
        keys = ("value", "unit", "dtype", "config", "forward_ms", "backward_ms",
                "backward_over_forward", "fd_probes", "fd_velocity_smooth_max_rel_err",
                "fd_mass_smooth_max_rel_err", "vjp", "vs_fp64_gradient")
        adj = adjoint_workload(args, probes=4, quiet=True)
        cfg["disk_bulge_adjoint"] = {k: adj[k] for k in keys if k in adj}
        log("adjoint FP64 done")
        adj = adjoint_workload(args, probes=4, quiet=True, precision="mixed")
        cfg["disk_bulge_adjoint_mixed"] = {k: adj[k] for k in keys if k in adj}
        log("adjoint mixed done")
        line["configs"] = cfg
    release_retired()
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(p, args)
    log("done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
