"""Benchmark: Barnes-Hut leapfrog steps on one MI355X.

Workload (default): Plummer sphere, N = 4,194,304, theta 0.5, softening 1e-3,
dt 1e-4, G = 1, 63-bit Morton keys.  One step = kick-drift + bounding box,
Morton keys, radix sort, leaf groups, Karras tree, radix->octree plan,
octree, mass/COM, walk layout, traversal, closing kick -- the reference's
Simulation::step (src/simulation.cpp:100-119) -- with inputs resident in HBM.

Prints one JSON line: Mparticle-steps/s of the mixed-FP32 pipeline, the FP64
pipeline's rate and the speed-up, the accuracy of the mixed forces against
the FP64 pipeline and an FP64 direct sum (at theta 0.3 / 0.5 / 0.8 for the
Plummer workload), momentum and energy drift over the timed steps, the
traversal's achieved VALU FLOP/s against the MI355X FP32 peak, and the CPU
oracle timed on a bounded sample of the same workload.

Energy: kinetic + potential, the potential from an FP64 walk over a tree
built at the state (Simulation.energy of an FP64 simulation of that state).
"""

import argparse
import json
import math
import os
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


def rel_err(a, b):
    a = np.stack(a)
    b = np.stack(b)
    r = np.sqrt(((a - b) ** 2).sum(0)) / np.sqrt((b ** 2).sum(0))
    return float(np.sqrt((r ** 2).mean())), float(r.max())


def momentum(s):
    return np.array([math.fsum(s["m"] * s[k]) for k in ("vx", "vy", "vz")])


def time_steps(nb, torch, sim, steps):
    """Runs `steps` steps; returns seconds and mean phase times (ms)."""
    phases = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        pt = sim.phase_times()
        if phases is None:
            phases = {k: 0.0 for k in pt}
        for k, v in pt.items():
            phases[k] += v / steps
    torch.cuda.synchronize()
    return time.perf_counter() - t0, phases


def traversal_traffic():
    """HBM bytes per traversal dispatch from the committed counter summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, calibrated against a copy of
    known size), or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "walk_mixed_plummer4m_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_dispatch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def energy_of(nb, state, args):
    """(E, K, U) of a state: an FP64 simulation of it, initialised (tree and
    forces), then its FP64 potential walk."""
    s = nb.Simulation(state, args.dt, args.theta, args.softening, gravitational_constant=args.g,
                      precision=nb.PRECISION_FP64, bits_per_axis=21)
    s.initialize()
    k, u = s.energy()
    s.close()
    return k + u, k, u


def accuracy_at(nb, torch, p, theta, args, targets, exact):
    """Mixed vs FP64 Barnes-Hut and vs the direct sum at one theta, with the
    traversal time of each precision (one force evaluation)."""
    out = {}
    acc = {}
    for name, prec in (("mixed", nb.PRECISION_MIXED), ("fp64", nb.PRECISION_FP64)):
        s = nb.Simulation(p, args.dt, theta, args.softening, gravitational_constant=args.g,
                          precision=prec, bits_per_axis=21)
        s.set_timing(True)
        s.initialize()
        acc[name] = s.accelerations()
        out[f"{name}_traversal_ms"] = round(s.phase_times()["traversal"], 3)
        s.close()
    r, mx = rel_err(acc["mixed"], acc["fp64"])
    out["rms_vs_fp64_bh"], out["max_vs_fp64_bh"] = float(f"{r:.4g}"), float(f"{mx:.4g}")
    r, mx = rel_err([a[targets] for a in acc["mixed"]], exact)
    out["rms_vs_direct"], out["max_vs_direct"] = float(f"{r:.4g}"), float(f"{mx:.4g}")
    r, _ = rel_err([a[targets] for a in acc["fp64"]], exact)
    out["fp64_rms_vs_direct"] = float(f"{r:.4g}")
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


def cpu_baseline(p, args):
    """The oracle (the reference algorithm in FP64 C++/OpenMP) on the same
    particles: one full rebuild plus the forces of a bounded sample of
    Morton-ordered targets, scaled to one step."""
    import oracle
    n = len(p["m"])
    sample = min(args.cpu_targets, n)
    sim = oracle.Sim(p, args.dt, args.theta, args.softening, args.g, 21)
    t0 = time.perf_counter()
    sim.rebuild()
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    sim.partial_forces(0, sample)
    t_pf = time.perf_counter() - t0
    t_forces = max(t_pf - t_build, 1e-9)
    step = t_build + t_forces * n / sample
    return {"value": n / step / 1e6, "unit": "Mparticle-steps/s",
            "cores": oracle.threads(), "kind": "port",
            "sample": f"one rebuild at N={n} ({t_build:.2f} s) + forces of {sample} "
                      f"Morton-ordered targets ({t_forces:.2f} s), scaled to N",
            "small_plummer_end_to_end": small_plummer_end_to_end()}


def small_plummer_end_to_end():
    """The plain CPU case end to end: Plummer N=16384, G=1, theta=0.5,
    eps=0.01, dt=1e-3, 10 KDK steps, 30-bit keys, on the oracle and on the
    GPU pipeline (whose FP64 states must equal the oracle's bit for bit)."""
    import oracle
    import nbody_amd as nb
    n, steps = 16384, 10
    p = nb.generate("plummer", n, 42)
    ref = oracle.Sim(p, 1e-3, 0.5, 0.01, 1.0, 10)
    ref.initialize()
    t0 = time.perf_counter()
    ref.step(steps)
    t_cpu = time.perf_counter() - t0
    out = {"n": n, "steps": steps, "cpu_ms_per_step": round(1e3 * t_cpu / steps, 3)}
    rs = ref.state()
    for name, prec in (("fp64", nb.PRECISION_FP64), ("mixed", nb.PRECISION_MIXED)):
        sim = nb.Simulation(p, 1e-3, 0.5, 0.01, gravitational_constant=1.0,
                            precision=prec, bits_per_axis=10)
        sim.initialize()
        import torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run(steps)
        torch.cuda.synchronize()
        out[f"gpu_{name}_ms_per_step"] = round(1e3 * (time.perf_counter() - t0) / steps, 3)
        st = sim.state()
        if prec == nb.PRECISION_FP64:
            out["fp64_state_bit_exact"] = all(np.array_equal(st[f], rs[f])
                                             for f in ("x", "y", "z", "vx", "vy", "vz"))
        else:
            dx = np.sqrt(sum((st[f] - rs[f]) ** 2 for f in ("x", "y", "z")))
            out["mixed_max_position_diff"] = float(dx.max())
        sim.close()
    return out


def adjoint_workload(args):
    """The differentiable workload: K leapfrog steps of the FP64 pipeline on
    the disk + bulge galaxy, reverse mode with checkpoints, and the gradient
    of L = sum_i |r_i(T) - r*_i|^2 (r* a perturbed copy of the final
    positions) with respect to the initial velocities and the masses,
    checked by central differences on a few coordinates."""
    import torch
    import nbody_amd as nb

    n, K, every = args.n, args.adjoint_steps, args.adjoint_every
    p = nb.generate("disk", n, 42)
    adj = nb.Adjoint(n, args.dt, args.theta, args.softening, K, every,
                     gravitational_constant=args.g)
    state = [torch.from_numpy(p[k]).cuda() for k in ("x", "y", "z", "vx", "vy", "vz")]
    m = torch.from_numpy(p["m"]).cuda()

    def final(mm, st):
        return np.stack([t.cpu().numpy() for t in adj.forward(mm, *st)[:3]])

    r_end = final(m, state)
    target = r_end + 1e-3 * np.random.default_rng(1).normal(size=r_end.shape)

    def loss(mm, st):
        # L(run) - L(base run), from the small difference of the final
        # positions: sum d (d + 2 (r_end - r*)), d = r(T) - r_end
        d = final(mm, st) - r_end
        return float((d * (d + 2.0 * (r_end - target))).sum())

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
    probes = []
    for k in range(args.adjoint_probes):
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
        probes.append({"what": "mass" if on_mass else f"v{'xyz'[c]}", "index": i,
                       "adjoint": ad, "fd": fd1,
                       "smooth": bool(abs(fd1 - fd2) <= 1e-4 * abs(fd1)),
                       "rel_err": abs(ad - fd1) / max(abs(fd1), 1e-300)})
    def worst(kind):
        e = [q["rel_err"] for q in probes if q["smooth"] and (q["what"] == "mass") == kind]
        return (len(e), max(e) if e else None)
    fwd_ms, bwd_ms = 1e3 * (t1 - t0), 1e3 * (t3 - t2)
    line = {
        "metric": "adjoint Mparticle-steps/s (forward + reverse), disk + bulge",
        "value": round(n * K / ((t1 - t0) + (t3 - t2)) / 1e6, 3),
        "unit": "Mparticle-steps/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": f"disk+bulge N={n} theta={args.theta} eps={args.softening} "
                               f"dt={args.dt} K={K} checkpoints every {every}",
                   "n_particles": n, "steps": K, "checkpoint_every": every},
        "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
        "backward_over_forward": round(bwd_ms / fwd_ms, 3),
        "loss": "sum |r(T) - r*|^2, r* = r(T) + 1e-3 N(0,1)",
        "fd_probes": probes,
        # probes whose steps h and 2h agree to 1e-4 (no opening decision
        # flipped within 2h), and their largest relative error
        "fd_velocity_smooth_max_rel_err": worst(False),
        "fd_mass_smooth_max_rel_err": worst(True),
    }
    adj.close()
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
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
    ap.add_argument("--direct-targets", type=int, default=8192)
    ap.add_argument("--cpu-targets", type=int, default=524288)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this size first (the "
                         "reference dispatch for FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--thetas", default="0.3,0.8",
                    help="extra opening angles for the accuracy table (Plummer only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-energy", action="store_true",
                    help="skip the FP64 energy check (it builds a second, FP64 simulation)")
    ap.add_argument("--adjoint", action="store_true",
                    help="measure the differentiable workload instead (FP64 forward + "
                         "reverse over --adjoint-steps steps; use with --n 262144 "
                         "--theta 0.5 --softening 0.01 --dt 1e-3)")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--adjoint-every", type=int, default=10)
    ap.add_argument("--adjoint-probes", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.adjoint:
        adjoint_workload(args)
        return

    import torch
    import nbody_amd as nb

    if args.calibrate_copy_mib > 0:
        words = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(words, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    n = args.n
    p = nb.generate(args.workload, n, 42)
    if args.particle_mass is not None:
        p["m"][:] = args.particle_mass
    common = dict(gravitational_constant=args.g, bits_per_axis=21)

    mixed = nb.Simulation(p, args.dt, args.theta, args.softening,
                          precision=nb.PRECISION_MIXED, **common)
    mixed.set_timing(True)
    mixed.initialize()
    a_mixed0 = mixed.accelerations()
    pp, pn, visits = mixed.count_interactions(with_visits=True)
    interactions = pp + pn

    # accuracy at t = 0 against the FP64 pipeline (bit-exact with the oracle)
    # and against an FP64 direct sum on sampled targets
    acc = {}
    fp64_rate = None
    if not args.no_fp64:
        f64 = nb.Simulation(p, args.dt, args.theta, args.softening,
                            precision=nb.PRECISION_FP64, **common)
        f64.set_timing(True)
        f64.initialize()
        a64 = f64.accelerations()
        acc["rms_vs_fp64_bh"], acc["max_vs_fp64_bh"] = rel_err(a_mixed0, a64)
    rng = np.random.default_rng(0)
    targets = np.sort(rng.choice(n, min(args.direct_targets, n), replace=False)).astype(np.uint32)
    dev = [torch.from_numpy(p[k]).cuda() for k in ("m", "x", "y", "z")]
    exact = [t.cpu().numpy() for t in
             nb.direct_sum(*dev, torch.from_numpy(targets.view(np.int32)).cuda(),
                           args.softening, args.g)]
    acc["rms_vs_direct"], acc["max_vs_direct"] = rel_err([a[targets] for a in a_mixed0], exact)
    if not args.no_fp64:
        acc["fp64_rms_vs_direct"], acc["fp64_max_vs_direct"] = rel_err(
            [a[targets] for a in a64], exact)
    del dev

    s0 = mixed.state()
    p0 = momentum(s0)
    e0 = None if args.no_energy else energy_of(nb, s0, args)
    for _ in range(args.warmup):
        mixed.step()
    secs, phases = time_steps(nb, torch, mixed, args.steps)
    s1 = mixed.state()
    p1 = momentum(s1)
    e1 = None if args.no_energy else energy_of(nb, s1, args)
    ms_per_step = 1e3 * secs / args.steps
    value = n * args.steps / secs / 1e6

    energy64 = None
    if not args.no_fp64:
        f64.step()  # warm-up
        f0 = f64.energy()
        secs64, phases64 = time_steps(nb, torch, f64, args.fp64_steps)
        fp64_rate = n * args.fp64_steps / secs64 / 1e6
        f1 = f64.energy()
        energy64 = (sum(f1) - sum(f0)) / abs(sum(f0))
        f64.close()

    by_theta = {}
    if args.workload == "plummer" and args.thetas:
        for th in (float(t) for t in args.thetas.split(",")):
            by_theta[str(th)] = accuracy_at(nb, torch, p, th, args, targets, exact)

    frames = frame_capture_timing(nb, torch, mixed, n)

    trav_ms = phases["traversal"]
    achieved = FLOPS_PER_INTERACTION * interactions / (trav_ms * 1e-3) / 1e12
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
        "fp64_value": None if fp64_rate is None else round(fp64_rate, 3),
        "speedup_vs_fp64": None if fp64_rate is None else round(value / fp64_rate, 3),
        "accuracy": {k: float(f"{v:.4g}") for k, v in acc.items()},
        "momentum_drift": float(np.abs(p1 - p0).max()),
        "energy": None if e0 is None else {
            "E_start": e0[0], "E_end": e1[0], "kinetic_start": e0[1],
            "potential_start": e0[2], "rel_drift": (e1[0] - e0[0]) / abs(e0[0]),
            "steps": args.warmup + args.steps,
            "fp64_rel_drift": energy64, "fp64_steps": args.fp64_steps},
        "accuracy_by_theta": by_theta,
        "phase_ms": {k: round(v, 4) for k, v in phases.items()},
        "fp64_phase_ms": None if fp64_rate is None else
        {k: round(v, 4) for k, v in phases64.items()},
        "frame_capture": frames,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 4),
                     "traffic": traversal_traffic() if (args.workload == "plummer" and
                                                        n == 4_194_304 and
                                                        args.theta == 0.5) else None,
                     "kernel": "walk_mixed_kernel",
                     "flops_per_interaction": FLOPS_PER_INTERACTION},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(p, args),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
