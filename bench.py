#!/usr/bin/env python3
"""One 3D lateral-viscosity step at lmax 128 to residual 1e-8 on one MI355X.

Workload (the 3D step configuration): the three-region Earth (fluid core to
3480 km, Maxwell mantle 5e20 / 1e22 Pa s split at 5701 km, elastic
lithosphere), 128 elements (70 / 56 / 2) x 6 GLL nodes, Gauss memory
points (280 memory nodes), lmax 128 on the 194 x 388 Gauss-Legendre grid,
eta = eta_ref exp(delta) with delta a seeded isotropic Gaussian random field
of angular power (1 + l/8)^-3 to l = 64, rms 0.7, AR(1) 0.9 across the
memory nodes; the exponential trapezoid, dt = 100 yr, the reference step
benchmark's zonal load rate; the fixed point to rtol 1e-8 with adaptive
Anderson mixing of depth 6. A "step" is one Viscoelastic3D::Step (StartStep,
the fixed point's passes, Finish) from the state the previous step left.

Prints one JSON line (the driver's contract). --gpus N > 1 exits: this build
runs on one GPU.
"""

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D lateral-viscosity steps/s at lmax128, tol 1e-8: learned+batched vs plain"
PEAK_HBM = 8.0e12       # B/s
PEAK_FP64 = 78.6e12     # flop/s, MI355X dense FP64 (vector and matrix)
# HBM bytes per launch (read + written), measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE on this build at lmax 128, scaled by a 1 GiB device copy (FETCH_SIZE
# x 2.0, WRITE_SIZE x 1.0); profiles/step_lmax128_hbm_traffic.md. Only valid for
# the default workload.
MEASURED_TRAFFIC = {
    "ring_pass": 1.4058e9,
    "response": 5.459e8,
    "legendre_analysis": 8.448e8,
    "legendre_synthesis": 7.159e8,
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--lmax", type=int, default=128)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--solver", choices=["fixed-point", "cg"], default="fixed-point",
                    help="how a step closes its system: the fixed point (the default, the "
                         "metric's) or preconditioned conjugate gradients")
    ap.add_argument("--solve-rtol", type=float, default=1e-11,
                    help="the conjugate gradients' relative residual (energy norm)")
    ap.add_argument("--realisations", type=int, default=1,
                    help="independent realisations (delta seeds) stepped side by side in one "
                         "operator (every pass kernel over all their memory nodes); value is "
                         "their aggregate steps/s, beside the same realisations one operator "
                         "at a time")
    ap.add_argument("--learned-train", type=int, default=64,
                    help="training steps (delta seeds) of the learned warm start's leg, "
                         "0 to skip it (the full learned configuration: 4000)")
    ap.add_argument("--learned-eval", type=int, default=6,
                    help="held-out delta seeds the learned leg evaluates (full: 200)")
    ap.add_argument("--learned-epochs", type=int, default=50)
    ap.add_argument("--extrapolate-order", type=int, default=4,
                    help="order of the extrapolated-start trajectory leg, 1 to 5 (0 skips it)")
    ap.add_argument("--side-by-side", type=int, default=4,
                    help="realisations of the leg that steps several delta seeds side by side "
                         "in one operator, plain and with extrapolated starts (0 skips it)")
    ap.add_argument("--uneven", type=int, default=0,
                    help="an uneven batch of this many independent realisations (lmax from "
                         "{32, 64, 96, 128, 192}, meshes of {64, 96, 128} elements x degree 5, "
                         "seeded), "
                         "stepped --steps times each by 4 worker threads; value is their "
                         "aggregate steps/s")
    ap.add_argument("--maxwell1d-steps", type=int, default=200,
                    help="steps of the batched 1D Maxwell evolution leg (lmax 256, 8 ice caps "
                         "ramped to zero, 100 yr; 0 skips it)")
    ap.add_argument("--maxwell1d-lmax", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph-leg", action="store_true",
                    help="skip the leg that replays every pass from a captured hipGraph")
    ap.add_argument("--transform-degrees", default="32,48,64,80,96,112",
                    help="the transform degrees of the leg that cuts the transforms with depth "
                         "(comma-separated; empty skips it)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run device copies of 1 GiB (the HBM byte counters' scale)")
    ap.add_argument("--band-solve", action="store_true",
                    help="only time the batched band Cholesky solve on one system of the "
                         "1D evolution's longest shape (n 1281, kd 17, 2 columns, row-major)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def work_models(op, lmax):
    """Algorithmic work per launch of the step's kernels: flops for the
    matrix-core kernels, HBM bytes for the sweeps."""
    tasks = op.export("tasks").reshape(-1, 5)
    first_node = {int(t[0]): int(t[3]) for t in tasks if int(t[1]) == 0}
    nodes, nt, nph = op.nodes, op.ntheta, op.nphi
    active = 0  # (block, m, l, node) coefficients kept by the degree truncation
    for spin in (-2, -1, 0):
        mlo = 0 if spin == 0 else -lmax
        for m in range(mlo, lmax + 1):
            for l in range(max(abs(m), abs(spin)), lmax + 1):
                active += nodes - first_node[l]
    legendre_flops = 4.0 * nt * active  # real table x complex coefficient, FMA = 2
    # longitude transforms: R sequences of P points per ring (spin 0: two
    # real rings per sequence); a direction costs 8 H kmax flops per sequence
    # (the symmetric split's cos and sin products over the H frequencies and
    # the kmax = floor((L + R - 1) / R) that the band limit leaves nonzero)
    R = 4 if nph % 4 == 0 else 2
    P = nph // R
    H = (P - 1) // 2
    kmax = min(H, (lmax + R - 1) // R)
    seq = nodes * nt * R * (1.0 + 1.0 + 0.5)  # the three blocks
    per_direction = seq * 8.0 * H * kmax
    # the pass: one launch, three blocks, both directions; the step's start
    # (two analyses) and finish (one synthesis): three launches each
    ring_pass_flops = 2.0 * per_direction
    ring_flops = 3.0 * per_direction / 6.0
    # the memory response: x = G f per degree task
    response_flops = sum(2.0 * t[4] * t[4] * 2 * (t[0] + 1) for t in tasks)
    S = op.spectral_doubles * 8.0
    # the pass's longitude stage moves the Fourier coefficients in and out
    # ([m][node][theta] complex: 2L + 1 orders for each complex block, L + 1
    # for spin 0) and reads the weight W once (one double per grid point):
    # its algorithmic bytes, 1.285 GB at lmax 128
    fourier = 16.0 * nodes * nt * (2 * (2 * lmax + 1) + (lmax + 1))
    ring_pass_bytes = 2.0 * fourier + 8.0 * nodes * nt * nph
    return {
        "legendre_synthesis": ("mfma", legendre_flops),
        "legendre_analysis": ("mfma", legendre_flops),
        "ring_pass": ("hbm", ring_pass_bytes),
        "ring_pass_flops": ("mfma", ring_pass_flops),
        "ring": ("mfma", ring_flops),
        "response": ("mfma", response_flops),
        "residual": ("hbm", 2.0 * S),
        "anderson": ("hbm", None),
    }


def physical_cores():
    """Physical cores of NUMA node 0 this process may run on (one per
    (package, core) pair of its CPUs), and the CPU counts: nproc and the
    affinity set."""
    allowed = os.sched_getaffinity(0)
    node_cpus = allowed
    try:
        text = open("/sys/devices/system/node/node0/cpulist").read().strip()
        node_cpus = set()
        for part in text.split(","):
            a, _, b = part.partition("-")
            node_cpus.update(range(int(a), int(b or a) + 1))
        node_cpus &= allowed
    except OSError:
        pass
    cores = set()
    for cpu in node_cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{cpu}/topology/"
            cores.add((open(base + "physical_package_id").read().strip(),
                       open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", cpu))
    return max(len(cores), 1), os.cpu_count() or 1, len(allowed)


def cpu_baseline(op, earth, dt, load, passes_per_step, rtol):
    """The oracle (the CPU restatement of the step: FP64, C with OpenMP) on the
    same workload, timed in a child process on the host's cores as the
    reference advises its runs (docs/running.md:9-55): OMP_NUM_THREADS the
    physical cores of one NUMA node (within this process's CPU share),
    OMP_PLACES=cores, OMP_PROC_BIND=close. One warm-up step, then the best of
    three steps, each capped at two passes and timed by phase: the per-pass
    and the once-per-step (StartStep + Finish) times give the CPU steps/s at
    the GPU run's mean passes per step (the same algorithm: the iteration
    counts agree pass for pass in the parity tests). The oracle's per-call
    set-up (band factorisations, tables, unit solves) is excluded, as the
    reference caches it per (dt, scheme)."""
    import subprocess
    import tempfile

    import numpy as np

    cores, nproc, allowed = physical_cores()
    share = int(os.environ.get("OMP_NUM_THREADS", cores))
    threads = max(1, min(cores, share))
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "workload.npz")
        np.savez(path, delta=op.delta_grid(), theta=op.theta, wtheta=op.wtheta,
                 lmax=op.lmax, dt=dt, rtol=rtol)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores",
                   OMP_PROC_BIND="close")
        log(f"cpu baseline: the oracle in a child process, {threads} threads "
            f"(node-0 physical cores {cores}, nproc {nproc}, affinity {allowed}) ...")
        run = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", path],
                             env=env, capture_output=True, text=True, timeout=900)
        if run.returncode != 0:
            raise RuntimeError("cpu baseline child failed:\n" + run.stderr[-2000:])
        r = json.loads(run.stdout.strip().splitlines()[-1])
    per_pass, once = r["per_pass_s"], r["once_s"]
    step_time = once + passes_per_step * per_pass
    return {
        "value": 1.0 / step_time,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "nproc": nproc,
        "sample": (f"the oracle (FP64 C/OpenMP restatement with direct longitude DFTs where the "
                   f"reference uses FFTW, so it overstates the reference's CPU time) on the full "
                   f"workload, {threads} threads (OMP_PLACES=cores, OMP_PROC_BIND=close; nproc "
                   f"{nproc}): best of {r['repeats']} steps capped at {r['passes']} passes after "
                   f"one warm-up, {per_pass:.2f} s/pass and {once:.2f} s StartStep+Finish "
                   f"(set-up {r['setup_s']:.1f} s per call excluded), scaled to "
                   f"{passes_per_step:.1f} passes/step"),
    }


def cpu_child(path):
    """The timed oracle steps of cpu_baseline, in this (child) process: no
    device work."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from step_problem import OracleStep, amplitudes, oracle_library

    import sl3dpp_amd.step as st

    w = np.load(path)
    lmax = int(w["lmax"])
    dt, rtol = float(w["dt"]), float(w["rtol"])
    earth = st.Earth(st.three_region_earth(70, 56, 2, 6))
    load, _ = st.zonal_load_rate(lmax, earth)
    oracle = oracle_library()
    p = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt, w["delta"],
                   theta=w["theta"], wtheta=w["wtheta"])
    cap, repeats = 2, 3
    best_pass = best_once = float("inf")
    setup = 0.0
    for k in range(1 + repeats):
        amp = amplitudes(load, lmax, (k + 1) * dt)
        info = p.step(dt, st.EXPONENTIAL_TRAPEZOID, amp, max_iterations=cap, rtol=rtol,
                      anderson_memory=6, anderson_rate=0.5)
        log(f"cpu child: step {k}: set-up {info.setup_ms / 1e3:.1f} s, start "
            f"{info.start_ms / 1e3:.2f} s, {info.iterations} passes {info.passes_ms / 1e3:.2f} s, "
            f"finish {info.finish_ms / 1e3:.2f} s")
        if k == 0:
            continue  # warm-up
        best_pass = min(best_pass, info.passes_ms / 1e3 / max(info.iterations, 1))
        best_once = min(best_once, (info.start_ms + info.finish_ms) / 1e3)
        setup = info.setup_ms / 1e3
    print(json.dumps({"per_pass_s": best_pass, "once_s": best_once, "repeats": repeats,
                      "passes": cap, "setup_s": setup}), flush=True)
    return 0


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--cpu-child":
        return cpu_child(sys.argv[2])
    if len(sys.argv) >= 3 and sys.argv[1] == "--uneven-cpu-child":
        return uneven_cpu_child(sys.argv[2:])
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported: the step "
              "would shard by radius (contiguous memory-node slices per GPU, one separator "
              "sum per pass), which is not built", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import sl3dpp_amd
    import sl3dpp_amd.step as st

    assert torch.cuda.is_available(), "bench.py needs the HIP device"
    if args.copy_calibration:
        # a plain copy of known size: 1 GiB read and 1 GiB written per launch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        # the practical HBM rate: the same copy timed
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(10):
            dst.copy_(src)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / 10
        log(f"copy calibration: 13 copies of 1 GiB; {ms:.3f} ms each, "
            f"{2 * 2 ** 30 / (ms * 1e-3) / 1e12:.2f} TB/s (read + write)")
        return 0
    if args.band_solve:
        return band_solve_probe()
    if args.realisations > 1:
        return batched(args)
    if args.uneven > 0:
        return uneven(args)
    lmax = args.lmax
    earth = st.Earth(st.three_region_earth(70, 56, 2, 6))
    length, tscale, mass = earth.scales()
    dt = 100.0 * st.YEAR / tscale
    cg = args.solver == "cg"
    opts = st.StepOptions.default(max_iterations=200, rtol=args.rtol, anderson_memory=6,
                                  anderson_rate=0.5, solve_rtol=args.solve_rtol,
                                  solver=st.CONJUGATE_GRADIENTS if cg else st.FIXED_POINT)
    log("building the operator ...")
    t0 = time.perf_counter()
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    coef = st.delta_coefficients(min(64, lmax), op.nodes, seed=2026, rms=0.7, ar=0.9,
                                 lmax=lmax)
    op.set_delta(op.delta_field(coef))
    load, _ = st.zonal_load_rate(lmax, earth)
    log(f"operator: {op.nodes} memory nodes, grid {op.ntheta} x {op.nphi}, "
        f"{time.perf_counter() - t0:.1f} s")
    state = op.zero_state()
    for k in range(args.warmup):
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        log(f"warmup step {k}: {info.iterations} passes ({info.mixed_passes} mixed), "
            f"residual {info.residual:.2e}, {info.start_ms + info.passes_ms + info.finish_ms:.1f} ms")
    torch.cuda.synchronize()
    sl3dpp_amd.synchronize()
    sl3dpp_amd.profile_reset()
    # inside the timed region only the dominant kernel carries events (its
    # average launch time is the roofline's); every kernel's events would
    # cost the step about 3%
    dominant = "ring_pass"
    sl3dpp_amd.profile_select([dominant])
    sl3dpp_amd.profile_enable(True)
    passes, mixed, converged = [], [], []
    torch.cuda.synchronize()
    sl3dpp_amd.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        passes.append(info.iterations)
        mixed.append(info.mixed_passes)
        converged.append(bool(info.converged))
        log(f"step {k}: {info.iterations} passes ({info.mixed_passes} mixed), residual "
            f"{info.residual:.2e}, start {info.start_ms:.2f} ms, passes {info.passes_ms:.2f} ms, "
            f"finish {info.finish_ms:.2f} ms")
    sl3dpp_amd.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sl3dpp_amd.profile_enable(False)
    timed_kernels = sl3dpp_amd.profile_read()
    plain_final = state.export("coefficients")
    # the per-kernel breakdown: two steps from rest with every kernel timed,
    # outside the timed region and on a state of their own (the timed
    # trajectory's final state is compared below)
    sl3dpp_amd.profile_reset()
    sl3dpp_amd.profile_select(None)
    sl3dpp_amd.profile_enable(True)
    probe = op.zero_state()
    for k in range(2):
        op.step(probe, dt, st.EXPONENTIAL_TRAPEZOID, load)
    sl3dpp_amd.profile_enable(False)
    kernels = sl3dpp_amd.profile_read()
    del probe
    steps_per_s = args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    # dominant kernel and its roofline
    models = work_models(op, lmax)
    top = max(kernels.items(), key=lambda kv: kv[1][0])
    if top[0] != dominant:
        log(f"note: the breakdown's largest kernel is {top[0]}, not {dominant}")
    name = dominant
    total_ms, launches = timed_kernels.get(dominant, (0.0, 0))
    roofline = None
    bound, work = models.get(name, ("hbm", None))
    traffic = MEASURED_TRAFFIC.get(name) if lmax == 128 else None
    avg_s = total_ms / max(launches, 1) / 1e3
    if work is not None and avg_s > 0:
        achieved = work / avg_s
        if bound == "mfma":
            roofline = {"bound": "mfma", "kernel": name, "achieved": achieved / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64, "traffic": traffic,
                        "work_per_launch": work, "avg_launch_ms": avg_s * 1e3}
        else:
            roofline = {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9,
                        "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM, "traffic": traffic,
                        "work_per_launch": work, "avg_launch_ms": avg_s * 1e3}
            flops = models.get(name + "_flops")
            if flops is not None:
                # the same launch against the FP64 matrix peak: the builder's
                # matrix-DFT products (O(P^2) per sequence), not an algorithmic
                # FFT count
                roofline["fp64_flops_per_launch"] = flops[1]
                roofline["fp64_frac"] = flops[1] / avg_s / PEAK_FP64
    breakdown = {k: {"total_ms": round(v[0], 3), "launches": v[1]} for k, v in kernels.items()}
    t = op.export("tuning")
    tuning = {"legendre_analysis_rows": int(t[0]), "legendre_synthesis_theta": int(t[1]),
              "analysis_ms": {str(r): round(float(v), 4)
                              for r, v in zip((16, 32, 64, 128), t[2:6])},
              "synthesis_ms": {str(r): round(float(v), 4)
                               for r, v in zip((32, 64, 112, 208), t[6:10])}}
    cpu = None
    if not args.no_cpu_baseline and not cg:
        cpu = cpu_baseline(op, earth, dt, load, float(np.mean(passes)), args.rtol)
    graphs = None
    if not cg and not args.no_graph_leg:
        graphs = graph_passes_leg(op, dt, load, args, steps_per_s, plain_final)
    cut = None
    degrees = [int(d) for d in args.transform_degrees.split(",") if d.strip()]
    degrees = [d for d in degrees if 2 <= d < lmax]
    if degrees and not cg:
        cut = transform_degrees_leg(earth, lmax, coef, opts, degrees, dt, load, args,
                                    steps_per_s, plain_final)
    extrapolated = None
    if args.extrapolate_order > 0 and not cg:
        extrapolated = extrapolated_leg(op, dt, load, args, steps_per_s, plain_final)
    side = None
    if args.side_by_side > 1 and not cg:
        side = side_by_side_leg(earth, lmax, dt, load, args, steps_per_s, plain_final)
    learned = None
    if args.learned_train > 0 and not cg:
        learned = learned_leg(op, dt, load, args)
    maxwell1d = None
    if args.maxwell1d_steps > 0:
        maxwell1d = maxwell1d_leg(args)
    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("one 3D lateral-viscosity step, lmax 128, grid 194x388, 128 elements "
                         "(70/56/2) x 6 GLL nodes, 280 Gauss memory nodes, delta GRF "
                         "(1+l/8)^-3 to l=64 rms 0.7 AR(1) 0.9, exponential trapezoid "
                         "dt 100 yr, zonal load rate, "
                         + (f"conjugate gradients, solve_rtol {args.solve_rtol:g}" if cg else
                            f"Anderson depth 6, rtol {args.rtol:g}")),
            "variant": ("preconditioned conjugate gradients (one realisation)" if cg else
                        "plain GPU fixed point (one realisation); the learned warm start's "
                        "first steps from rest are under 'learned'"),
            "lmax": lmax, "grid": [op.ntheta, op.nphi], "memory_nodes": op.nodes,
            "passes_per_step": passes, "mixed_passes": mixed, "converged": converged,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "tuning": tuning,
        "graph_passes": graphs,
        "transforms_cut_with_depth": cut,
        "extrapolated_start": extrapolated,
        "side_by_side": side,
        "learned": learned,
        "maxwell1d": maxwell1d,
        "kernels": breakdown,
    }
    print(json.dumps(out), flush=True)
    return 0


def transform_degrees_leg(earth, lmax, coef, opts, degrees, dt, load, args, plain_steps_per_s,
                          plain_final):
    """The timed trajectory again on an operator whose transforms are cut
    with depth (Options::transform_degrees): each memory element's nodes at
    the least listed degree at or above the largest degree reaching it, on a
    grid of that degree, with the same lateral field (its coefficients
    synthesised on every node's own grid). A different discretisation of
    the lateral product, so its final state is compared with the plain
    trajectory's, not equal to it."""
    import numpy as np
    import torch

    import sl3dpp_amd
    import sl3dpp_amd.step as st

    o = st.StepOptions()
    ctypes_copy = bytes(opts)
    ctypes.memmove(ctypes.addressof(o), ctypes_copy, len(ctypes_copy))
    o.transform_degrees = degrees
    op = st.Viscoelastic3D(earth, lmax, None, o)
    op.set_delta(op.delta_field(coef))
    state = op.zero_state()
    for k in range(args.warmup):
        op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    passes = []
    torch.cuda.synchronize()
    sl3dpp_amd.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        passes.append(info.iterations)
        assert info.converged, "a step with the transforms cut did not converge"
    sl3dpp_amd.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    final = state.export("coefficients")
    diff = float(np.max(np.abs(final - plain_final)) / np.max(np.abs(plain_final)))
    full_points = op.nodes * op.ntheta * op.nphi
    # the per-kernel split: two profiled steps from rest, outside the timed region
    sl3dpp_amd.profile_reset()
    sl3dpp_amd.profile_select(None)
    sl3dpp_amd.profile_enable(True)
    probe = op.zero_state()
    for k in range(2):
        op.step(probe, dt, st.EXPONENTIAL_TRAPEZOID, load)
    sl3dpp_amd.profile_enable(False)
    kernels = {k: {"total_ms": round(v[0], 3), "launches": v[1]}
               for k, v in sl3dpp_amd.profile_read().items()}
    out = {
        "degrees": degrees,
        "blocks": [{"lmax": k["lmax"], "nodes": k["nodes"], "grid": [k["ntheta"], k["nphi"]]}
                   for k in op.blocks],
        "grid_points_vs_full": op.grid_values / full_points,
        "steps_per_s": args.steps / elapsed,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "passes_per_step": passes,
        "speedup_vs_plain": (args.steps / elapsed) / plain_steps_per_s,
        "final_state_vs_plain": diff,
        "kernels": kernels,
        "what": ("the timed trajectory again from the zero state with the transforms cut with "
                 "depth (each memory element at the least listed degree covering the largest "
                 "degree that reaches it, on that degree's grid); final_state_vs_plain = max "
                 "|difference| / max |plain| of the final strain coefficients"),
    }
    log(f"transforms cut with depth {degrees}: {out['steps_per_s']:.2f} steps/s, "
        f"x{out['speedup_vs_plain']:.3f} of plain, grid points x{out['grid_points_vs_full']:.3f}, "
        f"final state vs plain {diff:.1e}")
    return out


def graph_passes_leg(op, dt, load, args, plain_steps_per_s, plain_final):
    """The timed trajectory again (zero state, the same warm-up and timed
    steps) with every pass replayed from a captured hipGraph; its final state
    equals the plain trajectory's bit for bit."""
    import numpy as np
    import torch

    import sl3dpp_amd
    import sl3dpp_amd.step as st

    op.set_options(graph_passes=1)
    try:
        state = op.zero_state()
        for k in range(args.warmup):
            op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        passes = []
        torch.cuda.synchronize()
        sl3dpp_amd.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            info = op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
            passes.append(info.iterations)
            assert info.converged, "a step of graph passes did not converge"
        sl3dpp_amd.synchronize()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        final = state.export("coefficients")
    finally:
        op.set_options(graph_passes=0)
    out = {
        "steps_per_s": args.steps / elapsed,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "passes_per_step": passes,
        "speedup_vs_plain": (args.steps / elapsed) / plain_steps_per_s,
        "final_state_equals_plain": bool(np.array_equal(final, plain_final)),
        "what": ("the timed trajectory again from the zero state with each pass (synthesis, "
                 "ring pass, analysis, force, response, strain) replayed from a hipGraph "
                 "captured on its first use"),
    }
    log(f"graph passes: {out['steps_per_s']:.2f} steps/s, x{out['speedup_vs_plain']:.3f} of "
        f"plain, final state equal: {out['final_state_equals_plain']}")
    return out


def extrapolated_leg(op, dt, load, args, plain_steps_per_s, plain_final):
    """The same trajectory (zero state, the same warm-up and timed steps)
    with each step warm-started from the Lagrange extrapolation in time of
    the last converged steps' strain coefficients (order 3: the last four),
    converged to the same tolerance; its final state against the plain
    trajectory's."""
    import numpy as np
    import torch

    import sl3dpp_amd
    import sl3dpp_amd.step as st

    state = op.zero_state()
    ws = st.ExtrapolatedStart(op, order=args.extrapolate_order)
    for k in range(args.warmup):
        ws.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    passes = []
    torch.cuda.synchronize()
    sl3dpp_amd.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        info = ws.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        passes.append(info.iterations)
        assert info.converged, "an extrapolated-start step did not converge"
    sl3dpp_amd.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    final = state.export("coefficients")
    diff = float(np.max(np.abs(final - plain_final)) / np.max(np.abs(plain_final)))
    # the same trajectory further on, once its history is full and the
    # transients of the first steps have decayed: steps 11-15 after 3
    # more untimed ones (the plain step's count is steady from the start)
    done = args.warmup + args.steps
    for k in range(max(0, 10 - done)):
        ws.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    later = []
    torch.cuda.synchronize()
    sl3dpp_amd.synchronize()
    t0 = time.perf_counter()
    for k in range(5):
        info = ws.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        later.append(info.iterations)
        assert info.converged, "an extrapolated-start step did not converge"
    sl3dpp_amd.synchronize()
    torch.cuda.synchronize()
    later_s = time.perf_counter() - t0
    out = {
        "order": args.extrapolate_order,
        "steps_per_s": args.steps / elapsed,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "passes_per_step": passes,
        "speedup_vs_plain": (args.steps / elapsed) / plain_steps_per_s,
        "final_state_vs_plain": diff,
        "later_steps": {"first": max(done, 10) + 1, "count": 5, "passes_per_step": later,
                        "steps_per_s": 5 / later_s,
                        "speedup_vs_plain": (5 / later_s) / plain_steps_per_s},
        "what": ("the timed trajectory again from the zero state, each step started from the "
                 "extrapolation in time of the last converged strains (Converge's start), "
                 "converged to the same tolerance; final_state_vs_plain = max |difference| / "
                 "max |plain| of the final strain coefficients"),
    }
    log(f"extrapolated start later on: {5 / later_s:.2f} steps/s, passes {later}")
    log(f"extrapolated start (order {args.extrapolate_order}): {out['steps_per_s']:.2f} steps/s, "
        f"passes {passes}, x{out['speedup_vs_plain']:.2f} of plain, final state vs plain {diff:.1e}")
    return out


def side_by_side_leg(earth, lmax, dt, load, args, plain_steps_per_s, plain_final):
    """The batched form: R realisations (delta seeds 2026 .. 2026 + R - 1,
    the first the timed trajectory's own) side by side in one operator, every
    pass kernel over all their memory nodes, each converging on its own.
    The same trajectory (zero state, the same warm-up and timed steps), first
    plain, then with every step started from the extrapolation in time of
    the last converged strains (all realisations' columns at once), with
    steps 11-15 as in the extrapolated leg. Rates are realisation-steps/s;
    ratios against the plain single-realisation value. The first
    realisation's plain final state must equal the timed trajectory's, bit
    for bit."""
    import numpy as np
    import torch

    import sl3dpp_amd
    import sl3dpp_amd.step as st

    R = args.side_by_side
    opts = st.StepOptions.default(max_iterations=200, rtol=args.rtol, anderson_memory=6,
                                  anderson_rate=0.5, realisations=R)
    op = st.Viscoelastic3D(earth, lmax, None, opts)
    N = op.radial_nodes
    coef = np.concatenate([st.delta_coefficients(min(64, lmax), N, seed=2026 + i, rms=0.7,
                                                 ar=0.9, lmax=lmax) for i in range(R)])
    op.set_delta(op.delta_field(coef))

    def timed(step, state, count):
        passes = []
        torch.cuda.synchronize()
        sl3dpp_amd.synchronize()
        t0 = time.perf_counter()
        for _ in range(count):
            info = step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
            assert info.converged, "a side-by-side step did not converge"
            passes.append([i.iterations for i in op.realisation_infos()])
        sl3dpp_amd.synchronize()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, passes

    state = op.zero_state()
    for _ in range(args.warmup):
        op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    plain_s, plain_passes = timed(op.step, state, args.steps)
    first = state.export("coefficients").reshape(-1)
    # the reference layout is per block [node][coefficient]: realisation 0's
    # share of each block is its first N nodes' rows
    same = True
    at, at1 = 0, 0
    for b in range(3):
        size = 2 * ((lmax + 1) ** 2 - (4 if b == 0 else 1) if b < 2 else (lmax + 1) * (lmax + 2) // 2)
        block = first[at:at + R * N * size].reshape(R * N, size)
        same = same and np.array_equal(block[:N].ravel(), plain_final[at1:at1 + N * size])
        at += R * N * size
        at1 += N * size
    del state
    state = op.zero_state()
    ws = st.ExtrapolatedStart(op, order=max(args.extrapolate_order, 1))
    for _ in range(args.warmup):
        ws.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    extra_s, extra_passes = timed(ws.step, state, args.steps)
    for _ in range(max(0, 10 - args.warmup - args.steps)):
        ws.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
    later_s, later_passes = timed(ws.step, state, 5)
    total = R * args.steps
    out = {
        "realisations": R,
        "plain_steps_per_s": total / plain_s,
        "plain_passes": plain_passes,
        "plain_speedup_vs_plain": total / plain_s / plain_steps_per_s,
        "first_equals_timed_trajectory": bool(same),
        "extrapolated_order": max(args.extrapolate_order, 1),
        "extrapolated_steps_per_s": total / extra_s,
        "extrapolated_passes": extra_passes,
        "extrapolated_speedup_vs_plain": total / extra_s / plain_steps_per_s,
        "later_steps": {"first": max(args.warmup + args.steps, 10) + 1, "count": 5,
                        "steps_per_s": 5 * R / later_s, "passes": later_passes,
                        "speedup_vs_plain": 5 * R / later_s / plain_steps_per_s},
        "what": ("R realisations side by side in one operator (delta seeds 2026..), the timed "
                 "trajectory plain and with extrapolated starts; rates in realisation-steps/s, "
                 "ratios against the single-realisation plain value"),
    }
    log(f"side by side ({R}): plain {out['plain_steps_per_s']:.2f} realisation-steps/s "
        f"(x{out['plain_speedup_vs_plain']:.2f}), extrapolated {out['extrapolated_steps_per_s']:.2f} "
        f"(x{out['extrapolated_speedup_vs_plain']:.2f}), later {out['later_steps']['steps_per_s']:.2f} "
        f"(x{out['later_steps']['speedup_vs_plain']:.2f}); first realisation equal: {same}")
    assert same, "the first side-by-side realisation differs from the timed trajectory"
    return out


def band_solve_probe(n=1281, kd=17, nrhs=2, reps=20):
    """The band Cholesky solve alone on one diagonally dominant system of
    the 1D evolution's longest shape; prints microseconds per launch and
    nanoseconds per row of the two sweeps (HIP events on the library
    stream)."""
    import ctypes

    import numpy as np
    import torch

    import sl3dpp_amd
    from sl3dpp_amd import check, lib

    I, I64 = ctypes.c_int, ctypes.c_int64
    rng = np.random.default_rng(1)
    ld = kd + 1
    ab = np.zeros((n, ld))
    ab[:, :kd] = 0.1 * rng.normal(size=(n, kd))
    ab[:, kd] = 4.0 + 0.2 * kd
    ab_dev = torch.tensor(ab.ravel(), device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    one = lambda v, t: (t * 1)(v)  # noqa: E731
    check(lib.sl3_band_cholesky_factor_batched(1, one(n, I), one(kd, I), one(0, I64),
                                               ctypes.c_void_p(ab_dev.data_ptr()),
                                               ctypes.c_void_p(info.data_ptr())))
    b = torch.randn(n * nrhs, dtype=torch.float64, device="cuda")
    args = (1, one(n, I), one(kd, I), one(0, I64), ctypes.c_void_p(ab_dev.data_ptr()),
            one(nrhs, I), one(0, I64), one(nrhs, I64), one(1, I64), ctypes.c_void_p(b.data_ptr()))
    check(lib.sl3_band_cholesky_solve_batched(*args))
    sl3dpp_amd.synchronize()
    sl3dpp_amd.profile_reset()
    sl3dpp_amd.profile_select(["band_solve"])
    sl3dpp_amd.profile_enable(True)
    for _ in range(reps):
        check(lib.sl3_band_cholesky_solve_batched(*args))
    sl3dpp_amd.profile_enable(False)
    total_ms, launches = sl3dpp_amd.profile_read()["band_solve"]
    us = 1e3 * total_ms / launches
    log(f"band solve n {n} kd {kd} nrhs {nrhs}: {us:.1f} us per launch, "
        f"{1e3 * us / (2 * n):.1f} ns per row")
    return 0


def ice_caps_pattern(lmax, earth, caps=8, seed=1956):
    """The surface mass of a seeded sum of Gaussian ice caps as complex
    coefficients per (l, m >= 0) of the orthonormal spherical harmonics
    (non-dimensional surface mass; (Re, Im) doubles at l(l+1)/2 + m). Cap
    k: thickness H_k exp(-(gamma / a_k)^2) of ice (917 kg m^-3) at angular
    distance gamma from a uniform random centre, a_k uniform in 5-15
    degrees, H_k uniform in 1-3 km. Its Legendre series h_l P_l(cos gamma)
    by Gauss-Legendre quadrature, rotated to the centre by the addition
    theorem: c_lm = h_l 4 pi / (2l + 1) conj(Y_lm(centre))."""
    import numpy as np
    from scipy.special import sph_harm_y

    length, time, mass = earth.scales()
    sigma = mass / length ** 2
    rng = np.random.default_rng(seed)
    x, w = np.polynomial.legendre.leggauss(4 * (lmax + 1))
    gamma = np.arccos(x)
    # P_l(x) for l = 0..lmax by the three-term recurrence
    P = np.zeros((lmax + 1, x.size))
    P[0] = 1.0
    if lmax > 0:
        P[1] = x
    for l in range(2, lmax + 1):
        P[l] = ((2 * l - 1) * x * P[l - 1] - (l - 1) * P[l - 2]) / l
    ls = np.arange(lmax + 1)
    out = np.zeros((lmax + 1) * (lmax + 2) // 2, dtype=np.complex128)
    ll = np.concatenate([np.full(l + 1, l) for l in ls])
    mm = np.concatenate([np.arange(l + 1) for l in ls])
    for _ in range(caps):
        a = np.radians(rng.uniform(5.0, 15.0))
        H = rng.uniform(1000.0, 3000.0)
        theta0 = np.arccos(rng.uniform(-1.0, 1.0))
        phi0 = rng.uniform(0.0, 2.0 * np.pi)
        h = 917.0 * H * np.exp(-(gamma / a) ** 2) / sigma
        hl = (2 * ls + 1) / 2.0 * (P * (w * h)).sum(axis=1)
        Y = sph_harm_y(ll, mm, theta0, phi0)
        out += (hl[ll] * 4.0 * np.pi / (2 * ll + 1)) * np.conj(Y)
    out[mm == 0] = out[mm == 0].real
    return np.stack([out.real, out.imag], axis=-1).ravel()


def maxwell1d_leg(args):
    """The batched 1D Maxwell evolution (ViscoelasticOperator's Step for
    every spheroidal degree, sl3_maxwell1d_*): the same Earth (70 / 56 / 2
    elements x 6 GLL nodes, 5e20 / 1e22 Pa s split at 5701 km, elastic
    lithosphere), lmax 256 with every order as columns, the exponential
    trapezoid, 100 yr steps under 8 seeded Gaussian ice caps ramped linearly
    to zero over the run; from rest with the elastic start (untimed), then
    the timed steps, each one replay of the captured step graph. Then the
    same steps as stream launches (the graph's gain), and 10 profiled
    stream-launched steps for the per-kernel split and the roofline of the
    largest."""
    import numpy as np
    import torch

    import sl3dpp_amd
    import sl3dpp_amd.step as st
    from sl3dpp_amd import radial

    lmax, steps = args.maxwell1d_lmax, args.maxwell1d_steps
    earth = st.Earth(st.three_region_earth(70, 56, 2, 6))
    length, tscale, mass = earth.scales()
    dt = 100.0 * st.YEAR / tscale
    t0 = time.perf_counter()
    op = radial.ViscoelasticOperator(earth, lmax)
    op.set_load(ice_caps_pattern(lmax, earth), np.linspace(1.0, 0.0, steps + 1))
    op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, 0)  # the elastic start, both flavours built
    op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, 1, graph=True)  # the graph captured
    setup_s = time.perf_counter() - t0
    log(f"maxwell1d: lmax {lmax}, {op.nodes} memory nodes, {op.memory_doubles} memory and "
        f"{op.solution_doubles} solution doubles, set-up {setup_s:.1f} s")
    timings = {}
    for graph in (True, False):
        op.reset()
        op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, steps, graph=graph)
        timings[graph] = time.perf_counter() - t0
        if graph:
            final = op.export("memory")
            surface = op.surface(radial.U).cpu().numpy()
    # the per-kernel split: 10 stream-launched steps with events
    op.reset()
    op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, 0)
    sl3dpp_amd.synchronize()
    sl3dpp_amd.profile_reset()
    sl3dpp_amd.profile_select(None)
    sl3dpp_amd.profile_enable(True)
    op.evolve(dt, st.EXPONENTIAL_TRAPEZOID, 10, graph=False)
    sl3dpp_amd.profile_enable(False)
    kernels = sl3dpp_amd.profile_read()
    M, S = 8.0 * op.memory_doubles, 8.0 * op.solution_doubles
    # algorithmic HBM bytes per launch: the force reads m and d and writes
    # every right-hand side; the solves read and write them once (the
    # factors, about 50 MB, left out); the update reads the right-hand
    # sides and reads and writes m and d
    work = {"maxwell_force": 2 * M + S, "band_solve": 2 * S, "maxwell_update": S + 4 * M}
    split = {}
    for name, (total_ms, launches) in kernels.items():
        entry = {"total_ms": round(total_ms, 3), "launches": launches}
        if name in work and launches > 0:
            # the band solves are two launches a step (Cholesky, LU)
            per_step = total_ms / 10.0 * 1e-3
            entry["hbm_frac"] = work[name] / per_step / PEAK_HBM
        split[name] = entry
    top = max((k for k in split if k in work), key=lambda k: split[k]["total_ms"])
    per_step = split[top]["total_ms"] / 10.0 * 1e-3
    total = timings[True]
    tuning = op.export("tuning")
    return {
        "workload": (f"batched 1D Maxwell evolution, lmax {lmax} (every order of every "
                     "spheroidal degree), 128 elements (70/56/2) x 6 GLL nodes, 280 Gauss "
                     f"memory nodes, exponential trapezoid, {steps} steps of 100 yr, 8 seeded "
                     "Gaussian ice caps (5-15 deg, 1-3 km) ramped linearly to zero"),
        "steps": steps,
        "steps_per_s": steps / total,
        "ms_per_step": 1e3 * total / steps,
        "seconds": total,
        "stream_launch_ms_per_step": 1e3 * timings[False] / steps,
        "graph_speedup": timings[False] / total,
        "setup_s": setup_s,
        "memory_doubles": op.memory_doubles,
        "solution_doubles": op.solution_doubles,
        "roofline": {"bound": "hbm", "kernel": top, "achieved": work[top] / per_step / 1e9,
                     "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": work[top] / per_step / PEAK_HBM, "work_per_step": work[top]},
        "kernels": split,
        "band_panel": {"rows": int(tuning[0]),
                       "ms_per_solve": {str(r): round(float(v), 4)
                                        for r, v in zip((4, 8, 16), tuning[1:])}},
        "final_memory_rms": float(np.sqrt(np.mean(final ** 2))),
        "surface_u_l2_m0": float(surface[2 * 3] * length),
    }


def learned_leg(op, dt, load, args):
    """The learned warm start on first steps from rest (the learned
    configuration's protocol): training triples made by the solver (one
    converged first step per delta seed 5000.., the delta distribution of
    the timed workload), the per-node network (delta coefficients to l 32 and
    their column mean, the radial response to l 8, radius -> the correction
    to l 32; hidden width 512, four layers, BF16 products, FP32
    accumulation) trained by Adam at lr 1e-3 on the library's kernels, then
    held-out seeds 9000.. stepped cold (the plain fixed point from the step's
    own start), from the radial reference alone, and from the learned start
    (guarded), each to the same tolerance. Steps/s of the cold and learned
    steps over the same seeds; the learned time includes inference."""
    import numpy as np

    import sl3dpp_amd.learned as ln
    import sl3dpp_amd.step as st

    t0 = time.perf_counter()
    ws = ln.WarmStart(op, dt, st.EXPONENTIAL_TRAPEZOID, load, lin=32, lrad=8, lcut=32,
                      width=512, layers=4, max_rows=1024, seed=7, context=True)
    sample = ln.delta_sampler(op)
    X, T, infos = ws.make_triples(sample, list(range(5000, 5000 + args.learned_train)), log=log)
    triples_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ws.fit(X, T, infos, epochs=args.learned_epochs, lr=1e-3, log=log)
    fit_s = time.perf_counter() - t0
    del X, T
    ln.evaluate(ws, sample, [8999], radial=False)  # untimed: first-call set-up
    records = ln.evaluate(ws, sample, list(range(9000, 9000 + args.learned_eval)), log=log)
    s = ln.summary(records)
    s.update({
        "train_steps": args.learned_train, "epochs": args.learned_epochs,
        "triples_s": round(triples_s, 2), "fit_s": round(fit_s, 2),
        "final_train_loss": ws.history[-1], "guard": ws.guard,
        "network": {"widths": ws.mlp.widths, "rows_per_step": ws.nodes,
                    "chunks": [[int(v) for v in t] for t in ws.mlp.chunks()]},
        "protocol": ("first steps from rest, held-out delta seeds; cold = the plain fixed point "
                     "from the step's own start; radial = from the delta-0 response alone"),
    })
    log("learned: " + json.dumps({k: v for k, v in s.items() if k != "network"}))
    return s


def uneven(args, workers=4, seed=48):
    """An uneven batch: realisation i has lmax and a radial mesh drawn from
    the sets above (seed 48), delta seed 3000 + i (rms 0.7, AR(1) 0.9, to
    l = min(64, lmax)), the benchmark's load, scheme and tolerance. The
    realisations of one (lmax, mesh) go side by side into one operator
    (every pass kernel over all their memory nodes, each realisation
    converging on its own); `workers` host threads (one library stream
    each) take these operators from a queue, the costliest first, and step
    each args.steps times after args.warmup untimed steps: the value. Beside
    it, each realisation in an operator of its own (made, stepped and freed
    one after another: all of them at once would not fit the GPU beside the
    side-by-side ones) takes the same steps, timed, and must end in the same
    state as its side-by-side self, bit for bit."""
    import queue
    import threading

    import numpy as np
    import torch

    import sl3dpp_amd.step as st

    rng = np.random.default_rng(seed)
    meshes = {64: (35, 27, 2), 96: (52, 42, 2), 128: (70, 56, 2)}
    earths = {e: st.Earth(st.three_region_earth(*meshes[e], 6)) for e in meshes}
    cases = []
    for i in range(args.uneven):
        lmax = int(rng.choice([32, 64, 96, 128, 192]))
        elements = int(rng.choice([64, 96, 128]))
        _, tscale, _ = earths[elements].scales()
        load, _ = st.zonal_load_rate(lmax, earths[elements])
        cases.append({"lmax": lmax, "elements": elements, "earth": earths[elements],
                      "dt": 100.0 * st.YEAR / tscale, "load": load, "seed": 3000 + i,
                      "passes": []})

    def options(realisations):
        return st.StepOptions.default(max_iterations=200, rtol=args.rtol, anderson_memory=6,
                                      anderson_rate=0.5, realisations=realisations)

    def coefficients(c, nodes):
        return st.delta_coefficients(min(64, c["lmax"]), nodes, seed=c["seed"], rms=0.7, ar=0.9,
                                     lmax=c["lmax"])

    def single(c):
        """Realisation c in an operator of its own."""
        op = st.Viscoelastic3D(c["earth"], c["lmax"], None, options(1))
        op.set_delta(op.delta_field(coefficients(c, op.nodes)))
        return op

    # the realisations of one (lmax, mesh) side by side
    t0 = time.perf_counter()
    groups = {}
    for i, c in enumerate(cases):
        groups.setdefault((c["lmax"], c["elements"]), []).append(i)
    batches = []
    for (lmax, elements), members in groups.items():
        first = cases[members[0]]
        op = st.Viscoelastic3D(earths[elements], lmax, None, options(len(members)))
        N = op.radial_nodes
        for i in members:
            cases[i]["nodes"] = N
        op.set_delta(op.delta_field(np.concatenate([coefficients(cases[i], N) for i in members])))
        state = op.zero_state()
        for _ in range(args.warmup):
            op.step(state, first["dt"], st.EXPONENTIAL_TRAPEZOID, first["load"])
        batches.append({"lmax": lmax, "elements": elements, "members": members, "op": op,
                        "state": state, "dt": first["dt"], "load": first["load"]})
    # the costliest first: the queue's tail is then short
    batches.sort(key=lambda b: -len(b["members"]) * b["lmax"] ** 3 * b["elements"])
    torch.cuda.synchronize()
    log(f"uneven batch: {len(cases)} realisations in {len(batches)} side-by-side operators "
        f"({[len(b['members']) for b in batches]}) set up in {time.perf_counter() - t0:.1f} s")

    work = queue.Queue()
    for b in batches:
        work.put(b)

    def worker():
        while True:
            try:
                b = work.get_nowait()
            except queue.Empty:
                return
            for _ in range(args.steps):
                b["op"].step(b["state"], b["dt"], st.EXPONENTIAL_TRAPEZOID, b["load"])
                for k, info in enumerate(b["op"].realisation_infos()):
                    cases[b["members"][k]]["passes"].append(info.iterations)
                    assert info.converged, "a realisation's step did not converge"

    threads = [threading.Thread(target=worker) for _ in range(workers)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    total = len(cases) * args.steps
    log(f"uneven batch: {total} steps in {elapsed:.2f} s side by side, {workers} workers "
        f"({total / elapsed:.1f} steps/s)")

    # each realisation alone: the same steps, timed one after another
    sequential, same = 0.0, True
    for b in batches:
        N = b["op"].radial_nodes
        side = b["state"].export("coefficients_device").reshape(-1, len(b["members"]) * N, 2)
        for k, i in enumerate(b["members"]):
            c = cases[i]
            op = single(c)
            state = op.zero_state()
            for _ in range(args.warmup):
                op.step(state, c["dt"], st.EXPONENTIAL_TRAPEZOID, c["load"])
            own = []
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                own.append(op.step(state, c["dt"], st.EXPONENTIAL_TRAPEZOID, c["load"]).iterations)
            torch.cuda.synchronize()
            sequential += time.perf_counter() - t0
            same = same and own == c["passes"] and np.array_equal(
                side[:, k * N:(k + 1) * N, :].ravel(), state.export("coefficients_device"))
            del state, op
    log(f"uneven batch: one at a time {total / sequential:.1f} steps/s, side by side "
        f"{total / elapsed:.1f} steps/s; equal to each alone: {same}")
    assert same, "a side-by-side realisation differs from its operator alone"
    del batches, work
    torch.cuda.empty_cache()
    cpu, errors = None, None
    if not args.no_cpu_baseline:
        cpu, errors = uneven_cpu_baseline(cases, args, single)
    out = {
        "metric": METRIC + " (uneven batch of independent realisations)",
        "value": total / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / total,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{len(cases)} independent realisations, lmax from {{32, 64, 96, 128, 192}} "
                         "and meshes of {64, 96, 128} elements x degree 5 (seed 48), delta "
                         "seeds 3000..; the realisations of each (lmax, mesh) side by side in "
                         f"one operator ({len(groups)} operators), stepped on worker threads' "
                         f"streams ({workers} workers)"),
            "realisations": [{"lmax": c["lmax"], "elements": c["elements"],
                              "passes": c["passes"]} for c in cases],
            "side_by_side_operators": [{"lmax": k[0], "elements": k[1], "realisations": len(v)}
                                       for k, v in groups.items()],
            "equal_to_each_alone": same,
            "one_at_a_time_steps_per_s": total / sequential,
            "speedup_vs_one_at_a_time": sequential / elapsed,
            "realisation_errors": errors,
        },
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def uneven_cpu_baseline(cases, args, single, cap=2):
    """The uneven batch's CPU baseline and per-realisation errors: for the
    first realisation drawn at each lmax, one step from rest capped at `cap`
    passes on the GPU (its operator, lateral field, load and tolerance) and
    the oracle's same capped step from the same inputs, in a child process
    on the host's cores (as cpu_baseline). The oracle's per-pass and
    StartStep + Finish times, scaled by each realisation's measured passes,
    give the CPU time of the whole batch; the relative gaps between the two
    capped steps (strain coefficients, memory and strain grids) are each
    sampled realisation's error against the FP64 CPU restatement. single(c)
    makes realisation c's own operator."""
    import subprocess
    import tempfile

    import numpy as np

    import sl3dpp_amd.step as st

    sample, seen = [], set()
    for i, c in enumerate(cases):
        if c["lmax"] not in seen:
            seen.add(c["lmax"])
            sample.append(i)
    cores, nproc, allowed = physical_cores()
    threads = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    with tempfile.TemporaryDirectory() as tmp:
        files = []
        for i in sample:
            c = cases[i]
            op = single(c)
            op.set_options(max_iterations=cap)
            state = op.zero_state()
            op.step(state, c["dt"], st.EXPONENTIAL_TRAPEZOID, c["load"])
            path = os.path.join(tmp, f"r{i}.npz")
            np.savez(path, delta=op.delta_grid(), theta=op.theta, wtheta=op.wtheta,
                     lmax=c["lmax"], elements=c["elements"], dt=c["dt"], rtol=args.rtol,
                     cap=cap, coef=state.export("coefficients"),
                     **{f"m{b}": state.export(f"m{b}") for b in range(3)},
                     **{f"d{b}": state.export(f"strain{b}") for b in range(3)})
            files.append(path)
            del state, op
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores",
                   OMP_PROC_BIND="close")
        log(f"uneven cpu baseline: the oracle on {len(files)} sampled realisations, "
            f"{threads} threads ...")
        # the child's progress goes straight to this process's stderr
        run = subprocess.run([sys.executable, os.path.abspath(__file__), "--uneven-cpu-child"]
                             + files, env=env, stdout=subprocess.PIPE, text=True, timeout=1800)
        if run.returncode != 0:
            raise RuntimeError("uneven cpu child failed")
        results = json.loads(run.stdout.strip().splitlines()[-1])
    per_lmax = {r["lmax"]: r for r in results}
    cpu_seconds = 0.0
    for c in cases:
        r = per_lmax[c["lmax"]]
        # the sampled realisation's per-pass cost scaled by mesh size
        scale = c["nodes"] / r["nodes"]
        for passes in c["passes"]:
            cpu_seconds += scale * (r["once_s"] + passes * r["per_pass_s"])
    total = sum(len(c["passes"]) for c in cases)
    errors = [{"lmax": r["lmax"], "elements": r["elements"], "passes": r["passes"],
               "coefficients": r["coefficients"], "memory": r["memory"], "strain": r["strain"]}
              for r in results]
    cpu = {
        "value": total / cpu_seconds,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "nproc": nproc,
        "sample": (f"the oracle (FP64 C/OpenMP, direct longitude DFTs) on one realisation per "
                   f"lmax ({len(results)}), one step from rest capped at {cap} passes, "
                   f"{threads} threads (OMP_PLACES=cores, OMP_PROC_BIND=close); per-pass and "
                   f"StartStep+Finish times scaled to every realisation's passes and memory "
                   f"nodes"),
    }
    return cpu, errors


def uneven_cpu_child(paths):
    """uneven_cpu_baseline's oracle steps, in this (child) process."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from step_problem import OracleStep, amplitudes, oracle_library

    import sl3dpp_amd.step as st

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    meshes = {64: (35, 27, 2), 96: (52, 42, 2), 128: (70, 56, 2)}
    oracle = oracle_library()
    out = []
    for path in paths:
        w = np.load(path)
        lmax, elements = int(w["lmax"]), int(w["elements"])
        dt, rtol, cap = float(w["dt"]), float(w["rtol"]), int(w["cap"])
        earth = st.Earth(st.three_region_earth(*meshes[elements], 6))
        load, _ = st.zonal_load_rate(lmax, earth)
        log(f"uneven cpu child: lmax {lmax}, {elements} elements ...")
        p = OracleStep(oracle, earth, lmax, st.EXPONENTIAL_TRAPEZOID, dt, w["delta"],
                       theta=w["theta"], wtheta=w["wtheta"])
        info = p.step(dt, st.EXPONENTIAL_TRAPEZOID, amplitudes(load, lmax, dt),
                      max_iterations=cap, rtol=rtol, anderson_memory=6, anderson_rate=0.5)
        r = {"lmax": lmax, "elements": elements, "nodes": p.nodes, "passes": info.iterations,
             "per_pass_s": info.passes_ms / 1e3 / max(info.iterations, 1),
             "once_s": (info.start_ms + info.finish_ms) / 1e3,
             "coefficients": rel(w["coef"], p.coef),
             "memory": max(rel(w[f"m{b}"], p.m[b]) for b in range(3)),
             "strain": max(rel(w[f"d{b}"], p.strain[b]) for b in range(3))}
        log(f"uneven cpu child: lmax {lmax}: {r['per_pass_s']:.2f} s/pass, errors "
            f"{r['coefficients']:.1e} {r['memory']:.1e} {r['strain']:.1e}")
        out.append(r)
    print(json.dumps(out), flush=True)
    return 0


def batched(args):
    """Independent realisations on one GPU side by side: one operator of R
    realisations (delta seeds 2026 + i) whose every pass kernel takes all
    their memory nodes as extra columns, each realisation converging on its
    own; value is the aggregate steps/s over the timed steps. Beside it, the
    same realisations as R operators stepped one after another (the same
    warm-up and timed steps from rest), whose states must equal the side-
    by-side realisations' bit for bit."""
    import numpy as np
    import torch

    import sl3dpp_amd.step as st

    lmax, R = args.lmax, args.realisations
    earth = st.Earth(st.three_region_earth(70, 56, 2, 6))
    length, tscale, mass = earth.scales()
    dt = 100.0 * st.YEAR / tscale
    load, _ = st.zonal_load_rate(lmax, earth)

    def options(realisations):
        return st.StepOptions.default(max_iterations=200, rtol=args.rtol, anderson_memory=6,
                                      anderson_rate=0.5, realisations=realisations)

    singles = []
    for i in range(R):
        op = st.Viscoelastic3D(earth, lmax, None, options(1))
        coef = st.delta_coefficients(min(64, lmax), op.nodes, seed=2026 + i, rms=0.7, ar=0.9,
                                     lmax=lmax)
        op.set_delta(op.delta_field(coef))
        singles.append(op)
    batch = st.Viscoelastic3D(earth, lmax, None, options(R))
    batch.set_delta(torch.cat([op.delta for op in singles]))
    log(f"{R} realisations at lmax {lmax}: grid {batch.ntheta} x {batch.nphi}, "
        f"{batch.radial_nodes} memory nodes each, side by side in one operator")

    def timed(op, state, record):
        for _ in range(args.warmup):
            op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            op.step(state, dt, st.EXPONENTIAL_TRAPEZOID, load)
            record(op)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    passes = [[] for _ in range(R)]
    state = batch.zero_state()
    elapsed = timed(batch, state, lambda op: [passes[r].append(i.iterations)
                                              for r, i in enumerate(op.realisation_infos())])
    alone, same = 0.0, True
    N = batch.radial_nodes
    side = state.export("coefficients_device").reshape(-1, R * N, 2)
    for r, op in enumerate(singles):
        own = []
        s1 = op.zero_state()
        alone += timed(op, s1, lambda o: own.append(o.realisation_infos()[0].iterations))
        same = same and own == passes[r] and np.array_equal(
            side[:, r * N:(r + 1) * N, :].ravel(), s1.export("coefficients_device"))
        del s1
    total = R * args.steps
    log(f"{total} steps in {elapsed:.3f} s side by side ({total / elapsed:.2f} steps/s), "
        f"{alone:.3f} s one operator at a time ({total / alone:.2f} steps/s); passes per step "
        f"{np.mean([p for q in passes for p in q]):.1f}; equal to each alone: {same}")
    assert same, "a side-by-side realisation differs from its operator alone"
    out = {
        "metric": METRIC + " (independent realisations side by side)",
        "value": total / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / total,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{R} independent 3D lateral-viscosity realisations (delta seeds "
                         f"2026..{2026 + R - 1}) at lmax {lmax}, side by side in one operator "
                         "(every pass kernel over all their memory nodes; each realisation "
                         "converging on its own); the rest as the single-realisation workload"),
            "lmax": lmax, "realisations": R, "passes_per_step": passes,
            "one_at_a_time_steps_per_s": total / alone,
            "speedup_vs_one_at_a_time": alone / elapsed,
            "equal_to_each_alone": same,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
