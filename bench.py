"""Benchmark of the ATJUP time step on one MI355X: cell-updates/s of the HIP path.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config full|dyn] [--no-fp32] [--no-cpu]

A step is one iteration of the time loop (planet/cJupiterModel.cpp:1010-1555) on the 41 x 181 x 361
shell: the physics block with the 200-sweep red-black pressure solve on the reference's cadence,
the RK4 step and the boundary passes. Iterations are numbered in 5-step windows 1..5 with the first
of each window treated as the first iteration of a run, i.e. three pressure solves per five steps,
exactly the five-step dynamics case. One cell-update = one interior cell (i 1..39, j 3..177,
k 0..359: 2,457,000 cells) advanced by one step.

The default workload (--config full) is the full-physics case: the k-omega closure, the saturation
adjustments, the grey radiation, the three-category precipitation and the chemistry sequence in the
physics block, from vapour at 0.5-1.0 x saturation. --config dyn is the dry dynamics case.

The JSON line carries, besides the throughput, the roofline of the dominant kernel (the red-black
colour kernel of the pressure solve), measured live with HIP events on the library's stream, and
the CPU baseline: the reference itself (compiled from its own sources into oracle/_ref, at -O2 and
at the -O0 of its own Makefile) timed on this host's cores over a sample of the same workload, in
child processes with the OpenMP thread count set and reported.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s FP32-refined vs FP64 at <=1e-5 rel err; adjoint/forward time"
INTERIOR_CELLS = 39 * 175 * 360          # RK4 pass A range (RungeKutta_Jup_Turb.cpp:388-412)


def interior_cells(g):
    """Cells one step integrates: i 1..im-2, j 3..jm-4, k 0..km-2 (RungeKutta_Jup_Turb.cpp:388-412)."""
    return (g.im - 2) * (g.jm - 6) * (g.km - 1)
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the colour pass from two rocprofv3 PMC runs (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, KiB per dispatch, the factors from kernels of known size in the same runs: a
# whole-field read and a plain fill), per grid; profiles/step_full_41x181x361_hbm_traffic.md holds
# the summary and the command
SWEEP_TRAFFIC_BYTES = {(41, 181, 361): (12522.8 * 2.0 + 10192.5) * 1024}
WORKLOAD = {
    "dyn": "dynamics-only FP64 time step",
    "full": "full-physics FP64 time step (k-omega closure, saturation adjustment of H2O/NH3/CH4, NH4SH "
            "equilibrium, grey radiation, three-category precipitation, chemistry sequence with the NH4SH "
            "flux limiter, 17 prognostic fields)",
}
DATA = {
    "dyn": "synthetic (seeded 12-jet zonal wind, parabolic T with 110 K clamp, ellipsoid obstacle)",
    "full": "synthetic (seeded 12-jet zonal wind, parabolic T with 110 K clamp, ellipsoid obstacle, H2O/NH3/CH4 "
            "vapour at seeded 0.5-1.0 x saturation, no condensate, k=1e-4 m2/s2, omega=1e-3 1/s)",
}
# the physics the profile of each workload runs in its physics block
PHYSICS = {"dyn": {}, "full": {"radiation": True, "precip": True, "chemistry": True, "nh4sh_eq": True}}
MOISTURE = {"dyn": "seeded", "full": "subsaturated"}


def poisson_launch_bytes(g, esz_p=8, omega=1.0):
    """Algorithmic HBM bytes of one red-black colour launch with perfect neighbour reuse.

    The sweep domain is i 1..im-2, j 1..jm-2, k 0..km-2 (PressureSolver.h:323-358). One colour pass
    reads the other colour's values (half the field), its own source (half) and its face flags
    (1 byte, half the cells) and writes its own half of p; with over-relaxation it also reads its
    own old value.
    """
    n = (g.im - 2) * (g.jm - 2) * (g.km - 1)
    per_cell = 0.5 * esz_p + 0.5 * 8 + 0.5 * 1 + 0.5 * esz_p
    if omega != 1.0:
        per_cell += 0.5 * esz_p
    return n * per_cell


def step_bytes_per_cell_update(config, esz, sweeps, solves_per_step, physics=None):
    """Algorithmic HBM bytes per cell-update of the whole step.

    RK4: every prognostic field is streamed 17 times a step (stage 0: input, accumulator write,
    output; stages 1-2: input, n level, accumulator read and write, output; stage 3: input, n level,
    accumulator read, output), every read-only input once a stage, plus the aux / closure outputs
    and the 1-byte face flags. Pressure solve: per call the source build (4 fields) and 3 field
    passes per sweep (1.5 per colour: the other colour, the own colour's source and value). Physics
    block per call, in field streams: mixture density 4, the Shapiro pass on p_dyn 6 (3 axes, a read
    and a write each); full profile also the saturation adjustment 3 x 10, the closure 25, the
    radiation 14, the precipitation 33, and the chemistry sequence 76 (its two kernels 10 + 13, the
    limiter 5, eight Shapiro passes 48); without the chemistry the limiter 5 and its Shapiro pass 6.
    """
    physics = physics or {}
    if config == "dyn":
        n_prog, n_in, n_out, phys = 4, 5, 5, 4 + 6
    else:
        n_prog, n_in, n_out = 17, 22, 8
        phys = 4 + 6 + 30 + 25
        phys += 14 if physics.get("radiation") else 0
        phys += 33 if physics.get("precip") else 0
        phys += 76 if physics.get("chemistry") else 5 + 6
        phys += 9 if physics.get("nh4sh_eq") else 0     # t, NH3, H2S, NH4SH read and written, rho read
    rk = esz * (17 * n_prog + 4 * n_in + n_out) + 4 * 1
    pois = 8 * (4 + 3 * sweeps) * solves_per_step
    return rk + pois + esz * phys * solves_per_step


def cpu_threads():
    """The host threads the CPU legs use: ATJUP_CPU_THREADS, else the 16 CPUs a one-GPU box grants a
    job (os.cpu_count() shows the whole host there), else fewer if the host has fewer."""
    env = os.environ.get("ATJUP_CPU_THREADS")
    return int(env) if env else min(16, os.cpu_count() or 1)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg_child(config, sweeps, steps, opt, threads):
    """One CPU-baseline sample in this (child) process: the reference compiled at -<opt>, or the
    oracle restatement where the reference is not built, over iterations 1..steps of the workload
    (the same physics and inputs as the GPU run), with `threads` OpenMP threads. Prints one JSON line."""
    import oracle
    from atjup_amd import abi, inputs
    G, Cn, K, env = abi.make_profile(config, sweeps=sweeps, **PHYSICS[config])
    st = inputs.make_state(config, seed=11, moisture=MOISTURE[config])
    ph = PHYSICS[config]
    if oracle.reference_available(opt):
        ref = oracle.Reference(env, opt=opt)
        ref.set_threads(threads)
        used = ref.max_threads()
        ref.load_state(st, Cn)
        ref.call("bc_all")
        ref.call("restore")
        t0 = time.perf_counter()
        for it in range(1, steps + 1):
            ref.step_physics(it, it == 1, satadj=True, turb=(config != "dyn"), radiation=ph.get("radiation", False),
                             precip=ph.get("precip", False), chemistry=ph.get("chemistry", False),
                             nh4sh_eq=ph.get("nh4sh_eq", False))
        dt = time.perf_counter() - t0
        kind = "reference"
    else:
        orc = oracle.Oracle(G, Cn, K, threads=threads)
        used = threads
        orc.load_state(st)
        orc.call("bc_after_step")
        orc.call("swap_nlevel")
        t0 = time.perf_counter()
        for it in range(1, steps + 1):
            orc.call("step", it, 1 if it == 1 else 0)
        dt = time.perf_counter() - t0
        kind = "port"
    print(json.dumps({"seconds": dt, "steps": steps, "threads": used, "kind": kind, "opt": opt}), flush=True)


def cpu_baseline(config, sweeps, steps_o2=10, steps_o0=5):
    """The reference on this host's cores at -O2 (the value) and at -O0 (its own Makefile's level),
    each in a child process of its own (the reference holds one knob profile per process)."""
    import subprocess
    threads = cpu_threads()
    legs = {}
    for opt, steps in (("O2", steps_o2), ("O0", steps_o0)):
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", opt, "--cpu-steps", str(steps),
               "--cpu-threads", str(threads), "--config", config, "--sweeps", str(sweeps)]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
        if r.returncode != 0:
            legs[opt] = {"failed": (r.stderr or r.stdout)[-400:]}
            continue
        legs[opt] = json.loads(r.stdout.strip().splitlines()[-1])
    o2 = legs["O2"]
    if "seconds" not in o2:
        return {"value": None, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
                "sample": f"failed: {o2.get('failed')}"}
    solves = sorted(it for it in range(1, o2["steps"] + 1) if it % 2 == 0 or it == 1)
    out = {
        "value": INTERIOR_CELLS * o2["steps"] / o2["seconds"], "unit": "cell-updates/s",
        "cores": o2["threads"], "kind": o2["kind"],
        "sample": f"{o2['steps']} steps (iterations 1..{o2['steps']}, {sweeps}-sweep solves on iterations "
                  f"{','.join(map(str, solves))}) of the same 41x181x361 workload, {o2['seconds']:.1f} s, "
                  f"{o2['threads']} OpenMP threads (omp_set_num_threads), g++ -O2 -fopenmp",
        "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
    }
    o0 = legs.get("O0", {})
    if "seconds" in o0:
        out["O0"] = {"value": INTERIOR_CELLS * o0["steps"] / o0["seconds"], "unit": "cell-updates/s",
                     "cores": o0["threads"], "kind": o0["kind"],
                     "sample": f"{o0['steps']} steps of the same workload, {o0['seconds']:.1f} s, g++ -O0 -fopenmp "
                               f"(the reference Makefile's flags)"}
    else:
        out["O0"] = o0
    return out


def adjoint_leg(sweeps, omega, steps=20, every=5, directions=16, eps=1.0e-6, g=None, precision=None, ref=None):
    """The 20-step jet-misfit gradient on the dynamics profile: the time of one plain forward
    (steps steps), of the checkpointed forward and of the reverse sweep, all between library
    barriers with the adjoint seed already on the device; then dL/dw0 against central differences
    (eps) of the FP64 forward along seeded directions. precision: the form of the forward and of
    its adjoint (FP64 by default; the FP32-refined form's adjoint keeps its fields in FP32). ref:
    the FP64 leg's result, whose gradient and difference quotients (the FP64 reference gradient)
    the FP32-refined leg is compared with instead of forming its own."""
    from atjup_amd import abi, inputs
    from atjup_amd.adjoint import MisfitRunner
    from atjup_amd.lib import Context

    precision = abi.PRECISION_FP64 if precision is None else precision
    g = g or inputs.Grid()
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=sweeps, omega=omega, grid=g)
    st = inputs.make_state("dyn", seed=11, g=g)
    ctx = Context(G, Cn, K, precision)
    ctx.load_state(st, skip=set(inputs.SPECIES))
    ctx.bc_after_step()
    ctx.swap_nlevel()
    ctx.mixture_density()
    init = {n: ctx.get(n) for n in ("t", "u", "v", "w", "p_dyn", "aux_u", "aux_v", "aux_w", "rho_mix")}
    target = inputs.zonal_jets(g, np.random.default_rng(12)) / 100.0   # a second seeded 12-jet profile
    run = MisfitRunner(ctx, init, target, nsteps=steps, every=every)
    w0 = init["w"]

    def clock(fn):
        ctx.sync()
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3

    run.load(w0)
    clock(lambda: [ctx.step(1 + n, n == 0) for n in range(steps)])        # warm-up
    # warm-up of the gradient as well (its checkpoint buffers are allocated on the first call and
    # kept for calls of the same shape), as the forward's warm-up above
    run.forward(w0)
    run.backward()
    run.load(w0)
    fwd_ms = clock(lambda: [ctx.step(1 + n, n == 0) for n in range(steps)])
    run.load(w0)
    ckpt_ms = clock(lambda: ctx.adj_forward(steps, every, 1))
    loss, lam = run.loss_final()
    ctx.adj_zero()
    for f, name in enumerate(("t", "u", "v", "w")):
        ctx.adj_set(name, lam[f])
    ctx.set_timing(True)
    ctx.read_timers()                                                    # clear
    bwd_ms = clock(lambda: ctx.adj_backward(None, download=False))
    tm = ctx.read_timers()
    ctx.set_timing(False)
    grad = ctx.adj_get("w")
    errs, fds, ads = [], [], []
    for d_seed in range(directions if ref is None else len(ref["_fds"])):
        d = np.random.default_rng(1000 + d_seed).standard_normal(g.shape)
        d[:, :, -1] = d[:, :, 0]
        fd = ((run.evaluate(w0 + eps * d) - run.evaluate(w0 - eps * d)) / (2 * eps)) if ref is None else ref["_fds"][d_seed]
        fds.append(fd)
        ad = float(np.sum(grad * d))
        errs.append(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
        ads.append(ad)
    rms = float(np.sqrt(np.mean(np.square(fds)))) if fds else 0.0
    rms_errs = [abs(fd - ad) / max(abs(fd), rms, 1e-300) for fd, ad in zip(fds, ads)]
    ctx.close()
    form = "FP64" if precision == abi.PRECISION_FP64 else "FP32-refined"
    out = {
        "workload": f"jet misfit at i = {g.im - 2} after {steps} {form} dynamics steps, {g.im}x{g.jm}x{g.km}, dL/dw0 "
                    f"({g.im * g.jm * g.km} values), checkpoints every {every} steps"
                    + ("" if ref is None else ", the adjoint in FP32 (the solve's FP32 correction sweeps "
                                              "transposed inside the FP64 refinement)"),
        "forward_ms": fwd_ms, "checkpointed_forward_ms": ckpt_ms, "backward_ms": bwd_ms,
        "backward_over_forward": bwd_ms / fwd_ms,
        "backward_phases_ms": {"recompute": tm["adj_recompute"][0], "adjoint_steps": tm["adj_step"][0],
                               "adjoint_sweeps": tm["adj_sweeps"][0], "forward_solves": tm["pressure"][0]},
        "gradient_over_forward": (ckpt_ms + bwd_ms) / fwd_ms,
        "loss": loss,
        "fd_directions": directions, "fd_eps": eps,
        "fd_max_rel_err": max(errs) if errs else None,
        "fd_median_rel_err": float(np.median(errs)) if errs else None,
        # a random direction can be nearly orthogonal to the gradient, its quotient then a small
        # fraction of the others and the relative error large on both sides alike: the criterion
        # holds |fd - ad| against max(|fd|, rms of the directions' quotients)
        "fd_max_err_vs_max_fd_rms": max(rms_errs) if rms_errs else None,
        "within_1e-4": bool(max(rms_errs) <= 1e-4) if rms_errs else None,
        "_grad": grad, "_fds": fds,
    }
    if ref is not None:
        # the FP32 gradient against the FP64 one (the FP64 leg's start is not rounded to FP32,
        # so this includes the rounding of the initial state)
        out["fd_of"] = "the FP64 forward (the FP64 leg's difference quotients)"
        out["vs_fp64_gradient_normwise"] = float(np.linalg.norm(grad - ref["_grad"]) / np.linalg.norm(ref["_grad"]))
        out["gradient_over_fp64_forward"] = (ckpt_ms + bwd_ms) / ref["forward_ms"]
        out["gradient_speedup_vs_fp64"] = (ref["checkpointed_forward_ms"] + ref["backward_ms"]) / (ckpt_ms + bwd_ms)
    return out


def _public(leg):
    """A leg's result without its private arrays."""
    return {k: v for k, v in leg.items() if not k.startswith("_")} if leg else leg


def semi_axes_leg(sweeps, omega, steps=20, every=5, rel_eps=1.0e-4, g=None, i_top=None):
    """The misfit's gradient with respect to the obstacle's semi-axes (a_phi_deg, b_the_deg,
    c_r_km = 20, 10, 30): the binary obstacle replaced by the smooth drag region of the same
    ellipsoid (all-fluid mask), dL/daxes from one checkpointed forward and reverse sweep, against
    central differences of that forward (eps = rel_eps x the semi-axis). i_top: the misfit's level
    (default im - 2, the benchmark's top level)."""
    from atjup_amd import abi, inputs
    from atjup_amd.adjoint import MisfitRunner
    from atjup_amd.lib import Context

    g = g or inputs.Grid()
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=sweeps, omega=omega, grid=g)
    st = inputs.make_state("dyn", seed=11, g=g, with_obstacle=False)
    ctx = Context(G, Cn, K, abi.PRECISION_FP64)
    ctx.load_state(st, skip=set(inputs.SPECIES))
    ctx.bc_after_step()
    ctx.swap_nlevel()
    ctx.mixture_density()
    init = {n: ctx.get(n) for n in ("t", "u", "v", "w", "p_dyn", "aux_u", "aux_v", "aux_w", "rho_mix")}
    target = inputs.zonal_jets(g, np.random.default_rng(12)) / 100.0
    axes0 = np.array([20.0, 10.0, 30.0])
    if i_top is not None:
        target = init["w"][i_top, :, :-1].mean(axis=1) * 0.9
    run = MisfitRunner(ctx, init, target, nsteps=steps, every=every, obstacle=abi.obstacle(g, *axes0), i_top=i_top)
    w0 = init["w"]
    ctx.sync()
    t0 = time.perf_counter()
    loss = run.forward(w0)
    run.backward()
    ctx.sync()
    grad_ms = (time.perf_counter() - t0) * 1e3
    ga = run.axes_grad.copy()
    fds, rels = [], []
    for p in range(3):
        eps = rel_eps * axes0[p]
        vals = []
        for sgn in (+1, -1):
            ax = axes0.copy()
            ax[p] += sgn * eps
            run.set_axes(ax)
            vals.append(run.evaluate(w0))
        fd = (vals[0] - vals[1]) / (2 * eps)
        fds.append(fd)
        rels.append(abs(fd - ga[p]) / max(abs(fd), abs(ga[p]), 1e-300))
    ctx.close()
    responds = any(fd != 0.0 for fd in fds)
    return {
        "workload": f"jet misfit at i = {run.i_top} after {steps} FP64 dynamics steps, {g.im}x{g.jm}x{g.km}, "
                    f"dL/d(a_phi_deg, b_the_deg, c_r_km) of the smooth drag region (20 deg x 10 deg x 30 km, width "
                    f"0.25, tau 0.05) in place of the binary obstacle, checkpoints every {every} steps",
        "loss": loss, "gradient_ms": grad_ms, "dL_daxes": ga.tolist(), "fd": fds, "fd_rel_eps": rel_eps,
        "fd_rel_err": rels if responds else None,
        # when every difference quotient is exactly 0 the misfit does not respond to the axes at
        # FP64 resolution within these steps; the adjoint's values are then rounding residue
        "misfit_responds": responds,
        "within_1e-4": bool(max(rels) <= 1e-4) if responds else None,
    }


# Algorithmic bytes per cell of one radiation call: t, p_stat, rho_mix, CH4, NH3 and the six
# condensates read once, epsilon, radiation and Q_rad written once (planet/Radiation.h:311-463).
RAD_BYTES_PER_CELL = (11 + 3) * 8


def radiation_leg(ctx, g, calls=20):
    """Radiation<P>::run on the GPU, the mean over `calls` calls from the library's HIP events."""
    ctx.radiation()
    ctx.sync()
    ctx.set_timing(True)
    ctx.read_timers()
    for _ in range(calls):
        ctx.radiation()
    ctx.sync()
    total_ms, launches = ctx.read_timers()["radiation"]
    ms = total_ms / calls
    nbytes = RAD_BYTES_PER_CELL * g.im * g.jm * g.km
    return {"workload": f"grey two-stream radiation of all {g.jm * g.km} columns, {g.im} levels",
            "ms_per_call": ms, "launches_per_call": launches / calls,
            "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / (ms * 1e-3) / 1e9,
            "hbm_frac": nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9)}


# Algorithmic bytes per cell of one precipitation call: t, p_stat, rho_mix, the nine H2O/NH3/CH4
# vapour/cloud/ice fields and NH4SH read once; the nine rain/snow/graupel fluxes, the nine rates,
# Q_precip and the NH4SH flux written once (planet/Precipitation.h:194-574).
PRECIP_BYTES_PER_CELL = (13 + 20) * 8


def precipitation_leg(ctx, g, calls=20):
    """Precipitation<P>::run on the GPU, the mean over `calls` calls from the library's HIP events."""
    ctx.precipitation()
    ctx.sync()
    ctx.set_timing(True)
    ctx.read_timers()
    for _ in range(calls):
        ctx.precipitation()
    ctx.sync()
    total_ms, launches = ctx.read_timers()["precipitation"]
    ms = total_ms / calls
    nbytes = PRECIP_BYTES_PER_CELL * g.im * g.jm * g.km
    return {"workload": f"three-category precipitation of H2O, NH3, CH4 and NH4SH settling, all {g.jm * g.km} "
                        f"columns, {g.im} levels",
            "ms_per_call": ms, "launches_per_call": launches / calls,
            "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / (ms * 1e-3) / 1e9,
            "hbm_frac": nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9)}


def chemistry_leg(ctx, g, calls=20):
    """The physics block's chemistry sequence (two kernels, eight smoothing passes and the flux
    limiter): wall time per call with the device synchronised, and the two kernels' HIP-event time."""
    ctx.chemistry()
    ctx.sync()
    ctx.set_timing(True)
    ctx.read_timers()
    t0 = time.perf_counter()
    for _ in range(calls):
        ctx.chemistry()
    ctx.sync()
    wall = (time.perf_counter() - t0) / calls
    k_ms, launches = ctx.read_timers()["chemistry"]
    return {"workload": f"DiffMassFluxJup, ChemMassRateJup, the NH4SH limiter and eight Shapiro passes, "
                        f"{g.im}x{g.jm}x{g.km}",
            "ms_per_call": wall * 1e3, "kernels_ms_per_call": k_ms / calls, "kernel_launches_per_call": launches / calls}


def mca_leg(ctx, g, calls=5):
    """The moist convective adjustment (one thread per column): HIP-event time of the first call on
    the stepped state, where nearly every column of the seeded state is unstable and is mixed, then
    per call on the adjusted state, where the parcel census finds few unstable pairs."""
    ctx.sync()
    ctx.set_timing(True)
    ctx.read_timers()
    ctx.mca()
    first_ms, _ = ctx.read_timers()["mca"]
    for _ in range(calls):
        ctx.mca()
    again_ms, launches = ctx.read_timers()["mca"]
    return {"workload": f"moist convective adjustment (H2O, NH3; H2S, NH4SH tracers) of all {g.jm * g.km} "
                        f"columns, {g.im} levels",
            "first_call_ms": first_ms, "adjusted_state_ms_per_call": again_ms / calls,
            "launches_per_call": launches / calls}


def reporting_leg(ctx, g, calls=20):
    """searchMinMax_3D of t and steadyQuery's continuity residual on the device: wall time per call
    including the few-byte result download and the stream synchronisation."""
    ctx.search_minmax_3d("t"); ctx.continuity_residual()
    t0 = time.perf_counter()
    for _ in range(calls):
        ctx.search_minmax_3d("t")
    mm = (time.perf_counter() - t0) / calls
    t0 = time.perf_counter()
    for _ in range(calls):
        ctx.continuity_residual()
    cr = (time.perf_counter() - t0) / calls
    field = g.im * g.jm * g.km * 8
    return {"workload": f"searchMinMax_3D of t and the continuity residual of u, v, w, {g.im}x{g.jm}x{g.km}",
            "minmax_ms_per_call": mm * 1e3, "minmax_GBs": field / mm / 1e9,
            "residual_ms_per_call": cr * 1e3, "residual_GBs": 3 * field / cr / 1e9}


def restart_leg(ctx, g):
    """save_state then load_state of the device-resident state (JUP2, 55 arrays and vel_star), in a
    temporary directory on the host's disk: the D2H/H2D transfers and the file I/O together."""
    import tempfile
    nbytes = 20 + (38 + 17) * g.im * g.jm * g.km * 8 + g.jm * g.km * 8
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "jup_restart_0.bin")
        ctx.sync()
        t0 = time.perf_counter()
        ctx.save_restart(path, 0)
        t1 = time.perf_counter()
        ctx.load_restart(path)
        ctx.sync()
        t2 = time.perf_counter()
        size = os.path.getsize(path)
    return {"workload": f"JUP2 restart file of the {g.im}x{g.jm}x{g.km} state ({size / 1e9:.2f} GB)",
            "bytes": size, "expected_bytes": nbytes, "save_s": t1 - t0, "load_s": t2 - t1,
            "save_GBs": size / (t1 - t0) / 1e9, "load_GBs": size / (t2 - t1) / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="full", choices=["dyn", "full"])
    ap.add_argument("--precision-compare", dest="precision_compare", action="store_true", default=True,
                    help="also time the FP32-refined form and report its per-field error against FP64 (default)")
    ap.add_argument("--no-fp32", dest="precision_compare", action="store_false")
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--omega", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward leg")
    ap.add_argument("--no-restart", action="store_true", help="skip the restart-file leg")
    ap.add_argument("--grid", default="standard", choices=["standard", "fine"],
                    help="standard: 41x181x361 (1 deg); fine: 81x361x721 (0.5 deg, no CPU baseline)")
    ap.add_argument("--headline-only", dest="headline_only", action="store_true",
                    help="time the headline FP64 step only (no FP32 form, adjoint, CPU baseline or component legs)")
    ap.add_argument("--fp32-only", dest="fp32_only", action="store_true",
                    help="time the FP32-refined step alone (for kernel traces); prints its own short JSON line")
    ap.add_argument("--adjoint-only", dest="adjoint_only", action="store_true",
                    help="time the adjoint leg alone, without difference quotients (for kernel traces)")
    ap.add_argument("--cpu-leg", dest="cpu_leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-steps", dest="cpu_steps", type=int, default=10, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", dest="cpu_threads", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_leg:
        # a child of cpu_baseline(): the CPU sample only, no GPU
        cpu_leg_child(args.config, args.sweeps, args.cpu_steps, args.cpu_leg, args.cpu_threads or cpu_threads())
        return
    if args.headline_only:
        args.precision_compare = False
        args.no_cpu = args.no_adjoint = args.no_restart = True
    if args.adjoint_only:
        from atjup_amd import abi, inputs
        g = inputs.Grid() if args.grid == "standard" else inputs.fine_grid()
        kw = {"steps": 50, "every": 10} if args.grid == "fine" else {}
        a64 = adjoint_leg(args.sweeps, args.omega, g=g, directions=0, **kw)
        out = {"adjoint": _public(a64),
               "adjoint_fp32": _public(adjoint_leg(args.sweeps, args.omega, g=g, directions=0,
                                                   precision=abi.PRECISION_FP32_REF, ref=a64, **kw))}
        if args.grid == "standard":
            out["semi_axes"] = semi_axes_leg(args.sweeps, args.omega, g=g)
            out["semi_axes_level3"] = semi_axes_leg(args.sweeps, args.omega, g=g, i_top=3)
        print(json.dumps(out), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU layout (phi slabs over xGMI) "
              "is a design, not an implementation. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # torch is deliberately not imported: its wheel bundles its own HIP and HSA runtimes, which
    # would sit beside the ROCm runtime the library links, with separate streams and events. The
    # barrier is the library's own device synchronisation; the kernel timings are HIP events the
    # library records on its own stream.
    from atjup_amd import abi, inputs
    from atjup_amd.lib import Context

    g = inputs.Grid() if args.grid == "standard" else inputs.fine_grid()
    physics = PHYSICS[args.config]
    G, Cn, K, _ = abi.make_profile(args.config, sweeps=args.sweeps, omega=args.omega, grid=g, **physics)
    st = inputs.make_state(args.config, seed=11, g=g, moisture=MOISTURE[args.config])
    cells = interior_cells(g)
    fine = args.grid == "fine"

    def timed_run(precision, rounded_start=False, knobs=None):
        """warmup + steps iterations from the seeded state; the steps timed between two barriers."""
        ctx = Context(G, Cn, knobs or K, precision)
        s0 = st
        if rounded_start:
            s0 = inputs.State(grid=st.grid, mask=st.mask,
                              fields={n: v.astype(np.float32).astype(np.float64) for n, v in st.fields.items()})
        ctx.load_state(s0, skip=set(inputs.SPECIES) if args.config == "dyn" else set())
        ctx.bc_after_step()
        ctx.swap_nlevel()
        n = 0

        def run(k):
            nonlocal n
            for _ in range(k):
                it = n % 5 + 1
                ctx.step(it, it == 1)
                n += 1

        run(args.warmup)
        ctx.sync()
        ctx.set_timing(True)
        ctx.read_timers()                               # clear
        ctx.sync()
        t0 = time.perf_counter()
        run(args.steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        return ctx, wall * 1e3, ctx.read_timers(), ctx.check_finite()

    if args.fp32_only:
        c32, ms32, tm32, ok32 = timed_run(abi.PRECISION_FP32_REF)
        print(json.dumps({"workload": WORKLOAD[args.config] + f", FP32-refined, {g.im}x{g.jm}x{g.km}",
                          "ms_per_step": ms32 / args.steps, "finite": bool(ok32),
                          "phases_ms_per_step": {k: v[0] / args.steps for k, v in tm32.items()},
                          "poisson_sweep_avg_launch_ms": tm32["poisson_sweeps"][0] / max(tm32["poisson_sweeps"][1], 1)}),
              flush=True)
        c32.close()
        return
    ctx, ms, tm, ok = timed_run(abi.PRECISION_FP64)
    sweep_ms, sweep_launches = tm["poisson_sweeps"]
    rk_ms, rk_launches = tm["rk_stages"]

    fp32 = None
    if args.precision_compare:
        c32, ms32, tm32, ok32 = timed_run(abi.PRECISION_FP32_REF)
        # the problem's own sensitivity at this precision: FP64 from the start rounded to FP32
        cr, _, _, _ = timed_run(abi.PRECISION_FP64, rounded_start=True)
        fields = abi.PROG if args.config == "full" else ("t", "u", "v", "w")

        # a field whose FP64 values all lie below FP32's smallest normal number (a condensate that
        # has evaporated to ~1e-60) has no FP32 representation to compare: it is reported with its
        # absolute difference and left out of the worst case
        fp32_tiny = float(np.finfo(np.float32).tiny)

        def rel(c_other):
            out = {}
            for name in fields + ("p_dyn",):
                a, b = c_other.get(name), ctx.get(name)
                d = a - b
                nb, mb = float(np.linalg.norm(b)), float(np.max(np.abs(b)))
                if mb < fp32_tiny:
                    out[name] = {"below_fp32_range": True, "fp64_max_abs": mb,
                                 "max_abs_diff": float(np.max(np.abs(d)))}
                    continue
                out[name] = {"normwise": float(np.linalg.norm(d)) / nb, "max_abs": float(np.max(np.abs(d))) / mb}
            return out
        errs = rel(c32)
        floor = rel(cr)
        cr.close()
        compared = {n: e for n, e in errs.items() if "normwise" in e}
        worst = max(max(e["normwise"], e["max_abs"]) for e in compared.values())
        dyn_names = [n for n in ("t", "u", "v", "w", "p_dyn") if n in compared]
        worst_dyn = max(max(compared[n]["normwise"], compared[n]["max_abs"]) for n in dyn_names)
        worst_norm = max(e["normwise"] for e in compared.values())
        fp32 = {
            "value": cells * args.steps / (ms32 * 1e-3), "unit": "cell-updates/s",
            "ms_per_step": ms32 / args.steps, "speedup_vs_fp64": ms / ms32, "finite": bool(ok32),
            "poisson_sweep_avg_launch_ms": tm32["poisson_sweeps"][0] / max(tm32["poisson_sweeps"][1], 1),
            "rk4_stage_avg_launch_ms": tm32["rk_stages"][0] / max(tm32["rk_stages"][1], 1),
            "rel_err_vs_fp64_after_steps": args.warmup + args.steps, "worst_rel_err": worst,
            "within_1e-5": bool(worst <= 1e-5),
            "worst_rel_err_t_u_v_w_p_dyn": worst_dyn, "within_1e-5_t_u_v_w_p_dyn": bool(worst_dyn <= 1e-5),
            "worst_normwise_rel_err": worst_norm, "rel_err": errs,
            "fp64_from_fp32_rounded_start_rel_err": floor,
        }
        c32.close()

    # the grey radiation and the precipitation (components beside the path: they run in the
    # physics block when the profile turns them on), timed on the state the FP64 run left
    legs = args.config == "full" and not args.headline_only
    rad = radiation_leg(ctx, g) if legs else None
    precip = precipitation_leg(ctx, g) if legs else None
    chem = chemistry_leg(ctx, g) if legs else None
    mca = mca_leg(ctx, g) if legs else None
    report = reporting_leg(ctx, g) if not args.headline_only else None
    rst = None
    if not args.no_restart:
        try:
            rst = restart_leg(ctx, g)
        except Exception as e:  # disk space on the host never decides the GPU number
            rst = {"workload": "JUP2 restart file", "failed": str(e)}

    # the same step with the radiation, precipitation and chemistry left out of the physics block
    # (massflux_*, Q_rad, Q_precip, S_precip_* frozen): the dynamics, closure and adjustments alone
    closure = None
    if legs:
        _, _, Kc, _ = abi.make_profile("full", sweeps=args.sweeps, omega=args.omega, grid=g)
        cc, msc, _, okc = timed_run(abi.PRECISION_FP64, knobs=Kc)
        cc.close()
        closure = {"workload": "the same FP64 step with radiation, precipitation, chemistry and the NH4SH "
                               "equilibrium off (their outputs frozen inputs)", "value": cells * args.steps / (msc * 1e-3),
                   "unit": "cell-updates/s", "ms_per_step": msc / args.steps, "finite": bool(okc),
                   "step_algorithmic_bytes_per_cell_update":
                       step_bytes_per_cell_update("full", 8, args.sweeps, 3.0 / 5.0, {})}

    value = cells * args.steps / (ms * 1e-3)
    launch_ms = sweep_ms / max(sweep_launches, 1)
    lb = poisson_launch_bytes(g, 8, args.omega)
    achieved = lb / (launch_ms * 1e-3) / 1e9
    solves_per_step = 3.0 / 5.0
    step_b = step_bytes_per_cell_update(args.config, 8, args.sweeps, solves_per_step, physics)

    adj = None
    if not args.no_adjoint:
        # the fine shell: 50 steps with checkpoints every 10, fewer difference quotients
        kw = {"steps": 50, "every": 10, "directions": 4} if fine else {}
        adj = adjoint_leg(args.sweeps, args.omega, g=g, **kw)
        # the FP32-refined forward and its FP32 adjoint on the same misfit, against the FP64 gradient
        adj["fp32"] = _public(adjoint_leg(args.sweeps, args.omega, g=g, precision=abi.PRECISION_FP32_REF, ref=adj,
                                          **kw))
        adj = _public(adj)
        if not fine:
            # the benchmark's top-level misfit, and one at level 3, inside the drag region's height,
            # where the loss responds to the semi-axes well above the rounding of its difference quotients
            adj["semi_axes"] = semi_axes_leg(args.sweeps, args.omega, g=g)
            adj["semi_axes_level3"] = semi_axes_leg(args.sweeps, args.omega, g=g, i_top=3)

    cpu = None
    if not args.no_cpu and not fine:
        try:
            cpu = cpu_baseline(args.config, args.sweeps)
        except Exception as e:  # the baseline never decides the GPU number
            cpu = {"value": None, "unit": "cell-updates/s", "cores": 0, "kind": "port", "sample": f"failed: {e}"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": DATA[args.config],
        "config": {
            "workload": WORKLOAD[args.config] + f", {g.im}x{g.jm}x{g.km} shell, RK4 + red-black Poisson "
                        f"({args.sweeps} sweeps, omega={args.omega}) on iterations 1,2,4 of each 5-step window",
            "grid": [g.im, g.jm, g.km], "sweeps": args.sweeps, "omega": args.omega,
            "precision": "fp64", "finite": bool(ok),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": SWEEP_TRAFFIC_BYTES.get((g.im, g.jm, g.km)),
            "kernel": "one red-black colour pass of the FP64 solve (k_rb_sweep2f<double, double, 2>, two cells "
                      "per thread, on the 41x181x361 shell; k_rb_sweep2 on 81x361x721)",
            "algorithmic_bytes_per_launch": lb,
            "avg_launch_ms": launch_ms,
            "launches": int(sweep_launches),
        },
        "phases_ms_per_step": {
            "poisson_sweeps": sweep_ms / args.steps,
            "rk4_stage_kernels": rk_ms / args.steps,
            "rk4_stage_avg_launch_ms": rk_ms / max(rk_launches, 1),
            "pressure_solve": tm["pressure"][0] / args.steps,
            "radiation": tm["radiation"][0] / args.steps,
            "precipitation": tm["precipitation"][0] / args.steps,
            "chemistry_kernels": tm["chemistry"][0] / args.steps,
        },
        "step_algorithmic_bytes_per_cell_update": step_b,
        "step_hbm_frac": value * step_b / (HBM_PEAK_GBS * 1e9),
        "gpu_event_ms_per_step": tm["step"][0] / args.steps,
        "cpu_baseline": cpu,
        "dynamics_closure_only": closure,
        "fp32_refined": fp32,
        "adjoint": adj,
        "radiation": rad,
        "precipitation": precip,
        "chemistry": chem,
        "mca": mca,
        "reporting": report,
        "restart": rst,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
