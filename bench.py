#!/usr/bin/env python3
"""Benchmark: RK4 (RK4-Crank-Nicolson) steps/s of the FP64 pseudospectral
vorticity solver on one MI355X.

Workload (the one the headline metric is quoted on): decaying turbulence on a
1024^2 grid, FP64, random-phase initial w_hat with E(k) ~ k^4 exp(-(k/8)^2)
normalised to E = 0.5, nu = 5e-5 (nu k^2 viscosity), Hou-Li filter in the
reference's integer-division form, RK4CN, dt = 5e-4, with the reference's
per-step max|w| blow-up check.  A "step" is one full RK4CN time step (four
nonlinear right-hand sides).  The state is resident in HBM before timing
starts.  Reported under "extra": the same step on a 4096^2 grid, the
forced Kolmogorov CN-AB4 step on a 4096^2 grid, the 64-member uneven-size
ensemble against sequential plain solves, and the discrete adjoint's cost
relative to a forward solve (512^2, 2000 steps).

The roofline object covers the dominant kernel, timed live with HIP events
recorded on the solver's own stream around every launch; its algorithmic
bytes per launch are stated in alg_bytes() below.  The CPU baseline is the
compiled C restatement of the reference's RK4CN step (oracle/ns2d_cpu.c,
FFTW3 interface of MKL, OpenMP on the host cores), timed on a bounded number
of steps of the same workload (median of 3).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "2d-navier-stokes_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
NU, DT, K0 = 5e-5, 5e-4, 8.0


def alg_bytes(nx: int, ny: int) -> dict:
    """Algorithmic HBM bytes per launch (FP64 complex = 16 B per mode, M = Nx*(Ny/2+1)).

    rpass: reads the four x-transformed fields G1..G4, writes N          -> 5*16*M
    kpass: reads N and w_hat, reads/writes the RK accumulator, writes the four
           G's (stage 1: acc write only; stage 4: acc read + w_hat write)
           -> (112 + 128 + 128 + 128)/4 * M on average over a step
    step:  4 rpass + 4 kpass = 816*M;  the byte model with separate inverse
           and forward column passes is 928*M.
    """
    M = nx * (ny // 2 + 1)
    return {"rpass": 80.0 * M, "kpass": 124.0 * M, "step": 816.0 * M, "step_model": 928.0 * M}


def kernel_signature(pass_name, n):
    """Template signature of the shape a new solver runs for this pass and
    length (kpass_kernel<L, C, T, KS, PS, OFF, PK, W, PF> /
    rpass_kernel<NY, TLR, PAIRS, W, PFR>), as rocprofv3 names the kernel."""
    import ns2d
    if pass_name == "kpass":
        v = ns2d.kpass_variants(n)[ns2d.default_variant(ns2d.PASS_COLUMN, n)]
        pk = 2 if v["pair_in_registers"] else v["packed_pair"]
        return (f"kpass_kernel<{n}, {v['cols']}, {v['threads']}, {v['splits']}, {v['pad_shift']}, "
                f"{v['line_offset']}, {pk}, {v['min_waves']}, {v['prefetch']}>")
    v = ns2d.rpass_variants(n)[ns2d.default_variant(ns2d.PASS_ROW, n)]
    pfr = 3 if (v["single_row"] and v["prefetch"]) else 2 if v["single_row"] else v["prefetch"]
    return f"rpass_kernel<{n}, {v['threads_per_pair']}, {v['pairs']}, {v['min_waves']}, {pfr}>"


def traffic_key(pass_name, n):
    """kernel_signature plus the device layout the library was compiled with
    (k-space x-tile K, 0 = column-major; x-tiles G and N of the mixed-space
    fields each pass writes for the other): "<sig> layout kKgGnN"."""
    import ns2d
    k, g, m = ns2d.layout()
    return f"{kernel_signature(pass_name, n)} layout k{k}g{g}n{m}"


def make_ic(nx, ny, seed=1):
    from ns2d import ics
    return ics.random_phase(nx, ny, seed=seed, k0=K0, e0=0.5, filt="houli_ref")


def make_solver(nx, ny):
    import ns2d
    cfg = ns2d.Config(Nx=nx, Ny=ny, nu=NU, hyper_flag=0, hyper_pow=1.0, alpha=0.0, drag_flag=0,
                      drag_pow=0.0, filter="houli_ref", integrator="RK4CN", check_every=1)
    return ns2d.Solver(cfg)


def time_gpu(nx, ny, steps, warmup):
    s = make_solver(nx, ny)
    s.set_state(make_ic(nx, ny))
    s.spectral_step(DT, warmup)
    s.sync()
    t0 = time.perf_counter()
    s.spectral_step(DT, steps)
    s.sync()
    t1 = time.perf_counter()
    # live per-kernel timing (HIP events on the solver's stream)
    kms, rms, tot = s.profile_steps(DT, max(5, min(steps, 50)))
    m = s.system_measures()
    s.close()
    return (t1 - t0), kms, rms, m


def cfl_dt(w_hat, nx, ny, nu, cfl=0.5):
    """GetTimestep's CFL step (solver.c:2381-2450) with every mode's velocity."""
    from ns2d import ics
    return ics.cfl_timestep(w_hat, nx, ny, nu, cfl)


def ensemble_leg(steps=500, seed=2024):
    """64 decaying-turbulence members of uneven size (24 x 256^2, 24 x 512^2,
    16 x 1024^2), k0 ~ U[4, 12] per member, 2/3 filter, RK4CN, dt per member
    from CFL 0.5 at t = 0; one ensemble launch sequence vs the same members
    solved one after another with the plain solver."""
    import ns2d
    from ns2d import ics
    rng = np.random.default_rng(seed)
    sizes = [256] * 24 + [512] * 24 + [1024] * 16
    nu = 5e-5
    cfgs, ws, dts = [], [], []
    for i, n in enumerate(sizes):
        k0 = float(rng.uniform(4.0, 12.0))
        w0 = ics.random_phase(n, n, seed=seed + i, k0=k0, e0=0.5, filt="23")
        cfgs.append(ns2d.Config(Nx=n, Ny=n, nu=nu, hyper_flag=0, hyper_pow=1.0, alpha=0.0, drag_flag=0,
                                drag_pow=0.0, filter="23", integrator="RK4CN", check_every=1))
        ws.append(w0)
        dts.append(cfl_dt(w0, n, n, nu))
    member_steps = steps * len(sizes)
    with ns2d.Ensemble(cfgs) as e:
        for m, w in zip(e.members, ws):
            m.set_state(w)
        e.step(dts, 2)
        e.sync()
        for m, w in zip(e.members, ws):
            m.set_state(w)
        t0 = time.perf_counter()
        e.step(dts, steps)
        e.sync()
        t_ens = time.perf_counter() - t0
        # the shapes the tuner chose per size class (first member of each)
        chosen = {}
        for i, n in enumerate(sizes):
            if n not in chosen:
                k, r = e.variants(i)
                kv, rv = ns2d.kpass_variants(n)[k], ns2d.rpass_variants(n)[r]
                chosen[n] = {"kpass": k, "rpass": r,
                             "kpass_shape": [kv[f] for f in ("cols", "threads", "splits", "pad_shift", "line_offset")],
                             "rpass_shape": [rv[f] for f in ("threads_per_pair", "pairs", "single_row")]}
    t_seq = 0.0
    for c, w, dt in zip(cfgs, ws, dts):
        with ns2d.Solver(c) as s:
            s.set_state(w)
            s.spectral_step(dt, 2)
            s.sync()
            s.set_state(w)
            t0 = time.perf_counter()
            s.spectral_step(dt, steps)
            s.sync()
            t_seq += time.perf_counter() - t0
    return {"ensemble_members": len(sizes), "ensemble_steps_each": steps,
            "ensemble_member_steps_per_s": member_steps / t_ens,
            "sequential_member_steps_per_s": member_steps / t_seq,
            "ensemble_vs_sequential": t_seq / t_ens,
            "ensemble_shapes": {str(n): v for n, v in chosen.items()}}


def adjoint_leg(n=512, steps=2000, ckpt=100, seed=5, fd_dirs=3):
    """dZ(T)/dw0 of 2000 RK4CN steps at 512^2 (nu = 1e-4, dt = 1e-3, 2/3,
    E(k) ~ k^4 exp(-(k/6)^2) projected to energy 0.5 and palinstrophy 5000),
    checkpoint every 100 steps with recompute; cost relative to one forward
    solve of the same length, and the gradient's agreement with central
    finite differences of the GPU forward map (Richardson-combined steps) along
    a few seeded directions."""
    import ns2d
    from ns2d import ics
    nu, dt = 1e-4, 1e-3
    cfg = ns2d.Config(Nx=n, Ny=n, nu=nu, hyper_flag=0, hyper_pow=1.0, alpha=0.0, drag_flag=0, drag_pow=0.0,
                      filter="23", integrator="RK4CN", quirks=0, check_every=0)
    w0 = ics.random_phase_ep(n, n, seed=seed, k0=6.0, e0=0.5, p0=5000.0)
    with ns2d.Solver(cfg) as s:
        s.set_state(w0)
        s.spectral_step(dt, 10)
        s.sync()
        s.set_state(w0)
        t0 = time.perf_counter()
        s.spectral_step(dt, steps)
        s.sync()
        t_fwd = time.perf_counter() - t0
        wT = s.get_state()
        c = np.full((1, n // 2 + 1), 2.0)
        c[0, 0] = c[0, -1] = 1.0
        lam = 2.0 * c * wT * (4 * np.pi ** 2) * (0.5 / float(n * n) ** 2)
        lam[0, 0] = 0.0
        s.vjp(w0, dt, 4 * ckpt, lam, ckpt_every=ckpt)   # warm-up (tapes, streams, graphs)
        t0 = time.perf_counter()
        g = s.vjp(w0, dt, steps, lam, ckpt_every=ckpt)
        t_adj = time.perf_counter() - t0

        def Z(w):
            s.set_state(w)
            s.spectral_step(dt, steps)
            return s.system_measures()["enstrophy"]

        rng = np.random.default_rng(17)
        amp = np.abs(w0)
        rels = []
        for _ in range(fd_dirs):
            d = amp * np.exp(2j * np.pi * rng.random(w0.shape))
            h = 1e-4 * np.linalg.norm(w0) / np.linalg.norm(d)
            D = lambda e: (Z(w0 + e * d) - Z(w0 - e * d)) / (2 * e)
            fd = (4.0 * D(h / 2) - D(h)) / 3.0
            an = float(np.sum(g.real * d.real + g.imag * d.imag))
            rels.append(abs(fd - an) / abs(an))
    return {"adjoint_grid": n, "adjoint_steps": steps, "adjoint_ckpt_every": ckpt,
            "forward_s": t_fwd, "vjp_s": t_adj, "adjoint_over_forward": t_adj / t_fwd,
            "grad_norm": float(np.linalg.norm(g)), "adjoint_ic": "E = 0.5, P = 5000 (k0 = 6 random phases)",
            "adjoint_fd_directions": fd_dirs, "adjoint_fd_max_rel_err": max(rels)}


def aux_profile():
    """The adjoint and measures kernels' workloads for a kernel trace: one
    vector-Jacobian product of 400 RK4CN steps at 512^2 (the adjoint leg's
    flow and initial condition, checkpoints every 100 steps) and 10 measure
    calls with the flux terms at 1024^2 (the bench flow's initial state)."""
    import ns2d
    from ns2d import ics
    n, dt, steps = 512, 1e-3, 400
    cfg = ns2d.Config(Nx=n, Ny=n, nu=1e-4, hyper_flag=0, hyper_pow=1.0, alpha=0.0, drag_flag=0, drag_pow=0.0,
                      filter="23", integrator="RK4CN", quirks=0, check_every=0)
    w0 = ics.random_phase_ep(n, n, seed=5, k0=6.0, e0=0.5, p0=5000.0)
    with ns2d.Solver(cfg) as s:
        s.set_state(w0)
        s.spectral_step(dt, steps)
        wT = s.get_state()
        c = np.full((1, n // 2 + 1), 2.0)
        c[0, 0] = c[0, -1] = 1.0
        lam = 2.0 * c * wT * (4 * np.pi ** 2) * (0.5 / float(n * n) ** 2)
        lam[0, 0] = 0.0
        t0 = time.perf_counter()
        g = s.vjp(w0, dt, steps, lam, ckpt_every=100)
        print(f"adjoint {n}^2 x {steps} steps: {time.perf_counter() - t0:.3f} s, |g| = "
              f"{float(np.linalg.norm(g)):.6e}", flush=True)
    N = 1024
    with make_solver(N, N) as s:
        s.set_state(make_ic(N, N))
        t0 = time.perf_counter()
        for _ in range(10):
            m = s.system_measures(with_flux=True)
        print(f"measures {N}^2 with flux x 10: {time.perf_counter() - t0:.3f} s, E = {m['energy']:.12e}",
              flush=True)


def kolmogorov_leg(n=4096, steps=40, warmup=5, seed=3):
    """Forced Kolmogorov flow at 4096^2: f = F sin(k_f y), k_f = 4, F = 1,
    linear drag alpha = 0.01, nu = 2e-5, 2/3 filter, CN-AB4 after 3 RK4CN
    start-up steps (inside the warm-up), dt = 2.5e-4, per-step max|w| check;
    low-amplitude (E = 1e-6) uniform-phase vorticity on 0 < |k| <= 16."""
    import ns2d
    from ns2d import ics
    cfg = ns2d.Config(Nx=n, Ny=n, nu=2e-5, hyper_flag=0, hyper_pow=1.0, alpha=0.01, drag_flag=1, drag_pow=0.0,
                      filter="23", integrator="AB4", forcing="KOLM", force_k=4, force_scale=1.0, check_every=1)
    w0 = ics.random_phase(n, n, seed=seed, e0=1e-6, filt="23", spectrum="flat", kmax=16.0)
    dt = 2.5e-4
    with ns2d.Solver(cfg) as s:
        s.set_state(w0)
        s.spectral_step(dt, warmup)
        s.sync()
        t0 = time.perf_counter()
        s.spectral_step(dt, steps)
        s.sync()
        el = time.perf_counter() - t0
        kms, rms, _ = s.profile_steps(dt, 10)
    sps = steps / el
    M = n * (n // 2 + 1)
    return {"kolm_ab4_4096_steps_per_s": sps, "kolm_ab4_4096_ms_per_step": 1e3 / sps,
            "kolm_ab4_4096_kpass_avg_ms": kms, "kolm_ab4_4096_rpass_avg_ms": rms,
            "kolm_ab4_4096_model272_frac": 272.0 * M * sps / 1e9 / HBM_PEAK_GBS}


def _host_threads():
    """Threads the host run may use: OMP_NUM_THREADS when set (the GPU box
    sets it to the box's CPU share), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nx, ny, budget_s=20.0):
    """The compiled CPU restatement (oracle/ns2d_cpu: the reference's RK4CN
    step in C, FFTW3 interface of MKL, OpenMP) on the host cores: steps/s of
    the same workload from the same initial condition, CLOCK_MONOTONIC around
    the stepping loop, median of 3 repetitions, with the per-step max|w|
    check (the reference's SystemCheck) and, separately, without it.  The step
    count per repetition is sized from a short calibration run so that the
    whole leg takes about budget_s."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "ns2d_cpu")
    threads = _host_threads()
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} not built: run __graft_entry__.build()")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), MKL_THREADING_LAYER="GNU")
    with tempfile.TemporaryDirectory() as tmp:
        icp = os.path.join(tmp, "ic.bin")
        make_ic(nx, ny).astype(np.complex128).tofile(icp)

        def run(steps, check, reps):
            r = subprocess.run([exe, str(nx), str(ny), repr(NU), "1.0", "houli_ref", repr(DT), str(steps),
                                str(check), str(reps), icp], capture_output=True, text=True, env=env, check=True)
            return json.loads(r.stdout.strip().splitlines()[-1])

        cal = run(2, 1, 1)
        per_step = 1.0 / cal["median_steps_per_s"]
        steps = max(2, int(budget_s / 6.0 / per_step))
        with_chk = run(steps, 1, 3)
        no_chk = run(steps, 0, 3)
    return {"value": with_chk["median_steps_per_s"], "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} RK4CN steps of the {nx}x{ny} workload per repetition, median of 3, incl. the "
                      f"per-step max|w| check (compiled C restatement, FFTW3 interface of MKL, "
                      f"{threads} OpenMP threads)",
            "reps_steps_per_s": with_chk["steps_per_s"],
            "without_check_steps_per_s": no_chk["median_steps_per_s"],
            "cpu_model": _cpu_model()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-4096", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with a 1 GiB device-to-device copy (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--aux-profile", action="store_true",
                    help="run only the adjoint (512^2, 400 steps, checkpoints every 100) and the "
                         "flux measures (1024^2, 10 calls) workloads, for a kernel trace; prints no "
                         "bench line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU (slab-decomposed) runs are not "
              "part of it -- use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.aux_profile:
        aux_profile()
        return
    nx = ny = args.n
    el, kms, rms, meas = time_gpu(nx, ny, args.steps, args.warmup)
    sps = args.steps / el
    ab = alg_bytes(nx, ny)
    if kms >= rms:
        dom, dom_ms, dom_bytes = "kpass", kms, ab["kpass"]
    else:
        dom, dom_ms, dom_bytes = "rpass", rms, ab["rpass"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    # HBM bytes per launch measured by rocprofv3 --pmc for this exact kernel
    # shape and device layout (profiles/traffic.json, keyed by the kernel's
    # template signature and the library's layout); null when the shape
    # running here has no such measurement
    traffic, traffic_src = None, None
    sig = traffic_key(dom, nx)
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f).get(sig)
        if rec:
            traffic, traffic_src = rec["bytes"], rec["source"]
    extra = {
        "kpass_avg_ms": kms, "rpass_avg_ms": rms,
        "step_alg_GBps": ab["step"] * sps / 1e9,
        "step_model928_GBps": ab["step_model"] * sps / 1e9,
        "energy_after": meas["energy"],
    }
    if not args.no_4096 and args.n != 4096:
        el4, k4, r4, _ = time_gpu(4096, 4096, 20, 3)
        s4 = 20 / el4
        ab4 = alg_bytes(4096, 4096)
        extra.update({"rk4_4096_steps_per_s": s4, "rk4_4096_ms_per_step": 1e3 / s4,
                      "rk4_4096_kpass_avg_ms": k4, "rk4_4096_rpass_avg_ms": r4,
                      "rk4_4096_step_alg_GBps": ab4["step"] * s4 / 1e9,
                      "rk4_4096_step_model928_frac": ab4["step_model"] * s4 / 1e9 / HBM_PEAK_GBS})
        extra.update(kolmogorov_leg())
        if not args.no_cpu:
            # the same RK4CN workload at 4096^2 on the host cores, a few steps
            c4 = cpu_baseline(4096, 4096, budget_s=30.0)
            extra.update({"cpu_4096": c4, "rk4_4096_vs_cpu": s4 / c4["value"]})
    if not args.no_ensemble:
        extra.update(ensemble_leg())
    if not args.no_adjoint:
        extra.update(adjoint_leg())
    out = {
        "metric": "RK4 steps/s at 1024^2 FP64 (RK4CN, decaying turbulence)",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / sps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-phase vorticity, E(k)~k^4 exp(-(k/8)^2), E=0.5)",
        "config": {"workload": f"decaying turbulence {nx}x{ny} FP64, RK4CN, Hou-Li (integer-division "
                               f"form), nu=5e-5, dt=5e-4, per-step max|w| check",
                   "grid": [nx, ny], "integrator": "RK4CN", "filter": "houli_ref", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": sig, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "alg_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(nx, ny, args.cpu_budget),
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    if args.calib_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print("calibration copy: 1 GiB device to device", flush=True)


if __name__ == "__main__":
    main()
