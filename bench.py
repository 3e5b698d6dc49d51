#!/usr/bin/env python3
"""Headline benchmark: the fused MUSCL-Hancock step on one MI355X.

Workload (one step = one godunov_unsplit iteration of the whole grid): 3D Euler blast,
256^3 cells on the unit cube, periodic boundaries, gamma 1.4, HLLC, MC slopes
(slope_type 2), CFL 0.5; rho = 1, v = 0, p = 0.1 everywhere and p = 10 inside a sphere
of radius 0.1 centred at (0.5, 0.5, 0.5). The state is resident in HBM before timing;
the timed region is `steps` iterations of the device-resident time loop (dt computed on
the device every step). `value` is interior cell-updates per second in FP64 (the
bit-exact path). The same steps are then run in FP32 bulk mode and its density/energy
L1 error against the FP64 result is reported, and the oracle (a CPU restatement of the
reference's OpenMP step) is timed on the host cores on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps 100] [--warmup 5]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay on the cores they start on (read by the OpenMP
# runtime when the oracle library loads); the thread count is the environment's
os.environ.setdefault("OMP_PROC_BIND", "close")

METRIC = ("cell-updates/s at 1 GPU, % HBM peak; batched runs/s + p50/p99 at FP32 vs FP64")
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hydro3d_step_fp64_256cube_traffic.txt")


def device_sync() -> None:
    """hipDeviceSynchronize (the device-wide barrier torch.cuda.synchronize() issues)."""
    hip = C.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def read_traffic():
    """HBM bytes per step launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE (see file)."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def gpu_run(ekp, U0, warmup, steps):
    from euler_kokkos_amd import native
    ctx = native.Context(ekp)
    try:
        ctx.upload(U0, native.EK_LAYOUT_SOA)
        b0 = ctx.budgets()
        t, it, _ = ctx.run(warmup, warmup + steps, 1e30)
        ctx.run(0, warmup + steps, 1e30, t, it)  # capture the step graph outside the timed region
        device_sync()
        t0 = time.perf_counter()
        t, it, dt = ctx.run(steps, warmup + steps, 1e30, t, it)
        device_sync()
        wall = time.perf_counter() - t0
        ev_ms, launches = ctx.last_run_time()
        U = ctx.download(native.EK_LAYOUT_SOA)
        b = ctx.budgets()
    finally:
        ctx.close()
    return dict(wall=wall, ev_ms=ev_ms, launches=launches, U=U, t=t, it=it, dt=dt, budgets=b, budgets0=b0)


def cpu_baseline(n, seconds):
    """The oracle (CPU restatement of the reference step, OpenMP) on this host's cores."""
    from oracle import pyoracle
    from euler_kokkos_amd.params import make_params
    ekp = make_params("Hydro_Muscl_3D", n, n, n, gamma0="1.4", cfl="0.5", slope_type="2",
                      riemann="hllc", bc=(3,) * 6)
    op = pyoracle.from_ek(ekp)
    U = pyoracle.blast(op, radius=0.1, center_x=0.5, center_y=0.5, center_z=0.5, density_in=1.0,
                       density_out=1.0, pressure_in=10.0, pressure_out=0.1)
    host = host_info()  # before the first parallel region binds this thread to one place
    t0 = time.perf_counter()
    pyoracle.run(op, U, 1, 1e30)
    one = time.perf_counter() - t0
    steps = int(max(1, min(20, seconds // max(one, 1e-3))))
    t0 = time.perf_counter()
    pyoracle.run(op, U, steps, 1e30)
    el = time.perf_counter() - t0
    cores = pyoracle.lib().eko_num_threads()
    cups = steps * n ** 3 / el
    return {"value": cups, "unit": "cell-updates/s", "cores": int(cores), "kind": "port",
            "sample": f"oracle (C restatement of the reference impl-0 OpenMP step, gcc -O3 "
                      f"-ffp-contract=off) on the same {n}^3 blast, {steps} steps after 1 untimed step, "
                      f"{cores} OpenMP threads",
            "ref_style_mcell_per_s": steps * (n + 4) ** 3 / el * 1e-6, **host}


_HOST = None
_HOST_CPUS = None


def unpin():
    """Gives this thread back the CPUs it started with. The OpenMP runtime binds the thread
    that runs its first parallel region (a CPU baseline) to one place (OMP_PROC_BIND=close),
    and threads started afterwards -- the serving loops -- would inherit that single CPU."""
    if _HOST_CPUS:
        os.sched_setaffinity(0, _HOST_CPUS)


def host_info():
    """CPU model and OpenMP placement of the host the CPU baseline runs on (taken once, at the
    first call, before any OpenMP region binds the calling thread to one place)."""
    global _HOST, _HOST_CPUS
    if _HOST is not None:
        return dict(_HOST)
    _HOST_CPUS = set(os.sched_getaffinity(0))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    _HOST = {"cpu_model": model, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
             "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
             "host_cpus_available": len(os.sched_getaffinity(0)), "host_cpus_online": os.cpu_count()}
    return dict(_HOST)


def mhd3d_c5(n, steps, with_cpu, cpu_seconds):
    """3D ideal-MHD blast (the serving stream's MHD job kind) on n^3, periodic, HLLD + CT,
    gamma 1.6666667, CFL 0.8, MC slopes, field B0 = 0.5 along (1,1,1): the five-kernel step
    pipeline in FP64 and FP32, the discrete max |div B| of the FP64 result, and the oracle
    on the host cores."""
    from euler_kokkos_amd import native, problems
    from euler_kokkos_amd.serving import job_params
    res = {"workload": f"3D MHD blast {n}^3 periodic, HLLD + CT, MC slopes, gamma 5/3, CFL 0.8, "
                       f"B0 0.5 along (1,1,1), {steps} steps", "grid": [n, n, n], "solver": "MHD_Muscl_3D"}
    out = {}
    U0 = None
    for name, prec in (("fp64", native.EK_FP64), ("fp32", native.EK_FP32)):
        ekp = job_params("mhd_blast", n, prec)
        if U0 is None:
            U0 = problems.mhd_blast3d(ekp, B0=0.5)
        ctx = native.Context(ekp)
        try:
            ctx.upload(U0, native.EK_LAYOUT_SOA)
            b0 = ctx.budgets()
            t, it, _ = ctx.run(2, steps + 2, 1e30)
            ctx.run(0, steps + 2, 1e30, t, it)  # capture the step graph outside the timed region
            device_sync()
            t0 = time.perf_counter()
            ctx.run(steps, steps + 2, 1e30, t, it)
            device_sync()
            wall = time.perf_counter() - t0
            ev_ms, launches = ctx.last_run_time()
            b1 = ctx.budgets()
            U = ctx.download(native.EK_LAYOUT_SOA)
        finally:
            ctx.close()
        A, B, C = U[5], U[6], U[7]
        divb = np.abs((np.roll(A, -1, axis=2) - A) + (np.roll(B, -1, axis=1) - B) + (np.roll(C, -1, axis=0) - C)).max()
        out[name] = U
        res[name] = {"value": steps * n ** 3 / wall, "unit": "cell-updates/s", "ms_per_step": wall / steps * 1e3,
                     "ms_per_step_events": ev_ms / max(launches, 1),
                     "achieved_GBps": n ** 3 * 2 * 8 * (8 if prec == native.EK_FP64 else 4)
                     / (ev_ms / max(launches, 1) * 1e-3) / 1e9,
                     "max_abs_divB_times_dx": float(divb),
                     "mass_drift_rel": float(abs(b1[0] - b0[0]) / abs(b0[0])),
                     "energy_drift_rel": float(abs(b1[1] - b0[1]) / abs(b0[1]))}
    res["fp32"]["l1_rho_rel"] = float(np.abs(out["fp32"][0] - out["fp64"][0]).mean() / np.abs(out["fp64"][0]).mean())
    res["fp32"]["l1_E_rel"] = float(np.abs(out["fp32"][1] - out["fp64"][1]).mean() / np.abs(out["fp64"][1]).mean())
    if with_cpu:
        from oracle import pyoracle
        ekp = job_params("mhd_blast", n, native.EK_FP64)
        op = pyoracle.from_ek(ekp)
        R = np.zeros(pyoracle.ref_shape(op))
        g = op.gw
        R[g:-g, g:-g, g:-g, :] = U0.transpose(3, 2, 1, 0)
        t0 = time.perf_counter()
        pyoracle.mhd3d_run(op, R, 1, 1e30)
        one = time.perf_counter() - t0
        k = int(max(1, min(20, cpu_seconds // max(one, 1e-3))))
        t0 = time.perf_counter()
        pyoracle.mhd3d_run(op, R, k + 1, 1e30, 0.0, 1)
        el = time.perf_counter() - t0
        cores = pyoracle.lib().eko_num_threads()
        res["cpu_baseline"] = {"value": k * n ** 3 / el, "unit": "cell-updates/s", "cores": int(cores),
                               "kind": "port", "sample": f"oracle 3D MHD step on the same {n}^3 blast, "
                               f"{k} steps after 1 untimed step, {cores} OpenMP threads", **host_info()}
    return res


def mhd_c3(steps, with_cpu, cpu_seconds):
    """2D ideal-MHD Orszag-Tang vortex, 1024^2, periodic, HLLD + constrained transport,
    gamma 1.6666667 (float-parsed), CFL 0.8, MC slopes: FP64 and FP32 bulk on the GPU, the
    discrete max |div B| of the FP64 result, and the oracle on the host cores."""
    from euler_kokkos_amd import native, problems
    from euler_kokkos_amd.params import make_params
    n = 1024
    kw = dict(gamma0="1.6666667", cfl="0.8", slope_type="2", riemann="hlld", smallr="1e-7",
              smallc="1e-7", bc=(3,) * 6)
    res = {"workload": f"2D MHD Orszag-Tang {n}^2 periodic, HLLD + CT, MC slopes, gamma 5/3, "
                       f"CFL 0.8, {steps} steps", "grid": [n, n], "solver": "MHD_Muscl_2D"}
    out = {}
    for name, prec in (("fp64", native.EK_FP64), ("fp32", native.EK_FP32)):
        ekp = make_params("MHD_Muscl_2D", n, n, precision=prec, **kw)
        U0 = problems.orszag_tang(ekp)
        ctx = native.Context(ekp)
        try:
            ctx.upload(U0, native.EK_LAYOUT_SOA)
            b0 = ctx.budgets()
            ctx.run(0, steps, 1e30)  # capture the step graph outside the timed region
            device_sync()
            t0 = time.perf_counter()
            t, it, dt = ctx.run(steps, steps, 1e30)
            device_sync()
            wall = time.perf_counter() - t0
            ev_ms, launches = ctx.last_run_time()
            b1 = ctx.budgets()
            U = ctx.download(native.EK_LAYOUT_SOA)
        finally:
            ctx.close()
        A, B = U[5], U[6]
        divb = np.abs((np.roll(A, -1, axis=1) - A) / ekp.dx + (np.roll(B, -1, axis=0) - B) / ekp.dy).max()
        out[name] = U
        res[name] = {"value": steps * n * n / wall, "unit": "cell-updates/s",
                     "ms_per_step": wall / steps * 1e3, "avg_launch_ms": ev_ms / max(launches, 1),
                     "achieved_GBps": n * n * 2 * 8 * (8 if prec == native.EK_FP64 else 4)
                     / (ev_ms / max(launches, 1) * 1e-3) / 1e9,
                     "max_abs_divB": float(divb), "max_abs_divB_times_dx": float(divb * ekp.dx),
                     "divB_definition":
                     "max over cells |(Bx(i+1,j)-Bx(i,j))/dx + (By(i,j+1)-By(i,j))/dy| on face fields",
                     "mass_drift_rel": float(abs(b1[0] - b0[0]) / abs(b0[0])),
                     "energy_drift_rel": float(abs(b1[1] - b0[1]) / abs(b0[1]))}
    d64, d32 = out["fp64"], out["fp32"]
    res["fp32"]["l1_rho_rel"] = float(np.abs(d32[0] - d64[0]).mean() / np.abs(d64[0]).mean())
    res["fp32"]["l1_E_rel"] = float(np.abs(d32[1] - d64[1]).mean() / np.abs(d64[1]).mean())
    if with_cpu:
        from oracle import pyoracle
        ekp = make_params("MHD_Muscl_2D", n, n, **kw)
        op = pyoracle.from_ek(ekp)
        R = pyoracle.orszag_tang(op)
        t0 = time.perf_counter()
        pyoracle.mhd_run(op, R, 1, 1e30)
        one = time.perf_counter() - t0
        k = int(max(1, min(50, cpu_seconds // max(one, 1e-3))))
        t0 = time.perf_counter()
        pyoracle.mhd_run(op, R, k + 1, 1e30, 0.0, 1)
        el = time.perf_counter() - t0
        cores = pyoracle.lib().eko_num_threads()
        res["cpu_baseline"] = {"value": k * n * n / el, "unit": "cell-updates/s", "cores": int(cores),
                               "kind": "port", "sample": f"oracle MHD step on the same {n}^2 Orszag-Tang, "
                               f"{k} steps after 1 untimed step, {cores} OpenMP threads", **host_info()}
    return res


def batch_c4(n_runs, steps, seed):
    """Batched 2D hydro: n_runs independent random blasts (edges 64..512) advanced together
    by one persistent, graph-replayed pool launch per step, FP32 state with FP64 dt and
    budgets, tiles chosen by the pool's autotuner; beside it the same batch in FP64 and
    the plain FP64 path one job at a time (each run its own context and ek_run)."""
    from euler_kokkos_amd import native, pool
    runs = {prec: pool.random_blast_batch(n_runs, seed, prec) for prec in (native.EK_FP32, native.EK_FP64)}
    cells = sum(p.nx * p.ny for p, _, _ in runs[native.EK_FP64])
    res = {"workload": f"{n_runs} independent 2D hydro blasts, edges uniform in {{64,96,...,512}}, "
                       f"seed {seed}, HLLC + MC, gamma 1.4, CFL 0.5, reflecting walls, {steps} steps each",
           "runs": n_runs, "interior_cells_per_step": cells, "steps": steps, "seed": seed}
    finals = {}
    for name, prec in (("fp32_batched", native.EK_FP32), ("fp64_batched", native.EK_FP64)):
        rl = runs[prec]
        P = pool.Pool(rl[0][0], sum(2 * p.nx * p.ny * 4 * (8 if prec == native.EK_FP64 else 4) + 4096
                                    for p, _, _ in rl) + (64 << 20), n_runs)
        try:
            ids = [P.admit(p, U, steps) for p, U, _ in rl]
            choice, tune_ms = P.autotune(3)
            device_sync()
            t0 = time.perf_counter()
            P.step(steps)
            device_sync()
            wall = time.perf_counter() - t0
            ms, launches = P.last_step_time()
            assert all(P.status(j)[1] == steps for j in ids)
            finals[name] = [P.download(j)[:2] for j in ids]  # density and total energy
        finally:
            P.close()
        res[name] = {"runs_per_s": n_runs / wall, "cell_updates_per_s": cells * steps / wall,
                     "wall_s": wall, "ms_per_pool_step": ms / max(launches, 1),
                     # every run has the same step count, so all finish at the last launch:
                     # time-to-solution is the batch's wall time for every run
                     "tts_p50_s": wall, "tts_p99_s": wall,
                     "tts_note": "all runs advance together and finish at the same launch",
                     "tile_choice": choice, "autotune_ms_per_step": tune_ms}
    # plain FP64, one job at a time (queue order = generation order)
    done, per = [], []
    tq = 0.0
    for p, U, _ in runs[native.EK_FP64]:
        ctx = native.Context(p)
        try:
            ctx.upload(U, native.EK_LAYOUT_SOA)
            ctx.run(0, steps, 1e30)  # graph capture outside the timed run, as for the pool
            device_sync()
            t0 = time.perf_counter()
            ctx.run(steps, steps, 1e30)
            device_sync()
            el = time.perf_counter() - t0
        finally:
            ctx.close()
        tq += el
        per.append(el)
        done.append(tq)
    res["fp64_one_at_a_time"] = {"runs_per_s": n_runs / tq, "cell_updates_per_s": cells * steps / tq,
                                 "wall_s": tq, "tts_p50_s": float(np.percentile(done, 50)),
                                 "tts_p99_s": float(np.percentile(done, 99)),
                                 "run_time_p50_s": float(np.percentile(per, 50)),
                                 "run_time_p99_s": float(np.percentile(per, 99))}
    for v, tag in ((0, "rho"), (1, "E")):
        l1 = [float(np.abs(a[v] - b[v]).mean() / np.abs(b[v]).mean())
              for a, b in zip(finals["fp32_batched"], finals["fp64_batched"])]
        res[f"fp32_l1_{tag}_rel_max"] = max(l1)
        res[f"fp32_l1_{tag}_rel_mean"] = float(np.mean(l1))
    res["fp32_batched_vs_fp64_one_at_a_time"] = (res["fp32_batched"]["cell_updates_per_s"]
                                                 / res["fp64_one_at_a_time"]["cell_updates_per_s"])
    return res


def progress(msg):
    """One line per finished stage on stderr (stdout carries only the JSON result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def serving_c5(n_jobs, rate, seed, steps):
    """Serving stream: seeded Poisson arrivals of 3D jobs (edges uniform in [64,192], `steps`
    steps each), admitted into device-resident FP32 job pools by HBM footprint and advanced
    together (continuous batching); beside it the plain FP64 path serving the same stream
    one job at a time."""
    from euler_kokkos_amd import native, serving
    res = {"workload": f"{n_jobs} 3D jobs, Poisson arrivals (mean {rate:g} jobs/s, seed {seed}), edges "
                       f"uniform in [64,192], {steps} steps each; half 3D MHD blasts (rho 1, p 0.1 / 10 in r < 0.1, "
                       f"uniform B = B0 (1,1,1)/sqrt(3), B0 uniform in [0.1,1], gamma 5/3, HLLD + CT, CFL 0.8), "
                       f"half hydro Kelvin-Helmholtz shear layers (HLLC + MC, gamma 1.4, CFL 0.5); periodic; "
                       f"FP32 state, FP64 dt",
           "kinds": ["mhd_blast", "hydro_kh"]}
    jobs = serving.poisson_stream(n_jobs, rate, seed, steps=steps, precision=native.EK_FP32)
    res["host_cpus_serving"] = len(os.sched_getaffinity(0))
    progress(f"serving: {n_jobs} job states built")
    for name, burst in (("stream", False), ("burst", True)):
        # stream: the Poisson arrival times (latency at the offered load); burst: every job
        # arrives at t = 0 (throughput at saturation)
        arr = [j.arrival for j in jobs]
        if burst:
            for j in jobs:
                j.arrival = 0.0
        S = serving.Server(128 << 30, native.EK_FP32)
        try:
            fp32 = S.serve(jobs)
        finally:
            S.close()
        progress(f"serving {name}: FP32 continuous batching done ({fp32.get('jobs_per_s', 0):.1f} jobs/s)")
        for j in jobs:
            j.done_t = -1.0
        fp64 = serving.serve_one_at_a_time(jobs, native.EK_FP64)
        progress(f"serving {name}: FP64 one at a time done ({fp64.get('jobs_per_s', 0):.1f} jobs/s)")
        for j, a in zip(jobs, arr):
            j.arrival = a
            j.done_t = -1.0
        res[name] = {"fp32_continuous_batching": fp32, "fp64_one_at_a_time": fp64,
                     "fp32_batched_vs_fp64_one_at_a_time": fp32["cell_updates_per_s"] / fp64["cell_updates_per_s"]}
    # accuracy of the served FP32 results against FP64, outside the timed runs: a served job's
    # result is bitwise its result alone (tests/test_serving_gpu.py), so each job is rerun alone
    # in FP32 and in FP64 and the relative L1 of density and total energy compared
    l1 = {"rho": [], "E": []}
    for j in jobs:
        fin = {}
        for prec in (native.EK_FP32, native.EK_FP64):
            ctx = native.Context(serving.job_params(j.kind, j.n, prec))
            try:
                ctx.upload(j.U0, native.EK_LAYOUT_SOA)
                ctx.run(j.steps, j.steps, 1e30)
                fin[prec] = ctx.download(native.EK_LAYOUT_SOA)
            finally:
                ctx.close()
        for v, tag in ((0, "rho"), (1, "E")):
            a, b = fin[native.EK_FP32][v], fin[native.EK_FP64][v]
            l1[tag].append(float(np.abs(a - b).mean() / np.abs(b).mean()))
    for tag, vals in l1.items():
        res[f"fp32_l1_{tag}_rel_max"] = max(vals)
        res[f"fp32_l1_{tag}_rel_mean"] = float(np.mean(vals))
    return res


def peak_source(dev):
    """Where the HBM peak comes from: MI355X's published 8 TB/s, beside what the runtime
    reports (memory clock x bus width; HBM3E moves 4 bits per pin per reported clock)."""
    derived = dev["mem_clock_khz"] * 1e3 * 4 * dev["bus_width_bits"] / 8 / 1e9
    return (f"published MI355X HBM3E peak {HBM_PEAK_GBS:.0f} GB/s; the runtime reports a "
            f"{dev['mem_clock_khz'] / 1e3:.0f} MHz memory clock on a {dev['bus_width_bits']}-bit bus, "
            f"{derived:.0f} GB/s at 4 transfers per reported clock")


def copy_calibration(gib=1.0):
    """A device-to-device copy of known size, for scaling rocprofv3 FETCH/WRITE counters."""
    hip = C.CDLL("libamdhip64.so")
    n = int(gib * (1 << 30))
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n}))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="grid edge (default: the 256^3 workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 bulk run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--no-mhd", action="store_true", help="skip the 2D MHD (Orszag-Tang) line")
    ap.add_argument("--mhd-steps", type=int, default=500)
    ap.add_argument("--no-mhd3d", action="store_true", help="skip the 3D MHD (blast) line")
    ap.add_argument("--mhd3d-n", type=int, default=128)
    ap.add_argument("--mhd3d-steps", type=int, default=30)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 2D hydro line")
    ap.add_argument("--batch-runs", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--batch-seed", type=int, default=2024)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-stream line")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-seed", type=int, default=77)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): the library has no multi-GPU "
              "path (a grid would shard into one subdomain per GPU with RCCL halo exchanges, "
              "and batched jobs by job).", file=sys.stderr)
        return 2
    if a.copy_calibration:
        copy_calibration()
        return 0

    host_info()
    from euler_kokkos_amd import native, problems
    from euler_kokkos_amd.params import make_params
    dev = native.device_info()
    n = a.n
    kw = dict(gamma0="1.4", cfl="0.5", slope_type="2", riemann="hllc", bc=(3,) * 6)
    ekp64 = make_params("Hydro_Muscl_3D", n, n, n, **kw)
    U0 = problems.blast(ekp64, radius="0.1", density_in="1", density_out="1", pressure_in="10",
                        pressure_out="0.1")
    r64 = gpu_run(ekp64, U0, a.warmup, a.steps)
    progress(f"headline FP64 {n}^3: {r64['wall'] / a.steps * 1e3:.4f} ms/step")
    cells = n ** 3
    value = a.steps * cells / r64["wall"]
    avg_ms = r64["ev_ms"] / max(r64["launches"], 1)
    alg_bytes = cells * 2 * 5 * 8
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r64["wall"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D Euler blast {n}^3 periodic, MUSCL-Hancock + HLLC, MC slopes, "
                               f"gamma 1.4, CFL 0.5, FP64 (bit-exact path)",
                   "grid": [n, n, n], "solver": "Hydro_Muscl_3D", "riemann": "hllc",
                   "slope_type": 2, "cfl": 0.5, "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "kernel": "hydro3d_step<double,16,8,256,HLLC,MC>",
                     "avg_launch_ms": avg_ms, "alg_bytes_per_launch": alg_bytes,
                     "peak_source": peak_source(dev)},
        "device": dev,
        # totals over the periodic grid before the warm-up and after the timed steps (FP64
        # compensated sums); the blast starts at rest, so momentum drift is given per unit mass
        "conservation": {"steps": a.warmup + a.steps,
                         "mass_drift_rel": abs(r64["budgets"][0] - r64["budgets0"][0]) / abs(r64["budgets0"][0]),
                         "energy_drift_rel": abs(r64["budgets"][1] - r64["budgets0"][1]) / abs(r64["budgets0"][1]),
                         "momentum_drift_per_mass": max(abs(r64["budgets"][v] - r64["budgets0"][v]) for v in (2, 3, 4))
                         / abs(r64["budgets0"][0])},
    }
    if not a.no_fp32:
        ekp32 = make_params("Hydro_Muscl_3D", n, n, n, precision=native.EK_FP32, **kw)
        r32 = gpu_run(ekp32, U0, a.warmup, a.steps)
        avg32 = r32["ev_ms"] / max(r32["launches"], 1)
        d64, d32 = r64["U"], r32["U"]
        out["fp32"] = {
            "value": a.steps * cells / r32["wall"], "ms_per_step": r32["wall"] / a.steps * 1e3,
            "speedup_vs_fp64": r64["wall"] / r32["wall"],
            "achieved_GBps": cells * 2 * 5 * 4 / (avg32 * 1e-3) / 1e9,
            "l1_rho": float(np.abs(d32[0] - d64[0]).mean()),
            "l1_E_rel": float(np.abs(d32[1] - d64[1]).mean() / np.abs(d64[1]).mean()),
        }
        progress("headline FP32 done")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, a.cpu_seconds)
        unpin()
        progress("CPU baseline done")
    else:
        out["cpu_baseline"] = None
    if not a.no_mhd:
        out["mhd2d"] = mhd_c3(a.mhd_steps, not a.no_cpu, a.cpu_seconds / 2)
        unpin()
        progress("2D MHD done")
    if not a.no_mhd3d:
        out["mhd3d"] = mhd3d_c5(a.mhd3d_n, a.mhd3d_steps, not a.no_cpu, a.cpu_seconds / 2)
        unpin()
        progress("3D MHD done")
    if not a.no_batch:
        out["batch2d"] = batch_c4(a.batch_runs, a.batch_steps, a.batch_seed)
        progress("batched 2D done")
    if not a.no_serve:
        out["serving3d"] = serving_c5(a.serve_jobs, a.serve_rate, a.serve_seed, 50)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
