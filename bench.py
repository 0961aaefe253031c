#!/usr/bin/env python3
"""Benchmark: cell-updates/s per RK3 step of the 8-field MHD solver at 512^3, FP32 and FP64.

One step = 3 x {periodic halo fill, fused solve + RK3 substep, buffer swap} over the whole box,
driven through the C ABI (acDevicePeriodicBoundconds, acDeviceIntegrateSubstep,
acDeviceSwapBuffers), the caller sequence of the reference's single-GPU step
(src/core/node.cc:729-806). Inputs are generated on the GPU (seeded splitmix64, the uniform-field
distribution) and stay resident in HBM; the timed region is K steps bracketed by device
synchronisation. `value` is the FP32 rate; the FP64 build is timed on the same workload.

Roofline of the dominant kernel (the solve): algorithmic bytes per substep launch are
16 n^3 s (substep 0: read 8 fields, write 8) or 24 n^3 s (substeps 1-2: also read the previous
stage), 64 n^3 s per step; achieved = those bytes / the solve launches' mean duration, measured
with HIP events recorded on the library's stream around every solve launch in the timed loop.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "cell-updates/s per RK3 step at 512^3, FP32 vs FP64, + % of roofline"
HBM_PEAK_GBS = 8000.0
# Vector peak in use for both precisions: packed FP32 is compiled out, and unpacked FP32 runs at
# the FP64 rate (the vendor's 157.3 TFLOP/s FP32 figure counts packed math).
VALU_PEAK_TFLOPS = 78.6
# Floating-point operations per cell of one substep of the reference formulation that the eight
# rates depend on (adds + multiplies + divides), counted by oracle/opcount.cc ("live" lines;
# tests/test_oracle_opcount.py checks these constants against it): the right-hand side, its
# exponentials (not in the FLOP figure), and the RK3 update of substeps 0, 1 and 2.
FLOP_RHS = 552 + 421 + 9
EXP_RHS = 7
FLOP_UPDATE = (24, 48, 48)
# what one timed RK3 step runs (acDeviceIntegrateStep's form), by whether the build reads the
# periodic images of the ghost cells
STEP_FORM = {False: "3 x {periodic fill, solve+RK3 substep, swap}",
             True: "3 x {solve+RK3 substep reading the periodic images of the ghost cells, swap}"}


from astaroth_amd.hiprt import Hip  # noqa: E402  (events on the library's own HIP runtime)


def run_gpu(double: bool, n: int, steps: int, warmup: int, dt: float, autotune: bool, hip: Hip,
            kernel_config=None, halo="fill", reduce_each_step=False, perturb=False):
    from astaroth_amd.api import Library
    from astaroth_amd.initial import device_init_mhd

    lib = Library(double)
    info = lib.config(n=n)
    t0 = time.time()
    dev = lib.device_create(0, info, autotune=autotune)
    tune_s = time.time() - t0
    try:
        if kernel_config is not None:
            dev.set_kernel_config(*kernel_config)
        device_init_mhd(dev, config_id=3 if perturb else 1, perturb_sin2z=perturb)
        stream = dev.stream_handle()
        nmin, nmax = (3, 3, 3), (n + 3, n + 3, n + 3)
        for _ in range(warmup):
            dev.integrate_step(dt)
        dev.synchronize()
        evs = [[hip.event() for _ in range(6)] for _ in range(steps)]
        hev = [[hip.event() for _ in range(6)] for _ in range(steps)]
        ev_begin, ev_end = hip.event(), hip.event()
        d = lib.dll
        from astaroth_amd.abi import int3

        zero, m = int3(0, 0, 0), int3(n + 6, n + 6, n + 6)
        s_min, s_max = int3(*nmin), int3(*nmax)
        # the step of acDeviceIntegrateStep: where the build reads the periodic images of the
        # ghost cells (FP64 default), no fill; otherwise fill + substep
        images = halo == "fill" and d.acDeviceSubstepReadsPeriodicImages(dev.handle) == 1
        hip.sync()
        wall0 = time.perf_counter()
        hip.record(ev_begin, stream)
        reductions = []
        for k in range(steps):
            e = evs[k]
            for s in range(3):
                hip.record(hev[k][2 * s], stream)
                if images:
                    pass
                elif halo == "fill":
                    d.acDevicePeriodicBoundconds(dev.handle, 0, zero, m)
                else:  # 26-block pack + unpack through a device buffer
                    d.acDeviceHaloExchangeSelf(dev.handle, 0)
                hip.record(hev[k][2 * s + 1], stream)
                hip.record(e[2 * s], stream)
                if images:
                    d.acDeviceIntegrateSubstepPeriodic(dev.handle, 0, s, lib.real(dt))
                else:
                    d.acDeviceIntegrateSubstep(dev.handle, 0, s, s_min, s_max, lib.real(dt))
                hip.record(e[2 * s + 1], stream)
                d.acDeviceSwapBuffers(dev.handle)
            if reduce_each_step:  # FP64 diagnostics (synchronous, as the reference's reductions)
                reductions.append(step_diagnostics(dev))
        hip.record(ev_end, stream)
        hip.sync()
        wall = time.perf_counter() - wall0
        total_ms = hip.elapsed(ev_begin, ev_end)
        solve_ms = [hip.elapsed(e[2 * s], e[2 * s + 1]) for e in evs for s in range(3)]
        halo_ms = [hip.elapsed(e[2 * s], e[2 * s + 1]) for e in hev for s in range(3)]
        step_ms = []
        prev = ev_begin
        for e in evs:
            step_ms.append(hip.elapsed(prev, e[5]))
            prev = e[5]
        cfg = dev.kernel_config()
        # sanity: the state stays finite
        mx = dev.reduce_scal("RTYPE_MAX", "VTXBUF_LNRHO", double=True)
        if not (mx == mx) or abs(mx) > 1.0:
            raise RuntimeError("state diverged: max lnrho = %r" % mx)
    finally:
        dev.destroy()
    s = 8 if double else 4
    cells = float(n) ** 3
    ms_per_step = total_ms / steps
    solve_mean = sum(solve_ms) / len(solve_ms)
    alg_bytes_per_launch = 64.0 * cells * s / 3.0  # mean over substeps 0, 1, 2
    achieved = alg_bytes_per_launch / (solve_mean * 1e-3) / 1e9
    return dict(
        ms_per_step=ms_per_step,
        value=cells / (ms_per_step * 1e-3),
        median_step_ms=statistics.median(step_ms),
        p90_step_ms=sorted(step_ms)[min(len(step_ms) - 1, int(0.9 * len(step_ms)))],
        solve_ms_mean=solve_mean,
        solve_ms_by_substep=[sum(solve_ms[i::3]) / steps for i in range(3)],
        solve_fraction=sum(solve_ms) / total_ms,
        achieved_gbs=achieved,
        alg_bytes_per_launch=alg_bytes_per_launch,
        wall_s=wall,
        autotune_s=tune_s,
        kernel_config=dict(variant=cfg.variant, tile_x=cfg.tile_x, tile_y=cfg.tile_y, z_depth=cfg.z_depth),
        reads_images=images,
        reductions=reductions,
        halo_ms_mean=sum(halo_ms) / len(halo_ms),
        # ghost shell read + written once per field (fill), or read, packed, unpacked (pack)
        halo_alg_bytes=float((n + 6) ** 3 - n ** 3) * 8 * s * (2 if halo == "fill" else 4),
    )


def valu_roofline(r, n):
    """Vector-ALU side of the solve launch: algorithmic FLOP (mean over the three substeps) per
    launch / its mean duration, against the unpacked vector peak."""
    flop_cell = FLOP_RHS + sum(FLOP_UPDATE) / 3.0
    tflops = n ** 3 * flop_cell / (r["solve_ms_mean"] * 1e-3) / 1e12
    return {"achieved": tflops, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / VALU_PEAK_TFLOPS, "flop_per_cell_substep": flop_cell,
            "exp_per_cell_substep": EXP_RHS, "flop_source": "oracle/opcount.cc (live operations)"}


def ceilings(r, n, double):
    """Step-rate ceilings (cell-updates/s) set by HBM (64 n^3 s bytes per RK3 step) and by the
    vector ALUs (3 x right-hand side + the three updates per cell and step), and the measured
    rate's fraction of the lower one."""
    s = 8 if double else 4
    hbm = HBM_PEAK_GBS * 1e9 / (64 * s)
    valu = VALU_PEAK_TFLOPS * 1e12 / (3 * FLOP_RHS + sum(FLOP_UPDATE))
    low = min(hbm, valu)
    return {"hbm": hbm, "valu": valu, "binding": "hbm" if hbm <= valu else "valu",
            "frac_of_binding": r["value"] / low, "unit": "cell-updates/s"}


def step_diagnostics(dev):
    """FP64 max/min/rms of |u|, |A| and lnrho over the interior (RMS finalised as sqrt(sum/N),
    as the reference's node layer does, node.cc:932-957), in one pass (acDeviceDiagnostics;
    the same values as nine acDeviceReduce*Double calls, read once)."""
    return dev.diagnostics()


def precision_study(n: int, steps: int, dt: float, hip: Hip):
    """512^3 precision study: the same perturbed uniform-field state (u_x += 1e-2 sin 2z) advanced
    `steps` RK3 steps by the FP32 and the FP64 library; per-field error of FP32 against FP64 over
    the interior, range-normalised and at the point of maximum absolute error."""
    import numpy as np
    from astaroth_amd.api import Library, HostMesh
    from astaroth_amd.initial import device_init_mhd
    from astaroth_amd.abi import VTXBUF_NAMES

    states, diag, times = {}, {}, {}
    for prec in ("fp64", "fp32", "fp32_compensated"):
        lib = Library(prec == "fp64")
        info = lib.config(n=n)
        dev = lib.device_create(0, info, autotune=False)
        try:
            if prec == "fp32_compensated":
                dev.set_compensated_integration(True)
            device_init_mhd(dev, config_id=3, perturb_sin2z=True)
            diag[prec] = []
            hip.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                dev.integrate_step(dt)
                diag[prec].append(step_diagnostics(dev))
            hip.sync()
            times[prec] = (time.perf_counter() - t0) / steps
            mesh = HostMesh(lib, info)
            dev.store_mesh(mesh)
            dev.synchronize()
            states[prec] = [f[3:-3, 3:-3, 3:-3] for f in mesh.fields]
        finally:
            dev.destroy()
    out = {}
    for mode in ("fp32", "fp32_compensated"):
        fields = {}
        worst = 0.0
        for q, name in enumerate(VTXBUF_NAMES):
            ref = states["fp64"][q]
            d = np.abs(states[mode][q].astype(np.float64) - ref)
            i = int(np.argmax(d))
            rng = float(ref.max() - ref.min())
            rn = float(d.flat[i]) / rng if rng > 0 else float(d.flat[i])
            at = float(d.flat[i]) / abs(float(ref.flat[i])) if ref.flat[i] != 0 else float("inf")
            fields[name] = {"max_abs": float(d.flat[i]), "range": rng, "range_normalised": rn,
                            "rel_at_max_abs_point": at}
            worst = max(worst, rn)
            del d
        ddiag = {}
        for key in ("uu", "aa", "lnrho"):
            last64, last32 = diag["fp64"][-1][key], diag[mode][-1][key]
            ddiag[key] = [abs(a - b) / abs(a) if a != 0 else abs(b) for a, b in zip(last64, last32)]
        out[mode] = dict(fields=fields, worst_range_normalised=worst, passed=worst <= 1e-5,
                         diagnostics_rel_diff_last_step=ddiag)
    r = dict(out["fp32"])
    r.update(target=1e-5, s_per_step=times, compensated=out["fp32_compensated"])
    return r


def copy_probe(double: bool, n: int, reps: int, hip: Hip):
    """Known-size copy kernel (acDeviceCopyVertexBuffer: one read and one write of m^3 cells per
    launch) for calibrating memory counters and as the measured copy bandwidth."""
    from astaroth_amd.api import Library

    lib = Library(double)
    info = lib.config(n=n)
    dev = lib.device_create(0, info, autotune=False)
    try:
        stream = dev.stream_handle()
        dev.copy_vertex_buffer(0)
        dev.synchronize()
        a, b = hip.event(), hip.event()
        hip.record(a, stream)
        for r in range(reps):
            dev.copy_vertex_buffer(r % 8)
        hip.record(b, stream)
        ms = hip.elapsed(a, b) / reps
    finally:
        dev.destroy()
    nbytes = 2.0 * (n + 6) ** 3 * (8 if double else 4)
    return dict(bytes_per_launch=nbytes, ms_per_launch=ms, gbs=nbytes / (ms * 1e-3) / 1e9)


def measured_traffic(prec: str, n: int, cfg: dict):
    """HBM bytes per solve launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summaries
    (profiles/*_hbm_traffic.md, counters scaled by a copy-kernel calibration), when one exists for
    this precision, grid and kernel configuration; otherwise None."""
    import glob
    import re

    key = "%d,%d,%d,%d" % (cfg["variant"], cfg["tile_x"], cfg["tile_y"], cfg["z_depth"])
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*_%dcubed_hbm_traffic.md" % n))):
        for line in open(path):
            m = re.match(r"traffic_(fp32|fp64) (\S+) bytes_per_launch=(\S+)", line.strip())
            if m and m.group(1) == prec and m.group(2) == key:
                return float(m.group(3)), os.path.relpath(path, REPO)
    return None, None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The oracle (the CPU restatement of the host model solver, acHostIntegrateStep's form) on the
    host cores: one RK3 step of the uniform-field box at 64^3 (FP64 and long double) and 256^3
    (FP64), single-threaded (the reference's host solver runs on one thread, its OpenMP pragmas
    commented out at modelsolver.c:1018,1028) and on all the cores this process may use (OpenMP
    over k planes, at most the thread budget the environment sets). `value` is the all-core
    256^3 FP64 rate."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import pyoracle
    from astaroth_amd.api import Library, real_params_double
    from astaroth_amd.initial import mhd_fields

    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or usable
    all_cores = max(1, min(usable, budget))
    runs = []

    initial = {}

    def timed(n, dtype, threads):
        threads = pyoracle.set_threads(threads)
        if n not in initial:
            initial.clear()
            initial[n] = mhd_fields(n, n, n, np.float64, config_id=1)
        fields = [f.astype(dtype) for f in initial[n]]
        rp = real_params_double(Library(True).config(n=n))
        t0 = time.perf_counter()
        pyoracle.integrate_steps(fields, rp, 1e-3, 1)
        dt = time.perf_counter() - t0
        del fields
        runs.append({"n": n, "dtype": np.dtype(dtype).name, "threads": threads, "steps": 1, "s": dt,
                     "cell_updates_per_s": n ** 3 / dt})
        return runs[-1]

    timed(32, np.float64, all_cores)  # untimed in effect: first touch of the thread pool
    runs.clear()
    for threads in (1, all_cores):
        timed(64, np.float64, threads)
        timed(64, np.longdouble, threads)
    best = timed(256, np.float64, all_cores)
    single = timed(256, np.float64, 1)
    return dict(value=best["cell_updates_per_s"], unit="cell-updates/s", cores=best["threads"], kind="port",
                sample="256^3 uniform-field box, 1 RK3 step, FP64 host-form model solver (oracle/), %d threads, "
                       "%.2f s" % (best["threads"], best["s"]),
                single_thread=single["cell_updates_per_s"], nproc=os.cpu_count(), usable_cores=usable,
                cpu_model=cpu_model(), runs=runs)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--autotune", action="store_true",
                    help="run the library's autotuner at device creation instead of using its default "
                         "configuration (the one measured fastest at 512^3, whose HBM traffic is profiled)")
    ap.add_argument("--no-autotune", action="store_true", help="the default; kept for older command lines")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-config", default=None,
                    help="force a solve configuration 'variant,tile_x,tile_y,z_depth' instead of autotuning")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="REPS",
                    help="only run REPS launches of the plain copy kernel and print its bandwidth")
    ap.add_argument("--workload", choices=["headline", "sweep256", "precision512", "ensemble96", "halo1024",
                                           "reductions", "boundaries", "node"],
                    default="headline",
                    help="headline: the 512^3 FP32 vs FP64 metric (default). sweep256: 256^3 with per-step "
                         "FP64 diagnostics and the tuning sweep. precision512: 512^3 FP32-vs-FP64 precision "
                         "study. ensemble96: 96-box FP32 ensemble, batched vs looped. halo1024: 1024^3 FP32 with "
                         "the 26-block pack/unpack halo. reductions: MAX/RMS of lnrho and |u| at --n (512^3). "
                         "boundaries: periodic fill and 26-block pack/unpack at --n (512^3). node: "
                         "acNodeIntegrate against acDeviceIntegrateStep, FP32, at --n (512^3)")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this path has no multi-GPU layer (single periodic box "
              "per device); run it with --gpus 1", file=sys.stderr)
        return 2
    hip = Hip()
    if a.copy_probe:
        for prec in (["fp32", "fp64"] if a.precision == "both" else [a.precision]):
            print(json.dumps(dict(copy_probe=prec, n=a.n, **copy_probe(prec == "fp64", a.n, a.copy_probe, hip))),
                  flush=True)
        return 0
    kcfg = tuple(int(v) for v in a.kernel_config.split(",")) if a.kernel_config else None
    if a.workload != "headline":
        return run_workload(a, kcfg, hip)
    res = {}
    for prec in (["fp32", "fp64"] if a.precision == "both" else [a.precision]):
        res[prec] = run_gpu(prec == "fp64", a.n, a.steps if prec == "fp32" else max(3, a.steps // 2),
                            a.warmup, a.dt, a.autotune and kcfg is None, hip, kcfg)
        print("# %s: %.3f ms/step, %.4g cell-updates/s, solve %.3f ms/launch, %.0f GB/s, config %s"
              % (prec, res[prec]["ms_per_step"], res[prec]["value"], res[prec]["solve_ms_mean"],
                 res[prec]["achieved_gbs"], res[prec]["kernel_config"]), file=sys.stderr, flush=True)
    main_prec = "fp32" if "fp32" in res else "fp64"
    r = res[main_prec]
    traffic, traffic_src = measured_traffic(main_prec, a.n, r["kernel_config"])
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_prec == "fp32" else "f64",
        "data": "synthetic: seeded splitmix64 uniform fields (lnrho~U(-1e-2,1e-2), u,A~U(-1e-3,1e-3), s=0)",
        "config": {
            "workload": "8-field compressible MHD, %d^3 periodic box, 6th-order stencils, 2N-RK3, dt=%g" % (a.n, a.dt),
            "grid": [a.n, a.n, a.n],
            "step": STEP_FORM[r["reads_images"]],
            "kernel_config": r["kernel_config"],
            "median_step_ms": r["median_step_ms"],
            "p90_step_ms": r["p90_step_ms"],
            "solve_fraction_of_step": r["solve_fraction"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "solve (fused MHD stencil + RK3 substep)",
            "achieved": r["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": r["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "alg_bytes_per_launch": r["alg_bytes_per_launch"],
            "solve_ms_per_launch": r["solve_ms_mean"],
            "valu": valu_roofline(r, a.n),
            "ceiling": ceilings(r, a.n, main_prec == "fp64"),
        },
    }
    if "fp64" in res and "fp32" in res:
        f64 = res["fp64"]
        line["fp64"] = {"value": f64["value"], "ms_per_step": f64["ms_per_step"],
                        "achieved_gbs": f64["achieved_gbs"], "frac": f64["achieved_gbs"] / HBM_PEAK_GBS,
                        "kernel_config": f64["kernel_config"], "solve_ms_per_launch": f64["solve_ms_mean"],
                        "valu": valu_roofline(f64, a.n), "ceiling": ceilings(f64, a.n, True),
                        "step": STEP_FORM[f64["reads_images"]]}
        line["fp32_over_fp64"] = r["value"] / f64["value"]
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


def run_workload(a, kcfg, hip):
    """The other measured configurations; each prints one JSON line of its own."""
    base = {"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "data": "synthetic: seeded splitmix64 uniform fields"}
    if a.workload == "precision512":
        n = a.n
        steps = 10
        r = precision_study(n, steps, a.dt, hip)
        line = dict(base, metric="FP32 vs FP64 range-normalised max error after %d RK3 steps at %d^3" % (steps, n),
                    value=r["worst_range_normalised"], unit="relative", higher_is_better=False, steps=steps,
                    warmup=0, dtype="f32 vs f64",
                    config={"workload": "precision study: uniform-field state + u_x 1e-2 sin(2z), %d^3, dt=%g" % (n, a.dt)},
                    precision=r)
        line["data"] = "synthetic: seeded splitmix64 perturbed fields (uniform fields + u_x 1e-2 sin 2z)"
        print(json.dumps(line), flush=True)
        return 0
    if a.workload == "halo1024":
        n = 1024 if a.n == 512 else a.n
        # The explicit grad-div (bulk + shear) viscosity limits dt ~ dx^2: dt = 1e-3 is inside the
        # RK3 stability region at 512^3 (nu_eff dt |lambda_max| = 2.0 < 2.5) but not at 1024^3, so
        # the step is scaled with dx^2 (same margin as 512^3). Throughput does not depend on dt.
        dt = a.dt * (512.0 / n) ** 2
        r = run_gpu(False, n, a.steps if a.steps != 20 else 5, 1, dt, False, hip, kcfg, halo="pack",
                    reduce_each_step=True)
        line = dict(base, metric="cell-updates/s per RK3 step at %d^3 FP32, 26-block pack/unpack halo" % n,
                    value=r["value"], unit="cell-updates/s", steps=a.steps if a.steps != 20 else 5, warmup=1,
                    ms_per_step=r["ms_per_step"], dtype="f32",
                    config={"workload": "8-field MHD, %d^3 periodic box, dt=%g, halo by packing the 26 ghost "
                                        "blocks' periodic images and unpacking them (self exchange), FP64 "
                                        "diagnostics after every step" % (n, dt),
                            "kernel_config": r["kernel_config"], "median_step_ms": r["median_step_ms"],
                            "p90_step_ms": r["p90_step_ms"], "solve_fraction_of_step": r["solve_fraction"],
                            "halo_ms_per_substep": r["halo_ms_mean"]},
                    roofline={"bound": "hbm", "kernel": "solve (fused MHD stencil + RK3 substep)",
                              "achieved": r["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": r["achieved_gbs"] / HBM_PEAK_GBS, "traffic": None,
                              "solve_ms_per_launch": r["solve_ms_mean"],
                              "valu": valu_roofline(r, n), "ceiling": ceilings(r, n, False)},
                    halo={"bytes_per_exchange": r["halo_alg_bytes"], "ms_per_exchange": r["halo_ms_mean"],
                          "achieved_gbs": r["halo_alg_bytes"] / (r["halo_ms_mean"] * 1e-3) / 1e9},
                    diagnostics_last_step=r["reductions"][-1] if r["reductions"] else None)
        print(json.dumps(line), flush=True)
        return 0
    if a.workload == "ensemble96":
        r = ensemble_benchmark(a.dt, hip)
        best = r["batched_graph"]
        line = dict(base, metric="box-cell-updates/s per RK3 step, 96-box FP32 ensemble (one graph per step)",
                    value=best["value"], unit="box-cell-updates/s", steps=r["steps"], warmup=r["warmup"],
                    ms_per_step=best["ms_per_step"], dtype="f32",
                    config={"workload": "96 independent periodic boxes, edges from {32,48,64,96,128} by seed "
                                        "(cubic and anisotropic), nu ~ U[1e-3, 1e-2] per box, dt=%g" % a.dt,
                            "boxes": r["boxes"], "total_cells": r["cells"], "kernel_config": r["kernel_config"]},
                    results=r)
        line["vs_looped"] = best["value"] / r["looped"]["value"]
        print(json.dumps(line), flush=True)
        return 0
    if a.workload == "reductions":
        r = reductions_benchmark(a.n, hip)
        best = max(r, key=lambda x: x["gbs"])
        line = dict(base, metric="HBM GB/s of the FP64-accumulated reductions at %d^3" % a.n, value=best["gbs"],
                    unit="GB/s", steps=5, warmup=1, dtype="f32/f64 data, f64 accumulation",
                    config={"workload": "MAX and RMS of lnrho (1 field) and |u| (3 fields), %d^3 uniform-field box"
                                        % a.n}, results=r)
        print(json.dumps(line), flush=True)
        return 0
    if a.workload == "node":
        r = node_benchmark(a.n, a.dt)
        line = dict(base, metric="cell-updates/s per RK3 step of acNodeIntegrate at %d^3 FP32" % a.n,
                    value=a.n ** 3 / (r["node_ms_per_step"] * 1e-3), unit="cell-updates/s", steps=5, warmup=1,
                    ms_per_step=r["node_ms_per_step"], dtype="f32",
                    config={"workload": "8-field MHD, %d^3 periodic box, Node step (the Device step between "
                                        "synchronisations) against the Device step" % a.n}, results=r)
        print(json.dumps(line), flush=True)
        return 0
    if a.workload == "boundaries":
        out = {prec: boundaries_benchmark(a.n, hip, prec == "fp64") for prec in ("fp32", "fp64")}
        line = dict(base, metric="HBM GB/s of the periodic fill and the 26-block pack/unpack at %d^3" % a.n,
                    value=out["fp32"]["periodic_fill"]["gbs"], unit="GB/s", steps=10, warmup=1, dtype="f32",
                    config={"workload": "ghost shell of 8 fields, %d^3 periodic box" % a.n}, results=out)
        print(json.dumps(line), flush=True)
        return 0
    if a.workload == "sweep256":
        n = 256 if a.n == 512 else a.n
        out = {}
        for prec in ("fp32", "fp64"):
            r = run_gpu(prec == "fp64", n, 10, 2, a.dt, True, hip, kcfg, reduce_each_step=True)
            out[prec] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "achieved_gbs": r["achieved_gbs"],
                         "frac": r["achieved_gbs"] / HBM_PEAK_GBS, "kernel_config": r["kernel_config"],
                         "diagnostics": r["reductions"]}
            out[prec]["sweep"] = tuning_sweep(prec == "fp64", n)
        line = dict(base, metric="cell-updates/s per RK3 step at %d^3, FP32 and FP64, per-step FP64 diagnostics" % n,
                    value=out["fp32"]["value"], unit="cell-updates/s", steps=10, warmup=2,
                    ms_per_step=out["fp32"]["ms_per_step"], dtype="f32",
                    config={"workload": "8-field MHD, %d^3 periodic box, autotuned" % n}, results=out)
        print(json.dumps(line), flush=True)
        return 0
    raise ValueError(a.workload)


def ensemble_benchmark(dt, hip, steps=10, warmup=2):
    """The 96-box ensemble advanced by the batched step (eager and graph) and by a loop of
    single-box launches; box-cell-updates/s = sum of interior cells / step time."""
    from astaroth_amd.api import Library
    from astaroth_amd.initial import ensemble_init_mhd, ensemble96_boxes

    lib = Library(False)
    boxes, nus = ensemble96_boxes()
    infos = [lib.config(n=b[0], ny=b[1], nz=b[2], AC_nu_visc=nu) for b, nu in zip(boxes, nus)]
    cells = float(sum(b[0] * b[1] * b[2] for b in boxes))
    out = {"boxes": boxes, "cells": cells, "steps": steps, "warmup": warmup}
    t0 = time.perf_counter()
    with lib.ensemble_create(infos, autotune=True) as ens:  # the library's tuner runs at creation
        out["autotune_s"] = time.perf_counter() - t0
        tuned = ens.kernel_config()
        depths = ens.box_z_depths()
        stream = ens.stream_handle()

        def timed(mode, nsteps):
            for b in range(len(infos)):
                ensemble_init_mhd(ens, b, config_id=4, box_id=b)
            ens.periodic_boundconds()
            for _ in range(warmup):
                ens.integrate_step(dt, mode=mode)
            ens.synchronize()
            e0, e1 = hip.event(), hip.event()
            hip.record(e0, stream)
            for _ in range(nsteps):
                ens.integrate_step(dt, mode=mode)
            hip.record(e1, stream)
            ms = hip.elapsed(e0, e1) / nsteps
            return {"ms_per_step": ms, "value": cells / (ms * 1e-3)}

        # fixed configurations for every box, beside the tuner's choice
        sweep = []
        for cfg in ((32, 4, 0), (32, 4, 28), (16, 8, 0), (16, 8, 28), (16, 16, 0), (16, 16, 28)):
            ens.set_kernel_config(*cfg)
            sweep.append([list(cfg), timed("batched", 3)["ms_per_step"]])
        sweep.sort(key=lambda x: x[1])
        ens.autotune()  # back to the tuned choice (kept for this composition)
        assert ens.kernel_config().as_tuple() == tuned.as_tuple() and ens.box_z_depths() == depths
        out["fixed_config_sweep"] = sweep
        out["kernel_config"] = [tuned.tile_x, tuned.tile_y]
        out["z_depth_per_size_class"] = {str(b[2]): d for b, d in zip(boxes, depths)}
        out["tuned_batched"] = timed("batched", 3)["ms_per_step"]
        out["batched_eager"] = timed("batched", steps)
        out["batched_graph"] = timed("graph", steps)
        out["looped"] = timed("looped", steps)
        d = ens.diagnostics()
        if not np.isfinite(d).all() or d[:, 0].max() > 1.0:
            raise RuntimeError("ensemble state diverged")
        out["diagnostics_max_u_range"] = [float(d[:, 0].min()), float(d[:, 0].max())]
    return out


def reductions_benchmark(n: int, hip: Hip, reps=5):
    """acDeviceReduceScalDouble / acDeviceReduceVecDouble, MAX and RMS of lnrho and of |u|, `reps`
    calls each on an FP32 and an FP64 Device holding the uniform-field box; each call is bracketed
    by HIP events on the library's stream (both launches and the 8-byte result copy). Algorithmic
    bytes: fields read x n^3 x sizeof(AcReal)."""
    from astaroth_amd.api import Library
    from astaroth_amd.initial import device_init_mhd

    out = []
    for double in (False, True):
        lib = Library(double)
        with lib.device_create(0, lib.config(n=n), autotune=False) as dev:
            device_init_mhd(dev)
            stream = dev.stream_handle()
            calls = [("scalar", "lnrho", 1, lambda r: dev.reduce_scal(r, "VTXBUF_LNRHO", double=True)),
                     ("vector", "|u|", 3, lambda r: dev.reduce_vec(r, "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ",
                                                                   double=True))]
            for kind, what, nfields, fn in calls:
                for rtype in ("RTYPE_MAX", "RTYPE_RMS"):
                    fn(rtype)  # warm
                    ms = []
                    for _ in range(reps):
                        a, b = hip.event(), hip.event()
                        hip.record(a, stream)
                        fn(rtype)
                        hip.record(b, stream)
                        hip.sync()
                        ms.append(hip.elapsed(a, b))
                    nbytes = float(nfields) * n ** 3 * (8 if double else 4)
                    med = statistics.median(ms)
                    out.append({"build": "fp64" if double else "fp32", "reduction": "%s %s %s" % (kind, rtype, what),
                                "fields": nfields, "median_ms": med, "bytes": nbytes,
                                "gbs": nbytes / (med * 1e-3) / 1e9})
                    print("# %-4s %-22s %.1f us, %.0f GB/s" % (out[-1]["build"], out[-1]["reduction"], med * 1e3,
                                                              out[-1]["gbs"]), file=sys.stderr, flush=True)
    return out


def node_benchmark(n: int, dt: float, steps=5, rounds=3):
    """acNodeIntegrate (the Device step between synchronisations) against
    acDeviceIntegrateStep on the same n^3 FP32 state, alternating `rounds` times between the two
    so that both see the same clocks; host wall time per RK3 step (both are synchronous calls)."""
    from astaroth_amd.api import Device, Library
    from astaroth_amd.initial import device_init_mhd

    lib = Library(False)
    info = lib.config(n=n)
    node_ms, dev_ms = [], []
    with lib.node_create(info, autotune=False) as node:
        # the Node's own Device, driven through the Device calls
        dev = Device.view(lib, info, node.query_device_configuration().devices[0])
        device_init_mhd(dev)
        images = lib.dll.acDeviceSubstepReadsPeriodicImages(dev.handle) == 1
        node.integrate(dt)
        dev.integrate_step(dt)
        dev.synchronize()
        for _ in range(rounds):
            t0 = time.perf_counter()
            for _ in range(steps):
                node.integrate(dt)
            node_ms.append((time.perf_counter() - t0) * 1e3 / steps)
            t0 = time.perf_counter()
            for _ in range(steps):
                dev.integrate_step(dt)
            dev.synchronize()
            dev_ms.append((time.perf_counter() - t0) * 1e3 / steps)
            print("# node %.2f ms/step, device %.2f ms/step" % (node_ms[-1], dev_ms[-1]), file=sys.stderr, flush=True)
    return {"node_ms_per_step": min(node_ms), "device_ms_per_step": min(dev_ms),
            "node_over_device": min(node_ms) / min(dev_ms), "device_reads_images": images,
            "all_node_ms": node_ms, "all_device_ms": dev_ms}


def boundaries_benchmark(n: int, hip: Hip, double: bool, reps=10):
    """The periodic ghost fill (acDevicePeriodicBoundconds) and the 26-block self exchange
    (acDeviceHaloExchangeSelf: pack + unpack), `reps` launches each, timed with HIP events on the
    library's stream. Algorithmic bytes: the ghost shell g = m^3 - n^3 cells of 8 fields, read and
    written once by the fill (2 x 8 g s); read, packed, unpacked and written by the exchange
    (4 x 8 g s)."""
    from astaroth_amd.api import Library
    from astaroth_amd.initial import device_init_mhd

    lib = Library(double)
    s = 8 if double else 4
    g = float((n + 6) ** 3 - n ** 3)
    res = {}
    with lib.device_create(0, lib.config(n=n), autotune=False) as dev:
        device_init_mhd(dev)
        stream = dev.stream_handle()
        for name, fn, nbytes in (("periodic_fill", dev.periodic_boundconds, 2 * 8 * g * s),
                                 ("pack_unpack_26", dev.halo_exchange_self, 4 * 8 * g * s)):
            fn()
            dev.synchronize()
            a, b = hip.event(), hip.event()
            hip.record(a, stream)
            for _ in range(reps):
                fn()
            hip.record(b, stream)
            hip.sync()
            ms = hip.elapsed(a, b) / reps
            res[name] = {"ms": ms, "bytes": nbytes, "gbs": nbytes / (ms * 1e-3) / 1e9,
                         "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
            print("# %s %d^3 %s: %.1f us, %.0f GB/s" % ("fp64" if double else "fp32", n, name, ms * 1e3,
                                                      res[name]["gbs"]), file=sys.stderr, flush=True)
    return res


def tuning_sweep(double: bool, n: int):
    """Mean substep time of every compiled configuration and z-depth (the autotuner's search)."""
    from astaroth_amd.api import Library
    from astaroth_amd.abi import KernelConfig
    from astaroth_amd.initial import device_init_mhd

    lib = Library(double)
    dev = lib.device_create(0, lib.config(n=n), autotune=False)
    res = []
    try:
        device_init_mhd(dev)
        for c in lib.kernel_configs():
            # the library tuner's candidates: powers of two for variant 1, whole 8-plane rounds for
            # variant 13, the box depth cut in 2, 3, 4 or 8 for the other lean marches
            depths = [0] if c.variant == 0 else ([8, 16, 32, 64, 128, 0] if c.variant == 1 else
                                                 [16, 32, 64, 128, 0] if c.variant == 13 else
                                                 [(n + 1) // 2, (n + 2) // 3, (n + 3) // 4, (n + 7) // 8, 0])
            for zd in depths:
                if zd >= n:
                    continue
                cc = KernelConfig(c.variant, c.tile_x, c.tile_y, zd)
                res.append([list(cc.as_tuple()), dev.time_kernel_config(cc, 3)])
    finally:
        dev.destroy()
    return sorted(res, key=lambda x: x[1])


if __name__ == "__main__":
    sys.exit(main())
