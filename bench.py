#!/usr/bin/env python3
"""Benchmark: Mcell-updates/s per RK stage of the 64-patch subsonic jet on one MI355X.

Workload: 256^3 core cells as 4x4x4 patches of 64^3 core cells with two seek
layers on every interior patch face (patches stored as 66 or 68 cells per axis,
19,248,832 stored cells), AUSM+ with MUSCL/van Albada, 3-stage RK
(alpha = 1/4, 1/2, 1), exchange after every stage, top-hat Mach-0.9 inflow on
the z=0 face (Mach-0.1 co-flow around it), transmissive elsewhere.  A "step"
is one RK step over all patches.

  value       FP32 batched form (one launch per stage over the tile work-list of
              all 64 patches), state resident in HBM, timed region = K steps.
  fp64_plain  FP64 plain per-patch form (one launch per patch per stage).
  fp32_per_patch the FP32 arithmetic launched per patch (the batching gain).
  fp32_ieee   the FP32 batched form in the reference's arithmetic (IEEE division
              and square root instead of the hardware reciprocal).
  reference_faces the FP32 batched form on the same jet with the other domain
              faces as the reference's JetFlux has them (far-field lateral
              faces, a slip wall at the far end).
  roofline    stage kernel: algorithmic bytes per launch (10 or 15 values of
              4 B per stored cell: stage 1 reads U^n only) / average launch time
              from HIP events on the context stream over the timed region.
  valu_roofline the same kernel against its compute ceiling: VALU wave
              instructions per launch (committed rocprofv3 summary) / CUs /
              launch time, against the FP32 VALU issue rate measured by
              drnum_amd/lib/valu_probe after the timed region.  Figures from a
              committed summary carry the kernel-source sha256 they were taken
              of; a summary of other sources is marked stale.
  pcie_inclusive the same steps with one upload of the whole state before them
              and one download after them (drn_patch_upload/download, pageable
              host buffers), as a caller that keeps the state on the host runs.
  cpu_baseline the oracle (C restatement of the reference iterator, OpenMP over
              patches) on this host, one RK step of the same workload; its
              single_thread entry is one stage of the same grid on one thread.

  accuracy    (--accuracy) both forms after the configuration's step count:
              FP32 density Linf against FP64, FP64-sum drift of mass and
              energy, and the momentum flux through two planes of the
              level-set case.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W] [--accuracy]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s per RK stage + % HBM peak; FP32-batched vs FP64-plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(w, dtype, launch_tables=True):
    from drnum_amd import hip, host

    hg = host.HostGrid.from_workload(w, fp64_geometry=True)
    hg.compute_dependencies(w.periodic)
    ctx = hip.Context(dtype)
    ctx.set_operator(w.recon, w.limiter, w.viscous)
    hg.to_device(ctx)
    if launch_tables:
        hg.upload_exchange(ctx)
    if w.spheres is not None:
        # slip level-set fill after every stage (the post operation of the RK step)
        hg.levelset_spheres(w.spheres, w.ls_inner_layers, 0, w.ls_min_inner_rel)
        hg.upload_levelset(ctx)
    for p in range(len(w.patches)):
        ctx.upload(p, 0, w.initial_state(p, dtype))
    return hg, ctx


def stage_bytes_per_step(cells, esize, n_stages):
    # stage 1 reads U^n once (src == base) and writes 5 values; later stages read src and base
    vals = 10 + 15 * (n_stages - 1)
    return cells * vals * esize


def run_form(w, dtype, steps, warmup, launch, label, tiling=None, arith=None):
    from drnum_amd import hip

    hg, ctx = setup(w, dtype)
    if arith is not None:
        ctx.set_arith(arith)
    cells = w.total_cells
    esize = np.dtype(dtype).itemsize
    S = len(w.alpha)
    ex_bytes, ls_bytes = exchange_fill_bytes(hg, w, esize)
    n_fused = 0
    if tiling is None:
        # the library's autotuner picks the stage work-list (i-range, XCD order, kernel) for this form
        tiling = ctx.autotune(launch)
        log("%s: autotuned tiling %d" % (label, tiling))
    else:
        ctx.set_tiling(tiling, launch)
        log("%s: tiling %d" % (label, tiling))
    # receivers the stage kernel fills itself (the fused z-face exchange): the
    # exchange after the stage moves the bytes of the others only (one
    # unit-weight donor each: 52 B in FP32)
    n_fused = ctx.exchange_fused_cells(launch)
    ex_bytes -= n_fused * ((4 + 5 * esize) + (4 + 6 * esize))
    for _ in range(warmup):
        ctx.rk_step(w.dt, w.alpha, launch)
    ctx.sync()
    ctx.timing(0)
    ctx.timing(1)
    ctx.timing(2)
    ctx.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.rk_step(w.dt, w.alpha, launch)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.timing_enable(False)
    st_ms, st_n = ctx.timing(0)
    ex_ms, ex_n = ctx.timing(1)
    ls_ms, ls_n = ctx.timing(2)
    wall = t1 - t0
    diag = ctx.diagnostics(0, w.dt)
    log("%s: %d steps %.3f s, stage kernel avg %.4f ms (%d), exchange avg %.4f ms (%d, %d receivers filled by "
        "the stage kernel), diag %s" % (label, steps, wall, st_ms, st_n, ex_ms, ex_n, n_fused, diag))
    if diag["nonfinite"] > 0:
        raise RuntimeError("%s: non-finite values after the timed steps" % label)
    res = dict(wall=wall, ms_per_step=1e3 * wall / steps, mcups=cells * S * steps / wall / 1e6,
               stage_ms=st_ms, stage_n=st_n, exch_ms=ex_ms, exch_n=ex_n, tiling=tiling,
               exch_bytes=ex_bytes, ls_ms=ls_ms, ls_n=ls_n, ls_bytes=ls_bytes, fused=n_fused)
    # per launch the stage kernel advances all stored cells once
    bytes_step = stage_bytes_per_step(cells, esize, S)
    res["stage_bytes_per_launch"] = bytes_step / S
    if launch == hip.DRN_LAUNCH_BATCHED:
        res["stage_gbs"] = res["stage_bytes_per_launch"] / (st_ms * 1e-3) / 1e9 if st_ms > 0 else 0.0
    else:
        # per-patch form: st_ms is one launch over all patches of one stage
        res["stage_gbs"] = res["stage_bytes_per_launch"] / (st_ms * 1e-3) / 1e9 if st_ms > 0 else 0.0
    # graph replay of the same form (batched only): captured steps, no per-launch host work
    if launch == hip.DRN_LAUNCH_BATCHED:
        # untimed: capture and instantiate the graphs of both starting field
        # maps of the step cycle, then replay warm ones
        ctx.rk_graph_run(w.dt, w.alpha, 2 * max(1, warmup))
        ctx.rk_graph_run(w.dt, w.alpha, 1)
        ctx.rk_graph_run(w.dt, w.alpha, 2)
        ctx.sync()
        tg0 = time.perf_counter()
        ctx.rk_graph_run(w.dt, w.alpha, steps)
        ctx.sync()
        tg1 = time.perf_counter()
        res["graph_ms_per_step"] = 1e3 * (tg1 - tg0) / steps
        res["graph_mcups"] = cells * S * steps / (tg1 - tg0) / 1e6
    # host hand-over through PCIe (drn_patch_upload / drn_patch_download of every
    # patch, pageable host buffers): one upload of the whole state before a run
    # and one download after it, as a caller that keeps the state on the host does
    if launch == hip.DRN_LAUNCH_BATCHED:
        host_states = [w.initial_state(p, dtype) for p in range(len(w.patches))]
        ctx.sync()
        tu0 = time.perf_counter()
        for p, a in enumerate(host_states):
            ctx.upload(p, 0, a)
        ctx.sync()
        tu1 = time.perf_counter()
        for p in range(len(w.patches)):
            ctx.download(p, 0)
        ctx.sync()
        tu2 = time.perf_counter()
        res["upload_s"], res["download_s"] = tu1 - tu0, tu2 - tu1
        res["state_bytes"] = sum(a.nbytes for a in host_states)
        del host_states
    # STREAM-style copy of one state buffer on the same device
    ctx.copy_field(0, 2)
    ctx.sync()
    ctx.timing(3)
    ctx.timing_enable(True)
    for _ in range(10):
        ctx.copy_field(0, 2)
    ctx.timing_enable(False)
    cp_ms, cp_n = ctx.timing(3)
    nbytes = 2 * 5 * ctx_vstride(ctx, w) * esize
    res["copy_gbs"] = nbytes / (cp_ms * 1e-3) / 1e9 if cp_ms > 0 else 0.0
    ctx.close()
    return res


def exchange_fill_bytes(hg, w, esize):
    """Algorithmic bytes of one exchange and of one level-set fill in the tables'
    device layout.  Exchange: per receiver its 4-byte index and 5 values written,
    per (donor index, weight) entry 4 B + one weight + 5 donor values read.
    Fill, per layer cell: 8 x (4 B index + weight + 5 values) gathered and 5
    values written to the accumulator, then the cell index, the normal (3), the
    accumulator (5) and the current momentum (3) read and 5 values written."""
    n_rec = n_pairs = 0
    for p in range(len(w.patches)):
        t = hg.tables(p)
        n_rec += len(t["receivers"])
        n_pairs += len(t["donor_index"])
    ex = n_rec * (4 + 5 * esize) + n_pairs * (4 + 6 * esize)
    n_ls = 0
    if w.spheres is not None:
        for p in range(len(w.patches)):
            for layer in range(w.ls_inner_layers):
                n_ls += len(hg.ls_cells(p, layer)["cells"])
    ls = n_ls * (8 * (4 + 6 * esize) + 5 * esize + 4 + 3 * esize + 5 * esize + 3 * esize + 5 * esize)
    return ex, ls


def kernel_roofline(ms, nbytes, traffic=None):
    """achieved GB/s of a launch moving nbytes algorithmic bytes in ms, against HBM peak"""
    if not ms or not nbytes:
        return None
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "ms": round(ms, 4), "algorithmic_bytes_per_launch": int(nbytes),
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic}


def measured_exchange_traffic():
    """HBM bytes per exchange of the 64-patch jet from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibration applied)."""
    try:
        lines = open(TRAFFIC_TABLE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        cols = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cols) == 5 and cols[0].startswith("k_exchange_single<float>"):
            try:
                return float(cols[3])
            except ValueError:
                return None
    return None


def _workloads():
    from drnum_amd import configs

    return {
        "jet": (configs.jet, "subsonic jet, 256^3 core as 4x4x4 patches of 64^3 core + 2 seek layers on interior "
                             "faces; AUSM+ MUSCL van Albada, 3-stage RK, exchange every stage"),
        "uneven": (configs.uneven_batch, "uneven batch: 200 patches of 16-96 cells per axis with 2-4 cell overlaps, "
                                         "20% nested at ratio 2; Mach 0.5 + density pulse"),
        "spheres": (configs.levelset_spheres, "level set: 192x128x128 core as 12 patches, 8 spheres of radius 6-14 "
                                              "cells, Mach 0.3 farfield, slip level-set fill every stage"),
        "kh": (configs.kelvin_helmholtz, "periodic Kelvin-Helmholtz box, 512^3 core as 8x8x8 patches of 64^3 + 2 "
                                         "seek layers on every face"),
        "shear": (lambda: configs.viscous_shear(core=64, blocks=(4, 4, 4)),
                  "viscous periodic shear layer, 256^3 core of 1 um cells at 1 bar / 300 K as 4x4x4 patches of 64^3 "
                  "+ 2 seek layers on every face; AUSM+ MUSCL van Albada + CompressibleViscFlux, 3-stage RK"),
    }


WORKLOADS = _workloads()

TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "stage_traffic_jet4x64_fp32.md")


def measured_traffic(tiling):
    """HBM bytes per stage launch of the FP32 stage kernel for a work-list choice,
    from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (counters scaled by the copy-kernel calibration).  None when not measured."""
    try:
        lines = open(TRAFFIC_TABLE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        cols = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cols) == 8 and cols[0].isdigit() and int(cols[0]) == tiling and cols[1].startswith("k_stage"):
            try:
                return float(cols[6])
            except ValueError:
                return None
    return None


VALU_TABLE = os.path.join(ROOT, "profiles", "stage_valu_jet4x64_fp32.md")
KERNEL_SOURCES = [os.path.join(ROOT, "drnum_amd", "csrc", f) for f in ("drn_kernels.h", "drn_physics.h", "drnhip.hip")]


def kernel_source_hash():
    """sha256 (first 16 hex digits) of the device library's sources, the build
    a committed rocprofv3 summary records as 'kernel source sha256'"""
    import hashlib

    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profile_hash(path):
    """the 'kernel source sha256: <hash>' line of a committed profile summary"""
    try:
        for ln in open(path).read().splitlines():
            if "kernel source sha256:" in ln:
                return ln.split("kernel source sha256:", 1)[1].split()[0].strip("`")
    except OSError:
        pass
    return None


def profile_current(path):
    """whether a committed profile summary was taken of the current kernel sources"""
    return profile_hash(path) == kernel_source_hash()
VALU_PROBE = os.path.join(ROOT, "drnum_amd", "lib", "valu_probe")


def measured_valu_instructions(tiling):
    """SQ_INSTS_VALU (wave instructions) per launch of the FP32 stage kernel for a
    work-list choice, from the committed rocprofv3 summary of this workload."""
    try:
        lines = open(VALU_TABLE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        cols = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cols) == 7 and cols[0].isdigit() and int(cols[0]) == tiling and cols[1].startswith("k_stage"):
            try:
                return float(cols[2].replace(",", ""))
            except ValueError:
                return None
    return None


def valu_ceiling():
    """The device's FP32 VALU issue rate (wave instructions per CU per ns) from
    drnum_amd/lib/valu_probe, run as a child process after the measurement."""
    if not os.path.exists(VALU_PROBE):
        return None
    try:
        r = subprocess.run([VALU_PROBE], capture_output=True, text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return None
    if r.returncode != 0:
        log("valu_probe failed:", r.stderr.strip())
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def plane_momentum_flux(w, rho_states, gi):
    """x-momentum flux (rho u^2 + p) * dy * dz summed over the core cells of the
    global i-plane gi (a control surface normal to the freestream), FP64."""
    total = 0.0
    for p, P in enumerate(w.patches):
        li = gi - P.gidx[0]
        if not (P.seek[0] <= li < P.ni - P.seek[1]):
            continue
        U = rho_states[p].reshape(5, P.ni, P.nj, P.nk)[:, li, P.seek[2]:P.nj - P.seek[3], P.seek[4]:P.nk - P.seek[5]]
        U = U.astype(np.float64)
        rho = U[0]
        u, v, ww = U[1] / rho, U[2] / rho, U[3] / rho
        T = (U[4] / rho - 0.5 * (u * u + v * v + ww * ww)) / 717.5
        pr = rho * 287.2 * T
        total += float(np.sum(rho * u * u + pr)) * P.dx(1) * P.dx(2)
    return total


def accuracy_report(w, steps):
    """The FP64 per-patch and the FP32 batched form of a workload after `steps` RK
    steps (outside any timed region): density Linf of FP32 against FP64 relative
    to max rho, the drift of the FP64 core sums of mass and energy in both forms,
    and for a level-set workload the x-momentum flux through an upstream and a
    downstream plane with the FP32 error."""
    from drnum_amd import hip

    out = {"steps": steps}
    rho64 = None
    planes = None
    if w.spheres is not None:
        nx = max(P.gidx[0] + P.ni - P.seek[1] for P in w.patches)
        planes = (nx // 12, nx - nx // 12)
    for dtype, launch, name in ((np.float64, hip.DRN_LAUNCH_PER_PATCH, "fp64_plain"),
                                (np.float32, hip.DRN_LAUNCH_BATCHED, "fp32_batched")):
        hg, ctx = setup(w, dtype)
        ctx.autotune(launch)
        s0 = ctx.core_sums(0)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.rk_step(w.dt, w.alpha, launch)
        ctx.sync()
        log("accuracy: %s %d steps in %.1f s" % (name, steps, time.perf_counter() - t0))
        s1 = ctx.core_sums(0)
        diag = ctx.diagnostics(0, w.dt)
        states = [ctx.download(p, 0) for p in range(len(w.patches))]
        ctx.close()
        r = {"mass_drift": float((s1[0] - s0[0]) / s0[0]), "energy_drift": float((s1[4] - s0[4]) / s0[4]),
             "nonfinite": diag["nonfinite"]}
        if planes:
            r["momentum_flux_planes"] = [plane_momentum_flux(w, states, gi) for gi in planes]
            r["momentum_flux_difference"] = r["momentum_flux_planes"][0] - r["momentum_flux_planes"][1]
        out[name] = r
        if dtype == np.float64:
            rho64 = [s[0].astype(np.float64) for s in states]
        else:
            scale = max(float(np.abs(a).max()) for a in rho64)
            out["rho_linf_fp32_vs_fp64"] = max(float(np.abs(s[0].astype(np.float64) - a).max())
                                               for s, a in zip(states, rho64)) / scale
        del states
    if planes:
        out["planes_global_i"] = list(planes)
        a, b = out["fp64_plain"], out["fp32_batched"]
        out["momentum_flux_rel_error"] = [abs(x - y) / abs(x) for x, y in zip(b["momentum_flux_planes"],
                                                                              a["momentum_flux_planes"])]
        out["momentum_flux_difference_rel_error"] = (abs(b["momentum_flux_difference"] - a["momentum_flux_difference"])
                                                     / abs(a["momentum_flux_difference"]))
    return out


def ctx_vstride(ctx, w):
    # cells per variable block: patches padded to 64 cells
    return sum((p.cells + 63) // 64 * 64 for p in w.patches)


def cpu_baseline(w, dtype, threads):
    from oracle import oracle as O

    og = O.OracleGrid(w, dtype, threads=threads)
    og.init_state()
    og.compute_dependencies()
    t0 = time.perf_counter()
    og.rk_step()
    t1 = time.perf_counter()
    cells = w.total_cells
    # the same grid on one host thread: one stage (CartesianIterator::compute
    # over every patch, no exchange), which keeps the sample near 5 s
    og.set_threads(1)
    og.copy_field(0, 1)
    t2 = time.perf_counter()
    og.compute(w.alpha[0] * w.dt)
    t3 = time.perf_counter()
    return dict(value=cells * len(w.alpha) / (t1 - t0) / 1e6, unit="Mcell-updates/s", cores=threads,
                kind="port", sample="1 RK step (3 stages + 6 exchanges) of the same 64-patch jet, %s oracle, "
                                    "OpenMP over patches, %.1f s" % (np.dtype(dtype).name, t1 - t0),
                single_thread=dict(value=cells / (t3 - t2) / 1e6, cores=1,
                                   sample="1 stage of the same grid, 1 thread, %.1f s" % (t3 - t2)),
                host_cpu=host_cpu_model(), host_logical_cpus=os.cpu_count())


def host_cpu_model():
    """the host CPU's model name as the kernel reports it (first processor)"""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=0, help="steps of the FP64 plain form (default: steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--blocks", type=int, default=4)
    ap.add_argument("--core", type=int, default=64)
    ap.add_argument("--tiling", type=int, default=None, help="fixed stage work-list choice (default: autotune)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="jet",
                    help="jet (the headline case) or one of the other configurations")
    ap.add_argument("--fp32-only", action="store_true", help="skip the FP64 per-patch form (profiling runs)")
    ap.add_argument("--accuracy", action="store_true",
                    help="after the measurement, run both forms for the configuration's step count and report FP32 "
                         "against FP64 (density Linf, mass / energy drift, level-set momentum flux)")
    ap.add_argument("--accuracy-steps", type=int, default=0, help="steps of the accuracy report (default: the "
                                                                   "configuration's)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (patches would be sharded across GPUs "
              "with one exchange per stage over xGMI); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from drnum_amd import configs, hip

    w = configs.jet(core=args.core, blocks=args.blocks) if args.workload == "jet" else WORKLOADS[args.workload][0]()
    log("workload %s: %d patches, %d stored cells, %d core cells" % (w.name, len(w.patches), w.total_cells,
                                                                      w.core_cells))
    f32 = run_form(w, np.float32, args.steps, args.warmup, hip.DRN_LAUNCH_BATCHED, "fp32-batched", args.tiling)
    f64 = f32pp = None
    if not args.fp32_only:
        f64 = run_form(w, np.float64, args.fp64_steps or args.steps, args.warmup, hip.DRN_LAUNCH_PER_PATCH,
                       "fp64-plain", args.tiling)
        # the batching gain in one precision: the same FP32 arithmetic, one launch per patch per stage
        f32pp = run_form(w, np.float32, args.steps, args.warmup, hip.DRN_LAUNCH_PER_PATCH, "fp32-per-patch",
                         args.tiling)
    cpu = None
    if not args.no_cpu and args.workload == "jet":
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(w, np.float32, threads)
        log("cpu baseline", cpu)

    # the reference's arithmetic: the same FP32 batched form with IEEE division
    # and square root (DRN_ARITH_IEEE), as the reference's CUDA build computes
    f32ieee = None
    if not args.fp32_only:
        f32ieee = run_form(w, np.float32, args.steps, args.warmup, hip.DRN_LAUNCH_BATCHED, "fp32-ieee-batched",
                           args.tiling, arith=hip.DRN_ARITH_IEEE)
    # the same jet with the other domain faces as the reference's JetFlux has
    # them: far-field functors on the lateral faces, a slip wall at the far end
    f32ref = None
    if args.workload == "jet" and not args.fp32_only:
        wref = configs.jet(core=args.core, blocks=args.blocks, faces="reference")
        f32ref = run_form(wref, np.float32, args.steps, args.warmup, hip.DRN_LAUNCH_BATCHED,
                          "fp32-batched-reference-faces", args.tiling)
    ach = f32["stage_gbs"]
    default_workload = args.workload == "jet" and (args.core, args.blocks) == (64, 4)
    # figures read from committed rocprofv3 summaries hold only for the kernel
    # sources they were taken of: a summary of another build is reported stale
    traffic_current = profile_current(TRAFFIC_TABLE)
    valu_current = profile_current(VALU_TABLE)
    traffic = measured_traffic(f32["tiling"]) if default_workload else None
    # compute ceiling of the same kernel: VALU instructions per launch (committed
    # profile of this workload) over the live launch time, against the probe
    valu = None
    insts = measured_valu_instructions(f32["tiling"]) if default_workload else None
    probe = valu_ceiling() if insts else None
    if insts and probe:
        rate = insts / probe["cus"] / (f32["stage_ms"] * 1e6)
        valu = {"bound": "valu", "achieved": round(rate, 3), "peak": probe["peak"], "unit": "wave-instr/CU/ns",
                "frac": round(rate / probe["peak"], 4), "instructions_per_launch": insts,
                "stale": not valu_current,
                "source": "SQ_INSTS_VALU from profiles/stage_valu_jet4x64_fp32.md (tiling %d); peak measured now by "
                          "drnum_amd/lib/valu_probe" % f32["tiling"]}
        log("valu", valu)
    out = {
        "metric": METRIC,
        "value": round(f32["mcups"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.workload][1] if (args.workload != "jet" or default_workload) else w.name,
            "patches": len(w.patches),
            "cells_stored": w.total_cells,
            "cells_core": w.core_cells,
            # the same timing over the unique core cells only (the value counts
            # every stored cell, seek layers included, as the reference advances them)
            "core_mcell_updates_per_s": round(f32["mcups"] * w.core_cells / w.total_cells, 2),
            "stages": len(w.alpha),
            "form": "fp32 batched (one stage launch over all patches)",
            "tiling": f32["tiling"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": traffic if traffic_current else None,
            "traffic_source": ("rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/stage_traffic_jet4x64_fp32.md, tiling %d"
                               % f32["tiling"]) if traffic else None,
            # a summary taken of other kernel sources: its figure, not used as traffic
            "traffic_stale": None if (traffic is None or traffic_current) else traffic,
            "kernel_source_sha256": kernel_source_hash(),
            "algorithmic_bytes_per_launch": round(f32["stage_bytes_per_launch"]),
        },
        "valu_roofline": valu,
        "fp64_plain": None if f64 is None else {
            "value": round(f64["mcups"], 2),
            "ms_per_step": round(f64["ms_per_step"], 4),
            "stage_kernel_gbs": round(f64["stage_gbs"], 1),
            "tiling": f64["tiling"],
            "fp32_batched_over_fp64_plain": round(f32["mcups"] / f64["mcups"], 3),
        },
        "fp32_ieee": None if f32ieee is None else {
            # the reference's arithmetic (IEEE division and square root) in the same batched form
            "value": round(f32ieee["mcups"], 2),
            "ms_per_step": round(f32ieee["ms_per_step"], 4),
            "stage_kernel_ms": round(f32ieee["stage_ms"], 4),
            "tiling": f32ieee["tiling"],
            "fast_over_ieee": round(f32["mcups"] / f32ieee["mcups"], 3),
            "ieee_over_fp64_plain": None if f64 is None else round(f32ieee["mcups"] / f64["mcups"], 3),
        },
        "reference_faces": None if f32ref is None else {
            # far-field lateral faces and a slip wall at the far end (JetFlux, jetflux.h:140-198)
            "value": round(f32ref["mcups"], 2),
            "ms_per_step": round(f32ref["ms_per_step"], 4),
            "stage_kernel_ms": round(f32ref["stage_ms"], 4),
            "tiling": f32ref["tiling"],
        },
        "graph": {
            "ms_per_step": round(f32["graph_ms_per_step"], 4),
            "value": round(f32["graph_mcups"], 2),
        },
        "fp32_per_patch": None if f32pp is None else {
            "value": round(f32pp["mcups"], 2),
            "ms_per_step": round(f32pp["ms_per_step"], 4),
            "tiling": f32pp["tiling"],
            "batched_over_per_patch": round(f32["mcups"] / f32pp["mcups"], 3),
        },
        "pcie_inclusive": {
            # never the value: the same K steps with one upload of the whole
            # state before them and one download after them
            "value": round(len(w.alpha) * w.total_cells * args.steps
                           / (f32["wall"] + f32["upload_s"] + f32["download_s"]) / 1e6, 2),
            "unit": "Mcell-updates/s",
            "upload_ms": round(1e3 * f32["upload_s"], 3),
            "download_ms": round(1e3 * f32["download_s"], 3),
            "state_bytes": f32["state_bytes"],
            "sample": "%d steps between one upload and one download of the FP32 state (pageable host buffers)"
                      % args.steps,
        },
        "stage_kernel_ms": round(f32["stage_ms"], 4),
        "exchange_ms": round(f32["exch_ms"], 4),
        "exchange_roofline": kernel_roofline(f32["exch_ms"], f32["exch_bytes"],
                                             measured_exchange_traffic() if default_workload and traffic_current
                                             else None),
        "exchange_fused_receivers": f32["fused"],
        "levelset_fill_roofline": kernel_roofline(f32["ls_ms"], f32["ls_bytes"]),
        "stream_copy_gbs": round(f32["copy_gbs"], 1),
        "cpu_baseline": cpu,
    }
    if args.accuracy:
        out["accuracy"] = accuracy_report(w, args.accuracy_steps or w.steps)
        log("accuracy", out["accuracy"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
