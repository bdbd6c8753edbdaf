"""Benchmark: structured TMOP block-Jacobi sweeps on one MI355X.

Default workload (the metric's configuration): the 5-block 2D O-grid — a
central 129 x 129 block bounded by arcs of the circle r = 0.5 inside four
129 x 65 ring blocks reaching the square [-2, 2]^2 (49 409 nodes, 48 897
free) — with 2 % of the free nodes displaced by 1.5 x their local edge
length (seed 2026). It is untangled with the delta-continuation schedule (20
sweeps per delta, omega 0.5), then smoothed by block-Jacobi TMOP sweeps
(omega 0.8, halo exchange every sweep). A step is one sweep over all free
DOFs; the metric is node-updates/s = free DOFs x steps / wall time, with the
state resident in HBM.

Also reported on that line: time to untangled + converged (relative energy
within 1e-6 of the pinned FP64 plain-method energy E*, min det > 0) for the
plain method, the plain method with the learned proposal (network trained
on held-out seeds first) and FAS V(4,4); the sweep kernel's roofline (FP32
flops per launch from the op-counted constants and kernel-counted line-search
trials / measured launch time); and the CPU oracle on a bounded sample of the
same workload, whose FP64 result is also the parity reference.

Other workloads (one JSON line each, not the default):
  --workload batched       64 independent multi-block grids (3-9 blocks, sides
                           33..513), one session, autotuned tiles, FAS V(4,4)
                           on 3 levels
  --workload cubed_sphere  7-block 65^3 cubed-sphere hex grid, FAS 3 levels,
                           plus the 33^3 Laplacian PCG

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--workload W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Op-counted flops per free-DOF update (oracle/flopcount.cpp over
# egg_amd/csrc/tmop_math.hpp; shape metric): patch metric pass + Newton
# solve, and one line-search trial.
F_METRIC = {2: 1244, 3: 17465}
F_TRIAL = {2: 272, 3: 3206}
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak
PEAK_HBM_GBS = 8000.0
SEED = 2026
CONV_RTOL = 1e-6
GRAPH_CHUNK = 256  # sweeps per captured hipGraph (5-8 % per sweep on the small grids)
# multi-sweep candidates (sweeps per launch, core tile rows, columns, kernel
# variant): the fine level's plain sweeps (k = 1: the one-sweep kernels) and
# the FAS coarse levels' smoothing calls; each tuner keeps the fastest
MS_FINE_CANDS = [(1, 0, 0, 0)] + [(k, ty, tx, v) for k in (2, 3, 4) for ty, tx in ((16, 8), (8, 16), (16, 16))
                                  for v in (5, 1, 2)]
MS_LEVEL_CANDS = [(k, ty, tx, v) for k in (2, 4, 8) for ty, tx in ((8, 8), (16, 8), (16, 16)) for v in (5, 1, 0)]


def _tune_multi(s, level_sweeps, max_levels=32, fine=True):
    """Multi-sweep tuners: the fine level's plain sweeps (40 sweeps per
    candidate, captured) and each FAS coarse level's smoothing call of
    level_sweeps sweeps. Returns what each picked against its one-sweep time."""
    out = {}
    if fine:
        ms = s.core.autotune_multi(40, 0.8, MS_FINE_CANDS)
        b = int(np.argmin(ms))
        out["fine"] = {"one_sweep_us_per_sweep": float(ms[0]) / 40 * 1e3,
                       "best_multi": list(MS_FINE_CANDS[1 + int(np.argmin(ms[1:]))]),
                       "best_multi_us_per_sweep": float(np.min(ms[1:])) / 40 * 1e3,
                       "picked": list(MS_FINE_CANDS[b])}
    lv = s.core.autotune_levels_multi(level_sweeps, 0.8, MS_LEVEL_CANDS, max_levels)
    levels = []
    for i in range(lv.shape[0]):
        info = s.core.multi_sweep_info(i + 1)
        levels.append({"one_sweep_us": float(lv[i, 0]) * 1e3, "best_us": float(np.min(np.where(lv[i] < 0, np.inf, lv[i]))) * 1e3,
                       "picked": [info["k"], info["ty"], info["tx"], info["kernel"]] if info["k"] else "one-sweep"})
    out["fas_levels"] = {"sweeps_per_call": level_sweeps, "levels": levels}
    return out
# HBM bytes per sweep launch by workload and tile variant, from two rocprofv3
# --pmc runs of this file (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch,
# median; profiles/sweep2d_ogrid129.md, batched64_fas.md,
# cubed_sphere65_fas_pcg.md, ogrid129_boundary_layer.md). A device copy of N bytes reads FETCH_SIZE =
# N/2 and WRITE_SIZE = N on this part, so FETCH is scaled by 2 and WRITE by 1.
PMC_KIB = {
    ("ogrid", 0): (1314.06, 548.44), ("ogrid", 1): (1466.69, 616.38), ("ogrid", 2): (1218.31, 517.69),
    ("ogrid", 3): (1072.38, 538.5), ("ogrid", 4): (1856.91, 617.75), ("ogrid", 5): (2058.19, 784.62),
    ("ogrid", 6): (1214.56, 517.69), ("ogrid", 7): (1306.81, 548.44), ("ogrid", 8): (1089.25, 549.5),
    ("ogrid", 9): (773.69, 502.5), ("ogrid", 10): (1085.0, 549.5), ("ogrid", 11): (918.38, 505.38),
    ("ogrid", 12): (1309.0, 548.06), ("ogrid", 13): (1310.0, 537.75),
    ("batched", 0): (590518.03, 277104.75), ("batched", 1): (761484.0, 325449.72),
    ("batched", 2): (526222.62, 252865.31), ("batched", 3): (508322.66, 276764.81),
    ("batched", 4): (1073595.0, 325607.78), ("batched", 5): (1284491.03, 422312.42),
    ("batched", 6): (525395.94, 252865.33), ("batched", 7): (589010.25, 277104.77),
    ("batched", 8): (509931.56, 277069.66), ("batched", 9): (376742.16, 252706.42),
    ("batched", 10): (508547.75, 277069.69), ("batched", 11): (435552.28, 252758.38),
    ("batched", 12): (588511.62, 276649.7), ("batched", 13): (588924.25, 276800.11),
    ("cubed_sphere", 0): (135327.31, 49889.66), ("cubed_sphere", 1): (185090.25, 69130.72),
    ("cubed_sphere", 2): (133390.06, 49889.95), ("cubed_sphere", 3): (140850.5, 51144.95),
    ("cubed_sphere", 4): (132029.03, 49050.84),
    ("ogrid_bl", 0): (12851.0, 613.5), ("ogrid_bl", 1): (12856.0, 613.28), ("ogrid_bl", 2): (12855.0, 613.34),
    ("ogrid_bl", 3): (12605.38, 555.81), ("ogrid_bl", 4): (12839.0, 613.16), ("ogrid_bl", 5): (12845.0, 613.12),
    ("ogrid_bl", 6): (12855.5, 613.31), ("ogrid_bl", 7): (12854.5, 613.38), ("ogrid_bl", 8): (12853.0, 613.12),
    ("ogrid_bl", 9): (12851.5, 617.44), ("ogrid_bl", 10): (12852.5, 613.0), ("ogrid_bl", 11): (12856.0, 613.09),
    ("ogrid_bl", 12): (12221.75, 522.84), ("ogrid_bl", 13): (12616.06, 555.88),
}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0


def _traffic(workload, variant):
    kib = PMC_KIB.get((workload, int(variant)))
    return None if kib is None else (kib[0] * FETCH_SCALE + kib[1] * WRITE_SCALE) * 1024.0


PMC_SWEEPS = 20  # sweeps per tile variant in a counter child run


def _pmc_child(args):
    """--pmc-child: the O-grid workload's plain sweeps for a rocprofv3 --pmc
    run of the parent: untangle, then PMC_SWEEPS one-launch sweeps with every
    tile variant in turn (the parent splits the shape-phase sweep dispatches
    into blocks of PMC_SWEEPS in variant order); prints the byte count of the
    store copy every first run makes (the counters' calibration)."""
    from egg_amd import grids
    from egg_amd.smoothing import cpp_untangle

    g = grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=SEED)
    ctx, bsc, s = _build(g)
    Xu, _, _, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    s.core.set_graph_chunk(0)
    nv = len(s.core.autotune(1, 0.8)[1])  # the variant count (and every variant's code warmed)
    for v in range(nv):
        s.core.set_tile_variant(v)
        s.set_X(Xu)
        s.run(PMC_SWEEPS, omega=0.8)
    plane = (int(s.core.info()["n_slots"]) + 63) // 64 * 64  # the store's plane stride (64-slot multiple)
    print(json.dumps({"variants": nv, "store_bytes": plane * 2 * 4}), flush=True)


def _live_traffic():
    """HBM bytes per sweep launch of every tile variant on the O-grid,
    measured in this run: two child runs of this file (--pmc-child) under
    rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE (one counter per run, each
    under a hard time limit), started before this process touches the GPU.
    Each counter is calibrated by its reading for the device copy of the
    store in the same run (N bytes, known). Returns ({variant: bytes},
    info) or (None, reason)."""
    import csv
    import glob
    import os
    import shutil
    import subprocess
    import tempfile

    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    reads = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="eggmi_pmc_")
        try:
            r = subprocess.run(["timeout", "-s", "KILL", "150", "rocprofv3", "--pmc", ctr, "--output-format", "csv",
                                "-d", d, "-o", "run", "--", sys.executable, str(ROOT / "bench.py"), "--pmc-child"],
                               capture_output=True, text=True, cwd="/tmp", env={**os.environ, "TMPDIR": "/tmp"})
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
            if r.returncode != 0 or not files or not lines:
                return None, f"rocprofv3 --pmc {ctr}: exit {r.returncode}, {len(files)} result files"
            child = json.loads(lines[-1])
            with open(files[0], newline="") as fh:
                rows = [x for x in csv.DictReader(fh) if x.get("Counter_Name", ctr) == ctr]
        finally:
            shutil.rmtree(d, ignore_errors=True)
        rows.sort(key=lambda x: int(x.get("Dispatch_Id", 0)))
        sweeps = [float(x["Counter_Value"]) for x in rows if "sweep_kernel<2, 0," in x["Kernel_Name"]
                  or "sweep_kernel_cpl<2, 0," in x["Kernel_Name"] or "sample_kernel<0," in x["Kernel_Name"]]
        copies = [float(x["Counter_Value"]) for x in rows if "copyBuffer" in x["Kernel_Name"]]
        nv = child["variants"]
        tail = sweeps[-nv * PMC_SWEEPS:]
        if len(tail) != nv * PMC_SWEEPS or not copies:
            return None, f"rocprofv3 --pmc {ctr}: {len(sweeps)} sweep dispatches, {len(copies)} copies"
        per_v = [float(np.median(tail[v * PMC_SWEEPS:(v + 1) * PMC_SWEEPS])) for v in range(nv)]
        scale = child["store_bytes"] / (max(copies) * 1024.0)
        reads[ctr] = (per_v, scale, max(copies))
    fetch, fs, fcopy = reads["FETCH_SIZE"]
    write, ws, wcopy = reads["WRITE_SIZE"]
    out = {v: fetch[v] * 1024.0 * fs + write[v] * 1024.0 * ws for v in range(len(fetch))}
    return out, {"fetch_kib": fetch, "write_kib": write, "fetch_scale": fs, "write_scale": ws,
                 "copy_fetch_kib": fcopy, "copy_write_kib": wcopy}
METRIC = "node-updates/s; time-to-untangled+converged, plain vs MLP-proposal, 1 GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _threads():
    n = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return n


def _build(g):
    from egg_amd.smoothing import (CppStructuredSweepSession, build_structured_context_from_block_maps,
                                   build_sweep_context)

    ctx = build_sweep_context(g.blocks, g.free, g.d)
    bsc = build_structured_context_from_block_maps(g.d, g.blocks)
    return ctx, bsc, CppStructuredSweepSession(ctx, bsc, g.X)


def _roofline(s, dim, n_free, steps, omega, k_avg, traffic=None, sweeps_per_launch=1.0):
    """Roofline of the sweep kernel: k_avg = its average launch time (HIP
    events on the session stream around the timed region, which holds only
    back-to-back sweep launches); flops per launch from the op-counted
    constants and the line-search trials the kernel counts in an untimed
    replay of the same trajectory (the sweep is deterministic)."""
    core = s.core
    X0 = s.get_X()
    core.set_counters(True)
    s.run(steps, omega=omega)
    core.set_counters(False)
    trials, updates, _ = (int(v) for v in core.last_counters(1)[0])
    s.set_X(X0)
    assert updates == n_free * steps, (updates, n_free * steps)
    tpu = trials / updates
    flops = n_free * sweeps_per_launch * (F_METRIC[dim] + tpu * F_TRIAL[dim])
    ach = flops / k_avg / 1e12
    kernel = f"sweep_kernel<{dim}D, shape>" if sweeps_per_launch == 1 else \
        f"multi_sweep_kernel<2D, shape> ({sweeps_per_launch:g} sweeps per launch)"
    return {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / PEAK_FP32_TFLOPS, "traffic": traffic, "kernel": kernel,
            "kernel_avg_us": k_avg * 1e6, "flops_per_launch": flops, "trials_per_update": tpu,
            "flops_per_update": f"{F_METRIC[dim]} + trials x {F_TRIAL[dim]}",
            "algorithmic_bytes_per_launch": n_free * sweeps_per_launch * 4 * dim * 2}


def _timed(s, n, omega):
    """n back-to-back sweeps: (energies, mindets, wall s, kernel avg s)."""
    X0 = s.get_X()
    t0 = time.perf_counter()
    e, m = s.run(n, omega=omega)
    wall = time.perf_counter() - t0
    ms, launches, _ = s.core.last_timing()
    return e, m, wall, ms * 1e-3 / launches, X0


def _converge(step, e_star, budget_s, rtol=CONV_RTOL):
    """Call step() (returns (energies, mindets, units)) until the last report
    is within rtol of e_star with min det > 0; returns (units, seconds)."""
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        e, m, u = step()
        done += u
        if abs(e[-1] - e_star) <= rtol * abs(e_star) and m[-1] > 0:
            return done, time.perf_counter() - t0, float(e[-1])
    return None, time.perf_counter() - t0, float(e[-1])


def _ttc_fixture(s, Xu, t_untangle, name, budget_s=60.0, fas_chunk=10, plain_chunk=1000, dim=2, ctx=None,
                 with_cpu=False):
    """Time to untangled + converged on a large workload against its pinned
    fixture (tests/golden/estar_<name>.json: the GPU iteration's fixed point,
    its energy recomputed in FP64, with the band a run counts as converged
    in), from the untangled state Xu: FAS V(4,4) on 3 levels (the
    workload's cycle) and the plain sweep, each within budget_s. With
    with_cpu, the FP64 energy / min det of each converged state are
    recomputed on the CPU (the oracle) against E* and the fixture's min det."""
    fix = ROOT / "tests" / "golden" / f"estar_{name}.json"
    if not fix.exists():
        return None
    f = json.loads(fix.read_text())
    e_star, rtol = f["E_star"], f.get("convergence_rtol", CONV_RTOL)
    s.set_X(Xu)
    out = {"E_star": e_star, "E_untangled": float(s.core.energy()[0]), "rtol": rtol, "untangle_s": t_untangle,
           "E_star_is": f["how"], "E_star_bias_bound": f.get("bias_bound"), "note": f.get("convergence_note")}

    def errors():
        if not with_cpu:
            return {}
        from oracle import oracle as O  # the oracle (CPU baseline and parity reference)

        E, md = O.energy_mindet(dim, "barrier", 0.0, ctx.energy_stencil, s.get_X().astype(np.float64))
        return {"rel_energy_err": abs(float(E) - e_star) / abs(e_star), "energy_fp64": float(E),
                "min_det_fp64": float(md), "min_det_err": abs(float(md) - f["min_det"])}

    for key, step in (("fas_v44_3_levels", lambda: (*s.run_fas(fas_chunk, nu_pre=4, nu_post=4, omega=0.8,
                                                                   max_levels=3), fas_chunk)),
                      ("plain", lambda: (*s.run(plain_chunk, omega=0.8, report_every=plain_chunk), plain_chunk))):
        s.set_X(Xu)
        n, t, E = _converge(step, e_star, budget_s, rtol)
        unit = "cycles" if key != "plain" else "sweeps"
        out[key] = {unit: n, "s": t + t_untangle if n else None, "E_end": E,
                    "within_budget_s": budget_s, **(errors() if n else {})}
        log(f"{name} {key}: {n} {unit}, {t:.3f}s to converged (E {E:.10g}, E* {e_star:.10g})")
    if out["fas_v44_3_levels"]["s"] and out["plain"]["s"]:
        out["speedup_fas_vs_plain"] = out["plain"]["s"] / out["fas_v44_3_levels"]["s"]
    return out


def _cpu_rates(run, units, n_sweeps):
    """CPU oracle timing in both precisions: per precision, one call of 5
    sweeps to warm the threads and caches, then the median of 3 timed calls
    of `n_sweeps` from the same start. run(prec, n) returns the oracle's
    output; `units` = DOF updates per sweep. Returns (rates, FP64 output,
    seconds per timed call)."""
    rates, secs, out64 = {}, {}, None
    for prec in ("f64", "f32"):
        run(prec, min(5, n_sweeps))
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            out = run(prec, n_sweeps)
            t.append(time.perf_counter() - t0)
        if prec == "f64":
            out64 = out
        secs[prec] = t
        rates[prec] = units * n_sweeps / float(np.median(t))
    return rates, out64, secs


def _cpu_entry(rates, threads, sample):
    best = max(rates, key=rates.get)
    return {"value": rates[best], "unit": "node-updates/s", "cores": threads, "kind": "port",
            "sample": f"{sample}, OpenMP C oracle {best.upper()} (f64 {rates['f64']:.4g}, f32 {rates['f32']:.4g} "
                      "node-updates/s; 5 warm-up sweeps, median of 3 timed runs)"}


def _error_triple(X, x_star, e_star, es, dim, extent):
    """FP64 errors of a converged GPU state X against the FP64 plain-method
    converged state: max-abs position error (against X*, when pinned),
    relative energy error (E of X recomputed in FP64 on the CPU, against
    E*) and min-det error (against min det of X*)."""
    from oracle import oracle as O  # the oracle (CPU baseline and parity reference)

    e, md = O.energy_mindet(dim, "barrier", 0.0, es, X)
    out = {"rel_energy_err": abs(e - e_star) / abs(e_star), "energy_fp64": e, "min_det_fp64": md}
    if x_star is not None:
        es_, ms_ = O.energy_mindet(dim, "barrier", 0.0, es, x_star)
        out.update(max_abs_pos_err=float(np.abs(X - x_star).max()),
                   max_abs_pos_err_rel_extent=float(np.abs(X - x_star).max() / extent),
                   min_det_err=abs(md - ms_))
    return out


def _config1(threads, n_sweeps=40, seed=1):
    """The CPU reference's own case: one 129 x 129 block on the unit square,
    interior perturbed by U(-0.3h, 0.3h), identity target shape metric,
    40 FP64 sweeps at omega 0.8 on the host cores, beside the same 40 sweeps
    on the GPU in FP32, with the FP64 error triple of the GPU result."""
    from egg_amd import grids
    from oracle import oracle as O  # the oracle: CPU baseline and parity reference

    g = grids.single_block_2d(129, seed=seed, amp=0.3)
    ctx, bsc, s = _build(g)
    patches = O.materialize_patches(ctx.wire, g.blocks, 2)
    rates, (Xo, eo, mo), secs = _cpu_rates(
        lambda prec, n: O.jacobi(2, "barrier", 0.0, 0.8, n, g.X, patches, ctx.energy_stencil, prec=prec,
                                 threads=threads), g.n_free, n_sweeps)
    s.run(n_sweeps, omega=0.8)
    s.set_X(g.X)
    t0 = time.perf_counter()
    e, m = s.run(n_sweeps, omega=0.8)
    t_gpu = time.perf_counter() - t0
    Xg = s.get_X()
    eg, mg = O.energy_mindet(2, "barrier", 0.0, ctx.energy_stencil, Xg)
    return {"workload": f"one 129x129 block on [0,1]^2, interior U(+-0.3h), seed {seed}, {n_sweeps} sweeps, omega 0.8",
            "free_dofs": g.n_free,
            "cpu": _cpu_entry(rates, threads, f"all {n_sweeps} sweeps"),
            "cpu_f64_s_per_run": secs["f64"],
            "gpu_node_updates_per_s": g.n_free * n_sweeps / t_gpu, "gpu_s": t_gpu,
            "parity": _parity(Xg, Xo, eg, eo[-1], mg, mo[-1], 1.0, sweeps=n_sweeps, energy_gpu_state=eg,
                              energy_oracle=float(eo[-1]),
                              energy_of_gpu_state="FP64 on the CPU (the energy nears 0 on this grid, so the "
                                                  "relative error is of a small number)")}


def _parity(Xg, Xo, eg, eo, mg, mo, extent, **extra):
    return {"max_abs_pos_err": float(np.abs(Xg - Xo).max()),
            "max_abs_pos_err_rel_extent": float(np.abs(Xg - Xo).max() / extent),
            "rel_energy_err": float(abs(eg - eo) / abs(eo)), "min_det_err": float(abs(mg - mo)),
            "oracle": "FP64 CPU oracle (oracle/egg_oracle.c)", **extra}


def _batched_sample(s, g, index, Xu, n_sweeps, target=300_000):
    """Parity and CPU baseline of the batched workload on one complete grid
    of the batch (the one closest to `target` free DOFs): the grids are
    independent, so the batch's sweeps restricted to that grid are that
    grid's own sweeps. The GPU runs the whole batch from Xu."""
    from egg_amd import grids
    from egg_amd.smoothing import build_sweep_context
    from oracle import oracle as O  # the oracle: CPU baseline and parity reference

    frees = [int(g.free[n0:n0 + nn].sum()) for (_, _, n0, nn) in index]
    gi = int(np.argmin([abs(f - target) for f in frees]))
    b0, nb, n0, nn = index[gi]
    sub = grids.Grid(2, [b - n0 for b in g.blocks[b0:b0 + nb]], Xu[n0:n0 + nn].copy(), g.free[n0:n0 + nn].copy())
    ctx = build_sweep_context(sub.blocks, sub.free, 2)
    patches = O.materialize_patches(ctx.wire, sub.blocks, 2)
    threads = _threads()
    rates, (Xo, eo, mo), _ = _cpu_rates(
        lambda prec, n: O.jacobi(2, "barrier", 0.0, 0.8, n, sub.X, patches, ctx.energy_stencil, prec=prec,
                                 threads=threads), frees[gi], n_sweeps)
    cpu = _cpu_entry(rates, threads, f"{n_sweeps} sweeps of grid {gi} of the batch ({nb} blocks, {frees[gi]} free "
                                     "DOFs) from the untangled state")
    X0 = s.get_X()
    s.set_X(Xu)
    s.run(n_sweeps, omega=0.8)
    Xg = s.get_X()[n0:n0 + nn]
    s.set_X(X0)
    eg, mg = O.energy_mindet(2, "barrier", 0.0, ctx.energy_stencil, Xg)  # the GPU state, FP64 on the CPU
    extent = float((sub.X.max(0) - sub.X.min(0)).max())
    par = _parity(Xg, Xo, eg, eo[-1], mg, mo[-1], extent, sweeps=n_sweeps, grid=gi, free_dofs=frees[gi],
                  energy_of_gpu_state="FP64 on the CPU over that grid's cells")
    return cpu, par


def _batched_proposal(s, X0, K, e_fas, t_fas):
    """The learned proposal in the batched workload: a network trained on
    pairs harvested from held-out batches (four grids each, seeds 11 and
    12, 10 snapshots 50 sweeps apart: ~10 M pairs) proposes in the fine
    smoothing of the same K FAS V(4,4) cycles; energy and time against the
    cycles without it."""
    from egg_amd import grids
    from egg_amd._cpp import cpp_core
    from egg_amd.smoothing import cpp_untangle

    net = cpp_core.ProposalMLP(2, 0, 12_000_000)
    t0 = time.perf_counter()
    for sd in (11, 12):
        gh, _ = grids.batched_2d(4, seed=sd)
        ctx_h, bsc_h, sh = _build(gh)
        cpp_untangle(ctx_h, bsc_h, gh.X, session=sh)
        sh.harvest(net, 10, every=50, horizon=2, omega=0.8)
        del sh
    t_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = net.train(4, batch=32768, lr=1e-3, seed=1)
    t_t = time.perf_counter() - t0
    s.set_proposal(net, mode="predict_correct", scale=0.5)
    s.set_X(X0)
    s.run_fas(1, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)  # warm-up (and capture)
    s.set_X(X0)
    t0 = time.perf_counter()
    e, m = s.run_fas(K, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
    t_p = time.perf_counter() - t0
    # the fine sweep with and without the proposal (both from X0, timed)
    s.set_X(X0)
    _, _, t_sp, _, _ = _timed(s, 10, 0.8)
    s.set_proposal(None)
    s.set_X(X0)
    _, _, t_s0, _, _ = _timed(s, 10, 0.8)
    return {"pairs": int(net.shape()["n_data"]), "harvest_s": t_h, "train_s": t_t,
            "loss_first_last": [float(hist[0]), float(hist[-1])], "cycles": K,
            "energy_after": float(e[-1]), "energy_after_without": float(e_fas), "final_min_det": float(m[-1]),
            "ms_per_cycle": t_p / K * 1e3, "ms_per_cycle_without": t_fas / K * 1e3,
            "ms_per_fine_sweep": t_sp / 10 * 1e3, "ms_per_fine_sweep_without": t_s0 / 10 * 1e3,
            "mode": "predict_correct", "scale": 0.5}


def _cubed_sample(s, g, ctx, Xu, every=40):
    """Parity and CPU baseline of the 3D workload on a DOF sample (every
    `every`-th DOF of the wire): one Jacobi sweep moves each sampled DOF
    exactly as the full sweep does (frozen neighbours), so one GPU sweep of
    the whole grid is compared on the sample; the energy and min det of the
    GPU state are recomputed in FP64 on the CPU over the whole grid."""
    from oracle import oracle as O  # the oracle: CPU baseline and parity reference

    nd = len(ctx.wire["groups"][0]["dof_idx"])
    sample = np.arange(0, nd, every)
    patches = O.materialize_patches(ctx.wire, g.blocks, 3, dofs=sample)
    threads = _threads()
    # an empty energy stencil: the sweep's own report would sum the whole grid
    no_report = {"num_samples": 0, "gc": np.zeros(0, np.int32), "W_inv": np.eye(3),
                 **{f"gn{k}": np.zeros(0, np.int32) for k in range(3)}, **{f"s{k}": np.zeros(0) for k in range(3)}}
    rates, (Xo, _, _), _ = _cpu_rates(
        lambda prec, n: O.jacobi(3, "barrier", 0.0, 0.8, n, Xu, patches, no_report, prec=prec, threads=threads),
        len(sample), 1)
    cpu = _cpu_entry(rates, threads, f"1 sweep of {len(sample)} DOFs (every {every}th) of the same grid from the "
                                     "untangled state")
    X0 = s.get_X()
    s.set_X(Xu)
    eg_gpu, mg_gpu = s.run(1, omega=0.8)
    Xg = s.get_X()
    s.set_X(X0)
    dofs = np.asarray(ctx.wire["groups"][0]["dof_idx"])[sample]
    ec, mc = O.energy_mindet(3, "barrier", 0.0, ctx.energy_stencil, Xg)
    extent = float((g.X.max(0) - g.X.min(0)).max())
    par = _parity(Xg[dofs], Xo[dofs], float(eg_gpu[-1]), ec, float(mg_gpu[-1]), mc, extent, sweeps=1,
                  sampled_dofs=int(len(sample)),
                  energy="GPU FP64 reduction vs the FP64 CPU recomputation of the same state")
    return cpu, par


def _cubed_proposal(s, Xu, n_plain=400):
    """The 3D learned proposal (85 -> 128 -> 128 -> 3) on the cubed sphere:
    trained on the GPU on pairs harvested from plain runs on a held-out
    tangle seed (6 snapshots 25 sweeps apart of its 1.75 M interior DOFs:
    10.5 M pairs); then, from the untangled state, sweeps and time for the
    plain method and for the proposal in every k-th sweep (replace mode) to
    reach the energy the plain method has after `n_plain` sweeps (the time
    to the pinned converged energy is measured by _ttc_fixture)."""
    from egg_amd import grids
    from egg_amd._cpp import cpp_core
    from egg_amd.smoothing import cpp_untangle

    net = cpp_core.ProposalMLP(3, 0, 11_000_000)
    t0 = time.perf_counter()
    gh = grids.tangle(grids.cubed_sphere_3d(65), frac=0.01, factor=1.5, seed=11)
    ctx_h, bsc_h, sh = _build(gh)
    cpp_untangle(ctx_h, bsc_h, gh.X, session=sh)
    sh.harvest(net, 6, every=25, horizon=2, omega=0.8)
    del sh
    t_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = net.train(10, batch=4096, lr=1e-3, seed=1)
    t_t = time.perf_counter() - t0
    X0 = s.get_X()
    s.set_X(Xu)
    e, _ = s.run(n_plain, omega=0.8, report_every=n_plain)
    target = float(e[-1])

    def to_target(budget=30.0):
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            e, m = s.run(16, omega=0.8, report_every=16)
            done += 16
            if e[-1] <= target and m[-1] > 0:
                return done, time.perf_counter() - t0
        return None, time.perf_counter() - t0

    runs = {}
    for every in (None, 4, 8, 16):
        s.set_X(Xu)
        if every:
            s.set_proposal(net, mode="replace", scale=1.0, every=every)
        s.run(16, omega=0.8)  # warm the kernels
        s.set_X(Xu)
        n, t = to_target()
        runs["plain" if every is None else f"every_{every}"] = {"sweeps": n, "s": t}
        log(f"3D proposal {every}: {n} sweeps, {t:.3f}s to E {target:.8g}")
        s.set_proposal(None)
    s.set_X(X0)
    return {"pairs": int(net.shape()["n_data"]), "harvest_s": t_h, "train_s": t_t,
            "loss_first_last": [float(hist[0]), float(hist[-1])], "held_out_seed": 11,
            "target": f"E after {n_plain} plain sweeps from the untangled state", "target_energy": target,
            "mode": "replace", "scale": 1.0, "to_target": runs}


def _train_proposal(make_grid, seeds, horizon=2, snaps=42, every=50, epochs=10):
    """Harvest (features, displacement over `horizon` plain sweeps) pairs of
    every interior DOF from plain runs on held-out tangle seeds (5 seeds x
    42 snapshots x 48.1 k DOFs = 10.1 M pairs) and train on the GPU."""
    from egg_amd._cpp import cpp_core

    net = cpp_core.ProposalMLP(2, 0, 10_500_000)
    t0 = time.perf_counter()
    from egg_amd.smoothing import cpp_untangle

    for sd in seeds:
        g = make_grid(sd)
        ctx, bsc, sh = _build(g)
        cpp_untangle(ctx, bsc, g.X, session=sh)
        sh.harvest(net, snaps, every=every, horizon=horizon, omega=0.8)
    t_h = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = net.train(epochs, batch=4096, lr=1e-3, seed=1)
    return net, {"pairs": int(net.shape()["n_data"]), "harvest_s": t_h, "train_s": time.perf_counter() - t0,
                 "loss_first_last": [float(hist[0]), float(hist[-1])], "held_out_seeds": list(seeds),
                 "horizon": horizon}


def bench_ogrid(args):
    from egg_amd import grids
    from egg_amd.smoothing import cpp_untangle

    make = lambda sd: grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=sd)  # noqa: E731
    t0 = time.perf_counter()
    g = make(SEED)
    ctx, bsc, s = _build(g)
    n_free = g.n_free
    log(f"setup {time.perf_counter() - t0:.1f}s  nodes={g.n_nodes} free={n_free} md0={grids.min_det(g):.3e}")

    t0 = time.perf_counter()
    Xu, md_u, outer, delta = cpp_untangle(ctx, bsc, g.X, session=s)
    t_untangle = time.perf_counter() - t0
    log(f"untangle {t_untangle:.3f}s outer={outer} min_det={md_u:.3e}")

    tune_best, tune_ms = s.core.autotune(20, 0.8)
    log(f"autotune: variant {tune_best}, ms per 20 sweeps {np.round(tune_ms, 4).tolist()}")
    lv_best, lv_ms = s.core.autotune_levels(16, 0.8)  # FAS coarse levels, from the untangled state
    log(f"level autotune: variants {lv_best}")
    multi = _tune_multi(s, 8)
    log(f"multi-sweep tuners: {multi}")
    if args.force_multi:  # time (and profile) a given fine-level multi-sweep setting instead of the tuner's
        s.core.set_multi_sweep(0, *args.force_multi)
        multi["forced_fine"] = list(args.force_multi)
    s.core.set_graph_chunk(args.graph_chunk)  # replay sweeps as captured hipGraph chunks
    s.run(args.warmup, omega=0.8)
    e, m, t_run, k_avg, X_start = _timed(s, args.steps, 0.8)
    value = n_free * args.steps / t_run
    spl = args.steps / s.core.last_timing()[1]  # sweeps per launch
    X_end = s.get_X()
    s.set_X(X_start)
    live, live_info = getattr(args, "live_traffic", None) or (None, "not run (--no-pmc)")
    if spl == 1 and live:
        traffic, source = live[int(tune_best)], {"source": "two rocprofv3 --pmc child runs of this bench "
                                                 "(FETCH_SIZE, WRITE_SIZE), calibrated by the store copy",
                                                 **live_info}
    else:
        traffic = _traffic("ogrid", tune_best) if spl == 1 else None
        source = {"source": "committed PMC table (profiles/sweep2d_ogrid129.md)", "live": str(live_info)}
    roof = _roofline(s, 2, n_free, args.steps, 0.8, k_avg, traffic, spl)
    roof["traffic_source"] = source
    s.set_X(X_end)
    log(f"timed {args.steps} sweeps {t_run * 1e3:.2f} ms -> {value:.4g} node-updates/s  E={e[-1]:.9g}")
    log(f"sweep kernel {roof['kernel_avg_us']:.2f} us/launch, {roof['trials_per_update']:.3f} trials/update "
        f"-> {roof['achieved']:.3f} TFLOP/s ({100 * roof['frac']:.1f}% of FP32 peak)")

    # time to untangled + converged, from the same tangled start
    ttc = None
    extent = float((g.X.max(0) - g.X.min(0)).max())
    fix = ROOT / "tests" / "golden" / "estar_ogrid129.json"
    xfix = ROOT / "tests" / "golden" / "xstar_ogrid129.npy"
    if fix.exists() and not args.no_converge:
        e_star = json.loads(fix.read_text())["E_star"]
        x_star = np.load(xfix, allow_pickle=False) if xfix.exists() else None
        ttc = {"E_star": e_star, "rtol": CONV_RTOL, "untangle_s": t_untangle,
               "errors": "each result's FP64 error triple: max-abs position vs the FP64 converged state X*, "
                         "relative energy (FP64 recomputation on the CPU) vs E*, min det vs min det of X*"}

        def errors(n):
            return _error_triple(s.get_X(), x_star, e_star, ctx.energy_stencil, 2, extent) if n else None

        def plain_step(chunk=1000):
            e, m = s.run(chunk, omega=0.8, report_every=chunk)
            return e, m, chunk

        s.set_X(Xu)
        n, t, E = _converge(plain_step, e_star, 60.0)
        ttc["plain"] = {"sweeps": n, "s": t + t_untangle if n else None, "E_end": E, **(errors(n) or {})}
        log(f"plain: {n} sweeps, {t:.3f}s to converge")

        # FAS: V(4,4) on 3 levels (the batched workload's cycle), on every
        # level, and V(8,8) on 4 levels with 16 coarsest sweeps at omega 0.8
        # and at 0.95 (the fastest of a sweep over nu, levels, coarsest sweeps
        # and omega on this grid; from omega 1.0 on the cycle no longer
        # reaches 1e-6 of E*)
        for name, nu, lev, nuc, om in (("fas_v44_3_levels", 4, 3, 32, 0.8), ("fas_v44_all_levels", 4, 32, 32, 0.8),
                                       ("fas_v88_4_levels_coarse16", 8, 4, 16, 0.8),
                                       ("fas_v88_4_levels_coarse16_omega095", 8, 4, 16, 0.95)):
            s.set_X(Xu)

            def fas_step(chunk=5, nu=nu, lev=lev, nuc=nuc, om=om):
                e, m = s.run_fas(chunk, nu_pre=nu, nu_post=nu, nu_coarse=nuc, omega=om, max_levels=lev)
                return e, m, chunk

            n, t, E = _converge(fas_step, e_star, 60.0)
            ttc[name] = {"cycles": n, "s": t + t_untangle if n else None, "E_end": E, **(errors(n) or {})}
            log(f"FAS {name}: {n} cycles, {t:.3f}s to converge")

        net, info = _train_proposal(make, (11, 12, 13, 14, 15))
        s.set_X(Xu)
        s.set_proposal(net, mode="predict_correct", scale=0.5)  # (selects a variant with a proposal kernel)
        n, t, E = _converge(plain_step, e_star, 60.0)
        ttc["mlp_proposal"] = {"sweeps": n, "s": t + t_untangle if n else None, "E_end": E,
                               "mode": "predict_correct", "scale": 0.5, **info, **(errors(n) or {})}
        log(f"MLP proposal: {n} sweeps, {t:.3f}s to converge ({info})")
        # the proposal in every 16th sweep only (replace mode), the plain
        # kernel of the tuned variant in between
        s.set_proposal(None)
        s.core.set_tile_variant(int(tune_best))
        s.set_X(Xu)
        s.set_proposal(net, mode="replace", scale=1.0, every=16)
        n, t, E = _converge(plain_step, e_star, 60.0)
        ttc["mlp_proposal_every_16"] = {"sweeps": n, "s": t + t_untangle if n else None, "E_end": E,
                                        "mode": "replace", "scale": 1.0, "every": 16, **(errors(n) or {})}
        log(f"MLP proposal every 16th sweep: {n} sweeps, {t:.3f}s to converge")
        # the proposal in the fine-level smoothing of the fastest FAS cycle:
        # every sweep, then the first sweep of each smoothing call
        for every, mode, scale in ((1, "predict_correct", 0.5), (2, "replace", 1.0)):
            s.set_X(Xu)
            s.set_proposal(net, mode=mode, scale=scale, every=every)
            n, t, E = _converge(
                lambda: (*s.run_fas(5, nu_pre=8, nu_post=8, nu_coarse=16, omega=0.8, max_levels=4), 5), e_star, 60.0)
            key = "fas_v88_4_levels_coarse16_mlp_proposal" + ("" if every == 1 else "_first_sweep")
            ttc[key] = {"cycles": n, "s": t + t_untangle if n else None, "E_end": E, "mode": mode, "scale": scale,
                        "proposes_in": "every fine sweep" if every == 1 else "the first sweep of each fine smoothing",
                        **(errors(n) or {})}
            log(f"FAS V(8,8) + MLP proposal ({key}): {n} cycles, {t:.3f}s to converge")
        s.set_proposal(None)
        s.core.set_tile_variant(int(tune_best))

    cpu = parity = config1 = None
    if not args.no_cpu:
        from oracle import oracle as O  # the oracle: CPU baseline and parity reference

        threads = _threads()
        n_cpu = args.cpu_sweeps
        patches = O.materialize_patches(ctx.wire, g.blocks, 2)
        rates, (Xo, eo, mo), _ = _cpu_rates(
            lambda prec, n: O.jacobi(2, "barrier", 0.0, 0.8, n, Xu, patches, ctx.energy_stencil, prec=prec,
                                     threads=threads), n_free, n_cpu)
        cpu = _cpu_entry(rates, threads, f"{n_cpu} sweeps of the same O-grid from the untangled state")
        log(f"cpu {cpu}")
        s.set_X(Xu)
        ep, mp = s.run(n_cpu, omega=0.8)
        Xg = s.get_X()
        eg, mg = O.energy_mindet(2, "barrier", 0.0, ctx.energy_stencil, Xg)
        parity = _parity(Xg, Xo, eg, eo[-1], mg, mo[-1], extent, sweeps=n_cpu,
                         energy_of_gpu_state="FP64 on the CPU")
        log(f"parity {parity}")
        config1 = _config1(threads)
        log(f"config 1 {config1}")

    return {
        "metric": METRIC, "value": value, "unit": "node-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_run / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D 5-block O-grid (129x129 disc + 4x 129x65 rings, r=0.5 in [-2,2]^2), "
                               "2% tangled by 1.5h, untangle then block-Jacobi TMOP sweeps, omega 0.8",
                   "free_dofs": n_free, "seed": SEED, "precision": "FP32 sweeps, FP64 reductions"},
        "multi_sweep_tuners": multi,
        "autotune": {"best_variant": int(tune_best), "ms_per_20_sweeps": [float(v) for v in tune_ms],
                     "fas_level_variants": [int(v) for v in lv_best],
                     "fas_level_ms_per_16_sweeps": np.round(lv_ms, 4).tolist()},
        "graph_chunk": args.graph_chunk,
        "time_to_untangled_s": t_untangle, "untangle_outer": outer,
        "time_to_untangled_converged": ttc,
        "final_energy": float(e[-1]), "final_min_det": float(m[-1]),
        "parity": parity, "roofline": roof, "cpu_baseline": cpu, "config1_cpu_reference_case": config1,
    }


def _fas_levels_counts(s, nu_pre, nu_post, max_levels):
    """DOF updates per level of one counted V-cycle (replayed from, and
    restored to, the current state)."""
    X0 = s.get_X()
    s.core.set_counters(True)
    s.run_fas(1, nu_pre=nu_pre, nu_post=nu_post, omega=0.8, max_levels=max_levels)
    s.core.set_counters(False)
    c = s.core.last_counters(max_levels)
    s.set_X(X0)
    return [int(v) for v in c[:, 1]]


def bench_batched(args):
    from egg_amd import grids
    from egg_amd.smoothing import cpp_untangle

    t0 = time.perf_counter()
    g, index = grids.batched_2d(64, seed=SEED)
    ctx, bsc, s = _build(g)
    n_free = g.n_free
    log(f"setup {time.perf_counter() - t0:.1f}s grids=64 blocks={len(g.blocks)} nodes={g.n_nodes} free={n_free}")
    t0 = time.perf_counter()
    Xu, md_u, outer, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    t_untangle = time.perf_counter() - t0
    log(f"untangle {t_untangle:.3f}s outer={outer} min_det={md_u:.3e}")
    t0 = time.perf_counter()
    best, ms = s.core.autotune(5, 0.8)
    lv_best, lv_ms = s.core.autotune_levels(4, 0.8, 3)
    log(f"autotune {time.perf_counter() - t0:.2f}s best={best} ms/5 sweeps={ms} levels {lv_best}")
    _, _, t_plain, k_avg, X_start = _timed(s, 20, 0.8)
    s.set_X(X_start)
    roof = _roofline(s, 2, n_free, 20, 0.8, k_avg, _traffic("batched", best))
    counts = _fas_levels_counts(s, 4, 4, 3)
    K = max(1, args.steps // 100)
    s.core.set_graph_chunk(args.graph_chunk)  # FAS cycles replayed as captured graphs
    X_fas = s.get_X()
    s.run_fas(1, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)  # warm-up (and capture)
    s.set_X(X_fas)
    t0 = time.perf_counter()
    e, m = s.run_fas(K, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
    t_fas = time.perf_counter() - t0
    fine = counts[0] * K / t_fas
    allv = sum(counts) * K / t_fas
    log(f"FAS V(4,4) x{K}: {t_fas * 1e3:.1f} ms, fine {fine:.4g} / all-level {allv:.4g} node-updates/s, "
        f"E {e[0]:.6g} -> {e[-1]:.6g}")
    proposal = ttc = None
    if not args.no_converge:
        proposal = _batched_proposal(s, X_fas, K, e[-1], t_fas)
        log(f"learned proposal {proposal}")
        s.set_proposal(None)
        ttc = _ttc_fixture(s, Xu, t_untangle, "batched64", fas_chunk=10, plain_chunk=1000, ctx=ctx,
                           with_cpu=not args.no_cpu)
    cpu = parity = None
    if not args.no_cpu:
        cpu, parity = _batched_sample(s, g, index, Xu, 50)
        log(f"cpu {cpu}")
        log(f"parity {parity}")
    return {
        "metric": METRIC, "value": fine, "unit": "node-updates/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": t_fas / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched: 64 independent 2D strips of 3-9 blocks, sides from {33,41,..,513}, "
                               "2% tangled by 1.5h; one session; FAS V(4,4) on 3 levels, autotuned tiles",
                   "free_dofs": n_free, "blocks": len(g.blocks), "seed": SEED},
        "value_is": "fine-level node-updates/s of FAS V(4,4) cycles",
        "all_level_node_updates_per_s": allv, "updates_per_cycle_by_level": counts,
        "plain_sweep_node_updates_per_s": n_free * 20 / t_plain,
        "autotune": {"best_variant": int(best), "ms_per_variant": [float(v) for v in ms],
                     "fas_level_variants": [int(v) for v in lv_best]},
        "time_to_untangled_s": t_untangle, "untangle_outer": outer,
        "energy_first_last": [float(e[0]), float(e[-1])], "final_min_det": float(m[-1]),
        "learned_proposal": proposal, "time_to_untangled_converged": ttc,
        "roofline": roof, "cpu_baseline": cpu, "parity": parity,
    }


def bench_cubed(args):
    from egg_amd import grids, pcg
    from egg_amd._cpp import cpp_core
    from egg_amd.smoothing import cpp_untangle

    t0 = time.perf_counter()
    g = grids.tangle(grids.cubed_sphere_3d(65), frac=0.01, factor=1.5, seed=SEED)
    ctx, bsc, s = _build(g)
    n_free = g.n_free
    log(f"setup {time.perf_counter() - t0:.1f}s nodes={g.n_nodes} free={n_free} md0={grids.min_det(g):.3e}")
    t0 = time.perf_counter()
    Xu, md_u, outer, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    t_untangle = time.perf_counter() - t0
    log(f"untangle {t_untangle:.3f}s outer={outer} min_det={md_u:.3e}")
    t0 = time.perf_counter()
    best, ms = s.core.autotune(3, 0.8)
    lv_best, lv_ms = s.core.autotune_levels(3, 0.8, 3)
    log(f"autotune {time.perf_counter() - t0:.2f}s best={best} ms/3 sweeps={ms} levels {lv_best}")
    n3 = args.steps // 20 or 1
    e, m, t_plain, k_avg, X_start = _timed(s, n3, 0.8)
    value = n_free * n3 / t_plain
    X_end = s.get_X()
    s.set_X(X_start)
    roof = _roofline(s, 3, n_free, n3, 0.8, k_avg, _traffic("cubed_sphere", best))
    s.set_X(X_end)
    counts = _fas_levels_counts(s, 4, 4, 3)
    K = 5
    t0 = time.perf_counter()
    ef, mf = s.run_fas(K, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
    t_fas = time.perf_counter() - t0
    log(f"plain {value:.4g} node-updates/s; FAS x{K} {t_fas * 1e3:.1f} ms E {ef[0]:.6g}->{ef[-1]:.6g}")
    rp, col, val = pcg.lattice_laplacian(33, 3)
    b = np.random.default_rng(SEED).standard_normal(33**3)
    pcg_runs = {}
    for op in ("csr", "lattice"):
        for rec in ("standard", "one_reduction"):
            best_ms = None
            for _ in range(3):  # the first solve of a kind pays its kernels' first launch
                if op == "csr":
                    x, iters, relres, status = pcg.pcg_solve(rp, col, val, b, rtol=1e-8, max_iter=1000,
                                                             recurrence=rec)
                else:
                    x, iters, relres, status = pcg.pcg_solve_lattice(33, 3, b, rtol=1e-8, max_iter=1000,
                                                                     recurrence=rec)
                solve_ms = cpp_core.last_pcg_ms()
                best_ms = solve_ms if best_ms is None else min(best_ms, solve_ms)
            pcg_runs[f"{op}_{rec}"] = {"iterations": int(iters), "relres": float(relres), "status": int(status),
                                      "solve_ms": best_ms, "us_per_iteration": best_ms * 1e3 / max(iters, 1)}
            log(f"PCG 33^3 {op} {rec}: {iters} iterations, relres {relres:.2e}, {best_ms:.3f} ms")
    proposal = None if args.no_converge else _cubed_proposal(s, Xu)
    ttc = None if args.no_converge else _ttc_fixture(s, Xu, t_untangle, "cubed_sphere65", fas_chunk=10,
                                                     plain_chunk=500, dim=3, ctx=ctx, with_cpu=not args.no_cpu)
    cpu = parity = None
    if not args.no_cpu:
        cpu, parity = _cubed_sample(s, g, ctx, Xu)
        log(f"cpu {cpu}")
        log(f"parity {parity}")
    return {
        "metric": METRIC, "value": value, "unit": "node-updates/s", "n_gpus": 1, "steps": n3,
        "warmup": 0, "ms_per_step": t_plain / n3 * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3D 7-block cubed sphere (65^3 blocks, r=1 in [-2,2]^3), 1% tangled by 1.5h",
                   "free_dofs": n_free, "seed": SEED},
        "fas_v44_3_levels": {"cycles": K, "ms_per_cycle": t_fas / K * 1e3, "updates_per_cycle_by_level": counts,
                             "energy_first_last": [float(ef[0]), float(ef[-1])]},
        "pcg_33cubed": pcg_runs,
        "learned_proposal_3d": proposal, "time_to_untangled_converged": ttc,
        "autotune": {"best_variant": int(best), "ms_per_variant": [float(v) for v in ms],
                     "fas_level_variants": [int(v) for v in lv_best]},
        "time_to_untangled_s": t_untangle, "untangle_outer": outer, "roofline": roof, "cpu_baseline": cpu,
        "parity": parity,
    }


def bench_ogrid_bl(args):
    """The config-2 O-grid (129/65, 2 % tangled) under a boundary-layer
    target: the ring blocks ask for wall-normal spacing growing from 0.3 of
    the tangential spacing at the circle (reference BoundaryLayerTarget), so
    every sample has its own anisotropic W^-1 and every DOF takes the
    stored-stencil path (no synthesized interior). Untangle, autotune, timed
    plain sweeps; roofline of the stored path (HBM: the wire's per-occurrence
    tables plus X); CPU oracle baseline and FP64 parity on the same sweeps."""
    from egg_amd import grids
    from egg_amd.smoothing import (CppStructuredSweepSession, build_structured_context_from_block_maps,
                                   build_sweep_context, cpp_untangle)

    t0 = time.perf_counter()
    g = grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=SEED)
    W = grids.ogrid_boundary_layer_w_inv(g, 129)
    ctx = build_sweep_context(g.blocks, g.free, 2, w_inv=W)
    bsc = build_structured_context_from_block_maps(2, g.blocks)
    # untangle under the isotropic target, as the reference's cpp_untangle
    # expects (egg/smoothing/cpp_backend.py:773-848), then smooth under the
    # boundary-layer target from the untangled state
    ctx_iso, _, s_iso = _build(g)
    Xu, md_u, outer, _ = cpp_untangle(ctx_iso, bsc, g.X, session=s_iso)
    t_untangle = time.perf_counter() - t0
    log(f"untangle (isotropic target) {t_untangle:.3f}s outer={outer} min_det={md_u:.3e}")
    s = CppStructuredSweepSession(ctx, bsc, Xu)
    info = s.core.info()
    n_free = g.n_free
    occ = int(np.asarray(ctx.wire["groups"][0]["P_of"]).sum())
    log(f"setup free={n_free} stored={info['n_stored']} synth={info['n_synth']} occurrences={occ}")
    best, ms = s.core.autotune(20, 0.8)
    log(f"autotune: variant {best}, ms per 20 sweeps {np.round(ms, 4).tolist()}")
    s.core.set_graph_chunk(args.graph_chunk)
    s.run(args.warmup, omega=0.8)
    e, m, t_run, k_avg, _ = _timed(s, args.steps, 0.8)
    value = n_free * args.steps / t_run
    # algorithmic bytes per sweep: each DOF's occurrence tables (corner 4 B,
    # neighbours 8 B, role/sign byte, W^-1 row 16 B), its slot index, and X
    # read + written (the neighbour reads hit the caches)
    nbytes = occ * (4 + 8 + 1 + 16) + n_free * (4 + 16)
    ach = nbytes / k_avg / 1e9
    kib = PMC_KIB.get(("ogrid_bl", int(best)))
    roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
            "traffic": None if kib is None else (kib[0] * FETCH_SCALE + kib[1] * WRITE_SCALE) * 1024.0,
            "kernel": "sweep_kernel<2D, shape> (stored-stencil workgroups only)", "kernel_avg_us": k_avg * 1e6,
            "algorithmic_bytes_per_launch": nbytes}
    log(f"timed {args.steps} sweeps {t_run * 1e3:.2f} ms -> {value:.4g} node-updates/s; {k_avg * 1e6:.2f} us/launch, "
        f"{ach:.0f} GB/s algorithmic")
    cpu = parity = None
    if not args.no_cpu:
        from oracle import oracle as O  # the oracle: CPU baseline and parity reference

        threads = _threads()
        n_cpu = 50
        patches = O.materialize_patches(ctx.wire, g.blocks, 2)
        rates, (Xo, eo, mo), _ = _cpu_rates(
            lambda prec, n: O.jacobi(2, "barrier", 0.0, 0.8, n, Xu, patches, ctx.energy_stencil, prec=prec,
                                     threads=threads), n_free, n_cpu)
        cpu = _cpu_entry(rates, threads, f"{n_cpu} sweeps of the same grid and target from the untangled state")
        s.set_X(Xu)
        ep, mp = s.run(n_cpu, omega=0.8)
        parity = _parity(s.get_X(), Xo, ep[-1], eo[-1], mp[-1], mo[-1], float((g.X.max(0) - g.X.min(0)).max()),
                         sweeps=n_cpu)
        log(f"cpu {cpu}\nparity {parity}")
    return {
        "metric": METRIC, "value": value, "unit": "node-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_run / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D 5-block O-grid 129/65, 2% tangled by 1.5h, boundary-layer target on the rings "
                               "(first layer 0.3 of the tangential spacing, growth 1.2, 8 layers): all DOFs "
                               "on stored stencils, omega 0.8",
                   "free_dofs": n_free, "stored_occurrences": occ, "seed": SEED},
        "autotune": {"best_variant": int(best), "ms_per_20_sweeps": [float(v) for v in ms]},
        "graph_chunk": args.graph_chunk, "time_to_untangled_s": t_untangle,
        "untangle": "isotropic target (config-2 session), then this target",
        "final_energy": float(e[-1]), "final_min_det": float(m[-1]),
        "parity": parity, "roofline": roof, "cpu_baseline": cpu,
    }


def ogrid_wall_samples(g, weight=0.5, eps=1e-6):
    """Directional samples on the O-grid's ring blocks (axis 1 index 0 on the
    circle): a stem from each wall node to the next ring node with the wall
    tangent as through axis, and the first ring row as parallel segments with
    the radial direction at their midpoint as normal."""
    from egg_amd.smoothing.directional import KIND_PARALLEL, KIND_STEM, DirectionalSamples

    rows = []
    for ids in g.blocks[1:]:
        ids = np.asarray(ids)
        for i in range(ids.shape[0]):
            c0, n1 = int(ids[i, 0]), int(ids[i, 1])
            t = np.array([-g.X[c0, 1], g.X[c0, 0]])
            rows.append((n1, c0, KIND_STEM, t / np.linalg.norm(t)))
            if i + 1 < ids.shape[0]:
                a, b = int(ids[i, 1]), int(ids[i + 1, 1])
                mid = 0.5 * (g.X[a] + g.X[b])
                rows.append((a, b, KIND_PARALLEL, mid / np.linalg.norm(mid)))
    nodes = np.full((len(rows), 4), -1, np.int32)
    nodes[:, :2] = [r[:2] for r in rows]
    ref = np.array([r[3] for r in rows])
    return DirectionalSamples(nodes, np.array([r[2] for r in rows], np.int32), ref, np.zeros_like(ref),
                              np.full(len(rows), float(weight)), eps)


def bench_ogrid_terms(args):
    """The config-2 O-grid (129/65, 2 % tangled) with the soft terms on top of
    the shape metric: block-interface C2 curvature windows across every ring
    and core interface, and directional stems and parallel segments at the
    circle wall. The DOFs carrying windows run in the special stored-DOF
    launch beside the sweep kernel, and the directional whole-sample table
    joins every energy report. Untangle (plain), timed sweeps, CPU oracle
    baseline and FP64 parity on the same sweeps."""
    from egg_amd import grids
    from egg_amd.smoothing import (CppStructuredSweepSession, build_structured_context_from_block_maps,
                                   build_sweep_context, cpp_untangle)

    t0 = time.perf_counter()
    g = grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=SEED)
    curv = grids.interface_c2_windows(g, 0.5)
    smp = ogrid_wall_samples(g)
    ctx = build_sweep_context(g.blocks, g.free, 2, curvature=curv, directional=smp)
    bsc = build_structured_context_from_block_maps(2, g.blocks)
    ctx_iso, _, s_iso = _build(g)
    Xu, md_u, outer, _ = cpp_untangle(ctx_iso, bsc, g.X, session=s_iso)
    t_untangle = time.perf_counter() - t0
    s = CppStructuredSweepSession(ctx, bsc, Xu)
    info = s.core.info()
    n_free = g.n_free
    best, ms = s.core.autotune(20, 0.8)
    s.core.set_graph_chunk(args.graph_chunk)
    s.run(args.warmup, omega=0.8)
    e, m, t_run, _, _ = _timed(s, args.steps, 0.8)
    value = n_free * args.steps / t_run
    # the same sweeps without the terms, same session shape and tile variant
    s0 = s_iso
    s0.set_X(Xu)
    s0.core.set_tile_variant(int(best))
    s0.core.set_graph_chunk(args.graph_chunk)
    s0.run(args.warmup, omega=0.8)
    _, _, t_plain, _, _ = _timed(s0, args.steps, 0.8)
    log(f"terms: stored={info['n_stored']} synth={info['n_synth']} curvature windows={len(curv[0])} "
        f"directional samples={smp.n}; {t_run / args.steps * 1e6:.2f} us/sweep vs {t_plain / args.steps * 1e6:.2f} "
        f"plain -> {value:.4g} node-updates/s")
    cpu = parity = None
    if not args.no_cpu:
        from oracle import oracle as O  # the oracle: CPU baseline and parity reference

        threads = _threads()
        n_cpu = 50
        patches = O.materialize_patches(ctx.wire, g.blocks, 2)
        O.jacobi(2, "barrier", 0.0, 0.8, 5, Xu, patches, ctx.energy_stencil, threads=threads)  # warm-up
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            Xo, eo, mo = O.jacobi(2, "barrier", 0.0, 0.8, n_cpu, Xu, patches, ctx.energy_stencil, threads=threads)
            t.append(time.perf_counter() - t0)
        cpu = {"value": n_free * n_cpu / float(np.median(t)), "unit": "node-updates/s", "cores": threads,
               "kind": "port", "sample": f"{n_cpu} sweeps of the same grid and terms from the untangled state, "
                                         "OpenMP C oracle F64 (5 warm-up sweeps, median of 3 timed runs)"}
        s.set_X(Xu)
        ep, mp = s.run(n_cpu, omega=0.8)
        parity = _parity(s.get_X(), Xo, ep[-1], eo[-1], mp[-1], mo[-1], float((g.X.max(0) - g.X.min(0)).max()),
                         sweeps=n_cpu)
        log(f"cpu {cpu}\nparity {parity}")
    return {
        "metric": METRIC, "value": value, "unit": "node-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_run / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D 5-block O-grid 129/65, 2% tangled by 1.5h, with block-interface curvature "
                               "windows (w 0.5) and directional wall stems and parallel segments (w 0.5), "
                               "omega 0.8",
                   "free_dofs": n_free, "special_dofs": int(info["n_stored"]), "curvature_windows": int(len(curv[0])),
                   "directional_samples": int(smp.n), "seed": SEED},
        "autotune": {"best_variant": int(best), "ms_per_20_sweeps": [float(v) for v in ms]},
        "graph_chunk": args.graph_chunk, "time_to_untangled_s": t_untangle,
        "plain_ms_per_step_same_variant": t_plain / args.steps * 1e3,
        "final_energy": float(e[-1]), "final_min_det": float(m[-1]),
        "parity": parity,
        "roofline": None,  # two kernels per sweep; per-kernel times from the rocprofv3 summary
        "cpu_baseline": cpu,
    }


def bench_mlp(args):
    """The learned proposal's kernels on their own: batched inference (BF16
    features from HBM) and training steps (forward + backward + Adam), 2D
    and 3D networks, on synthetic pairs."""
    from egg_amd._cpp import cpp_core

    out = {}
    rng = np.random.default_rng(SEED)
    for dim, (nin, inp, hid, nout) in ((2, (22, 32, 64, 2)), (3, (85, 96, 128, 3))):
        n = 1 << 20
        X = np.zeros((n, inp), np.float32)
        X[:, :nin] = rng.standard_normal((n, nin)).astype(np.float32)
        T = (0.5 * np.tanh(X[:, :nout])).astype(np.float32)
        net = cpp_core.ProposalMLP(dim, 0, n)
        net.data_add(X, T)
        fwd_flops = 2 * (nin * hid + hid * hid + hid * nout)
        net.forward(X[:4096])
        t0 = time.perf_counter()
        net.forward(X)
        t_f = time.perf_counter() - t0  # includes the host->device copy of X
        res = {"forward_samples_per_s_incl_upload": n / t_f}
        for batch in (4096, 32768):
            net.train(1, batch=batch, lr=1e-3, seed=0)
            t0 = time.perf_counter()
            hist = net.train(2, batch=batch, lr=1e-3, seed=1)
            t_t = time.perf_counter() - t0
            rate = 2 * n / t_t
            res[f"train_batch_{batch}"] = {"samples_per_s": rate, "tflops": rate * 3 * fwd_flops / 1e12,
                                           "frac_of_bf16_mfma_peak": rate * 3 * fwd_flops / 2.5e15,
                                           "loss_last": float(hist[-1])}
        res["flops_per_sample_forward"] = fwd_flops
        out[f"{dim}d"] = res
        log(f"MLP {dim}D: {res}")
    return {"metric": "MLP proposal kernels: training samples/s", "value": out["2d"]["train_batch_32768"]["samples_per_s"],
            "unit": "samples/s", "n_gpus": 1, "steps": 2, "warmup": 1, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
            "config": {"workload": "MLP 22-64-64-2 and 85-128-128-3, 2^20 pairs"}, "mlp": out,
            "roofline": None, "cpu_baseline": None}


def bench_fas(args):
    """FAS on the config-2 O-grid (129/65, 2 % tangled, seed 2026) alone:
    per-cycle wall time and kernel count of V(8,8) on 4 levels (16 coarsest
    sweeps) and V(4,4) on 3 levels from the untangled state, and the time to
    |E - E*| <= 1e-6 E*, min det > 0 (E* pinned in
    tests/golden/estar_ogrid129.json). Cycles replay as captured graphs unless
    --graph-chunk 0 (single launches, e.g. under a kernel tracer)."""
    from egg_amd import grids
    from egg_amd.smoothing import cpp_untangle

    g = grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=SEED)
    ctx, bsc, s = _build(g)
    Xu, md_u, outer, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    s.core.autotune(20, 0.8)
    lv_best, _ = s.core.autotune_levels(16, 0.8)
    multi = None if args.no_multi else _tune_multi(s, 8, fine=False)
    s.core.set_graph_chunk(args.graph_chunk)
    e_star = json.loads((ROOT / "tests" / "golden" / "estar_ogrid129.json").read_text())["E_star"]
    K = max(1, min(args.steps, 100))
    out = {}
    for name, nu, lev, nuc in (("v88_4_levels_coarse16", 8, 4, 16), ("v44_3_levels", 4, 3, 32)):
        s.set_X(Xu)
        s.run_fas(2, nu_pre=nu, nu_post=nu, nu_coarse=nuc, omega=0.8, max_levels=lev)  # warm-up (and capture)
        s.set_X(Xu)
        t0 = time.perf_counter()
        s.run_fas(K, nu_pre=nu, nu_post=nu, nu_coarse=nuc, omega=0.8, max_levels=lev)
        t = time.perf_counter() - t0
        kernels = int(s.core.fas_cycle_kernels())
        s.set_X(Xu)
        n, tc, E = _converge(lambda: (*s.run_fas(5, nu_pre=nu, nu_post=nu, nu_coarse=nuc, omega=0.8,
                                                  max_levels=lev), 5), e_star, 60.0)
        out[name] = {"ms_per_cycle": t / K * 1e3, "kernels_per_cycle": kernels or None, "cycles_to_converged": n,
                     "s_to_converged": tc if n else None, "E_end": E}
        log(f"FAS {name}: {out[name]}")
    r = out["v88_4_levels_coarse16"]
    return {
        "metric": "FAS V-cycle wall time (O-grid 129/65)", "value": r["ms_per_cycle"], "unit": "ms per V(8,8) cycle",
        "n_gpus": 1, "steps": K, "warmup": 2, "ms_per_step": r["ms_per_cycle"], "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2D 5-block O-grid 129/65, 2% tangled by 1.5h, seed 2026, untangled; FAS cycles",
                   "graph_chunk": args.graph_chunk, "fas_level_variants": [int(v) for v in lv_best],
                   "multi_sweep_tuners": multi},
        "cycles": out, "roofline": None, "cpu_baseline": None,
    }


def _control_case(g, ctrl):
    from egg_amd.smoothing import build_structured_context_from_block_maps, build_sweep_context
    from egg_amd.smoothing import CppStructuredSweepSession
    from egg_amd.smoothing.control import (boolean_sum_offset, boundary_ctrl_mask, build_control_wire, control_map,
                                           fit_control_net)

    shape = tuple(np.shape(g.blocks[0]))
    ctx = build_sweep_context(g.blocks, g.free, g.d)
    cm = control_map(shape, ctrl)
    X0 = g.X.reshape(shape + (g.d,))
    C0 = fit_control_net(cm, X0)
    wire = build_control_wire(cm, C0, ~boundary_ctrl_mask(ctrl), boolean_sum_offset(cm, C0, X0))
    s = CppStructuredSweepSession(ctx, build_structured_context_from_block_maps(g.d, g.blocks), g.X, control=wire)
    return ctx, wire, s


def bench_control(args):
    """The control-net reduced Gauss-Newton smoother (run_control) on single
    blocks whose optimum is not a uniform lattice: a 129 x 129 block under a
    wavy top edge (interior perturbed by 0.005 N(0, 1), h = 0.031) with a
    24 x 24 net, and a bulged 33^3 block with a 9^3 net; wall time to
    convergence (max |G| < 1e-8) with the default inexact-Newton forcing.
    CPU baseline and FP64 parity: the dense oracle loop on a 65 x 65 block
    of the same generator with a 12 x 12 net, the device on the same
    problem."""
    from egg_amd import grids

    runs = {}
    g3 = grids.single_block_3d(33, seed=2, amp=0.15)
    x, y, z = g3.X.T
    g3.X = g3.X + np.stack([0 * x, 0 * y, 0.3 * np.sin(np.pi * x) * np.sin(np.pi * y) * z], -1)
    for name, g, ctrl in (("2d_129_net24", grids.wavy_block_2d(129, perturb=0.005, seed=1), (24, 24)),
                          ("3d_33_net9", g3, (9, 9, 9))):
        ctx, wire, s = _control_case(g, ctrl)
        s.run_control(1)  # warm the kernels
        s.set_C(np.asarray(wire["C0"]).reshape(-1, g.d))
        t0 = time.perf_counter()
        r = s.run_control(30)
        t = time.perf_counter() - t0
        runs[name] = {"free_controls": int(len(wire["free_idx"])), "nodes": int(g.n_nodes), "s": t,
                      "iters": int(r["iters"]), "converged": bool(r["converged"]),
                      "pcg_iterations": [int(v) for v in r["pcg_iters"]],
                      "energy_first_last": [float(r["energies"][0]), float(r["energies"][-1])],
                      "final_min_det": float(r["mindets"][-1])}
        log(f"control {name}: {runs[name]}")
    cpu = parity = None
    if not args.no_cpu:
        from oracle import control as OC  # the oracle: CPU baseline and parity reference

        g = grids.wavy_block_2d(65, perturb=0.01, seed=1)
        ctx, wire, s = _control_case(g, (12, 12))
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            Xo, Co, ro = OC.run_control(2, "barrier", g.X, ctx.energy_stencil, wire, g.blocks[0], n_outer=30)
            times.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        rd = s.run_control(30, pcg_max_iter=1000, pcg_rtol=1e-12, pcg_forcing=False)
        t_dev = time.perf_counter() - t0
        cpu = {"value": float(np.median(times)), "unit": "s to converged", "cores": 1, "kind": "port",
               "sample": f"65 x 65 block, 12 x 12 net ({len(wire['free_idx'])} free controls), dense FP64 oracle "
                         f"loop ({ro['iters']} iterations), median of 3; the device on it: {t_dev:.4f} s"}
        eo, ed = np.asarray(ro["energies"]), np.asarray(rd["energies"])
        n = min(len(eo), len(ed))
        parity = {"iters_oracle_device": [int(ro["iters"]), int(rd["iters"])],
                  "max_rel_energy_err": float(np.max(np.abs(ed[:n] - eo[:n]) / np.abs(eo[:n]))),
                  "energies_oracle_device": [[float(v) for v in eo], [float(v) for v in ed]],
                  "max_abs_pos_err": float(np.abs(s.get_X() - Xo).max()),
                  "max_abs_ctrl_err": float(np.abs(s.get_C() - Co).max()),
                  "min_det_err": float(abs(rd["mindets"][-1] - ro["mindets"][-1])),
                  "oracle": "dense FP64 reduced Gauss-Newton (oracle/control.py)"}
        log(f"cpu {cpu}\nparity {parity}")
    r2 = runs["2d_129_net24"]
    return {
        "metric": "control-net reduced Gauss-Newton: wall time to converged", "value": r2["s"], "unit": "s",
        "n_gpus": 1, "steps": r2["iters"], "warmup": 1, "ms_per_step": r2["s"] / max(r2["iters"], 1) * 1e3,
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "129x129 block on [0,4]^2 under a wavy top edge, interior perturbed (seed 1), "
                               "24x24 cubic net, run_control(30) with inexact-Newton forcing; and a bulged 33^3 "
                               "block, 9^3 net"},
        "runs": runs, "cpu_baseline": cpu, "parity": parity, "roofline": None,
    }


def bench_control_topo(args):
    """The multi-block control-net smoother (run_control under a "blocks"
    wire): the 2D O-grid of 129 x 129 + 4 x 129 x 65 nodes (untangled
    generator state) with 24 x 24 and 24 x 12 cubic nets whose seam controls
    are shared (build_control_topology), and the 7 x 33^3 cubed sphere with
    9^3 nets; wall time to convergence (max |G| < 1e-8) with the default
    inexact-Newton forcing. CPU baseline and FP64 parity: the dense oracle
    loop (oracle/control.py run_control_topo) on the 33 x 33 + 4 x 33 x 17
    O-grid with 8 x 8 / 8 x 5 nets, relaxed by 300 oracle Jacobi sweeps
    (where the reduced Hessian is positive definite), the device on the
    same problem with a tight PCG."""
    from egg_amd import grids
    from egg_amd.smoothing import (CppStructuredSweepSession, build_structured_context_from_block_maps,
                                   build_sweep_context)
    from egg_amd.smoothing.control import build_control_topology, topology_mismatch

    def case(g, shapes):
        ctx = build_sweep_context(g.blocks, g.free, g.d)
        wire, _ = build_control_topology(g.blocks, g.X, shapes)
        s = CppStructuredSweepSession(ctx, build_structured_context_from_block_maps(g.d, g.blocks), g.X,
                                      control=wire)
        return ctx, wire, s

    runs = {}
    for name, g, shapes in (("2d_ogrid129_net24", grids.ogrid_2d(129, 65), [(24, 24)] + [(24, 12)] * 4),
                            ("3d_cubed33_net9", grids.cubed_sphere_3d(33), [(9, 9, 9)] * 7)):
        ctx, wire, s = case(g, shapes)
        s.run_control(1)  # warm the kernels
        s.set_C(np.asarray(wire["C0"]).reshape(-1, g.d))
        t0 = time.perf_counter()
        r = s.run_control(30)
        t = time.perf_counter() - t0
        C = s.get_C().reshape(-1)
        runs[name] = {"nets": len(wire["blocks"]), "stacked_controls": int(len(C) // g.d), "reduced_dofs": int(wire["nq"]),
                      "nodes": int(g.n_nodes), "s": t, "iters": int(r["iters"]), "converged": bool(r["converged"]),
                      "pcg_iterations": [int(v) for v in r["pcg_iters"]],
                      "energy_first_last": [float(r["energies"][0]), float(r["energies"][-1])],
                      "final_min_det": float(r["mindets"][-1]),
                      "watertight_mismatch": topology_mismatch(wire, g.blocks, C)}
        log(f"control_topo {name}: {runs[name]}")
    cpu = parity = None
    if not args.no_cpu:
        from oracle import control as OC  # the oracle: CPU baseline and parity reference
        from oracle import oracle as O

        g = grids.ogrid_2d(33, 17)
        ctx = build_sweep_context(g.blocks, g.free, 2)
        pats = O.materialize_patches(ctx.wire, g.blocks, 2)
        X, _, _ = O.jacobi(2, "barrier", 0.0, 0.8, 300, g.X, pats, ctx.energy_stencil, report_every=300, threads=8)
        g = grids.Grid(2, g.blocks, X.astype(np.float32).astype(np.float64), g.free)
        ctx, wire, s = case(g, [(8, 8)] + [(8, 5)] * 4)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            Xo, Co, ro = OC.run_control_topo(2, "barrier", g.X, ctx.energy_stencil, wire, g.blocks, n_outer=30)
            times.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        rd = s.run_control(30, pcg_max_iter=2000, pcg_rtol=1e-12, pcg_forcing=False)
        t_dev = time.perf_counter() - t0
        cpu = {"value": float(np.median(times)), "unit": "s to converged", "cores": 1, "kind": "port",
               "sample": f"33 x 33 + 4 x 33 x 17 O-grid, 8 x 8 / 8 x 5 nets ({wire['nq']} reduced DOFs), dense FP64 "
                         f"oracle loop ({ro['iters']} iterations), median of 3; the device on it: {t_dev:.4f} s"}
        eo, ed = np.asarray(ro["energies"]), np.asarray(rd["energies"])
        n = min(len(eo), len(ed))
        parity = {"iters_oracle_device": [int(ro["iters"]), int(rd["iters"])],
                  "max_rel_energy_err": float(np.max(np.abs(ed[:n] - eo[:n]) / np.abs(eo[:n]))),
                  "max_abs_pos_err": float(np.abs(s.get_X() - Xo).max()),
                  "max_abs_ctrl_err": float(np.abs(s.get_C().reshape(-1) - Co).max()),
                  "oracle": "dense FP64 multi-block reduced Gauss-Newton (oracle/control.py run_control_topo)"}
        log(f"cpu {cpu}\nparity {parity}")
    r2 = runs["2d_ogrid129_net24"]
    return {
        "metric": "multi-block control-net reduced Gauss-Newton: wall time to converged", "value": r2["s"],
        "unit": "s", "n_gpus": 1, "steps": r2["iters"], "warmup": 1,
        "ms_per_step": r2["s"] / max(r2["iters"], 1) * 1e3, "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2D O-grid 129x129 + 4x129x65 (untangled generator state), 24x24 / 24x12 cubic "
                               "nets with shared seam controls, run_control(30) with inexact-Newton forcing; and "
                               "the 7x33^3 cubed sphere with 9^3 nets"},
        "runs": runs, "cpu_baseline": cpu, "parity": parity, "roofline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline and parity leg")
    ap.add_argument("--no-converge", action="store_true", help="skip the time-to-converged runs")
    ap.add_argument("--cpu-sweeps", type=int, default=150)
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-pmc", action="store_true",
                    help="(--workload ogrid) roofline traffic from the committed PMC table instead of two live "
                         "rocprofv3 --pmc child runs")
    ap.add_argument("--no-multi", action="store_true", help="(--workload fas) keep one launch per sweep on every level")
    ap.add_argument("--force-multi", type=lambda v: tuple(int(x) for x in v.split(",")), default=None,
                    metavar="K,TY,TX,KERNEL",
                    help="(--workload ogrid) time the fine level's sweeps with this multi-sweep setting")
    ap.add_argument("--workload", default="ogrid", choices=("ogrid", "ogrid_bl", "ogrid_terms", "batched", "cubed_sphere",
                                                              "mlp", "control", "control_topo", "fas"))
    ap.add_argument("--graph-chunk", type=int, default=GRAPH_CHUNK,
                    help="sweeps per captured hipGraph replay (even; 0 = single launches, e.g. under a tracer)")
    args = ap.parse_args()
    if args.pmc_child:
        _pmc_child(args)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by blocks, one halo exchange per sweep, but only the one-GPU path is built)",
              file=sys.stderr)
        return 2
    if args.workload == "ogrid" and not args.no_pmc:  # before this process touches the GPU
        try:
            args.live_traffic = _live_traffic()
        except Exception as exc:  # the table stands in, and the line says why
            args.live_traffic = (None, f"{type(exc).__name__}: {exc}")
        log(f"live traffic: {args.live_traffic}")
    fn = {"ogrid": bench_ogrid, "ogrid_bl": bench_ogrid_bl, "ogrid_terms": bench_ogrid_terms, "batched": bench_batched, "cubed_sphere": bench_cubed,
          "mlp": bench_mlp, "control": bench_control, "control_topo": bench_control_topo,
          "fas": bench_fas}[args.workload]
    print(json.dumps(fn(args)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
