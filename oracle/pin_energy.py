"""Pin E*, the FP64 plain-method converged energy of a benchmark workload
(TEST INFRASTRUCTURE ONLY: writes a fixture under tests/golden/).

The workload is rebuilt from its generator and seed, untangled with the
delta-continuation schedule of the reference's cpp_untangle
(egg/smoothing/cpp_backend.py:773-848: delta_0 = 2 |min det|, delta *= 0.8,
20 untangle sweeps per delta at omega 0.5, stop once min det > 1e-9) and
then smoothed by FP64 block-Jacobi sweeps (omega 0.8) of the CPU oracle
until the energy changes by less than ``--stop`` (relative) over a window
of ``--window`` sweeps. E* is the energy at that point; the fixture also
records how it was reached.

  python -m oracle.pin_energy --workload ogrid129 [--threads 8]

The converged FP64 state is written next to it (xstar_<workload>.npy), the
reference positions the benchmark's converged results are compared with.

For the 129-node O-grid the plain method needs ~10^5 sweeps, hours of FP64
CPU time from the untangled state; the pinned fixture was made from a warm
start instead (same basin, the FP64 sweeps then settle the last digits):
  python -m oracle.pin_energy --workload ogrid129 --gpu-warm-start ws.npy   (GPU)
  python -m oracle.pin_energy --workload ogrid129 --start ws.npy --window 5000

The batched 64-grid workload (28.6 M DOFs) and the 7 x 65^3 cubed sphere
(1.82 M 3D DOFs) are out of the FP64 CPU oracle's reach (hours per thousand
sweeps). Their E* is the fixed point of the GPU iteration instead (FP32
positions, FP64 energy reductions): untangle, FAS V(4,4) cycles on 3 levels
until the energy changes by less than --stop over a chunk, then plain
sweeps until the same holds, the state's energy recomputed in FP64 on the
CPU. The same procedure is run on the 129-node O-grid first and compared
with its FP64 E*, which bounds how far such a fixed point sits from the FP64
one; the fixture records that bias:
  python -m oracle.pin_energy --workload batched64 --gpu-fixed-point        (GPU)
  python -m oracle.pin_energy --workload cubed_sphere65 --gpu-fixed-point   (GPU)
"""

from __future__ import annotations

import argparse
import json
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent


def workload(name):
    from egg_amd import grids

    if name == "ogrid129":
        return grids.tangle(grids.ogrid_2d(129, 65), frac=0.02, factor=1.5, seed=2026), 2026
    if name == "ogrid33":
        return grids.tangle(grids.ogrid_2d(33, 17), frac=0.02, factor=1.5, seed=7), 7
    if name == "batched64":
        return grids.batched_2d(64, seed=2026)[0], 2026
    if name == "cubed_sphere65":
        return grids.tangle(grids.cubed_sphere_3d(65), frac=0.01, factor=1.5, seed=2026), 2026
    raise ValueError(name)


def gpu_fixed_point(name, stop, max_chunks, log=print):
    """The GPU iteration's fixed point of a workload: (X, E_device, min det,
    history, cycles, sweeps, untangle outer iterations)."""
    from egg_amd.smoothing import (CppStructuredSweepSession, build_structured_context_from_block_maps,
                                   build_sweep_context, cpp_untangle)

    g, _ = workload(name)
    ctx = build_sweep_context(g.blocks, g.free, g.d)
    bsc = build_structured_context_from_block_maps(g.d, g.blocks)
    s = CppStructuredSweepSession(ctx, bsc, g.X)
    _, md, outer, _ = cpp_untangle(ctx, bsc, g.X, session=s)
    s.core.set_graph_chunk(256)
    hist, cycles, sweeps, E_prev = [], 0, 0, None
    for _ in range(max_chunks):  # FAS V(4,4) on 3 levels, 50 cycles a chunk
        e, m = s.run_fas(50, nu_pre=4, nu_post=4, omega=0.8, max_levels=3)
        cycles += 50
        E = float(e[-1])
        hist.append(("fas", cycles, E))
        log(f"{name}: {cycles} FAS cycles E={E:.15g} md={float(m[-1]):.3e}")
        if E_prev is not None and abs(E_prev - E) <= stop * abs(E):
            break
        E_prev = E
    E_prev, met = None, False
    for _ in range(max_chunks):  # then plain sweeps, 2000 a chunk
        e, m = s.run(2000, omega=0.8, report_every=2000)
        sweeps += 2000
        E = float(e[-1])
        hist.append(("plain", sweeps, E))
        log(f"{name}: +{sweeps} plain sweeps E={E:.15g} md={float(m[-1]):.3e}")
        if E_prev is not None and abs(E_prev - E) <= stop * abs(E):
            met = True
            break
        E_prev = E
    return g, ctx, s.get_X(), E, float(m[-1]), hist, cycles, sweeps, outer, met


def main():
    from egg_amd.smoothing import build_sweep_context

    from . import oracle as O

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ogrid129")
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--window", type=int, default=2000)
    ap.add_argument("--stop", type=float, default=1e-10)
    ap.add_argument("--max-sweeps", type=int, default=400000)
    ap.add_argument("--start", default=None,
                    help=".npy of a near-converged state (global order) to start the FP64 sweeps from, "
                         "instead of untangling from the generated grid")
    ap.add_argument("--x-only", action="store_true",
                    help="write only the converged FP64 state (tests/golden/xstar_<workload>.npy), keep the E* fixture")
    ap.add_argument("--gpu-warm-start", default=None, metavar="NPY",
                    help="(on a GPU machine) write a warm start to NPY and stop: the device untangle, "
                         "300 FAS V(4,4) cycles and 20000 plain sweeps at omega 0.8; then rerun with --start NPY")
    ap.add_argument("--gpu-fixed-point", action="store_true",
                    help="(on a GPU machine) pin E* as the GPU iteration's fixed point, its energy recomputed "
                         "in FP64 here, with the O-grid's fixed point against its FP64 E* as the bias bound")
    ap.add_argument("--max-chunks", type=int, default=200)
    ap.add_argument("--conv-rtol", type=float, default=1e-6,
                    help="(--gpu-fixed-point) the relative band around E* a run counts as converged in (wider "
                         "for a workload whose E* is the FP32 floor of a zero-energy optimum)")
    args = ap.parse_args()
    if args.gpu_fixed_point:
        gold = ROOT / "tests" / "golden"
        # the bias of the procedure where the FP64 fixed point is known
        e_ref = json.loads((gold / "estar_ogrid129.json").read_text())["E_star"]
        g0, ctx0, X0, _, _, _, c0, s0, _, _ = gpu_fixed_point("ogrid129", args.stop, args.max_chunks)
        e0, _ = O.energy_mindet(2, "barrier", 0.0, ctx0.energy_stencil, X0.astype(np.float64))
        bias = abs(float(e0) - e_ref) / e_ref
        print(f"ogrid129: GPU fixed point E={float(e0):.15g} against the FP64 E* {e_ref:.15g}: {bias:.3e} relative",
              flush=True)
        g, ctx, X, E_dev, md, hist, cycles, sweeps, outer, met = gpu_fixed_point(args.workload, args.stop,
                                                                                  args.max_chunks)
        E, md64 = O.energy_mindet(g.d, "barrier", 0.0, ctx.energy_stencil, X.astype(np.float64))
        out = {
            "workload": args.workload, "seed": workload(args.workload)[1], "E_star": float(E),
            "min_det": float(md64), "E_device_fp64_reduction": E_dev, "untangle_outer": outer,
            "fas_cycles": cycles, "plain_sweeps": sweeps, "stop_met": met,
            "stop": f"|dE|/E <= {args.stop} over a chunk (50 FAS V(4,4) cycles on 3 levels, then 2000 plain sweeps)",
            "history": hist[-6:],
            "convergence_rtol": args.conv_rtol,
            "how": ("fixed point of the GPU iteration (FP32 positions, FP64 energy reductions): untangle, FAS "
                    "V(4,4) on 3 levels, then plain sweeps at omega 0.8; E* = its energy recomputed in FP64 by "
                    "the CPU oracle. The FP64 CPU oracle cannot reach this workload's fixed point in hours."),
            "bias_bound": {"workload": "ogrid129", "gpu_fixed_point_energy": float(e0), "fp64_E_star": e_ref,
                           "relative": bias, "fas_cycles": c0, "plain_sweeps": s0},
        }
        dst = gold / f"estar_{args.workload}.json"
        dst.write_text(json.dumps(out, indent=1))
        print("wrote", dst, json.dumps({k: v for k, v in out.items() if k != "history"}), flush=True)
        return
    if args.gpu_warm_start:
        from egg_amd.smoothing import (CppStructuredSweepSession, build_structured_context_from_block_maps,
                                       cpp_untangle)

        g, _ = workload(args.workload)
        ctx = build_sweep_context(g.blocks, g.free, g.d)
        bsc = build_structured_context_from_block_maps(g.d, g.blocks)
        s = CppStructuredSweepSession(ctx, bsc, g.X)
        cpp_untangle(ctx, bsc, g.X, session=s)
        s.run_fas(300, nu_pre=4, nu_post=4, omega=0.8)
        s.run(20000, omega=0.8, report_every=20000)
        np.save(args.gpu_warm_start, s.get_X())
        print("wrote", args.gpu_warm_start)
        return

    g, seed = workload(args.workload)
    d = g.d
    ctx = build_sweep_context(g.blocks, g.free, d)
    P = O.materialize_patches(ctx.wire, g.blocks, d)
    es = ctx.energy_stencil
    X = g.X.astype(np.float64)
    t0 = time.time()
    _, md = O.energy_mindet(d, "barrier", 0.0, es, X)
    delta = 2.0 * max(abs(md), 1e-12)
    outer = 0
    if args.start:
        X = np.load(args.start, allow_pickle=False).astype(np.float64).reshape(X.shape)
        _, md = O.energy_mindet(d, "barrier", 0.0, es, X)
        assert md > 0.0, "start state is tangled"
    while md <= 1e-9 and outer < 60:
        X, _, ms = O.jacobi(d, "untangle", delta, 0.5, 20, X, P, es, report_every=20, threads=args.threads)
        md = float(ms[-1])
        outer += 1
        delta *= 0.8
    print(f"untangled: outer={outer} min det={md:.3e} ({time.time() - t0:.1f}s)", flush=True)
    E_prev = None
    sweeps = 0
    hist = []
    while sweeps < args.max_sweeps:
        X, es_, ms_ = O.jacobi(d, "barrier", 0.0, 0.8, args.window, X, P, es, report_every=args.window,
                               threads=args.threads)
        sweeps += args.window
        E = float(es_[-1])
        hist.append((sweeps, E))
        print(f"sweeps={sweeps} E={E:.15g} md={float(ms_[-1]):.3e} ({time.time() - t0:.0f}s)", flush=True)
        if E_prev is not None and abs(E_prev - E) <= args.stop * abs(E):
            break
        E_prev = E
    out = {
        "workload": args.workload,
        "seed": seed,
        "E_star": E,
        "min_det": float(ms_[-1]),
        "untangle_outer": outer,
        "sweeps": sweeps,
        "stop": f"|dE|/E <= {args.stop} over {args.window} sweeps",
        "history": hist[-10:],
        "how": ("FP64 CPU oracle block-Jacobi sweeps, omega 0.8, from " +
                (f"the near-converged state {Path(args.start).name}" if args.start else
                 "the delta-continuation untangle of the generated grid")),
    }
    np.save(ROOT / "tests" / "golden" / f"xstar_{args.workload}.npy", X)
    print("wrote the converged state", f"xstar_{args.workload}.npy")
    if args.x_only:
        return
    dst = ROOT / "tests" / "golden" / f"estar_{args.workload}.json"
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main()
