"""Benchmark: Newton iterations x Mcells/s of the pseudo-transient Newton-Krylov
step on one MI355X.  Default workload: the batched polar -- 32 cases of the 2D
SA RANS case (Karman-Trefftz 10 deg section, C-grid 512 x 128 cells, Re 6e6,
alpha U[-4, 12] deg, Mach U[0.3, 0.78], Roe + smooth van Albada, GMRES(30) x 2
with the line preconditioner) in one context, every launch shared by the
batch: 2.1M cells, FP64 (mixed precision reported beside it).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mixed]
    python bench.py --single            (the 512 x 128 case alone, M 0.73, alpha 2.5, Re 6.5e6)
    python bench.py --polar B           (B cases instead of 32)
    python bench.py --converge [CASE_DIR]
    python bench.py --serve N [--rate R] [--forces-tol T] [--in-flight K]
    python bench.py --wing [--steps K] [--warmup W]
    python bench.py --calibrate-copy

Prints one JSON line.  A step is one full Newton step of every case (residual,
wave speeds, line factor, GMRES, update) on the device; the state starts at
the freestream and W steps run untimed first.  The timed region is bracketed
by device synchronisation.  The roofline entry times the dominant kernel with
HIP events on the library's stream and prices it with the algorithmic bytes
per interior cell of the path it belongs to (residual 8 (2 nv + d^2 + (d+1)
viscous + 1 SA), line apply 3 m^2 (+3 SA) block entries + 16 nv, ...).  The
CPU baseline times the FP64 oracle on a bounded sample of the same work.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "jax-rans_amd")]

import numpy as np  # noqa: E402

CASE = ROOT / "cases" / "kt10_c_sa"
HBM_PEAK_GBS = 8000.0
CPU_SECONDS = 12.0     # Newton steps of the C++ oracle are timed for the CPU baseline until this long has passed


def cr_factor_values(lengths) -> tuple[int, int]:
    """Values the block cyclic reduction keeps per chain of each length L:
    32 per odd row, 48 per even row over all levels, 16 for the final row
    (SA: 2, 3, 1)."""
    flow = sa = 0
    for L in lengths:
        n, odd, even = int(L), 0, 0
        while n > 1:
            odd += n // 2
            even += (n + 1) // 2
            n //= 2
        flow += 32 * odd + 48 * even + 16
        sa += 2 * odd + 3 * even + 1
    return flow, sa


def algorithmic_bytes(name: str, ctx, m: int, restarts: int) -> float:
    """Compulsory bytes of one launch of a hot kernel (every array it must
    touch, once), averaged over the launches of a Newton step.  FP64 fields;
    the line factor in FP32 for ``mixed``."""
    at = ctx.atlas
    nv, nc = ctx.nvar, at.n_canvas
    N = nv * nc * 8                        # one canvas field, bytes
    fb = 4 if ctx.config.precision == "mixed" else 8
    sa = nv > 4
    visc = ctx.viscous
    lengths = np.diff(at.lines.start)
    if name == "precond_apply":
        flow, sav = cr_factor_values(lengths)
        if ctx.config.line_solver == "thomas" or lengths.max() > 512:
            flow, sav = 48 * lengths.sum(), 3 * lengths.sum()
        factor = (flow + (sav if sa else 0)) * fb
        # per cycle (launches per step = cycles x (m + 1)):
        # m - 1 applies read w, write V_j and z; one reads V_0, writes z;
        # each of these m also reads U and writes the tangent's dQ (ghost cells
        # through their nv x nv fill Jacobians; not with the point-vortex far
        # field, whose tangent runs the fill); the cycle end reads T and X, writes X
        fused = ctx.far_field == "characteristic"
        ghosts = len(at.stages[0].dest) + len(at.stages[1].dest)
        tangent_in = (2 * N + ghosts * nv * nv * 8) if fused else 0
        per_cycle = (m - 1) * 3 * N + 2 * N + m * tangent_in + 3 * N
        return ctx.n_cases * (factor + per_cycle / (m + 1))
    if name == "tangent_faces":
        # Q and dQ of every cell, both face-vector families, cell centres (viscous),
        # the tangent face fluxes of both families written
        return ctx.n_cases * nc * 8 * (2 * nv + 4 + (2 if visc else 0) + 2 * nv)
    if name == "tangent_fill":
        return ctx.n_cases * nc * 8 * (2 * nv + nv)       # U, dU read; dQ written
    if name == "tangent_cells":
        # face fluxes of both families, dU, V/dt; SA source: Q, dQ, face vectors,
        # volume, wall distance; J dU written
        extra = (2 * nv + 4 + 2) if sa else 0
        return ctx.n_cases * nc * 8 * (2 * nv + nv + 1 + extra + nv)
    jmean = (m - 1) / 2
    if name == "cgs2_dots":
        return ctx.n_cases * N * (jmean + 1 + 1)          # V_0..V_j, w
    if name in ("cgs2_update_dots", "cgs2_update_norm"):
        return ctx.n_cases * N * (jmean + 1 + 2)          # V_0..V_j, w read and written
    if name == "precond_factor":
        flow, sav = cr_factor_values(lengths)
        if ctx.config.line_solver == "thomas" or lengths.max() > 512:
            flow, sav = 48 * lengths.sum(), 3 * lengths.sum()
        slots = lengths.sum()
        return ctx.n_cases * (slots * 36 * 8 + (flow + (sav if sa else 0)) * fb)
    raise KeyError(name)


def path_bytes_per_cell(ctx) -> dict:
    """Compulsory bytes per interior cell of each hot path (every array
    touched once): residual reads U and the face vectors (d^2 values), writes
    R, plus cell centres and volume (d + 1) when viscous and the wall distance
    for SA; the tangent adds dU and V/dt; the line apply reads the 3 m^2
    block entries (+3 SA) of the factor and r, writes z; the factor reads U,
    the wave speeds and geometry and writes the factor."""
    nv, d = ctx.nvar, 2
    visc = ctx.viscous
    sa = nv > 4
    fb = 4 if ctx.config.precision == "mixed" else 8
    blocks = 3 * 16 + (3 if sa else 0)
    residual = 8 * (2 * nv + d * d + (d + 1 if visc else 0) + (1 if sa else 0))
    return {"residual": residual, "tangent": residual + 8 * (nv + 1),
            "precond_apply": blocks * fb + 16 * nv,
            "precond_factor": 8 * (nv + d + d * d + (2 if sa else 0)) + blocks * fb}


def algorithmic_bytes_per_cell(ctx) -> dict:
    """Algorithmic bytes per interior cell of each hot path, every array
    touched once (nv variables, d dimensions, m x m flow blocks): residual
    ``8 (2 nv + d^2 + (d + 1) viscous + 1 SA)``; tangent ``+ 8 (nv + 1)``;
    wave speeds ``8 (nv + d^2 + d + 1 [+ 1 SA])``; line factor ``8 (nv + d +
    d^2 [+ 2 SA]) + (3 m^2 [+ 3 SA]) b``; line apply ``(3 m^2 [+ 3 SA]) b + 16
    nv`` (b = 8 FP64, 4 FP32 factor); one GMRES(k) cycle of fused three-pass
    CGS2 ``8 nv (1.5 k (k - 1) + 7 k)``."""
    nv, d = ctx.nvar, ctx.atlas.dim
    visc, sa = ctx.viscous, ctx.nvar > d + 2
    m = d + 2
    fb = 4 if ctx.config.precision == "mixed" else 8
    k = ctx.config.krylov_subspace
    blocks = 3 * m * m + (3 if sa else 0)
    residual = 8 * (2 * nv + d * d + (d + 1 if visc else 0) + (1 if sa else 0))
    return {"residual": residual, "tangent": residual + 8 * (nv + 1),
            "wave_speeds": 8 * (nv + d * d + d + 1 + (1 if sa else 0)),
            "precond_factor": 8 * (nv + d + d * d + (2 if sa else 0)) + blocks * fb,
            "precond_apply": blocks * fb + 16 * nv,
            "arnoldi_cycle": 8 * nv * (1.5 * k * (k - 1) + 7 * k)}


def roofline_of(kernels: dict, ctx, m: int, active: float) -> dict:
    """The dominant kernel (largest device time per step) priced with the
    algorithmic bytes of its path: a line apply or factor launch processes
    every interior cell of every case once; a CGS2 pass launch is one of the
    3 m passes of a cycle; the tangent's kernels together make one tangent
    evaluation (their summed time is the divisor)."""
    per = algorithmic_bytes_per_cell(ctx)
    units = ctx.atlas.cells * ctx.n_cases
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    if dom.startswith("tangent"):
        parts = [k for k in kernels if k.startswith("tangent")]
        ms = sum(kernels[k]["ms"] for k in parts)
        alg, label = per["tangent"] * units * active, "+".join(parts)
    elif dom.startswith("cgs2"):
        ms = kernels[dom]["ms"]
        alg, label = per["arnoldi_cycle"] * units / (3 * m) * active, dom
    else:
        ms = kernels[dom]["ms"]
        alg, label = per[dom] * units, dom
    achieved = alg / (ms * 1e-3) / 1e9
    k = kernels[dom]
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": label, "kernel_ms": round(ms, 5),
            "algorithmic_bytes": int(alg), "bytes_per_cell_per_launch": round(alg / units, 2),
            "cells_per_launch": units,
            "fused_bytes": k["algorithmic_bytes"], "fused_GB/s": k["GB/s"]}


def path_rates(ctx) -> dict:
    """Device time of one residual, tangent, line apply and line factor (all of
    their launches, HIP events on the library's stream) and the achieved rate
    of their compulsory bytes."""
    ids = {"residual": "residual", "tangent": "tangent", "precond_apply": "precond_apply",
           "precond_factor": "precond_factor"}
    per = path_bytes_per_cell(ctx)
    out = {}
    for name, kid in ids.items():
        ms = ctx.profile(kid, reps=20)
        b = per[name] * ctx.atlas.cells * ctx.n_cases
        out[name] = {"ms": round(ms, 5), "bytes_per_cell": per[name], "GB/s": round(b / (ms * 1e-3) / 1e9, 1),
                     "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def measured_traffic(kernel: str, workload: str | None):
    """HBM bytes per launch of ``kernel`` on ``workload`` from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated
    on a plain copy), or None when it holds no line for that pair.  Lines of
    ``profiles/pmc_traffic.txt``: ``<workload> <kernel> <bytes per launch> ...``."""
    path = ROOT / "profiles" / "pmc_traffic.txt"
    if workload is None or not path.is_file():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) >= 3 and parts[0] == workload and parts[1] == kernel:
            try:
                return float(parts[2])
            except ValueError:
                return None
    return None


def cpu_baseline() -> dict:
    """The oracle's C++/OpenMP restatement (oracle/cpu) on the same 2D SA case,
    timed on the host's cores in a child process per component, as the
    reference's scripts/benchmark.py:102-113 does (residual, tangent, line
    preconditioner set-up and apply, Newton step; ns per cell), on the
    benchmark input (the freestream times 1 + 1e-3 sin(k)), and
    Newton steps from that state, at least two and until CPU_SECONDS have
    passed, for the throughput.  The
    polar's cases are independent, so the CPU's batch throughput is this
    per-case rate."""
    code = r"""
import os, sys, time, json
sys.path[:0] = [sys.argv[1] + '/jax-rans_amd', sys.argv[1]]
import numpy as np
from jaxrans_amd import load_case
from oracle import residual as OR, cpu as OC
case = load_case(sys.argv[1] + '/cases/kt10_c_sa')
at = case.grid()
st = OR.setup_for(at, case.flow, case.numerics)
c = OC.Case(st, case.implicit)
U0 = OR.U0_of(st)
U = U0 * (1.0 + 1e-3 * np.sin(np.arange(U0.size)).reshape(U0.shape))
dU = 1e-3 * np.cos(np.arange(U0.size)).reshape(U0.shape) * np.abs(U0)
def tm(f, n=3):
    f()
    t = time.perf_counter()
    for _ in range(n):
        f()
    return (time.perf_counter() - t) / n
per = {"residual": tm(lambda: c.residual(U)), "jvp": tm(lambda: c.jvp(U, dU)),
       "precond_setup": tm(lambda: c.precond_setup(U, 10.0)), "precond_apply": tm(lambda: c.precond_apply(dU))}
budget = float(sys.argv[2])
Uc, state = U.copy(), np.array([case.implicit.cfl, np.inf])
t = time.perf_counter()
steps = 0
while steps < 2 or (time.perf_counter() - t < budget and steps < 500):
    c.newton_step(Uc, state)
    steps += 1
dt = time.perf_counter() - t
print(json.dumps({"seconds": dt, "cells": at.cells, "steps": steps, "threads": OC.threads(),
                  "ns_per_cell": {k: v / at.cells * 1e9 for k, v in per.items()} | {"newton_step": dt / steps / at.cells * 1e9}}))
"""
    out = subprocess.run([sys.executable, "-c", code, str(ROOT), str(CPU_SECONDS)], capture_output=True, text=True,
                         timeout=900)
    if out.returncode != 0:
        return {"value": None, "unit": "Mcell-steps/s", "cores": os.cpu_count(), "kind": "port",
                "sample": "oracle C++ Newton step failed: " + out.stderr[-300:]}
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["cells"] * r["steps"] / r["seconds"] / 1e6, "unit": "Mcell-steps/s", "cores": r["threads"],
            "kind": "port", "ns_per_cell": {k: round(v, 1) for k, v in r["ns_per_cell"].items()},
            "sample": f"{r['steps']} FP64 Newton steps (GMRES(30) x 2, block-Thomas line preconditioner) of the "
                      f"oracle's C++/OpenMP restatement on one {r['cells']}-cell case of the workload (the polar's "
                      f"cases are independent), {r['seconds']:.1f} s on {r['threads']} threads; per-component "
                      f"ns per cell as scripts/benchmark.py:102-113"}


def polar_flows(flow, n: int):
    """The batched-polar freestreams: alpha ~ U[-4, 12] deg, Mach ~ U[0.3, 0.78],
    Re 6e6, drawn with numpy.random.default_rng(2)."""
    from dataclasses import replace
    rng = np.random.default_rng(2)
    alphas = rng.uniform(-4.0, 12.0, n)
    machs = rng.uniform(0.3, 0.78, n)
    return [replace(flow, alpha_deg=float(a), mach=float(m), reynolds=6e6) for a, m in zip(alphas, machs)]


def run_steps(case, atlas, precision: str, steps: int, warmup: int, batch: int = 0):
    from dataclasses import replace

    from jaxrans_amd.core.state import q_infinity, uniform_canvas
    from jaxrans_amd.device import context_for
    cfg = replace(case.implicit, precision=precision)
    flows = polar_flows(case.flow, batch) if batch else [case.flow]
    n = len(flows)
    ctx = context_for(atlas, case.flow, case.numerics, cfg, n_cases=n)
    qs = np.stack([q_infinity(f) for f in flows])
    ctx.set_freestream(qs, np.array([f.mu_inf for f in flows]))
    ctx.set_state(np.stack([uniform_canvas(q, atlas.shape, f.gamma) for q, f in zip(qs, flows)]))
    ctx.newton_init(np.full(n, cfg.cfl))
    if warmup:
        ctx.newton_block(warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    hist = ctx.newton_block(steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    return ctx, dt, hist



def wing_bench(steps: int, warmup: int) -> dict:
    """The 3D case: SA RANS on the synthetic swept, tapered wing (NACA
    thickness law t = 0.10, taper 0.55, LE sweep 28 deg, AR 7) in four
    point-matched blocks of 192 x 49 x 29 cells (1,091,328), M 0.8, alpha 2,
    Re 1e7, HLLC + minmod, GMRES(30) x 2, line preconditioner (block Thomas,
    5x5 flow blocks), FP64 and mixed; Newton steps from the freestream.  The
    grid goes through the case path: written as multi-block Plot3D, read back
    by load_case (domain.grid "file") and assembled."""
    import shutil
    import tempfile
    from dataclasses import replace

    from jaxrans_amd import load_case
    from jaxrans_amd.core.state import q_infinity, uniform_canvas
    from jaxrans_amd.device import context_for
    from jaxrans_amd.mesh.wing import write_wing_grid

    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory() as tmp:
        for name in ("case.json", "numerical.json"):
            shutil.copy(ROOT / "cases" / "wing_sa" / name, tmp)
        case = load_case(tmp)
        write_wing_grid(Path(tmp) / case.grid_spec.file)
        atlas = case.grid()
    print(f"[wing] grid written, read back from Plot3D and assembled: atlas {atlas.shape}, {atlas.cells} cells, "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    flow, num, spec = case.flow, case.numerics, case.grid_spec
    ref = {"area": spec.reference_area, "length": spec.reference_length, "centre": tuple(spec.moment_centre)}
    out = {"cells": atlas.cells}
    for precision in ("fp64", "mixed"):
        cfg = replace(case.implicit, precision=precision)
        ctx = context_for(atlas, flow, num, cfg)
        q = q_infinity(flow, dim=3)
        ctx.set_freestream(q, flow.mu_inf)
        ctx.set_state(uniform_canvas(q, atlas.shape, flow.gamma))
        ctx.newton_init(cfg.cfl)
        if warmup:
            ctx.newton_block(warmup)
        ctx.synchronize()
        t = time.perf_counter()
        hist = ctx.newton_block(steps)
        ctx.synchronize()
        dt = time.perf_counter() - t
        probe = ctx.probe_block(2)
        f = ctx.forces3(ref["area"], ref["length"], ref["centre"])[0]
        out[precision] = {"value": round(atlas.cells * steps / dt / 1e6, 4), "ms_per_step": round(1e3 * dt / steps, 3),
                          "residual_first": float(hist[0, 0]), "residual_last": float(hist[-1, 0]),
                          "cl": float(f[0]), "cd": float(f[1]),
                          "probe_us": {k: round(v["ms"] * 1e3, 1) for k, v in probe.items() if v["launches"]},
                          "probe_launches_per_step": {k: v["launches"] / 2 for k, v in probe.items()
                                                      if v["launches"]},
                          "device_bytes": ctx.device_bytes}
        print(f"[wing] {precision}: {json.dumps(out[precision])}", file=sys.stderr, flush=True)
        ctx.close()
    return out


def converge(case_dir: Path, grid=None, overrides=()) -> dict:
    """The case run to convergence through run() in FP64 and in mixed precision:
    iterations, wall time and the force differences (mixed - FP64).
    ``overrides``: case overrides (file.section.key=value) applied to both."""
    from jaxrans_amd import load_case, run
    out = {}
    for precision in ("fp64", "mixed"):
        case = load_case(case_dir, tuple(overrides) + (f"numerical.implicit.precision={json.dumps(precision)}",))
        grid = case.grid() if grid is None else grid
        t0 = time.perf_counter()

        def progress(it, last, orders, f, precision=precision):
            if it % (50 * case.block_iters) == 0:
                print(f"[converge {precision}] {it} iterations, residual {last:.3e} ({orders:.2f} orders), "
                      f"CL {f.cl:+.8f} CD {f.cd:+.8f}", file=sys.stderr, flush=True)

        res = run(case, grid, progress)
        dt = time.perf_counter() - t0
        out[precision] = {"reason": res.history.reason, "iterations": res.history.n_iters, "seconds": round(dt, 3),
                          "residual_drop_orders": round(float(res.history.orders_dropped), 3),
                          "cl": res.forces.cl, "cd": res.forces.cd, "cm": res.forces.cm,
                          "solves_per_min": round(60.0 / dt, 3)}
    f, m = out["fp64"], out["mixed"]
    out["mixed_minus_fp64"] = {"iterations": m["iterations"] - f["iterations"], "cl": m["cl"] - f["cl"],
                               "cd": m["cd"] - f["cd"], "cd_counts": (m["cd"] - f["cd"]) * 1e4, "cm": m["cm"] - f["cm"]}
    return out


def polar_forces(max_iters: int) -> dict:
    """The 32-case batched polar through run_polar in FP64 and in mixed
    precision for up to ``max_iters`` Newton steps (every case stops when all
    have converged, tol 1e-10): per case the residual drop, the CL spread over
    the last quarter of its blocks and mixed - FP64 in CL and CD."""
    from dataclasses import replace

    from jaxrans_amd import load_case, run_polar
    flows = polar_flows(load_case(CASE).flow, 32)
    alphas = [f.alpha_deg for f in flows]
    machs = [f.mach for f in flows]
    out = {}
    for precision in ("fp64", "mixed"):
        case = load_case(CASE, (f"numerical.implicit.precision={json.dumps(precision)}",
                                f"numerical.convergence.max_iters={max_iters}"))
        t0 = time.perf_counter()

        def progress(it, cl, precision=precision):
            if it % 500 == 0:
                print(f"[polar {precision}] {it} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

        polar = run_polar(case, alphas, machs=machs, reynolds=6e6, progress=progress)
        out[precision] = (polar, time.perf_counter() - t0)
        print(f"[polar {precision}] {polar.timing.iterations} steps in {out[precision][1]:.1f} s", file=sys.stderr,
              flush=True)
    f64, mixed = out["fp64"][0], out["mixed"][0]
    rows = []
    for k, (a, m) in enumerate(zip(alphas, machs)):
        rf, rm = f64.results[k], mixed.results[k]
        rows.append({"alpha": round(a, 3), "mach": round(m, 4), "orders_fp64": round(float(rf.history.orders_dropped), 2),
                     "orders_mixed": round(float(rm.history.orders_dropped), 2),
                     "cl_spread_fp64": float(rf.cl_spread), "cl": rf.forces.cl, "cd": rf.forces.cd,
                     "dcl": rm.forces.cl - rf.forces.cl, "dcd_counts": (rm.forces.cd - rf.forces.cd) * 1e4})
    return {"max_iters": max_iters, "seconds": {k: round(v[1], 1) for k, v in out.items()}, "cases": rows}


def _serve_grid(args):
    from jaxrans_amd import serve
    req, cfg = args
    return req.rid, serve.request_grid(req, cfg)


SERVE_MODES = {"fp64_serial": ("fp64", False), "fp64_served": ("fp64", True), "mixed_served": ("mixed", True),
               "mixed_serial": ("mixed", False)}


def serve_bench(n: int, rate: float, forces_tol: float, in_flight: int, max_iters: int, driver: str = "slots",
                modes=("fp64_serial", "fp64_served", "mixed_served"), slots: int = 16, start: int = 0) -> dict:
    """The first ``n`` requests of the 256-request stream (Poisson arrivals at
    ``rate`` per second, numpy default_rng(3)) served three ways: the plain
    FP64 path one solve at a time; served FP64; served mixed, ``in_flight``
    solves on the device together.  Grids are built beforehand in a process
    pool (CPU grid generation is not part of a solve's latency here).
    Solves end when their forces settle (CL, CD, CM each within
    ``forces_tol`` over the last quarter of the blocks) or at
    ``max_iters``."""
    import multiprocessing as mp

    from jaxrans_amd import serve
    reqs = serve.make_requests(256, rate)[start:start + n]
    if start:   # a later part of the stream: its clock starts at its first arrival
        from dataclasses import replace as _replace
        reqs = [_replace(r, arrival=r.arrival - reqs[0].arrival) for r in reqs]
    cfg = serve.default_config("fp64", max_iters=max_iters, forces_tol=forces_tol, max_in_flight=in_flight,
                               driver=driver, slots=slots)
    t0 = time.perf_counter()
    workers = max(1, min(16, os.cpu_count() or 1, n))
    with mp.get_context("fork").Pool(workers) as pool:
        atlases = dict(pool.map(_serve_grid, [(r, cfg) for r in reqs], chunksize=1))
    print(f"[serve] {n} grids built in {time.perf_counter() - t0:.1f} s on {workers} processes", file=sys.stderr,
          flush=True)
    runs = {}
    for label in modes:
        precision, served = SERVE_MODES[label]
        c = serve.with_precision(cfg, precision)
        t = time.perf_counter()
        def done(o, label=label):
            print(f"[serve {label}] request {o.rid} ({o.cells} cells): {o.reason} after {o.iterations} steps, "
                  f"CL {o.cl:+.7f}, latency {o.latency:.2f} s", file=sys.stderr, flush=True)

        outs = (serve.serve if served else serve.serve_serial)(reqs, atlases, c, on_done=done)
        wall = time.perf_counter() - t
        runs[label] = ({o.rid: o for o in outs}, serve.summary(outs, wall))
        print(f"[serve] {label}: {json.dumps(runs[label][1])}", file=sys.stderr, flush=True)
    base = runs[modes[0]][0]
    out = {}
    for label, (by_id, summ) in runs.items():
        dcl = max(abs(by_id[k].cl - base[k].cl) for k in base)
        dcd = max(abs(by_id[k].cd - base[k].cd) for k in base)
        dit = [by_id[k].iterations - base[k].iterations for k in base]
        # the force bar against the first mode (CL within 1e-6, CD within 0.01 count), over the requests
        # both runs counted (converged or settled)
        both = [k for k in base if by_id[k].reason in serve.COUNTED and base[k].reason in serve.COUNTED]
        within = [k for k in both if abs(by_id[k].cl - base[k].cl) <= 1e-6 and abs(by_id[k].cd - base[k].cd) <= 1e-6]
        out[label] = {**summ, f"max_abs_dcl_vs_{modes[0]}": dcl, f"max_abs_dcd_counts_vs_{modes[0]}": dcd * 1e4,
                      f"mean_diterations_vs_{modes[0]}": float(np.mean(dit)),
                      f"counted_in_both_vs_{modes[0]}": len(both), f"within_force_bar_vs_{modes[0]}": len(within)}
    for label in modes[1:]:
        out[f"solves_per_min_{label}_over_{modes[0]}"] = round(
            out[label]["solves_per_min"] / out[modes[0]]["solves_per_min"], 3)
    out["requests"] = [{"rid": r.rid, "cells": list(r.cells), "mach": round(r.mach, 4), "alpha": round(r.alpha_deg, 3),
                        "re": r.reynolds, **{lab: [by_id[r.rid].iterations, by_id[r.rid].reason,
                                                   round(by_id[r.rid].cl, 10), round(by_id[r.rid].cd, 10), round(by_id[r.rid].latency, 2)]
                                             for lab, (by_id, _) in runs.items()}} for r in reqs]
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="instead of the step benchmark: one plain 512 MiB device-to-device copy (after a "
                         "512 MiB eviction fill), the reference dispatch for calibrating rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--polar", type=int, metavar="B", default=0,
                    help="B cases of the batched polar instead of 32 (alpha uniform on [-4, 12] deg, Mach uniform "
                         "on [0.3, 0.78], Re 6e6, numpy default_rng(2)), every launch shared by the batch")
    ap.add_argument("--steps-only", action="store_true",
                    help="only the warmup and timed Newton steps (no probe, path timings, mixed run or CPU "
                         "baseline): the process to run under a profiler's counters")
    ap.add_argument("--single", action="store_true",
                    help="instead of the batched polar: the 512 x 128 case alone (M 0.73, alpha 2.5, Re 6.5e6)")
    ap.add_argument("--converge", metavar="CASE_DIR", nargs="?", const=str(CASE), default=None,
                    help="instead of the step benchmark: run CASE_DIR (default: the benchmark case) to "
                         "convergence in FP64 and mixed precision and print iterations and force differences")
    ap.add_argument("--set", action="append", default=[], metavar="FILE.SECTION.KEY=VALUE",
                    help="a case override, as load_case takes it (repeatable), for --converge and the step "
                         "benchmark, e.g. 'numerical.implicit.line_solver=\"thomas\"'")
    ap.add_argument("--wing", action="store_true",
                    help="instead of the 2D case: Newton steps of the 3D SA wing (four blocks, 1.09M cells)")
    ap.add_argument("--polar-forces", type=int, metavar="MAX_ITERS", default=0,
                    help="instead of the step benchmark: the 32-case polar through run_polar in FP64 and mixed "
                         "precision, per-case residual drop and mixed - FP64 forces")
    ap.add_argument("--serve", type=int, metavar="N", default=0,
                    help="instead of the step benchmark: the served stream, first N of its 256 requests")
    ap.add_argument("--rate", type=float, default=0.5, help="--serve: Poisson arrival rate, requests per second")
    ap.add_argument("--serve-from", type=int, default=0, metavar="K",
                    help="--serve: requests K .. K+N-1 of the stream, its clock started at request K's arrival "
                         "(a long plain-path baseline measured in parts)")
    ap.add_argument("--forces-tol", type=float, default=1e-5, help="--serve: settle tolerance on CL, CD, CM")
    ap.add_argument("--in-flight", type=int, default=8, help="--serve: solves on the device together")
    ap.add_argument("--max-iters", type=int, default=4000, help="--serve: Newton budget per solve")
    ap.add_argument("--driver", choices=("threads", "async", "slots"), default="slots",
                    help="--serve: a host thread per solve, one thread polling asynchronous blocks, or one batched "
                         "context per grid size with the requests in its slots (shared launches)")
    ap.add_argument("--slots", type=int, default=16, help="--serve --driver slots: slots per grid size (at most; "
                                                          "about 2^20 cells per size)")
    ap.add_argument("--serve-modes", default="fp64_serial,fp64_served,mixed_served",
                    help=f"--serve: comma-separated runs of {sorted(SERVE_MODES)}; the first is the reference "
                         f"for the force differences")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): this build has no multi-GPU path; "
              f"a batch of cases would shard over GPUs with no communication", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        import torch
        n = 64 * 2**20                                   # 512 MiB of float64
        src = torch.full((n,), 1.0, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        evict = torch.full((n,), 2.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        ok = bool(torch.equal(dst[:: 2**16], src[:: 2**16]))
        print(json.dumps({"calibration_copy_bytes": n * 8, "checked": ok, "evict_sum": float(evict[:8].sum())}),
              flush=True)
        return 0 if ok else 1
    if args.wing:
        w = wing_bench(args.steps, args.warmup)
        line = {"metric": "Newton iters x Mcells/s", "value": w["fp64"]["value"], "unit": "Mcell-steps/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": w["fp64"]["ms_per_step"],
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                "config": {"workload": "3D SA RANS, swept tapered wing (t 0.10, taper 0.55, LE sweep 28, AR 7), "
                                       "4 blocks of 192x49x29 cells, M 0.8, alpha 2, Re 1e7, HLLC + minmod, "
                                       "GMRES(30)x2, line_j (block Thomas, 5x5), FP64",
                           "cells": w["cells"]},
                "mixed_value": w["mixed"]["value"], "detail": w}
        print(json.dumps(line), flush=True)
        return 0
    if args.polar_forces:
        print(json.dumps({"metric": "mixed vs FP64 per case, batched polar", **polar_forces(args.polar_forces)}),
              flush=True)
        return 0
    if args.serve:
        line = {"metric": "solves/min + p50/p95/p99 latency, served stream", "n_gpus": 1,
                "config": {"workload": f"served stream: requests {args.serve_from}..{args.serve_from + args.serve - 1} "
                                       f"of 256, Poisson {args.rate}/s, "
                                       "Karman-Trefftz (x0 U[-0.12,-0.04], y0 U[0,0.06], tau U[5,15]), C-grids "
                                       "{257x65, 513x129, 769x193, 1025x257} nodes, M U[0.3,0.8], alpha U[-2,6], "
                                       "Re log-U[1e6,2e7], SA, Roe + van_albada_smooth, GMRES(30)x2, line_j",
                           "forces_tol": args.forces_tol, "max_iters": args.max_iters, "in_flight": args.in_flight,
                           "driver": args.driver, "slots": args.slots},
                **serve_bench(args.serve, args.rate, args.forces_tol, args.in_flight, args.max_iters, args.driver,
                              tuple(args.serve_modes.split(",")), args.slots, args.serve_from)}
        print(json.dumps(line), flush=True)
        return 0
    if args.converge:
        line = {"metric": "mixed vs FP64 at convergence", "case": args.converge, "n_gpus": 1,
                "overrides": list(args.set), **converge(Path(args.converge), overrides=args.set)}
        print(json.dumps(line), flush=True)
        return 0
    from jaxrans_amd import load_case
    case = load_case(CASE, tuple(args.set))
    atlas = case.grid()
    B = 0 if args.single else (args.polar or 32)
    ncases = max(1, B)
    cells = atlas.cells * ncases
    print(f"[bench] {ncases} case(s) on the {atlas.cells}-cell grid: {args.warmup} warmup + {args.steps} timed "
          f"Newton steps, FP64", file=sys.stderr, flush=True)
    ctx, dt, hist = run_steps(case, atlas, "fp64", args.steps, args.warmup, B)
    ms_per_step = 1e3 * dt / args.steps
    if args.steps_only:
        print(json.dumps({"workload": "polar" + str(B) if B else "single", "steps": args.steps,
                          "ms_per_step": round(ms_per_step, 3)}), flush=True)
        ctx.close()
        return 0
    value = cells * args.steps / dt / 1e6
    m, restarts = case.implicit.krylov_subspace, case.implicit.krylov_restarts
    F64 = ctx.forces()                   # after warmup + steps, as the mixed run below
    # Live probe: the same Newton steps again, every launch of the hot kernels
    # bracketed by HIP events on the library's stream.
    cyc0 = ctx.scalars()[:, 8].copy()
    probe = ctx.probe_block(args.steps)
    # A later GMRES cycle is launched when some case still needs it; cases that
    # met their tolerance exit its kernels at once.  The Arnoldi kernels' bytes
    # count the active (case, column) pairs: cycles run x m over cases x columns launched.
    columns = probe["tangent_faces"]["launches"]
    active = float((ctx.scalars()[:, 8] - cyc0).sum() * m / max(1, ctx.n_cases * columns))
    arnoldi = ("precond_apply", "tangent_faces", "tangent_fill", "tangent_cells", "cgs2_dots", "cgs2_update_dots",
               "cgs2_update_norm")
    kernels = {}
    for k, v in probe.items():
        if not v["launches"]:
            continue                       # not launched on this path (e.g. the fill, folded into the apply)
        alg = algorithmic_bytes(k, ctx, m, restarts) * (active if k in arnoldi else 1.0)
        kernels[k] = {"launches_per_step": v["launches"] / args.steps, "ms": round(v["ms"], 5),
                      "ms_per_step": round(v["ms"] * v["launches"] / args.steps, 4),
                      "algorithmic_bytes": int(alg), "GB/s": round(alg / (v["ms"] * 1e-3) / 1e9, 1)}
    roofline = roofline_of(kernels, ctx, m, active)
    roofline["traffic"] = measured_traffic(roofline["kernel"], "polar32" if B == 32 else "single" if not B else None)
    roofline["active_gmres_fraction"] = round(active, 4)
    paths = path_rates(ctx)
    scal = ctx.scalars()
    ctx.close()
    extra = {}
    if not args.no_mixed:
        print("[bench] the same steps in mixed precision", file=sys.stderr, flush=True)
        cm, dtm, histm = run_steps(case, atlas, "mixed", args.steps, args.warmup, B)
        Fm = cm.forces()
        extra = {"mixed_value": round(cells * args.steps / dtm / 1e6, 4),
                 "mixed_ms_per_step": round(1e3 * dtm / args.steps, 3),
                 "mixed_last_residual_max": float(np.nanmax(histm[-1])),
                 "mixed_minus_fp64_after_the_timed_steps": {
                     "max_abs_dcl": float(np.max(np.abs(Fm[:, 0] - F64[:, 0]))),
                     "max_abs_dcd_counts": float(np.max(np.abs(Fm[:, 1] - F64[:, 1])) * 1e4),
                     "note": "transient states after warmup + steps Newton steps from the freestream, not "
                             "converged solutions"}}
        cm.close()
    solve = {"thomas": "block Thomas", "cyclic_reduction": "block cyclic reduction"}[case.implicit.line_solver]
    if B:
        workload = (f"batched polar: {B} cases on the same Karman-Trefftz 10 deg C-grid 512x128 "
                    f"({atlas.cells} cells each), alpha U[-4,12] deg, M U[0.3,0.78] (numpy default_rng(2)), Re 6e6, "
                    f"SA-noft2, Roe + van_albada_smooth, GMRES(30)x2, line_j ({solve}), FP64, "
                    "one context: every launch shared by the batch")
    else:
        workload = ("2D SA RANS, Karman-Trefftz 10 deg, C-grid 512x128 (65536 cells), M 0.73, alpha 2.5, Re 6.5e6, "
                    f"Roe + van_albada_smooth, GMRES(30)x2, line_j ({solve}), FP64")
    line = {
        "metric": "Newton iters x Mcells/s", "value": round(value, 4), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "cells": cells, "cases": ncases, "krylov_subspace": m,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "kernels": kernels,
        "paths": paths,
        "residual_first_max": float(np.nanmax(hist[0])), "residual_last_max": float(np.nanmax(hist[-1])),
        "cfl_last_median": float(np.median(scal[:, 0])),
        **extra,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
