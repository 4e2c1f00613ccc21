"""Served stream of uneven solves on one GPU.

Requests (a Karman-Trefftz section, a C-grid size, Mach, alpha, Reynolds)
arrive over time.  The default driver batches them continuously: every grid
size (one topology) has one batched context whose slots hold requests of
that size, each slot with its own geometry (``jr_set_geometry``), freestream
and Newton control, so the requests of a size share every launch of a
Newton step, their states and Krylov bases resident in HBM; the contexts of
the four sizes run on their own streams (four, the box's hardware queues)
and overlap on the device.  Between blocks of Newton steps a finished slot
is refilled from the arrived requests of its size, or parked
(``jr_park_case``: its GMRES skipped, its state kept) when none is waiting.
Slots per size: about ``slot_cells`` cells per context, within the free
device memory less a reserve.  A solve retires when the driver's block test
of ``driver.py:161-201`` says converged, non-finite, stalled or out of
iterations, or when its forces have settled (CL, CD and CM each moving less
than ``forces_tol`` over the last quarter of the blocks,
``simulation.py:254-272``) -- the same tests, in the same order, as
``simulation.run``.

Every kernel of a batched step works case by case, so a request's result is
the same alone and in the mix (a GPU test checks it bit for bit).  The
other drivers give each request its own context and stream, marched by a
host thread per solve or by one thread polling asynchronous blocks.

The reference has no serving loop; its closest analogue is the slot
scheduler of ``src/jaxrans/sweep.py:172-327`` (slots refilled between
blocks, idle slots masked), which holds one grid and geometry for all slots;
per-slot geometry and one context per grid size lift that restriction.
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field, replace

import numpy as np

from .case.loader import resolve_conditions
from .core.config import FlowConditions, ImplicitConfig, NumericsConfig
from .core.state import q_infinity, uniform_canvas
from .device import Context, context_for, mem_info
from .geometry.airfoil import karman_trefftz
from .mesh.assemble import assemble
from .mesh.atlas import assign_conditions, build_atlas, patches
from .mesh.generate import c_grid
from .simulation import ExitReason, cl_spread

GRID_SIZES = ((256, 64), (512, 128), (768, 192), (1024, 256))   # (cells around, cells normal) of the C-grids
FORCE_REF = (1.0, (0.25, 0.0))     # unit chord, moments about the quarter chord


@dataclass(frozen=True)
class Request:
    """One solve: a Karman-Trefftz section (trailing-edge angle, circle
    centre), a C-grid of ``cells`` = (around, normal) cells, and a freestream."""

    rid: int
    arrival: float              # seconds after the stream starts
    te_angle_deg: float
    centre: tuple[float, float]
    cells: tuple[int, int]
    mach: float
    alpha_deg: float
    reynolds: float | None


def make_requests(n: int, rate: float, seed: int = 3, sizes=GRID_SIZES, viscous: bool = True) -> list[Request]:
    """The served stream: Poisson arrivals at ``rate`` per second; centre
    offsets x ~ U[-0.12, -0.04], y ~ U[0, 0.06]; trailing-edge angle ~ U[5, 15]
    deg; a grid drawn uniformly from ``sizes``; M ~ U[0.3, 0.8]; alpha ~
    U[-2, 6] deg; Re log-uniform on [1e6, 2e7] (None when not viscous)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    out = []
    for k in range(n):
        cx, cy = rng.uniform(-0.12, -0.04), rng.uniform(0.0, 0.06)
        tau = rng.uniform(5.0, 15.0)
        size = sizes[int(rng.integers(len(sizes)))]
        mach, alpha = rng.uniform(0.3, 0.8), rng.uniform(-2.0, 6.0)
        re = float(np.exp(rng.uniform(np.log(1e6), np.log(2e7))))
        out.append(Request(k, float(t[k]), float(tau), (float(cx), float(cy)), tuple(size), float(mach),
                           float(alpha), re if viscous else None))
    return out


@dataclass(frozen=True)
class ServeConfig:
    """Numerics of every served solve and the admission policy."""

    numerics: NumericsConfig
    implicit: ImplicitConfig
    turbulence: str | None = "sa_noft2"
    wall_spacing: float = 1e-6
    far_field: float = 100.0
    block_iters: int = 20
    max_iters: int = 4000
    tol: float = 1e-10
    forces_tol: float | None = 1e-5
    max_in_flight: int = 8
    reserve_bytes: int = 8 << 30          # device memory kept free
    driver: str = "slots"                 # "slots": one batched context per grid size, requests in its slots;
    #                                       "threads": a context and host thread per solve; "async": one
    #                                       thread polls every solve's context
    slots: int = 16                       # "slots": cases per batched context (at most) ...
    slot_cells: int = 1 << 20             # ... and about this many cells per batched context

    def __post_init__(self) -> None:
        if self.driver not in ("threads", "async", "slots"):
            raise ValueError(f"driver must be 'threads', 'async' or 'slots', got {self.driver!r}")


def request_grid(req: Request, cfg: ServeConfig):
    """The request's atlas: its section (unit chord, sharp trailing edge), a
    C-grid with 3/4 of the around-cells on the section and 1/8 on each wake
    side, no-slip (viscous) or slip wall and a characteristic far field."""
    around, normal = req.cells
    foil = karman_trefftz(req.te_angle_deg, req.centre)
    wake = around // 8
    mesh = assemble([c_grid(foil, around - 2 * wake, wake, cells_normal=normal, wall_spacing=cfg.wall_spacing,
                            far_field=cfg.far_field)])
    kinds = resolve_conditions({}, patches(mesh), request_flow(req, cfg))
    return build_atlas(assign_conditions(mesh, kinds), cfg.numerics.n_ghost)


def request_flow(req: Request, cfg: ServeConfig) -> FlowConditions:
    viscous = req.reynolds is not None
    return FlowConditions(mach=req.mach, alpha_deg=req.alpha_deg, reynolds=req.reynolds,
                          turbulence=cfg.turbulence if viscous else None)


@dataclass
class Solve:
    """An admitted request on its device context."""

    req: Request
    ctx: Context
    started: float
    peak: float = 0.0
    reason: str = ExitReason.MAX_ITERS
    iters: int = 0
    coefficients: list = field(default_factory=list)    # (CL, CD, CM) after each block
    last_residual: float = float("nan")

    def block_done(self, res: np.ndarray, forces: np.ndarray, cfg: ServeConfig) -> bool:
        """The block test of ``simulation.run_to_steady`` on one case
        (driver.py:161-201 plus the settle test of simulation.py:254-272);
        True when the solve ends."""
        col = res[:, 0]
        self.iters += col.size
        self.coefficients.append(tuple(float(v) for v in forces[0, :3]))
        self.last_residual = float(col[-1])
        finite = np.where(np.isfinite(col), col, 0.0)
        self.peak = max(self.peak, float(finite.max()))
        settled = cfg.forces_tol is not None and all(
            cl_spread(h) < cfg.forces_tol for h in np.transpose(self.coefficients))
        if not np.all(np.isfinite(col)):
            self.reason = ExitReason.NON_FINITE
        elif self.peak > 0 and col[-1] <= cfg.tol * self.peak:
            self.reason = ExitReason.CONVERGED
        elif col.size > 1 and np.all(col == col[0]):
            self.reason = ExitReason.STALLED
        elif settled:
            self.reason = ExitReason.SETTLED
        elif self.iters >= cfg.max_iters:
            self.reason = ExitReason.MAX_ITERS
        else:
            return False
        return True


@dataclass(frozen=True)
class Outcome:
    rid: int
    arrival: float
    started: float
    finished: float
    iterations: int
    reason: str
    cl: float
    cd: float
    cm: float
    cells: int
    orders: float               # log10(peak / last) of the residual

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def start_solve(req: Request, atlas, cfg: ServeConfig, now: float, enqueue: bool = True) -> Solve:
    """The request's context at the freestream; with ``enqueue`` its first
    block is enqueued asynchronously."""
    flow = request_flow(req, cfg)
    ctx = context_for(atlas, flow, cfg.numerics, cfg.implicit)
    q = q_infinity(flow)
    ctx.set_freestream(q, None if flow.inviscid else flow.mu_inf)
    ctx.set_state(uniform_canvas(q, atlas.shape, flow.gamma))
    ctx.newton_init(cfg.implicit.cfl)
    if enqueue:
        ctx.newton_block_async(cfg.block_iters, FORCE_REF)
    return Solve(req, ctx, now)


def march(s: Solve, cfg: ServeConfig) -> None:
    """Blocking blocks of Newton steps, each followed by the forces, until
    the block test ends the solve (what ``simulation.run`` does for a case)."""
    while True:
        res = s.ctx.newton_block(cfg.block_iters)
        if s.block_done(res, s.ctx.forces(FORCE_REF[0], FORCE_REF[1]), cfg):
            return


def finish_solve(s: Solve, now: float) -> Outcome:
    cl, cd, cm = s.coefficients[-1]
    orders = float(np.log10(s.peak / s.last_residual)) if s.peak > 0 and s.last_residual > 0 else 0.0
    out = Outcome(s.req.rid, s.req.arrival, s.started, now, s.iters, s.reason, cl, cd, cm, int(s.ctx.atlas.cells),
                  orders)
    s.ctx.close()
    return out


def serve(requests: list[Request], atlases: dict, cfg: ServeConfig, clock=time.perf_counter,
          on_done=None) -> list[Outcome]:
    """Run the stream: requests become eligible at their arrival time (the
    clock starts at the call), are admitted in arrival order while the free
    device memory less ``reserve_bytes`` holds their estimated footprint and
    fewer than ``max_in_flight`` solves run, and are stepped block by block
    until their driver test ends them.  ``atlases`` maps request ids to their
    grids (built beforehand).  ``on_done(outcome)`` is called as each solve
    ends.  Returns one outcome per request, in finishing order.

    With ``driver="threads"`` each admitted solve is marched by its own host
    thread through blocking blocks (the library call releases the
    interpreter lock, so the threads enqueue and wait concurrently, each on
    its context's stream); with ``driver="async"`` this thread enqueues
    every solve's blocks asynchronously and polls them.  With
    ``driver="slots"`` the solves share launches: requests of one grid size
    (one topology) are packed into the slots of one batched context, each
    slot with its own geometry, freestream and Newton control, refilled
    between blocks as slots free up (the reference's slot scheduler,
    ``src/jaxrans/sweep.py:172-327``, with per-slot grids)."""
    if cfg.driver == "threads":
        return _serve_threads(requests, atlases, cfg, clock, on_done)
    if cfg.driver == "slots":
        return _serve_slots(requests, atlases, cfg, clock, on_done)
    t0 = clock()
    waiting = sorted(requests, key=lambda r: r.arrival)
    running: list[Solve] = []
    done: list[Outcome] = []
    footprint: dict = {}          # grid size -> device bytes of a context (measured on the first)
    while waiting or running:
        now = clock() - t0
        # admit
        while waiting and waiting[0].arrival <= now and _admissible(waiting[0], len(running), footprint, cfg):
            req = waiting.pop(0)
            s = start_solve(req, atlases[req.rid], cfg, now)
            footprint.setdefault(req.cells, s.ctx.device_bytes)
            running.append(s)
        # step
        progressed = False
        for s in list(running):
            if not s.ctx.poll():
                continue
            progressed = True
            res, forces = s.ctx.wait()
            if s.block_done(res, forces, cfg):
                running.remove(s)
                done.append(finish_solve(s, clock() - t0))
                if on_done is not None:
                    on_done(done[-1])
            else:
                s.ctx.newton_block_async(cfg.block_iters, FORCE_REF)
        if not progressed:
            if not running and waiting:
                time.sleep(max(0.0, min(0.01, waiting[0].arrival - (clock() - t0))))
            else:
                time.sleep(0.0002)
    return done


def _admissible(req: Request, n_running: int, footprint: dict, cfg: ServeConfig) -> bool:
    """Room for ``req``: fewer than ``max_in_flight`` solves, and free device
    memory less the reserve holds the footprint measured for its grid size
    (the first solve of a size is admitted on the free memory alone)."""
    if n_running >= cfg.max_in_flight:
        return False
    if n_running == 0:
        return True
    free, _ = mem_info()
    return free - cfg.reserve_bytes >= footprint.get(req.cells, 0)


def _serve_threads(requests, atlases, cfg: ServeConfig, clock, on_done) -> list[Outcome]:
    t0 = clock()
    waiting = sorted(requests, key=lambda r: r.arrival)
    footprint: dict = {}
    finished: list = []
    cond = threading.Condition()
    done: list[Outcome] = []
    n_running = 0

    def work(s: Solve) -> None:
        try:
            march(s, cfg)
            out = finish_solve(s, clock() - t0)
        except BaseException as e:          # handed to the serving thread
            out = e
        with cond:
            finished.append(out)
            cond.notify()

    threads = []
    while waiting or n_running:
        now = clock() - t0
        while waiting and waiting[0].arrival <= now and _admissible(waiting[0], n_running, footprint, cfg):
            req = waiting.pop(0)
            s = start_solve(req, atlases[req.rid], cfg, now, enqueue=False)
            footprint.setdefault(req.cells, s.ctx.device_bytes)
            t = threading.Thread(target=work, args=(s,), name=f"solve-{req.rid}", daemon=True)
            threads.append(t)
            n_running += 1
            t.start()
        with cond:
            if not finished:
                timeout = None
                if waiting and waiting[0].arrival > now and n_running < cfg.max_in_flight:
                    timeout = waiting[0].arrival - now
                elif waiting and n_running < cfg.max_in_flight:
                    timeout = 0.01          # memory-bound admission: look again soon
                cond.wait(timeout)
            ready, finished[:] = list(finished), []
        for out in ready:
            n_running -= 1
            if isinstance(out, BaseException):
                raise out
            done.append(out)
            if on_done is not None:
                on_done(out)
    for t in threads:
        t.join()
    return done


def slots_for(cells: int, cfg: ServeConfig) -> int:
    """Slots of the batched context of a grid of ``cells`` cells: about
    ``slot_cells`` cells per context, at most ``slots``."""
    return int(max(1, min(cfg.slots, cfg.slot_cells // max(cells, 1))))


def _serve_slots(requests, atlases, cfg: ServeConfig, clock, on_done) -> list[Outcome]:
    """Continuous batching: one batched context per grid size (one
    topology: the same index plans, every slot with its own geometry,
    freestream and Newton control), each on its own stream and marched by
    its own host thread, so the sizes overlap on the device while the
    requests of one size share every launch.  ``slots_for`` slots per size
    (fewer if the free device memory less the reserve does not hold them).
    Between blocks a finished slot is refilled from the requests of its size
    that have arrived, or parked when none has: a parked slot's Newton steps
    skip its GMRES and keep its state (the reference's slot mask,
    ``src/jaxrans/sweep.py:203-207``)."""
    t0 = clock()
    by_size: dict = {}
    for r in sorted(requests, key=lambda r: r.arrival):
        by_size.setdefault(r.cells, []).append(r)
    lock = threading.Lock()
    done: list[Outcome] = []
    errors: list = []

    def run_size(queue: list) -> None:
        first = queue[0]
        flow0 = request_flow(first, cfg)
        atlas0 = atlases[first.rid]
        probe = context_for(atlas0, flow0, cfg.numerics, cfg.implicit, n_cases=1)
        per_case = probe.device_bytes
        probe.close()
        with lock:
            free, _ = mem_info()
            n = int(max(1, min(slots_for(atlas0.cells, cfg), len(queue),
                               (free - cfg.reserve_bytes) // max(per_case, 1))))
            ctx = context_for(atlas0, flow0, cfg.numerics, cfg.implicit, n_cases=n)
        qs, mus = np.zeros((n, 5)), np.full(n, 0.0 if flow0.inviscid else flow0.mu_inf)
        q0 = q_infinity(flow0)
        qs[:, :q0.size] = q0
        slot: list = [None] * n
        parked = [False] * n
        waiting = list(queue)
        ctx.set_freestream(qs, None if flow0.inviscid else mus)
        shape = atlas0.shape
        ctx.set_state(np.broadcast_to(uniform_canvas(q0, shape, flow0.gamma), (n,) + ctx.shape[1:]))
        ctx.newton_init(np.full(n, cfg.implicit.cfl))
        try:
            while waiting or any(s is not None for s in slot):
                now = clock() - t0
                changed = False
                for k in range(n):
                    if slot[k] is None and waiting and waiting[0].arrival <= now:
                        req = waiting.pop(0)
                        flow = request_flow(req, cfg)
                        q = q_infinity(flow)
                        qs[k] = 0.0
                        qs[k, :q.size] = q
                        mus[k] = 0.0 if flow.inviscid else flow.mu_inf
                        ctx.set_geometry(k, atlases[req.rid])
                        ctx.set_state_case(k, uniform_canvas(q, shape, flow.gamma))
                        ctx.newton_init_case(k, cfg.implicit.cfl)
                        if parked[k]:
                            ctx.park_case(k, False)
                            parked[k] = False
                        slot[k] = Solve(req, ctx, now)
                        changed = True
                    elif slot[k] is None and not parked[k]:
                        ctx.park_case(k, True)
                        parked[k] = True
                if changed:
                    ctx.set_freestream(qs, None if flow0.inviscid else mus)
                if not any(s is not None for s in slot):
                    time.sleep(max(0.0, min(0.01, waiting[0].arrival - (clock() - t0))))
                    continue
                res = ctx.newton_block(cfg.block_iters)
                forces = ctx.forces(FORCE_REF[0], FORCE_REF[1])
                for k in range(n):
                    s = slot[k]
                    if s is None:
                        continue
                    if s.block_done(res[:, k:k + 1], forces[k:k + 1], cfg):
                        cl, cd, cm = s.coefficients[-1]
                        orders = (float(np.log10(s.peak / s.last_residual))
                                  if s.peak > 0 and s.last_residual > 0 else 0.0)
                        out = Outcome(s.req.rid, s.req.arrival, s.started, clock() - t0, s.iters, s.reason, cl, cd,
                                      cm, int(atlases[s.req.rid].cells), orders)
                        slot[k] = None
                        with lock:
                            done.append(out)
                            if on_done is not None:
                                on_done(out)
        finally:
            ctx.close()

    def work(queue):
        try:
            run_size(queue)
        except BaseException as e:          # reported by the serving thread
            with lock:
                errors.append(e)

    threads = [threading.Thread(target=work, args=(q,), name=f"slots-{size}", daemon=True)
               for size, q in by_size.items()]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return done


def serve_serial(requests: list[Request], atlases: dict, cfg: ServeConfig, clock=time.perf_counter,
                 on_done=None) -> list[Outcome]:
    """The same stream through the plain path: one solve at a time, in
    arrival order, each marched by blocking Newton blocks and a force call
    per block (what ``simulation.run`` does for one case).  Same block test,
    same outcomes per request; latency includes the wait in the queue."""
    t0 = clock()
    done = []
    for req in sorted(requests, key=lambda r: r.arrival):
        wait = req.arrival - (clock() - t0)
        if wait > 0:
            time.sleep(wait)
        s = start_solve(req, atlases[req.rid], cfg, clock() - t0, enqueue=False)
        march(s, cfg)
        done.append(finish_solve(s, clock() - t0))
        if on_done is not None:
            on_done(done[-1])
    return done


COUNTED = (ExitReason.CONVERGED, ExitReason.SETTLED)


def summary(outcomes: list[Outcome], wall: float) -> dict:
    """Throughput and latency of a served stream.  ``solves_per_min`` and
    the latency percentiles count the solves that ended converged or settled
    (``COUNTED``); ``finished_per_min`` counts every solve, whatever ended it."""
    counted = [o for o in outcomes if o.reason in COUNTED]
    lat = np.array([o.latency for o in counted]) if counted else np.array([np.nan])
    return {"requests": len(outcomes), "counted": len(counted), "wall_s": round(wall, 3),
            "solves_per_min": round(60.0 * len(counted) / wall, 3) if wall > 0 else None,
            "finished_per_min": round(60.0 * len(outcomes) / wall, 3) if wall > 0 else None,
            "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
            "latency_p95_s": round(float(np.percentile(lat, 95)), 3),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
            "reasons": {r: int(sum(o.reason == r for o in outcomes)) for r in sorted({o.reason for o in outcomes})},
            "newton_steps": int(sum(o.iterations for o in outcomes)),
            "cell_steps_per_s": round(sum(o.iterations * o.cells for o in outcomes) / wall, 1) if wall > 0 else None}


def default_config(precision: str = "mixed", viscous: bool = True, **kw) -> ServeConfig:
    """GMRES(30) x 2, line preconditioner, Roe + smooth van Albada (the
    benchmark case's numerics), relaxed updates for RANS."""
    num = NumericsConfig(flux="roe", limiter="van_albada_smooth", solver="implicit")
    imp = ImplicitConfig(krylov_subspace=30, max_relative_update=0.2 if viscous else None, precision=precision)
    return ServeConfig(numerics=num, implicit=imp, turbulence="sa_noft2" if viscous else None, **kw)


def with_precision(cfg: ServeConfig, precision: str) -> ServeConfig:
    return replace(cfg, implicit=replace(cfg.implicit, precision=precision))
