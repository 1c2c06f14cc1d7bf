"""Continuous-batching static solve server.

Requests are independent static Newton solves (a mesh, a material, one face
clamped and a traction or pressure load), each a full static_implicit.py:44-295 solve
of the reference.  The server keeps every admitted request's state on the
device (its DevicePlan: CSR, staging, displacement, Newton direction, PCG
vectors) and schedules at the granularity of one PCG chunk.  Each admitted
request is a state machine (tangent -> PCG -> line search -> tangent ...);
one *tick* of the server

* launches the tangent + residual of every request that needs one and the
  line-search evaluations (directional derivative, energies, step norms) of
  requests whose direction is ready, and starts the PCG of requests whose
  linear system is new,
* queues one asynchronous copy of all of the tick's results, including the
  status of the previous PCG chunk,
* queues ONE batched chunk of Jacobi-PCG iterations (``BatchedPCG``: a single
  launch per CG stage for all requests in the PCG stage, each stopping on its
  own tolerance), which the device runs while the host waits for the copy,
  admits arrivals and advances each request's state.

A request is never held back by another's PCG: a small request runs its
Newton iterations while a large one is still iterating.  Finished requests
leave (their memory returns to the budget) and arrived requests are admitted
in arrival order while the HBM budget and the batch-size cap allow.  Admission
estimates a request's device footprint from its mesh sizes before anything is
allocated; building the request (mesh, sparsity pattern, uploads) is host work
done while the device runs the tick's PCG chunk, for a bounded time per tick,
with its uploads on a side stream that the main stream then waits for; a built
request joins at the next tick.

Every per-request reduction has a fixed shape that depends only on that
request, and every decision is the request's own, so each result is
bit-identical whether the request runs alone or in any mix (the test suite
checks this).

The Newton iteration is that of ``StaticImplicitSolver`` (Armijo c1 = 0.3 with
halving, full step when the predicted energy decrease is below rounding,
relative residual test, round-off step test, linear-extrapolation predictor
over load increments); ``precision`` is 'fp64' (FP64 tangent and PCG) or
'mixed' (FP64 element arithmetic, tangent stored in FP32, FP64 PCG vectors).
"""
from __future__ import annotations

import concurrent.futures
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as nat
from . import meshgen
from ._native import F32, F64, call
from .assemble import loads, materials
from .solver._linear_solver import BatchedPCG, PcgVectors

_OPPOSITE = {'x0': 'x1', 'x1': 'x0', 'y0': 'y1', 'y1': 'y0', 'z0': 'z1', 'z1': 'z0'}
LOADED_SET = 'loaded'


# ---------------------------------------------------------------------- requests
@dataclass
class SolveRequest:
    """One static solve.

    spec keys: ``etype`` ('C3D4' Kuhn-split tets or 'C3D8' bricks), ``cells``
    (nx, ny, nz), ``size`` (Lx, Ly, Lz), ``material`` ('NH', mu, kappa) or
    ('MR', C10, C01, D1), ``fixed`` face name ('x0' .. 'z1'), ``load``
    ('pressure', p) on the five other faces or ('traction', (tx, ty, tz)) on
    the opposite face, ``load_steps`` equal load increments.
    """
    rid: int
    spec: dict
    arrival: float = 0.0
    # filled by the server
    status: str = 'pending'
    GC: object = None
    t_admit: float = None
    t_done: float = None
    newton_iters: int = 0
    cg_iters: int = 0
    footprint_est: int = 0
    footprint: int = 0
    error: str = ''

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    def n_elements(self):
        nx, ny, nz = self.spec['cells']
        return nx * ny * nz * (6 if self.spec['etype'] == 'C3D4' else 1)


def request_mesh(spec):
    """The request's block, centred on the coordinate origin, with the face set
    ``LOADED_SET`` the request's load acts on.

    A traction acts on the face opposite the clamped one.  A pressure acts on
    the five faces other than the clamped one: the follower-pressure functional
    V = 1/3 sum w det[x, x_xi, x_eta] (pressure.py:134-159) of an open surface
    adds line loads on the surface's edge (an in-plane p d / 3 per unit length,
    d the distance of the face plane from the origin, and a normal one that
    cancels 2/3 of the net force), which sharpen under refinement.  Closed by the
    clamped face, whose perimeter does not move, the surface's edge terms do no
    work and the load is the exact follower pressure.  With the origin inside the
    block no face plane passes through it, where the reference's form of the
    derivative (det A A^-T) is undefined."""
    nx, ny, nz = spec['cells']
    L = np.asarray(spec['size'], dtype=np.float64)
    if spec['etype'] == 'C3D4':
        mesh = meshgen.kuhn_tet_block(nx, ny, nz, size=tuple(spec['size']))
    elif spec['etype'] == 'C3D8':
        mesh = meshgen.hex_block(nx, ny, nz, size=tuple(spec['size']))
    else:
        raise ValueError('unsupported element type %r' % (spec['etype'],))
    mesh['nodes'] = mesh['nodes'] - 0.5 * L
    if spec['load'][0] == 'pressure':
        faces = [f for f in _OPPOSITE if f != spec['fixed']]
    else:
        faces = [_OPPOSITE[spec['fixed']]]
    mesh['face_sets'][LOADED_SET] = [entry for f in faces for entry in mesh['face_sets'][f]]
    return mesh


def request_model(spec):
    """FEAController of a request, built through the public API."""
    from .workloads import _controller_from_mesh
    mesh = request_mesh(spec)
    m = spec['material']
    if m[0] == 'NH':
        mat = materials.NeoHookean(m[1], m[2])
    elif m[0] == 'MR':
        mat = materials.MooneyRivlin(m[1], m[2], m[3])
    else:
        raise ValueError('unsupported material %r' % (m[0],))
    kind, value = spec['load']
    if kind == 'pressure':
        load = lambda nm: loads.Pressure(nm, LOADED_SET, float(value))
    elif kind == 'traction':
        load = lambda nm: loads.Traction(nm, LOADED_SET, tuple(float(v) for v in value))
    else:
        raise ValueError('unsupported load %r' % (kind,))
    return _controller_from_mesh(mesh, mat, fixed_set=spec['fixed'], load=load), mesh


def estimate_footprint(spec, precision='mixed', amg_min_dof=None):
    """Device bytes of a request before it is built (mesh sizes only).

    Counts the staging of every element's and loaded facet's node-pair blocks
    (the largest term), the CSR (neighbour count per node: 27 for bricks, 15
    for Kuhn tets), the per-element maps and the nodal / free-DOF vectors, and,
    when its 3N DOF reach amg_min_dof, the multigrid hierarchy and its PCG
    vectors (about 600 bytes per DOF: 550 measured on the 418k-DOF beam)."""
    nx, ny, nz = spec['cells']
    N = (nx + 1) * (ny + 1) * (nz + 1)
    cells = {'x': ny * nz, 'y': nx * nz, 'z': nx * ny}
    faces = [f for f in _OPPOSITE if f != spec['fixed']] if spec['load'][0] == 'pressure' else [_OPPOSITE[spec['fixed']]]
    if spec['etype'] == 'C3D8':
        E, nen, nb, S, snen = nx * ny * nz, 8, 27, sum(cells[f[0]] for f in faces), 4
    else:
        E, nen, nb, S, snen = 6 * nx * ny * nz, 4, 15, 2 * sum(cells[f[0]] for f in faces), 3
    nnz = 9 * nb * N
    kstage = E * nen * nen * 9 * 8 + S * snen * snen * 9 * 8
    maps = E * nen * nen * 4 + 2 * E * nen * 4 + S * snen * snen * 4 + 2 * S * snen * 4
    csr = nnz * (4 + (8 if precision == 'fp64' else 4)) + 4 * N * 3 + 3 * 4 * nb * N
    vectors = 3 * N * (8 * 12 + 4)
    amg = 600 * 3 * N if (amg_min_dof is not None and 3 * N >= amg_min_dof) else 0
    return int(kstage + maps + csr + vectors + amg + 64 * 1024)


def random_requests(n, qps, seed=0, e_range=(2e3, 2e5), aspect_range=(1.0, 4.0), load_steps=2):
    """A Poisson stream of ``n`` static solve requests at ``qps`` requests/s.

    Each request: C3D4 (Kuhn) or C3D8 block, element count log-uniform in
    ``e_range``, side ratios Lx/Lz and Ly/Lz uniform in ``aspect_range``;
    NeoHookean (mu, kappa) or Mooney-Rivlin (C10, C01, D1) with parameters
    uniform within +-50% of mu = 1, kappa = 100 (C10 = mu/2 split 5:1 with
    C01, D1 = 2/kappa); a random face clamped and either a compressive
    follower pressure 0.01 kappa (+-50%, a volume change of about 1%) on the
    five other faces or, on the opposite face, a transverse dead traction sized
    for a tip deflection of about 5% of the length, at most 0.03 mu (+-50%).
    """
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / qps, size=n)) if qps > 0 else np.zeros(n)
    t = t - t[0] if n else t
    out = []
    lo, hi = np.log(e_range[0]), np.log(e_range[1])
    for i in range(n):
        etype = 'C3D4' if rng.random() < 0.5 else 'C3D8'
        E = float(np.exp(rng.uniform(lo, hi)))
        cells = E / (6.0 if etype == 'C3D4' else 1.0)
        L = np.array([rng.uniform(*aspect_range), rng.uniform(*aspect_range), 1.0])
        h = (np.prod(L) / cells) ** (1.0 / 3.0)
        nc = tuple(int(max(1, round(l / h))) for l in L)
        mu = rng.uniform(0.5, 1.5)
        kappa = 100.0 * rng.uniform(0.5, 1.5)
        if rng.random() < 0.5:
            mat = ('NH', mu, kappa)
        else:
            mat = ('MR', 0.5 * mu * 5.0 / 6.0, 0.5 * mu / 6.0, 2.0 / kappa)
        fixed = ['x0', 'x1', 'y0', 'y1', 'z0', 'z1'][int(rng.integers(6))]
        ax = 'xyz'.index(fixed[0])
        if rng.random() < 0.5:
            load = ('pressure', 0.01 * kappa * rng.uniform(0.5, 1.5))
        else:
            others = [a for a in range(3) if a != ax]
            d_ax = min(others, key=lambda a: L[a])          # bend about the thinner side
            E_y = 9.0 * kappa * mu / (3.0 * kappa + mu)
            tmag = min(0.05 * E_y * L[d_ax] ** 2 / (4.0 * L[ax] ** 2), 0.03 * mu) * rng.uniform(0.5, 1.5)
            tv = [0.0, 0.0, 0.0]
            tv[d_ax] = tmag * (1.0 if rng.random() < 0.5 else -1.0)
            load = ('traction', tuple(tv))
        spec = {'etype': etype, 'cells': nc, 'size': tuple(float(v) for v in L), 'material': mat,
                'fixed': fixed, 'load': load, 'load_steps': int(load_steps)}
        out.append(SolveRequest(rid=i, spec=spec, arrival=float(t[i])))
    return out


# ---------------------------------------------------------------------- device state
class _Reads:
    """Device float64 tensors collected for one host read (one synchronisation)."""

    def __init__(self):
        self.items = []

    def add(self, t):
        self.items.append(t.reshape(-1))
        return len(self.items) - 1

    def fetch(self):
        return self.fetch_async().wait()

    def fetch_async(self):
        """Queue one copy of every collected tensor into pinned host memory; ``wait()``
        on the result blocks until that copy (not later work) is done."""
        return _Pending(self)


class _Pending:
    def __init__(self, reads):
        self.sizes = [t.numel() for t in reads.items]
        self.host = None
        if self.sizes:
            dev = torch.cat(reads.items)
            self.host = torch.empty(dev.numel(), dtype=dev.dtype, pin_memory=True)
            self.host.copy_(dev, non_blocking=True)
            self.event = torch.cuda.Event()
            self.event.record()
        reads.items = []

    def wait(self):
        if self.host is None:
            return []
        self.event.synchronize()
        host = self.host.numpy()
        out, o = [], 0
        for s in self.sizes:
            out.append(host[o:o + s])
            o += s
        return out


@dataclass
class _Slot:
    req: SolveRequest
    fe: object
    plan: object
    U: torch.Tensor
    U_prev: torch.Tensor
    Unew: torch.Tensor
    dGC: torch.Tensor
    vec: PcgVectors
    r_ref: float
    k: int = 1
    steps: int = 1
    it_total: int = 0
    low_alpha: int = 0
    alpha: float = 0.0
    R: torch.Tensor = None
    K: torch.Tensor = None
    ls: dict = field(default_factory=dict)
    state: str = 'tangent'
    max_iter: int = 0
    rd: object = None
    amg: object = None          # multigrid hierarchy (requests of at least amg_min_dof free DOF)
    apcg: object = None         # its PCG workspace (float64 vectors)
    amg_job: object = None      # (future of its host set-up, submit time) until taken over

    @property
    def scale(self):
        return self.k / self.steps


def _stats_launch(plan, x, reads):
    """(sum x^2, max |x|) partials of a float64 device vector, queued for a read."""
    n = x.numel()
    nb = nat.lib().fea_vec_blocks(n)
    part = torch.empty(2 * nb, dtype=torch.float64, device=plan.device)
    call('fea_vec_stats', n, x, part, nat.stream_ptr())
    return reads.add(part)


def _stats_value(h):
    v = h.reshape(-1, 2)
    return float(v[:, 0].sum()), float(v[:, 1].max())


def _energy_launch(plan, U, reads):
    return reads.add(plan.energy_terms(U, 1.0).clone())


def _energy_value(plan, terms, scale):
    E = 0.0
    for gi, g in enumerate(plan.groups):
        if g.n == 0:
            continue
        if g.kind == 'elem':
            E += terms[gi]
        elif g.kind == 'pressure':
            E += float(g.load.pressure) * scale * terms[gi]
        elif g.kind == 'traction':
            E -= scale * terms[gi]
    return float(E)


# ---------------------------------------------------------------------- server
class SolveServer:
    """Continuous-batching server of static solves on one GPU.

    hbm_budget: bytes of device memory admitted requests may hold (estimated
    before admission); max_batch: most requests in flight; admit_budget: host
    seconds per tick spent building arrived requests (at least one is built);
    the remaining arguments are those of ``StaticImplicitSolver``.
    """

    def __init__(self, precision='mixed', hbm_budget=None, max_batch=64, rel_tol=1e-10, step_tol=1e-12,
                 cg_tol=1e-8, cg_max_iter=20000, max_newton=200, chunk=32, use_graph=True,
                 armijo_slack=1e-12, verbose=False, predictor=True, preconditioner='block', admit_budget=0.004,
                 amg_min_dof=30000, amg_chunk=8, amg_threads=None, amg_builders=1):
        nat.require_gpu()
        if precision not in ('fp64', 'mixed'):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        self.mat_dtype = F64 if precision == 'fp64' else F32
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.rel_tol, self.step_tol = float(rel_tol), float(step_tol)
        self.cg_tol, self.cg_max_iter = float(cg_tol), int(cg_max_iter)
        self.max_newton = int(max_newton)
        self.armijo_slack = float(armijo_slack)
        self.predictor = bool(predictor)
        self.verbose = verbose
        # node-block Jacobi by default (point Jacobi, the reference's scaling: 'jacobi')
        self.pcg = BatchedPCG(self.mat_dtype, F64, 'cuda', chunk=chunk, use_graph=use_graph, precond=preconditioner)
        # requests of at least amg_min_dof free DOF solve their linear systems with the
        # multigrid-preconditioned PCG (amg_chunk iterations per tick) instead of joining the
        # batch; None: every request joins the batch
        self.amg_min_dof = None if amg_min_dof is None else int(amg_min_dof)
        self.amg_chunk = int(amg_chunk)
        # host threads of a hierarchy's set-up: by default half the cores less two, so the tick
        # loop keeps its core (configured 16-request stream on 16 cores: 6 threads 3.33 solves/s,
        # 14 threads 3.13)
        self.amg_builders = max(1, int(amg_builders))
        self.amg_threads = amg_threads or max(1, len(os.sched_getaffinity(0)) // (2 * self.amg_builders) - 2)
        self.amg_setup_s = 0.0
        self._pool = None
        self.active: list[_Slot] = []
        self._side = torch.cuda.Stream()     # admission uploads (default priority)
        self._retired = []
        self.admit_budget = float(admit_budget)
        self.used = 0
        self.rounds = 0
        self.batch_sizes = []
        self.peak_used = 0

    # -------------------------------------------------------------- admission
    def _admit(self, req):
        """Build the request's plan and read its full-load residual norm (the reference
        of the relative residual test); returns the slot.  Runs with the admission stream
        current, so it waits for none of the main stream's work."""
        fe, _ = request_model(req.spec)
        fe.assembly.initialize()
        plan = fe.assembly.plan
        dev = plan.device
        U = torch.zeros(3 * plan.n_nodes, dtype=torch.float64, device=dev)
        slot = _Slot(req=req, fe=fe, plan=plan, U=U, U_prev=U.clone(), Unew=torch.empty_like(U),
                     dGC=torch.zeros(plan.n_free, dtype=torch.float64, device=dev),
                     vec=PcgVectors(plan.n_free, F64, dev, n_groups=plan.n_groups), r_ref=0.0,
                     steps=int(req.spec.get('load_steps', 1)))
        rd = _Reads()
        _stats_launch(plan, plan.residual(U, 1.0, F64), rd)
        rss, _ = _stats_value(rd.fetch()[0])          # waits for this stream's work only
        slot.r_ref = math.sqrt(rss)
        if self._uses_amg(plan.n_free) and plan.grp is not None and plan.n_groups != 0:
            # the multigrid hierarchy from the float64 tangent at u = 0 under the full load (as
            # the static solver builds it): its host set-up runs on the builder thread while
            # the request starts with the batched PCG; the tick loop takes it over once ready
            from .solver.amg import HostBuild
            _, K = plan.tangent(U, 1.0, F64)
            hb = HostBuild(plan, K, threads=self.amg_threads)
            del K
            slot.amg_job = (self._builder().submit(self._timed_build, hb), time.perf_counter())
        req.footprint = self._actual_footprint(slot)
        return slot

    @staticmethod
    def _timed_build(hb):
        t0 = time.perf_counter()
        h = hb.run()
        return h, time.perf_counter() - t0

    def _builder(self):
        if self._pool is None:
            self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=self.amg_builders)
        return self._pool

    def _take_amg(self, s):
        """Once its host set-up has finished, move the request's hierarchy to the device and
        give it a multigrid PCG workspace (float64 vectors)."""
        job = s.amg_job
        if job is None or not job[0].done():
            return
        from .solver.amg import AMG
        from .solver._linear_solver import PCG
        s.amg_job = None
        h, secs = job[0].result()
        p = s.plan
        s.amg = AMG(p, None, handle=h)
        self.amg_setup_s += secs + s.amg.setup_time
        s.apcg = PCG(p.n_free, p.nnz, p.rowptr, p.col, p.diag_slot, self.mat_dtype, p.device, vec_dtype=F64,
                     grp=p.grp, n_groups=p.spmv_groups, chunk=self.amg_chunk, precond='amg', spmv_col=p.spmv_col,
                     amg=s.amg)

    def _drop_amg(self, s):
        """A retired request's hierarchy still on the builder: free its handle when done."""
        job, s.amg_job = s.amg_job, None
        if job is not None:
            from .solver.amg import destroy_handle
            job[0].add_done_callback(lambda f: destroy_handle(f.result()[0]))

    def _uses_amg(self, n_free):
        return self.amg_min_dof is not None and n_free >= self.amg_min_dof

    def _actual_footprint(self, s):
        p = s.plan
        ts = [p.X, p.dof_map, p.rowptr, p.col, p.diag_slot, p.grp, p.blk_a, p.blk_b, p.blk_colpos, p.kseg,
              p.kseg_rest, p.rseg, p.rstage, p.R, s.U, s.U_prev, s.Unew, s.dGC]
        for g in p.groups:
            ts += [g.conn, g.dst_k, g.dst_r, g.colpos, g.order, g.dst_k_rest]
        nb = sum(t.numel() * t.element_size() for t in ts if t is not None)
        # F64 staging (the non-element groups' only, under direct assembly), CSR values
        stage = p.n_kstage_rest if p.direct_assembly else p.n_kstage
        nb += 9 * stage * 8 + p.nnz * (8 if self.precision == 'fp64' else 4)
        if s.amg is not None:
            a = s.apcg
            nb += s.amg.nbytes() + sum(t.numel() * t.element_size() for t in (a.x, a.r, a.z, a.p, a.q, a.b, a.dinv,
                                                                               a.ws))
        return int(nb + s.vec.nbytes())

    def _retire(self, slot, status, now, reason=''):
        req = slot.req
        req.status = status
        req.error = reason
        req.t_done = now
        req.newton_iters = slot.it_total
        if status == 'done':
            # an asynchronous copy: complete before the next read (and before run returns)
            gc = slot.plan.u_to_gc(slot.U)
            req.GC = torch.empty(gc.shape, dtype=gc.dtype, pin_memory=True)
            req.GC.copy_(gc, non_blocking=True)
        self.used -= req.footprint_est
        self._drop_amg(slot)
        self.active.remove(slot)
        # its buffers came from the admission stream; they are released after the next
        # read, when no queued main-stream work can still touch them
        self._retired.append(slot)

    # -------------------------------------------------------------- scheduling
    # Every admitted request is a small state machine advanced one stage per tick:
    #   'tangent'  -> tangent + residual norm          (read) -> 'tangent' | 'pcg_new' | done
    #   'pcg_new'  -> joins the batched PCG (start)    -> 'pcg'
    #   'pcg'      -> one batched chunk of iterations  (read) -> 'pcg' | 'ls'
    #   'ls'       -> direction.R, energies, step norms (read) -> 'tangent' | 'trial'
    #   'trial'    -> backtracking trial energy        (read) -> 'tangent' | 'trial'
    # A tick launches the work of every request, then reads all results at once.
    def _tangent(self, s):
        p = s.plan
        if self.precision == 'fp64':
            s.R, s.K = p.tangent(s.U, s.scale, F64)
        else:
            s.R, s.K = p.tangent(s.U, s.scale, F64, val_dtype=F32)

    def _advance_increment(self, s, now):
        """Increment converged: finish, or move to the next load level with the predictor."""
        if s.k >= s.steps:
            self._retire(s, 'done', now())
            return
        if self.predictor:
            # U <- 2U - U_prev, U_prev <- old U (linear extrapolation of converged states)
            step = s.U - s.U_prev
            s.U_prev.copy_(s.U)
            s.U.add_(step)
        s.k += 1
        s.state = 'tangent'

    def _pcg_entry(self, s):
        return (s.plan, s.K, s.vec, self.cg_tol, s.max_iter)

    def _launch(self, reads):
        """Queue the device work of every admitted request's current stage; the
        results land on ``reads``.  Returns what ``_decide`` needs."""
        by = {}
        for s in self.active:
            by.setdefault(s.state, []).append(s)
        for s in by.get('tangent', []):
            self._tangent(s)
            s.rd = _stats_launch(s.plan, s.R, reads)
        for s in by.get('ls', []):
            p = s.plan
            p.dof_axpy(s.U, 1.0, s.dGC, s.Unew)
            s.rd = (reads.add(p.dot(s.dGC, s.R).clone()), _energy_launch(p, s.U, reads),
                    _energy_launch(p, s.Unew, reads), _stats_launch(p, s.dGC, reads),
                    _stats_launch(p, s.Unew, reads))
        for s in by.get('trial', []):
            st = s.ls
            s.plan.dof_axpy(s.U, st['sign'] * st['alpha'], s.dGC, s.Unew)
            s.rd = _energy_launch(s.plan, s.Unew, reads)
        # status of the chunk queued at the end of the previous tick (complete in stream
        # order before anything queued now); a start does not touch it
        prev, prev_rd = self._chunk, None
        if prev:
            prev_rd = reads.add(self.pcg.status(len(prev)))
        prev_amg = [(s, reads.add(s.apcg.ws[:2].clone())) for s in self._chunk_amg]
        new = by.get('pcg_new', [])
        for s in new:
            self._take_amg(s)
        self.pcg.start([self._pcg_entry(s) for s in new if s.amg is None])
        for s in new:
            if s.amg is not None:
                # multigrid PCG: node-block inverses and lambda_max of the new tangent, then
                # r, z, p from b = -R and the warm start
                a = s.apcg
                a.set_matrix(s.K)
                s.amg.pcg_start(self.mat_dtype, s.K, a.dinv, s.vec.b, s.vec.x, a.r, a.z, a.p, a.ws)
            s.state = 'pcg'
        return by, prev, prev_rd, prev_amg

    def _launch_chunk(self):
        """One batched chunk of PCG iterations for every request in its PCG stage.  It is
        queued after the tick's read, so the device runs it while the host decides; its
        status is read at the next tick.  A request whose PCG finished in the previous
        chunk is still listed here (its result is not read yet): the chunk skips it."""
        in_pcg = [s for s in self.active if s.state == 'pcg' and s.amg is None]
        if in_pcg:
            self.pcg.iterate([self._pcg_entry(s) for s in in_pcg])
        self._chunk = in_pcg
        in_amg = [s for s in self.active if s.state == 'pcg' and s.amg is not None]
        for s in in_amg:
            a = s.apcg
            s.amg.pcg_iterate(self.mat_dtype, s.K, a.dinv, s.vec.x, a.r, a.z, a.p, a.q, a.ws, self.cg_tol,
                              s.max_iter, self.amg_chunk, None)
        self._chunk_amg = in_amg

    def _decide(self, tick, host, now):
        """Advance every request's state machine from the tick's results."""
        by, prev, prev_rd, prev_amg = tick
        for s in by.get('tangent', []):
            rss, _ = _stats_value(host[s.rd])
            if s.it_total >= self.max_newton or not math.isfinite(rss):
                self._retire(s, 'failed', now(), 'Newton limit %d reached or residual not finite '
                             '(increment %d, |R| %.3e)' % (self.max_newton, s.k, math.sqrt(rss)))
                continue
            if self.verbose:
                print('req %d inc %d it %d |R|/|R0| %.3e alpha %.3g' % (s.req.rid, s.k, s.it_total,
                      math.sqrt(rss) / s.r_ref, s.alpha), flush=True)
            if math.sqrt(rss) <= self.rel_tol * s.r_ref:
                self._advance_increment(s, now)
                continue
            # K dGC = -R, warm start from the previous direction (static_implicit.py:297-350)
            s.vec.b.copy_(s.R).neg_()
            s.vec.x.copy_(s.dGC)
            s.max_iter = 4 * self.cg_max_iter if s.it_total % 20 == 0 or s.low_alpha > 0 else self.cg_max_iter
            s.state = 'pcg_new'
        if prev:
            status = host[prev_rd].reshape(-1, 2)
            for s, (its, done) in zip(prev, status):
                if s.state != 'pcg':         # its finish was read at the previous tick
                    continue
                if done != 0.0:
                    s.dGC.copy_(s.vec.x)
                    s.req.cg_iters += int(its)
                    s.it_total += 1
                    s.state = 'ls'
        for s, i in prev_amg:
            its, done = host[i][0], host[i][1]
            if s.state == 'pcg' and done != 0.0:
                s.dGC.copy_(s.vec.x)
                s.req.cg_iters += int(its)
                s.it_total += 1
                s.state = 'ls'
        for s in by.get('ls', []):
            p = s.plan
            i_dd, i_e0, i_e1, i_d, i_u = s.rd
            deltaE = float(host[i_dd][0])
            e0 = _energy_value(p, host[i_e0], s.scale)
            sign = 1.0
            if deltaE > 0:
                sign, deltaE = -1.0, -deltaE
            if not math.isfinite(deltaE):
                self._retire(s, 'failed', now(), 'Newton direction not finite')
                continue
            s.ls = {'e0': e0, 'deltaE': deltaE, 'sign': sign, 'alpha': 1.0,
                    'dmax': _stats_value(host[i_d])[1], 'umax': _stats_value(host[i_u])[1]}
            if abs(deltaE) <= 1e-13 * abs(e0):
                # predicted decrease below the energy's rounding: full Newton step
                if sign < 0:
                    p.dof_axpy(s.U, -1.0, s.dGC, s.Unew)
                self._accept(s, now)
            elif sign > 0:
                self._verdict(s, self._armijo(s, _energy_value(p, host[i_e1], s.scale)), now)
            else:
                s.state = 'trial'
        for s in by.get('trial', []):
            self._verdict(s, self._armijo(s, _energy_value(s.plan, host[s.rd], s.scale)), now)

    def _verdict(self, s, verdict, now):
        if verdict == 'accept':
            self._accept(s, now)
        elif verdict == 'fail':
            self._retire(s, 'failed', now(), 'line search found no decrease')
        else:
            s.state = 'trial'

    def _armijo(self, s, e_new):
        """Armijo test of the trial at s.ls['alpha'] (c1 = 0.3); halves alpha on rejection."""
        st = s.ls
        slack = self.armijo_slack * abs(st['e0'])
        if math.isfinite(e_new) and e_new <= st['e0'] + 0.3 * st['deltaE'] * st['alpha'] + slack:
            return 'accept'
        st['alpha'] *= 0.5
        return 'fail' if st['alpha'] < 1e-12 else 'retry'

    def _accept(self, s, now):
        """U <- Unew (= U + sign alpha dGC); low-alpha bookkeeping; round-off step test."""
        st = s.ls
        alpha = st['alpha']
        s.U.copy_(s.Unew)
        s.alpha = alpha
        s.low_alpha = s.low_alpha + 1 if alpha < 0.01 else max(0, s.low_alpha - 5)
        s.state = 'tangent'
        if s.low_alpha > 10:
            self._retire(s, 'failed', now(), 'more than 10 steps with alpha < 0.01')
            return
        if alpha == 1.0 and st['sign'] > 0 and st['dmax'] <= self.step_tol * st['umax']:
            self._advance_increment(s, now)

    # -------------------------------------------------------------- driver
    def run(self, requests, realtime=True, on_round=None):
        """Serve ``requests`` (arrival offsets in seconds from the start).

        realtime: a request is admissible once its arrival time has passed on
        the wall clock (the server idles until the next arrival when nothing
        is in flight); otherwise every request is available at once.  Returns
        the requests with status, result GC (host), timings and counts."""
        queue = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        if not realtime:
            for r in queue:
                r.arrival = 0.0
        qi = 0
        main = torch.cuda.current_stream()
        self._chunk = []
        self._chunk_amg = []
        while qi < len(queue) or self.active:
            # one tick: the stage work of every request, one asynchronous read of its
            # results, then the next PCG chunk, which the device runs while the host admits
            # arrivals and decides
            reads = _Reads()
            tick = None
            if self.active:
                self.batch_sizes.append(len(self.active))
                tick = self._launch(reads)
            pending = reads.fetch_async()
            self._launch_chunk()
            # admission in arrival order while slots and HBM budget allow, for at most
            # admit_budget seconds per tick (at least one request): building a request is
            # host work that overlaps the chunk; its uploads go on a side stream that the
            # main stream then waits for, and it joins at the next tick
            admitted = []
            t = now()
            with torch.cuda.stream(self._side):
                while (qi < len(queue) and queue[qi].arrival <= t
                       and len(self.active) + len(admitted) < self.max_batch
                       and (not admitted or now() - t < self.admit_budget)):
                    r = queue[qi]
                    est = estimate_footprint(r.spec, self.precision, self.amg_min_dof)
                    if (self.active or admitted) and self.used + est > self.hbm_budget:
                        break
                    qi += 1
                    if est > self.hbm_budget:
                        r.status, r.error, r.t_done = 'rejected', 'footprint %d B exceeds the budget' % est, now()
                        continue
                    r.footprint_est, r.t_admit, r.status = est, now(), 'running'
                    admitted.append(self._admit(r))
                    self.used += est
                    self.peak_used = max(self.peak_used, self.used)
            if admitted:
                main.wait_stream(self._side)
            if tick is None and not admitted:
                if qi < len(queue):
                    time.sleep(max(0.0, min(queue[qi].arrival - now(), 0.05)))
                continue
            host = pending.wait()
            self._retired = []
            if tick is not None:
                self._decide(tick, host, now)
                self.rounds += 1
                if on_round is not None:
                    on_round(self)
            self.active.extend(admitted)
        torch.cuda.synchronize()
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None
        return requests


def solve_one_at_a_time(requests, precision='fp64', rel_tol=1e-10, step_tol=1e-12, cg_tol=1e-8,
                        cg_max_iter=20000, max_newton=200, armijo_slack=1e-12, predictor=True,
                        preconditioner='jacobi'):
    """The plain path: each request built and solved alone by ``StaticImplicitSolver``,
    in arrival order (``preconditioner`` as there: 'jacobi' is the reference's PCG).
    Returns (GC list, per-request seconds including the model build and, for 'amg',
    the multigrid set-up, FIFO latencies under the requests' arrival times)."""
    from .solver import StaticImplicitSolver
    out, secs = [], []
    for r in sorted(requests, key=lambda r: (r.arrival, r.rid)):
        t0 = time.perf_counter()
        fe, _ = request_model(r.spec)
        fe.solver = StaticImplicitSolver(maximum_iteration=max_newton, precision=precision, rel_tol=rel_tol,
                                         step_tol=step_tol, load_steps=int(r.spec.get('load_steps', 1)),
                                         cg_tol=cg_tol, cg_max_iter=cg_max_iter, armijo_slack=armijo_slack,
                                         predictor=predictor, preconditioner=preconditioner)
        fe.initialize()
        GC = fe.solver.solve(GC0=torch.zeros_like(fe.assembly.GC))
        torch.cuda.synchronize()
        secs.append(time.perf_counter() - t0)
        out.append((r.rid, None if GC is False else GC.cpu()))
        del fe
    order = sorted(requests, key=lambda r: (r.arrival, r.rid))
    end, lat = 0.0, []
    for r, t in zip(order, secs):
        end = max(end, r.arrival) + t
        lat.append(end - r.arrival)
    return dict(out), secs, lat
