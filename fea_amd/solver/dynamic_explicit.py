"""Central-difference explicit dynamics (reference: solver/dynamic_explicit.py:10-166).

Each step is one residual pass followed by the fused gather + update kernel
(a = -M^-1 R, v_{n+1/2} = v_{n-1/2} + dt a, u_{n+1} = u_n + dt v_{n+1/2}); the lumped
mass is the reference's row-sum of the consistent mass restricted to free DOFs
(dynamic_explicit.py:147-166).  When every group is an element group the residual
is reduced per tile of Morton-ordered elements in LDS (one float64 partial per tile
and node, about two per node) instead of staged per element node; node-to-surface
contact forces come from candidate lists searched with radius + skin and reused
while no contact node has moved skin / 2 (the host polls the displacement every few
steps and rebuilds; the kernels fall back to full scans when a rebuild comes late),
in three launches per step.

Additions: ``dt`` (fixed step) or ``cfl_factor`` (dt = factor * 2 / omega_max,
the central difference's stability limit for the discrete system, omega_max^2
the largest eigenvalue of M^-1 (K + K_contact) by power iteration on the
device); the reference uses the constant 0.8 * 5e-6 (dynamic_explicit.py:54-64).
``precision='fp32'`` runs the element force arithmetic in FP32 (the nodal
state, the update and the energy balance stay FP64).  ``overlap`` forms the
contact forces on a side stream during the element pass and ``use_graph`` replays
chunks of ``chunk`` steps from a captured HIP graph; both give the same results
bit for bit and, on the two-block benchmark, no speed-up (the element pass fills
the GPU, and the host enqueues a step faster than the GPU runs it), so both are off
by default.
"""
from __future__ import annotations

import ctypes
import time
from types import SimpleNamespace

import numpy as np
import torch

from .. import _native as nat
from .._native import F32, F64, call, ptr
from .basesolver import BaseSolver


class DynamicExplicitSolver(BaseSolver):

    def __init__(self, time_end: float = 1.0, time_per_storage: float = 1e-4, *, dt: float | None = None,
                 cfl_factor: float | None = None, verbose: bool = False, store: bool = True,
                 energy_interval: int | None = None, precision: str = 'fp64', overlap: bool = False,
                 use_graph: bool = False, chunk: int = 8, fuse_contact: bool = True) -> None:
        super().__init__()
        self.overlap = bool(overlap)
        self.use_graph = bool(use_graph)
        # with tiles and one contact set of reusable lists, a step is two launches (element
        # pass + contact penalty; node update + contact-node forces + tracking): the same bits
        # as the five launches of the separate path
        self.fuse_contact = bool(fuse_contact)
        # element tiles read their nodes from LDS: bit-identical, measured slower on the
        # two-block benchmark (15.8 vs 14.7 us per block), so off by default; the fused
        # contact step (fuse_contact) does not stage
        self.stage_nodes = False
        self.chunk = int(chunk)
        if precision not in ('fp64', 'fp32'):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        self._time_end = time_end
        self._time_per_storage = time_per_storage
        self._next_storage_time = 0.0
        self._GC_list = None
        self._GV_list = None
        self._GA_list = None
        self._time_list = []
        self._deltaT = 0.0
        self._dt_fixed = dt
        self._cfl = cfl_factor
        self.verbose = verbose
        self.store = store
        self.steps_done = 0
        self._contacts = []
        self.energy_interval = energy_interval
        self.energy_history = []

    def initialize(self, assembly):
        super().initialize(assembly=assembly)
        self.assembly.initialize_dynamic()
        self._GC_list, self._GV_list, self._GA_list, self._time_list = [], [], [], []
        self._next_storage_time = 0.0
        p = self.assembly.plan
        from ..assemble.loads.contact import NodeToSurfaceContact
        self._contacts = [l for l in self.assembly._loads.values() if isinstance(l, NodeToSurfaceContact)]
        self._fc_sum = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device=p.device) if len(self._contacts) > 1 else None
        m = p.lumped_mass()
        m = torch.where(m.abs() < 1e-12, torch.ones_like(m), m)     # dynamic_explicit.py:165
        self._minv = 1.0 / m
        self._minv_node = self._node_minv()
        self._setup_tiles()
        self._deltaT = self.estimate_critical_timestep()
        if self._deltaT <= 0:
            raise ValueError("Critical timestep must be positive.")

    def _setup_tiles(self):
        """Per-tile force reduction for the step loop when every group of the plan is an
        element group (fea_elem_residual_tiles + fea_explicit_update_tiles): each workgroup
        sums its Morton-ordered tile's element forces per node in LDS and writes one float64
        partial per (tile, node); the update adds a node's partials.  Plans with surface-load
        groups keep the node-major staging.  Launch arguments are bound once (the buffers
        never move)."""
        from ._explicit_tiles import group_tiles, node_partials
        self._tiles = None
        p = self.assembly.plan
        groups = [g for g in p.groups if g.n > 0]
        if not groups or any(g.kind != 'elem' for g in groups):
            return
        X = p.X.reshape(-1, 3).cpu().numpy()
        free = (p.dof_map.reshape(-1, 3).cpu().numpy() >= 0).any(axis=1)
        L = nat.lib()
        # groups of one element type and material are tiled, and launched, as one: a launch of
        # twice the workgroups overlaps one half's node loads with the other's arithmetic
        merged = {}
        for g in groups:
            key = (g.ntype, np.asarray(g.params, dtype=np.float64).tobytes())
            if key in merged:
                m = merged[key]
                m['conn_np'] = np.concatenate([m['conn_np'], g.conn_np])
            else:
                merged[key] = {'ntype': g.ntype, 'params': g.params, 'conn_np': g.conn_np}
        groups = [SimpleNamespace(n=int(m['conn_np'].shape[0]), **m) for m in merged.values()]
        tabs, base, qoff = [], 0, 0
        for g in groups:
            t = group_tiles(g.conn_np, X, free, L.fea_elem_tile_size(g.ntype), base)
            t['pslot_ptr'] = t['pslot_ptr'] + qoff
            base += t['pnode'].shape[0]
            qoff += t['pslot'].shape[0]
            tabs.append(t)
        if qoff >= 2 ** 31 or 3 * base >= 2 ** 31:
            return
        dev = p.device
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)
        pslot_ptr = i32(np.concatenate([t['pslot_ptr'][:-1] for t in tabs] + [[qoff]]))
        pslot = torch.as_tensor(np.concatenate([t['pslot'] for t in tabs]).view(np.int16), device=dev)
        pptr, plist = node_partials(np.concatenate([t['pnode'] for t in tabs]), p.n_nodes)
        keep = {'pslot_ptr': pslot_ptr, 'pslot': pslot, 'node_pptr': i32(pptr), 'node_plist': i32(plist),
                'partial': torch.zeros(3 * max(1, base), dtype=torch.float64, device=dev), 'conn': [], 'lbase': []}
        code = F32 if self.precision == 'fp32' else F64
        elem_args = []
        for g, t in zip(groups, tabs):
            conn = i32(g.conn_np[t['order']])
            lb = i32(t['tile_lbase'])
            keep['conn'].append(conn)
            keep['lbase'].append(lb)
            # the tile's nodes staged in LDS when, with the forces' LDS, a workgroup stays within
            # 64 KB; otherwise each element gathers its nodes
            tsz = L.fea_elem_tile_size(g.ntype)
            lds = tsz * g.ntype * 3 * (4 if code == F32 else 8) + 48 * t['max_tile_nodes']
            stage = None
            if self.stage_nodes and 0 < t['max_tile_nodes'] <= 2048 and lds <= 65536:
                stage = (i32(t['tnode_ptr']), i32(t['tnode']),
                         torch.as_tensor(t['lconn'].view(np.int16), device=dev))
                keep.setdefault('stage', []).append(stage)
            elem_args.append((g.ntype, code, g.n, conn.data_ptr(), p.X.data_ptr(), g.params.ctypes.data,
                              lb.data_ptr(), pslot_ptr.data_ptr(), pslot.data_ptr(), keep['partial'].data_ptr(),
                              stage[0].data_ptr() if stage else None, stage[1].data_ptr() if stage else None,
                              stage[2].data_ptr() if stage else None, t['max_tile_nodes'] if stage else 0))
        keep['params'] = [g.params for g in groups]
        keep['elem_args'] = elem_args
        keep['n_partials'] = base
        self._tiles = keep

    def estimate_critical_timestep(self, safety_factor=0.8) -> float:
        if self._dt_fixed is not None:
            return float(self._dt_fixed)
        if self._cfl is None:
            return safety_factor * 5e-6                             # dynamic_explicit.py:54-64
        return self._cfl * self._cfl_limit()

    def _cfl_limit(self, iters=80):
        """Stability limit 2 / omega_max of the central difference for the discrete
        system: omega_max^2 = largest eigenvalue of M^-1 (K + K_c), with K the
        tangent at the reference state, M the lumped mass and K_c the penalty
        stiffness of the contact slaves (penalty * tributary area on each of
        their DOFs).  Power iteration in the M inner product, whose Rayleigh
        quotients rise to omega_max^2 from below; the last one is raised by 2%."""
        p = self.assembly.plan
        dev = p.device
        _, K = p.tangent(torch.zeros(3 * p.n_nodes, dtype=torch.float64, device=dev), 1.0, F64)
        K = K.clone()
        m = 1.0 / self._minv
        kc = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device=dev)
        for c in self._contacts:
            nodes = torch.as_tensor(np.asarray(c.slave_nodes, dtype=np.int64), device=dev)
            area = torch.as_tensor(np.asarray(c.slave_area, dtype=np.float64), device=dev)
            for j in range(3):
                kc[3 * nodes + j] += c.penalty * area
        kc = p.u_to_gc(kc)
        v = torch.rand(p.n_free, dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(0))
        w = torch.empty_like(v)
        lam = 0.0
        for _ in range(iters):
            p.spmv(K, v, w, F64)
            w.add_(kc * v)
            lam = float(torch.dot(v, w) / torch.dot(v, m * v))
            v = w / m
            v = v / torch.linalg.vector_norm(v)
        return 2.0 / np.sqrt(1.02 * lam)

    def solve(self, GC0=None, GV0=None, *args, **kwargs) -> bool:
        """dynamic_explicit.py:66-145."""
        t_start = time.time()
        p = self.assembly.plan
        if GC0 is None:
            GC0 = self.assembly.GC.clone()
        if GV0 is None:
            GV0 = torch.zeros_like(self.assembly.GC)
        U = p.gc_to_u(torch.as_tensor(GC0, dtype=torch.float64, device=p.device).contiguous())
        vh = torch.as_tensor(GV0, dtype=torch.float64, device=p.device).clone()
        acc = torch.zeros_like(vh)
        dt = self._deltaT
        nsteps = int(np.ceil(self._time_end / dt - 1e-12))
        if self.store:
            self._GC_list.append(GC0)
            self._GV_list.append(GV0)
            self._time_list.append(0.0)
        t = 0.0
        self.energy_history = []
        if self.energy_interval:
            self.energy_history.append((0.0, self.energy_balance(U, vh, None, dt)))
        step = 0
        while step < nsteps:
            # steps up to the next event: a contact check (every 256 steps), an energy
            # record, a stored state (t accumulated as below) or the end
            m, tt = 0, t
            while True:
                m += 1
                tt += dt
                k = step + m
                if (k == nsteps or (self._contacts and k % 256 == 0)
                        or (self.energy_interval and k % self.energy_interval == 0)
                        or (self.store and tt >= self._next_storage_time)):
                    break
            self.run_steps(U, vh, acc, dt, m, first=(step == 0))
            step += m
            t = tt
            if self._contacts and (step % 256 == 0 or step == nsteps):
                self._check_contacts()          # candidate-list overflow, one sync per 256 steps
            if self.energy_interval and (step % self.energy_interval == 0 or step == nsteps):
                self.energy_history.append((t, self.energy_balance(U, vh, acc, dt)))
            if self.store and (t >= self._next_storage_time or step == nsteps):
                gv = vh - 0.5 * dt * acc
                self._GC_list.append(p.u_to_gc(U))
                self._GV_list.append(gv)
                self._GA_list.append(acc.clone())
                self._time_list.append(t)
                self._next_storage_time += self._time_per_storage
        self.steps_done = nsteps
        self.GC = p.u_to_gc(U)
        self.U, self.V_half, self.A = U, vh, acc
        self.assembly.GC = self.GC
        self.assembly.RGC = self.assembly._GC2RGC(self.GC)
        if self.verbose:
            print(f'Total steps: {nsteps}, Total time: {time.time() - t_start:.2f} s')
        return True

    def energy_balance(self, U, vh, acc, dt):
        """Float64 energies at the full step of U: kinetic 1/2 v.M.v with
        v = v_{n-1/2} + dt/2 a(U) (v = vh itself at the start, acc None), the
        potential energy of elements and loads, and the contact energy."""
        p = self.assembly.plan
        fc = torch.zeros(3 * p.n_nodes, dtype=torch.float64, device=p.device)
        ce = 0.0
        for c in self._contacts:
            st = c.state(p)
            fc += st.forces(U, search=True)
            ce += st.energy()
        if acc is None:
            v = vh
        else:
            R = p.residual(U).clone()
            a = -self._minv * (R - p.u_to_gc(fc))
            v = vh + 0.5 * dt * a
        ke = 0.5 * float(torch.sum(v * v / self._minv))
        pe = p.energy(U)
        return {'kinetic': ke, 'potential': pe, 'contact': ce, 'total': ke + pe + ce}

    def _contact_force(self, U, host_logic=True):
        p = self.assembly.plan

        def one(c):
            st = c.state(p)
            if c.search_interval == 1:
                # exact forces from reused candidate lists; the host polls and rebuilds them
                # step by step, or between graph replays (host_logic=False)
                return st.step_forces(U) if host_logic else st.eval(U)
            return st.forces(U, check=False)

        if len(self._contacts) == 1:
            return one(self._contacts[0])
        self._fc_sum.zero_()
        for c in self._contacts:
            call('fea_axpy', self._fc_sum.numel(), 1.0, F64, one(c), self._fc_sum, nat.stream_ptr())
        return self._fc_sum

    def _check_contacts(self):
        for c in self._contacts:
            c.state(self.assembly.plan).check()

    def step(self, U, vh, acc, dt, first=False, load_scale=1.0, store_acc=True):
        """One central-difference step at the current U (in place), contact lists managed
        step by step.  store_acc=False: the fused step may leave acc unwritten (a step reads
        no acceleration; run_steps stores it on its last step)."""
        self._step_kernels(U, vh, acc, dt, first, load_scale, host_logic=True, store_acc=store_acc)

    def _node_minv(self):
        """1/m per node (the value of its free DOFs, which share the row-sum lumped mass; 0 for
        a node without free DOFs), or None if some node's free DOFs differ."""
        p = self.assembly.plan
        dm = p.dof_map.reshape(-1, 3).long()
        mi = torch.where(dm >= 0, self._minv[dm.clamp(min=0)], torch.full_like(dm, float('nan'), dtype=torch.float64))
        first = torch.nan_to_num(mi, nan=0.0).amax(dim=1)          # lumped masses are positive
        free = dm >= 0
        if bool(((mi != first[:, None]) & free).any()):
            return None
        return first.contiguous()

    def _fused_contact(self):
        """The contact state of the fused two-launch step, or None (no tiles, not exactly one
        contact set of reusable lists, a side stream for contact, or fuse_contact off)."""
        if not (self.fuse_contact and self._tiles is not None and len(self._contacts) == 1 and not self.overlap
                and self._contacts[0].search_interval == 1 and self._minv_node is not None):
            return None
        return self._contacts[0].state(self.assembly.plan)

    def _step_fused(self, st, U, vh, acc, dt, first, host_logic, store_acc=True):
        """One step in two launches: fea_elem_residual_tiles_contact (the first element launch
        carries the contact penalty) and fea_explicit_update_tiles_contact.  dmax must hold the
        displacement at U: tracked by the previous step's update, by a rebuild, or here on the
        first step."""
        p = self.assembly.plan
        s = nat.stream_ptr()
        tl = self._tiles
        if host_logic:
            st.step_begin(U)
        cs = ctypes.addressof(st.contact_set())
        if first:
            nat.check_rc(nat.raw('fea_contact_track')(cs, p.X.data_ptr(), U.data_ptr(), s), 'fea_contact_track')
        u = U.data_ptr()
        for k, a in enumerate(tl['elem_args']):
            if k == 0:
                nat.check_rc(nat.raw('fea_elem_residual_tiles_contact')(a[0], a[1], a[2], a[3], a[4], u, a[5], a[6],
                                                                        a[7], a[8], a[9], cs, s),
                             'fea_elem_residual_tiles_contact')
            else:
                nat.check_rc(nat.raw('fea_elem_residual_tiles')(a[0], a[1], a[2], a[3], a[4], u, a[5], a[6], a[7],
                                                                a[8], a[9], None, None, None, 0, s),
                             'fea_elem_residual_tiles')
        nat.check_rc(nat.raw('fea_explicit_update_tiles_contact')(
            p.n_nodes, tl['node_pptr'].data_ptr(), tl['node_plist'].data_ptr(), p.dof_map.data_ptr(),
            tl['partial'].data_ptr(), self._minv_node.data_ptr(), st.node_mask().data_ptr(), cs, p.X.data_ptr(),
            st.fc.data_ptr(), 0.5 * dt if first else dt, dt, acc.data_ptr() if store_acc else None, vh.data_ptr(), u,
            s),
            'fea_explicit_update_tiles_contact')
        if host_logic:
            st.step_end()

    def _step_kernels(self, U, vh, acc, dt, first=False, load_scale=1.0, host_logic=False, store_acc=True):
        """The launches of one step.  With ``overlap`` the contact forces are formed on a side
        stream while the element pass runs (both only read U); the update waits for both.
        host_logic=False: the contact forces come from the stored lists with no host reads
        or rebuilds (the form captured in a HIP graph)."""
        st = self._fused_contact()
        if st is not None:
            return self._step_fused(st, U, vh, acc, dt, first, host_logic, store_acc)
        p = self.assembly.plan
        s = nat.stream_ptr()
        tl = self._tiles
        fc = None
        side = self._side_stream() if (self._contacts and self.overlap) else None
        if side is not None:
            main = torch.cuda.current_stream()
            side.wait_stream(main)                      # U of this step is written
            with torch.cuda.stream(side):
                fc = self._contact_force(U, host_logic)
        if tl is None:
            p.residual_stage(U, load_scale, F32 if self.precision == 'fp32' else F64)
        else:
            fn, u = nat.raw('fea_elem_residual_tiles'), U.data_ptr()
            for a in tl['elem_args']:
                nat.check_rc(fn(a[0], a[1], a[2], a[3], a[4], u, a[5], a[6], a[7], a[8], a[9], a[10], a[11],
                                a[12], a[13], s), 'fea_elem_residual_tiles')
        if side is not None:
            main.wait_stream(side)                      # fc is formed
        elif self._contacts:
            fc = self._contact_force(U, host_logic)
        if tl is None:
            call('fea_explicit_update', p.n_nodes, p.rseg, p.dof_map, p.rstage, self._minv, fc,
                 0.5 * dt if first else dt, dt, acc, vh, U, s)
        else:
            call('fea_explicit_update_tiles', p.n_nodes, tl['node_pptr'], tl['node_plist'], p.dof_map,
                 tl['partial'], self._minv, fc, 0.5 * dt if first else dt, dt, acc, vh, U, s)

    def _side_stream(self):
        st = getattr(self, '_side', None)
        if st is None or st.device != torch.device('cuda', torch.cuda.current_device()):
            st = self._side = torch.cuda.Stream()       # default priority
        return st

    def run_steps(self, U, vh, acc, dt, n, first=False):
        """``n`` steps (the first one with the half-step velocity when ``first``).  With
        ``use_graph`` (and reusable contact lists or no contact) steps run in chunks of
        ``chunk`` replayed from a HIP graph captured once per (U, vh, acc, dt); between
        replays the host reads the contact displacement poll of the chunk before last (the
        GPU keeps the queued chunk), rebuilds the lists when due and issues the next poll.
        The graph replays exactly the launches of ``step``: same results bit for bit."""
        n = int(n)
        fst = self._fused_contact()
        if fst is not None and not first and n > 0:
            fst.track(U)                                # U may come from elsewhere than a step
        if first and n > 0:
            self.step(U, vh, acc, dt, first=True, store_acc=(n == 1))
            n -= 1
        usable = self._tiles is not None and all(c.search_interval == 1 for c in self._contacts)
        if not (self.use_graph and usable) or n < 2 * self.chunk:
            for k in range(n):
                self.step(U, vh, acc, dt, store_acc=(k == n - 1))
            return
        K = self.chunk
        key = (U.data_ptr(), vh.data_ptr(), acc.data_ptr(), float(dt), K)
        if getattr(self, '_graph_key', None) != key:
            self._graph = None
            states = [c.state(self.assembly.plan) for c in self._contacts]
            for st in states:                           # lists valid for the capture's replays
                if st.rebuild_due(2 * K):
                    st.rebuild(U)
            if self._contacts and self.overlap:
                self._side_stream()
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.graph(g):
                for k in range(K):
                    self._step_kernels(U, vh, acc, dt, store_acc=(k == K - 1))
            torch.cuda.synchronize()
            self._graph, self._graph_key = g, key
        states = [c.state(self.assembly.plan) for c in self._contacts]
        pending = [[] for _ in states]                  # per contact set: polls of the last two chunks
        while n >= K:
            for st, pend in zip(states, pending):
                if len(pend) >= 2:                      # the chunk before last has finished
                    st._poll = pend.pop(0)
                    st.poll_read()
                if st.rebuild_due(2 * K):
                    st.rebuild(U)
                    pend.clear()
            self._graph.replay()
            for st, pend in zip(states, pending):
                st.steps += K
                st.poll_issue()
                pend.append(st._poll)
                st._poll = None
            n -= K
        for st in states:
            st._poll = None
        for k in range(n):
            self.step(U, vh, acc, dt, store_acc=(k == n - 1))
