"""Node-to-surface penalty contact (new law, configs with two bodies).

Slave nodes (the nodes of a surface set of one instance) are pushed out of
master T3 facets (a surface set of another instance, faces oriented outward as
the reference's element faces are, tetrahedral.py:14-19) with force -k A g n
when they penetrate (gap g < 0); k is a stiffness per unit area and A the
node's tributary area on its surface, so the law does not depend on the mesh
size.  The reference's contact (contact.py:11-1217) is a smooth
Gauss-point-pair barrier with a cKDTree broad phase; this load replaces its role
in explicit dynamics (whose reference force-only branches are broken,
contact.py:658-659, :1140-1142) with the uniform-grid search of the HIP library.
"""
import ctypes

import numpy as np
import torch

from ... import _native as nat
from ..._native import call
from .base import BaseLoad


class NodeToSurfaceContact(BaseLoad):

    def __init__(self, slave_instance: str, slave_surface: str, master_instance: str, master_surface: str,
                 penalty: float = 1e3, radius: float = None, depth_max: float = None, kmax: int = 64,
                 search_interval: int = 1, skin: float = None) -> None:
        super().__init__()
        self.slave_instance, self.slave_surface = slave_instance, slave_surface
        self.master_instance, self.master_surface = master_instance, master_surface
        self.penalty = float(penalty)
        self.radius = radius
        self.depth_max = depth_max
        self.kmax = int(kmax)
        self.search_interval = int(search_interval)
        self.skin = skin                # explicit steps: candidate lists reused within skin / 2 of motion
        self._state = None

    def initialize(self, assembly):
        super().initialize(assembly)
        si = assembly.get_instance(self.slave_instance)
        mi = assembly.get_instance(self.master_instance)
        soff = assembly.RGC_list_indexStart[si._RGC_index] // 3
        moff = assembly.RGC_list_indexStart[mi._RGC_index] // 3
        self.slave_nodes, self.slave_area = _tributary_areas(si, self.slave_surface)
        self.slave_nodes = self.slave_nodes + soff
        facets = [s for s in mi.surfaces.get_elements(self.master_surface) if s._elems.shape[0] > 0]
        if any(s.num_nodes_per_elem != 3 for s in facets):
            raise NotImplementedError('node-to-surface contact supports T3 master facets')
        if not facets:
            raise ValueError('master surface %r has no facets' % self.master_surface)
        self.master_conn = np.concatenate([s._elems.cpu().numpy() for s in facets]).astype(np.int64) + moff
        X = mi.nodes.detach().cpu().numpy()
        tri = X[self.master_conn - moff]
        edge = float(np.mean([np.linalg.norm(tri[:, (k + 1) % 3] - tri[:, k], axis=1) for k in range(3)]))
        # a slave within depth_max of a facet's plane whose projection lies inside the
        # facet is within depth_max + (largest vertex-centroid distance) of its centroid
        rho = float(np.linalg.norm(tri - tri.mean(axis=1, keepdims=True), axis=2).max())
        if self.depth_max is None:
            self.depth_max = edge
        if self.radius is None:
            self.radius = float(self.depth_max) + 1.25 * rho
        if self.skin is None:
            # a margin of 0.35 radius keeps the radius + skin candidate lists within kmax on
            # structured tet surfaces (at most 47 candidates of 64, against 26 within radius)
            self.skin = 0.35 * float(self.radius)
        self.origin = tri.reshape(-1, 3).min(0) - self.radius
        self._state = None

    def set_required_DoFs(self, RGC_remain_index):
        return RGC_remain_index

    def device_groups(self, assembly):
        return []

    def state(self, plan):
        if self._state is None or self._state.plan is not plan:
            self._state = ContactState(self, plan)
        return self._state

    def get_stiffness(self, RGC, if_onlyforce=False, *args, **kwargs):
        if not if_onlyforce:
            raise NotImplementedError('node-to-surface contact provides forces only (explicit dynamics)')
        plan = self._assembly.plan
        U = self._assembly._u_from(RGC=RGC)
        fc = self.state(plan).forces(U, search=True)
        idx = torch.arange(fc.numel(), device=fc.device)
        return idx, fc.clone()

    def get_potential_energy(self, RGC):
        plan = self._assembly.plan
        U = self._assembly._u_from(RGC=RGC)
        st = self.state(plan)
        st.forces(U, search=True)
        return -st.energy()


def _tributary_areas(instance, surface):
    """Unique nodes of a surface set and their tributary areas (a third of each
    adjacent facet's area, reference configuration; Q4 facets split in two T3)."""
    X = instance.nodes.detach().cpu().numpy()
    tris = []
    for s in instance.surfaces.get_elements(surface):
        e = s._elems.cpu().numpy()
        if s.num_nodes_per_elem == 3:
            tris.append(e)
        elif s.num_nodes_per_elem == 4:
            tris += [e[:, [0, 1, 2]], e[:, [0, 2, 3]]]
        else:
            tris.append(e[:, :3])           # T6: corner nodes only
    if not tris:
        raise ValueError('slave surface %r has no facets' % surface)
    tri = np.concatenate(tris).astype(np.int64)
    P = X[tri]
    a = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    nodes, inv = np.unique(tri.reshape(-1), return_inverse=True)
    area = np.bincount(inv.reshape(-1), weights=np.repeat(a / 3.0, 3), minlength=nodes.shape[0])
    return nodes, area


DMAX_SLOTS, DMAX_STRIDE = 64, 16      # FEA_CONTACT_DMAX_SLOTS / _STRIDE of include/fea_amd.h


def slots_max(dmax_words):
    """The tracked squared displacement from the dmax words (int64 bits of doubles >= 0)."""
    return float(np.asarray(dmax_words).view(np.float64)[::DMAX_STRIDE][:DMAX_SLOTS].max())


class ContactSet(ctypes.Structure):
    """fea_contact_set of include/fea_amd.h (sizes, constants and device buffers of
    one contact evaluation)."""
    _fields_ = [('n_slave', ctypes.c_int32), ('n_facets', ctypes.c_int32), ('n_cnode', ctypes.c_int32),
                ('kmax', ctypes.c_int32), ('table_bits', ctypes.c_int32), ('n_master', ctypes.c_int32),
                ('cell', ctypes.c_double), ('radius', ctypes.c_double), ('skin', ctypes.c_double),
                ('penalty', ctypes.c_double), ('depth_max', ctypes.c_double), ('origin', ctypes.c_double * 3)] + \
               [(n, ctypes.c_void_p) for n in ('slave', 'conn', 'cnode', 'cn_slave', 'cn_master', 'nf_ptr',
                                               'nf_facet', 'nf_vert', 'slave_area', 'centroid', 'force', 'weight',
                                               'x0', 'dmax', 'key', 'bucket_start', 'bucket_items', 'work', 'cand',
                                               'count', 'max_count', 'chosen', 'sup_max', 'inv_start', 'inv_items', 'm_ptr',
                                               'm_items', 'lists_ok')]


class ContactState:
    """Device buffers and the search / force pipeline of one contact pair set."""

    def __init__(self, load, plan):
        self.plan, self.load = plan, load
        dev = plan.device
        S, F = load.slave_nodes.shape[0], load.master_conn.shape[0]
        self.S, self.F = S, F
        self.bits = max(4, int(np.ceil(np.log2(max(2 * F, 16)))))
        nb = 1 << self.bits
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)
        self.slave = i32(load.slave_nodes)
        self.area = torch.as_tensor(np.ascontiguousarray(load.slave_area, dtype=np.float64), device=dev)
        self.conn = i32(load.master_conn)
        mnodes, inv = np.unique(load.master_conn.reshape(-1), return_inverse=True)
        inv = inv.reshape(-1, 3)
        order = np.argsort(inv.reshape(-1), kind='stable')
        fac = np.repeat(np.arange(F), 3)[order]
        vert = np.tile(np.arange(3), F)[order]
        ptr = np.concatenate([[0], np.cumsum(np.bincount(inv.reshape(-1), minlength=mnodes.shape[0]))])
        self.mnodes, self.nf_ptr, self.nf_facet, self.nf_vert = i32(mnodes), i32(ptr), i32(fac), i32(vert)
        z = lambda n, dt: torch.zeros(max(1, n), dtype=dt, device=dev)
        self.cent = z(3 * F, torch.float64)
        self.key = z(F, torch.int32)
        self.bstart = z(nb + 1, torch.int32)
        self.bitems = z(F, torch.int32)
        self.work = z(max(nb, F), torch.int32)
        self.cand = z(S * load.kmax, torch.int32)
        self.count = z(S, torch.int32)
        self.max_count = z(1, torch.int32)
        self.chosen = z(S, torch.int32)
        self.force = z(3 * S, torch.float64)
        self.weight = z(3 * S, torch.float64)
        self.sb_start = z(F + 1, torch.int32)
        self.sb_items = z(S, torch.int32)
        self.fc = z(3 * plan.n_nodes, torch.float64)[:3 * plan.n_nodes]
        self.epart = z(nat.lib().fea_contact_blocks(S), torch.float64)
        self.eout = z(1, torch.float64)
        self.origin = np.ascontiguousarray(load.origin, dtype=np.float64)
        self.steps = 0
        self.searched = False
        # contact nodes (slaves and master nodes, each once) with their slave / master index
        cnode = np.union1d(load.slave_nodes, mnodes)
        cn_slave = np.full(cnode.shape[0], -1, dtype=np.int32)
        cn_master = np.full(cnode.shape[0], -1, dtype=np.int32)
        cn_slave[np.searchsorted(cnode, load.slave_nodes)] = np.arange(S, dtype=np.int32)
        cn_master[np.searchsorted(cnode, mnodes)] = np.arange(mnodes.shape[0], dtype=np.int32)
        self.cnode, self.cn_slave, self.cn_master = i32(cnode), i32(cn_slave), i32(cn_master)
        # reused candidate lists (explicit steps): the rebuild's state and the host's poll of dmax
        self.x0 = z(3 * cnode.shape[0], torch.float64)
        self.vcand = z(S * load.kmax, torch.int32)
        self.vcount = z(S, torch.int32)
        self.dmax = z(DMAX_SLOTS * DMAX_STRIDE, torch.int64)     # slot k at [k * DMAX_STRIDE]
        self.sup_max = z(1, torch.int32)
        self.inv_start = z(F + 1, torch.int32)
        self.inv_items = z(S * load.kmax, torch.int32)
        self.m_ptr = z(mnodes.shape[0] + 1, torch.int32)
        self.m_items = z(6 * S * load.kmax, torch.int32)
        self.lists_ok = z(1, torch.int32)
        self._set = None
        self._need_rebuild = True
        self._pin = torch.zeros(DMAX_SLOTS * DMAX_STRIDE, dtype=torch.int64).pin_memory() \
            if torch.cuda.is_available() else None
        self._poll = None
        self._reading = None
        self._rebuilt_at = 0
        self.rebuilds = 0

    def search(self, U, check=True):
        """Rebuild the facet grid at U and find every slave's candidate facets.
        With check=False the overflow test is deferred to ``check()``."""
        L, s = self.load, nat.stream_ptr()
        call('fea_contact_grid_build', self.F, self.conn, self.plan.X, U, self.origin, float(L.radius),
             self.bits, self.cent, self.key, self.bstart, self.bitems, self.work, s)
        call('fea_contact_search', self.S, self.slave, self.plan.X, U, self.cent, self.bstart, self.bitems,
             self.origin, float(L.radius), self.bits, float(L.radius), L.kmax, self.cand, self.count,
             self.max_count, s)
        self.searched = True
        if check:
            self.check()

    def check(self):
        """Raise if any search since the last check found more than kmax candidates."""
        mx = int(self.max_count.item())
        if mx > self.load.kmax:
            self.max_count.zero_()
            raise RuntimeError('contact search: %d candidate facets for one node exceed kmax=%d '
                               '(reduce the search radius)' % (mx, self.load.kmax))

    def pairs(self):
        """Candidate (slave position, facet) pairs as a sorted int64 array [P, 2]."""
        cnt = self.count[:self.S].cpu().numpy()
        cand = self.cand[:self.S * self.load.kmax].cpu().numpy().reshape(self.S, self.load.kmax)
        out = [(s, f) for s in range(self.S) for f in cand[s, :cnt[s]]]
        return np.array(out, dtype=np.int64).reshape(-1, 2)

    def forces(self, U, search=None, check=True):
        L = self.load
        if search or (search is None and (not self.searched or self.steps % L.search_interval == 0)):
            self.search(U, check=check)
        self.steps += 1
        call('fea_contact_forces', self.S, self.slave, self.plan.X, U, self.conn, self.F, self.cand, self.count,
             L.kmax, L.penalty, self.area, float(L.depth_max), int(self.mnodes.numel()), self.mnodes, self.nf_ptr,
             self.nf_facet, self.nf_vert, self.chosen, self.force, self.weight, self.sb_start, self.sb_items,
             self.work, 3 * self.plan.n_nodes, self.fc, self.epart, nat.stream_ptr())
        return self.fc

    def contact_set(self):
        """The fea_contact_set of this pair set (built once; the buffers never move)."""
        if self._set is None:
            L = self.load
            if L.kmax > 64:
                raise ValueError('kmax above 64')
            if nat.lib().fea_contact_set_size() != ctypes.sizeof(ContactSet):
                raise nat.FeaNativeError('fea_contact_set layout mismatch')
            c = ContactSet()
            c.n_slave, c.n_facets, c.n_cnode = self.S, self.F, int(self.cnode.numel())
            c.n_master = int(self.mnodes.numel())
            c.kmax, c.table_bits = L.kmax, self.bits
            c.radius, c.skin = float(L.radius), float(L.skin)
            c.cell = c.radius + c.skin
            c.penalty, c.depth_max = float(L.penalty), float(L.depth_max)
            c.origin[:] = [float(v) for v in self.origin]
            for name, t in (('slave', self.slave), ('conn', self.conn), ('cnode', self.cnode),
                            ('cn_slave', self.cn_slave), ('cn_master', self.cn_master), ('nf_ptr', self.nf_ptr),
                            ('nf_facet', self.nf_facet), ('nf_vert', self.nf_vert), ('slave_area', self.area),
                            ('centroid', self.cent), ('force', self.force), ('weight', self.weight),
                            ('x0', self.x0), ('dmax', self.dmax), ('key', self.key), ('bucket_start', self.bstart),
                            ('bucket_items', self.bitems), ('work', self.work), ('cand', self.vcand),
                            ('count', self.vcount), ('max_count', self.max_count), ('chosen', self.chosen),
                            ('sup_max', self.sup_max), ('inv_start', self.inv_start),
                            ('inv_items', self.inv_items), ('m_ptr', self.m_ptr), ('m_items', self.m_items),
                            ('lists_ok', self.lists_ok)):
                setattr(c, name, t.data_ptr())
            self._set = c
        return self._set

    def rebuild(self, U):
        """Search with radius + skin at U and store the candidate and facet -> slave lists."""
        call('fea_contact_rebuild', ctypes.addressof(self.contact_set()), self.plan.X, U, nat.stream_ptr())
        self._need_rebuild = False
        self._poll = None
        self._rebuilt_at = self.steps
        self._reading = None
        self.rebuilds += 1

    def eval(self, U):
        """Penalty forces at U from the stored lists (fea_contact_eval, three launches): the
        forces of ``forces(U, search=True)``, written only at the contact nodes of ``fc`` (its
        other entries stay zero).  No host logic: capturable in a HIP graph."""
        call('fea_contact_eval', ctypes.addressof(self.contact_set()), self.plan.X, U, self.fc, self.epart,
             nat.stream_ptr())
        self.searched = True
        return self.fc

    def poll_issue(self):
        """Copy dmax (the largest squared contact-node displacement since the rebuild) to
        pinned memory behind the work queued so far."""
        self._pin.copy_(self.dmax, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._poll = (ev, self.steps)

    def poll_read(self):
        """Wait for the pending poll (if any) and keep its displacement reading."""
        if self._poll is None:
            return
        ev, at = self._poll
        ev.synchronize()
        self._poll = None
        self._reading = (at, float(np.sqrt(max(0.0, slots_max(self._pin.numpy())))))

    def rebuild_due(self, ahead):
        """Whether the lists should be rebuilt now: the displacement of the last reading,
        extrapolated at 1.5 times its mean rate since the rebuild to ``ahead`` steps from now
        (the next chance to rebuild), would pass 0.45 skin (the lists hold while it is at
        most skin / 2)."""
        if self._need_rebuild:
            return True
        if self._reading is None:
            return False
        at, d = self._reading
        rate = d / max(1, at - self._rebuilt_at)
        return d + 1.5 * rate * (self.steps - at + ahead) > 0.45 * self.load.skin

    def step_forces(self, U, poll_every=8):
        """``eval`` with the host's rebuild policy for step-by-step use: every ``poll_every``
        steps the previous poll is read (so the host runs at most ``poll_every`` steps ahead
        of its reading), the lists are rebuilt when ``rebuild_due``, and a new poll is issued.
        The kmax overflow check is left to ``check()``."""
        self.step_begin(U, poll_every)
        self.eval(U)
        self.step_end(poll_every)
        return self.fc

    def step_begin(self, U, poll_every=8):
        """The host's part of ``step_forces`` before the step's launches (poll read, rebuild)."""
        if self._need_rebuild or self.steps % poll_every == 0:
            self.poll_read()
            if self.rebuild_due(poll_every):
                self.rebuild(U)

    def step_end(self, poll_every=8):
        """The host's part of ``step_forces`` after the step's launches (count, poll issue)."""
        self.steps += 1
        if self.steps % poll_every == 0:
            self.poll_issue()

    def track(self, U):
        """Raise dmax by the contact nodes' displacement at U (fea_contact_track): what the
        explicit loop's fused launches need at a U they did not produce."""
        call('fea_contact_track', ctypes.addressof(self.contact_set()), self.plan.X, U, nat.stream_ptr())

    def node_mask(self):
        """uint32 bit mask over the plan's nodes marking the contact nodes (bit A % 32 of word
        A / 32), for fea_explicit_update_tiles_contact."""
        if getattr(self, '_cmask', None) is None:
            n = self.plan.n_nodes
            bits = np.zeros(32 * ((n + 31) // 32), dtype=bool)
            bits[self.cnode.cpu().numpy()] = True
            words = np.packbits(bits.reshape(-1, 32)[:, ::-1], axis=1, bitorder='big').view('>u4').astype(np.uint32)
            self._cmask = torch.as_tensor(words.reshape(-1).view(np.int32), device=self.plan.device)
        return self._cmask

    def energy(self):
        nb = nat.lib().fea_contact_blocks(self.S)
        call('fea_reduce_sum', nb, self.epart, self.eout, nat.stream_ptr())
        return float(self.eout.item())
