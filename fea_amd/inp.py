"""Abaqus INP ingestion (reference: FEA/inp.py:10-908, FEA/__init__.py:10-82).

``FEA_INP().read_inp(path)`` fills the same data model as the reference reader:

* ``part[name]`` with ``nodes`` (float64 tensor [N, 4]: 0-based node id, x, y, z),
  ``elems`` ({element type: int64 array [E, 1 + nen]} of 0-based element id and
  0-based node ids), ``sets_nodes`` / ``sets_elems`` ({name: set of 0-based ids}),
  ``surfaces`` ({name: [(0-based element ids, 0-based local face)]}), ``sections``
  ([[elset, material]]) and ``elems_material`` (float64 tensor [E_total, 5]:
  element id, density, material type, two material parameters);
* ``material[name]`` with ``type`` (1 NeoHookean from ``*Hyperelastic, neo hooke``
  as (mu, kappa) = (2 C10, 2 / D1), inp.py:460-466; 0 linear elastic (E, nu)) and
  ``density``;
* ``assemble.instances`` ({instance: (part, 4x4 transform)}) and the assembly-level
  sets and surfaces that name an instance (``*Nset/*Elset, instance=``,
  ``*Surface, type=ELEMENT``), keyed ``instance.name``.

``from_inp(inp)`` builds the controller as the reference does: one Part (and
Instance) per INP part, one element group per (element type, material type),
node sets and part-level surfaces.  Element types on the accelerated path are
C3D4, C3D8 and C3D10 (their hybrid / reduced variants read as the base type);
other types raise ``NotImplementedError`` at ``from_inp``.

The parser is keyword-driven: a line starting with ``*`` (and not ``**``)
opens a keyword with its comma-separated parameters; data lines follow until
the next keyword.  Node ids are assumed contiguous from 1 (as the reference
assumes: it indexes the node array by id - 1).
"""
from __future__ import annotations

import gzip

import numpy as np
import torch

_ELEM_ALIASES = {'C3D4': 'C3D4', 'C3D4H': 'C3D4', 'C3D8': 'C3D8', 'C3D8H': 'C3D8', 'C3D8R': 'C3D8',
                 'C3D10': 'C3D10', 'C3D10H': 'C3D10'}


def _keyword(line):
    """('*KEYWORD', {param: value}) of a keyword line (names upper-cased, values kept)."""
    parts = [p.strip() for p in line[1:].split(',')]
    name = parts[0].upper()
    params = {}
    for p in parts[1:]:
        if not p:
            continue
        if '=' in p:
            k, v = p.split('=', 1)
            params[k.strip().lower()] = v.strip()
        else:
            params[p.strip().lower()] = True
    return name, params


def _blocks(lines):
    """Yield (keyword, params, data lines) for every keyword of the file."""
    cur = None
    data = []
    for raw in lines:
        line = raw.strip()
        if not line or line.startswith('**'):
            continue
        if line.startswith('*'):
            if cur is not None:
                yield cur[0], cur[1], data
            cur = _keyword(line)
            data = []
        elif cur is not None:
            data.append(line)
    if cur is not None:
        yield cur[0], cur[1], data


def _ints(data):
    out = []
    for line in data:
        out.extend(int(v) for v in line.replace(',', ' ').split())
    return out


def _id_set(params, data):
    """0-based ids of an *Nset / *Elset (list or 'generate' ranges)."""
    vals = _ints(data)
    if params.get('generate'):
        ids = []
        for i in range(0, len(vals), 3):
            a, b, c = vals[i:i + 3]
            ids.extend(range(a, b + 1, c))
        vals = ids
    return {v - 1 for v in vals}


class FEA_INP:

    class Parts:
        def __init__(self) -> None:
            self.nodes = None
            self.elems: dict[str, np.ndarray] = {}
            self.sets_nodes: dict[str, set] = {}
            self.sets_elems: dict[str, set] = {}
            self.surfaces: dict[str, list] = {}
            self.sections: list = []
            self.elems_material = None

        def n_elements(self):
            return sum(int(e.shape[0]) for e in self.elems.values())

    class Materials:
        def __init__(self) -> None:
            self.type = None
            self.mat_para = None
            self.density = 0.0

    class Assembly:
        def __init__(self) -> None:
            self.instances: dict[str, tuple] = {}
            self.sets_nodes: dict[str, set] = {}
            self.sets_elems: dict[str, set] = {}
            self.surfaces: dict[str, list] = {}

    def __init__(self) -> None:
        self.part: dict[str, FEA_INP.Parts] = {}
        self.material: dict[str, FEA_INP.Materials] = {}
        self.assemble = FEA_INP.Assembly()

    # the reference's tests call it under this name too
    def Read_INP(self, path):
        return self.read_inp(path)

    def read_inp(self, path):
        opener = gzip.open if str(path).endswith('.gz') else open
        with opener(path, 'rt') as f:
            lines = f.readlines()
        self.part, self.material, self.assemble = {}, {}, FEA_INP.Assembly()
        scope, part, mat, instance = None, None, None, None
        for kw, params, data in _blocks(lines):
            if kw == 'PART':
                scope, part = 'part', FEA_INP.Parts()
                self.part[params['name']] = part
            elif kw == 'END PART':
                scope, part = None, None
            elif kw == 'ASSEMBLY':
                scope = 'assembly'
            elif kw == 'END ASSEMBLY':
                scope = None
            elif kw == 'INSTANCE' and scope == 'assembly':
                instance = params['name']
                T = np.eye(4)
                if data:
                    t = [float(v) for v in data[0].split(',') if v.strip()]
                    T[:3, 3] = t[:3]
                    if len(data) > 1:
                        raise NotImplementedError('rotated instances are outside the accelerated path')
                self.assemble.instances[instance] = (params['part'], T)
            elif kw == 'MATERIAL':
                mat = FEA_INP.Materials()
                self.material[params['name']] = mat
                scope = 'material'
            elif kw == 'HYPERELASTIC' and mat is not None:
                if 'neo hooke' not in params:
                    raise NotImplementedError('only *Hyperelastic, neo hooke is read')
                c10, d1 = [float(v) for v in data[0].split(',') if v.strip()][:2]
                mat.type, mat.mat_para = 1, [2.0 * c10, 2.0 / d1]
            elif kw == 'ELASTIC' and mat is not None:
                mat.type, mat.mat_para = 0, [float(v) for v in data[0].split(',') if v.strip()][:2]
            elif kw == 'DENSITY' and mat is not None:
                mat.density = float(data[0].split(',')[0])
            elif kw == 'NODE' and scope == 'part':
                arr = np.array([[float(v) for v in line.split(',')[:4]] for line in data], dtype=np.float64)
                arr[:, 0] -= 1.0
                if not np.array_equal(arr[:, 0], np.arange(arr.shape[0])):
                    raise ValueError('node ids must be contiguous from 1')
                part.nodes = torch.as_tensor(arr)
            elif kw == 'ELEMENT' and scope == 'part':
                etype = params['type'].upper()
                ids = np.array(_ints(data), dtype=np.int64) - 1
                nen = {'C3D4': 4, 'C3D4H': 4, 'C3D8': 8, 'C3D8H': 8, 'C3D8R': 8, 'C3D10': 10, 'C3D10H': 10,
                       'C3D6': 6, 'C3D15': 15, 'C3D20': 20, 'C3D20R': 20}.get(etype)
                if nen is None:
                    raise NotImplementedError('element type %s' % etype)
                arr = ids.reshape(-1, nen + 1)
                part.elems[etype] = np.concatenate([part.elems[etype], arr]) if etype in part.elems else arr
            elif kw == 'NSET':
                target = part if scope == 'part' else self.assemble
                name = params['nset'] if scope == 'part' or 'instance' not in params else \
                    '%s.%s' % (params['instance'], params['nset'])
                target.sets_nodes[name] = _id_set(params, data)
            elif kw == 'ELSET':
                target = part if scope == 'part' else self.assemble
                name = params['elset'] if scope == 'part' or 'instance' not in params else \
                    '%s.%s' % (params['instance'], params['elset'])
                target.sets_elems[name] = _id_set(params, data)
            elif kw == 'SURFACE':
                if str(params.get('type', 'ELEMENT')).upper() != 'ELEMENT':
                    continue
                target = part if scope == 'part' else self.assemble
                entries = []
                for line in data:
                    elset, face = [v.strip() for v in line.split(',')[:2]]
                    sets = target.sets_elems
                    key = elset
                    if key not in sets:
                        key = next((k for k in sets if k.endswith('.' + elset)), elset)
                    entries.append((np.array(sorted(sets[key]), dtype=np.int64), int(face[1:]) - 1))
                name = params['name']
                if scope != 'part':
                    inst = next((k.split('.')[0] for k in self.assemble.sets_elems if '.' in k), None)
                    name = '%s.%s' % (inst, name) if inst else name
                target.surfaces[name] = entries
            elif kw == 'SOLID SECTION' and scope == 'part':
                part.sections.append([params['elset'], params['material']])
        # per-element material table (inp.py:870-880)
        for p in self.part.values():
            n = p.n_elements()
            p.elems_material = -torch.ones([n, 5], dtype=torch.float64)
            for elset, mname in p.sections:
                idx = torch.tensor(sorted(p.sets_elems[elset]), dtype=torch.int64)
                m = self.material[mname]
                p.elems_material[idx, 0] = idx.to(torch.float64)
                p.elems_material[idx, 1] = float(m.density)
                p.elems_material[idx, 2] = float(m.type)
                p.elems_material[idx, 3] = float(m.mat_para[0])
                p.elems_material[idx, 4] = float(m.mat_para[1])
        return self


def from_inp(inp: FEA_INP, create_instance=True):
    """FEAController of an INP model (FEA/__init__.py:10-82)."""
    from .assemble import Assembly, Instance, Part, elements, materials
    from .controller import FEAController
    asm = Assembly()
    for pname, pd in inp.part.items():
        part = Part(pd.nodes[:, 1:].clone())
        for sname, ids in pd.sets_nodes.items():
            part.set_nodes[sname] = np.unique(np.array(sorted(ids), dtype=np.int64))
        asm.add_part(part=part, name=pname)
        if create_instance:
            asm.add_instance(instance=Instance(part), name=pname)
        for etype, arr in pd.elems.items():
            base = _ELEM_ALIASES.get(etype)
            if base is None:
                raise NotImplementedError('element type %s is outside the accelerated path' % etype)
            table = pd.elems_material[torch.as_tensor(arr[:, 0])]
            for mtype in torch.unique(table[:, 2].to(torch.int64)).tolist():
                sel = (table[:, 2].to(torch.int64) == mtype).numpy()
                if mtype != 1:
                    raise NotImplementedError('material type %d is outside the accelerated path' % mtype)
                mat = materials.initialize_materials(materials_type=mtype, materials_params=table[sel][:, 3:])
                el = elements.initialize_element(element_type=base, elems_index=torch.as_tensor(arr[sel, 0]),
                                                 elems=torch.as_tensor(arr[sel, 1:]), part=part)
                el.set_materials(mat)
                el.set_density(table[sel][:, 1])
                part.add_element(el)
        for sname, surf in pd.surfaces.items():
            part.add_surface_set(sname, [(ids, face) for ids, face in surf])
    fe = FEAController()
    fe.assembly = asm
    return fe
