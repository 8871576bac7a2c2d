"""Model exchange in Gridap's JSON format, so that the Julia side and the device read one
mesh (for example one jittered tetrahedral mesh).

Gridap writes a DiscreteModel as the Dict of an UnstructuredDiscreteModel
(src/Geometry/DiscreteModels.jl:383-390 -> src/Geometry/UnstructuredDiscreteModels.jl:
51-84) with three parts and a format version:
  grid     (src/Geometry/UnstructuredGrids.jl:198-231): node_coordinates (Dp values per
           node, node after node), Dp, cell_node_ids {data, ptrs} (1-based Table,
           src/Arrays/Tables.jl:1023-1034), reffes [{orders, extrusion, space, value}]
           (src/ReferenceFEs/CLagrangianRefFEs.jl:137-170), cell_type (1-based),
           orientation;
  topology (src/Geometry/UnstructuredGridTopologies.jl:576-607): vertex_coordinates,
           Dp, Dc, cell_type, polytopes [extrusions], face_vertices_d {data, ptrs} for
           d = 1..Dc, orientation;
  labeling (src/Geometry/FaceLabelings.jl:641-666): D, entities_d (entity of every
           d-face, d = 0..D), tags (entities of each tag), names;
  version  "v0.16" (src/Geometry/UnstructuredDiscreteModels.jl:2).
JSON itself (src/Io/Json.jl:22-52) is text; this module reads and writes it with
Python's json.  The d-faces of the topology are numbered by first touch over the cells
in the local face order of the reference polytope (the numbering of
src/Geometry/GridTopologies.jl:1184-1251), which is also the edge order the P2
numbering uses.
"""
import json

import numpy as np

from .geometry import (CartesianDiscreteModel, FaceLabeling, UnstructuredDiscreteModel,
                       _first_touch, ncube_face_id, ncube_num_faces)

VERSION = "v0.16"
HEX_AXIS, TET_AXIS = 1, 2

# Local d-faces (vertex lists, 0-based) of the reference polytopes in Gridap's face
# order (src/ReferenceFEs/ExtrusionPolytopes.jl:_polytopenfaces).
_LOCAL_FACES = {
    ("TET", 1): ((0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3)),
    ("TET", 2): ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)),
    ("TRI", 1): ((0, 1), (0, 2), (1, 2)),
    ("QUAD", 1): ((0, 1), (2, 3), (0, 2), (1, 3)),
    ("HEX", 1): ((0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (1, 3), (4, 6), (5, 7), (0, 4),
                 (1, 5), (2, 6), (3, 7)),
    ("HEX", 2): ((0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 4, 5), (2, 3, 6, 7), (0, 2, 4, 6),
                 (1, 3, 5, 7)),
}


def _poly(D, simplex):
    return ("TET" if D == 3 else "TRI") if simplex else ("HEX" if D == 3 else "QUAD")


def _table(rows):
    """Rows (n x k int array, 0-based) -> Gridap Table dict (1-based data and ptrs)."""
    rows = np.asarray(rows, dtype=np.int64)
    n, k = rows.shape
    return {"data": (rows.reshape(-1) + 1).tolist(),
            "ptrs": (np.arange(n + 1, dtype=np.int64) * k + 1).tolist()}


def _faces(cells, local):
    """First-touch d-faces: (cell -> face ids [C, nlocal], face vertices [nf, k])."""
    C = cells.shape[0]
    loc = np.array(local, dtype=np.int64)
    verts = cells[:, loc]                                       # [C, nlocal, k]
    key_rows = np.sort(verts, axis=2).reshape(-1, loc.shape[1])
    nv = int(cells.max()) + 1 if cells.size else 1
    k = loc.shape[1]
    if float(nv) ** k < 2.0 ** 62:                              # one integer key per face
        keys = key_rows @ (nv ** np.arange(k - 1, -1, -1, dtype=np.int64))
    else:
        _, keys = np.unique(key_rows, axis=0, return_inverse=True)
    ids, nf = _first_touch(np.asarray(keys).reshape(-1))
    fv = np.empty((nf, loc.shape[1]), dtype=np.int64)
    flat = verts.reshape(-1, loc.shape[1])
    first = np.full(nf, -1, dtype=np.int64)
    order = np.arange(len(ids))
    first[ids[::-1]] = order[::-1]                              # first touch of each face
    fv[:] = flat[first]
    return ids.reshape(C, len(local)), fv


def _box_entity(D, ijk_list, part):
    """Entity of a face of a Cartesian box lattice: the smallest box face holding all its
    vertices (lattice coordinates ijk_list: list of [D, n] arrays)."""
    n = ijk_list[0].shape[1]
    free = np.zeros(n, dtype=np.int64)
    anchor = np.zeros(n, dtype=np.int64)
    for d in range(D):
        lo = np.all([a[d] == 0 for a in ijk_list], axis=0)
        hi = np.all([a[d] == part[d] for a in ijk_list], axis=0)
        free |= (~(lo | hi)).astype(np.int64) << d
        anchor |= hi.astype(np.int64) << d
    table = np.zeros((1 << D, 1 << D), dtype=np.int64)
    for m in range(1 << D):
        for an in range(1 << D):
            table[m, an] = ncube_face_id(D, m, an & ~m) if m != (1 << D) - 1 else ncube_num_faces(D)
    return table[free, anchor & ~free]


def to_dict(model):
    """to_dict(model) of a CartesianDiscreteModel (QUAD/HEX) or of a simplexified
    model (TET): Gridap's UnstructuredDiscreteModel dictionary."""
    if isinstance(model, FileDiscreteModel):
        return model.source_dict()
    D = model.D
    simplex = isinstance(model, UnstructuredDiscreteModel)
    if not simplex and not isinstance(model, CartesianDiscreteModel):
        raise TypeError("to_dict needs a CartesianDiscreteModel or a simplexified model")
    xyz = np.asarray(model.node_coordinates(), dtype=np.float64)
    cells = np.asarray(model.cell_node_ids(), dtype=np.int64) - 1
    C = cells.shape[0]
    poly = _poly(D, simplex)
    ext = [TET_AXIS if simplex else HEX_AXIS] * D
    part = model.partition
    n0, n1 = part[0] + 1, part[1] + 1 if D > 1 else 1
    nid = np.arange(xyz.shape[0], dtype=np.int64)
    ijk = np.stack([nid % n0, (nid // n0) % n1, nid // (n0 * n1)][:D])
    labels = model.get_face_labeling()
    entities = {0: np.asarray(labels.vertex_entity, dtype=np.int64)}
    topo = {"vertex_coordinates": xyz.reshape(-1).tolist(), "Dp": D, "Dc": D,
            "cell_type": [1] * C, "polytopes": [ext], "orientation": bool(simplex)}
    for d in range(1, D):
        _, fv = _faces(cells, _LOCAL_FACES[(poly, d)])
        topo["face_vertices_%d" % d] = _table(fv)
        entities[d] = _box_entity(D, [ijk[:, fv[:, k]] for k in range(fv.shape[1])], part)
    topo["face_vertices_%d" % D] = _table(cells)
    entities[D] = np.full(C, ncube_num_faces(D), dtype=np.int64)
    lab = {"D": D, "tags": [list(map(int, e)) for e in labels.tag_entities],
           "names": list(labels.tag_names)}
    for d in range(D + 1):
        lab["entities_%d" % d] = entities[d].tolist()
    grid = {"node_coordinates": xyz.reshape(-1).tolist(), "Dp": D,
            "cell_node_ids": _table(cells),
            "reffes": [{"orders": [1] * D, "extrusion": ext, "space": "default",
                        "value": "Float64"}],
            "cell_type": [1] * C, "orientation": bool(simplex)}
    return {"grid": grid, "topology": topo, "labeling": lab, "version": VERSION}


def to_json(model):
    return json.dumps(to_dict(model))


def to_json_file(model, path):
    with open(path, "w") as f:
        json.dump(to_dict(model), f)


def _check_version(d):
    if "version" not in d:
        raise ValueError("Cannot convert Dict to UnstructuredDiscreteModel for legacy format")
    have = tuple(int(x) for x in str(d["version"]).lstrip("v").split("."))
    need = tuple(int(x) for x in VERSION.lstrip("v").split("."))
    if have < need:
        raise ValueError("Cannot convert Dict to UnstructuredDiscreteModel for outdated "
                         "format %s; the required format is %s" % (d["version"], VERSION))


def _rows(tab):
    data = np.asarray(tab["data"], dtype=np.int64) - 1
    ptrs = np.asarray(tab["ptrs"], dtype=np.int64) - 1
    k = np.diff(ptrs)
    if len(k) and not np.all(k == k[0]):
        raise NotImplementedError("tables with rows of different lengths (mixed cells)")
    return data.reshape(len(k), k[0] if len(k) else 0)


class FileDiscreteModel(UnstructuredDiscreteModel):
    """An UnstructuredDiscreteModel read from Gridap's dictionary (from_dict): node
    coordinates, cell nodes, the d-face tables and the face labeling as stored."""

    def __init__(self, d):
        _check_version(d)
        g, t, lab = d["grid"], d["topology"], d["labeling"]
        D = int(g["Dp"])
        if int(t["Dc"]) != D:
            raise NotImplementedError("models whose cells are of lower dimension than space")
        if len(g["reffes"]) != 1:
            raise NotImplementedError("one cell type per model")
        ext = tuple(int(e) for e in g["reffes"][0]["extrusion"])
        if any(int(o) != 1 for o in g["reffes"][0]["orders"]):
            raise NotImplementedError("first-order geometry only")
        self.D = D
        self.simplex = all(e == TET_AXIS for e in ext)
        if not self.simplex and not all(e == HEX_AXIS for e in ext):
            raise NotImplementedError("simplices or n-cubes only")
        self._dict = d
        self._xyz = np.asarray(g["node_coordinates"], dtype=np.float64).reshape(-1, D)
        self._cells = np.ascontiguousarray(_rows(g["cell_node_ids"]))
        self._face_vertices = {k: _rows(t["face_vertices_%d" % k]) for k in range(1, D)}
        self._entities = {k: np.asarray(lab["entities_%d" % k], dtype=np.int64)
                          for k in range(D + 1)}
        self._labels = FaceLabeling(D, self._entities[0].astype(np.int32))
        self._labels.tag_names = list(lab["names"])
        self._labels.tag_entities = [list(map(int, e)) for e in lab["tags"]]
        self._edges = None
        self.cart = None

    def source_dict(self):
        return self._dict

    def is_simplex(self):
        return self.simplex

    def is_affine_cartesian(self):
        return False

    def with_node_coordinates(self, xyz):
        d = json.loads(json.dumps(self._dict))
        d["grid"]["node_coordinates"] = np.asarray(xyz, dtype=np.float64).reshape(-1).tolist()
        return FileDiscreteModel(d)

    def jittered(self, seed, amplitude=0.2):
        raise NotImplementedError("jitter the model before writing it")

    def _cell_faces(self, dim):
        poly = _poly(self.D, self.simplex)
        loc = np.array(_LOCAL_FACES[(poly, dim)], dtype=np.int64)
        fv = self._face_vertices[dim]
        nv = self._xyz.shape[0]
        key = lambda rows: np.sort(rows, axis=-1) @ (nv ** np.arange(rows.shape[-1])[::-1])
        fkeys = key(fv)
        order = np.argsort(fkeys)
        ck = key(self._cells[:, loc])
        pos = np.searchsorted(fkeys[order], ck)
        ids = order[np.minimum(pos, len(order) - 1)]
        if not np.all(fkeys[ids] == ck):
            raise ValueError("a cell face is missing from the topology's face table")
        return ids

    def cell_edges(self):
        if self._edges is None:
            self._edges = self._cell_faces(1)
            self._nedges = self._face_vertices[1].shape[0]
            self._edge_vertices = self._face_vertices[1]
        return self._edges

    def vertex_entity(self):
        return self._entities[0]

    def edge_entity(self):
        return self._entities[1]

    def get_face_labeling(self):
        return self._labels


def from_dict(d):
    """from_dict(DiscreteModel, dict) (src/Geometry/DiscreteModels.jl:388-390)."""
    return FileDiscreteModel(d)


def from_json(s):
    return from_dict(json.loads(s))


def from_json_file(path):
    with open(path) as f:
        return from_dict(json.load(f))


DiscreteModelFromFile = from_json_file
