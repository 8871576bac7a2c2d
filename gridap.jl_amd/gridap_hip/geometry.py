"""Discrete models on the host: Gridap's CartesianDiscreteModel and its labels.

Mirrors src/Geometry/CartesianDiscreteModels.jl and src/Geometry/CartesianGrids.jl:
node ids over LinearIndices(partition .+ 1) (x fastest), cell nodes over
CartesianIndices((2,...,2)) from the cell's lower corner, and the face labeling
whose tags are "tag_XX" (one per reference n-cube face), "interior" and
"boundary".  Only what the assembly path needs is kept on the host; the device
evaluates the CartesianMap analytically.
"""
import numpy as np


def ncube_num_faces(D):
    return 3 ** D


def ncube_face_id(D, free_mask, anchor_bits):
    """1-based id of the reference n-cube face with free axes `free_mask` and
    fixed coordinates `anchor_bits`, in the face order of
    src/ReferenceFEs/ExtrusionPolytopes.jl:460-478 (dimension, then extrusion, then
    anchor, the last axis most significant)."""
    from math import comb
    d = bin(free_mask).count("1")
    offset = sum(comb(D, k) << (D - k) for k in range(d))
    rank = sum(1 for m in range(free_mask) if bin(m).count("1") == d)
    a, bit = 0, 0
    for ax in range(D):
        if free_mask >> ax & 1:
            continue
        if anchor_bits >> ax & 1:
            a |= 1 << bit
        bit += 1
    return offset + rank * (1 << (D - d)) + a + 1


class FaceLabeling:
    """Vertex entities and tags of a Cartesian model (_fill_cartesian_face_labeling!,
    src/Geometry/CartesianDiscreteModels.jl:137-170)."""

    def __init__(self, D, vertex_entity):
        self.D = D
        self.vertex_entity = vertex_entity  # int32 per vertex
        nf = ncube_num_faces(D)
        width = len(str(nf))
        self.tag_names = ["tag_%s" % str(i).zfill(width) for i in range(1, nf)]
        self.tag_names += ["interior", "boundary"]
        self.tag_entities = [[i] for i in range(1, nf)] + [[nf], list(range(1, nf))]

    def add_tag_from_tags(self, name, tags):
        """add_tag_from_tags!(labels, name, tags) (src/Geometry/FaceLabelings.jl): a new
        tag whose entities are the union of the given tags' entities."""
        if isinstance(tags, (str, int, np.integer)):
            tags = [tags]
        ents = sorted({e for t in tags for e in self.tag_to_entities(t)})
        if name in self.tag_names:
            raise ValueError("tag %r already exists" % name)
        self.tag_names.append(name)
        self.tag_entities.append(ents)

    def tag_to_entities(self, tag):
        if isinstance(tag, str):
            if tag not in self.tag_names:
                raise KeyError("unknown tag %r" % tag)
            return self.tag_entities[self.tag_names.index(tag)]
        t = int(tag)
        if not 1 <= t <= len(self.tag_entities):
            raise KeyError("unknown tag id %r" % tag)
        return self.tag_entities[t - 1]

    def face_tag_index(self, tags, d=0):
        """get_face_tag_index (src/Geometry/FaceLabelings.jl:411-427) for vertices:
        position (1-based) of the first tag holding the vertex's entity, 0 if none."""
        if d != 0:
            raise NotImplementedError("only vertex tags are needed by Q1 spaces")
        if isinstance(tags, (str, int, np.integer)):
            tags = [tags]
        out = np.zeros(len(self.vertex_entity), dtype=np.int8)
        for i, tag in enumerate(tags):
            ents = np.asarray(self.tag_to_entities(tag), dtype=np.int32)
            hit = np.isin(self.vertex_entity, ents) & (out == 0)
            out[hit] = i + 1
        return out


class CartesianDiscreteModel:
    """CartesianDiscreteModel(domain, partition; map=identity).

    domain = (x0, x1, y0, y1[, z0, z1]); partition = (nx, ny[, nz]).  A non-identity
    `map` moves the node coordinates (src/Geometry/CartesianGrids.jl:107-124); the
    cell map is then isoparametric instead of the analytic CartesianMap."""

    def __init__(self, domain, partition, map=None):
        D = len(partition)
        if D not in (2, 3) or len(domain) != 2 * D:
            raise ValueError("domain/partition must describe a 2-D or 3-D box")
        self.D = D
        self.partition = tuple(int(n) for n in partition)
        if min(self.partition) < 1:
            raise ValueError("partition entries must be >= 1")
        self.origin = np.array([float(domain[2 * d]) for d in range(D)])
        self.sizes = np.array([(float(domain[2 * d + 1]) - float(domain[2 * d])) / partition[d]
                               for d in range(D)])
        self.map = map
        self._xyz = None
        self._cell_nodes = None
        self._labels = None

    # sizes ----------------------------------------------------------------
    def num_cells(self):
        return int(np.prod(self.partition))

    def num_nodes(self):
        return int(np.prod([n + 1 for n in self.partition]))

    num_vertices = num_nodes

    def is_affine_cartesian(self):
        return self.map is None

    def is_simplex(self):
        return False

    # geometry ---------------------------------------------------------------
    def node_coordinates(self):
        if self._xyz is None:
            axes = [self.origin[d] + np.arange(self.partition[d] + 1) * self.sizes[d]
                    for d in range(self.D)]
            grids = np.meshgrid(*axes, indexing="ij")
            xyz = np.stack([g.ravel(order="F") for g in grids], axis=1)
            if self.map is not None:
                xyz = np.ascontiguousarray(np.asarray(self.map(xyz), dtype=np.float64))
            self._xyz = xyz
        return self._xyz

    def cell_node_ids(self):
        """1-based Int32 cell_node_ids, cells x 2^D (CartesianCellNodes)."""
        if self._cell_nodes is None:
            D, part = self.D, self.partition
            ci = np.indices(part).reshape(D, -1, order="F")  # cell index per axis
            nodes = np.zeros((ci.shape[1], 1 << D), dtype=np.int64)
            for a in range(1 << D):
                node = np.zeros(ci.shape[1], dtype=np.int64)
                stride = 1
                for d in range(D):
                    node += (ci[d] + ((a >> d) & 1)) * stride
                    stride *= part[d] + 1
                nodes[:, a] = node + 1
            self._cell_nodes = nodes.astype(np.int32)
        return self._cell_nodes

    # labels -------------------------------------------------------------------
    def get_face_labeling(self):
        if self._labels is None:
            D, part = self.D, self.partition
            idx = np.indices([n + 1 for n in part]).reshape(D, -1, order="F")
            free = np.zeros(idx.shape[1], dtype=np.int64)
            anchor = np.zeros(idx.shape[1], dtype=np.int64)
            for d in range(D):
                free |= ((idx[d] > 0) & (idx[d] < part[d])).astype(np.int64) << d
                anchor |= (idx[d] == part[d]).astype(np.int64) << d
            table = np.zeros((1 << D, 1 << D), dtype=np.int32)
            for m in range(1 << D):
                for a in range(1 << D):
                    table[m, a] = ncube_face_id(D, m, a & ~m)
            ent = table[free, anchor & ~free]
            ent[free == (1 << D) - 1] = ncube_num_faces(D)
            self._labels = FaceLabeling(D, ent.astype(np.int32))
        return self._labels


class Triangulation:
    """Triangulation(model): the cells of the model (volume integrals only)."""

    def __init__(self, model):
        self.model = model

    def num_cells(self):
        return self.model.num_cells()


def get_triangulation(model):
    return Triangulation(model)


class BoundaryTriangulation:
    """BoundaryTriangulation(model[, labels]; tags="boundary") on a Cartesian model
    (src/Geometry/BoundaryTriangulations.jl:122-183): the boundary facets whose entity
    (the box side holding them) belongs to the tags, each as (cell, local facet) with
    the facet in the n-cube face order (QUAD: y=0, y=1, x=0, x=1; HEX: z=0, z=1, y=0,
    y=1, x=0, x=1).  Facets are listed side by side, cells ascending on each side."""

    def __init__(self, model, labels=None, tags="boundary"):
        if not isinstance(model, CartesianDiscreteModel) or not model.is_affine_cartesian():
            raise NotImplementedError("boundary triangulations of Cartesian models only")
        self.model = model
        labels = labels if labels is not None else model.get_face_labeling()
        if isinstance(tags, (str, int, np.integer)):
            tags = [tags]
        ents = {e for t in tags for e in labels.tag_to_entities(t)}
        D, part = model.D, model.partition
        cells, sides = [], []
        ci = np.indices(part).reshape(D, -1, order="F")
        for side in range(2 * D):
            axis = D - 1 - side // 2
            hi = side & 1
            free = ((1 << D) - 1) & ~(1 << axis)
            ent = ncube_face_id(D, free, (hi << axis))
            if ent not in ents:
                continue
            on = ci[axis] == (part[axis] - 1 if hi else 0)
            c = np.nonzero(on)[0]
            cells.append(c)
            sides.append(np.full(len(c), side))
        self.face_cells = (np.concatenate(cells) if cells else np.zeros(0, np.int64)).astype(np.int64)
        self.face_sides = (np.concatenate(sides) if sides else np.zeros(0, np.int64)).astype(np.int64)

    def num_cells(self):
        return len(self.face_cells)

    def normals(self):
        """Outward unit normal of every face, [nfaces, D]."""
        D = self.model.D
        n = np.zeros((len(self.face_sides), D))
        axis = D - 1 - self.face_sides // 2
        n[np.arange(len(axis)), axis] = np.where(self.face_sides & 1, 1.0, -1.0)
        return n


# ---- simplicial models ------------------------------------------------------------

# Local vertex lists of the six tetrahedra of a HEX (x-fastest local node order),
# the oriented split that simplexify uses by default
# (src/ReferenceFEs/ExtrusionPolytopes.jl:290-299).
HEX_TO_TETS = ((0, 1, 2, 6), (0, 1, 4, 6), (1, 2, 3, 6), (1, 3, 6, 7), (1, 4, 5, 6), (1, 5, 6, 7))
# Local edges of TET as vertex pairs, in Gridap's face order
# (src/ReferenceFEs/ExtrusionPolytopes.jl:_polytopenfaces on (TET_AXIS,)*3).
TET_EDGES = ((0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3))


def _first_touch(keys):
    """Face ids (0-based) of a flat array of face keys visited in (cell, local face)
    order: the k-th distinct key met gets id k (the first-touch numbering of
    src/Geometry/GridTopologies.jl:1184-1251)."""
    uniq, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(uniq))
    return rank[inverse.reshape(-1)], len(uniq)


class UnstructuredDiscreteModel:
    """simplexify(CartesianDiscreteModel): tetrahedra (or triangles) on the
    Cartesian nodes with their grid topology (vertices = nodes, edges numbered by
    first touch) and the labels transferred from the Cartesian entities
    (src/Geometry/UnstructuredDiscreteModels.jl:85-311).  Node coordinates may be
    replaced (jittered) without changing the topology."""

    def __init__(self, cart, xyz=None):
        if cart.D != 3:
            raise NotImplementedError("simplexify is provided for 3-D Cartesian models")
        self.D = 3
        self.cart = cart
        self.partition = cart.partition
        self._xyz = cart.node_coordinates() if xyz is None else np.ascontiguousarray(xyz)
        hexn = cart.cell_node_ids().astype(np.int64) - 1           # cells x 8
        tets = hexn[:, np.array(HEX_TO_TETS)]                        # cells x 6 x 4
        self._cells = np.ascontiguousarray(tets.reshape(-1, 4))      # 0-based, hex-major
        self._edges = None
        self._edge_entity = None

    # sizes / geometry -----------------------------------------------------------
    def num_cells(self):
        return self._cells.shape[0]

    def num_nodes(self):
        return self._xyz.shape[0]

    num_vertices = num_nodes

    def is_affine_cartesian(self):
        return False

    def is_simplex(self):
        return True

    def node_coordinates(self):
        return self._xyz

    def cell_node_ids(self):
        return (self._cells + 1).astype(np.int32)

    def with_node_coordinates(self, xyz):
        m = UnstructuredDiscreteModel(self.cart, xyz)
        m._edges, m._edge_entity = self._edges, self._edge_entity
        return m

    def jittered(self, seed, amplitude=0.2):
        """Interior vertices moved by independent U(-a h_d, a h_d) per axis from a seeded
        generator (boundary vertices stay, so the domain is unchanged)."""
        rng = np.random.default_rng(seed)
        xyz = self.cart.node_coordinates().copy()
        ijk = self._node_ijk()
        interior = np.all([(ijk[d] > 0) & (ijk[d] < self.partition[d]) for d in range(3)], axis=0)
        h = self.cart.sizes
        shift = rng.uniform(-amplitude, amplitude, size=xyz.shape) * h[None, :]
        xyz[interior] += shift[interior]
        return self.with_node_coordinates(xyz)

    # topology --------------------------------------------------------------------
    def _node_ijk(self):
        n = np.arange(self.num_nodes(), dtype=np.int64)
        n0, n1 = self.partition[0] + 1, self.partition[1] + 1
        return np.stack([n % n0, (n // n0) % n1, n // (n0 * n1)])

    def cell_edges(self):
        """cells x 6 edge ids (0-based), first-touch order."""
        if self._edges is None:
            c = self._cells
            nn = self.num_nodes()
            ends = [(np.minimum(c[:, a], c[:, b]), np.maximum(c[:, a], c[:, b])) for a, b in TET_EDGES]
            keys = np.stack([lo * nn + hi for lo, hi in ends], axis=1).reshape(-1)
            ids, ne = _first_touch(keys)
            self._edges = ids.reshape(-1, 6)
            self._nedges = ne
            # the two end vertices of every edge
            ev = np.empty((ne, 2), dtype=np.int64)
            flat = keys
            ev[ids, 0] = flat // nn
            ev[ids, 1] = flat % nn
            self._edge_vertices = ev
        return self._edges

    def num_edges(self):
        self.cell_edges()
        return self._nedges

    def edge_vertices(self):
        self.cell_edges()
        return self._edge_vertices

    # labels ------------------------------------------------------------------------
    def vertex_entity(self):
        return self.cart.get_face_labeling().vertex_entity

    def edge_entity(self):
        """Entity of each edge: the smallest Cartesian box face holding both ends (a
        hex edge keeps its label, a face diagonal gets the face's, the body diagonal
        the interior)."""
        if self._edge_entity is None:
            ev = self.edge_vertices()
            ijk = self._node_ijk()
            a, b = ijk[:, ev[:, 0]], ijk[:, ev[:, 1]]
            free = np.zeros(len(ev), dtype=np.int64)
            anchor = np.zeros(len(ev), dtype=np.int64)
            for d in range(3):
                lo = (a[d] == 0) & (b[d] == 0)
                hi = (a[d] == self.partition[d]) & (b[d] == self.partition[d])
                free |= (~(lo | hi)).astype(np.int64) << d
                anchor |= hi.astype(np.int64) << d
            table = np.zeros((8, 8), dtype=np.int32)
            for m in range(8):
                for an in range(8):
                    table[m, an] = ncube_face_id(3, m, an & ~m)
            self._edge_entity = table[free, anchor & ~free]
        return self._edge_entity

    def get_face_labeling(self):
        return self.cart.get_face_labeling()


def simplexify(model):
    """simplexify(model) (src/Geometry/DiscreteModels.jl:374-379) for Cartesian models."""
    if isinstance(model, UnstructuredDiscreteModel):
        return model
    return UnstructuredDiscreteModel(model)
