"""FE spaces on the host: the DOF numbering whose cell_dof_ids the engine consumes.

Q1 spaces on grids whose nodes are the vertices use Gridap's CLagrangianFESpace
numbering (src/FESpaces/FESpaceFactories.jl:56-85 -> src/FESpaces/
CLagrangianFESpaces.jl:155-379): node-major, components interleaved per node,
free ids 1.. and Dirichlet ids -1.. in node order; local dofs are
component-blocked (ldof = lnode + nnodes_cell*(comp-1),
src/ReferenceFEs/LagrangianDofBases.jl:77-96).
"""
import numpy as np

from .geometry import CartesianDiscreteModel, Triangulation, UnstructuredDiscreteModel
from .reference_fes import LagrangianRefFE


class FESpace:
    def __init__(self, model, reffe, conformity="H1", dirichlet_tags=(), dirichlet_masks=None,
                 trian=None, labels=None):
        if isinstance(model, Triangulation):
            model = model.model
        if not isinstance(model, (CartesianDiscreteModel, UnstructuredDiscreteModel)):
            raise TypeError("FESpace needs a CartesianDiscreteModel or a simplexified one")
        if not isinstance(reffe, LagrangianRefFE):
            raise TypeError("FESpace needs a ReferenceFE(lagrangian, T, order)")
        if conformity not in ("H1", ":H1"):
            raise NotImplementedError("only H1-conforming spaces are provided")
        self.simplex = bool(model.is_simplex())
        self.conforming = self.simplex or reffe.order == 2
        if self.simplex:
            if reffe.order not in (1, 2) or reffe.ncomp != 1:
                raise NotImplementedError("simplicial spaces: scalar P1 and P2")
        elif reffe.order == 2:
            if model.D not in (2, 3):
                raise NotImplementedError("Q2 spaces are provided on 2-D and 3-D Cartesian models")
        elif reffe.order != 1:
            raise NotImplementedError("n-cube spaces of order > 2 are not provided")
        self.model = model
        self.reffe = reffe
        self.order = reffe.order
        self.ncomp = reffe.ncomp
        if isinstance(dirichlet_tags, (str, int, np.integer)):
            dirichlet_tags = [dirichlet_tags]
        self.dirichlet_tags = list(dirichlet_tags)
        labels = labels if labels is not None else model.get_face_labeling()
        if self.simplex:
            self._number_simplex(labels)
            return
        if self.order == 2:
            self._number_q2(labels, dirichlet_masks)
            return
        node_to_tag = labels.face_tag_index(self.dirichlet_tags, 0)
        ntags = len(self.dirichlet_tags)
        if dirichlet_masks is None:
            masks = np.ones((max(ntags, 1), self.ncomp), dtype=bool)
        else:
            masks = np.array(dirichlet_masks, dtype=bool).reshape(ntags, self.ncomp)
        self._number(node_to_tag, masks)

    def _number(self, node_to_tag, masks):
        nn, nc = len(node_to_tag), self.ncomp
        tag = node_to_tag.astype(np.int64)
        is_d = np.zeros((nn, nc), dtype=bool)
        has = tag > 0
        is_d[has] = masks[tag[has] - 1]
        flat = is_d.ravel()  # node-major, component-minor
        ids = np.empty(flat.shape, dtype=np.int64)
        ids[~flat] = np.arange(1, np.count_nonzero(~flat) + 1)
        ids[flat] = -np.arange(1, np.count_nonzero(flat) + 1)
        self.node_and_comp_to_dof = ids.reshape(nn, nc).astype(np.int32)
        self.nfree = int(np.count_nonzero(~flat))
        self.ndirichlet = int(np.count_nonzero(flat))
        node_of = np.repeat(np.arange(1, nn + 1, dtype=np.int32), nc)
        comp_of = np.tile(np.arange(1, nc + 1, dtype=np.int16), nn)
        self.free_dof_to_node = node_of[~flat]
        self.free_dof_to_comp = comp_of[~flat]
        self.dirichlet_dof_to_node = node_of[flat]
        self.dirichlet_dof_to_comp = comp_of[flat]
        cn = self.model.cell_node_ids() - 1
        n2d = self.node_and_comp_to_dof
        if nc == 1:
            cd = n2d[cn, 0]
        else:
            cd = np.concatenate([n2d[cn, k] for k in range(nc)], axis=1)
        self.cell_dof_ids = np.ascontiguousarray(cd, dtype=np.int32)

    def _number_simplex(self, labels):
        """Conforming numbering on the simplicial model (src/FESpaces/
        ConformingFESpaces.jl:367-575, _split_face_own_dofs_into_free_and_dirichlet_
        generic!): faces dimension by dimension in face-id order -- vertices, then
        (P2) edges -- each owning one dof; free ids 1.., Dirichlet ids -1.. in that
        order.  Cell dofs: the 4 vertex dofs, then the 6 edge dofs in TET edge order."""
        m = self.model
        ents = set()
        for tag in self.dirichlet_tags:
            ents.update(labels.tag_to_entities(tag))
        ents = np.array(sorted(ents), dtype=np.int32)
        vdir = np.isin(m.vertex_entity(), ents)
        face_dir = [vdir]
        if self.order == 2:
            face_dir.append(np.isin(m.edge_entity(), ents))
        isd = np.concatenate(face_dir)
        ids = np.empty(len(isd), dtype=np.int64)
        ids[~isd] = np.arange(1, np.count_nonzero(~isd) + 1)
        ids[isd] = -np.arange(1, np.count_nonzero(isd) + 1)
        nv = m.num_vertices()
        cells = m.cell_node_ids().astype(np.int64) - 1
        cd = ids[cells]
        if self.order == 2:
            cd = np.concatenate([cd, ids[nv + m.cell_edges()]], axis=1)
        self.cell_dof_ids = np.ascontiguousarray(cd, dtype=np.int32)
        self.nfree = int(np.count_nonzero(~isd))
        self.ndirichlet = int(np.count_nonzero(isd))
        # coordinates of every dof node: vertices, then edge midpoints (affine cells)
        xyz = m.node_coordinates()
        if self.order == 2:
            ev = m.edge_vertices()
            xyz = np.concatenate([xyz, 0.5 * (xyz[ev[:, 0]] + xyz[ev[:, 1]])])
        self.free_dof_coords = xyz[~isd]
        self.dirichlet_dof_coords = xyz[isd]
        self.node_and_comp_to_dof = None

    def _number_q2(self, labels, dirichlet_masks):
        """Conforming (vector) Q2 numbering on a Cartesian quadrilateral or hexahedral
        grid (src/FESpaces/ConformingFESpaces.jl:367-636): the grid topology numbers
        edges (and quad faces) by first touch; faces are visited by dimension
        (vertices, edges, quads, cells) in id order, each owning one node and its ncomp
        dofs (node-major); free ids 1.., Dirichlet ids -1.. in that order, a component
        of a tagged face being Dirichlet when its tag's mask says so.  Local dofs are
        component-blocked: node + nn * comp, nodes = the vertices, edges, (quad faces)
        and the interior in Gridap's face order (nn = 9 or 27)."""
        from .geometry import _first_touch, ncube_face_id
        m = self.model
        D = m.D
        nc = self.ncomp
        cells = m.cell_node_ids().astype(np.int64) - 1            # C x 2^D, x-fastest
        nn = m.num_nodes()
        C = cells.shape[0]
        if D == 2:
            ncube_edges = ((0, 1), (2, 3), (0, 2), (1, 3))
            edge_axes = (0, 0, 1, 1)
            ncube_faces, face_axes = (), ()
        else:
            ncube_edges = ((0, 1), (2, 3), (4, 5), (6, 7), (0, 2), (1, 3), (4, 6), (5, 7),
                           (0, 4), (1, 5), (2, 6), (3, 7))
            edge_axes = (0,) * 4 + (1,) * 4 + (2,) * 4
            ncube_faces = ((0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 4, 5), (2, 3, 6, 7),
                           (0, 2, 4, 6), (1, 3, 5, 7))
            face_axes = (2, 2, 1, 1, 0, 0)
        nle, nlf = len(ncube_edges), len(ncube_faces)
        ekeys = np.stack([np.minimum(cells[:, a], cells[:, b]) * 3 + ax
                          for (a, b), ax in zip(ncube_edges, edge_axes)], axis=1).reshape(-1)
        cell_edges, ne = _first_touch(ekeys)
        cell_edges = cell_edges.reshape(C, nle)
        if nlf:
            fkeys = np.stack([cells[:, f[0]] * 3 + ax
                              for f, ax in zip(ncube_faces, face_axes)], axis=1).reshape(-1)
            cell_faces, nf = _first_touch(fkeys)
            cell_faces = cell_faces.reshape(C, nlf)
        else:
            cell_faces, nf = np.zeros((C, 0), dtype=np.int64), 0
        # entity of every face = the box face holding all its vertices
        part = m.partition
        npts = [p + 1 for p in part]

        def ijk(v):
            out, r = [], v
            for d in range(D):
                out.append(r % npts[d])
                r = r // npts[d]
            return out
        nmask = 1 << D
        table = np.zeros((nmask, nmask), dtype=np.int32)
        for mk in range(nmask):
            for an in range(nmask):
                table[mk, an] = ncube_face_id(D, mk, an & ~mk)

        def entity(vs):                                   # vs: list of vertex-id arrays
            L = [ijk(v) for v in vs]
            free = np.zeros(len(vs[0]), dtype=np.int64)
            anchor = np.zeros(len(vs[0]), dtype=np.int64)
            for d in range(D):
                lo = np.all([l[d] == 0 for l in L], axis=0)
                hi = np.all([l[d] == part[d] for l in L], axis=0)
                free |= (~(lo | hi)).astype(np.int64) << d
                anchor |= hi.astype(np.int64) << d
            return table[free, anchor & ~free]

        he = np.array(ncube_edges)
        first_e = np.zeros(ne, dtype=np.int64)
        first_e[cell_edges.reshape(-1)] = np.arange(C * nle)
        ce, le = first_e // nle, first_e % nle
        ev = np.stack([cells[ce, he[le, 0]], cells[ce, he[le, 1]]], 1)
        ents = [m.get_face_labeling().vertex_entity.astype(np.int64), entity([ev[:, 0], ev[:, 1]])]
        if nlf:
            hf = np.array(ncube_faces)
            first_f = np.zeros(nf, dtype=np.int64)
            first_f[cell_faces.reshape(-1)] = np.arange(C * nlf)
            cf, lf = first_f // nlf, first_f % nlf
            fv = np.stack([cells[cf, hf[lf, k]] for k in range(4)], 1)
            ents.append(entity([fv[:, k] for k in range(4)]))
        interior = 3 ** D   # the n-cube's own face id (9 for QUAD, 27 for HEX)
        ents.append(np.full(C, interior, dtype=np.int64))
        # tag index per face (first tag holding the entity), then per component
        tags = self.dirichlet_tags
        ntags = len(tags)
        masks = (np.ones((max(ntags, 1), nc), dtype=bool) if dirichlet_masks is None else
                 np.array(dirichlet_masks, dtype=bool).reshape(ntags, nc))
        isd = []
        for e in ents:
            t = np.zeros(len(e), dtype=np.int64)
            for k, tag in enumerate(tags):
                hit = np.isin(e, labels.tag_to_entities(tag)) & (t == 0)
                t[hit] = k + 1
            dm = np.zeros((len(e), nc), dtype=bool)
            dm[t > 0] = masks[t[t > 0] - 1]
            isd.append(dm.reshape(-1))                    # node-major, comp-minor
        isd = np.concatenate(isd)
        ids = np.empty(len(isd), dtype=np.int64)
        ids[~isd] = np.arange(1, np.count_nonzero(~isd) + 1)
        ids[isd] = -np.arange(1, np.count_nonzero(isd) + 1)
        off = np.cumsum([0, nn, ne, nf])                  # first node of each dimension
        node_face = np.concatenate([cells, off[1] + cell_edges, off[2] + cell_faces,
                                    off[3] + np.arange(C)[:, None]], axis=1)   # C x nn_loc
        cd = np.concatenate([ids[node_face * nc + k] for k in range(nc)], axis=1)
        self.cell_dof_ids = np.ascontiguousarray(cd, dtype=np.int32)
        self.nfree = int(np.count_nonzero(~isd))
        self.ndirichlet = int(np.count_nonzero(isd))
        # coordinates of every Q2 node (vertices, edge / face / cell midpoints), per dof:
        # the means of their vertices, the Q1 cell map at the reference midpoints
        xyz = m.node_coordinates()
        parts = [xyz, xyz[ev].mean(1)]
        if nlf:
            parts.append(xyz[fv].mean(1))
        parts.append(xyz[cells].mean(1))
        nodes = np.concatenate(parts)
        comp = np.tile(np.arange(nc), len(nodes))
        pts = np.repeat(nodes, nc, axis=0)
        self.free_dof_coords, self.free_dof_comp = pts[~isd], comp[~isd]
        self.dirichlet_dof_coords, self.dirichlet_dof_comp = pts[isd], comp[isd]
        self.node_and_comp_to_dof = None

    def get_cell_dof_ids(self):
        return self.cell_dof_ids

    def num_free_dofs(self):
        return self.nfree

    def num_dirichlet_dofs(self):
        return self.ndirichlet

    def get_free_dof_ids(self):
        return range(1, self.nfree + 1)

    def get_triangulation(self):
        return Triangulation(self.model)

    def dirichlet_values(self):
        return np.zeros(self.ndirichlet)


def TestFESpace(model, reffe, **kwargs):
    return FESpace(model, reffe, **kwargs)


class TrialFESpace:
    """TrialFESpace(V, g): V with Dirichlet values g evaluated at the Dirichlet
    nodes (Lagrangian dofs are point evaluations)."""

    def __init__(self, V, g=None):
        self.space = V
        self.g = g
        vals = np.zeros(V.ndirichlet)
        if g is not None and V.ndirichlet and getattr(V, "conforming", False):
            if callable(g):
                gv = np.asarray(g(V.dirichlet_dof_coords), dtype=np.float64)
                if gv.ndim > 1:                           # vector data: pick each dof's comp
                    gv = gv[np.arange(V.ndirichlet), V.dirichlet_dof_comp]
                vals = gv.reshape(-1)
            else:
                vals = np.full(V.ndirichlet, float(g))
        elif g is not None and V.ndirichlet:
            xyz = V.model.node_coordinates()
            nodes = V.dirichlet_dof_to_node - 1
            if callable(g):
                gv = np.asarray(g(xyz[nodes]), dtype=np.float64)
                if gv.ndim > 1:
                    gv = gv[np.arange(len(nodes)), V.dirichlet_dof_to_comp - 1]
                vals = gv.reshape(-1)
            else:
                vals[:] = float(g)
        self._dvals = vals

    def __getattr__(self, name):
        return getattr(self.space, name)

    def dirichlet_values(self):
        return self._dvals


class FEFunction:
    """Free and Dirichlet dof values of a function of a (trial) space."""

    def __init__(self, space, free_values, dirichlet_values=None):
        self.space = space
        self.free_values = np.asarray(free_values)
        self.dirichlet_values = (space.dirichlet_values() if dirichlet_values is None
                                 else np.asarray(dirichlet_values))

    def get_free_dof_values(self):
        return self.free_values

    def nodal_values(self):
        """Values at the nodes (scalar Q1): free dofs and Dirichlet data merged."""
        V = self.space
        n2d = V.node_and_comp_to_dof[:, 0].astype(np.int64)
        out = np.where(n2d > 0, self.free_values[np.maximum(n2d, 1) - 1],
                       self.dirichlet_values[np.maximum(-n2d, 1) - 1] if V.ndirichlet
                       else 0.0)
        return out


def interpolate(u, U):
    """interpolate(u, U): the FE function of U whose free dofs take u's values at their
    nodes (Lagrangian dofs are point evaluations) and whose Dirichlet dofs are U's
    Dirichlet values (src/FESpaces/FESpaceInterface.jl, interpolate)."""
    V = U.space if isinstance(U, TrialFESpace) else U
    if getattr(V, "conforming", False):
        x, comp = V.free_dof_coords, getattr(V, "free_dof_comp", None)
    else:
        x = V.model.node_coordinates()[V.free_dof_to_node - 1]
        comp = V.free_dof_to_comp - 1 if V.ncomp > 1 else None
    vals = np.asarray(u(x), dtype=np.float64)
    if vals.ndim > 1:
        vals = vals[np.arange(len(vals)), comp]
    return FEFunction(U, vals.reshape(-1))
