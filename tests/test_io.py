"""Model exchange in Gridap's JSON format (no GPU).

The writer follows to_dict of src/Geometry/UnstructuredDiscreteModels.jl:51-57 (grid,
topology, labeling, version) with the parts' layouts from UnstructuredGrids.jl:198-208,
UnstructuredGridTopologies.jl:576-590, FaceLabelings.jl:641-652 and Tables.jl:1023-1028;
the reader follows the from_dict methods next to them.  Checks: the key sets of every
part, exact round trips (coordinates bit for bit, tables, entities, tags), the version
checks of check_dict (UnstructuredDiscreteModels.jl:59-75), face counts in closed form,
and that FE spaces built on the read model number their dofs exactly as on the model
that was written (Q1 on hexahedra, P1/P2 on jittered tetrahedra).
"""
import json

import numpy as np
import pytest

import gridap_hip as G
from gridap_hip import io as GIO


def _tet_model(n=3, seed=4):
    return G.simplexify(G.CartesianDiscreteModel((0, 1) * 3, (n, n, n))).jittered(seed)


def test_keys_follow_gridap_to_dict():
    d = GIO.to_dict(_tet_model())
    assert set(d) == {"grid", "topology", "labeling", "version"}
    assert d["version"] == "v0.16"
    assert set(d["grid"]) == {"node_coordinates", "Dp", "cell_node_ids", "reffes",
                              "cell_type", "orientation"}
    assert set(d["grid"]["cell_node_ids"]) == {"data", "ptrs"}
    assert d["grid"]["reffes"] == [{"orders": [1, 1, 1], "extrusion": [2, 2, 2],
                                    "space": "default", "value": "Float64"}]
    assert set(d["topology"]) == {"vertex_coordinates", "Dp", "Dc", "cell_type", "polytopes",
                                  "face_vertices_1", "face_vertices_2", "face_vertices_3",
                                  "orientation"}
    assert set(d["labeling"]) == {"D", "entities_0", "entities_1", "entities_2",
                                  "entities_3", "tags", "names"}
    dh = GIO.to_dict(G.CartesianDiscreteModel((0, 1) * 3, (2, 2, 2)))
    assert dh["grid"]["reffes"][0]["extrusion"] == [1, 1, 1]


@pytest.mark.parametrize("part", [(3, 2), (2, 3, 2)])
def test_cartesian_face_counts_and_roundtrip(part):
    model = G.CartesianDiscreteModel((0, 1) * len(part), part)
    s = GIO.to_json(model)
    d = json.loads(s)
    D = len(part)
    nx = part
    if D == 2:
        ne = nx[0] * (nx[1] + 1) + nx[1] * (nx[0] + 1)
    else:
        ne = (nx[0] * (nx[1] + 1) * (nx[2] + 1) + nx[1] * (nx[0] + 1) * (nx[2] + 1) +
              nx[2] * (nx[0] + 1) * (nx[1] + 1))
        nf = ((nx[0] + 1) * nx[1] * nx[2] + nx[0] * (nx[1] + 1) * nx[2] +
              nx[0] * nx[1] * (nx[2] + 1))
        assert len(d["topology"]["face_vertices_2"]["ptrs"]) - 1 == nf
    assert len(d["topology"]["face_vertices_1"]["ptrs"]) - 1 == ne
    m2 = GIO.from_json(s)
    np.testing.assert_array_equal(m2.node_coordinates(), model.node_coordinates())
    np.testing.assert_array_equal(m2.cell_node_ids(), model.cell_node_ids())
    np.testing.assert_array_equal(m2.vertex_entity(),
                                  model.get_face_labeling().vertex_entity)
    # boundary faces carry box-side entities, interior faces the interior entity
    ent = np.asarray(d["labeling"]["entities_%d" % (D - 1)])
    assert np.count_nonzero(ent == 3 ** D) == (ne if D == 2 else nf) - 2 * sum(
        np.prod(part) // part[k] for k in range(D))
    # a Q1 space on the read model numbers its dofs as on the Cartesian model
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, 1)
    V1 = G.TestFESpace(model, reffe, dirichlet_tags=["tag_01", "boundary"][1:])
    V2 = G.TestFESpace(m2, reffe, dirichlet_tags=["boundary"])
    np.testing.assert_array_equal(V1.cell_dof_ids, V2.cell_dof_ids)
    assert (V1.nfree, V1.ndirichlet) == (V2.nfree, V2.ndirichlet)


@pytest.mark.parametrize("order", [1, 2])
def test_tet_model_roundtrip_numbering(tmp_path, order):
    model = _tet_model(4, 9)
    path = tmp_path / "model.json"
    GIO.to_json_file(model, str(path))
    m2 = G.DiscreteModelFromFile(str(path)) if hasattr(G, "DiscreteModelFromFile") else \
        GIO.DiscreteModelFromFile(str(path))
    assert m2.node_coordinates().tobytes() == model.node_coordinates().tobytes()
    np.testing.assert_array_equal(m2.cell_node_ids(), model.cell_node_ids())
    np.testing.assert_array_equal(m2.cell_edges(), model.cell_edges())
    np.testing.assert_array_equal(m2.edge_entity(), model.edge_entity())
    reffe = G.ReferenceFE(G.lagrangian, G.Float64, order)
    for tags in (["boundary"], ["tag_21", "tag_01"]):
        V1 = G.TestFESpace(model, reffe, dirichlet_tags=tags)
        V2 = G.TestFESpace(m2, reffe, dirichlet_tags=tags)
        np.testing.assert_array_equal(V1.cell_dof_ids, V2.cell_dof_ids)
        assert (V1.nfree, V1.ndirichlet) == (V2.nfree, V2.ndirichlet)
        np.testing.assert_array_equal(V1.free_dof_coords, V2.free_dof_coords)
    # writing the read model again gives the same document
    assert GIO.to_json(m2) == path.read_text()


def test_version_checks():
    d = GIO.to_dict(_tet_model(2))
    legacy = {k: v for k, v in d.items() if k != "version"}
    with pytest.raises(ValueError, match="legacy"):
        GIO.from_dict(legacy)
    old = dict(d, version="v0.15")
    with pytest.raises(ValueError, match="outdated"):
        GIO.from_dict(old)
    GIO.from_dict(dict(d, version="v0.17"))
