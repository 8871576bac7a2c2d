"""gridap_hip: Gridap's assembly API over the MI355X assembly engine.

    model = CartesianDiscreteModel((0,1,0,1,0,1), (128,128,128))
    reffe = ReferenceFE(lagrangian, Float64, 1)
    V = TestFESpace(model, reffe, dirichlet_tags="boundary")
    U = TrialFESpace(V)
    Ω = Triangulation(model); dΩ = Measure(Ω, 2)
    κ = CellField(kappa_per_cell, Ω)
    a = lambda u, v: integrate(κ * inner(gradient(v), gradient(u)), dΩ)
    l = lambda v: integrate(v * 1.0, dΩ)
    op = AffineFEOperator(a, l, U, V)
    uh = solve(op)

Geometry, reference elements and dof numbering stay on the host (Gridap's
role); element integration, the CSC pattern, the scatter and the solver run on
the device through libgridap_hip.so (include/gridap_hip.h).
"""
from ._lib import GridapHipError
from .assemblers import (AffineFEOperator, LinearFESolver, PCGSolver, SparseMatrixAssembler,
                         SparseMatrixCSC, assemble_matrix, assemble_matrix_and_vector,
                         assemble_matrix_and_vector_inplace, assemble_matrix_inplace,
                         assemble_vector, assemble_vector_inplace, get_matrix, get_vector,
                         solve)
from .celldata import (CellField, DomainContribution, IsotropicStress, Measure, PLaplacian,
                       compose, divergence, dot, get_fe_basis, get_normal_vector,
                       get_trial_fe_basis, grad,
                       gradient, inner, integral, integrate, symmetric_gradient, ε)
from .fespaces import FEFunction, FESpace, TestFESpace, TrialFESpace, interpolate
from .geometry import (BoundaryTriangulation, CartesianDiscreteModel, FaceLabeling, Triangulation,
                       UnstructuredDiscreteModel, get_triangulation, simplexify)
from .reference_fes import Float64, Quadrature, ReferenceFE, VectorValue, lagrangian
from .nonlinear import (FEOperator, FESolver, NewtonRaphsonSolver, jacobian, residual,
                        residual_and_jacobian, solve_inplace)
from .io import DiscreteModelFromFile, from_json, from_json_file, to_json, to_json_file
from .multifield import (BlockMultiFieldStyle, BlockSparseMatrix, BlockVector,
                         ConsecutiveMultiFieldStyle, MultiFieldFESpace)

__all__ = [
    "AffineFEOperator", "CartesianDiscreteModel", "CellField", "DomainContribution",
    "FEFunction", "FESpace", "FaceLabeling", "Float64", "GridapHipError", "LinearFESolver",
    "Measure", "PCGSolver", "Quadrature", "ReferenceFE", "SparseMatrixAssembler",
    "SparseMatrixCSC", "TestFESpace", "TrialFESpace", "Triangulation",
    "UnstructuredDiscreteModel", "VectorValue", "simplexify",
    "assemble_matrix", "assemble_matrix_and_vector", "assemble_vector", "divergence", "dot",
    "assemble_matrix_inplace", "assemble_matrix_and_vector_inplace", "assemble_vector_inplace",
    "get_fe_basis", "get_matrix", "get_triangulation", "get_trial_fe_basis", "get_vector",
    "grad", "gradient", "inner", "integral", "integrate", "lagrangian", "solve",
    "IsotropicStress", "symmetric_gradient", "ε",
    "FEOperator", "FESolver", "NewtonRaphsonSolver", "PLaplacian", "compose", "jacobian",
    "residual", "residual_and_jacobian", "solve_inplace",
    "BoundaryTriangulation", "get_normal_vector", "interpolate",
    "DiscreteModelFromFile", "from_json", "from_json_file", "to_json", "to_json_file",
    "BlockMultiFieldStyle", "BlockSparseMatrix", "BlockVector", "ConsecutiveMultiFieldStyle",
    "MultiFieldFESpace",
]
