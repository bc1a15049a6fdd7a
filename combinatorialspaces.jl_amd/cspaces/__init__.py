"""cspaces: the MI355X DEC hot path of CombinatorialSpaces.jl behind its own
FastDEC / Multigrid call signatures.  Compute runs in libdec_hip.so (HIP,
gfx950); this package is the host-side mirror of the reference interface."""
from ._lib import DecError, LIB_PATH, lib  # noqa: F401
from .fastdec import (  # noqa: F401
    HIP, Barycenter, Circumcenter, DiagonalHodge, GeometricHodge, Mesh, Operator, apply_many, avg01_mat,
    dec_boundary, dec_c_wedge_product_, dec_differential, dec_dual_derivative, dec_hodge_star,
    dec_inv_hodge_star, dec_wedge_product, dec_wedge_product_dd, dec_wedge_product_dp, dec_wedge_product_pd,
    d0_p0_interpolation, dual_laplacian, dualize, fast_laplace_beltrami, fill_uniform, flat_sharp_mat, icosphere,
    interior_product_dd, kuhn_box, kuhn_box_arrays, lie_derivative_dd, make_sphere, make_sphere_arrays,
    mesh3_from_arrays, mesh_from_arrays, mesh_from_tetrahedra, mesh_from_triangles, p0_d0_interpolation, p2_d2_interpolation, p3_d3_interpolation,
    subdivision, triangulated_grid, ℒ_dd)
from .meshio import (  # noqa: F401
    load_json_acset, load_mesh, read_json_acset, read_mesh, read_obj, read_stl, save_obj, write_json_acset, write_obj)
from .multigrid import (  # noqa: F401
    BinarySubdivision, DirectMode, GalerkinMode, MGData, MultigridData, cg, dec_Delta_inv, full_multigrid,
    gauss_seidel, block_gauss_seidel, jacobi,
    multigrid_mu_cycles, multigrid_vcycles, multigrid_wcycles, solve)
from .batch import HeatBatch  # noqa: F401
from .learned import LearnedSmoother, learned_stats, make_pairs, sh_basis  # noqa: F401
