"""The C-ABI library loads on the CPU and exports every symbol include/dec_hip.h
declares (no compute calls: those need a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "dec_hip.h")).read()
    return set(re.findall(r"^(?:int|const char \*)\s*(dec_\w+)\(", hdr, re.M))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("dec_mesh_create", "dec_op_create", "dec_op_apply", "dec_mg_create", "dec_mg_cycles",
              "dec_mg_solve", "dec_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from cspaces import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing
    assert set(_lib.SIGNATURES) == declared_symbols()


def test_version_and_error_without_gpu():
    from cspaces import _lib

    lib = _lib.lib()
    assert lib.dec_version().startswith(b"dec_hip")
    # argument errors are reported before any device work
    h = ctypes.c_void_p()
    assert lib.dec_mesh_icosphere(-1, ctypes.byref(h)) == 1
    assert b"k must be" in lib.dec_last_error()


def test_unsupported_degrees_are_reported_on_the_host():
    """The mirror refuses degree pairs the reference refuses (FastDEC.jl:361-372),
    with the reference's message, before touching the library."""
    import cspaces as cs
    import pytest

    with pytest.raises(cs.DecError, match="currently supported primal-dual wedges are 0-1, 1-0, and 1-1"):
        cs.dec_wedge_product_pd(1, 0, None)
    with pytest.raises(cs.DecError, match="currently supported dual-primal wedges are 0-1, 1-0, and 1-1"):
        cs.dec_wedge_product_dp(0, 1, None)


def test_new_entry_points_reject_bad_arguments_before_device_work():
    """Argument errors of the triangle-list mesh, the multigrid mode and the
    solve statistics are reported with a status and a message, without a GPU."""
    import numpy as np
    from cspaces import _lib

    lib = _lib.lib()
    h = ctypes.c_void_p()
    pts = np.zeros((3, 3))
    bad = np.array([[0, 1, 3]], np.int32)                       # vertex 3 of 3 points
    p = pts.ctypes.data_as(ctypes.c_void_p)
    assert lib.dec_mesh_from_triangles(3, 1, p, bad.ctypes.data_as(ctypes.c_void_p), 1, 0, ctypes.byref(h)) == 1
    assert b"out of range" in lib.dec_last_error()
    good = np.array([[1, 2, 3]], np.int32)
    assert lib.dec_mesh_from_triangles(3, 1, p, good.ctypes.data_as(ctypes.c_void_p), 1, 2, ctypes.byref(h)) == 1
    assert b"index_base" in lib.dec_last_error()
    assert lib.dec_mesh_from_triangles(3, 1, None, None, 1, 0, ctypes.byref(h)) == 1
    assert lib.dec_mg_create_mode(None, 2, 0, 0, 0, 2, 0, ctypes.byref(h)) == 1
    it, rr = ctypes.c_int32(), ctypes.c_double()
    assert lib.dec_op_solve_info(None, ctypes.byref(it), ctypes.byref(rr), None) == 1


def test_tetrahedral_entry_points_reject_bad_arguments_before_device_work():
    """The 3-D mesh constructors and accessors report argument errors with a
    status and a message, without a GPU."""
    import numpy as np
    from cspaces import _lib

    lib = _lib.lib()
    h = ctypes.c_void_p()
    pts = np.zeros((4, 3))
    p = pts.ctypes.data_as(ctypes.c_void_p)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    bad = np.array([[0, 1, 2, 4]], np.int32)                    # vertex 4 of 4 points
    assert lib.dec_mesh3_from_tets(4, 1, p, vp(bad), 1, 0, ctypes.byref(h)) == _lib.DEC_ERR_ARG
    assert b"out of range" in lib.dec_last_error()
    rep = np.array([[0, 1, 1, 2]], np.int32)
    assert lib.dec_mesh3_from_tets(4, 1, p, vp(rep), 1, 0, ctypes.byref(h)) == _lib.DEC_ERR_ARG
    assert b"repeated vertex" in lib.dec_last_error()
    good = np.array([[1, 2, 3, 4]], np.int32)
    assert lib.dec_mesh3_from_tets(4, 1, p, vp(good), 1, 2, ctypes.byref(h)) == _lib.DEC_ERR_ARG
    assert b"index_base" in lib.dec_last_error()
    assert lib.dec_mesh3_from_tets(4, 1, None, None, 1, 0, ctypes.byref(h)) == _lib.DEC_ERR_ARG
    z = np.zeros(1, np.int32)
    args = [1, 1, 1, 1] + [None] * 9 + [None, None, None, p, 0, ctypes.byref(h)]
    assert lib.dec_mesh3_create(*args) == _lib.DEC_ERR_ARG
    assert b"NULL array" in lib.dec_last_error()
    args = [4, 1, 1, 1, vp(z), vp(z), vp(z), vp(z), vp(z), vp(np.full(1, 5, np.int32)), vp(z), vp(z), vp(z),
            None, None, None, p, 0, ctypes.byref(h)]
    assert lib.dec_mesh3_create(*args) == _lib.DEC_ERR_ARG
    assert b"tetrahedron triangle out of range" in lib.dec_last_error()
    dim = ctypes.c_int32()
    assert lib.dec_mesh_dim(None, ctypes.byref(dim)) == _lib.DEC_ERR_ARG
    assert lib.dec_mesh3_counts(None, None, None, None, None) == _lib.DEC_ERR_ARG
    assert lib.dec_mesh3_export(None, None, None, None, None, None) == _lib.DEC_ERR_ARG
    assert lib.dec_mesh3_export_dual(None, None, None, None, None, None, None, None) == _lib.DEC_ERR_ARG
