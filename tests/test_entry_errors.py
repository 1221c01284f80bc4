"""CPU tests of the argument checks of the widened entry points (fluid, transport, radiative
heat, magnetostatics): each bad call returns RS_ERR_INVALID_ARG / RS_ERR_UNSUPPORTED with a
message before any device work, so these run without a GPU."""
import ctypes as C

import numpy as np
import pytest

INVALID, UNSUPPORTED = 1, 8


def call(rs, name, *args):
    lib = rs.load()
    rc = getattr(lib, name)(*args)
    return rc, (lib.rs_last_error() or b"").decode() if hasattr(lib, "rs_last_error") else ""


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def box(rs, n=2):
    xyz, tets = rs.kuhn_box(n, n, n)
    return np.ascontiguousarray(xyz.reshape(-1)), np.ascontiguousarray(tets.reshape(-1).astype(np.uint32))


def test_assembly_rejects_unsupported_dofs_per_node(rs):
    xyz, tets = box(rs)
    nn, ne = xyz.size // 3, tets.size // 4
    for dpn in (0, 2, 5):
        fixed = np.zeros(nn * max(dpn, 1), np.uint8)
        h = C.c_void_p()
        rc, msg = call(rs, "rs_assembly_create", nn, ptr(xyz), ne, ptr(tets), dpn, ptr(fixed), C.byref(h))
        assert rc == INVALID, dpn
        assert "dofs_per_node" in msg


def test_assembly_rejects_bad_node_ids(rs):
    xyz, tets = box(rs)
    nn, ne = xyz.size // 3, tets.size // 4
    bad = tets.copy()
    bad[5] = nn
    h = C.c_void_p()
    rc, msg = call(rs, "rs_assembly_create", nn, ptr(xyz), ne, ptr(bad), 4, ptr(np.zeros(4 * nn, np.uint8)),
                   C.byref(h))
    assert rc == INVALID and "out of range" in msg


def test_entry_points_reject_null_handles(rs):
    conf_f = rs.FluidConf(1.0, 1.0, 1.0, 0)
    conf_t = rs.TransportConf(1.0, 1.0, 0)
    d = C.c_void_p(0)
    checks = [
        ("rs_assemble_fluid_device", (None, C.byref(conf_f), d, d, d, d, d, d, d)),
        ("rs_fluid_update_device", (None, d, 0, 1.0, d, d, d, None)),
        ("rs_assemble_transport_device", (None, C.byref(conf_t), d, d, d, d, d, d)),
        ("rs_transport_update_device", (None, d, d, None)),
        ("rs_assemble_magnetostatics_device", (None, d, d, d)),
        ("rs_assemble_radiosity_device", (None, d, d, d, d, d)),
        ("rs_radiosity_process_device", (None, d, d)),
        ("rs_radiosity_matrix", (None, None)),
    ]
    for name, args in checks:
        rc, _ = call(rs, name, *args)
        assert rc == INVALID, name


def radiosity_args(rs):
    # two triangles, two patches
    xyz = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0], np.float64)
    tris = np.array([0, 1, 2, 1, 3, 2], np.uint32)
    return xyz, tris


def test_radiosity_create_checks_the_patch_book(rs):
    xyz, tris = radiosity_args(rs)
    h = C.c_void_p()
    cases = [
        (np.array([1, 1, 2], np.uint32), np.array([0, 1], np.uint32), "patch_ptr[0]"),
        (np.array([0, 2, 1], np.uint32), np.array([0, 1], np.uint32), "decreases"),
        (np.array([0, 1, 2], np.uint32), np.array([0, 2], np.uint32), "out of range"),
        (np.array([0, 1, 2], np.uint32), np.array([1, 1], np.uint32), "two patches"),
    ]
    for pp, pe, what in cases:
        rc, msg = call(rs, "rs_radiosity_create", 2, 4, ptr(xyz), 2, ptr(tris), ptr(pp), ptr(pe), C.byref(h))
        assert rc == INVALID and what in msg, (what, msg)
    bad_tris = np.array([0, 1, 2, 1, 3, 4], np.uint32)
    rc, msg = call(rs, "rs_radiosity_create", 2, 4, ptr(xyz), 2, ptr(bad_tris), ptr(np.array([0, 1, 2], np.uint32)),
                   ptr(np.array([0, 1], np.uint32)), C.byref(h))
    assert rc == INVALID and "node id" in msg
    rc, msg = call(rs, "rs_radiosity_create", 70000, 4, ptr(xyz), 2, ptr(tris),
                   ptr(np.zeros(70001, np.uint32)), ptr(np.array([0, 1], np.uint32)), C.byref(h))
    assert rc == UNSUPPORTED and "2^32" in msg
