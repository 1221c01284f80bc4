"""CPU tests of the library's host side: the C ABI loads and exports every declared symbol,
and the host-only model preparation (mesh, fields, node book) is right.  No GPU calls."""
import os
import re

import numpy as np
import pytest

from helpers import kuhn_box_np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "range_solver_amd.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s+\**\s*(rs_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(rs):
    lib = rs.load()
    names = declared_symbols()
    assert len(names) >= 40
    for name in names:
        assert hasattr(lib, name), name
    # the ctypes binding covers the same set
    assert set(rs.exported_symbols()) == set(names)


def test_spmv_shape_count_matches_header(rs):
    """rs_matrix_spmv_plan writes RS_SPMV_SHAPES tuning times: the header, the binding's shape
    tables and the library (a static_assert in the library) must agree."""
    with open(os.path.join(ROOT, "include", "range_solver_amd.h")) as f:
        m = re.search(r"#define\s+RS_SPMV_SHAPES\s+(\d+)", f.read())
    assert m and int(m.group(1)) == rs.SPMV_SHAPES
    assert len(rs.SPMV_SHAPE_NAMES) == rs.SPMV_SHAPES == len(rs.SPMV_COLUMN_BYTES)


def test_no_device_is_reported_cleanly(rs):
    rs.load()
    if rs.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(rs.RSError) as e:
        rs.init(0)
    assert e.value.code == 5  # RS_ERR_NO_DEVICE


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 4), (4, 4, 4)])
def test_kuhn_mesh(rs, dims):
    nx, ny, nz = dims
    xyz, tets = rs.kuhn_box(nx, ny, nz, 2.0, 1.0, 0.5)
    assert xyz.shape == ((nx + 1) * (ny + 1) * (nz + 1), 3)
    assert tets.shape == (6 * nx * ny * nz, 4)
    P = xyz[tets]
    vol = np.einsum("eij,eij->e", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])[:, None, :]
                    .repeat(1, 1), (P[:, 3] - P[:, 0])[:, None, :]) / 6.0
    assert np.all(vol > 0)
    assert vol.sum() == pytest.approx(1.0, rel=1e-13)
    # conforming: every face is shared by at most 2 tets, interior faces by exactly 2
    faces = np.sort(np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]],
                                    tets[:, [0, 1, 2]]]), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert counts.max() == 2
    n_boundary = (counts == 1).sum()
    assert n_boundary == 4 * (nx * ny + ny * nz + nx * nz)
    # same mesh as an independent numpy construction (same node numbering, same tets)
    xyz2, tets2 = kuhn_box_np(nx, ny, nz, 2.0, 1.0, 0.5)
    np.testing.assert_allclose(xyz, xyz2, rtol=0, atol=1e-15)
    assert np.array_equal(np.sort(tets, axis=1), np.sort(tets2, axis=1))


def test_plane_faces(rs):
    xyz, tets = rs.kuhn_box(4, 2, 3, 4.0, 1.0, 1.0)
    tris = rs.plane_faces(xyz, tets, 0, 4.0)
    assert tris.shape == (2 * 2 * 3, 3)
    assert np.all(np.abs(xyz[tris][:, :, 0] - 4.0) < 1e-12)
    area = 0.5 * np.linalg.norm(np.cross(xyz[tris[:, 1]] - xyz[tris[:, 0]], xyz[tris[:, 2]] - xyz[tris[:, 0]]), axis=1)
    assert area.sum() == pytest.approx(1.0, rel=1e-13)


def test_fields_deterministic_and_distributed(rs):
    a = rs.field_uniform(7, 200000)
    b = rs.field_uniform(7, 200000)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, rs.field_uniform(8, 200000))
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 5e-3 and abs(a.var() - 1 / 12) < 2e-3
    # prefix property: value i depends only on (seed, i)
    assert np.array_equal(rs.field_uniform(7, 1000), a[:1000])
    g = np.log(rs.field_lognormal(3, 400000, 0.0, 2.0))
    assert abs(g.mean()) < 0.02 and abs(g.std() - 2.0) < 0.02


def test_node_book_closed_form(rs, orc):
    n = 6
    xyz, tets = rs.kuhn_box(n, n, n)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    book, n_dof = rs.node_book(xyz.shape[0], tets, 1, fixed)
    assert n_dof == (n + 1) ** 3 - 2 * (n + 1) ** 2
    free = book != rs.RS_EOD
    assert np.array_equal(book[free], np.arange(n_dof, dtype=np.uint32))
    book2, n2 = orc.node_book(xyz.shape[0], tets, 1, fixed)
    assert n2 == n_dof and np.array_equal(book, book2)


def test_config_counts(rs):
    """Exact node/tet/unknown counts of the 20^3 heat configuration."""
    xyz, tets = rs.kuhn_box(20, 20, 20)
    assert xyz.shape[0] == 9261 and tets.shape[0] == 48000
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    _, n_dof = rs.node_book(xyz.shape[0], tets, 1, fixed)
    assert n_dof == 8379


@pytest.mark.parametrize("dpn,bins,max_coarse", [(1, (16, 16, 16), 512), (1, (10, 7, 5), 40), (3, (9, 9, 9), 100)])
def test_hierarchy_by_bins_matches_oracle(rs, orc, dpn, bins, max_coarse):
    """The host builder of the nested bins against the numpy restatement in the oracle (uneven
    bins, a 3-DOF book with fixed nodes)."""
    import learned_oracle as lo
    xyz, tets = rs.kuhn_box(16, 16, 16)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    book, _ = orc.node_book(xyz.shape[0], tets, dpn, fixed)
    maps = rs.hierarchy_by_bins(xyz, book, dpn, bins, max_coarse=max_coarse)
    ref = lo.bins_hierarchy(xyz, book, dpn, bins, max_coarse)
    assert len(maps) == len(ref) >= 1
    for m, r in zip(maps, ref):
        assert np.array_equal(m, r)
    # ids are in first-appearance order: the first occurrences of 0, 1, 2, ... are increasing
    for r in ref:
        first = np.unique(r, return_index=True)[1]
        assert np.all(np.diff(first) > 0)


@pytest.mark.parametrize("bins,max_coarse", [((9, 9, 9), 100), ((8, 3, 5), 300)])
def test_hierarchy_by_bins_per_component_matches_oracle(rs, orc, bins, max_coarse):
    """RS_AGG_COMPONENTS: the host builder against the oracle's per-component nested bins on a
    3-DOF book with a clamped face."""
    import learned_oracle as lo
    xyz, tets = rs.kuhn_box(16, 8, 8, 4.0, 1.0, 1.0)
    fixed = np.repeat(np.abs(xyz[:, 0]) < 1e-12, 3).astype(np.uint8)
    book, _ = orc.node_book(xyz.shape[0], tets, 3, fixed)
    maps = rs.hierarchy_by_bins(xyz, book, 3, bins, max_coarse=max_coarse, by_component=True)
    ref = lo.bins_hierarchy(xyz, book, 3, bins, max_coarse, by_component=True)
    assert len(maps) == len(ref) >= 2
    for m, r in zip(maps, ref):
        assert np.array_equal(m, r)
    assert int(maps[-1].max()) + 1 <= max_coarse


def test_hierarchy_by_bins_nested(rs, orc):
    """Nested coordinate bins: each level map halves the bins; composing the maps gives the
    direct binning of the finest positions at the coarser resolution."""
    xyz, tets = rs.kuhn_box(16, 16, 16)
    x = xyz[:, 0]
    fixed = ((np.abs(x) < 1e-12) | (np.abs(x - 1) < 1e-12)).astype(np.uint8)
    book, n_dof = orc.node_book(xyz.shape[0], tets, 1, fixed)
    maps = rs.hierarchy_by_bins(xyz, book, 1, (16, 16, 16), max_coarse=512)
    assert len(maps) == 2 and maps[0].size == n_dof
    sizes = [int(m.max()) + 1 for m in maps]
    assert sizes[1] == 512 and maps[1].size == sizes[0]
    direct16, n16 = rs.aggregates_by_bins(xyz, book, 1, (16, 16, 16))
    direct8, n8 = rs.aggregates_by_bins(xyz, book, 1, (8, 8, 8))
    assert np.array_equal(maps[0], direct16) and sizes[0] == n16
    assert np.array_equal(maps[1][maps[0]], direct8) and n8 == 512
    # every aggregate of a 2x2x2 binning holds at most 8 nodes
    assert np.bincount(maps[1]).max() <= 8
    with pytest.raises(rs.RSError):
        rs.hierarchy_by_bins(xyz, book, 1, (16, 16, 16), max_coarse=512, max_levels=1)
