"""On-disk formats on the host: the VTK XML writers against their reader, the isosurface
mesher against geometry fixed by construction, cell-to-point averaging by brute force, the .pvd
collection (reference validation/test_export_bundle.py:37-49) and the probes' cell resolution."""

from __future__ import annotations

import math

import numpy as np
import pytest


def test_image_data_round_trip(tmp_path):
    from waam_solver_amd.export.vtkxml import read_vtk_arrays, write_image_data

    rng = np.random.default_rng(3)
    n = (5, 4, 3)
    arrays = {"Temperature_K": rng.random(60, dtype=np.float32) * 3000,
              "Cell_Flags": rng.integers(0, 5, 60).astype(np.int32),
              "f64": rng.random(60), "u8": rng.integers(0, 255, 60).astype(np.uint8)}
    p = tmp_path / "a.vti"
    write_image_data(p, n, 0.4, (1.5, 0.0, 0.0), arrays)
    attrs, back = read_vtk_arrays(p)
    assert attrs == {"extent": [0, 5, 0, 4, 0, 3], "origin": [1.5, 0.0, 0.0],
                     "spacing": [0.4, 0.4, 0.4]}
    assert list(back) == list(arrays)
    for k, v in arrays.items():
        assert back[k].dtype == v.dtype
        np.testing.assert_array_equal(back[k], v)
    with pytest.raises(ValueError):
        write_image_data(p, n, 0.4, (0, 0, 0), {"short": np.zeros(59, np.float32)})


def test_poly_data_round_trip(tmp_path):
    from waam_solver_amd.export.vtkxml import read_vtk_arrays, write_poly_data, write_poly_points

    pts = np.arange(12, dtype=np.float32).reshape(4, 3)
    write_poly_points(tmp_path / "p.vtp", pts, {"state": np.array([1, 2, 1, 2], np.int32)})
    attrs, back = read_vtk_arrays(tmp_path / "p.vtp")
    assert attrs["Points"] == 4 and attrs["Verts"] == 4 and attrs["Polys"] == 0
    np.testing.assert_array_equal(back["Points"].reshape(4, 3), pts)
    np.testing.assert_array_equal(back["Verts/connectivity"], np.arange(4))
    np.testing.assert_array_equal(back["Verts/offsets"], np.arange(1, 5))
    tri = np.array([[0, 1, 2], [0, 2, 3]])
    write_poly_data(tmp_path / "t.vtp", pts, {"v": np.ones(4, np.float32)}, triangles=tri)
    attrs, back = read_vtk_arrays(tmp_path / "t.vtp")
    assert attrs["Polys"] == 2 and attrs["Verts"] == 0
    np.testing.assert_array_equal(back["Polys/connectivity"], tri.reshape(-1))
    np.testing.assert_array_equal(back["Polys/offsets"], [3, 6])


def test_cell_to_point_is_the_mean_of_the_sharing_cells():
    from waam_solver_amd.export.isosurface import cell_to_point

    rng = np.random.default_rng(0)
    a = rng.random((4, 3, 5))
    p = cell_to_point(a)
    assert p.shape == (5, 4, 6)
    for x in range(5):
        for y in range(4):
            for z in range(6):
                cells = [a[i, j, k] for i in (x - 1, x) for j in (y - 1, y) for k in (z - 1, z)
                         if 0 <= i < 4 and 0 <= j < 3 and 0 <= k < 5]
                assert p[x, y, z] == pytest.approx(np.mean(cells), abs=1e-15)


def _mesh_checks(verts, tris):
    """Closed, consistently oriented 2-manifold: every directed edge once, its reverse once."""
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    directed = set(map(tuple, e))
    assert len(directed) == len(e), "a directed edge appears twice (orientation flips)"
    assert all((b, a) in directed for a, b in directed), "an edge has one triangle (a crack)"
    V, E, F = len(verts), len(e) // 2, len(tris)
    return V - E + F


def test_marching_tetrahedra_sphere_is_closed_oriented_and_encloses_its_volume():
    from waam_solver_amd.export.isosurface import marching_tetrahedra

    n, R = 24, 7.3
    c = np.array([11.7, 12.2, 11.9])
    x, y, z = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    r = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    P = 1.0 - r / (2 * R)            # 0.5 on the sphere, inside > 0.5
    verts, tris, data = marching_tetrahedra(P, 0.5, {"r": r})
    assert _mesh_checks(verts, tris) == 2  # Euler characteristic of a sphere
    # every vertex lies on a grid edge or diagonal between one point inside and one outside
    rv = np.linalg.norm(verts - c, axis=1)
    assert np.all(np.abs(rv - R) < 0.5)
    # interpolated point data: r itself is linear in P, so the interpolation returns R exactly
    np.testing.assert_allclose(data["r"], R, rtol=1e-12)
    # signed volume by the divergence theorem, positive for outward normals
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    vol = np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum() / 6.0
    assert vol == pytest.approx(4.0 / 3.0 * math.pi * R ** 3, rel=0.02)


def test_marching_tetrahedra_plane_is_exact():
    """A level set of a linear field is a plane: every vertex lies on it to rounding."""
    from waam_solver_amd.export.isosurface import marching_tetrahedra

    x, y, z = np.meshgrid(np.arange(9), np.arange(7), np.arange(8), indexing="ij")
    a = np.array([0.11, -0.07, 0.23])
    P = a[0] * x + a[1] * y + a[2] * z
    lvl = 0.77
    verts, tris, _ = marching_tetrahedra(P, lvl, {})
    assert len(tris) > 0
    np.testing.assert_allclose(verts @ a, lvl, atol=1e-12)
    # normals point down the gradient (from inside, P > level, outward)
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    nrm = np.cross(p1 - p0, p2 - p0)
    keep = np.linalg.norm(nrm, axis=1) > 1e-9
    assert np.all(nrm[keep] @ a < 0)
    empty = marching_tetrahedra(np.zeros((3, 3, 3)), 0.5, {"q": np.zeros(27)})
    assert empty[1].shape == (0, 3) and empty[2]["q"].size == 0


def test_write_pvd_relative_paths_and_times(tmp_path):
    from waam_solver_amd.export.bundle import write_pvd

    seq = tmp_path / "seq"
    v = []
    for n, name in ((0, "volume_step_000001.vti"), (1, "volume_step_000051.vti")):
        d = seq / f"frame_{n:04d}"
        d.mkdir(parents=True)
        (d / name).write_text("stub")
        v.append(str(d / name))
    write_pvd(seq / "sequence.pvd", v, times_s=[2e-5, 1.02e-3])
    text = (seq / "sequence.pvd").read_text()
    assert 'timestep="2e-05" file="frame_0000/volume_step_000001.vti"' in text
    assert 'timestep="0.00102" file="frame_0001/volume_step_000051.vti"' in text
    write_pvd(seq / "frames.pvd", v)
    assert 'timestep="1" file="frame_0001/volume_step_000051.vti"' in (seq / "frames.pvd").read_text()


class _G:
    def __init__(self, layout):
        self.nx, self.ny, self.nz, self.dx, self.layout = 7, 5, 4, 3e-4, layout


class _T:
    def __init__(self, layout, off=0.0):
        self.grid = _G(layout)
        self._window_offset_x_m = off


@pytest.mark.parametrize("layout", [0, 1])
def test_probe_cells_follow_the_window_and_the_layout(layout):
    from waam_solver_amd.export.probes import ProbeRecorder, ProbeSpec, _linear_cell

    tw = _T(layout)
    rec = ProbeRecorder()
    rec.add_grid(3, 2, 1, tw, "a")
    rec.add_world_mm(0.1, 0.65, 0.95, tw)
    assert [p.name for p in rec.probes] == ["a", "probe_1"]
    assert rec.probes[0].resolve(tw) == (3, 2, 1)
    assert rec.probes[1].resolve(tw) == (0, 2, 3)
    # the window moves 2 cells: the world-fixed probe moves 2 cells back, then leaves
    tw._window_offset_x_m = 2 * 3e-4
    assert rec.probes[0].resolve(tw) == (1, 2, 1)
    assert rec.probes[1].resolve(tw) is None
    assert ProbeSpec("z", 0.0, 0.0, 10.0).resolve(tw) is None
    shape = (4, 5, 7) if layout == 1 else (7, 5, 4)
    idx = (1, 2, 6) if layout == 1 else (6, 2, 1)
    assert _linear_cell(tw.grid, (6, 2, 1)) == np.ravel_multi_index(idx, shape)
    assert _linear_cell(tw.grid, None) == -1
    rec2 = ProbeRecorder.from_job_list([{"name": "tc1", "i": 1, "j": 1, "k": 1},
                                        {"x_mm": 0.5, "y_mm": 0.5, "z_mm": 0.5}], tw)
    assert [p.name for p in rec2.probes] == ["tc1", "probe_1"]


def test_marching_tetrahedra_level_on_grid_points():
    """A plate top on a cell face: the averaged point values are exactly the level on the face
    plane, so every vertex is a grid point of that plane and no triangle is degenerate."""
    from waam_solver_amd.export.isosurface import cell_to_point, marching_tetrahedra

    phi = np.zeros((6, 5, 8))
    phi[:, :, :3] = 1.0
    P = cell_to_point(phi)
    verts, tris, data = marching_tetrahedra(P, 0.5, {"phi": P})
    assert len(tris) > 0
    np.testing.assert_array_equal(verts[:, 2], 3.0)
    assert np.all(verts[:, :2] == np.round(verts[:, :2]))
    np.testing.assert_array_equal(data["phi"], 0.5)
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    area = np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1) / 2
    assert np.all(area > 0)
    # the triangles tile the 6 x 5 plate top once, all facing +z (out of the metal)
    assert area.sum() == pytest.approx(30.0)
    assert np.all(np.cross(p1 - p0, p2 - p0)[:, 2] > 0)


def test_sequence_schedule():
    """Frames after step 0, every multiple of every_n and the last step; at the frame cap the
    next due frame ends the run right after its step (reference export/bundle.py:117-170)."""
    from waam_solver_amd.export.bundle import _schedule

    assert _schedule(5, 2, 200) == (5, [0, 2, 4])
    assert _schedule(6, 2, 200) == (6, [0, 2, 4, 5])
    assert _schedule(10, 3, 2) == (7, [0, 3])
    assert _schedule(10, 3, 0) == (1, [])
    assert _schedule(1, 100, 200) == (1, [0])
    assert _schedule(0, 1, 5) == (0, [])
