"""The committed reference-run fixtures and the mesh writer's hole layout are well formed
(CPU only)."""
import numpy as np

from conftest import GOLDEN


def test_gsis_adiabatic_fixture():
    g = np.load(GOLDEN / "gsis_adiabatic.npz")
    assert bool(g["c0_converged"])
    h = g["c0_residual_history"]
    assert len(h) == int(g["c0_iterations"]) and h[-1] < 1e-8 <= h[-2]
    # the adiabatic side walls leave int T dA = 1/2 (test_deviations.py:63-65)
    assert abs(float(g["c0_mass_history"][-1]) - 0.5) < 1e-3
    assert g["c0_temp"].shape == g["c0_qx"].shape == g["c0_qy"].shape == (200,)


def test_hole_mesh_topology(pkg, tmp_path):
    """The square with a central circular hole: an annulus (V - E + T = 0),
    every edge on one or two triangles, the hole wall closed (each of its
    nodes on exactly two hole edges), no triangle centroid inside the
    circle, and the plain square unchanged when no hole is asked for."""
    from pbte_amd.mesh import read_gmsh22, square_mesh_text, write_square_mesh
    path = write_square_mesh(tmp_path / "hole.msh", 24, jitter=0.25, seed=0, hole=(0.5, 0.5, 0.2))
    g = read_gmsh22(path)
    et = np.asarray(g.elem_type)
    phy = np.asarray(g.elem_phyid)
    tris = g.elem_nodes[np.flatnonzero(et == 2)]
    hole = g.elem_nodes[np.flatnonzero(phy == 16), :2]
    assert len(hole) > 0 and g.physical_names[(1, 16)].startswith("Hole")
    edges = {}
    for t in tris:
        for k in range(3):
            e = tuple(sorted((int(t[k]), int(t[(k + 1) % 3]))))
            edges[e] = edges.get(e, 0) + 1
    assert set(edges.values()) <= {1, 2}
    used = np.unique(tris)
    assert used.size == g.nodes.shape[0]               # no orphan nodes
    assert g.nodes.shape[0] - len(edges) + len(tris) == 0
    deg = np.bincount(hole.reshape(-1), minlength=g.nodes.shape[0])
    assert set(np.unique(deg[deg > 0]).tolist()) == {2}
    c = g.nodes[tris].mean(axis=1)
    assert np.all(np.hypot(c[:, 0] - 0.5, c[:, 1] - 0.5) > 0.2)
    # each hole line runs as its triangle's edge: the domain on its left
    for a, b in hole:
        t = tris[np.flatnonzero(np.any(tris == a, axis=1) & np.any(tris == b, axis=1))]
        assert len(t) == 1
        k = list(t[0]).index(a)
        assert t[0][(k + 1) % 3] == b
    assert square_mesh_text(24, jitter=0.25, seed=0) == square_mesh_text(24, jitter=0.25, seed=0, hole=None)
