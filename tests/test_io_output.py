"""The run's files: sampling grid, point location, field sampling, Tecplot
and RunTime.txt layouts (io_output.py:167-308 of the reference), pinned by
closed forms, a brute-force scan and fields the DG basis holds exactly."""
import numpy as np
import pytest

from conftest import CASES


@pytest.fixture(scope="module")
def problem(pkg, tmp_path_factory):
    from pbte_amd.mesh import write_square_mesh
    from pbte_amd.problem import build_problem
    d = tmp_path_factory.mktemp("io")
    write_square_mesh(d / "j.msh", 7, jitter=0.25, seed=5)
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    c.mesh.file = str(d / "j.msh")
    return build_problem(c)


def test_sampling_grid_closed_form(pkg):
    from pbte_amd.io_output import sampling_grid
    px, py = sampling_grid()
    assert px.shape == py.shape == (109,) and np.array_equal(px, py)
    assert px[0] == 0.0 and px[-1] == 1.0
    assert px[4] == 0.01 and px[104] == 0.99          # smootherstep(1/2) = 1/2, times 0.02
    ramp = lambda a: 0.02 * (6 * a**5 - 15 * a**4 + 10 * a**3)
    np.testing.assert_allclose(px[:5], [ramp(k / 8) for k in range(5)], rtol=0, atol=1e-17)
    np.testing.assert_allclose(np.diff(px[5:104]), 0.0098, rtol=0, atol=1e-15)
    np.testing.assert_allclose(px + px[::-1], 1.0, rtol=0, atol=1e-15)
    assert np.all(np.diff(px) > 0)
    ux, uy = sampling_grid(11, 5)
    assert np.array_equal(ux, np.linspace(0, 1, 11)) and np.array_equal(uy, np.linspace(0, 1, 5))


def _scan(mesh, px, py):
    """The first element in index order whose three signed sub-areas are
    >= -1e-12: a plain loop over points and elements."""
    xs, ys = mesh.nodes[mesh.tri_nodes, 0], mesh.nodes[mesh.tri_nodes, 1]
    out = np.full((px.size, py.size), -1)
    for i, x in enumerate(px):
        for j, y in enumerate(py):
            for k in range(mesh.n_tris):
                a = [0.5 * (x * (ys[k, m] - ys[k, (m + 1) % 3]) - y * (xs[k, m] - xs[k, (m + 1) % 3])
                            + xs[k, m] * ys[k, (m + 1) % 3] - xs[k, (m + 1) % 3] * ys[k, m]) for m in range(3)]
                if min(a) >= -1e-12:
                    out[i, j] = k
                    break
    return out


def test_locate_points_is_the_first_match_scan(pkg, problem):
    from pbte_amd.io_output import locate_points, sampling_grid
    mesh = problem.mesh
    px, py = sampling_grid(23, 19)
    px = np.concatenate([px, [-0.1, 1.2, 0.5]])      # outside, and a line of shared edges
    py = np.concatenate([py, [0.5, 1.0 / 7.0]])
    got = locate_points(mesh, px, py)
    assert np.array_equal(got, _scan(mesh, px, py))
    assert np.all(got[-3:-1, :] == -1)
    assert np.all(got[:-3, :-2] >= 0)


def _nodal(problem, f):
    """DOFs of f at each element's nodes (the basis is nodal)."""
    mesh, basis = problem.mesh, problem.basis
    xs, ys = mesh.nodes[mesh.tri_nodes, 0], mesh.nodes[mesh.tri_nodes, 1]
    xi, eta = basis.node_tri[:, 0], basis.node_tri[:, 1]
    X = xs[:, :1] + (xs[:, 1:2] - xs[:, :1]) * xi + (xs[:, 2:3] - xs[:, :1]) * eta
    Y = ys[:, :1] + (ys[:, 1:2] - ys[:, :1]) * xi + (ys[:, 2:3] - ys[:, :1]) * eta
    return f(X, Y).T                                    # [P, N]


def test_sampled_fields_reproduce_polynomials(pkg, problem):
    """A quadratic is exact in the P2 basis: sampled at any located point it
    is the polynomial itself; the stress rows combine as stated."""
    from pbte_amd.io_output import _fields, locate_points, sample_dofs, sampling_grid
    px, py = sampling_grid()
    triid = locate_points(problem.mesh, px, py)
    X, Y = np.meshgrid(px, py, indexing="ij")
    fs = [lambda x, y: 1 + 2 * x - 3 * y, lambda x, y: x * y - 0.5 * x * x, lambda x, y: 0 * x + 4.0]
    got = sample_dofs(problem.mesh, problem.basis, [_nodal(problem, f) for f in fs], px, py, triid)
    for f, g in zip(fs, got):
        np.testing.assert_allclose(g, f(X, Y), rtol=0, atol=1e-12)
    P = problem.basis.ndof_tri
    rows = [_nodal(problem, lambda x, y, k=k: x + k * y) for k in range(7)]
    uq = np.concatenate(rows, axis=0)
    out = _fields(problem.mesh, problem.basis, rows[0], rows[1], rows[2], px, py, triid, uq, P)
    lxx, lxy, lyx, lyy = (X + k * Y for k in (3, 4, 5, 6))
    np.testing.assert_allclose(out[3], -4 / 3 * lxx + 2 / 3 * lyy, rtol=0, atol=1e-11)
    np.testing.assert_allclose(out[4], -lxy - lyx, rtol=0, atol=1e-11)
    np.testing.assert_allclose(out[5], 2 / 3 * lxx - 4 / 3 * lyy, rtol=0, atol=1e-11)


def test_sample_fields_takes_angular_moments(pkg, problem):
    """f(d, x) = g(x) (1 + cx_d): T = g sum(w)/cv, qx = g sum(cx^2 w)."""
    from pbte_amd.io_output import locate_points, sample_dofs, sample_fields, sampling_grid
    cx, cy, w = problem.vel.flat()
    g = _nodal(problem, lambda x, y: 1 + x * y)        # [P, N]
    vdf = g.T[None, :, :] * (1 + cx)[:, None, None]
    px, py = sampling_grid(15, 15)
    triid = locate_points(problem.mesh, px, py)
    pT, pqx, pqy, z0, z1, z2 = sample_fields(problem.mesh, problem.basis, vdf, cx, cy, w, 2.0,
                                             px, py, triid)
    (pg,) = sample_dofs(problem.mesh, problem.basis, [g], px, py, triid)
    np.testing.assert_allclose(pT, pg * np.sum(w) / 2.0, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(pqx, pg * np.sum(cx * cx * w), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(pqy, pg * np.sum(cy * cx * w), rtol=0, atol=1e-12)
    assert not (z0.any() or z1.any() or z2.any())


def test_tecplot_and_runtime_log_layouts(pkg, tmp_path):
    from pbte_amd.io_output import append_runtime_log, write_tecplot
    px, py = np.array([0.0, 0.5, 1.0]), np.array([0.25, 0.75])
    a = np.arange(6.0).reshape(3, 2)
    path = write_tecplot(tmp_path / "f.dat", px, py, [a, -a], ["T", "qx"])
    lines = path.read_text().splitlines()
    assert lines[0] == ' VARIABLES="x","y","T","qx"'
    assert lines[1] == " ZONE I = 3 J = 2"
    assert len(lines) == 2 + 6 and all(len(s) == 1 + 4 * 16 for s in lines[2:])
    # x runs fastest; value a[i, j] at (px[i], py[j])
    vals = np.array([[float(s[1 + 16 * k:1 + 16 * (k + 1)]) for k in range(4)] for s in lines[2:]])
    assert np.array_equal(vals[:, 0], np.tile(px, 2)) and np.array_equal(vals[:, 1], np.repeat(py, 3))
    assert np.array_equal(vals[:, 2], a.T.ravel()) and np.array_equal(vals[:, 3], -a.T.ravel())
    assert lines[2] == " " + "    0.000000E+00    2.500000E-01    0.000000E+00   -0.000000E+00"
    log = append_runtime_log(tmp_path / "RunTime.txt", "GSIS", 1e-3, 1e-4, 200, 4, 16, 33, 12.34)
    append_runtime_log(log, "CIS", 10.0, 0.5, 51200, 16, 64, 1234567, 98765.0)
    assert log.read_text().splitlines() == [
        " GSIS  1.00E-03  1.00E-04   200     4    16      33                12.3",
        "  CIS  1.00E+01  5.00E-01 51200    16    64 1234567             98765.0"]
