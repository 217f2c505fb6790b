"""The face integrals of ``pbte_face_integrals`` (host threads) against the
oracle's numpy restatement (``oracle/face_integrals.py``): bit for bit at
degrees 1 and 2 (no monomial power above 2), to a few units in the last
place at degrees 3 and 4 (libm's pow against numpy's); and both against
what the mathematics fixes: the face basis sums to one, so summing
int_tri_fc_fc over it gives int phi_m ds, which on a wall face is
int_tri_tri_fc's column; the partition of unity of the element basis
makes sum_m int phi_m ds the face length."""
import numpy as np
import pytest

from conftest import ROOT


def _setup(deg, mesh_name, periodic):
    from pbte_amd.basis import build_basis
    from pbte_amd.integration import face_partners
    from pbte_amd.mesh import build_mesh
    from pbte_amd.quadrature import gauss_legendre_core
    names = ["SWall", "NWall", "Master", "Slave"] if periodic else ["SWall", "NWall", "EWall", "WWall"]
    codes = [1, 1, 3, 3] if periodic else [1, 1, 1, 1]
    mesh = build_mesh(ROOT / "tests/golden/meshes" / mesh_name, [11, 12, 13, 14], codes, names,
                      [0, 0, -1.0, 1.0] if periodic else None, [0, 0, 0, 0] if periodic else None)
    basis = build_basis(deg)
    yq, lpq = gauss_legendre_core(15)
    other, ox, oy = face_partners(mesh, np.asarray(codes), [0, 0, -1.0, 1.0] if periodic else [0] * 4,
                                  [0] * 4)
    return mesh, basis, yq, lpq, other, ox, oy


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
@pytest.mark.parametrize("mesh_name,periodic", [("J_Nx6_jit25_s0.msh", False), ("C1_Nx6_Ny6.msh", True)])
def test_face_integrals_match_oracle(deg, mesh_name, periodic):
    from oracle.face_integrals import face_integrals
    from pbte_amd.integration import build_integrals
    mesh, basis, yq, lpq, other, ox, oy = _setup(deg, mesh_name, periodic)
    codes = [1, 1, 3, 3] if periodic else [1, 1, 1, 1]
    from pbte_amd.constants import NP_TRI_BY_DEG
    ig = build_integrals(mesh, basis, NP_TRI_BY_DEG[deg], 15, bc_codes=codes,
                         bc_xoff=[0, 0, -1.0, 1.0] if periodic else None, bc_yoff=None)
    want_fc, want_tt = face_integrals(mesh, basis.nodfun_tri, basis.nodfun_fc, basis.mono, yq, lpq, other, ox, oy)
    if deg <= 2:
        np.testing.assert_array_equal(ig.int_tri_fc_fc, want_fc)
        np.testing.assert_array_equal(ig.int_tri_tri_fc, want_tt)
    else:
        # entries that cancel to ~0 differ by rounding of the largest terms
        for got, want in ((ig.int_tri_fc_fc, want_fc), (ig.int_tri_tri_fc, want_tt)):
            np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    # face basis is a partition of unity: sum over it gives int phi_m ds
    edge = ig.int_tri_fc_fc.sum(axis=1)                     # [P, N, 3]
    wall = other < 0
    col = ig.int_tri_tri_fc[:, :, :, 0].transpose(2, 0, 1)  # [P, N, 3]
    np.testing.assert_allclose(edge[:, wall], col[:, wall], rtol=1e-12, atol=1e-15)
    # element basis is a partition of unity: sum_m int phi_m ds = face length
    length = mesh.face_len[mesh.tri_faces]
    np.testing.assert_allclose(edge.sum(axis=0), length, rtol=1e-12)
    # element x neighbour on a shared face: sum over both bases is the length
    two = ~wall
    np.testing.assert_allclose(ig.int_tri_tri_fc.sum(axis=(2, 3))[two], length[two], rtol=1e-12)


def test_face_integrals_threads_do_not_change_bits():
    from pbte_amd.integration import build_integrals
    mesh, basis, *_ = _setup(3, "J_Nx6_jit25_s0.msh", False)
    a = build_integrals(mesh, basis, 12, 15, bc_codes=[1, 1, 1, 1], threads=1)
    b = build_integrals(mesh, basis, 12, 15, bc_codes=[1, 1, 1, 1], threads=7)
    np.testing.assert_array_equal(a.int_tri_fc_fc, b.int_tri_fc_fc)
    np.testing.assert_array_equal(a.int_tri_tri_fc, b.int_tri_tri_fc)
