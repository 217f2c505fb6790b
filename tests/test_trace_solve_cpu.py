"""The GSIS trace factorisation: the reference's COLAMD call for small
systems; beyond the threshold the minimum-degree ordering with diagonal
pivots, which must give the same solution, and which falls back to the
reference's call when a diagonal pivot fails."""
import numpy as np

from conftest import CASES


def _trace_matrix(pkg):
    from pbte_amd.hdg import assemble_global_matrix, build_ba_sol, build_local_operators
    from pbte_amd.problem import build_problem
    c = pkg.Case.from_yaml(CASES / "F2_1_1.yaml")
    c.scheme.accflag = 1
    pb = build_problem(c)
    st = np.asarray(c.scheme.stabilisation, dtype=float).copy()
    if c.scheme.scale_by_hmin:
        st[1] /= pb.mesh.hmin
        st[2] /= pb.mesh.hmin
    _, aa_trace = build_local_operators(pb.mesh, pb.integrals, c.ndof_tri, c.ndof_fc, c.flow.cv,
                                        c.flow.tau_r, c.tau_c, st, c.scheme.acc_variant)
    ba = build_ba_sol(pb.mesh, pb.integrals, c.ndof_tri, c.ndof_fc, st, pb.bcdata.bc_type,
                      c.scheme.acc_variant)
    K, _ = assemble_global_matrix(pb.mesh, pb.integrals, ba, aa_trace, c.ndof_fc, pb.bcdata.bc_type,
                                  drop_tol=1e-30, periodic_flip=pb.bcdata.periodic_flip)
    return K


def test_trace_orderings_agree_and_fall_back(pkg, monkeypatch):
    from pbte_amd import acceleration
    K = _trace_matrix(pkg)
    b = np.random.default_rng(3).standard_normal(K.shape[0])
    lu0, o0 = acceleration.factorise_trace(K)
    assert o0 == "COLAMD"                          # below the threshold: the reference's call
    monkeypatch.setattr(acceleration, "_COLAMD_MAX", 0)
    lu1, o1 = acceleration.factorise_trace(K)
    assert o1 == "MMD_AT_PLUS_A"
    x0, x1 = lu0.solve(b), lu1.solve(b)
    assert np.abs(x1 - x0).max() <= 1e-11 * np.abs(x0).max()
    # diagonal-pivot factors that do not solve the system (forced here: SuperLU
    # itself pivots on an exactly zero diagonal) give way to the reference's call
    import scipy.sparse.linalg as sla
    real = sla.splu

    class Wrong:
        def solve(self, rhs):
            return np.zeros_like(rhs)

    def splu(A, permc_spec=None, **kw):
        return Wrong() if permc_spec == "MMD_AT_PLUS_A" else real(A, permc_spec=permc_spec, **kw)

    monkeypatch.setattr(sla, "splu", splu)
    lu2, o2 = acceleration.factorise_trace(K)
    assert o2 == "COLAMD"
    assert np.array_equal(lu2.solve(b), x0)
