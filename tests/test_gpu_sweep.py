"""GPU parity of the HIP sweep path against the C oracle and the reference's
golden vectors.  Every comparison feeds both sides the same seeded inputs.

Tolerances: binary64 device path vs oracle <= 1e-12 relative (the device
uses compact upwind face blocks and a split source, which reassociate sums);
binary32 sweeps <= 2e-5 relative.
"""
import copy

import numpy as np
import pytest

from conftest import CASES, GOLDEN

pytestmark = pytest.mark.gpu


def _case(pkg, tag="F1_1_1", **kw):
    c = pkg.Case.from_yaml(CASES / f"{tag}.yaml")
    for k, v in kw.items():
        obj, attr = k.split("__")
        setattr(getattr(c, obj), attr, v)
    return c


def _pair(pkg, case, **skw):
    import torch  # noqa: F401
    from oracle.cpu import Oracle

    from pbte_amd.problem import build_problem
    print(f"case mesh={case.mesh.file} deg={case.dg.deg} dirs={case.velmesh.npole}x"
          f"{case.velmesh.nazim} prec={case.performance.precision} {skw}", flush=True)
    s = pkg.Solver(case, **skw)
    o = Oracle(build_problem(case))
    assert o.factorise() == 0
    return s, o


def _seed_state(s, o, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    vdf = rng.standard_normal((o.D, o.N, o.P))
    ts, qx, qy = (rng.standard_normal((o.P, o.N)) for _ in range(3))
    fw = rng.standard_normal((o.F, o.P)) * 0.1
    o.vdf[:] = vdf
    o.ts[:], o.qxs[:], o.qys[:] = ts, qx, qy
    o.flux_wall[:] = fw
    s.vdf = vdf
    s.mom.set_dofs(ts, qx, qy)
    s.ctx.flux_wall.copy_(torch.from_numpy(fw))
    return rng


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("tag", ["F1_1_1", "F2_1_1", "F3_1_1"])
def test_single_sweep_fp64_matches_oracle(pkg, tag):
    s, o = _pair(pkg, _case(pkg, tag))
    _seed_state(s, o, seed=1)
    s.ctx.sweep(s.mom, s.vdf_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 1e-12


def test_single_sweep_periodic_buffer(pkg):
    import torch
    s, o = _pair(pkg, _case(pkg, "F3_1_1"))
    rng = _seed_state(s, o, seed=2)
    pb = rng.standard_normal((o.D, o.F, o.P))
    o.pbuf[:] = pb
    s.ctx.use_pbuf = True
    s.ctx.ensure_pbuf(torch.float64).copy_(torch.from_numpy(pb))
    s.ctx.sweep(s.mom, s.vdf_d)
    o.sweep(use_pbuf=True)
    assert _rel(s.vdf, o.vdf) <= 1e-12


@pytest.mark.parametrize("deg,mesh", [(1, "A1_Nx11_Ny11"), (3, "A1_Nx6_Ny6"), (4, "A1_Nx6_Ny6")])
def test_single_sweep_degrees(pkg, deg, mesh):
    c = _case(pkg, "F2_1_1", dg__deg=deg)
    c.mesh.file = f"../meshes/{mesh}.msh"
    s, o = _pair(pkg, c)
    _seed_state(s, o, seed=3)
    s.ctx.sweep(s.mom, s.vdf_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 1e-12


@pytest.mark.parametrize("bundle", [1, 2, 4])
def test_bundles_are_bitwise_identical(pkg, bundle):
    """The level schedule and direction bundling never change a bit."""
    c = _case(pkg, "F1_1_1")
    s1, o = _pair(pkg, c)
    _seed_state(s1, o, seed=4)
    s1.ctx.sweep(s1.mom, s1.vdf_d)
    ref = s1.vdf
    s2 = pkg.Solver(c, max_bundle=bundle, block=128)
    _seed_state(s2, o, seed=4)
    s2.ctx.sweep(s2.mom, s2.vdf_d)
    assert np.array_equal(s2.vdf, ref)


@pytest.mark.parametrize("operator", ["stored", "fused"])
def test_single_sweep_fp32(pkg, operator):
    c = _case(pkg, "F2_1_1", performance__precision="fp32", performance__fp32_operator=operator)
    s, o = _pair(pkg, c)
    assert s.ctx.fused == (operator == "fused")
    _seed_state(s, o, seed=5)
    s.ctx.sweep(s.mom, s.vdf32_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 2e-5


@pytest.mark.parametrize("tag,deg", [("F1_1_1", 1), ("F1_1_1", 2), ("F1_1_1", 3), ("F2_1_1", 2),
                                     ("F3_1_1", 1), ("F3_1_1", 2), ("F3_1_1", 3)])
def test_fused_fp32_sweep_matches_oracle(pkg, tag, deg):
    """The fused binary32 sweep (every element operator built from the
    reference-element matrices and factorised in the sweep, nothing stored
    in binary32) against the oracle's binary64 sweep: thermalising,
    adiabatic and periodic walls, degrees 1-3, two sweeps in a row."""
    c = _case(pkg, tag, performance__precision="fp32", dg__deg=deg)
    s, o = _pair(pkg, c)
    assert s.ctx.fused and s.ctx.mode == "precomputed" and s.ctx._precomputed == {"fp64"}
    assert s.ctx.sweep_path("fp32") == "fused"
    _seed_state(s, o, seed=31)
    for _ in range(2):
        s.ctx.sweep(s.mom, s.vdf32_d)
        o.sweep()
        assert _rel(s.vdf, o.vdf) <= 2e-5


@pytest.mark.parametrize("bundle,depth", [(1, 3), (2, 3), (4, 3), (2, 2), (1, 4), (2, 4)])
def test_fused_sweep_bitwise_across_schedules(pkg, bundle, depth):
    """The fused sweep's per-item arithmetic does not depend on the
    schedule: bundles of 1, 2 or 4 directions and pipeline depths 2, 3 and
    4 give the same bits."""
    import torch
    c = _case(pkg, "F3_1_1", performance__precision="fp32")
    outs = []
    for b, dp in ((1, 3), (bundle, depth)):
        s = pkg.Solver(c, max_bundle=b, sweep_variant=dp)
        assert s.ctx.fused
        rng = np.random.default_rng(41)
        s.vdf = rng.standard_normal(s.sweep_vdf.shape)
        s.mom.set_dofs(*(rng.standard_normal((s.ndof, s.n_tris)) for _ in range(3)))
        s.ctx.flux_wall.copy_(torch.from_numpy(rng.standard_normal(s.ctx.flux_wall.shape)))
        for _ in range(2):
            s.ctx.sweep(s.mom, s.sweep_vdf)
        outs.append(s.vdf)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("tag,deg,bundle", [("F1_1_1", 1, 1), ("F1_1_1", 2, 2), ("F2_1_1", 2, 1),
                                            ("F3_1_1", 2, 2), ("F3_1_1", 1, 1)])
def test_split_fused_sweep_is_the_one_wave_fused_sweep(pkg, tag, deg, bundle):
    """The split kernel (the second half of the workgroup factorising the
    next level's operators into LDS while the first half solves) gives the
    one-wave fused sweep's bits: thermalising, adiabatic and periodic walls
    (previous-sweep reads), degrees 1 and 2, one and two directions per
    workgroup, two sweeps in a row."""
    import torch
    c = _case(pkg, tag, performance__precision="fp32", dg__deg=deg)
    outs = []
    for mode in ("split", True):
        s = pkg.Solver(c, max_bundle=bundle)
        assert s.ctx.set_fused(mode)
        assert s.ctx.sweep_path("fp32") == ("fused-split" if mode == "split" else "fused")
        rng = np.random.default_rng(51)
        s.vdf = rng.standard_normal(s.sweep_vdf.shape)
        s.mom.set_dofs(*(rng.standard_normal((s.ndof, s.n_tris)) for _ in range(3)))
        s.ctx.flux_wall.copy_(torch.from_numpy(rng.standard_normal(s.ctx.flux_wall.shape)))
        for _ in range(2):
            s.ctx.sweep(s.mom, s.sweep_vdf)
        outs.append(s.vdf)
    assert np.array_equal(outs[0], outs[1])


def test_fused_jittered_and_hole_meshes(pkg, tmp_path):
    """Fused binary32 sweeps on a jittered mesh and on the square with a hot
    hole (P3) against the oracle."""
    from pbte_amd.mesh import write_square_mesh
    write_square_mesh(tmp_path / "j.msh", 24, jitter=0.25, seed=0)
    c = _case(pkg, "F1_1_1", velmesh__npole=4, velmesh__nazim=16, performance__precision="fp32")
    c.mesh.file = str(tmp_path / "j.msh")
    s, o = _pair(pkg, c, max_bundle=2)
    assert s.ctx.fused
    _seed_state(s, o, seed=43)
    s.ctx.sweep(s.mom, s.vdf32_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 2e-5
    c = _hole_case(pkg, tmp_path / "hole.msh", deg=3)
    s, o = _pair(pkg, c)
    assert s.ctx.fused
    _seed_state(s, o, seed=44)
    s.ctx.sweep(s.mom, s.vdf32_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 2e-5


def test_moments_wallflux_residual_match_oracle(pkg):
    s, o = _pair(pkg, _case(pkg, "F2_1_1"))
    _seed_state(s, o, seed=6)
    s.ctx.moments(s.vdf_d, s.mom)
    o.moments()
    for a, b in ((s.mom.ts, o.ts), (s.mom.qxs, o.qxs), (s.mom.qys, o.qys), (s.mom.temp, o.temp),
                 (s.mom.qx, o.qx), (s.mom.qy, o.qy)):
        assert _rel(a, b) <= 1e-14
    s.ctx.wall_flux(s.vdf_d)
    o.wall_flux()
    walls = np.flatnonzero(np.abs(o.flux_wall).sum(axis=1) > 0)
    assert walls.size > 0
    assert _rel(s.ctx.flux_wall.cpu().numpy()[walls], o.flux_wall[walls]) <= 1e-13
    num_g, den_g = s.ctx.transport_residual_parts(s.mom, s.vdf_d)
    num_o, den_o = o.transport_residual()
    assert _rel(num_g, num_o) <= 1e-10
    assert _rel(den_g, den_o) <= 1e-10


@pytest.mark.parametrize("tag", ["F1_1_1", "F2_1_1", "F3_1_1"])
def test_cis_reproduces_golden_history(pkg, tag):
    """The reference's own converged source iteration (33/75/70 steps)."""
    c = _case(pkg, tag)
    print(f"case {tag} cis", flush=True)
    s = pkg.Solver(c)
    rec = s.run()
    g = np.load(GOLDEN / f"cis_{tag}.npz")
    assert rec.iterations == len(g["residual_history"])
    assert rec.converged
    np.testing.assert_allclose(rec.residual_history, g["residual_history"], rtol=1e-8, atol=0)
    np.testing.assert_allclose(rec.mass_history, g["mass_history"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(rec.temp, g["temp"], rtol=0, atol=1e-12)
    assert _rel(s.vdf[g["vdf_dirs"]], g["vdf"]) <= 1e-10


def test_frozen_cis_tau1_deg3(pkg):
    """FROZEN (1.0, 1e5, 3, 20, 40, CIS) -> (24, 0.24968704672388894)."""
    c = _case(pkg, "F1_1_1", flow__tau_r=1.0, flow__tau_n=1e5, dg__deg=3, velmesh__npole=20,
              velmesh__nazim=40)
    print("case frozen CIS tau1 deg3 20x40", flush=True)
    rec = pkg.Solver(c).run()
    assert rec.iterations == 24
    assert abs(rec.mass - 0.24968704672388894) <= 1e-12


def test_jittered_mesh_sweep(pkg, tmp_path):
    from pbte_amd.mesh import write_square_mesh
    write_square_mesh(tmp_path / "j.msh", 24, jitter=0.25, seed=0)
    c = _case(pkg, "F1_1_1", velmesh__npole=4, velmesh__nazim=16)
    c.mesh.file = str(tmp_path / "j.msh")
    s, o = _pair(pkg, c, max_bundle=2)
    _seed_state(s, o, seed=7)
    s.ctx.sweep(s.mom, s.vdf_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 1e-12


@pytest.mark.parametrize("tag,bundle,cap,depth", [("F1_1_1", 1, 0, 3), ("F2_1_1", 2, 0, 3), ("F3_1_1", 4, 0, 3),
                                                 ("F2_1_1", 1, 0, 2), ("F3_1_1", 2, 0, 2),
                                                 ("F1_1_1", 1, 16, 3), ("F3_1_1", 2, 2, 2)])
def test_pipelined_sweep_is_bitwise_plain_sweep(pkg, tag, bundle, cap, depth):
    """The pipelined kernel (LDS trace pool, or the global fallback when a
    capped pool overflows) computes the plain level kernel's values bit for
    bit, in both precisions."""
    import torch
    for prec in ("fp64", "fp32"):
        c = _case(pkg, tag, performance__precision=prec, performance__fp32_operator="stored")
        outs = []
        for variant in (depth, 1):
            print(f"case {tag} {prec} variant {variant} bundle {bundle} cap {cap}", flush=True)
            s = pkg.Solver(c, max_bundle=bundle, sweep_variant=variant, lds_slots_cap=cap)
            if variant != 1 and cap:
                info = s.ctx.schedule_info()
                print(info, flush=True)
                assert info["pool_slots"] <= cap and info["global_bundles"] > 0
            rng = np.random.default_rng(11)
            vdf = rng.standard_normal(s.sweep_vdf.shape)
            s.vdf = vdf
            s.mom.set_dofs(*(rng.standard_normal((s.ndof, s.n_tris)) for _ in range(3)))
            s.ctx.flux_wall.copy_(torch.from_numpy(rng.standard_normal(s.ctx.flux_wall.shape)))
            for _ in range(2):
                s.ctx.sweep(s.mom, s.sweep_vdf)
            outs.append(s.vdf)
        assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("prec,operator,bundle,use_pbuf", [("fp64", "stored", 1, False), ("fp64", "stored", 2, True),
                                                            ("fp32", "fused", 1, False), ("fp32", "fused", 2, True),
                                                            ("fp32", "stored", 2, False), ("fp32", "stored", 1, True)])
def test_staged_reads_are_bitwise_global_reads(pkg, prec, operator, bundle, use_pbuf):
    """The periodic partners swept later (previous-sweep values) and the
    explicit periodic buffer's entries, staged in LDS by the sweep's
    prologue, give the global reads' values bit for bit."""
    import torch
    c = _case(pkg, "F3_1_1", performance__precision=prec, performance__fp32_operator=operator)
    s = pkg.Solver(c, max_bundle=bundle)
    info = s.ctx.staging()
    print(prec, operator, bundle, use_pbuf, info, s.ctx.sweep_path(prec), flush=True)
    assert info["on"] and info["staged_periodic_reads"] > 0 and info["staged_old_reads"] > 0
    assert info["slots"] > 0 and info["slots_pbuf"] > 0
    outs = []
    for on in (True, False):
        s.ctx.staging(on)
        rng = np.random.default_rng(23)
        s.vdf = rng.standard_normal(s.sweep_vdf.shape)
        s.mom.set_dofs(*(rng.standard_normal((s.ndof, s.n_tris)) for _ in range(3)))
        s.ctx.flux_wall.copy_(torch.from_numpy(rng.standard_normal(s.ctx.flux_wall.shape)))
        if use_pbuf:
            pb = s.ctx.ensure_pbuf(s.sweep_vdf.dtype)
            pb.copy_(torch.from_numpy(rng.standard_normal(tuple(pb.shape))))
            s.ctx.use_pbuf = True
        try:
            for _ in range(2):
                s.ctx.sweep(s.mom, s.sweep_vdf)
        finally:
            s.ctx.use_pbuf = False
        outs.append(s.vdf)
    assert s.ctx.staging(True)["on"]
    assert np.array_equal(outs[0], outs[1])


def test_broken_cycle_sweep_matches_oracle(pkg):
    """Inflow lagged by a broken cycle reads the previous sweep's value (the
    reference order emits the element before its upwind neighbour)."""
    import torch
    from oracle.cpu import Oracle

    from pbte_amd.moments import Moments
    from pbte_amd.ordering import build_sweep_order
    from pbte_amd.problem import build_problem
    from pbte_amd.sweep import SweepContext
    c = _case(pkg, "F1_1_1", velmesh__npole=2, velmesh__nazim=4)
    c.mesh.file = "../meshes/A1_Nx6_Ny6.msh"
    pb = build_problem(c)
    cx, cy, _ = pb.vel.flat()
    m = pb.mesh
    a = 10
    b = int(m.tri_neighbour[a][m.tri_neighbour[a] >= 0][0])
    la = int(np.flatnonzero(m.tri_neighbour[a] == b)[0])
    lb = int(np.flatnonzero(m.tri_neighbour[b] == a)[0])
    n = -np.array([cx[0], cy[0]]) / np.hypot(cx[0], cy[0])
    m.tri_normal[a, la] = n
    m.tri_normal[b, lb] = n
    pb.order = build_sweep_order(m, pb.vel, on_cycle="break")
    assert pb.order.cycles_broken.sum() >= 1
    o = Oracle(pb)
    o.factorise()
    rng = np.random.default_rng(12)
    o.vdf[:] = rng.standard_normal(o.vdf.shape)
    o.ts[:], o.qxs[:], o.qys[:] = (rng.standard_normal(o.ts.shape) for _ in range(3))
    for variant in (0, 1):
        print(f"case broken cycle variant {variant}", flush=True)
        ctx = SweepContext.build(c, m, pb.integrals, pb.vel, pb.order, pb.bcdata, variant=variant)
        mom = Moments.zeros(o.P, o.N)
        mom.set_dofs(o.ts, o.qxs, o.qys)
        vdf = torch.from_numpy(o.vdf.copy()).cuda()
        ctx.sweep(mom, vdf)
        got = vdf.cpu().numpy()
        ref = Oracle(pb)
        ref.factorise()
        ref.vdf[:], ref.ts[:], ref.qxs[:], ref.qys[:] = o.vdf, o.ts, o.qxs, o.qys
        ref.sweep()
        assert _rel(got, ref.vdf) <= 1e-12


@pytest.mark.parametrize("tag,deg", [("F1_1_1", 2), ("F2_1_1", 1), ("F3_1_1", 3)])
def test_onthefly_sweep_matches_oracle(pkg, tag, deg):
    """performance.precompute_inverse: false selects the reference's
    on-the-fly mode (sweep.py:104-110, _kernels.py:242-269): every element
    operator is rebuilt and factorised where it is used, nothing stored."""
    c = _case(pkg, tag, performance__precompute_inverse=False, dg__deg=deg)
    s, o = _pair(pkg, c)
    assert s.ctx.mode == "onthefly"
    _seed_state(s, o, seed=13)
    s.ctx.sweep(s.mom, s.vdf_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 1e-12


@pytest.mark.parametrize("tag,deg,prec", [("F1_1_1", 3, "fp32"), ("F2_1_1", 2, "fp32"), ("F3_1_1", 1, "fp32"),
                                          ("F2_1_1", 3, "fp64")])
def test_mixed_mode_refined_sweep_matches_oracle(pkg, tag, deg, prec):
    """The "mixed" factor mode (binary64 factors do not fit, binary32 ones
    do; fp32 and fp64 runs): the binary64 sweep refines the stored
    binary32 inverse against the matrix-free operator; thermalising,
    adiabatic and periodic walls, degrees 1-3, against the oracle."""
    import torch
    c = _case(pkg, tag, performance__precision=prec, dg__deg=deg)
    from pbte_amd.problem import build_problem
    pb = build_problem(c)
    nbytes = float(pb.vel.ndir) * pb.mesh.n_tris * c.ndof_tri ** 2 * 8.0
    s, o = _pair(pkg, c, max_bytes=0.9 * nbytes)    # binary32 factors fit, binary64 ones do not
    assert s.ctx.mode == "mixed"
    _seed_state(s, o, seed=17)
    # an fp32 solver's vdf is its binary32 tensor: seed the binary64 one
    # too (the periodic mesh's broken cycles read the previous sweep)
    s.vdf_d.copy_(torch.from_numpy(o.vdf))
    s.ctx.sweep(s.mom, s.vdf_d)
    o.sweep()
    assert _rel(s.vdf_d.cpu().numpy(), o.vdf) <= 1e-12


def test_onthefly_sweep_fp32(pkg):
    c = _case(pkg, "F2_1_1", performance__precompute_inverse=False, performance__precision="fp32")
    s, o = _pair(pkg, c)
    assert s.ctx.mode == "onthefly"
    _seed_state(s, o, seed=14)
    s.ctx.sweep(s.mom, s.vdf32_d)
    o.sweep()
    assert _rel(s.vdf, o.vdf) <= 2e-5


def test_onthefly_cis_reproduces_golden_history(pkg):
    """The on-the-fly mode reaches the reference's converged source
    iteration (adiabatic walls: the wall emission is refreshed every step)."""
    c = _case(pkg, "F2_1_1", performance__precompute_inverse=False)
    print("case F2_1_1 cis onthefly", flush=True)
    s = pkg.Solver(c)
    assert s.ctx.mode == "onthefly"
    rec = s.run()
    g = np.load(GOLDEN / "cis_F2_1_1.npz")
    assert rec.iterations == len(g["residual_history"])
    np.testing.assert_allclose(rec.residual_history, g["residual_history"], rtol=1e-8, atol=0)
    np.testing.assert_allclose(rec.temp, g["temp"], rtol=0, atol=1e-12)


def _hole_case(pkg, path, n=20, deg=3):
    from pbte_amd.mesh import write_square_mesh
    write_square_mesh(path, n, jitter=0.25, seed=0, hole=(0.5, 0.5, 0.2))
    walls = [{"name": nm, "phyid": p, "type": "thermalising", "temp": 0.0}
             for nm, p in (("SWall", 11), ("NWall", 12), ("EWall", 13), ("WWall", 14))]
    walls.append({"name": "Hole", "phyid": 16, "type": "thermalising", "temp": 1.0})
    return pkg.Case.from_dict({
        "iteration": {"tol": 1e-8, "tmax": 100},
        "scheme": {"accflag": 0, "method": "source"},
        "velmesh": {"npole": 4, "nazim": 8},
        "dg": {"deg": deg},
        "flow": {"cv": 1.0, "vg": 1.0, "tau_r": 0.05, "tau_n": 0.005},
        "mesh": {"file": str(path)},
        "boundaries": walls,
        "restart": {"enabled": False, "error_on_stale": False},
        "output": {"field": False, "runtime_log": False, "run_record": False},
        "performance": {"precision": "fp32"},
    }, basedir=str(path.parent))


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_hole_mesh_mixed_factor_mode_matches_oracle(pkg, tmp_path, deg):
    """The square with a central hot hole (the 362^2 P3 1024-ordinate
    layout, here 20^2 with 32 ordinates) in the "mixed" factor mode that
    layout needs: binary32 factors stored, the binary64 sweep refining
    the binary32 inverse against the matrix-free binary64 operator.  Both
    sweeps against the oracle (which factorises every operator)."""
    import torch
    c = _hole_case(pkg, tmp_path / f"hole{deg}.msh", deg=deg)
    from pbte_amd.problem import build_problem
    pb = build_problem(c)
    nbytes = float(pb.vel.ndir) * pb.mesh.n_tris * c.ndof_tri ** 2 * 8.0
    s, o = _pair(pkg, c, max_bytes=0.9 * nbytes)   # binary32 (0.5 nbytes) fits, binary64 (1.0) does not
    assert s.ctx.mode == "mixed"
    assert int(np.count_nonzero(np.asarray(pb.bcdata.bc_type) >= 0)) >= 5
    _seed_state(s, o, seed=21)
    s.vdf_d.copy_(torch.from_numpy(o.vdf))
    s.ctx.sweep(s.mom, s.vdf_d)
    o.sweep()
    assert _rel(s.vdf_d.cpu().numpy(), o.vdf) <= 1e-12
    _seed_state(s, o, seed=22)
    s.ctx.sweep(s.mom, s.vdf32_d)
    o.sweep()
    assert _rel(s.vdf32_d.to(torch.float64).cpu().numpy(), o.vdf) <= 2e-5


@pytest.mark.parametrize("deg", [1, 3])
def test_pipelined_sweep_is_bitwise_plain_sweep_hole_mesh(pkg, tmp_path, deg):
    """P = 3 and P = 10 (the inverse read in the solve, not prefetched) on the
    square with a hole: the pipelined kernel gives the plain level kernel's
    bits in both precisions."""
    import torch
    for prec in ("fp64", "fp32"):
        c = _hole_case(pkg, tmp_path / f"hole{deg}.msh", n=16, deg=deg)
        c.performance.precision = prec
        c.performance.fp32_operator = "stored"
        outs = []
        for variant in (3, 1):
            print(f"case hole deg {deg} {prec} variant {variant}", flush=True)
            s = pkg.Solver(c, sweep_variant=variant)
            rng = np.random.default_rng(12)
            s.vdf = rng.standard_normal(s.sweep_vdf.shape)
            s.mom.set_dofs(*(rng.standard_normal((s.ndof, s.n_tris)) for _ in range(3)))
            for _ in range(2):
                s.ctx.sweep(s.mom, s.sweep_vdf)
            outs.append(s.vdf)
        assert np.array_equal(outs[0], outs[1])
