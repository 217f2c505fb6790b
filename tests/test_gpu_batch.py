"""Batched sweeps and lockstep source iteration of several problems: the
same bits as each problem alone (the batch only merges the sweep launches)."""
import numpy as np
import pytest

from conftest import CASES

pytestmark = pytest.mark.gpu


def _case(pkg, tag, **kw):
    c = pkg.Case.from_yaml(CASES / f"{tag}.yaml")
    for k, v in kw.items():
        obj, attr = k.split("__")
        setattr(getattr(c, obj), attr, v)
    return c


def _cases(pkg, tmp_path, prec):
    from pbte_amd.mesh import write_square_mesh
    write_square_mesh(tmp_path / "j.msh", 12, jitter=0.25, seed=3)
    jit = _case(pkg, "F1_1_1", velmesh__npole=4, velmesh__nazim=12, performance__precision=prec)
    jit.mesh.file = str(tmp_path / "j.msh")
    return [_case(pkg, "F1_1_1", performance__precision=prec),
            _case(pkg, "F2_1_1", velmesh__npole=4, velmesh__nazim=16, performance__precision=prec),
            _case(pkg, "F3_1_1", performance__precision=prec),
            jit]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batched_sweep_is_bitwise_individual_sweeps(pkg, tmp_path, prec):
    """Four problems (thermalising, adiabatic and periodic walls, a jittered
    mesh; different ordinate sets) swept in one launch and one at a time."""
    import torch

    from pbte_amd.batch import SweepBatch
    cases = _cases(pkg, tmp_path, prec)
    print(f"case batch sweep {prec}", flush=True)
    solvers = [pkg.Solver(c) for c in cases]
    rng = np.random.default_rng(31)
    start = []
    for s in solvers:
        v = rng.standard_normal(s.sweep_vdf.shape)
        m = [rng.standard_normal((s.ndof, s.n_tris)) for _ in range(3)]
        fw = rng.standard_normal(s.ctx.flux_wall.shape)
        start.append((v, m, fw))

    def load(k):
        s = solvers[k]
        v, m, fw = start[k]
        s.vdf = v
        s.mom.set_dofs(*m)
        s.ctx.flux_wall.copy_(torch.from_numpy(fw))

    alone = []
    for k, s in enumerate(solvers):
        load(k)
        for _ in range(2):
            s.ctx.sweep(s.mom, s.sweep_vdf)
        alone.append(s.vdf)
    for k in range(len(solvers)):
        load(k)
    b = SweepBatch([s.ctx for s in solvers])
    for _ in range(2):
        b.sweep([s.mom for s in solvers], [s.sweep_vdf for s in solvers])
    b.close()
    for k, s in enumerate(solvers):
        assert np.array_equal(s.vdf, alone[k]), k


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_run_source_batch_is_each_solver_alone(pkg, tmp_path, prec):
    """Lockstep CIS of four problems gives each problem's own source
    iteration: the same iteration count, residual and mass histories and
    moments, bit for bit (fp32: the source iteration with the switch to
    binary64 sweeps at its own step)."""
    from pbte_amd.batch import run_source_batch
    cases = _cases(pkg, tmp_path, prec)
    for c in cases:
        c.iteration.tmax = 60
    print(f"case batch cis {prec}", flush=True)
    alone = [pkg.Solver(c)._run_source() for c in cases]
    recs = run_source_batch([pkg.Solver(c) for c in cases])
    for a, r in zip(alone, recs):
        assert r.iterations == a.iterations and r.converged == a.converged
        np.testing.assert_array_equal(r.residual_history, a.residual_history)
        np.testing.assert_array_equal(r.mass_history, a.mass_history)
        np.testing.assert_array_equal(r.temp_dofs, a.temp_dofs)
        assert r.sweep_count == a.sweep_count


def test_run_source_batch_mixed_degrees_and_gsis(pkg, tmp_path):
    """Lockstep source iteration over problems of degrees 1 and 2 (one sweep
    launch per degree), two of them GSIS (``accflag: 1``), fp32 with the
    binary64 finish: each problem's source iteration alone, bit for bit."""
    from pbte_amd.batch import run_source_batch
    cases = _cases(pkg, tmp_path, "fp32")
    cases[1].dg.deg = 1
    cases[3].dg.deg = 1
    for k in (0, 1):
        cases[k].scheme.accflag = 1
    for c in cases:
        c.iteration.tmax = 40
    print("case batch mixed degrees and gsis", flush=True)
    alone = [pkg.Solver(c)._run_source() for c in cases]
    recs = run_source_batch([pkg.Solver(c) for c in cases])
    for a, r in zip(alone, recs):
        assert r.scheme == a.scheme and r.deg == a.deg
        assert r.iterations == a.iterations and r.converged == a.converged
        np.testing.assert_array_equal(r.residual_history, a.residual_history)
        np.testing.assert_array_equal(r.mass_history, a.mass_history)
        np.testing.assert_array_equal(r.temp_dofs, a.temp_dofs)


def test_solvers_built_on_threads_are_the_solvers_built_alone(pkg, tmp_path):
    """``build_solvers`` sets up four problems (degrees 1 and 2, two with
    GSIS) on four host threads at once; their lockstep fp32 source
    iteration is each problem's iteration from a solver built alone, bit
    for bit."""
    from pbte_amd.batch import build_solvers, run_source_batch
    cases = _cases(pkg, tmp_path, "fp32")
    cases[1].dg.deg = 1
    for k in (0, 2):
        cases[k].scheme.accflag = 1
    for c in cases:
        c.iteration.tmax = 30
    print("case solvers built on threads", flush=True)
    alone = [pkg.Solver(c)._run_source() for c in cases]
    solvers = build_solvers(cases, workers=4)
    assert [s.n_tris for s in solvers] == [a.n_tris for a in alone]
    recs = run_source_batch(solvers)
    for a, r in zip(alone, recs):
        assert r.iterations == a.iterations and r.converged == a.converged
        np.testing.assert_array_equal(r.residual_history, a.residual_history)
        np.testing.assert_array_equal(r.temp_dofs, a.temp_dofs)


def test_run_source_batch_matches_oracle(pkg, tmp_path):
    """Every problem of a lockstep batch -- jittered meshes at degrees 1 and 2
    with different ordinate sets and walls, fp64 CIS -- against the oracle's
    own source iteration of that problem: residual and mass histories and
    the final temperature DOFs."""
    from oracle.cpu import Oracle
    from pbte_amd.batch import run_source_batch
    from pbte_amd.mesh import write_square_mesh
    from pbte_amd.problem import build_problem
    write_square_mesh(tmp_path / "j1.msh", 10, jitter=0.25, seed=5)
    write_square_mesh(tmp_path / "j2.msh", 14, jitter=0.25, seed=6)
    cases = []
    for mesh, deg, npole, nazim, tag in ((tmp_path / "j1.msh", 1, 4, 12, "F1_1_1"),
                                         (tmp_path / "j2.msh", 2, 4, 8, "F1_1_1"),
                                         (tmp_path / "j1.msh", 2, 6, 8, "F2_1_1"),
                                         (None, 1, 4, 16, "F3_1_1")):
        c = _case(pkg, tag, velmesh__npole=npole, velmesh__nazim=nazim, dg__deg=deg,
                  iteration__tmax=40, iteration__tol=0.0)
        if mesh is not None:
            c.mesh.file = str(mesh)
        cases.append(c)
    print("case batch vs oracle", flush=True)
    recs = run_source_batch([pkg.Solver(c) for c in cases])
    for c, r in zip(cases, recs):
        orc = Oracle(build_problem(c))
        orc.factorise()
        res, mass = orc.run_cis(tol=0.0, tmax=40)
        assert r.iterations == 40
        # the step residual falls to ~1e-10 in 40 steps: relative 1e-8, absolute 1e-14
        np.testing.assert_allclose(r.residual_history, res, rtol=1e-8, atol=1e-14)
        np.testing.assert_allclose(r.mass_history, mass, rtol=0, atol=1e-12)
        err = np.abs(r.temp_dofs - orc.ts).max() / np.abs(orc.ts).max()
        assert err <= 1e-11, err


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_falls_back_for_contexts_off_the_pipelined_kernel(pkg, tmp_path, prec):
    """Contexts whose sweeps do not run the pipelined kernel -- factors built
    on the fly (``precompute_inverse: false``) or the plain level kernel
    (``sweep_variant=1``) -- are swept one by one inside the batch: every
    problem still gets the iterates of its own source iteration, bit for
    bit."""
    from pbte_amd.batch import run_source_batch
    cases = _cases(pkg, tmp_path, prec)[:3]
    cases[0].performance.precompute_inverse = False
    for c in cases:
        c.iteration.tmax = 30
    print(f"case batch fallback {prec}", flush=True)

    def solvers():
        return [pkg.Solver(cases[0]), pkg.Solver(cases[1], sweep_variant=1), pkg.Solver(cases[2])]

    alone = [s._run_source() for s in solvers()]
    batch = solvers()
    assert batch[0].ctx.mode == "onthefly"
    recs = run_source_batch(batch)
    for a, r in zip(alone, recs):
        assert r.iterations == a.iterations and r.converged == a.converged
        np.testing.assert_array_equal(r.residual_history, a.residual_history)
        np.testing.assert_array_equal(r.temp_dofs, a.temp_dofs)


def test_run_refined_batch_is_each_refined_run_alone(pkg, tmp_path):
    """The fp32 defect-correction iteration of five problems in lockstep --
    degrees 1 and 2, thermalising, adiabatic and periodic walls (explicit
    periodic buffers in the batched launch), two of them GSIS -- gives each
    problem's own refined run, bit for bit: iteration counts, residual and
    mass histories, binary64 steps, moments."""
    from pbte_amd.batch import run_refined_batch
    cases = _cases(pkg, tmp_path, "fp32")
    cases.append(_case(pkg, "F3_1_1", performance__precision="fp32", dg__deg=1))
    cases[1].dg.deg = 1
    for k in (0, 2):
        cases[k].scheme.accflag = 1
    for c in cases:
        c.iteration.tmax = 400
    print("case batch refined", flush=True)
    alone = [pkg.Solver(c).run() for c in cases]
    for a in alone:
        assert a.diagnostics["precision_scheme"] == "defect-correction"
    calls = []
    solvers = [pkg.Solver(c) for c in cases]
    recs = run_refined_batch(solvers, callback=lambda k, step, res, s: calls.append(k))
    for a, r in zip(alone, recs):
        assert r.scheme == a.scheme and r.iterations == a.iterations and r.converged == a.converged
        np.testing.assert_array_equal(r.residual_history, a.residual_history)
        np.testing.assert_array_equal(r.mass_history, a.mass_history)
        np.testing.assert_array_equal(r.temp_dofs, a.temp_dofs)
        np.testing.assert_array_equal(r.qx, a.qx)
        assert r.diagnostics["fp64_steps"] == a.diagnostics["fp64_steps"]
    assert len(calls) == sum(r.iterations for r in recs)
