"""Solver.write_outputs after a device run: the files the reference writes,
with the fields sampled from the moment DOFs in HBM equal to the fields
sampled from the host distribution (the reference's route)."""
import json

import numpy as np
import pytest

from conftest import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("accflag", [0, 1])
def test_write_outputs_after_a_run(pkg, tmp_path, accflag):
    from pbte_amd.io_output import locate_points, sample_fields, sampling_grid
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    c.scheme.accflag = accflag
    c.iteration.tmax = 40
    c.output.dir = str(tmp_path / "out")
    c.output.field = c.output.run_record = c.output.runtime_log = True
    c.output.npz = c.output.tecplot = True
    print(f"case write_outputs F1_1_1 accflag {accflag}", flush=True)
    s = pkg.Solver(c)
    rec = s.run()
    written = s.write_outputs(rec)
    assert set(written) == {"run_record", "runtime_log", "field_npz", "field_dat"}
    rr = json.loads((tmp_path / "out" / "run_record.json").read_text())
    assert rr["iterations"] == rec.iterations and rr["mass"] == rec.mass
    line = (tmp_path / "out" / "RunTime.txt").read_text().splitlines()[-1]
    assert line.startswith(" GSIS" if accflag else "  CIS") and int(line.split()[6]) == rec.iterations
    f = np.load(tmp_path / "out" / "field.npz")
    px, py = sampling_grid()
    assert np.array_equal(f["px"], px) and np.array_equal(f["triid"], locate_points(s.mesh, px, py))
    uq = s.acc.full_uq().cpu().numpy() if accflag else None
    cx, cy, w = s.vel.flat()
    ref = sample_fields(s.mesh, s.basis, s.vdf, cx, cy, w, c.flow.cv, px, py, f["triid"], uq=uq)
    for k, name in enumerate(["T", "qx", "qy", "Nxx", "Nxy", "Nyy"]):
        np.testing.assert_allclose(f[name], ref[k], rtol=0, atol=1e-12 * max(1.0, np.abs(ref[k]).max()))
    assert np.abs(f["T"]).max() > 0.1 and (np.abs(f["Nxx"]).max() > 0) == bool(accflag)
    dat = (tmp_path / "out" / "field.dat").read_text().splitlines()
    assert len(dat) == 2 + 109 * 109


@pytest.mark.parametrize("accflag", [0, 1])
def test_restart_round_trip(pkg, tmp_path, accflag):
    """save_restart after some steps, then a solver with restart.enabled
    starts from that distribution (driver.py:100-139): the file holds the
    distribution bit for bit and the run record names it with its digest
    (the moments start at zero, as in the reference).  A file of the run's name
    with restart off is refused (FileExistsError); a file of the wrong size
    raises ValueError."""
    import hashlib
    import shutil
    c = pkg.Case.from_yaml(CASES / "F2_1_1.yaml")
    c.scheme.accflag = accflag
    c.velmesh.npole, c.velmesh.nazim = 4, 8
    c.iteration.tol = 0.0
    work = tmp_path / "case"
    work.mkdir()
    shutil.copy(CASES / "F2_1_1.yaml", work / "case.yaml")
    c.basedir = str(work)
    c.mesh.file = str((CASES / c.mesh.file).resolve())
    print(f"case restart round trip accflag {accflag}", flush=True)
    c.iteration.tmax = 20
    s = pkg.Solver(c)
    s.run()
    path = s.save_restart()
    assert path.name == "VDFP2T200NP4NA8.out"
    v20 = np.asarray(s.vdf).copy()
    assert path.read_bytes() == v20.astype("<f8").tobytes()
    with pytest.raises(FileExistsError):
        pkg.Solver(c)
    c.restart.enabled = True
    c.iteration.tmax = 10
    r = pkg.Solver(c)
    np.testing.assert_array_equal(np.asarray(r.vdf), v20)
    rec = r.run()
    assert rec.restart["enabled"] and rec.restart["path"] == str(path)
    assert rec.restart["sha256"] == hashlib.sha256(path.read_bytes()).hexdigest()
    path.write_bytes(path.read_bytes()[:-8])
    with pytest.raises(ValueError):
        pkg.Solver(c)


@pytest.mark.parametrize("precompute", [True, False])
def test_factorisation_count(pkg, precompute):
    """The run record's work counter (driver.py:84-85, 155-156): D N element
    factorisations once with stored operators, D N per sweep on the fly."""
    c = pkg.Case.from_yaml(CASES / "F1_1_1.yaml")
    c.velmesh.npole, c.velmesh.nazim = 4, 8
    c.iteration.tol, c.iteration.tmax = 0.0, 5
    c.performance.precompute_inverse = precompute
    print(f"case factorisation count precompute {precompute}", flush=True)
    s = pkg.Solver(c)
    rec = s.run()
    dn = s.ndir * s.n_tris
    assert rec.factorisation_count == (dn if precompute else 5 * dn)
    assert rec.sweep_count == 5
