"""Sweeps at the size of the benchmark's default workload (256 x 256
jittered channel, 131 072 P2 triangles, 512 ordinates, hot and cold walls,
a periodic pair), checked through properties that do not depend on the
size: results that must agree bit for bit between kernels and schedules,
exact scaling by two, and the binary32 sweep's distance from the binary64
one.  The oracle finishes at this size only in minutes per sweep; it pins
the same kernels on the small meshes of the other test files."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def channel(pkg, tmp_path_factory):
    from pbte_amd.mesh import write_square_mesh
    work = tmp_path_factory.mktemp("full")
    mesh = work / "channel_256.msh"
    write_square_mesh(mesh, 256, layout="channel", jitter=0.25, seed=0)
    return pkg.Case.from_dict({
        "iteration": {"tol": 1e-8, "tmax": 10},
        "scheme": {"accflag": 0, "method": "source"},
        "velmesh": {"npole": 16, "nazim": 32},
        "dg": {"deg": 2},
        "flow": {"cv": 1.0, "vg": 1.0, "tau_r": 1e-3, "tau_n": 1e-4},
        "mesh": {"file": str(mesh)},
        "boundaries": [
            {"name": "SWall", "phyid": 11, "type": "thermalising", "temp": 0.0},
            {"name": "NWall", "phyid": 12, "type": "thermalising", "temp": 1.0},
            {"name": "Master", "phyid": 13, "type": "periodic", "xoff": -1.0},
            {"name": "Slave", "phyid": 14, "type": "periodic", "xoff": 1.0},
        ],
        "restart": {"enabled": False, "error_on_stale": False},
        "output": {"field": False, "runtime_log": False, "run_record": False},
        "performance": {"precision": "fp32"},
    }, basedir=str(work))


def _inputs(s, seed):
    import torch
    rng = np.random.default_rng(seed)
    P, N = s.ndof, s.n_tris
    mom = [rng.standard_normal((P, N)) for _ in range(3)]
    vdf = rng.standard_normal((s.ndir, N, P))
    fw = torch.from_numpy(rng.standard_normal(tuple(s.ctx.flux_wall.shape)))
    return mom, vdf, fw


def _sweep(s, mom, vdf, fw, dtype, homogeneous=False, scale=1.0):
    import torch
    s.mom.set_dofs(*(scale * m for m in mom))
    s.ctx.flux_wall.copy_(fw)
    out = torch.from_numpy(scale * vdf).to(device="cuda", dtype=dtype)
    s.ctx.sweep(s.mom, out, homogeneous=homogeneous)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_full_size_fused_sweep(pkg, channel):
    """Fused binary32 sweep at full size: the periodic partners staged in
    LDS give the global reads' bits; scaling the moments and the previous
    distribution by two scales a homogeneous sweep by two exactly; the
    binary32 result lies within 1e-4 (relative L2, over every value) of the
    binary64 sweep of the same inputs."""
    import torch
    print("case full-size channel 256, fused fp32", flush=True)
    s = pkg.Solver(channel)
    assert s.ctx.fused and s.ctx.sweep_path("fp32") == "fused"
    info = s.ctx.staging()
    assert info["on"] and info["staged_old_reads"] > 0
    mom, vdf, fw = _inputs(s, 5)
    staged = _sweep(s, mom, vdf, fw, torch.float32)
    s.ctx.staging(False)
    direct = _sweep(s, mom, vdf, fw, torch.float32)
    s.ctx.staging(True)
    assert np.array_equal(staged, direct)
    one = _sweep(s, mom, vdf, fw, torch.float32, homogeneous=True)
    two = _sweep(s, mom, vdf, fw, torch.float32, homogeneous=True, scale=2.0)
    assert np.array_equal(2.0 * one, two)
    ref = _sweep(s, mom, vdf, fw, torch.float64)
    err = np.linalg.norm(staged - ref) / np.linalg.norm(ref)
    print(f"binary32 vs binary64 sweep: relative L2 {err:.3e}", flush=True)
    assert err <= 1e-4


def test_full_size_pipelined_fp64_is_plain_sweep(pkg, channel):
    """Binary64 at full size: the pipelined kernel (pool, stage slots or
    global reads as the LDS allows) gives the plain level kernel's values
    bit for bit."""
    import torch
    case = channel
    case.performance.precision = "fp64"
    outs = []
    try:
        for variant in (0, 1):
            print(f"case full-size channel 256, fp64 variant {variant}", flush=True)
            s = pkg.Solver(case, sweep_variant=variant)
            mom, vdf, fw = _inputs(s, 9)
            outs.append(_sweep(s, mom, vdf, fw, torch.float64))
            del s
            torch.cuda.empty_cache()
    finally:
        case.performance.precision = "fp32"
    assert np.array_equal(outs[0], outs[1])
