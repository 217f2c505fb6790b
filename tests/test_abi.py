"""The C-ABI library loads without a device and exports every entry point
that include/pbte_hip.h declares; host-only entry points run here."""
import re

import numpy as np

from conftest import ROOT


def _declared():
    text = (ROOT / "include" / "pbte_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbte_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree(pkg):
    from pbte_amd import _native
    names = _declared()
    assert len(names) >= 14
    assert set(names) == set(_native.EXPORTED)


def test_library_exports_every_declared_symbol(pkg):
    from pbte_amd import _native
    lib = _native.lib()
    for name in _declared():
        assert hasattr(lib, name), name


def test_ctx_create_rejects_bad_dims_without_device(pkg):
    import ctypes as C

    from pbte_amd import _native
    dims = _native.PbteDims(ndir=4, n_tris=10, nd=7, nf=3, n_faces=20, nbc=1)
    su = _native.PbteSetup()
    err = C.create_string_buffer(256)
    h = _native.lib().pbte_ctx_create(C.byref(dims), C.byref(su), None, err, 256)
    assert not h
    assert b"unsupported" in err.value


def test_build_order_runs_on_host(pkg):
    from pbte_amd.mesh import build_mesh
    from pbte_amd.ordering import build_sweep_order
    from pbte_amd.velocity import build_velocity_mesh
    m = build_mesh(ROOT / "tests/golden/meshes/A1_Nx6_Ny6.msh", [11, 12, 13, 14], [1, 1, 1, 1])
    so = build_sweep_order(m, build_velocity_mesh(4, 8))
    assert so.order.shape == (32, m.n_tris)
    for row in so.order:
        assert np.array_equal(np.sort(row), np.arange(m.n_tris))
