"""The allocation-free operator entry point against the allocating one, through
the C ABI, on two calls in a row (the second with fewer bands), plus its
size limit and argument checks."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_wfc


def _register(gpu, gs):
    from oracle.dfpt import sqvc_init
    gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))


@pytest.mark.gpu
@pytest.mark.parametrize("key", ["3h2", "si8"])
def test_noalloc_operator_matches_allocating_entry(gpu, key, request):
    import torch
    from west_amd.lib import load, ptr
    from oracle.hamiltonian import apply_sternheimerop
    gs = request.getfixturevalue(f"gs_{key}")
    _register(gpu, gs)
    lib = load()
    alpha = np.array([gs.alpha_pv, 0.0])
    for m in (7, 3):        # second call smaller: stale scratch contents must not leak
        print(f"case noalloc {key} m={m}", flush=True)
        psi = random_wfc(gs, m, seed=70 + m)
        e = np.linspace(-0.6, 0.2, m)
        ref_entry = gpu.apply_sternheimerop_to_m_wfcs(gs.nbndval, psi, e, gs.alpha_pv)
        d_psi = torch.from_numpy(psi).to("cuda")
        d_e = torch.from_numpy(e).to("cuda")
        d_out = torch.full_like(d_psi, complex(np.nan, np.nan))
        rc = lib.wh_apply_sternheimerop_to_m_wfcs_noalloc(gs.nbndval, ptr(d_psi), ptr(d_out), ptr(d_e), ptr(alpha), m)
        assert rc == 0, lib.wh_last_error().decode()
        out = d_out.cpu().numpy()
        scale = np.abs(ref_entry).max()
        assert np.abs(out - ref_entry).max() / scale < 1e-13
        ref = apply_sternheimerop(gs, psi, e, gs.alpha_pv)
        assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-12


@pytest.mark.gpu
def test_noalloc_operator_limits(gpu, gs_si8):
    import torch
    from west_amd.lib import load, ptr
    gs = gs_si8
    _register(gpu, gs)
    lib = load()
    alpha = np.array([gs.alpha_pv, 0.0])
    # host arrays are refused (they would need staging buffers)
    psi = random_wfc(gs, 2, seed=3)
    out = np.zeros_like(psi)
    e = np.zeros(2)
    rc = lib.wh_apply_sternheimerop_to_m_wfcs_noalloc(gs.nbndval, ptr(psi), ptr(out), ptr(e), ptr(alpha), 2)
    assert rc < 0 and b"device arrays" in lib.wh_last_error()
    # more bands than the 64 MiB scratch holds: a clear error before any launch
    m = (64 << 20) // (32 * gs.basis.npw) + 8
    d_psi = torch.zeros((m, gs.npwx), dtype=torch.complex128, device="cuda")
    d_out = torch.zeros_like(d_psi)
    d_e = torch.zeros(m, dtype=torch.float64, device="cuda")
    rc = lib.wh_apply_sternheimerop_to_m_wfcs_noalloc(gs.nbndval, ptr(d_psi), ptr(d_out), ptr(d_e), ptr(alpha), m)
    assert rc < 0 and b"scratch" in lib.wh_last_error()
    assert float(d_out.abs().max()) == 0.0
