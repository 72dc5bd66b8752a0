"""Wfreq's Lanczos chains on the GPU against the oracle (the Gamma
deflated Lanczos with full re-orthogonalisation, the brak products, and the
chi(i omega) update of dmati built from them), through the C ABI."""
import numpy as np
import pytest

from conftest import random_wfc


def _pc(gs, f):
    from oracle.hamiltonian import apply_alpha_pc
    return apply_alpha_pc(gs, f, 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("key", ["3h2", "si8"])
def test_lanczos_chains_vs_oracle(gpu, key, request):
    from oracle.dfpt import sqvc_init
    from oracle.lanczos import solve_deflated_lanczos_w_full_ortho as lanczos_ref
    from west_amd import wfreq
    gs = request.getfixturevalue(f"gs_{key}")
    gpu.set_ground_state(gs, sqvc=sqvc_init(gs.basis, gs.npwx))
    b = _pc(gs, random_wfc(gs, 5, seed=81))
    nls = 16
    print(f"case lanczos {key}", flush=True)
    a, bt, q, bn = wfreq.solve_deflated_lanczos_w_full_ortho(gs.nbndval, b, nls)
    ar, btr, qr, bnr = lanczos_ref(gs, gs.nbndval, b, nls)
    np.testing.assert_allclose(bn, bnr, rtol=1e-13)
    assert np.abs(a - ar).max() < 1e-10 * np.abs(ar).max()
    assert np.abs(bt - btr).max() < 1e-10 * np.abs(btr).max()
    assert np.abs(q - qr).max() < 1e-9 * np.abs(qr).max()


@pytest.mark.gpu
def test_chi_imaginary_frequency_pipeline_vs_oracle(gpu, gs_si8):
    """One occupied band's contribution to chi(i omega) in the PDEP basis:
    dvpsi = P_c [psi_v (sqvc dv_i)(r)] (solve_wfreq.f90:398-455), the chains,
    the brak products, the tridiagonal eigensolve and the dmati update."""
    from oracle.dfpt import sqvc_init
    from oracle.fft_at_gamma import single_fwfft_gamma, single_invfft_gamma
    from oracle import lanczos as ref
    from west_amd import wfreq
    gs = gs_si8
    sq = sqvc_init(gs.basis, gs.npwx)
    gpu.set_ground_state(gs, sqvc=sq)
    b = gs.basis
    npdep, nls, iv = 4, 20, gs.nbndval - 1
    pdep = random_wfc(gs, npdep, seed=5)
    psic = single_invfft_gamma(b, b.npw, gs.evc[iv:iv + 1])[0]
    pert_r = single_invfft_gamma(b, b.npw, pdep * sq[None, :])
    dvpsi = single_fwfft_gamma(b, b.npw, gs.npwx, psic[None, :] * pert_r)
    dvpsi = _pc(gs, dvpsi)
    freqs = np.array([0.0, 0.25, 1.5])
    out = {}
    for name, mod in (("gpu", wfreq), ("oracle", ref)):
        if name == "gpu":
            a, bt, q, bn = wfreq.solve_deflated_lanczos_w_full_ortho(gs.nbndval, dvpsi, nls)
            brak = wfreq.get_brak(dvpsi, q)
        else:
            a, bt, q, bn = ref.solve_deflated_lanczos_w_full_ortho(gs, gs.nbndval, dvpsi, nls)
            brak = ref.get_brak(gs, dvpsi, q)
        diag = np.zeros((npdep, nls))
        bks = np.zeros_like(brak)
        for ip in range(npdep):
            diag[ip], bks[ip] = mod.diago_lanczos(bn[ip], a[ip], bt[ip], brak[ip])
        out[name] = mod.dmati_update(np.zeros((npdep, npdep, len(freqs))), bks, diag, gs.et[iv], 1.0, freqs,
                                     gs.weight, b.omega)
    err = np.abs(out["gpu"] - out["oracle"]).max() / np.abs(out["oracle"]).max()
    print(f"case chi(i omega) pipeline: rel err {err:.2e}", flush=True)
    assert err < 1e-9
