"""Gamma FFT helpers of the oracle against the textbook DFT and the
identities they must satisfy (FFT_kernel/fft_at_gamma.f90)."""
import numpy as np

from oracle.fft_at_gamma import (double_fwfft_gamma, double_invfft_gamma, fwfft, invfft, single_fwfft_gamma,
                                 single_invfft_gamma)
from oracle.hamiltonian import glbrak_gamma
from west_amd.gsphere import make_gamma_basis


def _brute_dft(nr, b, sign):
    n1, n2, n3 = nr
    i = np.arange(n1)[:, None]
    e1 = np.exp(sign * 2j * np.pi * i * i.T / n1)
    j = np.arange(n2)[:, None]
    e2 = np.exp(sign * 2j * np.pi * j * j.T / n2)
    k = np.arange(n3)[:, None]
    e3 = np.exp(sign * 2j * np.pi * k * k.T / n3)
    a = b.reshape(n3, n2, n1)
    a = np.einsum("zyx,xX->zyX", a, e1)
    a = np.einsum("zyx,yY->zYx", a, e2)
    a = np.einsum("zyx,zZ->Zyx", a, e3)
    return a.ravel()


def test_invfft_fwfft_conventions():
    rng = np.random.default_rng(0)
    nr = (4, 6, 5)
    b = rng.normal(size=120) + 1j * rng.normal(size=120)
    # invfft = sum exp(+iGr) unnormalised, first grid index fastest
    assert np.allclose(invfft(nr, b), _brute_dft(nr, b, +1), atol=1e-12)
    # fwfft = (1/N) sum exp(-iGr)
    assert np.allclose(fwfft(nr, b), _brute_dft(nr, b, -1) / 120, atol=1e-13)


def _basis():
    return make_gamma_basis(np.diag([7.0, 8.0, 7.5]), 9.0)


def _coeffs(b, m, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(m, b.npw)) + 1j * rng.normal(size=(m, b.npw))
    a[:, 0] = a[:, 0].real
    return a


def test_real_functions_and_packing():
    b = _basis()
    a = _coeffs(b, 2, 1)
    s1 = single_invfft_gamma(b, b.npw, a[0:1])[0]
    s2 = single_invfft_gamma(b, b.npw, a[1:2])[0]
    assert np.abs(s1.imag).max() < 1e-12 * np.abs(s1).max()
    d = double_invfft_gamma(b, b.npw, a[0:1], a[1:2])[0]
    assert np.allclose(d.real, s1.real, atol=1e-12)
    assert np.allclose(d.imag, s2.real, atol=1e-12)


def test_round_trip_and_parseval():
    b = _basis()
    a = _coeffs(b, 2, 2)
    psic = double_invfft_gamma(b, b.npw, a[0:1], a[1:2])
    b1, b2 = double_fwfft_gamma(b, b.npw, b.npw, psic)
    assert np.allclose(b1[0], a[0], atol=1e-12) and np.allclose(b2[0], a[1], atol=1e-12)
    r = single_fwfft_gamma(b, b.npw, b.npw, single_invfft_gamma(b, b.npw, a[0:1]))
    assert np.allclose(r[0], a[0], atol=1e-12)
    # Parseval: (1/N) sum_r psi(r)^2 = <a|a> (Gamma inner product)
    s = single_invfft_gamma(b, b.npw, a[0:1])[0].real
    nrm = glbrak_gamma(a[0:1], a[0:1], b.npw, 2)[0, 0]
    assert np.isclose(np.sum(s ** 2) / b.nnr, nrm, rtol=1e-12)


def test_g0_last_write_wins():
    """At G = 0 nl == nlm; the conj(a1 - i a2) store is the one kept (fft_at_gamma.f90:61-62)."""
    b = _basis()
    a1 = np.zeros((1, b.npw), complex)
    a2 = np.zeros((1, b.npw), complex)
    a1[0, 0] = 1.0 + 2.0j
    a2[0, 0] = 3.0 + 5.0j
    psic = double_invfft_gamma(b, b.npw, a1, a2)[0]
    expect = np.conj(a1[0, 0] - 1j * a2[0, 0])
    assert np.allclose(psic, expect)
