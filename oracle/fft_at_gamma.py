"""Gamma-point FFT helpers (restatement of ``FFT_kernel/fft_at_gamma.f90``).

Grid layout: QE's, first index fastest, ``ir = i + nr1*(j + nr2*k)``; a grid
array of length ``nnr`` is viewed as shape ``(nr3, nr2, nr1)``.

QE conventions pinned by the tests: ``invfft`` is the unnormalised
``sum_G c(G) exp(+iG.r)``; ``fwfft`` is ``(1/N) sum_r f(r) exp(-iG.r)``.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

WORKERS = -1


def invfft(nr, b: np.ndarray) -> np.ndarray:
    """In place semantics of QE ``invfft`` on one or a batch of grids (..., nnr)."""
    n1, n2, n3 = nr
    shp = b.shape
    a = b.reshape(shp[:-1] + (n3, n2, n1))
    out = sfft.ifftn(a, axes=(-3, -2, -1), norm="forward", workers=WORKERS)
    return out.reshape(shp)


def fwfft(nr, a: np.ndarray) -> np.ndarray:
    n1, n2, n3 = nr
    shp = a.shape
    x = a.reshape(shp[:-1] + (n3, n2, n1))
    out = sfft.fftn(x, axes=(-3, -2, -1), norm="forward", workers=WORKERS)
    return out.reshape(shp)


def double_invfft_gamma(basis, n, a1, a2):
    """``fft_at_gamma.f90:30-70``.  a1, a2: (..., nx) -> psic (..., nnr).

    b(nl) = a1 + i a2 ; b(nlm) = conj(a1 - i a2)  (the nlm store is second,
    so at G = 0, where nl == nlm, the conjugated value is the one kept)."""
    a1 = np.asarray(a1)
    a2 = np.asarray(a2)
    lead = a1.shape[:-1]
    b = np.zeros(lead + (basis.nnr,), dtype=np.complex128)
    nl = basis.nl[:n] - 1
    nlm = basis.nlm[:n] - 1
    b[..., nl] = a1[..., :n] + 1j * a2[..., :n]
    b[..., nlm] = np.conj(a1[..., :n] - 1j * a2[..., :n])
    return invfft(basis.nr, b)


def double_fwfft_gamma(basis, n, nx, a):
    """``fft_at_gamma.f90:72-120``.  a: (..., nnr) -> b1, b2 (..., nx)."""
    f = fwfft(basis.nr, a)
    nl = basis.nl[:n] - 1
    nlm = basis.nlm[:n] - 1
    fp = (f[..., nl] + f[..., nlm]) * 0.5
    fm = (f[..., nl] - f[..., nlm]) * 0.5
    lead = a.shape[:-1]
    b1 = np.zeros(lead + (nx,), dtype=np.complex128)
    b2 = np.zeros(lead + (nx,), dtype=np.complex128)
    b1[..., :n] = fp.real + 1j * fm.imag
    b2[..., :n] = fp.imag - 1j * fm.real
    return b1, b2


def single_invfft_gamma(basis, n, a1):
    """``fft_at_gamma.f90:122-160``: b(nl) = a ; b(nlm) = conj(a)."""
    a1 = np.asarray(a1)
    lead = a1.shape[:-1]
    b = np.zeros(lead + (basis.nnr,), dtype=np.complex128)
    b[..., basis.nl[:n] - 1] = a1[..., :n]
    b[..., basis.nlm[:n] - 1] = np.conj(a1[..., :n])
    return invfft(basis.nr, b)


def single_fwfft_gamma(basis, n, nx, a):
    """``fft_at_gamma.f90:162-203``: b1 = fwfft(a)(nl), zero beyond n."""
    f = fwfft(basis.nr, a)
    lead = a.shape[:-1]
    b1 = np.zeros(lead + (nx,), dtype=np.complex128)
    b1[..., :n] = f[..., basis.nl[:n] - 1]
    return b1
