"""G-sphere and FFT-grid generation, pinned by the counts recorded in the
reference test-suite outputs (the only reference fixtures reachable without QE)."""
import numpy as np
import pytest

from west_amd.gsphere import count_sphere, cubic, fft_dims, good_fft_order, grid_index, make_gamma_basis

# test-suite/test001/ref/wstat.json:75,90,143-147 (SiH4, 20 bohr cube, Gamma, 25/100 Ry)
# test-suite/test014/ref/wstat.json:75,90,145 (fcc cell, Gamma)
# test-suite/test004/ref/wstat.json:76,91,144 (fcc Si, k-points: full spheres at k = 0)
FCC14 = np.array([[0.0, 6.7402562440539509, 6.7402562440539509],
                  [6.7402562440539509, 0.0, 6.7402562440539509],
                  [6.7402562440539509, 6.7402562440539509, 0.0]])
FCC04 = np.array([[-5.1306064283589663, 0.0, 5.1306064283589663],
                  [0.0, 5.1306064283589663, 5.1306064283589663],
                  [-5.1306064283589663, 5.1306064283589663, 0.0]])


@pytest.mark.parametrize("at,half,npw,ngm,nr", [
    (cubic(20.0), True, 8440, 67522, (64, 64, 64)),
    (FCC14, True, 646, 5209, (32, 32, 32)),
    (FCC04, False, 537, 4573, (24, 24, 24)),
])
def test_reference_counts(at, half, npw, ngm, nr):
    assert count_sphere(at, 25.0, half) == npw
    assert count_sphere(at, 100.0, half) == ngm
    assert fft_dims(at, 100.0) == nr


@pytest.mark.parametrize("L,npw,nr", [(18.0, 6147, 60), (20.52, 9139, 72), (30.78, 30723, 100)])
def test_workload_spheres(L, npw, nr):
    b = make_gamma_basis(cubic(L), 25.0)
    assert b.npw == npw
    assert b.nr == (nr, nr, nr)


def test_good_fft_order():
    assert [good_fft_order(n) for n in (63, 57, 65, 97, 107)] == [64, 60, 72, 100, 108]


def test_index_maps():
    b = make_gamma_basis(cubic(9.0), 12.0)
    assert b.gstart == 2 and np.all(b.mill[0] == 0)
    assert b.nl[0] == b.nlm[0] == 1
    # nlm is the index of -G, both 1-based, first index fastest
    assert np.array_equal(b.nl - 1, grid_index(b.mill, b.nr))
    assert np.array_equal(b.nlm - 1, grid_index(-b.mill, b.nr))
    # half space: no G together with -G (except G = 0)
    s = {tuple(m) for m in b.mill}
    assert all(tuple(-m) not in s for m in b.mill[1:])
    # ascending |G|^2
    assert np.all(np.diff(b.g2kin) >= -1e-12)
    assert np.allclose(b.g2kin, np.einsum("ij,ij->i", b.gvec, b.gvec))
