"""Normal-mode tunes of a one-turn matrix (bmx_normal_mode_tunes, host only): the convention of
make_N (bmad/modules/mode3_mod.f90: eigen_decomp_6mat, order_evecs_by_plane_dominance),
pinned by the reference's mode3_test (its one-turn matrix t6.dat and the tunes make_N prints,
tests/golden/mode3/) and by closed forms: uncoupled blocks built from Twiss parameters give
their phase advances in [0, 2 pi) (above pi too), a symplectic change of basis close to the
identity keeps the tunes and their plane order, a 6x6 one-turn matrix with the rf off gives
its transverse tunes with n_dim = 4, and a defocusing block has no tune."""
import math
import os

import numpy as np
import pytest
from scipy import linalg

HERE = os.path.dirname(os.path.abspath(__file__))


def _block(beta, alpha, mu):
    g = (1 + alpha * alpha) / beta
    c, s = math.cos(mu), math.sin(mu)
    return np.array([[c + alpha * s, beta * s], [-g * s, c - alpha * s]])


def _S(n=6):
    S = np.zeros((n, n))
    for i in range(n // 2):
        S[2 * i, 2 * i + 1] = 1
        S[2 * i + 1, 2 * i] = -1
    return S


def test_make_n_tunes_of_the_reference_matrix(bmx):
    """mode3_test's t6.dat: the three tunes make_N prints ("make_N:ord_by_dom_tunes", REL 1e-6)."""
    t6 = np.loadtxt(os.path.join(HERE, "golden", "mode3", "t6.dat"))
    line = [ln for ln in open(os.path.join(HERE, "golden", "mode3", "make_n.correct"))
            if ln.startswith('"make_N:ord_by_dom_tunes"')][0]
    want = [float(v) for v in line.split()[3:6]]
    got = bmx.tracking.normal_mode_tunes(t6, 6)
    assert np.allclose(got, want, rtol=1e-6, atol=0), (got, want)


@pytest.mark.parametrize("mus", [(1.1, 2.3, 0.05), (5.2, 0.4, 6.1), (3.0, 3.5, 1.0)])
def test_uncoupled_blocks(bmx, mus):
    """Block-diagonal matrices from Twiss parameters: the tunes are the blocks' phase advances,
    x first, then y, then z."""
    M = linalg.block_diag(_block(12.0, -1.5, mus[0]), _block(4.0, 0.7, mus[1]),
                          _block(30.0, 0.2, mus[2]))
    got = bmx.tracking.normal_mode_tunes(M, 6)
    assert np.allclose(got, mus, rtol=0, atol=1e-12), got


def test_coupled_matrix_keeps_its_tunes(bmx):
    """N R N^-1 with R the uncoupled rotation and N = expm(S H) a symplectic coupling (H
    symmetric, small): the same tunes in the same plane order, to 1e-11."""
    rng = np.random.default_rng(3)
    mus = (1.7, 4.1, 0.3)
    R = linalg.block_diag(_block(8.0, 0.5, mus[0]), _block(3.0, -0.4, mus[1]),
                          _block(20.0, 0.0, mus[2]))
    H = rng.standard_normal((6, 6)) * 0.05
    H = H + H.T
    N = linalg.expm(_S() @ H)
    assert np.abs(N.T @ _S() @ N - _S()).max() < 1e-13
    M = N @ R @ np.linalg.inv(N)
    got = bmx.tracking.normal_mode_tunes(M, 6)
    assert np.allclose(got, mus, rtol=0, atol=1e-11), got


def test_transverse_tunes_with_the_rf_off(bmx):
    """A 6x6 one-turn matrix with the rf off (dispersion into x and px, the path length
    depending on x, px and pz, the pz row unity) has no longitudinal tune; n_dim = 4 takes the
    tunes of its x-y block."""
    M = np.eye(6)
    M[:4, :4] = linalg.block_diag(_block(10.0, 1.0, 2.2), _block(5.0, -0.3, 1.3))
    M[0, 5], M[1, 5] = 0.8, -0.02
    M[4, 0], M[4, 1], M[4, 5] = 0.01, 0.3, -5.0
    got = bmx.tracking.normal_mode_tunes(M, 4)
    assert np.allclose(got, (2.2, 1.3), rtol=0, atol=1e-12), got


def test_unstable_matrix_has_no_tune(bmx):
    """A block with |trace| > 2 (real eigenvalues): BMX_ERR_UNSTABLE."""
    M = np.eye(6)
    M[:2, :2] = [[math.cosh(0.5), 2 * math.sinh(0.5)], [math.sinh(0.5) / 2, math.cosh(0.5)]]
    M[2:4, 2:4] = _block(5.0, 0.0, 1.0)
    with pytest.raises(bmx.native.BmxError) as ei:
        bmx.tracking.normal_mode_tunes(M, 4)
    assert ei.value.code == bmx.native.ERR_UNSTABLE
