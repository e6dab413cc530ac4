"""The algebra behind the lubrication step's fused inverse and factor
(pbx_inverse_and_factor): with J the reversal permutation and
J R J = L L^T, the lower Cholesky factor of kBT inv(R) is sqrt(kBT) J L^{-T} J
and kBT inv(R) = kBT J L^{-T} L^{-1} J.  Checked in numpy against
numpy.linalg.cholesky(kBT inv(R)), including the identity pad the device
matrices carry.  No GPU."""
import numpy as np
import pytest


def _spd(n, seed, cond=1e3):
    rng = np.random.RandomState(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, np.log10(cond), n)
    return (Q * ev) @ Q.T


@pytest.mark.parametrize("n,np_", [(30, 30), (45, 64), (100, 128)])
def test_reversed_factor_gives_the_cholesky_factor_of_the_inverse(n, np_):
    kBT = 4.1e-2
    R = _spd(n, n)
    # the device layout: order np_, identity pad after the leading n x n block
    Rp = np.eye(np_)
    Rp[:n, :n] = R
    J = np.eye(np_)[::-1]
    Rrev = J @ Rp @ J
    # reversed pad carries kBT on its diagonal (so D and B keep identity pads)
    pad = np.arange(np_) < np_ - n
    Rrev[pad, pad] = kBT
    L = np.linalg.cholesky(Rrev)
    X = np.linalg.inv(L)
    B = np.sqrt(kBT) * J @ X.T @ J
    D = kBT * J @ (X.T @ X) @ J
    D_ref = kBT * np.linalg.inv(R)
    B_ref = np.linalg.cholesky(D_ref)
    assert np.allclose(np.tril(B), B)
    assert np.abs(B[:n, :n] - B_ref).max() <= 1e-11 * np.abs(B_ref).max()
    assert np.abs(D[:n, :n] - D_ref).max() <= 1e-11 * np.abs(D_ref).max()
    assert np.allclose(B[n:, n:], np.eye(np_ - n)) and np.allclose(D[n:, n:], np.eye(np_ - n))
    assert np.abs(B[n:, :n]).max(initial=0.0) == 0.0
    # the inverses of B's diagonal 64-tiles come from L's reversed tiles
    T = np_ // 64
    for k in range(T):
        Bk = B[64 * k:64 * k + 64, 64 * k:64 * k + 64]
        Lt = L[64 * (T - 1 - k):64 * (T - k), 64 * (T - 1 - k):64 * (T - k)]
        Wk = (Lt[::-1, ::-1]).T / np.sqrt(kBT)
        assert np.abs(Wk @ Bk - np.eye(64)).max() <= 1e-10
