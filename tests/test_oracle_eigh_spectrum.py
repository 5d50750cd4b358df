"""Oracle pinning: the symmetric eigensolver against numpy.linalg.eigh golden
fixtures (tests/golden/eigh_fixtures.npz) and the C2R transform against
numpy.fft.irfft(norm="forward") fixtures and closed forms (cuFFT's C2R
contract, calcIR.cu:790-791)."""

import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _eigh_cases():
    z = np.load(os.path.join(GOLD, "eigh_fixtures.npz"))
    names = sorted({k[:-2] for k in z.files})
    return [(n, z[n + "_A"], z[n + "_w"], z[n + "_V"]) for n in names]


def _projectors_agree(w, V, Vn, tol):
    # compare the spectral projectors of each group of (near-)equal eigenvalues
    i = 0
    n = len(w)
    scale = max(np.abs(w).max(), 1.0)
    while i < n:
        j = i + 1
        while j < n and w[j] - w[j - 1] < 1e-8 * scale:
            j += 1
        P = V[:, i:j] @ V[:, i:j].T
        Pn = Vn[:, i:j] @ Vn[:, i:j].T
        if np.abs(P - Pn).max() > tol:
            return False
        i = j
    return True


@pytest.mark.parametrize("case", _eigh_cases(), ids=lambda c: c[0])
def test_eigh_fp64_matches_lapack_fixture(case):
    name, A, wn, Vn = case
    w, V = O.eigh(O.F64, A)
    scale = max(np.abs(wn).max(), 1.0)
    assert np.abs(w - wn).max() < 1e-12 * scale * max(1, len(w))
    assert _projectors_agree(w, V, Vn, 1e-9)
    assert np.abs(V.T @ V - np.eye(len(w))).max() < 1e-12
    assert np.abs(A @ V - V * w).max() < 1e-11 * scale


@pytest.mark.parametrize("case", _eigh_cases(), ids=lambda c: c[0])
def test_eigh_fp32_matches_lapack_fixture(case):
    name, A, wn, Vn = case
    w, V = O.eigh(O.F32, A)
    scale = max(np.abs(wn).max(), 1.0)
    n = len(w)
    assert np.abs(w - wn).max() < 2e-6 * scale * max(1, np.sqrt(n))
    assert np.abs(V.T @ V - np.eye(n)).max() < 2e-6 * max(1, np.sqrt(n))


def test_eigh_uses_lower_triangle():
    # MAGMA "Upper" on the column-major view reads the row-major lower triangle
    rng = np.random.default_rng(2)
    a = rng.standard_normal((12, 12))
    L = np.tril(a)
    S = L + np.tril(a, -1).T
    junk = L + np.triu(rng.standard_normal((12, 12)), 1) * 100
    w1, _ = O.eigh(O.F64, S)
    w2, _ = O.eigh(O.F64, junk)
    assert np.allclose(w1, w2, rtol=1e-13, atol=1e-12)


def _irfft_cases():
    z = np.load(os.path.join(GOLD, "irfft_fixtures.npz"))
    keys = sorted(k[2:] for k in z.files if k.startswith("X_"))
    return [(k, z["X_" + k], z["y_" + k]) for k in keys]


@pytest.mark.parametrize("case", _irfft_cases(), ids=lambda c: c[0])
def test_c2r_matches_irfft_fixture(case):
    key, X, y = case
    n = int(key.split("_")[0])
    out = O.c2r(n, X)
    assert np.abs(out - y).max() < 1e-10 * max(1.0, np.abs(y).max())


def test_c2r_damped_exponential_closed_form():
    # X_j = r^j (j < J): y_k = 1 + 2 Re sum_{j=1}^{J-1} (r e^{i 2 pi k/n})^j
    n, J, r = 4096, 150, np.exp(-0.01 / (2 * 0.26))
    X = r ** np.arange(J)
    y = O.c2r(n, X)
    z = r * np.exp(2j * np.pi * np.arange(n) / n)
    geo = z * (1 - z ** (J - 1)) / (1 - z)
    assert np.allclose(y, 1 + 2 * geo.real, rtol=0, atol=1e-9)


def test_c2r_ignores_imaginary_parts_of_dc_and_nyquist():
    n = 16
    X = np.zeros(9, np.complex128)
    X[0] = 2 + 5j
    X[8] = 3 - 7j
    y = O.c2r(n, X)
    assert np.allclose(y, 2 + 3 * (-1.0) ** np.arange(n), atol=1e-13)
    # inputs beyond n/2 are not read
    X2 = np.concatenate([X, np.ones(5)])
    assert np.allclose(O.c2r(n, X2), y, atol=1e-13)
