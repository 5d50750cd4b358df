"""Writes the golden fixtures of tests/golden/ (run from the repository root:
python tests/golden/make_golden.py).

  eigh_fixtures.npz   numpy.linalg.eigh (LAPACK syevd via numpy) of seeded
                      symmetric matrices: random, graded, degenerate and a
                      32-chromophore exciton Hamiltonian built from a seeded
                      synthetic frame with the project's own field/map/TDC
                      formulas written out below in numpy.
  irfft_fixtures.npz  numpy.fft.irfft(X, n, norm="forward"): the unnormalised
                      complex-to-real inverse DFT of cuFFT's C2R contract, for
                      zero-padded correlation-function-like inputs.
"""

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def eigh_cases():
    rng = np.random.default_rng(20261202)
    cases = {}
    for n in (1, 2, 8, 33, 64):
        a = rng.standard_normal((n, n))
        cases["random%d" % n] = (a + a.T) * 40.0
    d = np.geomspace(1e-2, 1e3, 24) * rng.choice([-1, 1], 24)
    q, _ = np.linalg.qr(rng.standard_normal((24, 24)))
    cases["graded24"] = (q * d) @ q.T
    q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    lam = np.repeat([-250.0, 10.0, 300.0], 10)
    cases["degenerate30"] = (q * lam) @ q.T
    t = np.diag(rng.uniform(-400, 400, 40)) + np.diag(rng.uniform(-50, 50, 39), 1)
    cases["tridiagonal40"] = t + np.triu(t, 1).T
    return cases


def exciton_hamiltonian():
    """A 32-chromophore Hamiltonian (16 waters) in the form of calcIR.cu:1306-1404
    with randomly oriented molecules on a 4x4x1 grid, E drawn uniformly."""
    rng = np.random.default_rng(7)
    nmol = 16
    box = 1.2
    o = np.array([[0.15 + 0.3 * (i % 4), 0.15 + 0.3 * (i // 4), 0.6] for i in range(nmol)])
    u = rng.standard_normal((2 * nmol, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    E = rng.uniform(-0.02, 0.05, 2 * nmol)
    w = 3670.2 - 3541.7 * E - 152677.0 * E * E
    x = 0.19285 - 1.7261e-5 * w
    p = 1.6466 + 5.7692e-4 * w
    mup = 0.1646 + 11.39 * E + 63.41 * E * E
    r = np.repeat(o, 2, axis=0) + 0.067 * u
    n = 2 * nmol
    H = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                H[i, j] = w[i] - 3400.0
            elif i // 2 == j // 2:
                H[i, j] = (-1361.0 + 27165.0 * (E[i] + E[j])) * x[i] * x[j] - 1.887 * p[i] * p[j]
            else:
                d = r[i] - r[j]
                d -= box * np.round(d / box)
                rr = np.linalg.norm(d)
                dh = d / rr
                rb = rr * 18.8973
                H[i, j] = ((u[i] @ u[j] - 3 * (u[i] @ dh) * (u[j] @ dh)) / rb ** 3
                           * x[i] * x[j] * mup[i] * mup[j] * 2.1947463e5)
    return H


def main():
    out = {}
    cases = eigh_cases()
    cases["exciton32"] = exciton_hamiltonian()
    for name, A in cases.items():
        w, V = np.linalg.eigh(A)
        out[name + "_A"] = A
        out[name + "_w"] = w
        out[name + "_V"] = V
    np.savez_compressed(os.path.join(HERE, "eigh_fixtures.npz"), **out)

    rng = np.random.default_rng(99)
    ff = {}
    for n, nx in ((2048, 150), (64, 10), (63, 10), (16, 9), (17, 9), (12, 12)):
        j = np.arange(nx)
        X = np.exp(-j / 40.0) * np.exp(1j * 0.3 * j) + 0.01 * (rng.standard_normal(nx) + 1j * rng.standard_normal(nx))
        Xf = np.zeros(n // 2 + 1, np.complex128)
        m = min(nx, n // 2 + 1)
        Xf[:m] = X[:m]
        ff["X_%d_%d" % (n, nx)] = X
        ff["y_%d_%d" % (n, nx)] = np.fft.irfft(Xf, n, norm="forward")
    np.savez_compressed(os.path.join(HERE, "irfft_fixtures.npz"), **ff)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
