"""Oracle pinning: Skinner maps (calcIR.cu:1203-1304) and the exciton
Hamiltonian (calcIR.cu:1306-1404) by closed forms.

Map coefficients as published (Gruenbaum et al. 2013 for map 0; Li and
Skinner 2010 for map 1), OH maps for H2O and HOD/D2O, OD maps for HOD/H2O
and D2O; the transition-dipole derivative and intramolecular coupling maps
do not depend on the species."""

import numpy as np
import pytest

import oracle_lib as O

BOHR = 18.8973
HARTREE_CM = 2.1947463e5

#            omega: a0, a1, a2              x10: b0, b1             p10: c0, c1
MAPS = {
    (0, "OH"): ((3670.2, -3541.7, -152677.0), (0.19285, -1.7261e-5), (1.6466, 5.7692e-4)),
    (0, "OD"): ((2767.8, -2630.3, -102601.0), (0.16593, -2.0632e-5), (2.0475, 8.9108e-4)),
    (1, "OH"): ((3732.9, -3519.8, -153520.0), (0.19318, -1.7248e-5), (1.6102, 5.8697e-4)),
    (1, "OD"): ((2748.2, -2572.2, -102980.0), (0.16598, -2.0752e-5), (1.9813, 9.1419e-4)),
}
MUP = {0: (0.1646, 11.39, 63.41), 1: (0.1622, 10.381, 137.6)}
WINTRA = {0: (-1361.0, 27165.0, -1.887), 1: (-1360.8, 27171.0, -1.887)}
SPECIES_KIND = {0: "OH", 1: "OH", 2: "OD", 3: "OD"}


def _two_waters(sep=1.0):
    """Molecule A at the origin with OH bonds along +x (H1) and +y (H2);
    molecule B at (sep, 0, 0) with OH bonds along +x (H1) and +z (H2).  The
    O...H distances exceed the field cutoff, so every E is exactly 0."""
    r = 0.09572
    a = [[0, 0, 0], [r, 0, 0], [0, r, 0], [0.01, 0.01, 0.0]]
    b = [[sep, 0, 0], [sep + r, 0, 0], [sep, 0, r], [sep + 0.01, 0, 0.01]]
    return np.array(a + b, np.float32)


@pytest.mark.parametrize("imap", [0, 1])
@pytest.mark.parametrize("ispecies", [0, 1, 2, 3])
def test_site_maps_closed_form(imap, ispecies):
    x = _two_waters()
    box = np.array([4.0, 4.0, 4.0], np.float32)
    cfg = O.config(imodel=0, ispecies=ispecies, imap=imap)
    E = np.array([0.0, 0.01, -0.02, 0.035])
    s = O.sites(O.F64, cfg, x, box, E)
    (a0, a1, a2), (b0, b1), (c0, c1) = MAPS[(imap, SPECIES_KIND[ispecies])]
    m0, m1, m2 = MUP[imap]
    w = a0 + a1 * E + a2 * E * E
    xx = b0 + b1 * w
    p = c0 + c1 * w
    mup = m0 + m1 * E + m2 * E * E
    assert np.allclose(s[:, O.SITE_W], w, rtol=1e-14)
    assert np.allclose(s[:, O.SITE_X], xx, rtol=1e-14)
    assert np.allclose(s[:, O.SITE_P], p, rtol=1e-14)
    assert np.allclose(s[:, O.SITE_MUP], mup, rtol=1e-14)
    # E = 0 gives the constant terms
    assert s[0, O.SITE_W] == pytest.approx(a0)
    # unit OH vectors: +x, +y (molecule A), +x, +z (molecule B)
    u = s[:, O.SITE_U:O.SITE_U + 3]
    assert np.allclose(u, [[1, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-7)
    # mu = u mu' x ; alpha = (4.6 u u^T + I) x
    assert np.allclose(s[:, O.SITE_MU:O.SITE_MU + 3], u * (mup * xx)[:, None], rtol=1e-12, atol=1e-15)
    a = s[:, O.SITE_A:O.SITE_A + 6]
    uu = np.stack([u[:, 0] ** 2, u[:, 1] ** 2, u[:, 2] ** 2, u[:, 0] * u[:, 1], u[:, 1] * u[:, 2], u[:, 2] * u[:, 0]], 1)
    iso = np.array([1, 1, 1, 0, 0, 0])
    assert np.allclose(a, (4.6 * uu + iso) * xx[:, None], rtol=1e-12, atol=1e-12)
    # TDM 0.067 nm from O along the bond (min-imaged)
    assert np.allclose(s[1, O.SITE_R:O.SITE_R + 3], [0, 0.067, 0], atol=1e-7)
    # FP32 form agrees to float precision
    s32 = O.sites(O.F32, cfg, x, box, E)
    assert np.allclose(s32[:, O.SITE_W], w, rtol=1e-6)


@pytest.mark.parametrize("imap", [0, 1])
def test_hamiltonian_closed_form_two_molecules(imap):
    x = _two_waters(sep=1.0)
    box = np.array([4.0, 4.0, 4.0], np.float32)
    cfg = O.config(imodel=0, ispecies=0, imap=imap, avef=3400.0)
    e = O.eproj(O.F64, cfg, x, box)
    assert np.all(e == 0.0)  # everything beyond the cutoff
    H = O.hamiltonian(O.F64, cfg, x, box, e)
    (a0, _, _), (b0, b1), (c0, c1) = MAPS[(imap, "OH")]
    xx = b0 + b1 * a0
    p = c0 + c1 * a0
    mup = MUP[imap][0]
    wi0, wi1, wi2 = WINTRA[imap]
    # diagonal: omega - avef
    assert np.allclose(np.diag(H), a0 - 3400.0, rtol=1e-14)
    # intramolecular: (wi0 + wi1 (E_i+E_j)) x_i x_j + wi2 p_i p_j with E = 0
    assert H[0, 1] == pytest.approx(wi0 * xx * xx + wi2 * p * p, rel=1e-13)
    assert H[2, 3] == pytest.approx(wi0 * xx * xx + wi2 * p * p, rel=1e-13)
    # TDC A0-B0: both along +x, TDMs at (0.067,0,0) and (1.067,0,0):
    # (u.u - 3 (u.r)(u.r)) / r^3 = (1 - 3) / r^3
    r = 1.0 * BOHR
    pref = (xx * mup) ** 2 * HARTREE_CM
    assert H[0, 2] == pytest.approx(-2.0 / r ** 3 * pref, rel=1e-6)
    # TDC A0-B1: u_A = x, u_B = z, d = (0.067,0,0) - (1,0,0.067) = (-0.933, 0, -0.067)
    d = np.array([0.067 - 1.0, 0.0, -0.067])
    rn = np.linalg.norm(d)
    dh = d / rn
    val = (0.0 - 3 * dh[0] * dh[2]) / (rn * BOHR) ** 3 * pref
    assert H[0, 3] == pytest.approx(val, rel=1e-6)
    # TDC A1-B0: u_A = y is perpendicular to d = (-1, 0.067, 0)/|..| ... use the formula directly
    d = np.array([0.0, 0.067, 0.0]) - np.array([1.067, 0.0, 0.0])
    rn = np.linalg.norm(d)
    dh = d / rn
    val = (0.0 - 3 * dh[1] * dh[0]) / (rn * BOHR) ** 3 * pref
    assert H[1, 2] == pytest.approx(val, rel=1e-6)
    assert np.allclose(H, H.T, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ispecies", [1, 2])
def test_hamiltonian_hod_species_is_diagonal(ispecies):
    from calcir_amd import synth_trajectory

    xyz, box, _ = synth_trajectory(32, 1, 3)
    cfg = O.config(imodel=1, ispecies=ispecies, imap=0, avef=2500.0)
    e = O.eproj(O.F64, cfg, xyz[0], box[0])
    H = O.hamiltonian(O.F64, cfg, xyz[0], box[0], e)
    assert np.count_nonzero(H - np.diag(np.diag(H))) == 0
    s = O.sites(O.F64, cfg, xyz[0], box[0], e)
    assert np.allclose(np.diag(H), s[:, O.SITE_W] - 2500.0)


def test_hamiltonian_symmetric_and_species_maps_differ():
    from calcir_amd import synth_trajectory

    xyz, box, _ = synth_trajectory(32, 1, 4)
    H0 = O.hamiltonian(O.F64, O.config(ispecies=0), xyz[0], box[0], O.eproj(O.F64, O.config(), xyz[0], box[0]))
    H3 = O.hamiltonian(O.F64, O.config(ispecies=3), xyz[0], box[0], O.eproj(O.F64, O.config(), xyz[0], box[0]))
    assert np.allclose(H0, H0.T, rtol=1e-12, atol=1e-10)
    assert not np.allclose(np.diag(H0), np.diag(H3))


def test_coefficient_override_closed_form_and_identity():
    # overridden coefficients enter the same closed-form polynomials, and an
    # override equal to the table row reproduces the table run exactly
    x = _two_waters(sep=0.5)  # O...H within the cutoff: the coupling is nonzero
    box = np.array([4.0, 4.0, 4.0], np.float32)
    rng = np.random.default_rng(5)
    coef = np.array([3600.0, -3000.0, -1.4e5, 0.2, -1.5e-5, 1.7, 6e-4, 0.17, 11.0, 70.0, -1300.0, 26000.0, -2.0])
    coef *= 1 + 0.05 * rng.standard_normal(13)
    cfg = O.config(imodel=0, ispecies=0, imap=0, coef=coef)
    E = np.array([0.0, 0.01, -0.02, 0.035])
    s = O.sites(O.F64, cfg, x, box, E)
    w = coef[0] + coef[1] * E + coef[2] * E * E
    assert np.allclose(s[:, O.SITE_W], w, rtol=1e-14)
    assert np.allclose(s[:, O.SITE_X], coef[3] + coef[4] * w, rtol=1e-14)
    assert np.allclose(s[:, O.SITE_P], coef[5] + coef[6] * w, rtol=1e-14)
    assert np.allclose(s[:, O.SITE_MUP], coef[7] + coef[8] * E + coef[9] * E * E, rtol=1e-14)
    # intramolecular coupling of molecule A's two chromophores
    e = O.eproj(O.F64, cfg, x, box)
    s = O.sites(O.F64, cfg, x, box, e)
    H = O.hamiltonian(O.F64, cfg, x, box, e)
    xx, pp, ee = s[:, O.SITE_X], s[:, O.SITE_P], e
    h01 = (coef[10] + coef[11] * (ee[0] + ee[1])) * xx[0] * xx[1] + coef[12] * pp[0] * pp[1]
    assert H[0, 1] == pytest.approx(h01, rel=1e-13)
    # override == table row: identical pipeline output
    from cases import frames
    xyz, bx, times = frames(16, 24, 3)
    base = dict(imodel=1, ispecies=0, imap=0, ntcfpoints=6, nsamples=2, sampleEvery=0.05, nzeros=58,
                omegaStart=2800.0, omegaStop=4000.0, avef=3400.0)
    table = np.array([3670.2, -3541.7, -152677.0, 0.19285, -1.7261e-5, 1.6466, 5.7692e-4, 0.1646, 11.39, 63.41,
                      -1361.0, 27165.0, -1.887])
    a = O.run(O.F64, O.config(**base), xyz, bx, times)
    b = O.run(O.F64, O.config(coef=table, **base), xyz, bx, times)
    assert np.array_equal(a["ftcf_ir"], b["ftcf_ir"]) and np.array_equal(a["tcf_vv"], b["tcf_vv"])
    c = O.run(O.F64, O.config(coef=table * 1.01, **base), xyz, bx, times)
    assert not np.allclose(a["ftcf_ir"], c["ftcf_ir"])
