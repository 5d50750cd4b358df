"""Oracle pinning of the generalisations beyond the reference: triclinic
boxes (the reference uses only the box diagonal, calcIR.cu:404-405), one OH
chromophore per molecule (the reference hard-codes chrom%2,
calcIR.cu:1061-1070) and three-site water (the reference reads atom 3 as the
M site, calcIR.cu:1124-1158).

Pinned by: reduction to the reference case (zero shear, nchrom_mol 2 rows),
brute-force nearest-image Coulomb sums and couplings, invariance under
lattice translations of single molecules, and the TIP4P M geometry."""

import itertools

import numpy as np
import pytest

import oracle_lib as O
from cases import frames, triclinic_frames

BOHR = 18.8973
HARTREE_CM = 2.1947463e5


_IMG = np.array(list(itertools.product(range(-2, 3), repeat=3)), np.float64)


def _nearest(d, M):
    """Nearest lattice images of separations d [..., 3] (brute force over
    5^3 images)."""
    d = np.asarray(d, np.float64)
    cand = d[..., None, :] + _IMG @ M  # [..., 125, 3]
    k = np.argmin((cand * cand).sum(-1), axis=-1)
    return np.take_along_axis(cand, k[..., None, None], axis=-2)[..., 0, :]


def _cfg(**kw):
    base = dict(imodel=1, ispecies=0, imap=0, cutoff=0.7831)
    base.update(kw)
    return O.config(**base)


def test_zero_shear_triclinic_is_the_rectangular_oracle():
    xyz, box, times = frames(32, 3, 5)
    b9 = np.zeros((3, 3), np.float32)
    np.fill_diagonal(b9, box[1])
    for prec in (O.F32, O.F64):
        e3 = O.eproj(prec, _cfg(), xyz[1], box[1])
        e9 = O.eproj(prec, _cfg(triclinic=1), xyz[1], b9)
        assert np.array_equal(e3, e9)
        H3 = O.hamiltonian(prec, _cfg(), xyz[1], box[1], e3)
        H9 = O.hamiltonian(prec, _cfg(triclinic=1), xyz[1], b9, e9)
        assert np.array_equal(H3, H9)


def test_triclinic_field_and_coupling_against_brute_force_images():
    # 512 waters: L = 2.48 nm, so the 0.7831 nm cutoff stays well inside the
    # region where the sequential triclinic minimum image is the nearest image
    xyz, b9, _ = triclinic_frames(512, 2, 11)
    x, M = xyz[1].astype(np.float64), b9[1].astype(np.float64)
    cfg = _cfg(triclinic=1, imodel=0)  # M read from the trajectory
    e = O.eproj(O.F64, cfg, xyz[1], b9[1])
    s = O.sites(O.F64, cfg, xyz[1], b9[1], e)
    H = O.hamiltonian(O.F64, cfg, xyz[1], b9[1], e)
    nmol = x.shape[0] // 4
    xm = x.reshape(nmol, 4, 3)
    rng = np.random.default_rng(1)
    for n in rng.choice(2 * nmol, 48, replace=False):
        mol, h = n // 2, 1 + n % 2
        hpos = xm[mol, h]
        u = _nearest(hpos - xm[mol, 0], M)
        u /= np.linalg.norm(u)
        keep = np.linalg.norm(_nearest(xm[:, 0] - hpos, M), axis=-1) <= 0.7831
        keep[mol] = False
        ef = np.zeros(3)
        for i, q in ((1, 0.52), (2, 0.52), (3, -1.04)):
            d = _nearest(hpos - xm[keep, i], M) * BOHR
            ef += (q * d / (np.linalg.norm(d, axis=-1) ** 3)[:, None]).sum(0)
        assert e[n] == pytest.approx(ef @ u, rel=1e-9, abs=1e-12)
    # transition-dipole coupling from the nearest image of the TDM positions,
    # for pairs within 0.4 of the box height
    for _ in range(60):
        i, j = rng.integers(0, 2 * nmol, 2)
        if i // 2 == j // 2:
            continue
        d = _nearest(s[i, O.SITE_R:O.SITE_R + 3] - s[j, O.SITE_R:O.SITE_R + 3], M)
        r = np.linalg.norm(d)
        if r > 0.4 * M[2, 2]:
            continue
        dn = d / r
        ui, uj = s[i, O.SITE_U:O.SITE_U + 3], s[j, O.SITE_U:O.SITE_U + 3]
        k = ((ui @ uj) - 3 * (ui @ dn) * (uj @ dn)) / (r * BOHR) ** 3
        k *= s[i, O.SITE_X] * s[j, O.SITE_X] * s[i, O.SITE_MUP] * s[j, O.SITE_MUP] * HARTREE_CM
        assert H[i, j] == pytest.approx(k, rel=1e-10, abs=1e-12)


@pytest.mark.parametrize("vec", [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, -1, 1)])
def test_lattice_translation_of_one_molecule_changes_nothing(vec):
    xyz, b9, _ = triclinic_frames(48, 2, 23)
    x, M = xyz[1].copy(), b9[1]
    cfg = _cfg(triclinic=1)
    e0 = O.eproj(O.F64, cfg, x, M)
    H0 = O.hamiltonian(O.F64, cfg, x, M, e0)
    shift = (np.array(vec, np.float64) @ M.astype(np.float64)).astype(np.float32)
    y = x.copy()
    y[4 * 7:4 * 8] += shift
    e1 = O.eproj(O.F64, cfg, y, M)
    H1 = O.hamiltonian(O.F64, cfg, y, M, e1)
    # float32 coordinates move by one rounding when shifted
    assert np.abs(e1 - e0).max() <= 1e-5 * np.abs(e0).max()
    assert np.abs(H1 - H0).max() <= 1e-4 * np.abs(H0).max()


def test_three_site_water_builds_the_tip4p_m_site():
    # the synthetic frames carry M on the HOH bisector; with model tip4p2005 it
    # is moved to 0.015 nm from O (calcIR.cu:1124-1140), which is where a
    # three-site water's built M sits
    xyz, box, _ = frames(64, 2, 31)
    x4 = xyz[1]
    x3 = x4.reshape(-1, 4, 3)[:, :3, :].reshape(-1, 3).copy()
    e4 = O.eproj(O.F64, _cfg(imodel=1), x4, box[1])
    e3 = O.eproj(O.F64, _cfg(imodel=1, natom_mol=3), x3, box[1])
    assert np.abs(e3 - e4).max() <= 1e-6 * np.abs(e4).max()
    H4 = O.hamiltonian(O.F64, _cfg(imodel=1), x4, box[1], e4)
    H3 = O.hamiltonian(O.F64, _cfg(imodel=1, natom_mol=3), x3, box[1], e3)
    assert np.abs(H3 - H4).max() <= 1e-5 * np.abs(H4).max()


def test_one_oh_per_molecule_is_the_h1_rows():
    xyz, box, _ = frames(32, 2, 41)
    e2 = O.eproj(O.F64, _cfg(), xyz[1], box[1])
    e1 = O.eproj(O.F64, _cfg(nchrom_mol=1), xyz[1], box[1])
    assert np.array_equal(e1, e2[0::2])
    H2 = O.hamiltonian(O.F64, _cfg(), xyz[1], box[1], e2)
    H1 = O.hamiltonian(O.F64, _cfg(nchrom_mol=1), xyz[1], box[1], e1)
    assert np.array_equal(H1, H2[0::2, 0::2])
