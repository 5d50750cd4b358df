"""Oracle pinning of the molecule layouts the reference accepts and of the
generalisations beyond it.

Reference semantics: H1 or H2 by chrom % 2 whatever nchrom_mol is
(calcIR.cu:1061-1070, 1266-1275, 1360-1370), and charged sites = atoms
1..natom_mol-1 with +0.52 on atoms 1, 2 and -1.04 on atom 3
(calcIR.cu:1117-1160).  Generalisations: triclinic boxes (the reference uses
only the box diagonal, calcIR.cu:404-405), the oh_per_molecule rule (H1 for
nchrom_mol 1) and the msite_bisector rule (a built M site for three-site
water).

Pinned by: hand-placed collinear molecules against the Coulomb closed form,
a brute-force numpy field sum, reduction to the nchrom_mol 2 / natom_mol 4
case, brute-force nearest-image sums and couplings, lattice-translation
invariance and the TIP4P M geometry."""

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


def test_msite_bisector_builds_the_tip4p_m_site_for_three_site_water():
    # the synthetic frames carry M on the HOH bisector; with model tip4p2005 it
    # is moved to 0.015 nm from O (calcIR.cu:1124-1140), which is where the
    # msite_bisector rule builds a three-site water's negative site
    xyz, box, _ = frames(64, 2, 31)
    x4 = xyz[1]
    x3 = x4.reshape(-1, 4, 3)[:, :3, :].reshape(-1, 3).copy()
    e4 = O.eproj(O.F64, _cfg(imodel=1), x4, box[1])
    e3 = O.eproj(O.F64, _cfg(imodel=1, natom_mol=3, msite_bisector=1), x3, box[1])
    assert np.abs(e3 - e4).max() <= 1e-6 * np.abs(e4).max()
    H4 = O.hamiltonian(O.F64, _cfg(imodel=1), x4, box[1], e4)
    H3 = O.hamiltonian(O.F64, _cfg(imodel=1, natom_mol=3, msite_bisector=1), x3, box[1], e3)
    assert np.abs(H3 - H4).max() <= 1e-5 * np.abs(H4).max()
    # the rule changes nothing for four-site water
    assert np.array_equal(O.eproj(O.F64, _cfg(imodel=1, msite_bisector=1), x4, box[1]), e4)


def test_oh_per_molecule_with_one_oh_is_the_h1_rows():
    xyz, box, _ = frames(32, 2, 41)
    e2 = O.eproj(O.F64, _cfg(), xyz[1], box[1])
    e1 = O.eproj(O.F64, _cfg(nchrom_mol=1, oh_per_molecule=1), xyz[1], box[1])
    assert np.array_equal(e1, e2[0::2])
    H2 = O.hamiltonian(O.F64, _cfg(), xyz[1], box[1], e2)
    H1 = O.hamiltonian(O.F64, _cfg(nchrom_mol=1, oh_per_molecule=1), xyz[1], box[1], e1)
    assert np.array_equal(H1, H2[0::2, 0::2])
    # with two OH per molecule the rule is the reference's
    assert np.array_equal(O.eproj(O.F64, _cfg(oh_per_molecule=1), xyz[1], box[1]), e2)


def test_reference_one_oh_alternates_h1_h2_by_molecule_parity():
    # calcIR.cu:1061-1070: chrom % 2 picks the H, and with nchrom_mol 1 chrom is
    # the molecule index, so molecule n is the stretch of H(1 + n % 2): the rows
    # 2n + n % 2 of the two-OH case, field and Hamiltonian alike
    xyz, box, _ = frames(32, 2, 41)
    idx = np.array([2 * n + n % 2 for n in range(32)])
    for prec in (O.F64, O.F32):
        e2 = O.eproj(prec, _cfg(), xyz[1], box[1])
        e1 = O.eproj(prec, _cfg(nchrom_mol=1), xyz[1], box[1])
        assert np.array_equal(e1, e2[idx])
        H2 = O.hamiltonian(prec, _cfg(), xyz[1], box[1], e2)
        H1 = O.hamiltonian(prec, _cfg(nchrom_mol=1), xyz[1], box[1], e1)
        # every off-diagonal pair is intermolecular in both layouts
        assert np.array_equal(H1, H2[np.ix_(idx, idx)])
    s = O.sites(O.F64, _cfg(nchrom_mol=1), xyz[1], box[1], e1)
    x = xyz[1].astype(np.float64).reshape(32, 4, 3)
    for n in (0, 1, 2, 7):
        u = x[n, 1 + n % 2] - x[n, 0]
        u -= float(box[1, 0]) * np.round(u / float(box[1, 0]))
        assert np.allclose(s[n, O.SITE_U:O.SITE_U + 3], u / np.linalg.norm(u), atol=1e-6)


def test_reference_three_site_water_field_is_the_h_charges_brute_force():
    # calcIR.cu:1117-1160 with natom_mol 3: atoms 1 and 2 (+0.52 each) are the
    # only charged sites.  Brute-force numpy sum over the molecules whose O is
    # within the cutoff of the H, minimum image on each difference.
    xyz, box, _ = frames(48, 2, 5)
    nmol = 48
    x4 = xyz[1].astype(np.float64).reshape(nmol, 4, 3)
    x3 = np.ascontiguousarray(xyz[1].reshape(nmol, 4, 3)[:, :3, :].reshape(-1, 3))
    L = float(box[1, 0])
    for model in (0, 1):
        e = O.eproj(O.F64, _cfg(imodel=model, natom_mol=3), x3, box[1])
        for chrom in range(0, 2 * nmol, 5):
            n = chrom // 2
            h = x4[n, 1 + chrom % 2]
            u = h - x4[n, 0]
            u -= L * np.round(u / L)
            u /= np.linalg.norm(u)
            ef = np.zeros(3)
            for m in range(nmol):
                if m == n:
                    continue
                d = x4[m, 0] - h
                d -= L * np.round(d / L)
                if np.linalg.norm(d) > 0.7831:
                    continue
                for i in (1, 2):
                    dr = h - x4[m, i]
                    dr = (dr - L * np.round(dr / L)) * BOHR
                    ef += 0.52 * dr / np.linalg.norm(dr) ** 3
            assert e[chrom] == pytest.approx(ef @ u, rel=1e-9, abs=1e-12)
        # the same molecules with their M sites carry the extra -1.04 terms
        e4 = O.eproj(O.F64, _cfg(imodel=model), xyz[1], box[1])
        assert np.abs(e4 - e).max() > 1e-3


def test_reference_atoms_past_the_fourth_carry_no_charge():
    # calcIR.cu:1150-1158 adds a term only for atoms 1, 2 (H) and 3 (M); a
    # fifth atom per molecule (a TIP5P-like layout) changes nothing
    xyz, box, _ = frames(32, 2, 9)
    x4 = xyz[1].reshape(32, 4, 3)
    extra = x4[:, :1, :] + np.float32(0.05)
    x5 = np.ascontiguousarray(np.concatenate([x4, extra], axis=1).reshape(-1, 3))
    for model in (0, 1):
        e4 = O.eproj(O.F64, _cfg(imodel=model), xyz[1], box[1])
        e5 = O.eproj(O.F64, _cfg(imodel=model, natom_mol=5), x5, box[1])
        assert np.array_equal(e4, e5)


def test_unsupported_layouts_are_refused():
    xyz, box, _ = frames(32, 2, 9)
    x2 = np.ascontiguousarray(xyz[1].reshape(32, 4, 3)[:, :2, :].reshape(-1, 3))
    with pytest.raises(RuntimeError):
        O.eproj(O.F64, _cfg(natom_mol=2), x2, box[1])
    with pytest.raises(RuntimeError):
        O.eproj(O.F64, _cfg(nchrom_mol=3, oh_per_molecule=1), xyz[1], box[1])
