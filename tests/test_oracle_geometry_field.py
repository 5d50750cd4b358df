"""Oracle pinning: minimum image, frame selection and the field projection
(calcIR.cu:1020-1168, 1457-1460, 378-379) by closed forms on hand-placed
molecules."""

import numpy as np
import pytest

import calcir_amd as C
import oracle_lib as O

BOHR = 18.8973


@pytest.mark.parametrize("dx,L,expect", [(0.3, 1.0, 0.3), (0.7, 1.0, -0.3), (-0.7, 1.0, 0.3), (0.5, 1.0, -0.5),
                                         (-0.5, 1.0, 0.5), (1.5, 1.0, -0.5), (2.25, 1.0, 0.25),
                                         (-2.75, 2.0, -0.75), (0.0, 3.0, 0.0)])
def test_minimage_closed_form_half_away_from_zero(dx, L, expect):
    # dx - L*round(dx/L) with round half away from zero
    assert O.lib().orc_minimage(O.F64, dx, L) == pytest.approx(expect, abs=1e-15)
    assert O.lib().orc_minimage(O.F32, dx, L) == pytest.approx(expect, abs=1e-6)
    assert C.lib().cir_minimage(dx, L) == pytest.approx(expect, abs=1e-6)


def test_frame_index_reference_arithmetic():
    # desired = s*sampleEvery + beginTime + dt*k (float), frame = round(desired/frame_dt)
    p = C.default_params(sampleEvery=0.2, dt=0.01, beginTime=0.5)
    cfg = O.config(sampleEvery=0.2, dt=0.01, beginTime=0.5)
    import ctypes
    t = ctypes.c_float()
    f = O.lib().orc_frame_index(ctypes.byref(cfg), 3, 7, 0.01, ctypes.byref(t))
    assert f == 117 and abs(t.value - 1.17) < 1e-6
    fc, tc = C.frame_index(p, 3, 7, 0.01)
    assert fc == 117 and tc == t.value
    assert C.frame_index(p, 0, 0, 0.01)[0] == 50


def _water_x_probe(h_at=(0.0, 0.0, 0.0)):
    """Molecule 0: O at -0.09572 x from H1, H1 at h_at (chromophore 0 points
    along +x).  Returns a 4-atom block."""
    h = np.array(h_at)
    return np.array([h + [-0.09572, 0, 0], h, h + [-0.12, 0.09, 0.0], h + [-0.09, 0.01, 0.0]])


def _line_molecule(xo, xh1, xh2, xm, y=0.0):
    """A (non-physical) 'molecule' whose charged sites sit on the x axis."""
    return np.array([[xo, y, 0.0], [xh1, y, 0.0], [xh2, y, 0.0], [xm, y, 0.0]])


def _field0(x, L=5.0, model=0, cutoff=0.7831, prec=O.F64):
    cfg = O.config(imodel=model, cutoff=cutoff)
    return O.eproj(prec, cfg, x.astype(np.float32), np.array([L, L, L], np.float32))


def _closed(xs, qs, h=0.0, L=None):
    # field at a hydrogen at (h, 0, 0) from charges q at (x, 0, 0), projected on
    # +x, a.u.; coordinates are the float32 values the trajectory holds
    out = 0.0
    for x, q in zip(xs, qs):
        d = float(np.float32(h)) - float(np.float32(x))
        if L is not None:
            d -= L * np.round(d / L)
        out += q * d * BOHR / (abs(d) * BOHR) ** 3
    return out


def test_field_closed_form_collinear_charges():
    x = np.vstack([_water_x_probe(), _line_molecule(0.33, 0.30, 0.40, 0.35)])
    e = _field0(x)
    expect = _closed([0.30, 0.40, 0.35], [0.52, 0.52, -1.04])
    assert e[0] == pytest.approx(expect, rel=1e-12)
    # the same in exact decimals (float32 coordinates and the partial
    # cancellation of the three charges leave ~2e-6 relative)
    assert e[0] == pytest.approx(-(0.52 / 0.09 + 0.52 / 0.16 - 1.04 / 0.1225) / BOHR ** 2, rel=1e-5)
    # FP32 form
    assert _field0(x, prec=O.F32)[0] == pytest.approx(expect, rel=2e-6)


def test_field_cutoff_is_between_other_oxygen_and_this_hydrogen():
    # O of molecule 1 at 0.78 nm from H (included) vs 0.79 nm (excluded);
    # its charged sites stay at the same places.
    inc = np.vstack([_water_x_probe(), _line_molecule(0.78, 0.30, 0.40, 0.35)])
    exc = np.vstack([_water_x_probe(), _line_molecule(0.79, 0.30, 0.40, 0.35)])
    expect = _closed([0.30, 0.40, 0.35], [0.52, 0.52, -1.04])
    assert _field0(inc)[0] == pytest.approx(expect, rel=1e-12)
    assert _field0(exc)[0] == 0.0
    # a different cutoff moves the boundary
    assert _field0(inc, cutoff=0.75)[0] == 0.0


def test_field_excludes_own_molecule():
    x = _water_x_probe()
    e = _field0(x)
    assert np.all(e == 0.0)


def test_field_m_site_relocated_for_tip4p2005_e3b3():
    # model 1 moves M to O + 0.015 nm along O->M: here from 0.35 to 0.345
    x = np.vstack([_water_x_probe(), _line_molecule(0.33, 0.30, 0.40, 0.35)])
    e0 = _field0(x, model=0)[0]
    e1 = _field0(x, model=1)[0]
    assert e0 == pytest.approx(_closed([0.30, 0.40, 0.35], [0.52, 0.52, -1.04]), rel=1e-12)
    # M' = O + 0.015 (M - O)/|M - O| from the float32 O and M (not rounded again)
    mprime = float(np.float32(0.33)) + 0.015
    expect = _closed([0.30, 0.40], [0.52, 0.52]) + (-1.04) * (-mprime) * BOHR / (mprime * BOHR) ** 3
    assert e1 == pytest.approx(expect, rel=1e-12)
    assert e1 == pytest.approx(_closed([0.30, 0.40, 0.345], [0.52, 0.52, -1.04]), rel=1e-5)


def test_field_periodic_image_invariance():
    rng = np.random.default_rng(1)
    xyz, box, _ = C.synth_trajectory(32, 1, 5)
    x = xyz[0].astype(np.float64)
    L = float(box[0, 0])
    e = _field0(x, L=L)
    x2 = x.copy()
    for m in rng.choice(32, 8, replace=False):
        shift = L * rng.integers(-2, 3, 3)
        x2[4 * m:4 * m + 4] += shift  # a whole molecule by box vectors
    e2 = _field0(x2, L=L)
    # shifted float32 coordinates round differently at the 1e-7 nm level
    # (1e-7 a.u. shifts omega by < 1e-3 cm^-1)
    assert np.allclose(e, e2, rtol=1e-5, atol=1e-7)


def test_field_minimum_image_across_the_boundary():
    # molecule 1 sits across the box edge: its sites at x = L - 0.1 etc. are
    # seen at x = -0.1 from a hydrogen at the origin
    L = 2.0
    probe = _water_x_probe(h_at=(0.0, 0.0, 0.0)) + [1.0, 0, 0]
    probe -= [1.0, 0, 0]
    x = np.vstack([probe, _line_molecule(L - 0.13, L - 0.10, L - 0.20, L - 0.15)])
    e = _field0(x, L=L)
    assert e[0] == pytest.approx(_closed([L - 0.10, L - 0.20, L - 0.15], [0.52, 0.52, -1.04], L=L), rel=1e-12)
    assert e[0] == pytest.approx(_closed([-0.10, -0.20, -0.15], [0.52, 0.52, -1.04]), rel=1e-5)


def _field(x, natom_mol=4, nchrom_mol=2, model=0, L=5.0, **rules):
    cfg = O.config(imodel=model, natom_mol=natom_mol, nchrom_mol=nchrom_mol, **rules)
    return O.eproj(O.F64, cfg, np.ascontiguousarray(x, np.float32), np.array([L, L, L], np.float32))


def test_field_one_oh_per_molecule_alternates_h_by_parity_closed_form():
    # nchrom_mol 1: chromophore n is molecule n and calcIR.cu:1061-1070 takes
    # its H by n % 2.  Molecule 0 -> H1 at x = 0 (OH along +x); molecule 1 ->
    # H2 at x = 0.30 (OH along -x).  All sites on the x axis.
    m0 = _line_molecule(-0.10, 0.00, -0.25, -0.11)
    m1 = _line_molecule(0.33, 0.45, 0.30, 0.34)
    e = _field(np.vstack([m0, m1]), nchrom_mol=1)
    assert e[0] == pytest.approx(_closed([0.45, 0.30, 0.34], [0.52, 0.52, -1.04], h=0.0), rel=1e-12)
    assert e[1] == pytest.approx(-_closed([0.00, -0.25, -0.11], [0.52, 0.52, -1.04], h=0.30), rel=1e-12)
    # the oh_per_molecule rule takes molecule 1's H1 at 0.45 (OH along +x)
    v = _field(np.vstack([m0, m1]), nchrom_mol=1, oh_per_molecule=1)
    assert v[0] == e[0]
    assert v[1] == pytest.approx(_closed([0.00, -0.25, -0.11], [0.52, 0.52, -1.04], h=0.45), rel=1e-12)


def test_field_three_site_water_closed_form():
    # natom_mol 3: calcIR.cu:1117 loops i = 1, 2 only, so the two H charges are
    # the whole field; the msite_bisector rule adds -1.04 at O + 0.015 nm along
    # the HOH bisector (here both OH point along +x: at 0.33 + 0.015)
    m0 = _line_molecule(-0.10, 0.00, -0.25, 0.0)[:3]
    m1 = _line_molecule(0.33, 0.45, 0.40, 0.0)[:3]
    for model in (0, 1):
        e = _field(np.vstack([m0, m1]), natom_mol=3, model=model)
        assert e[0] == pytest.approx(_closed([0.45, 0.40], [0.52, 0.52], h=0.0), rel=1e-12)
        b = _field(np.vstack([m0, m1]), natom_mol=3, model=model, msite_bisector=1)
        mprime = float(np.float32(0.33)) + 0.015
        expect = _closed([0.45, 0.40], [0.52, 0.52], h=0.0) + (-1.04) * (-mprime) * BOHR / (mprime * BOHR) ** 3
        assert b[0] == pytest.approx(expect, rel=1e-12)
