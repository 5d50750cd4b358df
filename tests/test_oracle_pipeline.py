"""Oracle pinning of the whole reference control flow (calcIR.cu:360-919) on
small synthetic systems: literal F-matrix vs vector propagation, the
uncoupled (HOD) closed form, C_IR(0) = sum |mu|^2, the Raman TCFs at t = 0 as
orientational averages, the spectral density of uncoupled chromophores as a
Lorentzian sum, histogram normalisation, IPR closed forms and FP32 vs
FP64."""

import numpy as np
import pytest

import calcir_amd as C
import oracle_lib as O
from cases import frames, rel

HBAR = 5.308837367


def _cfg(**kw):
    base = dict(imodel=1, ispecies=0, imap=0, avef=3400.0, ntcfpoints=20, nsamples=3, sampleEvery=0.05,
                nzeros=492, omegaStart=1500.0, omegaStop=4500.0, omegaStep=5, nthreads=4)
    base.update(kw)
    return O.config(**base)


def test_literal_and_vector_propagation_agree():
    xyz, box, times = frames(16, 40, 21)
    a = O.run(O.F64, _cfg(propagation=O.PROP_LITERAL), xyz, box, times)
    b = O.run(O.F64, _cfg(propagation=O.PROP_VECTOR), xyz, box, times)
    for k in ("tcf_ir", "tcf_vv", "tcf_vh", "ftcf_ir", "ftcf_vv", "ftcf_vh"):
        assert rel(a[k], b[k]) < 1e-12, k
    for k in ("sw", "pw"):
        assert np.allclose(a[k], b[k], rtol=1e-13, atol=0)


def test_uncoupled_closed_form_and_ipr():
    # HOD: diagonal H, F(t_k) = diag(exp(+i sum_{j=1..k} (w(t_j) - avef) dt/hbar))
    xyz, box, times = frames(16, 40, 21)
    cfg = _cfg(ispecies=1, nsamples=1, ntcfpoints=15)
    r = O.run(O.F64, cfg, xyz, box, times)
    phase = 0.0
    ref = []
    for k in range(15):
        e = O.eproj(O.F64, cfg, xyz[k], box[k])
        s = O.sites(O.F64, cfg, xyz[k], box[k], e)
        mu = s[:, O.SITE_MU:O.SITE_MU + 3]
        if k == 0:
            mu0 = mu
        else:
            phase = phase + (s[:, O.SITE_W] - cfg.avef) * np.float32(cfg.dt) / HBAR
        ref.append(np.sum(np.sum(mu * mu0, axis=1) * np.exp(1j * phase)))
    ref = np.array(ref) * np.exp(-np.arange(15) * np.float32(cfg.dt) / (2 * np.float32(cfg.t1)))
    assert rel(r["tcf_ir"], ref) < 1e-12
    # localised modes: IPR and molecular IPR are 1 wherever a bin is populated
    m = r["pw"] > 0
    assert np.allclose(r["rw"][m], 1.0, rtol=1e-12)
    assert np.allclose(r["rmw"][m], 1.0, rtol=1e-12)


def test_tcf_at_zero_is_sum_of_squared_dipoles():
    xyz, box, times = frames(16, 40, 21)
    cfg = _cfg()
    r = O.run(O.F64, cfg, xyz, box, times)
    tot = 0.0
    for smp in range(cfg.nsamples):
        f = int(round(smp * 0.05 / 0.01))
        s = O.sites(O.F64, cfg, xyz[f], box[f], O.eproj(O.F64, cfg, xyz[f], box[f]))
        tot += np.sum(s[:, O.SITE_MU:O.SITE_MU + 3] ** 2)
    assert r["tcf_ir"][0].real == pytest.approx(tot / cfg.nsamples, rel=1e-12)
    assert r["tcf_ir"][0].imag == 0.0


def test_histogram_and_spectral_density_normalisation():
    xyz, box, times = frames(16, 40, 21)
    cfg = _cfg(omegaStart=-20000.0, omegaStop=30000.0)
    r = O.run(O.F64, cfg, xyz, box, times)
    N = 32
    binned = 1.0 - r["hist_skipped"] / (N * cfg.nsamples * cfg.ntcfpoints)
    assert np.sum(r["pw"]) * cfg.omegaStep == pytest.approx(binned, rel=1e-12)
    # Lorentzians of unit area weighted by |V^T mu|^2 (sum = |mu|^2): the
    # window leaves tails of 2 gamma / (pi D) ~ 3e-4 outside, so sum(Sw)*step
    # ~ mean over samples of sum |mu|^2 = C_IR(0)
    assert np.sum(r["sw"]) * cfg.omegaStep == pytest.approx(r["tcf_ir"][0].real, rel=1e-3)
    # IPR lies between 1 and N
    m = r["pw"] > 0
    assert np.all((r["rw"][m] >= 1 - 1e-12) & (r["rw"][m] <= N + 1e-9))


def test_fp32_form_within_1e4_of_fp64():
    xyz, box, times = frames(16, 40, 21)
    a = O.run(O.F64, _cfg(), xyz, box, times)
    b = O.run(O.F32, _cfg(), xyz, box, times)
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        assert rel(b[k], a[k]) < 1e-4, k


def test_lineshape_of_uncoupled_static_chromophores():
    # identical frames, HOD: each chromophore oscillates at (w_i - avef); the
    # C2R of the decayed TCF peaks at the frequencies avef + (w - avef)
    xyz, box, _ = frames(16, 1, 21)
    x = np.repeat(xyz[:1], 200, axis=0)
    bx = np.repeat(box[:1], 200, axis=0)
    times = (np.arange(200) * 0.010).astype(np.float32)
    cfg = _cfg(ispecies=1, nsamples=1, ntcfpoints=200, nzeros=3896, t1=0.26)
    r = O.run(O.F64, cfg, x, bx, times)
    ft = r["ftcf_ir"]
    n = cfg.ntcfpoints + cfg.nzeros
    factor = 2 * np.pi * HBAR / (np.float32(cfg.dt) * n)
    k = int(np.argmax(ft))
    # C2R index -> frequency as the lineshape writer maps it (calcIR.cu:899-910)
    peak = cfg.avef - k * factor if k < n // 2 else cfg.avef + (n - k) * factor
    s = O.sites(O.F64, cfg, x[0], bx[0], O.eproj(O.F64, cfg, x[0], bx[0]))
    w = s[:, O.SITE_W]
    wgt = np.sum(s[:, O.SITE_MU:O.SITE_MU + 3] ** 2, axis=1)
    # the spectrum's global maximum lies within a linewidth (hbar/2T1) of a
    # chromophore frequency
    assert np.min(np.abs(w - peak)) < HBAR / (2 * 0.26)
    # and the total intensity matches C(0): sum_k y_k = n * Re c_0
    assert np.sum(ft) == pytest.approx(n * r["tcf_ir"][0].real, rel=1e-9)
    assert r["tcf_ir"][0].real == pytest.approx(np.sum(wgt), rel=1e-12)


def test_raman_tcfs_at_zero_are_the_orientational_averages():
    # F(0) = I, so VV(0) and VH(0) are sums over chromophores of the isotropic
    # averages <a_zz^2> = (2 tr(a^2) + (tr a)^2) / 15 and
    # <a_zx^2> = (3 tr(a^2) - (tr a)^2) / 30 (calcIR.cu:736-745 combine the
    # component products into exactly these); both are rotation invariants, so
    # a swapped coefficient or a misplaced tensor component fails here
    xyz, box, times = frames(16, 40, 21)
    cfg = _cfg(nsamples=1, ntcfpoints=4)
    r = O.run(O.F64, cfg, xyz, box, times)
    s = O.sites(O.F64, cfg, xyz[0], box[0], O.eproj(O.F64, cfg, xyz[0], box[0]))
    a = s[:, O.SITE_A:O.SITE_A + 6]  # xx, yy, zz, xy, yz, zx
    tr = a[:, 0] + a[:, 1] + a[:, 2]
    tr2 = np.sum(a[:, :3] ** 2, axis=1) + 2 * np.sum(a[:, 3:] ** 2, axis=1)
    assert r["tcf_vv"][0].real == pytest.approx(np.sum(2 * tr2 + tr ** 2) / 15, rel=1e-12)
    assert r["tcf_vh"][0].real == pytest.approx(np.sum(3 * tr2 - tr ** 2) / 30, rel=1e-12)
    assert r["tcf_vv"][0].imag == 0.0 and r["tcf_vh"][0].imag == 0.0
    # and VH(0) is positive: the map's 4.6 u u^T term makes every alpha anisotropic
    assert r["tcf_vh"][0].real > 0


def test_spectral_density_of_uncoupled_chromophores_is_a_lorentzian_sum():
    # HOD: H is diagonal, each eigenvector a unit vector, so
    # Sw(om) = (1/nsamples) sum_samples sum_i |mu_i|^2 (gamma/pi) / ((om - w_i)^2 + gamma^2)
    # over frame 0 of each sample, gamma = hbar / (2 T1) (calcIR.cu:458-482, 1414-1444, 832)
    xyz, box, times = frames(16, 40, 21)
    cfg = _cfg(ispecies=1, nsamples=2)
    r = O.run(O.F64, cfg, xyz, box, times)
    gamma = HBAR / (float(np.float32(cfg.t1)) * 2.0)
    om = np.array([np.float32(cfg.omegaStart + np.float32(cfg.omegaStep * i)) for i in range(len(r["sw"]))],
                  np.float64)
    expect = np.zeros_like(om)
    for smp in range(cfg.nsamples):
        f = int(round(smp * 0.05 / 0.01))
        s = O.sites(O.F64, cfg, xyz[f], box[f], O.eproj(O.F64, cfg, xyz[f], box[f]))
        m2 = np.sum(s[:, O.SITE_MU:O.SITE_MU + 3] ** 2, axis=1)
        wi = s[:, O.SITE_W]
        expect += np.sum(m2[None, :] * gamma / np.pi / ((om[:, None] - wi[None, :]) ** 2 + gamma ** 2), axis=1)
    expect /= cfg.nsamples
    np.testing.assert_allclose(r["sw"], expect, rtol=1e-9, atol=1e-12 * expect.max())
