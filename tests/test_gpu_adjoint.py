"""Gradients of the IR lineshape with respect to the 13 map coefficients
(cir_forward_ls / cir_backward_ls) against central finite differences of
the CPU oracle's FP64 pipeline at overridden coefficients, against the GPU
FP64 forward at a larger size, and FP32 against FP64.

The loss is the L2 distance between the IR lineshape rows and a
target lineshape made at seeded +-2% perturbed coefficients.  Gradients are
compared as relative sensitivities coef_c * dL/dcoef_c, which share one
scale across coefficients of very different sizes."""

import numpy as np
import pytest

import calcir_amd as C
import oracle_lib as O
from calcir_amd.adjoint import LineshapeSession, lineshape_map
from cases import frames
from conftest import announce

pytestmark = pytest.mark.gpu

KW = dict(model="tip4p2005", ntcfpoints=12, nsamples=3, sampleEvery=0.1, avef=3400.0, omegaStart=2800.0,
          omegaStop=4000.0, omegaStep=5, nzeros=500, t1=0.26, dt=0.010)


def _loss_grad(p, ft, target):
    idx, _, scale = lineshape_map(p)
    ls = ft[idx] * scale
    r = ls - target
    g = np.zeros_like(ft)
    np.add.at(g, idx, 2.0 * r * scale)
    return float(r @ r), g


def _target(p, xyz, box, times, coef, seed):
    pert = coef * (1 + 0.02 * np.random.default_rng(seed).choice([-1.0, 1.0], coef.size))
    idx, _, scale = lineshape_map(p)
    o = O.run(O.F64, O.config_from_params(p, coef=pert), xyz, box, times)
    return o["ftcf_ir"][idx] * scale


def _oracle_fd(p, xyz, box, times, coef, target, h=1e-4):
    idx, _, scale = lineshape_map(p)
    g = np.zeros(coef.size)
    for c in range(coef.size):
        d = h * abs(coef[c])
        vals = []
        for sgn in (1.0, -1.0):
            cf = coef.copy()
            cf[c] += sgn * d
            ft = O.run(O.F64, O.config_from_params(p, coef=cf), xyz, box, times)["ftcf_ir"]
            r = ft[idx] * scale - target
            vals.append(float(r @ r))
        g[c] = (vals[0] - vals[1]) / (2 * d)
    return g


def _check_sens(coef, ga, gf, tol):
    sa, sf = coef * ga, coef * gf
    big = np.abs(sf).max()
    assert big > 0
    err = np.abs(sa - sf)
    assert err.max() <= tol * big, (err / big, sa, sf)
    for c in range(coef.size):
        if abs(sf[c]) > 0.05 * big:
            assert err[c] <= tol * abs(sf[c]) * 10, (C.COEF_NAMES[c], sa[c], sf[c])


@pytest.mark.parametrize("species,imap", [("H2O", 0), ("HOD/D2O", 1), ("D2O", 1)])
def test_adjoint_matches_oracle_finite_differences(species, imap):
    xyz, box, times = frames(64, 40, 2024)
    p = C.default_params(precision="fp64", species=species, map=imap, **KW)
    coef = C.map_coefs(imap, C.resolve_species(species))
    target = _target(p, xyz, box, times, coef, 7)
    announce("adjoint fp64 %s map %d: forward/backward" % (species, imap))
    sess = LineshapeSession(p, xyz, box, times)
    ft = sess.forward(coef)
    # the forward is the oracle's IR lineshape
    o = O.run(O.F64, O.config_from_params(p), xyz, box, times)
    assert np.abs(ft - o["ftcf_ir"]).max() <= 1e-10 * np.abs(o["ftcf_ir"]).max()
    L, dL = _loss_grad(p, ft, target)
    ga = sess.backward(dL)
    sess.close()
    gf = _oracle_fd(p, xyz, box, times, coef, target)
    _check_sens(coef, ga, gf, 1e-4)
    if species.startswith("HOD"):
        # uncoupled: the intramolecular coefficients cannot matter
        assert np.all(ga[C.COEF_NAMES.index("wi0"):] == 0.0)


def test_forward_ls_equals_run_and_gpu_fd_at_n512():
    # N = 512: finite differences of the GPU FP64 forward itself on the map
    # coefficients of omega, mu' and x10
    xyz, box, times = frames(256, 30, 99)
    kw = dict(KW, nsamples=2, ntcfpoints=10)
    p = C.default_params(precision="fp64", species="H2O", **kw)
    coef = C.map_coefs(0, 0)
    announce("adjoint fp64 N=512 forward vs cir_run")
    sess = LineshapeSession(p, xyz, box, times)
    ft = sess.forward(coef)
    ref = C.run(p, xyz, box, times)["ftcf_ir"]
    assert np.abs(ft - ref).max() <= 1e-11 * np.abs(ref).max()
    idx, _, scale = lineshape_map(p)
    target = ft[idx] * scale * 1.05
    L, dL = _loss_grad(p, ft, target)
    ga = sess.backward(dL)
    names = ("w0", "w1", "w2", "mu0", "mu1", "x0", "x1")
    for name in names:
        c = C.COEF_NAMES.index(name)
        d = 1e-4 * abs(coef[c])
        vals = []
        for sgn in (1.0, -1.0):
            cf = coef.copy()
            cf[c] += sgn * d
            announce("adjoint fp64 N=512 fd %s %+d" % (name, sgn))
            r = sess.forward(cf)[idx] * scale - target
            vals.append(float(r @ r))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(ga[c] - fd) <= 1e-4 * abs(fd), (name, ga[c], fd)
    sess.close()


def test_adjoint_fp32_against_fp64_and_autograd():
    torch = pytest.importorskip("torch")
    from calcir_amd.adjoint import IRLineshape

    # seed 1001: a trajectory on which the reference algorithm's FP32 and FP64
    # spectra agree to ~2e-6 (some seeds hold close contacts whose FP32
    # rounding alone moves the spectrum by ~4e-4, which an L2 loss amplifies)
    xyz, box, times = frames(128, 40, 1001)
    coef = C.map_coefs(0, 0)
    grads = {}
    for prec in ("fp64", "fp32"):
        announce("adjoint %s N=256" % prec)
        p = C.default_params(precision=prec, species="H2O", **KW)
        sess = LineshapeSession(p, xyz, box, times)
        ct = torch.tensor(coef, dtype=torch.float64, requires_grad=True)
        ls = IRLineshape.apply(ct, sess)
        if prec == "fp64":
            target = (ls.detach() * 0.9).clone()
        loss = ((ls - target) ** 2).sum()
        loss.backward()
        grads[prec] = ct.grad.numpy().copy()
        # the autograd path is the session's backward (a fresh forward: the
        # eigensolver's atomics reorder sums, so agreement is to rounding)
        ft = sess.forward(coef)
        _, dL = _loss_grad(p, ft, target.numpy())
        tol = 1e-9 if prec == "fp64" else 1e-3
        assert np.abs(sess.backward(dL) * coef - grads[prec] * coef).max() <= tol * np.abs(grads[prec] * coef).max()
        sess.close()
    s64, s32 = coef * grads["fp64"], coef * grads["fp32"]
    assert np.abs(s32 - s64).max() <= 1e-3 * np.abs(s64).max(), (s32, s64)


def test_backward_needs_forward_and_coefficients_round_trip():
    xyz, box, times = frames(16, 24, 3)
    p = C.default_params(precision="fp64", species="H2O", **dict(KW, ntcfpoints=6, nsamples=2, sampleEvery=0.05))
    sess = LineshapeSession(p, xyz, box, times)
    with pytest.raises(C.CalcirError):
        sess.backward(np.zeros(sess.nfft))
    assert np.array_equal(sess.coefs(), C.map_coefs(0, 0))
    cf = C.map_coefs(0, 0) * 1.01
    sess.forward(cf)
    assert np.array_equal(sess.coefs(), cf)
    # cir_run with the same override gives the same IR spectrum
    ft = sess.forward()
    r = C.run(p, xyz, box, times, coefs=cf)["ftcf_ir"]
    assert np.abs(ft - r).max() <= 1e-12 * np.abs(r).max()
    sess.close()


def test_recomputed_eigensystems_give_the_stored_gradient():
    # adjoint_store 2 recomputes each chunk's eigensystems in the reverse
    # sweep (the memory-bounded form); chunk 8 makes several chunks
    xyz, box, times = frames(64, 40, 2024)
    coef = C.map_coefs(0, 0)
    g = {}
    for store in (1, 2):
        announce("adjoint fp64 adjoint_store=%d" % store)
        p = C.default_params(precision="fp64", species="H2O", chunk=8, adjoint_store=store, **KW)
        sess = LineshapeSession(p, xyz, box, times)
        ft = sess.forward(coef)
        idx, _, scale = lineshape_map(p)
        dL = np.zeros_like(ft)
        np.add.at(dL, idx, np.sin(np.arange(idx.size) * 0.1) * scale)
        g[store] = (ft, sess.backward(dL))
        sess.close()
    assert np.abs(g[1][0] - g[2][0]).max() <= 1e-12 * np.abs(g[1][0]).max()
    s1, s2 = coef * g[1][1], coef * g[2][1]
    assert np.abs(s1 - s2).max() <= 1e-10 * np.abs(s1).max()


def test_lineshape_gradient_1024oh_workload():
    # 512 TIP4P/2005 waters (1024 OH), 400 frames, origins every 10 frames
    # (26 samples of 150 points); loss: L2 between the IR lineshape and one
    # made at seeded +-2% perturbed coefficients.  Adjoint (FP64) against
    # central differences (h = 1e-4 relative) of the FP64 forward on the 7
    # coefficients of omega = a0 + a1 E + a2 E^2, mu' = b0 + b1 E,
    # x10 = c0 + c1 omega; FP32 adjoint against the FP64 gradient.
    xyz, box, times = frames(512, 400, 1004)
    kw = dict(species="H2O", model="tip4p2005", map=0, ntcfpoints=150, nsamples=26, sampleEvery=0.1,
              avef=3400.0, omegaStart=2800.0, omegaStop=4000.0, omegaStep=5, nzeros=1898, t1=0.26, dt=0.010)
    coef = C.map_coefs(0, 0)
    pert = coef * (1 + 0.02 * np.random.default_rng(1004).choice([-1.0, 1.0], coef.size))
    p64 = C.default_params(precision="fp64", **kw)
    idx, _, scale = lineshape_map(p64)
    announce("1024 OH: target at perturbed coefficients")
    sess = LineshapeSession(p64, xyz, box, times)
    target = sess.forward(pert)[idx] * scale
    announce("1024 OH: fp64 forward/backward")
    ft = sess.forward(coef)
    L, dL = _loss_grad(p64, ft, target)
    g64 = sess.backward(dL)
    for name in ("w0", "w1", "w2", "mu0", "mu1", "x0", "x1"):
        c = C.COEF_NAMES.index(name)
        d = 1e-4 * abs(coef[c])
        vals = []
        for sgn in (1.0, -1.0):
            cf = coef.copy()
            cf[c] += sgn * d
            announce("1024 OH: fd %s %+d" % (name, sgn))
            r = sess.forward(cf)[idx] * scale - target
            vals.append(float(r @ r))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(g64[c] - fd) <= 1e-4 * abs(fd), (name, g64[c], fd)
    sess.close()
    announce("1024 OH: fp32 forward/backward")
    p32 = C.default_params(precision="fp32", **kw)
    sess = LineshapeSession(p32, xyz, box, times)
    ft32 = sess.forward(coef)
    _, dL32 = _loss_grad(p32, ft32, target)
    g32 = sess.backward(dL32)
    sess.close()
    s64, s32 = coef * g64, coef * g32
    # measured agreement ~1.4e-7 of the largest sensitivity
    assert np.abs(s32 - s64).max() <= 1e-5 * np.abs(s64).max(), (s32 / s64)


def test_autograd_interleaved_forwards_differentiate_their_own_coefficients():
    # two apply() calls on one session before either backward: each backward
    # must give the gradient at its own coefficients (the session's context
    # holds only the last forward pass, so the older one is re-run), on the
    # device of the coefficient tensor
    torch = pytest.importorskip("torch")
    from calcir_amd.adjoint import IRLineshape

    xyz, box, times = frames(32, 40, 5)
    p = C.default_params(precision="fp64", species="H2O", **KW)
    ca = C.map_coefs(0, 0)
    cb = ca * 1.02
    sess = LineshapeSession(p, xyz, box, times)
    idx, _, scale = lineshape_map(p)
    want = {}
    for name, cf in (("a", ca), ("b", cb)):
        ft = sess.forward(cf)
        dL = np.zeros_like(ft)
        np.add.at(dL, idx, scale)  # d(sum of rows)/d ftcf
        want[name] = sess.backward(dL)
    for dev in ("cpu", "cuda"):
        announce("adjoint interleaved apply on %s" % dev)
        ta = torch.tensor(ca, dtype=torch.float64, device=dev, requires_grad=True)
        tb = torch.tensor(cb, dtype=torch.float64, device=dev, requires_grad=True)
        la = IRLineshape.apply(ta, sess)
        lb = IRLineshape.apply(tb, sess)
        sess.forward(cb * 0.99)  # an unrelated forward in between
        la.sum().backward()
        lb.sum().backward()
        assert ta.grad.device.type == dev and tb.grad.device.type == dev
        for name, t in (("a", ta), ("b", tb)):
            g = t.grad.cpu().numpy()
            assert np.abs(g - want[name]).max() <= 1e-9 * np.abs(want[name]).max(), name
    assert np.abs(want["a"] - want["b"]).max() > 1e-4 * np.abs(want["a"]).max()
    sess.close()
