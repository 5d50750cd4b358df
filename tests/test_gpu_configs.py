"""The benchmark workloads at their own sizes (calcIR.cu:360-763 driver loop):
the 2048-OH box through the whole pipeline against the FP64 CPU oracle, the
8192-OH box's eigenpairs and re-unitarised propagation, and the 256-OH box
with its 0.75 nm cutoff."""

import os

import numpy as np
import pytest
import torch

import calcir_amd as C
import oracle_lib as O
from calcir_amd import stages as S
from cases import rel, rel_l2
from conftest import announce

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
BASE = dict(species="H2O", model="tip4p2005", avef=3400.0, omegaStart=2800.0, omegaStop=4000.0, omegaStep=5,
            t1=0.26, dt=0.010)


def _oracle_threads():
    return max(1, min(16, os.cpu_count() or 1))


def _spectra_close(g, o, tol):
    out = {}
    for k in ("tcf_ir", "tcf_vv", "tcf_vh"):
        out[k] = rel(g[k], o[k])
    for k in ("ftcf_ir", "ftcf_vv", "ftcf_vh"):
        out[k] = rel_l2(g[k], o[k])
    out["sw"] = rel(g["sw"], o["sw"])
    print("[errors] " + " ".join("%s=%.2e" % kv for kv in out.items()), flush=True)
    for k, v in out.items():
        assert v < tol, (k, v, tol)


def test_2048_oh_box_pipeline_against_fp64_oracle():
    # 1024 waters (N = 2048), L = 3.128 nm, cutoff 0.7831 nm, seed 1002; two
    # overlapping origins one frame apart, 6 points each (7 distinct frames,
    # 12 steps).  The CPU oracle in FP64 with the vector-form propagation is
    # the reference: batched FP32 within 1e-4, batched BF16x3 within 1e-3.
    xyz, box, times = C.synth_trajectory(1024, 8, 1002, box_nm=3.128)
    kw = dict(BASE, ntcfpoints=6, nzeros=2042, nsamples=2, sampleEvery=0.01, cutoff=0.7831)
    p64 = C.default_params(precision="fp64", mode="batched", **kw)
    cfg = O.config_from_params(p64)
    cfg.nthreads = _oracle_threads()
    announce("2048 OH: FP64 oracle, %d threads" % cfg.nthreads)
    o = O.run(O.F64, cfg, xyz, box, times)
    for prec, tol in (("fp32", 1e-4), ("bf16x3", 1e-3)):
        announce("2048 OH: batched %s" % prec)
        g = C.run(C.default_params(precision=prec, mode="batched", **kw), xyz, box, times)
        assert g["nsteps"] == 12 and g["nframes_used"] == 7
        _spectra_close(g, o, tol)


def test_8192_oh_eigenpairs_residual_and_orthogonality():
    # one frame of the 8192-OH box (4096 waters, L = 4.965 nm, seed 1005):
    # the FP32 eigensolver's residual max |H v - lambda v| and orthogonality
    # max |V^T V - I|, formed in FP64 on the device
    xyz, box, _ = C.synth_trajectory(4096, 1, 1005, box_nm=4.965)
    announce("8192 OH: H and eigh (FP32)")
    x_d, b_d = torch.from_numpy(xyz).to(DEV), torch.from_numpy(box).to(DEV)
    sg = S.eproj_sites(x_d, b_d)
    H = S.hamiltonian(sg, b_d, avef=3400.0)
    n = H.shape[1]
    assert n == 8192
    Hd = H[0, :, :n].double().t().contiguous()
    w, V = S.eigh(H.clone())
    w = w[0].to(DEV).double()
    Vd = V[0, :, :n].double().contiguous()
    res = (Hd @ Vd - Vd * w[None, :]).abs().max().item()
    orth = (Vd.t().contiguous() @ Vd - torch.eye(n, dtype=torch.float64, device=DEV)).abs().max().item()
    wn = torch.linalg.eigvalsh(Hd)
    dw = (w - wn).abs().max().item()
    scale = wn.abs().max().item()
    print("[errors] residual=%.3e orth=%.3e dlambda=%.3e scale=%.1f" % (res, orth, dw, scale), flush=True)
    # measured: residual 2.7e-3 cm^-1 and eigenvalue error 1.3e-3 cm^-1 for
    # |lambda| up to 1032 cm^-1 (about 2 eps ||H|| sqrt(n)), orthogonality 7e-6
    assert torch.all(w[1:] >= w[:-1])
    assert res < 1e-2, res
    assert orth < 5e-5, orth
    assert dw < 1e-2, dw


def test_8192_oh_reunitarised_propagation():
    # the 8192-OH box over 26 frames, one origin x 26 points: BF16x3 with the
    # propagated state re-unitarised every 25 points against plain FP32 (the
    # FP32 path is held to 1e-4 of the FP64 pipeline at this size elsewhere)
    xyz, box, times = C.synth_trajectory(4096, 26, 1005, box_nm=4.965)
    kw = dict(BASE, ntcfpoints=26, nzeros=2022, nsamples=1, sampleEvery=0.01, cutoff=0.7831)
    announce("8192 OH: 26 frames FP32")
    g32 = C.run(C.default_params(precision="fp32", mode="batched", **kw), xyz, box, times)
    announce("8192 OH: 26 frames BF16x3, reunitarise_every 25")
    gr = C.run(C.default_params(precision="bf16x3", mode="batched", reunitarise_every=25, **kw), xyz, box, times)
    assert gr["nsteps"] == 26
    _spectra_close(gr, g32, 1e-3)


@pytest.mark.parametrize("prec,mode,tol", [("fp32", "batched", 1e-4), ("fp32", "plain", 1e-4),
                                           ("bf16x3", "batched", 1e-3)])
def test_256_oh_box_with_its_075_nm_cutoff(prec, mode, tol):
    # 128 waters (N = 256), L = 1.564 nm, cutoff 0.75 nm, seed 1001; three
    # origins 10 frames apart, 30 points each, against the FP64 oracle
    xyz, box, times = C.synth_trajectory(128, 60, 1001, box_nm=1.564)
    kw = dict(BASE, ntcfpoints=30, nzeros=2018, nsamples=3, sampleEvery=0.1, cutoff=0.75)
    p = C.default_params(precision=prec, mode=mode, **kw)
    announce("256 OH cutoff 0.75: %s %s" % (prec, mode))
    g = C.run(p, xyz, box, times)
    o = O.run(O.F64, O.config_from_params(p), xyz, box, times)
    assert g["nsteps"] == 90
    _spectra_close(g, o, tol)
