"""Per-stage device calls of libcalcir on torch tensors (current torch stream).

Each function validates shapes on the host before launching, so that a kernel
never sees buffers smaller than its grid assumes.  Layouts follow
include/calcir.h:
  sites      [nframes, NSITE, N]            (field, maps, OH vector, TDM, mu, alpha)
  H          [nframes, N, ld]  column-major per frame (symmetric: H[f, j, i] = H_ij)
  V          [nframes, N, ld]  row-major per frame: V[f, i, k] = component i of mode k
  w          [nframes, N]      float64, ascending
"""

import ctypes

import torch

from . import NSITE, PREC_BF16X3, PREC_FP32, PREC_FP64, SITE_MU, check, lib

_vp = ctypes.c_void_p


def _p(t):
    return _vp(t.data_ptr()) if t is not None else None


def _stream():
    return _vp(torch.cuda.current_stream().cuda_stream)


def _prec_of(dtype):
    return PREC_FP64 if dtype == torch.float64 else PREC_FP32


def _need(cond, msg):
    if not cond:
        raise ValueError(msg)


def eproj_sites(xyz, box, natom_mol=4, nchrom_mol=2, imodel=1, ispecies=0, imap=0, cutoff=0.7831,
                dtype=torch.float32):
    """Field projection + site records for a batch of frames (get_eproj_GPU and
    the per-row part of get_kappa_GPU, calcIR.cu:1020-1304)."""
    _need(xyz.is_cuda and xyz.dtype == torch.float32 and xyz.dim() == 3 and xyz.shape[2] == 3,
          "xyz must be a cuda float32 [nframes, natoms, 3] tensor")
    nframes, natoms = xyz.shape[0], xyz.shape[1]
    _need(natoms % natom_mol == 0, "natoms must be a multiple of natom_mol")
    _need(box.shape == (nframes, 3) and box.dtype == torch.float32 and box.is_cuda, "box must be [nframes, 3]")
    xyz, box = xyz.contiguous(), box.contiguous()
    n = natoms // natom_mol * nchrom_mol
    sites = torch.empty((nframes, NSITE, n), dtype=dtype, device=xyz.device)
    check(lib().cir_eproj_sites(_prec_of(dtype), nframes, natoms, natom_mol, nchrom_mol, imodel, ispecies, imap,
                                cutoff, _p(xyz), _p(box), _p(sites), _stream()))
    return sites


def hamiltonian(sites, box, nchrom_mol=2, ispecies=0, imap=0, avef=3415.2, ld=None):
    """Exciton Hamiltonian (calcIR.cu:1306-1404) from site records."""
    _need(sites.is_cuda and sites.dim() == 3 and sites.shape[1] == NSITE, "sites must be [nframes, NSITE, N]")
    nframes, n = sites.shape[0], sites.shape[2]
    ld = ld or (n + 3) // 4 * 4
    _need(ld >= n, "ld must be >= N")
    _need(box.shape == (nframes, 3), "box must be [nframes, 3]")
    H = torch.empty((nframes, n, ld), dtype=sites.dtype, device=sites.device)
    check(lib().cir_hamiltonian(_prec_of(sites.dtype), nframes, n, ld, nchrom_mol, ispecies, imap, avef,
                                _p(box.contiguous()), _p(sites.contiguous()), _p(H), _stream()))
    return H


def eigh(H, n=None, prec=None):
    """Batched eigensolve of H [batch, n, ld] (destroyed).  Returns (w, V).
    prec=PREC_BF16X3 (float32 H) forms the back-transform products with
    split-BF16x3 MFMA."""
    _need(H.is_cuda and H.dim() == 3 and H.is_contiguous(), "H must be a contiguous cuda [batch, n, ld] tensor")
    batch, nn, ld = H.shape
    n = n or nn
    _need(nn == n and ld >= n and ld % 4 == 0, "H must be [batch, n, ld>=n], ld a multiple of 4")
    if prec is None:
        prec = _prec_of(H.dtype)
    _need(prec == _prec_of(H.dtype) or (prec == PREC_BF16X3 and H.dtype == torch.float32),
          "prec must match H's dtype (or be PREC_BF16X3 for float32)")
    V = torch.zeros_like(H)
    w = torch.empty((batch, n), dtype=torch.float64, device=H.device)
    ws = lib().cir_eigh_workspace(prec, batch, n, ld)
    work = torch.empty(ws, dtype=torch.uint8, device=H.device)
    check(lib().cir_eigh_batched(prec, batch, n, ld, _p(H), _p(V), _p(w), _p(work), _stream()))
    return w, V


def ipr_hist(V, w, nomega, avef, omegaStart, omegaStep, nchrom_mol=2, weight=None):
    """Pw, Rw, Rmw accumulators (unnormalised, calcIR.cu:494-538) and the count
    of modes outside the window."""
    batch, n, ld = V.shape
    _need(w.shape == (batch, n) and w.dtype == torch.float64, "w must be float64 [batch, n]")
    acc = torch.zeros((3, nomega), dtype=torch.float64, device=V.device)
    skipped = torch.zeros(1, dtype=torch.int64, device=V.device)
    if weight is not None:
        _need(weight.shape == (batch,) and weight.dtype == torch.float64, "weight must be float64 [batch]")
    check(lib().cir_ipr_hist(_prec_of(V.dtype), batch, n, ld, nchrom_mol, avef, omegaStart, omegaStep, nomega,
                             _p(V), _p(w), _p(weight), _p(acc[0]), _p(acc[1]), _p(acc[2]), _p(skipped),
                             _stream()))
    return acc, skipped


def spectral_density(V, w, sites, nomega, t1, avef, omegaStart, omegaStep, weight=None):
    """Sw accumulator (calcIR.cu:458-482, 1414-1444)."""
    batch, n, ld = V.shape
    _need(sites.shape == (batch, NSITE, n), "sites must be [batch, NSITE, n]")
    sw = torch.zeros(nomega, dtype=torch.float64, device=V.device)
    work = torch.empty((batch, n), dtype=torch.float64, device=V.device)
    check(lib().cir_spectral_density(_prec_of(V.dtype), batch, n, ld, t1, avef, omegaStart, omegaStep, nomega,
                                     _p(V), _p(w), _p(sites), _p(weight), _p(sw), _p(work), _stream()))
    return sw


def propagate(V, w, Y, ncols, dt, prec=None):
    """Y <- V exp(+i diag(w) dt/hbar) V^T Y for Y [n, ldy] (Re columns [0,ncols),
    Im columns [ncols, 2ncols)).  V: [n, ld] (one frame), w: [n] float64."""
    n, ld = V.shape
    _need(Y.shape[0] == n and Y.shape[1] >= 2 * ncols and Y.dtype == V.dtype and Y.is_contiguous(),
          "Y must be [n, ldy >= 2*ncols] of V's dtype")
    _need(w.shape == (n,) and w.dtype == torch.float64, "w must be float64 [n]")
    if prec is None:
        prec = _prec_of(V.dtype)
    _need((prec == PREC_FP64) == (V.dtype == torch.float64), "prec does not match the dtype")
    work = torch.empty(lib().cir_propagate_workspace(prec, n, ncols), dtype=torch.uint8, device=V.device)
    check(lib().cir_propagate(prec, n, ld, ncols, Y.shape[1], dt, _p(V.contiguous()), _p(w.contiguous()), _p(Y),
                              _p(Y), _p(work), _stream()))
    return Y


def correlate(sites, Y, nslot=1):
    """IR, VV, VH values (calcIR.cu:628-745) for each slot: returns complex128
    [nslot, 3]."""
    n = sites.shape[-1]
    _need(sites.shape[-2] == NSITE and Y.shape[0] == n and Y.shape[1] >= 18 * nslot, "bad shapes")
    out = torch.zeros((nslot, 6), dtype=torch.float64, device=Y.device)
    target = torch.arange(nslot, dtype=torch.int32, device=Y.device)
    check(lib().cir_correlate(_prec_of(Y.dtype), n, nslot, Y.shape[1], None, _p(target), _p(sites.contiguous()),
                              _p(Y), _p(out), _stream()))
    o = out.cpu()
    return torch.complex(o[:, 0::2], o[:, 1::2])


def c2r(X, nfft):
    """cuFFT C2R contract of X (complex128 [nspec, nx]) to [nspec, nfft]."""
    nspec, nx = X.shape
    Xr = torch.view_as_real(X.to(torch.complex128).contiguous()).contiguous()
    y = torch.empty((nspec, nfft), dtype=torch.float64, device=X.device)
    check(lib().cir_spectrum_c2r(nspec, nfft, nx, _p(Xr), _p(y), _stream()))
    return y


def gemm(A, B, prec=PREC_FP32, transA=False, transB=False):
    """C = op(A) op(B) with the library's batched product (2-D or 3-D inputs)."""
    squeeze = A.dim() == 2
    if squeeze:
        A, B = A.unsqueeze(0), B.unsqueeze(0)
    A, B = A.contiguous(), B.contiguous()
    bt = A.shape[0]
    if transA:
        K, M = A.shape[1], A.shape[2]
        a_m, a_k = 1, A.shape[2]
    else:
        M, K = A.shape[1], A.shape[2]
        a_m, a_k = A.shape[2], 1
    if transB:
        N = B.shape[1]
        _need(B.shape[2] == K, "inner dimensions differ")
        b_k, b_n = 1, B.shape[2]
    else:
        N = B.shape[2]
        _need(B.shape[1] == K, "inner dimensions differ")
        b_k, b_n = B.shape[2], 1
    C = torch.zeros((bt, M, N), dtype=A.dtype, device=A.device)
    check(lib().cir_gemm(prec, M, N, K, bt, _p(A), a_m, a_k, A.shape[1] * A.shape[2], _p(B), b_k, b_n,
                         B.shape[1] * B.shape[2], _p(C), N, 1, M * N, 1.0, 0.0, _stream()))
    return C[0] if squeeze else C


__all__ = ["eproj_sites", "hamiltonian", "eigh", "ipr_hist", "spectral_density", "propagate", "correlate",
           "c2r", "gemm", "PREC_FP32", "PREC_FP64", "PREC_BF16X3"]
