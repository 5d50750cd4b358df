"""cir_propagate_noalloc (the propagation step with its work buffer in the
library's fixed device array) against cir_propagate with a workspace, through
the C ABI: two calls in a row, the second with a smaller input, in each
precision form; shapes beyond the array are refused."""

import ctypes

import numpy as np
import pytest

import calcir_amd as C
from conftest import announce

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _case(torch, n, ncols, dtype, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ld = (n + 3) & ~3
    V = np.zeros((n, ld))
    V[:, :n] = q
    w = rng.uniform(-300.0, 300.0, n)
    ldy = 2 * ncols + 4
    Y = rng.standard_normal((n, ldy))
    dev = torch.device("cuda")
    return (torch.tensor(V, dtype=dtype, device=dev), ld, torch.tensor(w, dtype=torch.float64, device=dev),
            torch.tensor(Y, dtype=dtype, device=dev), ldy)


@pytest.mark.parametrize("prec", [C.PREC_FP32, C.PREC_BF16X3, C.PREC_FP64])
def test_noalloc_propagation_equals_the_workspace_form(prec):
    torch = pytest.importorskip("torch")
    L = C.lib()
    dtype = torch.float64 if prec == C.PREC_FP64 else torch.float32
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dt = np.float32(0.010)
    for n, ncols, seed in ((1024, 72, 1), (384, 27, 2)):  # the second call is smaller
        announce("propagate_noalloc prec=%d n=%d ncols=%d" % (prec, n, ncols))
        V, ld, w, Y, ldy = _case(torch, n, ncols, dtype, seed)
        ref = torch.empty_like(Y)
        out = torch.full_like(Y, float("nan"))
        work = torch.empty(L.cir_propagate_workspace(prec, n, ncols), dtype=torch.uint8, device=V.device)
        C.check(L.cir_propagate(prec, n, ld, ncols, ldy, dt, _p(V), _p(w), _p(Y), _p(ref), _p(work), stream))
        C.check(L.cir_propagate_noalloc(prec, n, ld, ncols, ldy, dt, _p(V), _p(w), _p(Y), _p(out), stream))
        torch.cuda.synchronize()
        a = ref[:, :2 * ncols].double().cpu().numpy()
        b = out[:, :2 * ncols].double().cpu().numpy()
        assert np.isfinite(b).all()
        # FP64 runs the same split in a fixed order; FP32 and BF16x3 split K
        # with float atomics, whose order changes the last bits
        tol = 1e-12 if prec == C.PREC_FP64 else 1e-5
        assert np.abs(a - b).max() <= tol * np.abs(a).max(), np.abs(a - b).max()
        # in place (Yin == Yout), as the batched pipeline calls it
        Yi = Y.clone()
        C.check(L.cir_propagate_noalloc(prec, n, ld, ncols, ldy, dt, _p(V), _p(w), _p(Yi), _p(Yi), stream))
        torch.cuda.synchronize()
        c = Yi[:, :2 * ncols].double().cpu().numpy()
        assert np.abs(a - c).max() <= tol * np.abs(a).max()


def test_noalloc_refuses_shapes_beyond_its_array():
    torch = pytest.importorskip("torch")
    L = C.lib()
    cap = L.cir_propagate_noalloc_max_bytes()
    assert cap == 32 << 20
    n, ncols = 8192, 1024
    assert L.cir_propagate_workspace(C.PREC_FP64, n, ncols) > cap
    V = torch.zeros(8, dtype=torch.float64, device="cuda")  # never read: refused before any launch
    st = L.cir_propagate_noalloc(C.PREC_FP64, n, n, ncols, 2 * ncols, np.float32(0.01), _p(V), _p(V), _p(V), _p(V),
                                 None)
    assert st == 1 and b"32 MiB" in L.cir_last_error()
