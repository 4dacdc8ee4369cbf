"""GPU rank-predictor network (FP32 HIP forward/backward/Adam) against the
FP64 numpy restatement.  Tolerances are FP32-level: forward 1e-4 relative to
the output scale, gradients 2e-3 relative to the largest gradient entry."""
import ctypes as C

import numpy as np
import pytest

import mlp_ref as M

pytestmark = pytest.mark.gpu


def _lib():
    import paratt_amd as P
    L = P.lib()
    L.ptt_hip_mlp_forward.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.ptt_hip_mlp_grad.argtypes = [C.c_int] + [C.c_void_p] * 5
    L.ptt_hip_mlp_adam.argtypes = [C.c_void_p] * 4 + [C.c_float, C.c_int]
    return L


def _data(ns, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((ns, M.NIN))
    Y = np.tanh(X[:, 0]) + 0.3 * X[:, 2] * X[:, 3]
    return X, Y


def test_nparam_matches_layout():
    assert _lib().ptt_hip_mlp_nparam() == M.NPARAM


@pytest.mark.parametrize("ns", [1, 15, 16, 17, 1000])
def test_forward_and_grad(ns):
    from gpu_util import DevBuf
    L = _lib()
    X, Y = _data(ns, ns)
    p = M.init(5).astype(np.float32).astype(np.float64)
    Xf = X.astype(np.float32).astype(np.float64)
    Yf = Y.astype(np.float32).astype(np.float64)
    out_ref, _ = M.forward(p, Xf)
    loss_ref, g_ref = M.loss_grad(p, Xf, Yf)
    dX, dY, dP = DevBuf(Xf, np.float32), DevBuf(Yf, np.float32), DevBuf(p, np.float32)
    dO = DevBuf(np.zeros(ns), np.float32)
    dG = DevBuf(np.full(M.NPARAM, 7.0), np.float32)  # overwritten by the call
    dL = DevBuf(np.full(1, 7.0), np.float32)
    L.ptt_hip_mlp_forward(ns, dX.ptr, dP.ptr, dO.ptr)
    L.ptt_hip_mlp_grad(ns, dX.ptr, dY.ptr, dP.ptr, dG.ptr, dL.ptr)
    L.ptt_hip_sync()
    out = dO.get().astype(np.float64)
    g = dG.get().astype(np.float64)
    loss = float(dL.get()[0])
    scale = max(1.0, np.max(np.abs(out_ref)))
    assert np.max(np.abs(out - out_ref)) <= 1e-4 * scale
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, loss_ref)
    assert np.max(np.abs(g - g_ref)) <= 2e-3 * np.max(np.abs(g_ref))
    for b in (dX, dY, dP, dO, dG, dL):
        b.free()


def test_adam_training_matches_reference():
    from gpu_util import DevBuf
    L = _lib()
    X, Y = _data(256, 9)
    Xf = X.astype(np.float32).astype(np.float64)
    Yf = Y.astype(np.float32).astype(np.float64)
    p = M.init(6).astype(np.float32).astype(np.float64)
    pr, m, v = p.copy(), np.zeros_like(p), np.zeros_like(p)
    dX, dY, dP = DevBuf(Xf, np.float32), DevBuf(Yf, np.float32), DevBuf(p, np.float32)
    dG, dL = DevBuf(np.zeros(M.NPARAM), np.float32), DevBuf(np.zeros(1), np.float32)
    dM, dV = DevBuf(np.zeros(M.NPARAM), np.float32), DevBuf(np.zeros(M.NPARAM), np.float32)
    l0 = M.loss_grad(pr, Xf, Yf)[0]
    for t in range(1, 51):
        _, g = M.loss_grad(pr, Xf, Yf)
        M.adam(pr, g, m, v, 1e-3, t)
        L.ptt_hip_mlp_grad(256, dX.ptr, dY.ptr, dP.ptr, dG.ptr, dL.ptr)
        L.ptt_hip_mlp_adam(dP.ptr, dG.ptr, dM.ptr, dV.ptr, C.c_float(1e-3), t)
    L.ptt_hip_sync()
    pg = dP.get().astype(np.float64)
    lg = M.loss_grad(pg, Xf, Yf)[0]
    lr_ = M.loss_grad(pr, Xf, Yf)[0]
    assert lg < 0.8 * l0
    assert abs(lg - lr_) <= 0.05 * lr_
    # Adam steps are sign-like, so parameters agree to ~lr per step at most
    assert np.max(np.abs(pg - pr)) < 50 * 1e-3
    for b in (dX, dY, dP, dG, dL, dM, dV):
        b.free()
