"""Learned rank prediction for the step-truncation composites.

Pipeline (all compute on the GPU through libparatt_amd.so):
  1. collect(): run the plain method (full orthogonalization + SVD truncation
     at every bond) on a seeded family of reaction-Burgers initial conditions
     with logging on; every rounding of every truncation site yields, per
     bond, the 18 features of its previous rounding (16 log singular values
     around the previous cut, bond index, previous rank) and the new rank.
  2. train(): fit the 18 -> 64 -> 64 -> 64 -> 1 GELU network with the
     library's FP32 forward/backward/Adam kernels (ptt_hip_mlp_*).
  3. enable(): load the parameters and switch the composites to learned
     mode: each truncation predicts every bond's rank, sizes a randomized
     range finder (predicted rank + 8 columns, one power iteration) with it,
     and keeps the sketched step only when its rank is certified equal to
     the plain method's; otherwise the plain step runs.

Initial-condition family: u0 = prod_i exp(-(x_i - c_i)^2 / w_i) on the
periodic grid x_j = 2 pi j / n, with c_i = pi + 0.3 g_i and
w_i = 0.5 (1 + 0.3 v_i), g_i ~ N(0, 1), v_i ~ U(-1, 1) drawn from
numpy.random.default_rng(seed); seed None is the unperturbed condition
(c_i = pi, w_i = 0.5).
"""
import ctypes as C
import math
import os

import numpy as np

from . import DevTT, Paratt, lib, sync

NFEAT = 18
HERE = os.path.dirname(os.path.abspath(__file__))
PARAM_FILE = os.path.join(HERE, "data", "rank_predictor_d20_n256_eps1e-8.npy")


def _L():
    L = lib()
    PP = C.POINTER(Paratt)
    L.paratt_amd_rxn_burgers_setup.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
    L.paratt_amd_rk4_st.argtypes = [PP, PP, C.c_void_p, C.c_double, C.c_double, C.c_int]
    L.paratt_amd_learned_set_params.argtypes = [C.POINTER(C.c_float), C.c_int]
    L.paratt_amd_learned_mode.argtypes = [C.c_int]
    L.paratt_amd_learned_log.argtypes = [C.c_int, C.c_int]
    L.paratt_amd_learned_log_fetch.restype = C.c_int
    L.paratt_amd_learned_log_fetch.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]
    L.paratt_amd_learned_counts.argtypes = [C.POINTER(C.c_long)]
    L.ptt_hip_mlp_nparam.restype = C.c_int
    L.ptt_hip_mlp_forward.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    L.ptt_hip_mlp_grad.argtypes = [C.c_int] + [C.c_void_p] * 5
    L.ptt_hip_mlp_adam.argtypes = [C.c_void_p] * 4 + [C.c_float, C.c_int]
    L.ptt_hip_malloc.restype = C.c_void_p
    L.ptt_hip_malloc.argtypes = [C.c_size_t]
    L.ptt_hip_free.argtypes = [C.c_void_p]
    L.ptt_hip_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.ptt_hip_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    return L


class _F32:
    """A device float32 buffer."""

    def __init__(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        self.n = a.size
        self.ptr = _L().ptt_hip_malloc(max(a.nbytes, 4))
        if a.size:
            _L().ptt_hip_h2d(self.ptr, a.ctypes.data, a.nbytes)

    def get(self):
        out = np.empty(self.n, dtype=np.float32)
        _L().ptt_hip_d2h(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr:
            _L().ptt_hip_free(self.ptr)
            self.ptr = None


def initial_condition(d, n, seed=None):
    """Rank-1 device TT of the seeded Gaussian family (see module doc)."""
    xs = 2 * np.pi * np.arange(n) / n
    if seed is None:
        c = np.full(d, np.pi)
        w = np.full(d, 0.5)
    else:
        g = np.random.default_rng(seed)
        c = np.pi + 0.3 * g.standard_normal(d)
        w = 0.5 * (1.0 + 0.3 * g.uniform(-1.0, 1.0, d))
    u = DevTT([n] * d, [1] * (d + 1))
    L = lib()
    for k in range(d):
        v = np.ascontiguousarray(np.exp(-(xs - c[k]) ** 2 / w[k]))
        L.paratt_amd_core_upload(u.tt, k, v.ctypes.data_as(C.POINTER(C.c_double)))
    return u


def setup(n, nu=0.01, eps=1e-8, cap=64):
    _L().paratt_amd_rxn_burgers_setup(nu, 2 * math.pi / n, eps, cap)


def rk4(u, steps, dt=1e-3, eps=1e-8, cap=64):
    """Advance the device TT u by `steps` step-truncation RK4 steps (in place)."""
    L = _L()
    field = C.cast(L.paratt_amd_rxn_burgers_field, C.c_void_p)
    un = DevTT(u.n, u.r)
    for _ in range(steps):
        L.paratt_amd_rk4_st(u.p, un.p, field, dt, eps, cap)
        L.copy_rank_change_tt(u.tt, un.tt)
    sync()
    return u


def collect(seeds, steps, d=20, n=256, warmup=0, capacity=400000):
    """Features/targets from plain runs, one initial condition per seed."""
    L = _L()
    setup(n)
    L.paratt_amd_learned_mode(0)
    X = np.zeros((capacity, NFEAT), dtype=np.float32)
    Y = np.zeros(capacity, dtype=np.float32)
    tot = 0
    for s in seeds:
        L.paratt_amd_learned_reset_history()
        u = initial_condition(d, n, s)
        if warmup:
            rk4(u, warmup)
        L.paratt_amd_learned_log(1, capacity - tot)
        rk4(u, steps)
        m = L.paratt_amd_learned_log_fetch(X[tot:].ctypes.data_as(C.POINTER(C.c_float)),
                                           Y[tot:].ctypes.data_as(C.POINTER(C.c_float)),
                                           capacity - tot)
        L.paratt_amd_learned_log(0, 0)
        tot += m
        u.free()
    return X[:tot], Y[:tot]


_params_loaded = [False]


def _has_params():
    return _params_loaded[0]


def init_params(seed=0):
    """He-scaled random initial parameters (layout of the HIP kernels)."""
    g = np.random.default_rng(seed)
    nh = 64
    parts = [g.standard_normal((NFEAT, nh)) / math.sqrt(NFEAT), 0.1 * g.standard_normal(nh),
             g.standard_normal((nh, nh)) / math.sqrt(nh), 0.1 * g.standard_normal(nh),
             g.standard_normal((nh, nh)) / math.sqrt(nh), 0.1 * g.standard_normal(nh),
             g.standard_normal(nh) / math.sqrt(nh), np.zeros(1)]
    p = np.concatenate([q.ravel() for q in parts]).astype(np.float32)
    assert p.size == _L().ptt_hip_mlp_nparam()
    return p


def train(X, Y, iters=3000, lr=2e-3, seed=0, P0=None, log_every=0):
    """Full-batch Adam on the GPU; returns (parameters, loss history)."""
    L = _L()
    P = init_params(seed) if P0 is None else np.asarray(P0, dtype=np.float32)
    npar = P.size
    dX, dY, dP = _F32(X), _F32(Y), _F32(P)
    dG, dL = _F32(np.zeros(npar)), _F32(np.zeros(1))
    dM, dV = _F32(np.zeros(npar)), _F32(np.zeros(npar))
    hist = []
    ns = int(len(Y))
    for t in range(1, iters + 1):
        L.ptt_hip_mlp_grad(ns, dX.ptr, dY.ptr, dP.ptr, dG.ptr, dL.ptr)
        # cosine-decayed step
        lrt = lr * 0.5 * (1.0 + math.cos(math.pi * t / iters))
        L.ptt_hip_mlp_adam(dP.ptr, dG.ptr, dM.ptr, dV.ptr, C.c_float(lrt), t)
        if t == 1 or t == iters or (log_every and t % log_every == 0):
            hist.append((t, float(dL.get()[0])))
    out = dP.get()
    for b in (dX, dY, dP, dG, dL, dM, dV):
        b.free()
    return out, hist


def predict(P, X):
    """Network outputs (rank / 64) for feature rows X."""
    L = _L()
    dX, dP, dO = _F32(X), _F32(P), _F32(np.zeros(len(X)))
    L.ptt_hip_mlp_forward(len(X), dX.ptr, dP.ptr, dO.ptr)
    out = dO.get()
    for b in (dX, dP, dO):
        b.free()
    return out


def save(P, path=PARAM_FILE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(P, dtype=np.float32))


def load(path=PARAM_FILE):
    return np.load(path, allow_pickle=False).astype(np.float32)


def enable(P, mode=1):
    """Learned mode 1: network predictions; 2: previous rank as prediction
    (the heuristic baseline); 0: plain."""
    L = _L()
    P = np.ascontiguousarray(P, dtype=np.float32)
    L.paratt_amd_learned_set_params(P.ctypes.data_as(C.POINTER(C.c_float)), P.size)
    _params_loaded[0] = True
    L.paratt_amd_learned_mode(mode)


def disable():
    _L().paratt_amd_learned_mode(0)


def counts():
    c = (C.c_long * 3)()
    _L().paratt_amd_learned_counts(c)
    return {"roundings_with_prediction": c[0], "sketch_attempted": c[1], "sketch_accepted": c[2]}


def _main():
    """`python -m paratt_amd.learned [train|ab]`: train the committed model
    (seeds 0-49, two warm-up and two logged RK4 steps each, 5000 training
    rows) or compare plain / learned / previous-rank modes on seed 100."""
    import json
    import sys
    import time

    what = sys.argv[1] if len(sys.argv) > 1 else "ab"
    d, n = 20, 256
    if what == "train":
        X, Y = collect(range(50), steps=2, d=d, n=n, warmup=2)
        idx = np.random.default_rng(0).permutation(len(Y))
        tr, te = idx[:5000], idx[5000:10000]
        Pm, hist = train(X[tr], Y[tr], iters=4000, lr=3e-3, log_every=500)
        err = np.abs(np.rint(predict(Pm, X[te]) * 64) - Y[te] * 64)
        save(Pm)
        print(json.dumps({"samples": int(len(Y)), "train_rows": 5000, "loss": hist,
                          "heldout_within_1": float(np.mean(err <= 1))}))
        return
    Pm = load()
    out = {}
    for mode in (0, 1, 2):
        setup(n)
        lib().paratt_amd_learned_reset_history()
        enable(Pm, mode)
        u = initial_condition(d, n, 100)
        rk4(u, 3)
        c0 = counts()
        ts = []
        for _ in range(8):
            t0 = time.perf_counter()
            rk4(u, 1)
            ts.append(time.perf_counter() - t0)
        c1 = counts()
        out[("plain", "learned", "previous-rank")[mode]] = {
            "ms_per_step": round(1e3 * float(np.median(ts)), 2), "ranks": list(u.r),
            "sketch_attempted": c1["sketch_attempted"] - c0["sketch_attempted"],
            "sketch_accepted": c1["sketch_accepted"] - c0["sketch_accepted"]}
        u.free()
    disable()
    print(json.dumps(out))


if __name__ == "__main__":
    _main()
