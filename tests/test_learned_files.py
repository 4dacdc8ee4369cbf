"""The committed rank-predictor parameters and held-out spectra: layout and
sanity (CPU)."""
import os

import numpy as np

import mlp_ref as M

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_parameter_file_layout():
    p = np.load(os.path.join(ROOT, "paratt_amd", "data", "rank_predictor_d20_n256_eps1e-8.npy"),
                allow_pickle=False)
    assert p.dtype == np.float32 and p.shape == (M.NPARAM,)
    assert np.all(np.isfinite(p))


def test_heldout_predictions_with_fp64_network():
    """The FP64 restatement of the network reproduces the committed model's
    accuracy on the held-out spectra (no GPU)."""
    p = np.load(os.path.join(ROOT, "paratt_amd", "data", "rank_predictor_d20_n256_eps1e-8.npy"),
                allow_pickle=False).astype(np.float64)
    f = np.load(os.path.join(HERE, "golden", "rank_predictor_heldout.npz"), allow_pickle=False)
    X, Y = f["X"].astype(np.float64), f["Y"].astype(np.float64)
    assert X.shape[1] == 18
    out, _ = M.forward(p, X)
    err = np.abs(np.rint(out * 64) - Y * 64)
    assert np.mean(err <= 1) >= 0.98
