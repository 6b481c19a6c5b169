"""The surrogate's probe correction on the CPU: the host restatement the training fits
(fe2.correction) equals the oracle's (tests/fe2_oracle.py), and the least-squares fit
(fe2_train.fit_correction) recovers planted weights from noise-free samples."""
import numpy as np

import fe2_oracle as fo
from ceos_amd import fe2, fe2_train


def _samples(n_rve, per_rve, seed):
    rng = np.random.default_rng(seed)
    R = rng.normal(scale=1e-2, size=(n_rve, 36))
    U = fe2.random_stretch(rng, n_rve * per_rve, 0.85, 1.15)
    ids = np.repeat(np.arange(n_rve), per_rve)
    return rng, R, U, ids


def test_host_correction_equals_oracle():
    rng, R, U, ids = _samples(5, 4, 1)
    W = rng.normal(size=(6, 37, 21))
    dS = fe2.correction(W, R[ids], U)
    for k in range(len(U)):
        ref, _ = fo.probe_correction(W, R[ids[k]], U[k], fe2.PROBE_EP)
        assert np.allclose(dS[k], ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_fit_recovers_planted_weights():
    rng, R, U, ids = _samples(64, 40, 2)
    W = rng.normal(size=(6, 37, 21))
    S_nn = rng.normal(size=(len(U), 6))
    S = S_nn + fe2.correction(W, R[ids], U)
    Wfit = fe2_train.fit_correction({"R": R, "U": U, "ids": ids, "S": S, "S_nn": S_nn}, alpha=1e-14)
    assert np.abs(Wfit - W).max() < 1e-6 * np.abs(W).max()
