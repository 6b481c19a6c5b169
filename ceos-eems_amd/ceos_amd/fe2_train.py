"""Train the FE^2 surrogate on the GPU from batched RVE solves.

    python -m ceos_amd.fe2_train --out surrogate.npz [--data-out data.npz | --load-data data.npz]

1. Training microstructures: n_rve seeded RVEs (sides drawn from {8, 12, 16}, 20 % spherical
   inclusions, Neo-Hookean phases; seed SEED + 1, so the macro workload's RVEs are unseen).
2. Descriptors: C_lin of every RVE (condensed tangent at F = I, ceos_rve_solve with tangent).
3. Samples: every RVE walks `steps` stretches U (principal stretches reflected in [0.8, 1.25],
   principal axes turned by small random rotations; small steps keep each solve at a few Newton
   iterations); each step is one batched solve from the previous converged state, and every
   converged state is a sample (U, C_lin, S = U^-1 P).
   Hyperelastic RVEs are path independent, so walk samples are samples of the map U -> S.
4. Targets: t = (S - S_base) e_ref^2 / |E|^2, normalized, with S_base the Neo-Hookean base of
   fe2.base_stress; inputs from the library's own feature kernel (ceos_surrogate_features_device).
5. Training: Adam on the library's BF16-MFMA training kernels (ceos_mlp_train_step_weighted),
   cosine learning-rate decay; per-sample weights turn the loss into the squared relative error
   of the predicted stress (near-isochoric states carry stresses of shear-modulus size, far
   below the bulk-dominated ones, and would otherwise be fitted only in absolute terms).
6. Test: unseen microstructures (seed SEED + 2) at stretches drawn independently from the
   training distribution (solved from the undeformed state), evaluated through the inference
   path (ceos_surrogate_eval_device); the error of the base alone (the method without the
   network) is reported beside it.

The result file holds the parameters, the normalization and hull constants and the test errors
(numpy .npz, no pickled objects).
"""
import argparse
import math
import time

import numpy as np

from . import fe2, gpu, hiprt, rve

SEED_TRAIN = rve.SEED + 1
SEED_TEST = rve.SEED + 2


def log(*a):
    print(*a, flush=True)


def make_batch(n_rve, seed):
    sides, ph = rve.batch(n_rve, seed)
    b = gpu.RveBatch(sides, ph, gpu.MODEL_NH, fe2.PROPS)
    b.set_precision(gpu.K_FP32_STORED)
    b.set_tangent_rtol(1e-8)
    return b, sides


def S_from(U, P):
    S = np.linalg.solve(U, P)
    return fe2.to_voigt(0.5 * (S + S.transpose(0, 2, 1)))


def generate(n_rve, steps, seed, rng, sig_lam, sig_rot, lo=fe2.LAM_LO, hi=fe2.LAM_HI, sig_logj=None):
    b, sides = make_batch(n_rve, seed)
    t = time.time()
    Clin = fe2.descriptors(b, pcg_rtol=1e-8)
    log("descriptors of %d RVEs: %.1f s" % (n_rve, time.time() - t))
    U = fe2.random_stretch(rng, n_rve, lo, hi)
    if sig_logj is not None:
        lam, Q = np.linalg.eigh(U)
        U = np.einsum("nij,nj,nkj->nik", Q, fe2.isochoric(rng, lam, sig_logj, lo, hi), Q)
    Us, ids, Ss = [], [], []
    for k in range(steps):
        if k:
            U = fe2.walk_stretch(rng, U, sig_lam, sig_rot, lo, hi, sig_logj)
        t = time.time()
        P, _, info = b.solve(U, check=False)
        ok = info[:, 0] == 0
        Us.append(U[ok])
        ids.append(np.nonzero(ok)[0])
        Ss.append(S_from(U[ok], P[ok]))
        log("step %d/%d: %.1f s, converged %d/%d, newton %.1f, cg %.0f" % (
            k + 1, steps, time.time() - t, ok.sum(), n_rve, info[:, 1].mean(), info[:, 2].mean()))
    return Clin, sides, np.concatenate(Us), np.concatenate(ids), np.concatenate(Ss)


def test_set(n_rve, seed, rng, lo=fe2.LAM_LO, hi=fe2.LAM_HI, sig_logj=None):
    b, sides = make_batch(n_rve, seed)
    Clin = fe2.descriptors(b, pcg_rtol=1e-8)
    U = fe2.random_stretch(rng, n_rve, lo, hi)
    if sig_logj is not None:
        lam, Q = np.linalg.eigh(U)
        U = np.einsum("nij,nj,nkj->nik", Q, fe2.isochoric(rng, lam, sig_logj, lo, hi), Q)
    t = time.time()
    P, _, info = b.solve(U, check=False)
    ok = info[:, 0] == 0
    log("test set: %d RVEs solved from the undeformed state in %.1f s (%d converged)" % (n_rve, time.time() - t,
                                                                                        ok.sum()))
    return Clin[ok], sides[ok], U[ok], S_from(U[ok], P[ok])


def targets(Clin_pts, U, S):
    E = fe2.green_lagrange(U)
    e2 = np.sum(E * E, axis=(1, 2))
    t = (S - fe2.base_stress(Clin_pts, U)) * (fe2.E_REF ** 2 / np.maximum(e2, 1e-300))[:, None]
    return t, e2


def base_cfg(Clin_train, margin=0.25):
    d_mean = Clin_train.mean(0)
    d_std = Clin_train.std(0)
    z = (Clin_train - d_mean) / d_std
    lo, hi = z.min(0), z.max(0)
    w = hi - lo
    return {"u_scale": fe2.U_SCALE, "lam_lo": fe2.LAM_LO, "lam_hi": fe2.LAM_HI, "e_ref": fe2.E_REF,
            "d_mean": d_mean, "d_std": d_std, "d_lo": lo - margin * w, "d_hi": hi + margin * w,
            "t_mean": np.zeros(6), "t_std": np.ones(6)}


def features(mlp, cfg, U, Clin_pts, chunk):
    """Network inputs through the library's feature kernel (F = U)."""
    sur = gpu.Surrogate(mlp, cfg, chunk)
    n = len(U)
    X = np.zeros((n, fe2.N_IN), dtype=np.float32)
    dF = hiprt.DeviceArray(9 * chunk)
    dC = hiprt.DeviceArray(21 * chunk)
    dX = hiprt.DeviceArray(fe2.N_IN * chunk, np.float32)
    for a in range(0, n, chunk):
        m = min(chunk, n - a)
        dF.view(0, 9 * m).upload(np.ascontiguousarray(U[a:a + m].transpose(0, 2, 1).reshape(-1)))
        dC.view(0, 21 * m).upload(Clin_pts[a:a + m].reshape(-1))
        sur.features_device(m, dF.ptr, dC.ptr, dX.ptr)
        X[a:a + m] = dX.view(0, fe2.N_IN * m).numpy().reshape(m, fe2.N_IN)
    sur.close()
    return X


def evaluate(mlp, cfg, U, Clin_pts, chunk):
    """Surrogate S (Voigt) and hull flags through ceos_surrogate_eval_device."""
    sur = gpu.Surrogate(mlp, cfg, chunk)
    n = len(U)
    S = np.zeros((n, 6))
    flags = np.zeros(n, dtype=np.int32)
    dF = hiprt.DeviceArray(9 * chunk)
    dC = hiprt.DeviceArray(21 * chunk)
    dext = hiprt.DeviceArray(27 * chunk)
    dfl = hiprt.DeviceArray(chunk, np.int32)
    for a in range(0, n, chunk):
        m = min(chunk, n - a)
        dF.view(0, 9 * m).upload(np.ascontiguousarray(U[a:a + m].transpose(0, 2, 1).reshape(-1)))
        dC.view(0, 21 * m).upload(Clin_pts[a:a + m].reshape(-1))
        sur.eval_device(m, dF.ptr, dC.ptr, dext.ptr, dfl.ptr)
        S[a:a + m] = dext.view(0, 27 * m).numpy().reshape(m, 27)[:, :6]
        flags[a:a + m] = dfl.view(0, m).numpy()
    sur.close()
    return S, flags


def relative_weights(phi, t_std, S, clip=50.0):
    """Per-sample, per-output loss weights that make the weighted squared error of the
    normalized targets the squared relative error of S (Voigt, shear counted twice):
    |dS|^2 / |S|^2 = sum_o w_o (phi t_std_o dy_o)^2 / |S|^2.  Normalized to mean 1, clipped."""
    vw = np.array([1, 1, 1, 2, 2, 2.0])
    s2 = np.sum(vw * S * S, axis=1)
    W = vw[None, :] * (phi[:, None] * t_std[None, :]) ** 2 / np.maximum(s2, 1e-300)[:, None]
    W /= W.mean()
    return np.minimum(W, clip).astype(np.float32)


def train(X, T, epochs, batch, lr0, seed, W=None, log_every=25, fp32=False, lr_floor=1e-2):
    n = len(X)
    mlp = gpu.Mlp(fe2.N_IN, fe2.WIDTH, fe2.N_OUT, max_batch=batch, seed=seed)
    mlp.set_training_fp32(fp32)
    dX = hiprt.DeviceArray(n * fe2.N_IN, np.float32)
    dT = hiprt.DeviceArray(n * fe2.N_OUT, np.float32)
    dW = hiprt.DeviceArray(n * fe2.N_OUT, np.float32) if W is not None else None
    rng = np.random.default_rng(seed)
    nb = n // batch
    total = epochs * nb
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        perm = rng.permutation(n)
        dX.upload(X[perm])
        dT.upload(T[perm])
        if dW is not None:
            dW.upload(W[perm])
        for b in range(nb):
            lr = lr_floor * lr0 + 0.5 * (lr0 - lr_floor * lr0) * (1.0 + math.cos(math.pi * step / total))
            want = (b == nb - 1) and (ep % log_every == 0 or ep == epochs - 1)
            loss = mlp.train_step(dX.view(b * batch * fe2.N_IN, batch * fe2.N_IN),
                                  dT.view(b * batch * fe2.N_OUT, batch * fe2.N_OUT), batch, lr, want_loss=want,
                                  dW=dW.view(b * batch * fe2.N_OUT, batch * fe2.N_OUT) if dW is not None else None)
            step += 1
            if want:
                log("epoch %d/%d  loss %.3e  lr %.2e  %.1f s" % (ep + 1, epochs, loss, lr, time.time() - t0))
    return mlp, time.time() - t0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rves", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=196)
    ap.add_argument("--sig-lam", type=float, default=0.02)
    ap.add_argument("--sig-rot", type=float, default=0.06)
    ap.add_argument("--test-rves", type=int, default=1024)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", required=True)
    ap.add_argument("--data-out", default=None)
    ap.add_argument("--load-data", action="append", default=None,
                    help="train on saved samples (repeatable: the sets are combined)")
    ap.add_argument("--lam-lo", type=float, default=fe2.LAM_LO, help="stretch range of the generated walks")
    ap.add_argument("--lam-hi", type=float, default=fe2.LAM_HI)
    ap.add_argument("--isochoric", type=float, default=None,
                    help="near-isochoric walks: log J ~ N(0, this) (nearly incompressible composites deform so)")
    ap.add_argument("--seed-offset", type=int, default=0,
                    help="offset of the training / test microstructure seeds (a new set of RVEs)")
    ap.add_argument("--margin", type=float, default=0.25,
                    help="descriptor hull: training range widened by this fraction on each side")
    ap.add_argument("--train-fp32", action="store_true", help="FP32 training GEMMs (reference mode)")
    ap.add_argument("--lr-floor", type=float, default=1e-2, help="final learning rate / initial")
    ap.add_argument("--blend", type=float, default=None,
                    help="loss weights (1 + blend W_rel) / (1 + blend), W_rel the relative-error weights")
    ap.add_argument("--loss", choices=["relative", "plain"], default="plain",
                    help="relative: weights that make the loss the squared relative stress error")
    a = ap.parse_args(argv)
    rng = np.random.default_rng(a.seed)
    t_all = time.time()
    if a.load_data:
        sets = []
        for path in a.load_data:
            with np.load(path, allow_pickle=False) as z:
                sets.append({k: z[k] for k in z.files})
        t_gen = float(sum(float(d["t_gen"]) for d in sets))
        offs = np.cumsum([0] + [len(d["Clin"]) for d in sets])
        d = {"Clin": np.concatenate([x["Clin"] for x in sets]),
             "U": np.concatenate([x["U"] for x in sets]),
             "ids": np.concatenate([x["ids"] + o for x, o in zip(sets, offs)]),
             "S": np.concatenate([x["S"] for x in sets]),
             "Clin_test": np.concatenate([x["Clin_test"] for x in sets]),
             "U_test": np.concatenate([x["U_test"] for x in sets]),
             "S_test": np.concatenate([x["S_test"] for x in sets]),
             "test_set": np.concatenate([np.full(len(x["U_test"]), k) for k, x in enumerate(sets)])}
    else:
        t = time.time()
        Clin, sides, U, ids, S = generate(a.rves, a.steps, SEED_TRAIN + a.seed_offset, rng, a.sig_lam, a.sig_rot,
                                          a.lam_lo, a.lam_hi, a.isochoric)
        Ct, st, Ut, St = test_set(a.test_rves, SEED_TEST + a.seed_offset, rng, a.lam_lo, a.lam_hi, a.isochoric)
        t_gen = time.time() - t
        d = {"Clin": Clin, "sides": sides, "U": U, "ids": ids, "S": S, "Clin_test": Ct, "sides_test": st,
             "U_test": Ut, "S_test": St, "t_gen": t_gen, "lam_range": np.array([a.lam_lo, a.lam_hi])}
        if a.data_out:
            np.savez(a.data_out, **d)
        d["test_set"] = np.zeros(len(Ut), dtype=np.int64)
    Clin, U, ids, S = d["Clin"], d["U"], d["ids"], d["S"]
    Cp = Clin[ids]
    t_raw, e2 = targets(Cp, U, S)
    keep = e2 > 1e-6
    log("samples %d (kept %d), RVEs %d, generation %.1f s" % (len(U), keep.sum(), len(Clin), t_gen))
    cfg = base_cfg(Clin, a.margin)
    cfg["t_mean"] = t_raw[keep].mean(0)
    cfg["t_std"] = t_raw[keep].std(0)
    T = ((t_raw[keep] - cfg["t_mean"]) / cfg["t_std"]).astype(np.float32)
    W = relative_weights(e2[keep] / fe2.E_REF ** 2, cfg["t_std"], S[keep]) if a.loss == "relative" else None
    if a.blend is not None:
        Wr = relative_weights(e2[keep] / fe2.E_REF ** 2, cfg["t_std"], S[keep])
        W = ((1.0 + a.blend * Wr) / (1.0 + a.blend)).astype(np.float32)
    feat_mlp = gpu.Mlp(fe2.N_IN, fe2.WIDTH, fe2.N_OUT, max_batch=a.batch, seed=1)
    X = features(feat_mlp, cfg, U[keep], Cp[keep], a.batch)
    feat_mlp.close()
    mlp, t_train = train(X, T, a.epochs, a.batch, a.lr, a.seed, W, fp32=a.train_fp32, lr_floor=a.lr_floor)
    # held-out microstructures and stretches
    Ct, Ut, St = d["Clin_test"], d["U_test"], d["S_test"]
    S_nn, flags = evaluate(mlp, cfg, Ut, Ct, a.batch)
    err = fe2.stress_rel_error(S_nn, St)
    err_lin = fe2.stress_rel_error(fe2.base_stress(Ct, Ut), St)
    inside = flags == 0
    E = fe2.green_lagrange(Ut)
    norm_err = np.linalg.norm((S_nn - St)[inside] * np.sqrt([1, 1, 1, 2, 2, 2])) / \
        np.linalg.norm(St[inside] * np.sqrt([1, 1, 1, 2, 2, 2]))
    for k in np.unique(d["test_set"]):
        m = inside & (d["test_set"] == k)
        log("test set %d (%d points inside): surrogate mean %.4f p99 %.4f max %.4f; base alone mean %.4f" % (
            k, m.sum(), err[m].mean(), np.percentile(err[m], 99), err[m].max(), err_lin[m].mean()))
    log("test (%d points, %d inside the hull): surrogate rel. error mean %.4f p99 %.4f max %.4f, "
        "aggregate %.4f; base alone (no network) mean %.4f max %.4f" % (
            len(Ut), inside.sum(), err[inside].mean(), np.percentile(err[inside], 99), err[inside].max(), norm_err,
            err_lin[inside].mean(), err_lin[inside].max()))
    np.savez(a.out, params=mlp.params(), width=np.int64(fe2.WIDTH),
             **{k: np.asarray(v, dtype=np.float64) for k, v in cfg.items()},
             n_samples=np.int64(keep.sum()), n_train_rves=np.int64(len(Clin)), seed_train=np.int64(SEED_TRAIN),
             seed_test=np.int64(SEED_TEST), epochs=np.int64(a.epochs), batch=np.int64(a.batch),
             t_generate=np.float64(t_gen), t_train=np.float64(t_train),
             test_points=np.int64(len(Ut)), test_inside=np.int64(inside.sum()),
             test_err_mean=np.float64(err[inside].mean()), test_err_p99=np.float64(np.percentile(err[inside], 99)),
             test_err_max=np.float64(err[inside].max()), test_err_aggregate=np.float64(norm_err),
             test_err_base_mean=np.float64(err_lin[inside].mean()),
             test_err_base_max=np.float64(err_lin[inside].max()),
             test_max_strain=np.float64(np.sqrt(np.sum(E * E, axis=(1, 2))).max()))
    log("saved %s (%.1f s total)" % (a.out, time.time() - t_all))


if __name__ == "__main__":
    main()
