"""Learned neighbour radius: Python front end of the gpunb_mlp_* entry points
(an MLP 8 -> 64 -> 64 -> 1 with BF16 weights and activations and FP32
accumulation, run and trained by the library's HIP kernels), the training
pairs logged by the host driver, and the train / evaluate pipeline.

Pipeline (``experiment``):
  1. a plain run (no network) of one cluster realisation logs (features,
     target) pairs with ``hermite_ac --log-train``;
  2. the network is trained on the GPU (Adam, minibatches) on those pairs;
  3. an independent realisation is run with the plain radius rules and with
     the learned radius (``--rs mlp``; the Ahmad-Cohen rule is the fallback
     for a body whose learned radius gave a list outside [NNBMIN, NNBMAX] or
     an overflow), and the runs are compared: neighbour-count deviation from
     NNBOPT, list overflows, fallbacks, pair interactions per second and
     |dE/E|.

Feature order (hermite_ac.cpp): log10 local density, NNB/NNBOPT, log10 |a|,
log10 |jerk|, log10 r, log10 m/<m>, log2(1/dtr), log10 RS; target log10 of the
RS giving NNBOPT neighbours at the next regular step.
"""
from __future__ import annotations

import ctypes
import json
import os
import tempfile
import time

import numpy as np

from . import gpunb

NI, NH = 8, 64
NP = NH * NI + NH + NH * NH + NH + NH + 1
REC = NI + 2                       # logged record: features, target, censored
_fp = ctypes.POINTER(ctypes.c_float)
_declared = False


def _lib():
    global _declared
    lib = gpunb.load()            # binds the gpunb_mlp_* signatures
    if not _declared:
        if lib.gpunb_mlp_nparams() != NP:
            raise gpunb.GpunbError("gpunb_mlp_nparams mismatch")
        _declared = True
    return lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(shape) if shape is not None else a


def _pf(a):
    return a.ctypes.data_as(_fp)


def _rc(rc):
    if rc != 0:
        gpunb._check()


def init(seed: int = 1) -> None:
    _rc(_lib().gpunb_mlp_init(int(seed)))


def set_norm(xmean, xstd, ymean: float, ystd: float) -> None:
    xm, xs = _f32(xmean, (NI,)), _f32(xstd, (NI,))
    _rc(_lib().gpunb_mlp_set_norm(_pf(xm), _pf(xs), float(ymean), float(ystd)))


def get_params() -> np.ndarray:
    p = np.empty(NP, np.float32)
    _rc(_lib().gpunb_mlp_get_params(_pf(p)))
    return p


def set_params(p) -> None:
    p = _f32(p, (NP,))
    _rc(_lib().gpunb_mlp_set_params(_pf(p)))


def predict(X) -> np.ndarray:
    X = _f32(X).reshape(-1, NI)
    y = np.empty(X.shape[0], np.float32)
    if X.shape[0]:
        _rc(_lib().gpunb_mlp_predict(X.shape[0], _pf(X), _pf(y)))
    return y


def grad(X, t):
    X = _f32(X).reshape(-1, NI)
    t = _f32(t, (X.shape[0],))
    g = np.empty(NP, np.float32)
    loss = ctypes.c_double()
    _rc(_lib().gpunb_mlp_grad(X.shape[0], _pf(X), _pf(t), _pf(g),
                              ctypes.byref(loss)))
    return loss.value, g


def train(X, t, epochs: int, batch: int = 1024, lr: float = 1e-3,
          seed: int = 1) -> float:
    X = _f32(X).reshape(-1, NI)
    t = _f32(t, (X.shape[0],))
    loss = ctypes.c_double()
    _rc(_lib().gpunb_mlp_train(X.shape[0], _pf(X), _pf(t), int(epochs),
                               int(batch), float(lr), int(seed),
                               ctypes.byref(loss)))
    return loss.value


def read_log(path):
    """(features (n, 8), target (n,), censored (n,)) from a training log."""
    a = np.fromfile(path, dtype=np.float32)
    if a.size % REC:
        raise ValueError(f"{path}: size is not a multiple of {REC} floats")
    a = a.reshape(-1, REC)
    return a[:, :NI].copy(), a[:, NI].copy(), a[:, NI + 1] > 0.5


def fit(F, t, epochs=30, batch=1024, lr=2e-3, seed=1, val_frac=0.1,
        log=None):
    """Normalise, train on the GPU, report the validation error.  Returns a
    dict with the norm, the parameters and the validation statistics (RMS and
    fraction within 10 % of the target radius)."""
    rng = np.random.default_rng(seed)
    n = F.shape[0]
    perm = rng.permutation(n)
    nv = max(int(val_frac * n), 1)
    va, tr = perm[:nv], perm[nv:]
    xm = F[tr].mean(0)
    xs = np.maximum(F[tr].std(0), 1e-3)
    ym, ys = float(t[tr].mean()), float(max(t[tr].std(), 1e-3))
    init(seed)
    set_norm(xm, xs, ym, ys)
    t0 = time.perf_counter()
    losses = []
    for e in range(epochs):
        losses.append(train(F[tr], t[tr], 1, batch, lr, seed + e))
        if log:
            log(f"epoch {e}: train loss {losses[-1]:.5f}")
    t_train = time.perf_counter() - t0
    yv = predict(F[va]).astype(np.float64)
    err = yv - t[va]
    base = F[va, 7] + np.log10(np.cbrt(1.0 / np.maximum(F[va, 1], 1e-3)))
    return {
        "norm": (xm, xs, ym, ys), "params": get_params(), "n_train": int(tr.size),
        "n_val": int(nv), "epochs": epochs, "train_seconds": t_train,
        "loss_first": losses[0], "loss_last": losses[-1],
        "val_rms_dex": float(np.sqrt(np.mean(err ** 2))),
        "val_within_10pct": float(np.mean(np.abs(err) < np.log10(1.1))),
        # The unclamped volume rule RS (NNBOPT/NNB)^(1/3) as a reference point.
        "volume_rule_rms_dex": float(np.sqrt(np.mean((base - t[va]) ** 2))),
    }


def save_weights(path, params, norm) -> None:
    xm, xs, ym, ys = norm
    np.concatenate([_f32(params, (NP,)), _f32(xm, (NI,)), _f32(xs, (NI,)),
                    _f32([ym, ys])]).astype(np.float32).tofile(path)


def summarise(recs):
    """Run summary from the driver's last energy-check record."""
    last = recs[-1]
    pairs = last["reg_pairs"] + last["irr_pairs"]
    return {
        "dE_E": max(abs(r["dE_E"]) for r in recs),
        "overflows": last["overflows"],
        "overflow_rate": last["overflows"] / max(last["reg_steps"], 1),
        "nnb_dev": last["nnb_dev"], "nnb_mean": last["nnb_mean"],
        "learned_steps": last["learned_steps"],
        "fallbacks": last["fallbacks"],
        "reg_steps": last["reg_steps"], "irr_steps": last["irr_steps"],
        "reg_pairs": last["reg_pairs"], "irr_pairs": last["irr_pairs"],
        "wall": last["wall"], "interactions_per_s": pairs / last["wall"],
        "t_mlp": last["t_mlp"],
    }


def experiment(n=16384, t_log=0.25, t_run=0.25, seed_train=31, seed_eval=37,
               log_max=2_000_000, epochs=20, nnbopt=48, etai=0.05, etar=0.1,
               timeout=1200, workdir=None, log=None, nnbmax=0):
    """Train on one segregated Plummer realisation, then run another with
    three radius rules: ``ac`` (the reference's Ahmad-Cohen rule, which lowers
    the membership outside the half-mass radius), ``simple`` (volume scaling
    towards NNBOPT, the same target as the network's, without learning) and
    ``learned``.  The pairs are logged from a ``simple`` run, whose lists
    straddle NNBOPT, so about half of the targets are exact rather than
    censored.  ``nnbmax`` > 0 caps the lists of the three compared runs
    (the logging run keeps the default), to measure overflow rates.  Files go
    to ``workdir`` (a temporary directory, removed afterwards, when None).
    Returns a dict (fit statistics, the three runs)."""
    if workdir is None:
        with tempfile.TemporaryDirectory(prefix="rsnet_") as td:
            return experiment(n, t_log, t_run, seed_train, seed_eval, log_max,
                              epochs, nnbopt, etai, etar, timeout, td, log,
                              nnbmax)
    from . import ics, integrate
    # Energy checks every dtadj (a power of two no longer than either run).
    dtadj = 0.125
    while dtadj > min(t_log, t_run):
        dtadj *= 0.5
    td = workdir
    lp = os.path.join(td, "pairs.bin")
    wp = os.path.join(td, "rsnet.bin")
    ct = ics.plummer_segregated(n, seed=seed_train)
    t0 = time.perf_counter()
    integrate.run(ct.m, ct.x, ct.v, t_end=t_log, etai=etai, etar=etar,
                  dtadj=dtadj, nnbopt=nnbopt, timeout=timeout,
                  extra=("--rs", "simple", "--log-train", lp, "--log-max",
                         str(log_max)), progress=log)
    t_logrun = time.perf_counter() - t0
    F, t, cens = read_log(lp)
    if log:
        log(f"logged {F.shape[0]} pairs ({cens.mean():.3f} censored) in "
            f"{t_logrun:.1f} s")
    st = fit(F, t, epochs=epochs, log=log)
    save_weights(wp, st["params"], st["norm"])
    ce = ics.plummer_segregated(n, seed=seed_eval)
    runs = {}
    for name, extra in (("ac", ("--rs", "ac")), ("simple", ("--rs", "simple")),
                        ("learned", ("--rs", "mlp", "--mlp", wp))):
        if nnbmax > 0:
            extra = extra + ("--nnbmax", str(nnbmax))
        recs = integrate.run(ce.m, ce.x, ce.v, t_end=t_run, etai=etai,
                             etar=etar, dtadj=dtadj, nnbopt=nnbopt,
                             timeout=timeout, extra=extra, progress=log)
        runs[name] = summarise(recs)
        if log:
            log(f"{name}: {runs[name]}")
    fit_stats = {k: v for k, v in st.items() if k not in ("params", "norm")}
    fit_stats["pairs_logged"] = int(F.shape[0])
    fit_stats["censored_frac"] = float(cens.mean())
    fit_stats["log_run_seconds"] = t_logrun
    return {"n": n, "nnbmax": nnbmax or None, "fit": fit_stats, "runs": runs,
            "workdir": td}


def kernel_bench(n_train=2_000_000, n_pred=1_048_576, epochs=3, batch=1024,
                 seed=5):
    """Seconds per training epoch and per prediction pass of the network
    kernels on synthetic features (host arrays, as the driver passes them)."""
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n_train, NI)).astype(np.float32)
    t = (0.3 * X[:, 7] - 0.1 * X[:, 0] + 0.05 * np.sin(X[:, 2])).astype(np.float32)
    init(seed)
    set_norm(X.mean(0), X.std(0), float(t.mean()), float(t.std()))
    train(X[:batch], t[:batch], 1, batch)              # warm-up
    t0 = time.perf_counter()
    loss = train(X, t, epochs, batch, 2e-3, seed)
    t_train = (time.perf_counter() - t0) / epochs
    Xp = X[:n_pred]
    predict(Xp[:1024])
    t0 = time.perf_counter()
    predict(Xp)
    t_pred = time.perf_counter() - t0
    return {"train_pairs": n_train, "batch": batch, "epochs": epochs,
            "seconds_per_epoch": t_train, "loss": loss,
            "predictions": n_pred, "seconds_per_prediction_pass": t_pred,
            "flops_per_forward": 2 * (NI * NH + NH * NH + NH)}


def main(argv=None):
    """python -m nbody6ppgpu_amd.rsnet [--n N] [--t T] [--epochs E]: the
    experiment on a mass-segregated Plummer cluster with a Kroupa IMF; prints
    progress lines, then one JSON object (fit statistics and the three runs)."""
    import argparse
    import sys
    ap = argparse.ArgumentParser(prog="python -m nbody6ppgpu_amd.rsnet")
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--t", type=float, default=0.25,
                    help="N-body time units of the logging and compared runs")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--log-max", type=int, default=2_000_000)
    ap.add_argument("--nnbmax", type=int, default=0,
                    help="list cap of the compared runs (0: the driver default)")
    ap.add_argument("--kernel-bench", action="store_true",
                    help="time the network kernels alone on synthetic pairs "
                         "(2M training pairs, 3 epochs; 1M predictions)")
    a = ap.parse_args(argv)
    if a.kernel_bench:
        print(json.dumps(kernel_bench()), flush=True)
        return
    t0 = time.perf_counter()
    out = experiment(n=a.n, t_log=a.t, t_run=a.t, epochs=a.epochs,
                     log_max=a.log_max, nnbmax=a.nnbmax,
                     log=lambda s: print(f"[{time.perf_counter() - t0:7.1f}s] {s}",
                                         file=sys.stderr, flush=True))
    out.pop("workdir")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
