"""Pins of the CPU network checker (oracle/mlp_ref.py): BF16 rounding against
torch's float -> bfloat16 conversion, backward() against central finite
differences of the unrounded FP64 network (a wrong index, transpose or
dropped term in any layer fails), the FP32-emulating forward against the
FP64 one, a known answer, and the first Adam step's closed form."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def rand_params(rng, scale=1.0):
    p = np.zeros(M.NP)
    p[M.OW1:M.OB1] = rng.normal(0, scale * np.sqrt(2.0 / 72), M.NH * M.NI)
    p[M.OB1:M.OW2] = rng.normal(0, 0.1, M.NH)
    p[M.OW2:M.OB2] = rng.normal(0, scale * np.sqrt(2.0 / 128), M.NH * M.NH)
    p[M.OB2:M.OW3] = rng.normal(0, 0.1, M.NH)
    p[M.OW3:M.OB3] = rng.normal(0, np.sqrt(2.0 / 65), M.NH)
    p[M.OB3] = 0.3
    return p


def norm_for(X, t):
    return X.mean(0), X.std(0) + 1e-3, float(t.mean()), float(t.std() + 1e-3)


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    a = np.concatenate([
        rng.normal(0, 10, 5000).astype(np.float32),
        # exact ties: 1 + k/256 + 1/512 (halfway between two BF16 values)
        (1.0 + np.arange(64) / 128.0 + 1.0 / 256.0).astype(np.float32),
        np.array([0.0, -0.0, 1e-38, 3.4e38, -2.5], np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(M.bf16(a), want)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (40, M.NI))
    t = rng.normal(0, 1, 40)
    p = rand_params(rng)
    norm = norm_for(X, t)
    L, g = M.loss_and_grad(p, norm, X, t, emulate=False, bf16_round=False)
    # indices from every block, including block edges
    idx = np.unique(np.concatenate([
        [M.OW1, M.OB1 - 1, M.OB1, M.OW2 - 1, M.OW2, M.OB2 - 1, M.OB2,
         M.OW3 - 1, M.OW3, M.OB3 - 1, M.OB3],
        rng.integers(0, M.NP, 60)]))
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        Lp, _ = M.loss_and_grad(pp, norm, X, t, emulate=False, bf16_round=False)
        Lm, _ = M.loss_and_grad(pm, norm, X, t, emulate=False, bf16_round=False)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g).max()), (k, fd, g[k])


def test_emulated_forward_close_to_fp64():
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (300, M.NI))
    t = rng.normal(0, 1, 300)
    p = rand_params(rng).astype(np.float32)
    norm = norm_for(X, t)
    ye, _ = M.forward(p, norm, X, emulate=True)
    yd, _ = M.forward(p, norm, X, emulate=False)
    err = np.abs(ye - yd)
    assert np.median(err) < 1e-5
    assert err.max() < 5e-2 * norm[3]


def test_known_answer_bias_only():
    p = np.zeros(M.NP, np.float32)
    p[M.OB3] = 1.2345                   # bf16 -> 1.234375
    X = np.random.default_rng(4).normal(0, 1, (7, M.NI))
    norm = (np.zeros(8), np.ones(8), 0.5, 2.0)
    y, _ = M.forward(p, norm, X)
    assert np.all(y == np.float32(1.234375) * np.float32(2.0) + np.float32(0.5))


def test_first_adam_step_closed_form():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1, 100)
    g = rng.normal(0, 1, 100)
    p1, _, _ = M.adam_step(p, g, np.zeros(100), np.zeros(100), 1, 1e-3)
    # m_hat = g, v_hat = g^2: each parameter moves by lr * sign(g).
    assert np.allclose(p1 - p, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12,
                       atol=1e-15)
