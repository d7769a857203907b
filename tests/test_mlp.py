"""Surrogate MLP: the FP64 CPU model (oracle/mlp_oracle.py) against finite differences,
and the HIP MFMA kernels (through the C ABI) against the CPU model: forward outputs,
gradients (BF16 arithmetic model: tight; plain FP64: loose), and training on flight data."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402
import oracle_lib as O  # noqa: E402


def _params(seed, scale=0.15):
    sl, n = M.param_slices()
    rng = np.random.default_rng(seed)
    p = (rng.normal(0, scale, n)).astype(np.float32)
    (a, b), _ = sl[0]
    W1 = p[a:b].reshape(M.K0, M.HID)
    W1[M.NIN:] = 0
    (a, b), (c, d) = sl[3]
    W4 = p[a:b].reshape(M.HID, M.NOP)
    W4[:, M.NOUT:] = 0
    p[c + M.NOUT:d] = 0
    return p


def _data(n, seed):
    lo, hi = O.bounds()
    rng = np.random.default_rng(seed)
    g = (lo + rng.random((n, 35), dtype=np.float32) * (hi - lo)).astype(np.float32)
    t = np.stack([rng.normal(5, 3, n), rng.normal(2, 3, n), rng.normal(0, 4, n), (rng.random(n) < 0.3)], 1)
    return g, t.astype(np.float32), lo, hi


def test_fp64_gradient_matches_finite_differences():
    g, t, lo, hi = _data(24, 1)
    p = _params(2).astype(np.float64)
    loss, grad = M.loss_grad(p, g, t, lo, hi, bf=False)
    rng = np.random.default_rng(3)
    sl, n = M.param_slices()
    live = np.ones(n, bool)
    (a, b), _ = sl[0]
    live[a:b] = np.tile(np.arange(M.K0) < M.NIN, (M.HID, 1)).T.ravel()
    (a, b), (c, d) = sl[3]
    live[a:b] = np.tile(np.arange(M.NOP) < M.NOUT, M.HID)
    live[c:d] = np.arange(M.NOP) < M.NOUT
    idx = rng.choice(np.nonzero(live)[0], 60, replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy(); q[i] += h
        lp, _ = M.loss_grad(q, g, t, lo, hi, bf=False)
        q[i] -= 2 * h
        lm, _ = M.loss_grad(q, g, t, lo, hi, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])


@pytest.mark.gpu
def test_mlp_forward_matches_cpu_model():
    import planes_amd as PA
    g, t, lo, hi = _data(200, 4)
    p = _params(5)
    S = PA.Surrogate(max_batch=256)
    S.set_params(p)
    pred = S.predict(g)
    out, _ = M.forward(p, g, lo, hi, bf=True)
    ref = np.concatenate([10 * out[:, :3], 1 / (1 + np.exp(-out[:, 3:4]))], 1)
    err = np.abs(pred[:, :4] - ref) / (np.abs(ref) + 0.05)
    assert err.max() < 2e-3, err.max()


@pytest.mark.gpu
def test_mlp_prediction_in_chunks_matches_cpu_model():
    """Prediction runs in one fused pass of 128-row blocks: rows at the start, across block
    boundaries and in the last, partial block (70000 = 546 x 128 + 112) of a 70000-row call
    equal the CPU model's (BF16 mimic), and a row's prediction does not depend on the call it
    is part of."""
    import planes_amd as PA
    g, t, lo, hi = _data(70000, 8)
    p = _params(9)
    S = PA.Surrogate(max_batch=512)
    S.set_params(p)
    pred = S.predict(g)
    rows = np.r_[0:64, 65500:65600, 69936:70000]
    out, _ = M.forward(p, g[rows], lo, hi, bf=True)
    ref = np.concatenate([10 * out[:, :3], 1 / (1 + np.exp(-out[:, 3:4]))], 1)
    err = np.abs(pred[rows, :4] - ref) / (np.abs(ref) + 0.05)
    assert err.max() < 2e-3, err.max()
    assert np.array_equal(S.predict(g[65500:65600]), pred[65500:65600])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [512, 300])
def test_mlp_gradients_vs_cpu_model(n):
    """Gradients of the fused training pass (forward, loss, input gradients in 32-row blocks;
    weight gradients over the batch) against the CPU model; n = 300 leaves 212 zero rows in
    the 512-row padded batch, which must contribute nothing."""
    import planes_amd as PA
    g, t, lo, hi = _data(n, 6)
    p = _params(7)
    S = PA.Surrogate(max_batch=1024)
    S.set_params(p)
    loss, grad = S.loss_grad(g, t)
    lb, gb = M.loss_grad(p.astype(np.float64), g, t, lo, hi, bf=True)
    l64, g64 = M.loss_grad(p.astype(np.float64), g, t, lo, hi, bf=False)
    sl, _ = M.param_slices()
    for L, ((a, b), (c, d)) in enumerate(sl):
        for lo_, hi_ in ((a, b), (c, d)):
            x, y, z = grad[lo_:hi_], gb[lo_:hi_], g64[lo_:hi_]
            rel_b = np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-12)
            rel_64 = np.linalg.norm(x - z) / max(np.linalg.norm(z), 1e-12)
            cos = float(x @ z / max(np.linalg.norm(x) * np.linalg.norm(z), 1e-30))
            print(f"layer {L} [{lo_}:{hi_}] rel err vs BF16 model {rel_b:.2e}, vs FP64 {rel_64:.2e}, cos {cos:.5f}")
            assert rel_b < 5e-3 and rel_64 < 5e-2 and cos > 0.998
    assert abs(loss - lb) <= 1e-3 * abs(lb) and abs(loss - l64) <= 2e-2 * abs(l64)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1024, 480])
def test_mlp_adam_step_matches_gradient(n):
    """One training batch (pr_mlp_train: the weight-gradient kernel takes each parameter's
    Adam step itself) equals an FP64 Adam step applied to the gradient pr_mlp_loss_grad
    returns for the same batch; the padding parameters (W1 rows >= 35, W4 and b4 columns
    >= 4) stay zero, and predictions afterwards use the stepped weights."""
    import planes_amd as PA
    g, t, lo, hi = _data(n, 12)
    p = _params(13)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    S = PA.Surrogate(max_batch=1024)
    S.set_params(p)
    _, grad = S.loss_grad(g, t)
    S.train(g, t, epochs=1, batch=n, lr=lr, shuffle_seed=None)
    after = S.params()
    gd = grad.astype(np.float64)
    m = (1 - b1) * gd
    v = (1 - b2) * gd * gd
    want = p.astype(np.float64) - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    sl, _ = M.param_slices()
    pad = np.zeros(len(p), bool)
    (a, b), _ = sl[0]
    pad[a:b] = (np.arange(b - a) // M.HID) >= M.NIN
    (a, b), (c, d) = sl[3]
    pad[a:b] = (np.arange(b - a) % M.NOP) >= M.NOUT
    pad[c:d] = np.arange(d - c) >= M.NOUT
    assert np.all(after[pad] == 0.0)
    live = ~pad
    # the first Adam step moves each parameter by ~lr sign(g); FP32 against FP64 rounding
    step_gpu = after[live].astype(np.float64) - p[live]
    step_ref = want[live] - p[live]
    assert np.max(np.abs(step_gpu - step_ref)) <= 1e-6 + 1e-3 * lr, np.max(np.abs(step_gpu - step_ref))
    assert np.mean(np.abs(step_ref) > 0.5 * lr) > 0.7  # most parameters moved (dead ReLU units: zero gradient)
    S2 = PA.Surrogate(max_batch=1024)
    S2.set_params(after)
    assert np.array_equal(S.predict(g[:64]), S2.predict(g[:64]))


@pytest.mark.gpu
def test_mlp_learns_flight_fitness():
    """Trained on 4096 genomes x 8 flights, the surrogate ranks 2048 held-out printable
    genomes well above chance (the same recipe reaches ~0.6 with the FP64 CPU model) and
    identifies unprintable designs."""
    import planes_amd as PA
    n_tr, n_te, k = 4096, 2048, 8
    lo, hi = PA.gene_bounds()
    rng = np.random.default_rng(11)
    g = (lo + rng.random((n_tr + n_te, 35), dtype=np.float32) * (hi - lo)).astype(np.float32)
    pl = PA.Planes(n_tr + n_te, k, (n_tr + n_te) * k)
    fit, _ = pl.evaluate(g, np.arange(1, k + 1, dtype=np.int32), PA.SimParams.reference(max_steps=4800))
    tg = M.fitness_targets(fit)
    S = PA.Surrogate(max_batch=4096, seed=3)
    l0, _ = S.loss_grad(g[:4096], tg[:4096])
    S.train(g[:n_tr], tg[:n_tr], epochs=20, batch=256, lr=2e-3)
    l1, _ = S.loss_grad(g[:4096], tg[:4096])
    pred = S.predict(g[n_tr:])
    reg = 0.65 * pred[:, 0] + 0.25 * pred[:, 1] + 0.10 * pred[:, 2]
    true_score = O.robust_scores(fit[n_tr:], n_te, k)
    ok = tg[n_tr:, 3] < 0.5
    ra = np.argsort(np.argsort(reg[ok])); rb = np.argsort(np.argsort(true_score[ok]))
    rho = np.corrcoef(ra, rb)[0, 1]
    acc = np.mean((pred[:, 3] > 0.5) == (tg[n_tr:, 3] > 0.5))
    mae = np.mean(np.abs(reg[ok] - true_score[ok]))
    base = max(np.mean(~ok), np.mean(ok))  # accuracy of always guessing the majority class
    print(f"surrogate: loss {l0:.3f} -> {l1:.3f}, held-out Spearman {rho:.3f}, oversize accuracy {acc:.3f} "
          f"(majority class {base:.3f}), MAE {mae:.2f} m")
    assert l1 < 0.5 * l0 and rho > 0.45 and acc > 0.8 and acc > base + 0.1
