"""Candidate-pool GA (pr_pool_*, the surrogate pre-screening workload) through the C ABI.

The pool GA has no reference counterpart (the reference's GA breeds on the host; its
operators are pinned in test_gpu_parity.py and test_host.py).  Pinned here:
  - the mutation generator: parents and Box-Muller normals restated from their integer
    hashes (parents exact, genes within float libm rounding), the normals N(0, 1) and the
    parents uniform over the elite by their distributions;
  - the elite update: the fly set's simulated scores equal an independent evaluation of
    the same genomes under the generation's seeds (bit for bit), and the new elite is the
    fly set's best n_elite by that score;
  - surrogate pre-screening: the fly set is the elite, then the exploration sample (hashed
    per candidate), then the best-predicted candidates, in the order of a stable sort of
    those keys; the reported rank correlation, mean absolute error (members both call
    flyable) and oversize-flag accuracy are those over the exploration sample;
  - the replay-buffer targets equal the CPU model's fitness statistics of the flights, taken
    relative to the batch's upper medians over its flyable designs (and absolute with the
    normalisation off);
  - reproducibility: the same seed gives the same generations bit for bit.
"""
import os
import sys

import numpy as np
import pytest
from scipy import stats

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

U32 = np.uint32


@pytest.fixture(scope="module")
def PA():
    import planes_amd
    print("device:", planes_amd.device_name(), flush=True)
    return planes_amd


def mix32(x):
    x = np.asarray(x, np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def hash3(a, b, c):
    a, b, c = (np.asarray(v, np.uint64) & 0xFFFFFFFF for v in (a, b, c))
    inner = mix32((c + 0x27D4EB2F) & 0xFFFFFFFF)
    mid = mix32(((b * 0x85EBCA6B) & 0xFFFFFFFF) ^ inner)
    return mix32(((a * 0x9E3779B9) & 0xFFFFFFFF) ^ mid)


def unit24(h):
    return (np.asarray(h, np.uint64) >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)


def mutation_draws(seed, gen, n_elite, n_pool, ng=35):
    """Parents and standard normals of candidates n_elite..n_pool-1 (FP64 Box-Muller)."""
    i = np.arange(n_elite, n_pool, dtype=np.uint64)
    par = (hash3(seed, gen, i) % n_elite).astype(np.int64)
    z = np.zeros((len(i), ng))
    for j in range(0, ng, 2):
        h1 = hash3(seed ^ 0x5BD1E995, gen * 64 + j, i)
        h2 = mix32((h1 + 0x68E31DA4) & 0xFFFFFFFF)
        u1 = ((h1 >> 8).astype(np.float64) + 1.0) / 16777216.0
        u2 = (h2 >> 8).astype(np.float64) / 16777216.0
        r = np.sqrt(-2.0 * np.log(u1))
        z[:, j] = r * np.cos(2 * np.pi * u2)
        if j + 1 < ng:
            z[:, j + 1] = r * np.sin(2 * np.pi * u2)
    return par, z


def independent_scores(PA, genomes, gen, k, params):
    seeds = (gen * k + np.arange(k) + 1).astype(np.int32)
    pl = PA.Planes(len(genomes), k, len(genomes) * k)
    fit, _ = pl.evaluate(genomes, seeds, params)
    pl.close()
    fit = np.asarray(fit, np.float32).reshape(len(genomes), k)
    return fit, O.robust_scores(fit.reshape(-1), len(genomes), k)


PARAMS = dict(gust_rms_max=4.0, warp_max_deg=1.0, wind_sd=2.0, max_steps=1200)


def test_plain_generation_mutation_and_elite(PA):
    n_elite, n_pool, k, seed, sigma = 32, 4096, 4, 7, 0.05
    params = PA.SimParams.config(**PARAMS)
    g0, _, _ = PA.family_population(n_elite, seed=0)
    ga = PA.PoolGA(n_elite, n_pool, n_pool, k, buf_cap=n_pool, sigma_frac=sigma, seed=seed)
    ga.set_elite(g0)
    gen = 3
    r = ga.generation(gen, params)
    last = ga.last()
    pool, fly, sc = last["pool"], last["fly"], last["scores"]
    assert np.array_equal(fly, pool)  # plain: the whole pool flies
    assert np.array_equal(pool[:n_elite], g0)
    lo, hi = PA.gene_bounds()
    assert (pool >= lo).all() and (pool <= hi).all()
    # mutation generator
    par, z = mutation_draws(seed, gen, n_elite, n_pool)
    rng_ = (hi - lo).astype(np.float64)
    want = np.clip(g0[par].astype(np.float64) + sigma * rng_ * z, lo, hi)
    err = np.abs(pool[n_elite:] - want) / rng_
    print(f"mutants vs FP64 restatement: max |err| / range {err.max():.2e}")
    assert err.max() < 2e-6
    assert stats.kstest(z.reshape(-1), "norm").pvalue > 1e-3
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    cnt = np.bincount(par, minlength=n_elite)
    assert stats.chisquare(cnt).pvalue > 1e-3
    # elite update: simulated scores = an independent evaluation, new elite = best of them
    _, want_sc = independent_scores(PA, fly, gen, k, params)
    assert np.array_equal(sc, want_sc)
    E = ga.elite()
    top = np.sort(sc)[::-1][:n_elite]
    rows = {fly[i].tobytes(): i for i in range(len(fly))}
    e_idx = np.array([rows[e.tobytes()] for e in E])
    assert np.array_equal(np.sort(sc[e_idx])[::-1], top)
    assert r["best"] == pytest.approx(float(top[0]), abs=0) and r["n_fly"] == n_pool
    assert r["elite_mean"] == pytest.approx(float(np.mean(top.astype(np.float64))), rel=1e-12)
    assert r["flights"] == n_pool * k + 8 * 128
    _, hs = ga.champion()
    assert r["holdout"] == pytest.approx(hs, abs=0) and r["best_holdout"] == r["holdout"]


def test_pool_generations_reproducible(PA):
    n_elite, n_pool, k = 16, 1024, 4
    params = PA.SimParams.config(**PARAMS)
    g0, _, _ = PA.family_population(n_elite, seed=1)
    res = []
    for _ in range(2):
        ga = PA.PoolGA(n_elite, n_pool, n_pool, k, buf_cap=n_pool, sigma_frac=0.05, seed=11)
        ga.set_elite(g0)
        rs = [ga.generation(g, params) for g in range(3)]
        res.append((ga.elite(), [(x["best"], x["elite_mean"], x["holdout"]) for x in rs], ga.champion()[0]))
    (e0, r0, c0), (e1, r1, c1) = res
    assert np.array_equal(e0, e1) and r0 == r1 and np.array_equal(c0, c1)


def test_surrogate_prescreening_selection_and_targets(PA):
    n_elite, n_pool, n_fly, k, seed, explore = 32, 8192, 1024, 4, 5, 0.25
    params = PA.SimParams.config(**PARAMS)
    g0, _, _ = PA.family_population(n_elite, seed=2)
    S = PA.Surrogate(max_batch=1024, seed=3)  # untrained weights: any fixed predictor will do
    ga = PA.PoolGA(n_elite, n_pool, n_fly, k, buf_cap=4 * n_fly, sigma_frac=0.05, seed=seed)
    ga.set_elite(g0)
    ga.set_explore(explore)
    gen = 2
    r = ga.generation(gen, params, S, train_epochs=0)
    last = ga.last(targets=True)
    pool, fly, sc, tg = last["pool"], last["fly"], last["scores"], last["targets"]
    assert len(fly) == n_fly and r["n_fly"] == n_fly
    pred = S.predict(pool)[:, 4]
    i = np.arange(n_pool)
    p_explore = np.float32(explore) * np.float32(n_fly - n_elite) / np.float32(n_pool - n_elite)
    is_exp = (i >= n_elite) & (unit24(hash3(seed ^ 0x2545F491, gen, i)) < p_explore)
    keys = np.where(i < n_elite, np.float32(-3.0e38), np.where(is_exp, np.float32(-2.0e38), -pred)).astype(np.float32)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(fly, pool[order[:n_fly]])
    n_exp = int(is_exp.sum())
    print(f"fly set: {n_elite} elite, {n_exp} exploration, {n_fly - n_elite - n_exp} best-predicted")
    assert 0 < n_exp < n_fly - n_elite
    # simulated scores and targets of the fly set vs an independent evaluation
    fit, want_sc = independent_scores(PA, fly, gen, k, params)
    assert np.array_equal(sc, want_sc)
    T = M.fitness_targets(fit)
    ok = T[:, 3] == 0
    med = np.sort(T[ok, :3], axis=0)[ok.sum() // 2]  # the upper median of each head
    want_tg = T.copy()
    want_tg[:, :3] = T[:, :3] - med  # float32, as the library subtracts
    assert np.array_equal(tg, want_tg)
    off = 0.65 * float(med[0]) + 0.25 * float(med[1]) + 0.10 * float(med[2])  # the predicted score's offset
    # Spearman over the exploration sample (positions n_elite .. n_elite + n_exp of the fly set)
    ex = slice(n_elite, n_elite + n_exp)
    a, b = pred[order[:n_fly]][ex], sc[ex]
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    rho = np.corrcoef(ra, rb)[0, 1]
    assert r["rank_corr"] == pytest.approx(rho, abs=1e-9)
    pa, pb = a == -100.0, b == -100.0
    both = ~pa & ~pb
    assert r["mae"] == pytest.approx(float(np.mean(np.abs(a[both].astype(np.float64) + off - b[both]))), rel=1e-6)
    assert r["oversize_acc"] == pytest.approx(float(np.mean(pa == pb)), abs=1e-12)
    # the best-predicted part really is the best predicted outside the elite and the sample
    rest = ~is_exp & (i >= n_elite)
    chosen = order[n_elite + n_exp:n_fly]
    assert pred[chosen].min() >= np.sort(pred[rest])[::-1][len(chosen) - 1]
    # with the normalisation off the stored targets are the absolute statistics
    ga2 = PA.PoolGA(n_elite, n_pool, n_fly, k, buf_cap=4 * n_fly, sigma_frac=0.05, seed=seed)
    ga2.set_elite(g0)
    ga2.set_explore(explore)
    ga2.set_target_norm(False)
    ga2.generation(gen, params, S, train_epochs=0)
    last2 = ga2.last(targets=True)
    assert np.array_equal(last2["fly"], fly) and np.array_equal(last2["targets"], T)


def test_add_training_buffer_trains_like_the_relative_targets(PA):
    """pr_pool_add_training turns distances into the batch's targets relative to its upper
    medians over the flyable designs (absolute with the normalisation off): training on the
    pool's replay buffer gives the network bit for bit the weights that training on those
    targets directly gives (same initial weights, batches in buffer order)."""
    n, k = 2048, 6
    g, _, _ = PA.family_population(n, seed=8)
    rng = np.random.default_rng(4)
    samples = rng.normal(6.0, 3.0, (n, k)).astype(np.float32)
    samples[rng.random(n) < 0.07] = -100.0  # oversize designs: every sample -100
    T = M.fitness_targets(samples)
    ok = T[:, 3] == 0
    rel = T.copy()
    rel[:, :3] = T[:, :3] - np.sort(T[ok, :3], axis=0)[ok.sum() // 2]
    d_g = PA.DeviceBuffer(g.nbytes)
    d_g.upload(np.ascontiguousarray(g, np.float32))
    for norm, want in ((True, rel), (False, T)):
        ga = PA.PoolGA(8, n, n, k, buf_cap=n, sigma_frac=0.05, seed=1)
        ga.set_target_norm(norm)
        ga.add_training(d_g.ptr, samples)
        S1 = PA.Surrogate(max_batch=1024, seed=3)
        ga.train(S1, epochs=1, lr=1e-3)
        S2 = PA.Surrogate(max_batch=1024, seed=3)
        S2.train(g, want, epochs=1, batch=1024, lr=1e-3, shuffle_seed=None)
        assert np.array_equal(S1.params(), S2.params()), norm
        del ga


def test_surrogate_parent_delta_keys(PA):
    """After a generation has simulated the elite, a mutant's selection key is its parent's
    simulated score plus the network's predicted change from parent to mutant: the second
    generation's fly set is the stable-sort order of exactly those keys (parents from the
    mutation hash, the elite's scores from the first generation's flights), and with the
    option off it is the order of the predicted scores themselves.  With key weights
    (1, 0.25, -0.25) (pr_pool_set_key_weights; powers of two, so every product in the key is
    exact) the order is that of parent score + 0.25 x the predicted change."""
    n_elite, n_pool, n_fly, k, seed, explore = 32, 8192, 1024, 4, 9, 0.25
    params = PA.SimParams.config(**PARAMS)
    g0, _, _ = PA.family_population(n_elite, seed=4)
    S = PA.Surrogate(max_batch=1024, seed=5)
    for delta in (True, False, "weighted"):
        ga = PA.PoolGA(n_elite, n_pool, n_fly, k, buf_cap=4 * n_fly, sigma_frac=0.05, seed=seed)
        ga.set_elite(g0)
        ga.set_explore(explore)
        ga.set_parent_delta(bool(delta))
        if delta == "weighted":
            ga.set_key_weights(1.0, 0.25, -0.25)
        gen = 3
        ga.generation(gen, params, S, train_epochs=0)
        last = ga.last()
        fly1, sc1 = last["fly"], last["scores"]
        E = ga.elite()
        rows = {fly1[i].tobytes(): i for i in range(len(fly1))}
        escore = sc1[np.array([rows[e.tobytes()] for e in E])].astype(np.float32)
        ga.generation(gen + 1, params, S, train_epochs=0)
        last = ga.last()
        pool, fly = last["pool"], last["fly"]
        assert np.array_equal(pool[:n_elite], E)
        pred = S.predict(pool)[:, 4].astype(np.float32)
        i = np.arange(n_pool)
        par = (hash3(seed, gen + 1, i) % n_elite).astype(np.int64)
        p_explore = np.float32(explore) * np.float32(n_fly - n_elite) / np.float32(n_pool - n_elite)
        is_exp = (i >= n_elite) & (unit24(hash3(seed ^ 0x2545F491, gen + 1, i)) < p_explore)
        if delta == "weighted":
            s = escore[par] + np.float32(0.25) * (pred - pred[par])
        else:
            s = escore[par] + (pred - pred[par]) if delta else pred
        keys = np.where(i < n_elite, np.float32(-3.0e38), np.where(is_exp, np.float32(-2.0e38), -s)).astype(np.float32)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(fly, pool[order[:n_fly]]), delta
        del ga


def test_surrogate_generations_reproducible(PA):
    """Surrogate mode end to end -- predict, select, fly, add the flown pairs to the replay
    buffer, train -- gives the same elites, holdout scores, losses and network weights on a
    second run: the buffer is filled on the null stream and the network trains on its own
    stream, and the pool orders the two (an unordered version trained on partly copied
    pairs and drifted from run to run)."""
    n_elite, n_pool, n_fly, k = 16, 16384, 4096, 4
    params = PA.SimParams.config(**PARAMS)
    g0, _, _ = PA.family_population(n_elite, seed=4)
    res = []
    for _ in range(2):
        S = PA.Surrogate(max_batch=1024, seed=3)
        ga = PA.PoolGA(n_elite, n_pool, n_fly, k, buf_cap=4 * n_fly, sigma_frac=0.05, seed=13)
        ga.set_elite(g0)
        ga.set_explore(0.25)
        rs = [ga.generation(g, params, S, train_epochs=2, lr=1e-3) for g in range(3)]
        res.append((ga.elite(), S.params(), [(r["holdout"], r["loss"], r["rank_corr"]) for r in rs]))
    (e0, p0, r0), (e1, p1, r1) = res
    print("surrogate generations (holdout, loss, rank corr):", r0)
    assert np.array_equal(e0, e1) and np.array_equal(p0, p1) and r0 == r1


def test_selection_without_network_uniform_and_parent_score(PA):
    """The surrogate's ablation baselines fly the same budget chosen without the network:
    "uniform" -- the elite, then the mutants in the order of a stable sort of hashed uniform
    keys; "parent" -- the elite, then the mutants ordered by their parent's simulated score
    (descending, ties in pool order), the parent drawn as the mutation hash draws it and the
    scores those the previous generation simulated; before any generation, "parent" falls back
    to the uniform keys.  The flown scores are those of an independent evaluation."""
    n_elite, n_pool, n_fly, k, seed = 32, 8192, 1024, 4, 11
    params = PA.SimParams.config(**PARAMS)
    g0, _, _ = PA.family_population(n_elite, seed=6)
    i = np.arange(n_pool)
    for mode in ("uniform", "parent"):
        ga = PA.PoolGA(n_elite, n_pool, n_fly, k, buf_cap=4 * n_fly, sigma_frac=0.05, seed=seed)
        ga.set_elite(g0)
        ga.set_selection(mode)
        gen = 2
        r = ga.generation(gen, params)
        assert r["n_fly"] == n_fly and r["flights"] == n_fly * k + 8 * 128
        last = ga.last()
        pool, fly, sc = last["pool"], last["fly"], last["scores"]
        u = unit24(hash3(seed ^ 0x6A09E667, gen, i))
        keys = np.where(i < n_elite, np.float32(-3.0e38), -u).astype(np.float32)
        assert np.array_equal(fly, pool[np.argsort(keys, kind="stable")[:n_fly]]), mode
        _, want = independent_scores(PA, fly, gen, k, params)
        assert np.array_equal(sc, want)
        E = ga.elite()
        rows = {fly[j].tobytes(): j for j in range(len(fly))}
        escore = sc[np.array([rows[e.tobytes()] for e in E])].astype(np.float32)
        ga.generation(gen + 1, params)
        last = ga.last()
        pool, fly = last["pool"], last["fly"]
        if mode == "uniform":
            u = unit24(hash3(seed ^ 0x6A09E667, gen + 1, i))
            keys = np.where(i < n_elite, np.float32(-3.0e38), -u).astype(np.float32)
        else:
            par = (hash3(seed, gen + 1, i) % n_elite).astype(np.int64)
            keys = np.where(i < n_elite, np.float32(-3.0e38), -escore[par]).astype(np.float32)
        assert np.array_equal(fly, pool[np.argsort(keys, kind="stable")[:n_fly]]), mode
        if mode == "parent":  # the chosen mutants descend from the best-scoring parents
            par = (hash3(seed, gen + 1, i) % n_elite).astype(np.int64)
            chosen = np.argsort(keys, kind="stable")[n_elite:n_fly]
            assert escore[par[chosen]].min() >= np.sort(escore)[len(escore) // 2]
        del ga
