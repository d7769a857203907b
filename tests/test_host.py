"""Host side of the product on the CPU (no GPU calls): the C ABI exports every entry
point the headers declare, argument errors fail loudly, and the GA's ordering/RNG work
is bit-exact with the oracle's restatement of sim.cu:315-490 (both use libstdc++)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import planes_amd as PA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions():
    src = open(os.path.join(ROOT, "include", "planes_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = _declared_c_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(PA.LIB, n), n
    out = subprocess.run(["nm", "-DC", "--defined-only", PA.LIB_PATH], capture_output=True, text=True).stdout
    for sig in ("run_ga(int, int, int, char const*, char const*)",
                "evaluate_design(Genome const*, int, int, char const*)",
                "select_finalist(Genome const*, int, int, int, char const*)",
                "panelbuf_alloc(int)", "panelbuf_free(void*)", "gpu_prepare(Genome const*, void*, int)",
                "gpu_sim(void*, int, int, float*, float8*, int, int, float, int)"):
        assert sig in out, sig


def test_argument_errors_fail_loudly():
    h = C.c_void_p()
    rc = PA.LIB.pr_create(C.byref(h), 0, 0, 1, 1, 0, 0)
    assert rc != 0 and b"capacities" in PA.LIB.pr_last_error()
    with pytest.raises(PA.PlanesError):
        PA._ck(-1, "probe")
    # pool entry points check their handle before touching the GPU
    assert PA.LIB.pr_pool_set_key_weights(None, 1.0, 1.0, -1.0) == -1
    assert b"pr_pool_set_key_weights" in PA.LIB.pr_pool_last_error()
    assert PA.LIB.pr_pool_set_parent_delta(None, 1) == -1


def test_gene_bounds_match_reference_table():
    lo, hi = PA.gene_bounds()
    olo, ohi = O.bounds()
    assert np.array_equal(lo, olo) and np.array_equal(hi, ohi)
    assert np.all(lo < hi)


def test_rank_is_std_sort_order_with_ties():
    rng = np.random.default_rng(0)
    for pop in (2, 17, 1000, 4096, 100_000):
        acc = rng.integers(-3, 4, pop).astype(np.float32)  # heavy ties
        assert np.array_equal(PA.ga_rank(acc), O.ga_rank(acc))
        acc = rng.normal(size=pop).astype(np.float32)
        idx = PA.ga_rank(acc)
        assert np.array_equal(idx, O.ga_rank(acc))
        assert np.all(np.diff(acc[idx]) <= 0)


@pytest.mark.parametrize("n", [33, 8193, 40_000, 250_000])
def test_rank_matches_std_sort_on_ordered_and_tied_patterns(n):
    """The ranking's introsort (parallel above 32 768 genomes) against std::sort of the
    indices: ascending, descending, all equal, organ pipe, two values, signed zeros, a third
    at -100 (unprintable), uniform."""
    rng = np.random.default_rng(n)
    i = np.arange(n)
    pats = {
        "ascending": i.astype(np.float32),
        "descending": (n - i).astype(np.float32),
        "equal": np.ones(n, np.float32),
        "organ": np.minimum(i, n - i).astype(np.float32),
        "two": rng.integers(0, 2, n).astype(np.float32),
        "zeros": np.where(i % 2 == 1, np.float32(0.0), np.float32(-0.0)).astype(np.float32),
        "unprintable": np.where(i % 3 == 0, -100.0, rng.uniform(0, 15, n)).astype(np.float32),
        "uniform": rng.uniform(-5, 15, n).astype(np.float32),
    }
    for name, acc in pats.items():
        assert np.array_equal(PA.ga_rank(acc), O.ga_rank(acc)), name


def test_random_genomes_and_breeding_match_oracle_sequence():
    pop, gens = 256, 6
    ga = PA.GA(42)
    ho = O.lib().or_ga2_new(42)
    try:
        g = ga.random(pop)
        go = np.zeros(pop * 35, np.float32)
        O.lib().or_ga2_random(ho, go, pop)
        assert np.array_equal(g.reshape(-1), go)
        rng = np.random.default_rng(1)
        for gen in range(gens):
            acc = rng.normal(size=pop).astype(np.float32)
            idx = PA.ga_rank(acc)
            nxt = ga.breed(g, acc, idx, gen, gens)
            nxo = np.zeros(pop * 35, np.float32)
            O.lib().or_ga2_breed(ho, g.reshape(-1), acc, idx, pop, gen, gens, nxo)
            assert np.array_equal(nxt.reshape(-1), nxo), gen
            lo, hi = PA.gene_bounds()
            assert np.all(nxt >= lo) and np.all(nxt <= hi)
            g = nxt
    finally:
        O.lib().or_ga2_free(ho)


def test_host_robust_score_matches_oracle():
    rng = np.random.default_rng(2)
    for n in (1, 3, 8, 12, 128, 300):
        row = rng.normal(5, 4, n).astype(np.float32)
        row[rng.random(n) < 0.1] = -100.0
        assert PA.LIB.pr_robust_score_host(row, n) == O.robust_scores(row, 1, n)[0]


def test_sanitize_clamps_like_gset():
    lo, hi = PA.gene_bounds()
    g = np.stack([lo - 1, hi + 1, (lo + hi) / 2]).astype(np.float32)
    s = PA.ga_sanitize(g)
    assert np.array_equal(s[0], lo) and np.array_equal(s[1], hi) and np.array_equal(s[2], g[2])


def _design_stats_f32(x):
    """evaluate_design's summary restated in float32 (sim.cu:250-256, 291-299): moments summed
    over the sorted samples with one division per term, interpolated order statistics."""
    v = np.sort(np.asarray(x, np.float32))
    n = len(v)
    fn = np.float32(n)
    m1 = np.float32(0)
    m2 = np.float32(0)
    for a in v:
        m1 = np.float32(m1 + a / fn)
        m2 = np.float32(m2 + np.float32(a * a) / fn)

    def pct(q):
        x = np.float32(q) * np.float32(n - 1)
        lo, hi = int(np.floor(x)), int(np.ceil(x))
        t = np.float32(x - np.float32(lo))
        return np.float32(v[lo] * (np.float32(1) - t) + v[hi] * t)

    w = max(1, n // 10)
    cv = np.float32(0)
    for a in v[:w]:
        cv = np.float32(cv + a / np.float32(w))
    sd = np.float32(np.sqrt(np.float32(max(np.float32(0), np.float32(m2 - m1 * m1)))))
    return [v[0], pct(.05), pct(.10), pct(.25), pct(.5), m1, sd, pct(.75), pct(.9), pct(.95), v[-1], cv,
            np.float32((v < 0).sum())]


@pytest.mark.parametrize("n", [8, 37, 128, 4096])
def test_design_stats_match_float32_restatement(n):
    rng = np.random.default_rng(n)
    x = (rng.normal(8, 4, n) * (rng.random(n) > 0.05) - 2 * (rng.random(n) < 0.03)).astype(np.float32)
    got = PA.design_stats(x)
    want = _design_stats_f32(x)
    for k, w in zip(PA.AUDIT_FIELDS, want):
        assert np.float32(got[k]) == w, (k, got[k], w)


def test_design_stats_closed_forms():
    """Samples 0..n-1 in any order: min 0, max n-1, quantile q = q (n-1), no failures; with
    k negatives, failures = k and CVaR10 = the mean of the worst n/10."""
    n = 1001
    x = np.random.default_rng(0).permutation(np.arange(n, dtype=np.float32))
    s = PA.design_stats(x)
    assert s["min"] == 0 and s["max"] == n - 1 and s["failures"] == 0
    for q, k in ((.05, "p05"), (.25, "p25"), (.5, "median"), (.9, "p90")):
        assert s[k] == pytest.approx(q * (n - 1), abs=1e-3)
    assert s["mean"] == pytest.approx((n - 1) / 2, rel=1e-5)
    assert s["sd"] == pytest.approx(np.sqrt((n * n - 1) / 12.0), rel=1e-3)
    assert s["cvar10"] == pytest.approx(np.mean(np.arange(100)), rel=1e-5)
    x[:7] = -1.0
    assert PA.design_stats(x)["failures"] == 7
