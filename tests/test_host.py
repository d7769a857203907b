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


def test_gene_bounds_match_reference_table():
    lo, hi = PA.gene_bounds()
    olo, ohi = O.bounds()
    assert np.array_equal(lo, olo) and np.array_equal(hi, ohi)
    assert np.all(lo < hi)


def test_rank_is_std_sort_order_with_ties():
    rng = np.random.default_rng(0)
    for pop in (2, 17, 1000, 4096):
        acc = rng.integers(-3, 4, pop).astype(np.float32)  # heavy ties
        assert np.array_equal(PA.ga_rank(acc), O.ga_rank(acc))
        acc = rng.normal(size=pop).astype(np.float32)
        idx = PA.ga_rank(acc)
        assert np.array_equal(idx, O.ga_rank(acc))
        assert np.all(np.diff(acc[idx]) <= 0)


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
