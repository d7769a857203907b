"""The sim.h entry points at run time, called the way the reference's own callers call them,
and the chunked flight driver's state carry (sim.cu:110-114, 208-211, 237-248).

tests/native/shim_driver.cpp (built by planes-rl_amd/csrc/Makefile, target tests) is a C++
program that evaluates one generation through panelbuf_alloc / gpu_prepare / gpu_sim exactly
as run_ga does (sim.cu:411-456): one gpu_sim per seed with vary = 0, then for each of the top
eight a 128-copy batch flown with vary = 1 from seed 100003.  Its distances are checked against
the FP64 oracle, its ranking against the oracle's robust scores and std::sort order, its
holdout scores against the oracle's, and all of it against the batched pr_gen_evaluate.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "native", "_bin", "shim_driver")


@pytest.fixture(scope="module")
def PA():
    import planes_amd
    print("device:", planes_amd.device_name(), flush=True)
    return planes_amd


def spearman(a, b):
    ra = np.argsort(np.argsort(a, kind="stable"), kind="stable").astype(np.float64)
    rb = np.argsort(np.argsort(b, kind="stable"), kind="stable").astype(np.float64)
    return np.corrcoef(ra, rb)[0, 1]


def run_driver(genomes, nseeds, gen, tmp_path):
    gpath, opath = tmp_path / "g.bin", tmp_path / "o.bin"
    gpath.write_bytes(np.ascontiguousarray(genomes, np.float32).tobytes())
    print(f"shim driver: {len(genomes)} genomes x {nseeds} seeds, generation {gen}", flush=True)
    r = subprocess.run([DRIVER, str(gpath), str(len(genomes)), str(nseeds), str(gen), str(opath)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    b = opath.read_bytes()
    pop, ns, k = struct.unpack("<3i", b[:12])
    o = 12
    samples = np.frombuffer(b, np.float32, pop * ns, o).reshape(pop, ns)
    o += 4 * pop * ns
    idx = np.frombuffer(b, np.int32, pop, o)
    o += 4 * pop
    hold = np.frombuffer(b, np.float32, k, o)
    o += 4 * k
    champ = struct.unpack("<i", b[o:o + 4])[0]
    return samples, idx, hold, champ


def test_sim_h_shims_evaluate_a_generation_like_run_ga(PA, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/native/_bin/shim_driver not built (make -C planes-rl_amd/csrc tests)")
    pop, nseeds, gen = 384, 4, 2
    gs = O.random_genomes(pop, 23)
    samples, idx, hold, champ = run_driver(gs, nseeds, gen, tmp_path)
    seeds = np.array([gen * nseeds + s + 1 for s in range(nseeds)], np.int32)

    # host logic of the caller (sim.cu:439-441) = the oracle's robust scores and std::sort order
    acc = O.robust_scores(samples, pop, nseeds)
    assert np.array_equal(idx, O.ga_rank(acc))

    # distances against the FP64 oracle (reference preset, 7200-step cap: gpu_sim's settings)
    gi = np.repeat(np.arange(pop, dtype=np.int32), nseeds)
    f64, _, _ = O.fly(gs, gi, np.tile(seeds, pop), prec=64, nthreads=min(16, os.cpu_count() or 1))
    gpu = samples.reshape(-1).astype(np.float64)
    rel = np.abs(gpu - f64) / np.maximum(np.abs(f64), 1e-3)
    print(f"  vs FP64: {np.mean(rel < 0.01) * 100:.2f}% within 1%, spearman {spearman(gpu, f64):.6f}", flush=True)
    assert np.mean(rel < 0.01) >= 0.97 and spearman(gpu, f64) >= 0.99
    assert np.array_equal(gpu == -100.0, f64 == -100.0)

    # the holdout: each of the top eight under seeds 100003 + p * 7919 (vary = 1, sim.cu:64,452)
    hseeds = (100003 + 7919 * np.arange(128)).astype(np.int32)
    for h in range(len(hold)):
        fh, _, _ = O.fly(gs[idx[h]][None], np.zeros(128, np.int32), hseeds, prec=64,
                         nthreads=min(16, os.cpu_count() or 1))
        want = float(O.robust_scores(fh.astype(np.float32), 1, 128)[0])
        assert abs(hold[h] - want) <= 0.01 * max(abs(want), 1.0), (h, hold[h], want)
    assert champ == idx[int(np.argmax(hold))]

    # the same generation through the batched evaluator: equal up to the reassociation of the
    # panel sums where the two batch shapes fly a flight with different lane mappings
    E = PA.Generation(pop, nseeds)
    E.upload(gs)
    r = E.evaluate(pop, seeds, PA.SimParams.reference(), want_samples=True)
    b = r["samples"].reshape(-1).astype(np.float64)
    relb = np.abs(gpu - b) / np.maximum(np.abs(b), 1e-3)
    print(f"  vs pr_gen_evaluate: {np.mean(gpu == b) * 100:.1f}% bit-identical, {np.mean(relb < 1e-3) * 100:.1f}% "
          f"within 1e-3", flush=True)
    assert np.mean(relb < 1e-3) >= 0.95 and np.array_equal(gpu == -100.0, b == -100.0)
    if np.array_equal(samples, r["samples"]):
        assert np.array_equal(idx, r["idx"])
    top = np.sort(hold)[::-1]
    if len(top) > 1 and top[0] - top[1] > 0.01 * abs(top[0]):
        assert r["champion"] == champ


def test_chunked_flights_carry_state_exactly(PA):
    """Flight state carried across 1-s launches (sim.cu:110-114, 208-211): with a steady wind
    (no gusts, so the chunk clock rule t = step0 * dt does not enter) and one lane mapping per
    bucket, a 6-s flight in 600-, 1200-, 1800- or 3600-step launches gives the same distances
    and step counts bit for bit."""
    gs, res, _ = PA.family_population(300, seed=8)
    seeds = np.arange(1, 5, dtype=np.int32)
    out = []
    for chunk in (600, 1200, 1800, 3600):
        params = PA.SimParams.config(gust_rms_max=0.0, warp_max_deg=1.0, wind_sd=2.0, max_steps=3600, chunk=chunk)
        pl = PA.Planes(len(gs), 4, 4 * len(gs), max_ch=int(res[:, 0].max()), max_sp=int(res[:, 1].max()))
        pl.set_tail_mapping(0)
        pl.set_split(False)
        fit, steps = pl.evaluate(gs, seeds, params, res=res)
        out.append((chunk, fit.reshape(-1).copy(), steps.reshape(-1).copy(), pl.stats()["fly_launches"]))
        pl.close()
    _, f0, s0, _ = out[0]
    assert (s0 > 600).sum() > 100  # many flights cross launch boundaries
    for chunk, f, s, nl in out[1:]:
        assert np.array_equal(f, f0), chunk
        assert np.array_equal(s, s0), chunk
    assert out[0][3] > out[-1][3]  # fewer, longer launches
