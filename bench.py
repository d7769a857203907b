#!/usr/bin/env python3
"""Benchmark of the fitness-evaluation path on one MI355X.

One step = one GA generation's evaluation (pr_gen_evaluate): meshing + mass properties of
every genome, scenario generation, every (genome, seed) flight at 600 Hz in 1-s chunks,
tail-aware robust scores, ranking, and the top-8 re-test on the 128-scenario holdout
suite.  Breeding and file I/O are outside the timed region (the metric's definition).

value = plane-flight-seconds per wall-second = sum of integrated steps / 600 over the
timed steps (training and holdout flights) / wall time, genomes resident in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "planes-rl_amd"))

METRIC = "plane-flight-sec/s at 600 Hz; GA gens/hour at 100k×12 env, plain vs surrogate"
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak
FLOP_PANEL, FLOP_PLANE = 68, 222  # algorithmic FLOPs per panel-step and per plane-step
# HBM bytes per k_fly_all launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled
# by a 1 GiB copy's readings) on this workload; the counters cannot be read inside a timed run
TRAFFIC = {"c3": (3.11e8, "profiles/k_fly_all_c3_hbm_traffic.md")}  # per population launch

WORKLOADS = {
    # name: pop, seeds per genome, flight cap [steps], gust RMS max, warp max [deg], steady wind sd
    "c1": dict(pop=64, nseeds=8, max_steps=4800, gust=2.0, warp_deg=1.0, wind_sd=2.0,
               desc="64 genomes x 8 seeded launch scenarios, config preset, 8 s cap"),
    "c2": dict(pop=10000, nseeds=8, max_steps=9000, gust=2.0, warp_deg=1.0, wind_sd=2.0, graph=True,
               desc="one GA generation: pop 10000 x 8 seeds, config preset, 15 s cap, top-8 x 128 holdout"),
    "c3": dict(pop=100000, nseeds=12, max_steps=7200, gust=4.0, warp_deg=1.0, wind_sd=2.0, families=True, autotune=True,
               desc="one GA generation: pop 100000 x 12 environments, mixed families (delta/elliptical/plank/"
                    "biplane, centre/twin/triple/outboard fins), triangles log-uniform 200..900, gusts <= 4 m/s "
                    "RMS, 12 s cap, top-8 x 128 holdout"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def seeds_for(gen, nseeds):
    return np.array([gen * nseeds + s + 1 for s in range(nseeds)], np.int32)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(genomes, res, w, budget_s):
    """The oracle's FP32 transcription on this host's cores over a bounded sample of the
    same workload (same genome distribution and mesh resolutions, scenario preset and cap):
    random (genome, seed) flights pulled from an atomic counter by all host threads, the
    sample grown until it holds about budget_s of work; then the 1-thread rate on a
    sample of about budget_s / 5."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    nth = os.cpu_count() or 1
    rng = np.random.default_rng(1)
    kw = dict(prec=32, preset=1, gust=w["gust"], warp=float(np.deg2rad(w["warp_deg"])), wsd=w["wind_sd"],
              max_steps=w["max_steps"])
    rs = np.ascontiguousarray(res if res is not None else np.tile([[5, 8]], (len(genomes), 1)), np.int32)

    def run(n, threads):
        gi = rng.integers(0, len(genomes), n).astype(np.int32)
        gi.sort()  # a thread keeps its prepared genome while consecutive flights share it
        sd = rng.integers(1, 1 + w["nseeds"] * 1000, n).astype(np.int32)
        _, steps, wall = O.fly_res(genomes, rs, gi, sd, nthreads=threads, **kw)
        return int(steps.sum()), wall

    def sized(threads, budget):
        n = max(64, 4 * threads)
        while True:
            steps, wall = run(n, threads)
            if wall >= 0.6 * budget or n >= 4_000_000:
                return n, steps, wall
            n = int(n * min(8.0, budget / max(wall, 1e-3)))

    n, steps, wall = sized(nth, budget_s)
    n1, steps1, wall1 = sized(1, budget_s / 5)
    return {"value": steps / 600.0 / wall, "unit": "plane-flight-s/s", "cores": nth, "kind": "port",
            "single_thread_value": steps1 / 600.0 / wall1, "cpu": cpu_model(),
            "sample": f"{n} random (genome, seed) flights of the same workload (per-genome mesh resolution), FP32 "
                      f"oracle (g++ -O3 -ffp-contract=off), {nth} threads, {wall:.1f} s wall, {steps / 600:.0f} "
                      f"plane-flight-s; 1 thread: {n1} flights, {wall1:.1f} s"}


def fp64_check(genomes, res, w, samples, seeds):
    """The oracle checking the timed run's own results (this leg only, beside the CPU baseline):
    the distances the last timed generation flew for its first genomes (read back from the
    device after the timed region, as that generation's mapping produced them) against the
    FP64 oracle flying the same (genome, seed) flights on the host; per-flight distance error,
    the rank agreement of the flights and the -100 (unprintable) flags."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    n = len(samples)
    g = np.ascontiguousarray(genomes[:n], np.float32)
    rs = np.ascontiguousarray(res[:n] if res is not None else np.tile([[5, 8]], (n, 1)), np.int32)
    gpu = np.asarray(samples, np.float64).reshape(-1)
    gi = np.repeat(np.arange(n, dtype=np.int32), len(seeds))
    sd = np.tile(seeds, n).astype(np.int32)
    f64, _, wall = O.fly_res(g, rs, gi, sd, nthreads=os.cpu_count() or 1, prec=64, preset=1, gust=w["gust"],
                             warp=float(np.deg2rad(w["warp_deg"])), wsd=w["wind_sd"], max_steps=w["max_steps"])
    f64 = np.asarray(f64, np.float64)
    rel = np.abs(gpu - f64) / np.maximum(np.abs(f64), 1e-3)
    ra = np.argsort(np.argsort(gpu, kind="stable"), kind="stable")
    rb = np.argsort(np.argsort(f64, kind="stable"), kind="stable")
    return {"planes": n, "flights": int(len(gpu)), "source": "the last timed generation's own samples",
            "seeds": [int(seeds[0]), int(seeds[-1])], "within_1pct": round(float(np.mean(rel < 0.01)), 4),
            "median_rel_err": float(f"{np.median(rel):.3g}"), "p99_rel_err": float(f"{np.quantile(rel, 0.99):.3g}"),
            "spearman": round(float(np.corrcoef(ra, rb)[0, 1]), 6),
            "unprintable_flags_equal": bool(np.array_equal(gpu == -100.0, f64 == -100.0)),
            "fp64_wall_s": round(wall, 1)}


def relative_targets(tg):
    """The regression targets (mean, CVaR, worst) of one batch flown under one scenario set,
    relative to the batch's medians over its flyable designs (pr_pool_set_target_norm)."""
    tg = tg.copy()
    ok = tg[:, 3] == 0
    if ok.any():  # the upper median (element n // 2 in order), as the library takes it
        tg[:, :3] -= np.sort(tg[ok, :3], axis=0)[ok.sum() // 2]
    return tg


def pretrain_surrogate(PA, S, n_pairs, params, source="c2", runs=8, batch=250000, seeds_per=8, epochs=4, lr=2e-3,
                       norm=False):
    """Setup (untimed, reported): about n_pairs (genome, fitness) pairs, on whose
    [mean, CVaR25, worst, oversize] targets the network trains.
    source "c2" (the configuration's "generated by config 2 runs"): `runs` independent runs
    of the reference GA at config 2's settings -- pop 10 000 x 8 seeds (gen g flies seeds
    8 g + 1 .. 8 g + 8), config preset with gusts <= 2 m/s RMS, 15 s cap, reference mesh --
    each for n_pairs / (runs x 10 000) generations; every genome a generation evaluates is
    one pair with its 8-seed statistics.
    source "family": n_pairs designs of the family generator, each flown under seeds_per
    seeds with the workload's scenario preset (params), in batches.
    norm: each batch's targets relative to its medians (relative_targets)."""
    t0 = time.perf_counter()
    gs, tg = [], []
    if source == "c2":
        w = WORKLOADS["c2"]
        pop, ns = w["pop"], w["nseeds"]
        gens = max(1, n_pairs // (runs * pop))
        p2 = PA.SimParams.config(gust_rms_max=w["gust"], warp_max_deg=w["warp_deg"], wind_sd=w["wind_sd"],
                                 max_steps=w["max_steps"])
        E = PA.Generation(pop, ns)
        for run in range(runs):
            ga = PA.GA(seed=101 + run)
            g = ga.random(pop)
            for gen in range(gens):
                E.upload(g)
                r = E.evaluate(pop, seeds_for(gen, ns), p2, holdout_k=0, want_samples=True)
                gs.append(g)
                t = robust_stats(r["samples"])[0]
                tg.append(relative_targets(t) if norm else t)
                g = ga.breed(g, r["acc"], r["idx"], gen, gens)
        desc = (f"config 2 runs: {runs} GA runs x {gens} generations, pop {pop} x {ns} seeds, config preset "
                f"(gusts <= {w['gust']} m/s RMS), {w['max_steps'] // 600} s cap")
        per = ns
    else:
        seeds = seeds_for(900, seeds_per)
        for b0 in range(0, n_pairs, batch):
            n = min(batch, n_pairs - b0)
            g, r, _ = PA.family_population(n, seed=31 + b0 // batch)
            pl = PA.Planes(n, seeds_per, n * seeds_per, int(r[:, 0].max()), int(r[:, 1].max()))
            fit, _ = pl.evaluate(g, seeds, params, res=r)
            pl.close()
            gs.append(g)
            t = robust_stats(np.asarray(fit).reshape(n, seeds_per))[0]
            tg.append(relative_targets(t) if norm else t)
        desc, per = "family generator, config preset", seeds_per
    sim_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    G, T = np.concatenate(gs), np.concatenate(tg)
    loss = S.train(G, T, epochs=epochs, batch=4096, lr=lr)
    return {"pairs": int(len(G)), "seeds_per_design": per, "source": desc, "relative_targets": bool(norm),
            "sim_s": round(sim_s, 2),
            "train_s": round(time.perf_counter() - t0, 2), "epochs": epochs, "final_loss": round(float(loss), 4)}


ARMS = {
    # name: (flies, selection) -- what each arm of the surrogate A/B flies per generation
    "plain": "the whole pool (no pre-screen)",
    "surrogate": "the elite + the network's choice: parent's simulated score + predicted parent-to-mutant change, "
                 "with an exploration sample",
    "strict": "the elite + the top candidates by the network's predicted score alone (no exploration sample)",
    "uniform": "the elite + a uniform random sample of the mutants (no network)",
    "parent": "the elite + the mutants of the best-scoring parents (parent's simulated score as the key, no network)",
}


def surrogate_ab(PA, n_pool=1_000_000, n_fly=100_000, n_elite=10_000, reps=3, gens=4, sur_gens=None, pretrain=2_000_000,
                 explore=0.5, buffer=300_000, mlp_batch=4096, train_epochs=None, lr=None, boot_epochs=16,
                 arms=("plain", "surrogate", "strict", "uniform", "parent"), delta_mode=True, pretrained_state=None,
                 pretrain_source="c2", target_norm=True, key_weights=None):
    """GA generations/hour with surrogate pre-screening against the plain batched sim, and
    against the same fly budget chosen without the network.  One generation: a pool of n_pool
    Gaussian mutants (sigma 5 % of each gene's range) of an n_elite elite; every arm flies its
    fly set under 12 environments (gusts <= 4 m/s RMS, 12 s cap), keeps the best simulated
    genomes as the next elite and re-flies its top 8 on the 128-scenario holdout suite.  The
    arms (ARMS): "plain" flies all n_pool candidates; the others fly n_fly.  The network is
    pretrained once on about `pretrain` sim pairs (pretrain_surrogate: config 2 GA runs by
    default; setup, untimed, its cost reported and also added to the network arms' time to
    the plain champion), and each network run starts from those weights and fine-tunes on its own
    flights.  With pretrain 0 a network arm's first generation is a plain one whose flights
    bootstrap the network (timed).  Every arm runs `reps` GA seeds from the same initial
    elite; the plain arm `gens` generations, the others `sur_gens` (default 3 x gens).
    Reported per arm: generations/hour, champions (best holdout score so far) at the plain
    arm's generation count and at the end, and per GA seed the first generation whose champion
    reaches 99 % of that seed's plain champion, with the wall time of both arms to that point."""
    sur_gens = sur_gens or 3 * gens
    params = PA.SimParams.config(gust_rms_max=4.0, warp_max_deg=1.0, wind_sd=2.0, max_steps=7200)
    g0, _, _ = PA.family_population(n_elite, seed=0)
    pretrained = pretrain > 0
    net_arms = [a for a in arms if a in ("surrogate", "strict")]
    lr = lr if lr is not None else (1e-4 if pretrained else 2e-3)
    tepochs = train_epochs if train_epochs is not None else (2 if pretrained else 4)
    pre, p0, pre_s = None, None, 0.0
    if pretrained and net_arms:
        if pretrained_state is not None:  # weights pretrained by an earlier call (same settings)
            pre, p0 = pretrained_state
        else:
            S = PA.Surrogate(max_batch=mlp_batch, seed=3)
            pre = pretrain_surrogate(PA, S, pretrain, params, source=pretrain_source, norm=target_norm)
            p0 = S.params()
            log(f"[surrogate] pretraining: {pre}")
        pre_s = pre["sim_s"] + pre["train_s"]
    runs = {a: [] for a in arms}
    for rep in range(reps):
        for arm in arms:
            ga = PA.PoolGA(n_elite, n_pool, n_pool if arm == "plain" else n_fly, 12, buf_cap=buffer, sigma_frac=0.05,
                           seed=7 + rep)
            ga.set_elite(g0)
            S, out = None, []
            if arm in ("uniform", "parent"):
                ga.set_selection(arm)
            elif arm in ("surrogate", "strict"):
                S = PA.Surrogate(max_batch=mlp_batch, seed=3)
                ga.set_explore(explore if arm == "surrogate" else 0.0)
                ga.set_parent_delta(delta_mode if arm == "surrogate" else False)
                if arm == "surrogate" and key_weights is not None:
                    ga.set_key_weights(*key_weights)
                ga.set_target_norm(target_norm)
                if pretrained:
                    S.set_params(p0)
                else:
                    ga.set_collect(True)
                    r = ga.generation(0, params, None)
                    ga.set_collect(False)
                    t0 = time.perf_counter()
                    r["loss"] = ga.train(S, epochs=boot_epochs, lr=lr)
                    r["wall_ms"] += 1000 * (time.perf_counter() - t0)
                    out.append(r)
            n_gens = gens if arm == "plain" else max(gens, sur_gens)
            for gen in range(len(out), n_gens):
                r = ga.generation(gen, params, S, train_epochs=tepochs, lr=lr)
                out.append(r)
                log(f"[{arm} {rep}] gen {gen}: best {r['best']:.3f} elite {r['elite_mean']:.3f} holdout "
                    f"{r['holdout']:.3f} (best {r['best_holdout']:.3f}) {r['wall_ms']:.0f} ms (fly {r['fly_ms']:.0f}, "
                    f"mlp {r['mlp_ms']:.0f})" + (f" rank-corr {r['rank_corr']:.3f} MAE {r['mae']:.2f} m" if S else ""))
            runs[arm].append(out)
            del ga
    summary = {}
    for arm, reps_out in runs.items():
        allg = [r for g in reps_out for r in g]
        wall = sum(r["wall_ms"] for r in allg) / 1000.0
        fin = lambda key: [r[key] for r in allg if r[key] == r[key]]
        d = {"flies": ARMS[arm], "gens": len(allg), "wall_s": round(wall, 3),
             "gens_per_hour": round(3600.0 * len(allg) / wall, 1), "flights_per_gen": allg[-1]["flights"],
             "champion_at_plain_gens": [round(g[min(gens, len(g)) - 1]["best_holdout"], 4) for g in reps_out],
             "champion_final": [round(g[-1]["best_holdout"], 4) for g in reps_out],
             "mlp_ms_per_gen": round(float(np.mean([r["mlp_ms"] for r in allg])), 1)}
        if arm in net_arms:
            d["mean_rank_corr"] = round(float(np.mean(fin("rank_corr"))), 3) if fin("rank_corr") else None
            d["mean_mae_m"] = round(float(np.mean(fin("mae"))), 3) if fin("mae") else None
        summary[arm] = d
    if "plain" in runs:
        for arm in arms:
            if arm == "plain":
                continue
            per = []
            for rep in range(reps):
                pg, sg = runs["plain"][rep], runs[arm][rep]
                target, wp = pg[-1]["best_holdout"], sum(r["wall_ms"] for r in pg) / 1000.0
                ws, hit = 0.0, None
                for i, r in enumerate(sg):
                    ws += r["wall_ms"] / 1000.0
                    if r["best_holdout"] >= 0.99 * target:
                        hit = {"gen": i + 1, "wall_s": round(ws, 3)}
                        break
                m = {"plain_champion": round(target, 4), "plain_wall_s": round(wp, 3), "reached_at": hit,
                     "speedup": round(wp / hit["wall_s"], 2) if hit else None}
                if hit and arm in net_arms and pre_s > 0:
                    m["speedup_with_pretraining"] = round(wp / (hit["wall_s"] + pre_s), 2)
                per.append(m)
            hits = [m["speedup"] for m in per if m["speedup"] is not None]
            summary[arm]["time_to_plain_champion"] = {
                "per_seed": per, "reached": f"{len(hits)}/{len(per)}",
                "speedup_median": float(np.median(hits)) if hits else None}
            summary[arm]["gens_per_hour_vs_plain"] = round(summary[arm]["gens_per_hour"] /
                                                           summary["plain"]["gens_per_hour"], 2)
            summary[arm]["champion_ratio_at_plain_gens"] = round(
                float(np.mean(summary[arm]["champion_at_plain_gens"])) /
                float(np.mean(summary["plain"]["champion_at_plain_gens"])), 4)
    # what the network adds over the same fly budget chosen without it
    gain = {}
    if "surrogate" in summary:
        s_end = float(np.mean(summary["surrogate"]["champion_final"]))
        for base in ("uniform", "parent"):
            if base in summary:
                b_end = float(np.mean(summary[base]["champion_final"]))
                gain[f"vs_{base}"] = {"champion_final_ratio": round(s_end / b_end, 4),
                                      "champion_final_diff_m": round(s_end - b_end, 3),
                                      "per_seed_diff_m": [round(a - b, 3) for a, b in
                                                          zip(summary["surrogate"]["champion_final"],
                                                              summary[base]["champion_final"])]}
    for arm in net_arms:
        w = [(r["w_parent"], r["w_pred"], r["w_parent_pred"]) for g in runs[arm] for r in g
             if r.get("w_parent") == r.get("w_parent")]
        if w:  # least-squares weights of the mutants' scores on (parent score, prediction, parent's prediction)
            summary[arm]["mean_key_fit"] = [round(float(v), 3) for v in np.mean(np.array(w), 0)]
    return {"config": {"pool": n_pool, "fly": n_fly, "elite": n_elite, "environments": 12, "ga_seeds": reps,
                       "plain_gens": gens, "other_gens": max(gens, sur_gens), "explore": explore,
                       "relative_targets": bool(target_norm),
                       "key_weights": list(key_weights) if key_weights is not None else [1.0, 1.0, -1.0],
                       "fine_tuning": {"lr": lr, "epochs_per_gen": tepochs, "buffer": buffer},
                       "scenarios": "config preset, gusts <= 4 m/s RMS, warp +-1 deg, 12 s cap"},
            "pretraining": pre, "arms": summary, "network_gain": gain, "_pretrained_state": (pre, p0)}


def pool_ab(PA, args):
    """C4 as its own bench line: the surrogate A/B (surrogate_ab) with the command-line sizes;
    value = the surrogate arm's generations/hour."""
    ab = surrogate_ab(PA, n_pool=args.pool, n_fly=args.fly, n_elite=args.elite, reps=args.reps, gens=args.gens,
                      sur_gens=args.sur_gens, pretrain=args.pretrain, pretrain_source=args.pretrain_source,
                      explore=args.explore, buffer=args.buffer, target_norm=args.target_norm,
                      mlp_batch=args.mlp_batch, train_epochs=args.train_epochs, lr=args.lr,
                      boot_epochs=args.boot_epochs, key_weights=args.key_weights,
                      **({"arms": tuple(args.arms.split(","))} if args.arms else {}))
    ab.pop("_pretrained_state")
    sur = ab["arms"]["surrogate"]
    line = {"metric": METRIC, "value": sur["gens_per_hour"], "unit": "GA generations/hour",
            "n_gpus": 1, "steps": args.gens * args.reps, "warmup": 0, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 sim, bf16 surrogate", "data": "synthetic",
            "config": {"workload": f"C4: pool {args.pool} Gaussian mutants (sigma 5% of range) of a {args.elite} "
                                   f"elite, surrogate flies the top {args.fly}; 12 environments, gusts <= 4 m/s RMS, "
                                   f"12 s cap; {args.reps} GA seeds", "parallelism": "single GPU"},
            "surrogate_ab": ab}
    print(json.dumps(line), flush=True)


def roofline_of(stats_list):
    """Dominant-kernel roofline from pr_stats dicts: the fused flight kernel's algorithmic
    FLOPs (68 per panel-step + 222 per plane-step) over its HIP-event-timed launches."""
    fl = {"flight_steps": 0, "panel_steps": 0}
    ms, nl = 0.0, 0
    for st in stats_list:
        ms += st["fused_ms"]
        nl += st["fused_launches"]
        for b in st["buckets"]:
            if b["fused"]:
                fl["flight_steps"] += b["flight_steps"]
                fl["panel_steps"] += b["panel_steps"]
    flops = FLOP_PANEL * fl["panel_steps"] + FLOP_PLANE * fl["flight_steps"]
    achieved = flops / max(ms * 1e-3, 1e-12) / 1e12
    return {"bound": "valu", "kernel": "k_fly_all", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
            "avg_launch_ms": round(ms / max(1, nl), 4), "launches": nl, "flops_per_launch": round(flops / max(1, nl))}


def audit_bench(PA, args):
    """C5 as its own bench line."""
    print(json.dumps(audit_run(PA, args.steps, args.warmup, args.cpu_seconds)), flush=True)


def audit_run(PA, steps, warmup, cpu_seconds, oracle_check=None):
    """C5: robustness audit of a champion + 1 024 designs on 4 096 scenarios each (seed 9001 +
    s*7919 by scenario), gusts <= 5 m/s RMS, warp +-1 deg, 20 s cap; one step = one batched
    audit (prepare, scenarios, every flight, per-design statistics on the host).  The CPU leg
    (oracle_check, default: when a CPU baseline is asked for) runs the oracle: the FP64 check
    of the audit champion and the CPU baseline (cpu_seconds > 0)."""
    n_gen, n_scen = 1025, 4096
    genomes, res, _ = PA.family_population(n_gen, seed=9)
    params = PA.SimParams.config(gust_rms_max=5.0, warp_max_deg=1.0, wind_sd=2.0, max_steps=12000)
    pl = PA.Planes(n_gen, n_scen, n_gen * n_scen, max_ch=int(res[:, 0].max()), max_sp=int(res[:, 1].max()))
    pl.upload_genomes(genomes, res)
    for _ in range(warmup):
        pl.audit(None, n_scen, 9001, params, upload=False)
    PA.LIB.pr_device_sync()
    fsteps, stats = 0, []
    t0 = time.perf_counter()
    for _ in range(steps):
        st, dist = pl.audit(None, n_scen, 9001, params, upload=False)
        s = pl.stats()
        stats.append(s)
        fsteps += s["flight_steps"]
    wall = time.perf_counter() - t0
    mean_col = PA.AUDIT_FIELDS.index("mean")
    best = int(np.argmax(st[:, mean_col]))
    log(f"{steps} audits of {n_gen} x {n_scen} in {wall:.3f} s; best mean {st[best, mean_col]:.2f} m "
        f"(design {best}), failures {int(st[best, PA.AUDIT_FIELDS.index('failures')])}")
    value = fsteps / 600.0 / wall
    rank = surrogate_ranking(PA, genomes, res, dist, params)
    log(f"surrogate ranking of the audited designs: {rank}")
    pl.close()
    cpu, f64 = None, None
    if oracle_check if oracle_check is not None else cpu_seconds > 0:
        f64 = audit_fp64_check(genomes, res, best, dist)
        log(f"FP64 check {f64}")
    if cpu_seconds > 0:
        wc5 = dict(gust=5.0, warp_deg=1.0, wind_sd=2.0, max_steps=12000, nseeds=n_scen)
        cpu = cpu_baseline(genomes, res, wc5, cpu_seconds)
        log(f"cpu baseline {cpu['value']:.1f} pfs/s")
    out = {"metric": METRIC, "value": round(value, 2), "unit": "plane-flight-s/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(1000 * wall / steps, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "C5: robustness audit, 1025 designs (family mix) x 4096 scenarios (seed 9001 + "
                                  "s*7919), gusts <= 5 m/s RMS, warp +-1 deg, 20 s cap",
                      "designs": n_gen, "scenarios": n_scen, "max_steps": 12000, "parallelism": "single GPU"},
           "audits_per_hour": round(3600.0 * steps / wall, 1), "roofline": roofline_of(stats),
           "cpu_baseline": cpu, "fp64_check": f64, "surrogate_ranking": rank}
    return out


def robust_stats(samples):
    """Per design, from its scenario distances: [mean, CVaR of the worst quarter, worst,
    all-oversize flag] (the surrogate's targets) and the robust score 0.65 mean + 0.25 CVaR
    + 0.10 worst (sim.cu:340-357)."""
    x = np.asarray(samples, np.float64)
    srt = np.sort(x, axis=1)
    tail = (x.shape[1] + 3) // 4
    mean, cvar, worst = x.mean(1), srt[:, :tail].mean(1), srt[:, 0]
    over = np.all(x == -100.0, axis=1).astype(np.float64)
    return np.stack([mean, cvar, worst, over], 1).astype(np.float32), 0.65 * mean + 0.25 * cvar + 0.10 * worst


def surrogate_ranking(PA, genomes, res, dist, params, n_train=400000, epochs=10):
    """The audited designs ranked by the surrogate against their simulated ranks (robust
    score over every audit scenario).  The network is trained on sim flights of a separate
    population from the same family generator (n_train designs x 8 seeds)."""
    g_tr, r_tr, _ = PA.family_population(n_train, seed=11)
    seeds = seeds_for(500, 8)
    pl = PA.Planes(len(g_tr), len(seeds), len(g_tr) * len(seeds), int(r_tr[:, 0].max()), int(r_tr[:, 1].max()))
    fit, _ = pl.evaluate(g_tr, seeds, params, res=r_tr)
    pl.close()
    tg, _ = robust_stats(np.asarray(fit).reshape(len(g_tr), len(seeds)))
    S = PA.Surrogate(max_batch=4096, seed=3)
    t0 = time.perf_counter()
    loss = S.train(g_tr, tg, epochs=epochs, batch=1024, lr=2e-3)
    train_s = time.perf_counter() - t0
    pred = S.predict(genomes)[:, 4].astype(np.float64)
    _, sim = robust_stats(dist)
    ra = np.argsort(np.argsort(pred, kind="stable"), kind="stable")
    rb = np.argsort(np.argsort(sim, kind="stable"), kind="stable")
    top = lambda v, k: set(np.argsort(-v, kind="stable")[:k].tolist())
    return {"train_pairs": int(len(g_tr)), "train_seeds": int(len(seeds)), "epochs": epochs, "train_s": round(train_s, 2),
            "final_loss": round(float(loss), 4), "spearman": round(float(np.corrcoef(ra, rb)[0, 1]), 4),
            "mean_abs_rank_err": round(float(np.mean(np.abs(ra - rb))), 1),
            "top10_overlap": len(top(pred, 10) & top(sim, 10)), "top100_overlap": len(top(pred, 100) & top(sim, 100)),
            "sim_best_design": int(np.argmax(sim)), "sim_best_rank_by_surrogate": int(len(sim) - 1 - ra[np.argmax(sim)])}


def audit_fp64_check(genomes, res, best, dist, n_scen=256):
    """The oracle checking the GPU results (beside the CPU baseline): the audit champion's first n_scen
    scenarios in FP64 against the batched GPU audit's distances, per scenario."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    g = np.ascontiguousarray(genomes[best:best + 1], np.float32)
    rs = np.ascontiguousarray(res[best:best + 1], np.int32)
    seeds = (9001 + 7919 * np.arange(n_scen, dtype=np.int64)).astype(np.int32)
    f64, _, wall = O.fly_res(g, rs, np.zeros(n_scen, np.int32), seeds, nthreads=os.cpu_count() or 1, prec=64,
                             preset=1, gust=5.0, warp=float(np.deg2rad(1.0)), wsd=2.0, max_steps=12000)
    gpu = np.asarray(dist[best, :n_scen], np.float64)
    rel = np.abs(gpu - f64) / np.maximum(np.abs(np.asarray(f64, np.float64)), 1e-3)
    return {"design": int(best), "scenarios": n_scen, "within_1pct": round(float(np.mean(rel < 0.01)), 4),
            "median_rel_err": float(f"{np.median(rel):.3g}"), "p99_rel_err": float(f"{np.quantile(rel, 0.99):.3g}"),
            "fp64_wall_s": round(wall, 1)}


def generation_run(PA, wname, args, steps, warmup, cpu_seconds, oracle_check=None):
    """C1-C3: `steps` timed GA-generation evaluations (prepare, scenarios, every flight, the
    robust scores, the top-8 x 128 holdout) of one population after `warmup` untimed ones;
    returns the bench line (without the surrogate A/B).  The CPU leg after the timed region
    (oracle_check, default: when a CPU baseline is asked for) runs the oracle: the FP64 check of
    the timed results and the CPU baseline (cpu_seconds > 0)."""
    w = WORKLOADS[wname]
    res = None
    if w.get("families"):
        genomes, res, _ = PA.family_population(w["pop"], seed=0)
    else:
        lo, hi = PA.gene_bounds()
        rng = np.random.default_rng(0)
        genomes = (lo + rng.random((w["pop"], 35), dtype=np.float32) * (hi - lo)).astype(np.float32)
    params = PA.SimParams.config(gust_rms_max=w["gust"], warp_max_deg=w["warp_deg"], wind_sd=w["wind_sd"],
                                 max_steps=w["max_steps"])
    max_ch = int(res[:, 0].max()) if res is not None else 0
    max_sp = int(res[:, 1].max()) if res is not None else 0
    E = PA.Generation(w["pop"], w["nseeds"], max_ch, max_sp)
    E.upload(genomes, res)
    if args.unfused:
        E.set_fused(False)
    if args.list_order != "size":
        E.set_list_order(False)
    if args.no_split:
        E.set_split(False)
    elif args.split_holdout:
        E.set_split(True, holdout=True)
    if args.graph is None:
        args.graph = w.get("graph", False)
    if args.graph:
        E.set_graph(True)
    if args.tail_mode is not None:
        E.set_tail_mapping(args.tail_mode)
    tuned = None
    if args.buckets:
        spec = args.buckets if args.buckets.lstrip().startswith("{") else open(args.buckets).read()
        b = json.loads(spec)
        tuned = (list(b["nt_max"]), list(b["L"]), list(b["P"]))
        E.set_buckets(*tuned)
    elif args.autotune if args.autotune is not None else w.get("autotune", False):
        # one-time setup outside the timed region: the mapping per triangle-count bucket from
        # timing every compiled (L, P) on a sample of this population's flights
        t0 = time.perf_counter()
        tuned = PA.autotune(genomes, res, seeds_for(20000, w["nseeds"]), params)
        E.set_buckets(*tuned)
        log(f"autotuned buckets {list(zip(*tuned))} in {time.perf_counter() - t0:.1f} s")

    for k in range(warmup):
        E.evaluate(w["pop"], seeds_for(10000 + k, w["nseeds"]), params)
    PA.LIB.pr_device_sync()
    fsteps = 0
    bucket = {}
    fused = {"flight_steps": 0, "panel_steps": 0, "ms": 0.0, "launches": 0}
    fused_graph = {"flight_steps": 0, "panel_steps": 0, "ms": 0.0, "launches": 0}
    t0 = time.perf_counter()
    pipelined = not args.no_pipeline
    for k in range(steps):
        if pipelined:
            # generation k's holdout flies on its own stream while generation k+1's
            # population flies; the last holdout is collected inside the timed region
            E.evaluate_async(w["pop"], seeds_for(k, w["nseeds"]), params)
            st_list = [E.stats()[0]]
            if k > 0:
                r = E.collect()
                st_list.append(E.stats()[1])
            if k == steps - 1:
                r = E.collect()
                st_list.append(E.stats()[1])
        else:
            r = E.evaluate(w["pop"], seeds_for(k, w["nseeds"]), params)
            st_list = E.stats()
        for i, st in enumerate(st_list):
            fsteps += st["flight_steps"]
            for b in st["buckets"]:
                key = (b["L"], b["P"], b["tail"])
                acc = bucket.setdefault(key, {"flight_steps": 0, "panel_steps": 0, "fly_ms": 0.0, "launches": 0,
                                              "fused": b["fused"]})
                for f in ("flight_steps", "panel_steps", "fly_ms", "launches"):
                    acc[f] += b[f]
            if i > 0 and pipelined:
                continue  # a holdout batch overlapping other work: its launch times include sharing
            fused["ms"] += st["fused_ms"]
            fused["launches"] += st["fused_launches"]
            fused_graph["ms"] += st["fly_ms"]
            fused_graph["launches"] += 1
            fused_graph["flight_steps"] += st["flight_steps"]
            fused_graph["panel_steps"] += st["panel_steps"]
            for b in st["buckets"]:
                if b["fused"]:
                    fused["flight_steps"] += b["flight_steps"]
                    fused["panel_steps"] += b["panel_steps"]
            # small chunks: k_fly_split, timed inside the fused kernel's event bracket
            fused["flight_steps"] += st["split_flight_steps"]
            fused["panel_steps"] += st["split_panel_steps"]
            fused["split_steps"] = fused.get("split_steps", 0) + st["split_flight_steps"]
    PA.LIB.pr_device_sync()
    wall = time.perf_counter() - t0
    # the last timed generation's distances for its first genomes, for the FP64 check below
    # (still on the device; read after the timed region)
    timed_samples = E.read_samples(0, min(256, w["pop"]), w["nseeds"])
    timed_seeds = seeds_for(steps - 1, w["nseeds"])
    pfs = fsteps / 600.0
    value = pfs / wall
    ms_per_step = 1000.0 * wall / steps
    log(f"{steps} generations in {wall:.3f} s, {pfs:.0f} plane-flight-s, last holdout {r['holdout']:.3f}")

    # dominant kernel: the fused flight kernel (one launch per 1-s chunk) unless every
    # bucket was flown by its own launches; algorithmic FLOPs per launch / its HIP-event
    # timed average duration on the context stream. Pipelined runs time the population
    # launches only (the holdout batches run concurrently on their own stream).
    if args.graph:
        # replayed graphs carry no per-launch events: the graph's own events time the whole
        # flight sequence of each population evaluation (its k_init and list compactions
        # are < 0.2 % of it)
        name, fl, ms, nl = "k_fly_all (graph replays)", fused_graph, fused_graph["ms"], fused_graph["launches"]
    elif fused["launches"]:
        name, fl, ms, nl = "k_fly_all", fused, fused["ms"], fused["launches"]
        if fused.get("split_steps"):
            name += " + k_fly_split (small chunks)"
    else:
        key, fl = max(bucket.items(), key=lambda kv: kv[1]["fly_ms"])
        name, ms, nl = f"k_fly<{key[0]},{key[1]}>" + (" (tail mapping)" if key[2] else ""), fl["fly_ms"], fl["launches"]
    flops = FLOP_PANEL * fl["panel_steps"] + FLOP_PLANE * fl["flight_steps"]
    achieved = flops / (ms * 1e-3) / 1e12
    if pipelined and fused["launches"]:
        name += " (population flights)"
    roof = {"bound": "valu", "kernel": name, "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": None, "avg_launch_ms": round(ms / max(1, nl), 4), "launches": nl,
            "flops_per_launch": round(flops / max(1, nl))}
    if name.startswith("k_fly_all") and wname in TRAFFIC:
        roof["traffic"], roof["traffic_source"] = TRAFFIC[wname]
    log(f"  k_fly_all: {fused['launches']} launches, {fused['ms']:.2f} ms")
    last = E.stats()[0]
    for k, ch in enumerate(last["chunks"]):
        fl = FLOP_PANEL * ch["panel_steps"] + FLOP_PLANE * ch["flight_steps"]
        rate = f"{fl / ch['ms'] / 1e9:.1f} TFLOP/s" if ch["ms"] > 0 else "untimed (graph replay)"
        log(f"  last generation chunk {k}: {ch['flights']} flights, {ch['flight_steps']} flight-steps, "
            f"{ch['ms']:.2f} ms, {rate}")
    for kk, bb in sorted(bucket.items()):
        log(f"  {'(fused) ' if bb['fused'] else ''}k_fly<{kk[0]},{kk[1]}>{' tail' if kk[2] else ''}: "
            f"{bb['launches']} own launches, {bb['fly_ms']:.2f} ms, "
            f"{bb['flight_steps']} plane-steps, {bb['panel_steps']} panel-steps")

    cpu, f64 = None, None
    if oracle_check if oracle_check is not None else cpu_seconds > 0:
        f64 = fp64_check(genomes, res, w, timed_samples, timed_seeds)
        log(f"FP64 check: {f64}")
    if cpu_seconds > 0:
        log("cpu baseline ...")
        cpu = cpu_baseline(genomes, res, w, cpu_seconds)
        log(f"cpu baseline {cpu['value']:.1f} pfs/s on {cpu['cores']} threads")

    E.close()
    out = {"metric": METRIC, "value": round(value, 2), "unit": "plane-flight-s/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(ms_per_step, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{wname.upper()}: {w['desc']}", "pop": w["pop"], "seeds": w["nseeds"],
                      "max_steps": w["max_steps"], "holdout": "top-8 x 128 scenarios", "parallelism": "single GPU",
                      "buckets": [list(b) for b in zip(*tuned)] if tuned else "default table",
                      "launch": "hipGraph replay per evaluation" if args.graph else "direct launches"},
           "gens_per_hour_eval": round(3600.0 / (ms_per_step / 1000.0), 1),
           "roofline": roof, "cpu_baseline": cpu, "fp64_check": f64}
    return out


def generation_options(**kw):
    """The c1-c3 run options of the command line, at their defaults unless given."""
    o = dict(unfused=False, list_order="size", no_split=False, split_holdout=False, graph=None, tail_mode=None,
             buckets=None, autotune=None, no_pipeline=False)
    o.update(kw)
    return argparse.Namespace(**o)


def other_workloads(PA, oracle_check):
    """C2 and C5 measured in the same process after the C3 line's own work, at their default
    settings (C2: 10 timed generations after 2 warm-up ones, graph replays; C5: 2 timed audits
    after 1), each with its dominant-kernel roofline and, in the CPU leg (oracle_check), an FP64
    check of its own timed results."""
    out = {}
    t0 = time.perf_counter()
    c2 = generation_run(PA, "c2", generation_options(), 10, 2, 0, oracle_check)
    keep = ("value", "unit", "steps", "warmup", "ms_per_step", "gens_per_hour_eval", "roofline", "fp64_check")
    out["c2"] = {k: c2[k] for k in keep}
    out["c2"]["workload"] = c2["config"]["workload"]
    out["c2"]["wall_s"] = round(time.perf_counter() - t0, 1)
    t0 = time.perf_counter()
    c5 = audit_run(PA, 2, 1, 0, oracle_check)
    keep = ("value", "unit", "steps", "warmup", "ms_per_step", "audits_per_hour", "roofline", "fp64_check",
            "surrogate_ranking")
    out["c5"] = {k: c5[k] for k in keep}
    out["c5"]["workload"] = c5["config"]["workload"]
    out["c5"]["wall_s"] = round(time.perf_counter() - t0, 1)
    log(f"other workloads: C2 {out['c2']['value']:.0f} pfs/s, C5 {out['c5']['value']:.0f} pfs/s")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c4", "c5"])
    ap.add_argument("--pool", type=int, default=1_000_000, help="c4: candidate pool per generation")
    ap.add_argument("--fly", type=int, default=100_000, help="c4: candidates flown after pre-screening")
    ap.add_argument("--elite", type=int, default=10_000, help="c4: elite size")
    ap.add_argument("--gens", type=int, default=6, help="c4: generations of the plain arm (and of the equal-"
                                                          "generation champion comparison)")
    ap.add_argument("--sur-gens", type=int, default=None,
                    help="c4: generations of the surrogate arm (default 3 x --gens), for the time it takes to "
                         "reach the plain arm's champion")
    ap.add_argument("--buffer", type=int, default=300_000, help="c4: surrogate replay buffer")
    ap.add_argument("--explore", type=float, default=0.5, help="c4: fraction of the fly set drawn at random")
    ap.add_argument("--target-norm", action=argparse.BooleanOptionalAction, default=True,
                    help="surrogate targets of each batch relative to its medians (the scenario set's common "
                         "effect removed); --no-target-norm: absolute targets")
    ap.add_argument("--boot-epochs", type=int, default=16, help="c4: surrogate epochs on the bootstrap generation")
    ap.add_argument("--train-epochs", type=int, default=None,
                    help="c4: surrogate epochs per generation (default 2 after pretraining, else 4)")
    ap.add_argument("--lr", type=float, default=None,
                    help="c4: surrogate learning rate (Adam; default 1e-4 fine-tuning after pretraining, else 2e-3)")
    ap.add_argument("--reps", type=int, default=5, help="c4: GA seeds per arm")
    ap.add_argument("--arms", default=None,
                    help="c4: comma-separated arms to run (plain,surrogate,strict,uniform,parent; default all)")
    ap.add_argument("--key-weights", type=lambda t: tuple(float(x) for x in t.split(",")), default=None,
                    help="c4: surrogate arm's key weights a,b,c: a * parent score + b * prediction + "
                         "c * parent's prediction (default 1,1,-1)")
    ap.add_argument("--mlp-batch", type=int, default=4096, help="c4: surrogate training batch")
    ap.add_argument("--pretrain-source", choices=["c2", "family"], default="c2",
                    help="surrogate pretraining pairs: config 2 GA runs (the configuration's source) or the family "
                         "generator")
    ap.add_argument("--pretrain", type=int, default=2_000_000,
                    help="c4: sim pairs the surrogate is pretrained on (setup; 0: bootstrap generation instead)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget (0 = skip)")
    ap.add_argument("--surrogate-ab", action=argparse.BooleanOptionalAction, default=None,
                    help="c3: after the timed generations, the surrogate A/B at 100k x 12 (plain vs surrogate "
                         "gens/hour, time to the plain champion, and the same fly budget chosen without the "
                         "network) in the same line (default on for c3)")
    ap.add_argument("--ab-reps", type=int, default=3, help="surrogate A/B in the c3 line: GA seeds per arm")
    ap.add_argument("--ab-gens", type=int, default=4, help="surrogate A/B in the c3 line: plain-arm generations")
    ap.add_argument("--ab-sur-gens", type=int, default=None,
                    help="surrogate A/B in the c3 line: generations of the other arms (default 3 x --ab-gens)")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=None,
                    help="c1-c3: flight-kernel mapping per bucket from the autotuner (setup, untimed; "
                         "default on for c3, whose configuration asks for it)")
    ap.add_argument("--buckets", default=None,
                    help="c1-c3: flight-kernel mapping per bucket, as the autotuner's JSON "
                         "({\"nt_max\": [...], \"L\": [...], \"P\": [...]}), inline or a file path; replaces the autotuner")
    ap.add_argument("--graph", action=argparse.BooleanOptionalAction, default=None,
                    help="c1-c3: each evaluation's flight sequence (k_init, chunks, compactions) as a replayed hipGraph "
                         "(default on for c2, whose configuration asks for it; the roofline then times whole graphs)")
    ap.add_argument("--tail-mode", type=int, default=None, choices=[0, 1, 2],
                    help="c1-c3: tail-mapping mode of every flight context (pr_set_tail_mapping)")
    ap.add_argument("--no-split", action="store_true",
                    help="c1-c3: fly small chunks like the others instead of one flight per 2- or 4-wave workgroup")
    ap.add_argument("--split-holdout", action="store_true",
                    help="c1-c3: the small-chunk mapping for the holdout batches too (default: population only)")
    ap.add_argument("--list-order", choices=["size", "flight"], default="size",
                    help="c1-c3: chunk 0's work lists with genomes in order of mesh size (default) or flight order")
    ap.add_argument("--unfused", action="store_true", help="one flight launch per bucket instead of the fused kernel")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="c1-c3: finish each generation's holdout before the next generation starts")
    ap.add_argument("--other-workloads", action=argparse.BooleanOptionalAction, default=None,
                    help="c3: after the surrogate A/B, C2 (10 timed generations) and C5 (2 timed audits) in the "
                         "same line, each with its roofline and FP64 check (default on for c3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout "
              f"(all-gather of fitness samples over RCCL) is not run here", file=sys.stderr)
        sys.exit(2)

    import planes_amd as PA

    if args.workload == "c4":
        log("device:", PA.device_name())
        return pool_ab(PA, args)
    if args.workload == "c5":
        log("device:", PA.device_name())
        return audit_bench(PA, args)
    log("device:", PA.device_name())
    out = generation_run(PA, args.workload, args, args.steps, args.warmup, args.cpu_seconds)
    ab = None
    if args.surrogate_ab if args.surrogate_ab is not None else args.workload == "c3":
        t0 = time.perf_counter()
        ab = surrogate_ab(PA, reps=args.ab_reps, gens=args.ab_gens, sur_gens=args.ab_sur_gens,
                          pretrain_source=args.pretrain_source, target_norm=args.target_norm)
        ab.pop("_pretrained_state")
        ab["wall_s"] = round(time.perf_counter() - t0, 1)
        log(f"surrogate A/B: {json.dumps(ab['arms'].get('surrogate', {}))[:400]}")
    if ab is not None:
        out["surrogate_ab"] = ab
    if args.other_workloads if args.other_workloads is not None else args.workload == "c3":
        out["other_workloads"] = other_workloads(PA, args.cpu_seconds > 0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
