"""The reference-compatible CLI and file formats (main.cu:36-105, sim.cu:250-520) through
the planes_amd binary: evolve writes history.csv / best.bin / pop.bin / archive.bin with the
reference's layouts and checkpoints, is deterministic, warm-starts from pop.bin; evaluate
writes the 28-column robustness CSV; select-best writes the finalist CSV and best-final.bin."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "planes-rl_amd", "planes_amd")

pytestmark = pytest.mark.gpu


def run(*args, cwd):
    r = subprocess.run([BIN, *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def read_pop(path):
    b = open(path, "rb").read()
    n = struct.unpack("<i", b[:4])[0]
    assert len(b) == 4 + 140 * n
    return np.frombuffer(b[4:], np.float32).reshape(n, 35)


def test_evolve_outputs_and_determinism(tmp_path):
    out1, out2 = tmp_path / "a", tmp_path / "b"
    for o in (out1, out2):
        s = run("evolve", "--pop", "64", "--gens", "26", "--seeds", "2", "--out", str(o), cwd=tmp_path)
    assert "best holdout score" in s
    hist = (out1 / "history.csv").read_text().splitlines()
    assert hist[0] == "gen,train_best,holdout128,mean,p90" and len(hist) == 27
    for f in ("history.csv", "best.bin", "pop.bin", "archive.bin"):
        assert (out1 / f).read_bytes() == (out2 / f).read_bytes(), f
    assert len((out1 / "best.bin").read_bytes()) == 140
    pop = read_pop(out1 / "pop.bin")
    arch = read_pop(out1 / "archive.bin")
    assert pop.shape == (64, 35) and arch.shape == (26, 35)
    lo, hi = O.bounds()
    assert np.all(pop >= lo) and np.all(pop <= hi)
    # warm start from the saved population (clamped to bounds)
    s = run("evolve", "--pop", "64", "--gens", "2", "--seeds", "2", "--out", str(tmp_path / "c"), "--init",
            str(out1 / "pop.bin"), cwd=tmp_path)
    assert "warm start: 64 genomes" in s


def test_evaluate_csv_and_inspect(tmp_path):
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_presets.npz"), allow_pickle=False)["genomes"][0]
    best = tmp_path / "best.bin"
    best.write_bytes(g.astype(np.float32).tobytes())
    s = run("evaluate", "--best", str(best), "--n", "64", "--seed", "9001", "--o", str(tmp_path / "r.csv"),
            cwd=tmp_path)
    assert "robustness audit: 64 GPU scenarios" in s and "CVaR10" in s
    lines = (tmp_path / "r.csv").read_text().splitlines()
    assert len(lines) == 65 and len(lines[0].split(",")) == 28
    assert lines[0].startswith("scenario,seed,distance_m,speed_mps")
    row = lines[1].split(",")
    assert int(row[1]) == 9001
    sc = O.scenario(9001)
    assert abs(float(row[3]) - sc[0]) < 1e-4  # speed column = the scenario's launch speed
    s = run("inspect", "--best", str(best), cwd=tmp_path)
    m, printable, _, _ = O.mass(g)
    assert "printable" in s and f"{m[0] * 1000:6.1f}g" in s


def test_select_best(tmp_path):
    gs = np.load(os.path.join(ROOT, "tests", "golden", "ref_presets.npz"), allow_pickle=False)["genomes"]
    pop = tmp_path / "cands.bin"
    pop.write_bytes(struct.pack("<i", len(gs)) + gs.astype(np.float32).tobytes())
    s = run("select-best", "--pop-file", str(pop), "--n", "128", "--seed", "1900001", "--out", str(tmp_path / "o"),
            cwd=tmp_path)
    assert "independent finalist winner" in s
    rows = (tmp_path / "o" / "finalist-selection.csv").read_text().splitlines()
    assert rows[0] == "candidate,robust_score,mean,cvar25,worst" and len(rows) == 6
    scores = [float(r.split(",")[1]) for r in rows[1:]]
    win = (tmp_path / "o" / "best-final.bin").read_bytes()
    assert win == gs[int(np.argmax(scores))].astype(np.float32).tobytes()


def test_presets_seed_population_and_descriptions(tmp_path):
    """presets (designs.cu:58-73): seeds.bin holds the five preset genomes bit for bit in the
    pop.bin layout (a warm start for evolve --init), and each description line (designs.cu:45-54)
    equals the one formatted from the oracle's mesh and mass properties."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "ref_presets.npz"), allow_pickle=False)
    gs, names = d["genomes"], d["names"]
    s = run("presets", "--out", str(tmp_path / "p"), cwd=tmp_path)
    assert np.array_equal(read_pop(tmp_path / "p" / "seeds.bin"), gs)
    lines = s.splitlines()
    for i, (g, name) in enumerate(zip(gs, names)):
        m, printable, _, _ = O.mass(g)
        V, _ = O.mesh(g)
        e = V.max(0) - V.min(0)
        want = (f"{name:<18} {m[0] * 1000:6.1f}g  {e[0] * 1000:3.0f}x{e[1] * 1000:3.0f}x{e[2] * 1000:3.0f} mm  "
                f"CG=({m[2] * 1000:5.1f},{m[3] * 1000:5.1f},{m[4] * 1000:5.1f}) mm  "
                f"{'printable' if printable else 'TOO LARGE'}")
        assert lines[i] == want, (lines[i], want)
    assert "preset seed population ->" in s
    s = run("evolve", "--pop", "16", "--gens", "1", "--seeds", "2", "--out", str(tmp_path / "e"), "--init",
            str(tmp_path / "p" / "seeds.bin"), cwd=tmp_path)
    assert "warm start: 5 genomes" in s
