"""The drop-in boundary at the source level: the reference's own CLI (main.cu + designs.cu +
stl.cu, compiled in place from the reference checkout by oracle/Makefile's `refcli` target)
builds with include/planes_sim.h standing in for src/sim.h and links against
libplanes_hip.so in place of sim.cu (sim.h:1-13).

CPU tests (this container, where the checkout exists): the header compiles after common.h
as sim.h and on its own; the built CLI takes run_ga / evaluate_design / select_finalist from
the library and runs its host-only modes (presets, inspect) with the oracle's results.
GPU test: the prebuilt CLI (it travels with the tree; the checkout does not) runs the GPU
modes -- evolve, evaluate, select-best -- and writes byte-identical files to planes_amd's,
which calls the same library entry points from its own main.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "planes_ref_cli")
PR_CLI = os.path.join(ROOT, "planes-rl_amd", "planes_amd")
LIB = os.path.join(ROOT, "planes-rl_amd", "libplanes_hip.so")
HEADER = os.path.join(ROOT, "include", "planes_sim.h")

needs_checkout = pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference checkout absent on this machine")


def presets():
    d = np.load(os.path.join(ROOT, "tests", "golden", "ref_presets.npz"), allow_pickle=False)
    return d["genomes"], d["names"]


def describe_line(name, g):
    """designs.cu:45-54's description line, formatted from the oracle's mesh and mass."""
    m, printable, _, _ = O.mass(g)
    V, _ = O.mesh(g)
    e = V.max(0) - V.min(0)
    return (f"{name:<18} {m[0] * 1000:6.1f}g  {e[0] * 1000:3.0f}x{e[1] * 1000:3.0f}x{e[2] * 1000:3.0f} mm  "
            f"CG=({m[2] * 1000:5.1f},{m[3] * 1000:5.1f},{m[4] * 1000:5.1f}) mm  "
            f"{'printable' if printable else 'TOO LARGE'}")


def read_pop(path):
    b = open(path, "rb").read()
    n = struct.unpack("<i", b[:4])[0]
    assert len(b) == 4 + 140 * n
    return np.frombuffer(b[4:], np.float32).reshape(n, 35)


@needs_checkout
def test_header_compiles_as_sim_h_after_common_h(tmp_path):
    """main.cu includes common.h, then sim.h (main.cu:1-2): with the header as sim.h, the
    Genome of common.h:61 is used and float8 is defined once.  On its own (no common.h) the
    header defines the same 140-byte record.  Both compile as C++17 with no diagnostics."""
    dropin = tmp_path / "dropin"
    dropin.mkdir()
    (dropin / "sim.h").write_text(f'#include "{HEADER}"\n')
    for src in ("main.cu", "designs.cu"):
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Werror", "-I", str(dropin), "-I-", "-I", REF_SRC,
                            "-x", "c++", os.path.join(REF_SRC, src)], capture_output=True, text=True)
        errs = [ln for ln in r.stderr.splitlines() if "obsolete option" not in ln]
        assert r.returncode == 0 and not errs, r.stderr
    alone = tmp_path / "alone.cpp"
    alone.write_text(f'#include "{HEADER}"\n#include "{HEADER}"\n'
                     "static_assert(sizeof(Genome) == 140 && sizeof(float8) == 32, \"layout\");\n"
                     "int main() { Genome g{}; g.center_chord = 1; return (int)g.span; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", str(alone)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


@needs_checkout
def test_reference_cli_links_against_library_and_runs_host_modes(tmp_path):
    """The reference's CLI, built in place against libplanes_hip.so, takes its GPU entry
    points from the library, and its host-only modes run: presets writes the five preset
    genomes bit for bit (the fixture) with the oracle's description lines; inspect prints the
    oracle's line for a best.bin; a renderer mode reports that it is unavailable."""
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "refcli"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    und = subprocess.run(["nm", "-DC", "--undefined-only", REF_CLI], capture_output=True, text=True).stdout
    defined = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True).stdout
    for sig in ("run_ga(int, int, int, char const*, char const*)",
                "evaluate_design(Genome const*, int, int, char const*)",
                "select_finalist(Genome const*, int, int, int, char const*)"):
        assert sig in und and sig in defined, sig
    needed = subprocess.run(["readelf", "-d", REF_CLI], capture_output=True, text=True).stdout
    assert "libplanes_hip.so" in needed

    gs, names = presets()
    out = subprocess.run([REF_CLI, "presets", "--out", str(tmp_path / "p")], cwd=tmp_path, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert np.array_equal(read_pop(tmp_path / "p" / "seeds.bin"), gs)
    desc = [ln for ln in out.stdout.splitlines() if not ln.startswith("stl:")]
    for i, (g, name) in enumerate(zip(gs, names)):
        assert desc[i] == describe_line(name, g), (desc[i], describe_line(name, g))

    best = tmp_path / "best.bin"
    best.write_bytes(gs[3].astype(np.float32).tobytes())
    out = subprocess.run([REF_CLI, "inspect", "--best", str(best)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == describe_line(str(best), gs[3])

    out = subprocess.run([REF_CLI, "hero", "--best", str(best)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "renderer is not part of this build" in out.stderr


@pytest.mark.gpu
def test_reference_cli_gpu_modes_match_planes_amd(tmp_path):
    """The reference's main.cu driving the library on the GPU: evolve (run_ga, sim.cu:393-520),
    evaluate (evaluate_design, sim.cu:258-312) and select-best (select_finalist, sim.cu:359-391)
    through the prebuilt reference CLI write the same bytes as planes_amd, whose own main calls
    the same entry points with the same arguments."""
    if not os.path.exists(REF_CLI):
        pytest.skip("reference CLI not built (oracle/Makefile refcli needs the reference checkout)")
    gs, _ = presets()
    best = tmp_path / "best.bin"
    best.write_bytes(gs[1].astype(np.float32).tobytes())
    cands = tmp_path / "cands.bin"
    cands.write_bytes(struct.pack("<i", len(gs)) + gs.astype(np.float32).tobytes())
    outs = {}
    for tag, exe in (("ref", REF_CLI), ("pr", PR_CLI)):
        d = tmp_path / tag
        d.mkdir()
        # output paths relative to the run directory, so that the printouts name the same files
        for args in (["evolve", "--pop", "64", "--gens", "3", "--seeds", "2", "--out", "ga"],
                     ["evaluate", "--best", str(best), "--n", "64", "--seed", "9001", "--o", "r.csv"],
                     ["select-best", "--pop-file", str(cands), "--n", "128", "--seed", "1900001", "--out", "sel"]):
            print(f"{tag}: {' '.join(args[:1])}", flush=True)
            r = subprocess.run([exe, *args], cwd=d, capture_output=True, text=True, timeout=300)
            assert r.returncode == 0, r.stderr
            outs.setdefault(tag, []).append(r.stdout)
    for rel in ("ga/history.csv", "ga/best.bin", "ga/pop.bin", "ga/archive.bin", "r.csv",
                "sel/finalist-selection.csv", "sel/best-final.bin"):
        a, b = (tmp_path / "ref" / rel).read_bytes(), (tmp_path / "pr" / rel).read_bytes()
        assert a == b, rel
    assert len((tmp_path / "ref" / "r.csv").read_text().splitlines()) == 65
    for k in range(3):  # run_ga's, evaluate_design's and select_finalist's printouts
        assert outs["ref"][k] == outs["pr"][k]
