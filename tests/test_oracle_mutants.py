"""Planted mistakes in the oracle (on the CPU): each mutant below is the oracle's source with
one plausible slip -- a sign, an operand order, a transposed rotation, a rate, a scale, a
range, an index, a shared lane -- compiled as its own library; the pins that should catch it
(tests/test_oracle_pins.py; tests/test_oracle_config_preset.py for the configurations'
scenario preset; tests/test_oracle_physics.py for the pose recorder) are run against that
build and must fail.  An unmodified build compiled
the same way passes them (the control)."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "planes_oracle.cpp")
PINS = os.path.join(ROOT, "tests", "test_oracle_pins.py")
CFG = os.path.join(ROOT, "tests", "test_oracle_config_preset.py")
CFG_FN = "static void scenario_cfg("  # config-preset mistakes are planted in its body
BM_FN = "static float lane_normal("
PHYS = os.path.join(ROOT, "tests", "test_oracle_physics.py")

GA2 = "void or_ga2_breed"  # breeding mistakes are planted in the persistent-distribution GA only

# name: (text to replace, replacement, section marker the text must follow or None, -k expression)
MUTANTS = {
    "pitch_sign": ("qfrom(mk<R>(0, 1, 0), -pitch)", "qfrom(mk<R>(0, 1, 0), pitch)", None, "attitude"),
    "force_pitch_sign": ("-(pitch + R(d[S_FPITCH]))", "-(pitch - R(d[S_FPITCH]))", None, "attitude"),
    "lever_arm_sign": ("crs(grip - F.shift, Jb)", "crs(grip + F.shift, Jb)", None, "grip"),
    "cross_order": ("crs(grip - F.shift, Jb)", "crs(Jb, grip - F.shift)", None, "grip"),
    "rotation_not_transposed": ("apply_t(qrows(s.q), mul(s.vel, F.mp.mass))", "apply(qrows(s.q), mul(s.vel, F.mp.mass))",
                                None, "grip"),
    "spin_clamp": ("if (w0 > R(35.0f)) s.wb = mul(s.wb, R(35.0f) / w0);",
                   "if (w0 > R(30.0f)) s.wb = mul(s.wb, R(30.0f) / w0);", None, "clamp"),
    "grid_column": ("oy = R((float)((plane_index % cols) - cols / 2) * sc.spread);",
                    "oy = R((float)((plane_index % cols) - cols / 2 + 1) * sc.spread);", None, "clamp"),
    "grand_range": ("LO[i] + u(rng) * (HI[i] - LO[i])", "LO[i] + u(rng) * HI[i]", None, "grand"),
    "elite_count": ("int elite = pop / 64;", "int elite = pop / 32;", GA2, "elite"),
    "immigrant_rate": ("if (G->u01(rng) < 0.05f)", "if (G->u01(rng) < 0.04f)", GA2, "identical_parents"),
    "mutation_rate": ("if (G->u01(rng) < 0.25f)", "if (G->u01(rng) < 0.20f)", GA2, "identical_parents"),
    "mutation_scale": ("G->nrm(rng) * 0.08f", "G->nrm(rng) * 0.07f", GA2, "identical_parents"),
    "anneal": ("float anneal = 1.0f - 0.6f * gen / gens;", "float anneal = 1.0f - 0.5f * gen / gens;", GA2,
               "identical_parents"),
    "missing_clamp": ("clampset(out, j, v);", "out[j] = v;", GA2, "identical_parents"),
    "blend_range": ("float a = -0.1f + 1.2f * G->u01(rng);", "float a = -0.1f + 1.1f * G->u01(rng);", GA2,
                    "tournament"),
    "tournament_of_three": ("for (int k = 0; k < 3; k++) { int c = (int)(G->u01(rng) * pop);",
                            "for (int k = 0; k < 2; k++) { int c = (int)(G->u01(rng) * pop);", GA2, "tournament"),
    "tournament_picks_worse": ("if (acc[c] > acc[b]) b = c;", "if (acc[c] < acc[b]) b = c;", GA2, "tournament"),
    "deck_u_step": ("R u = R((float)(i - rs.sp) / (float)rs.sp);", "R u = R((float)(i - rs.sp) / (float)(rs.sp + 1));",
                    None, "refinement"),
    "deck_xc_fixed": ("R xc = R((float)j / (float)rs.ch);", "R xc = R((float)j / 5.0f);", None, "refinement"),
    "deck_triangle_index": ("int a = base + i * (rs.ch + 1) + j, b = a + rs.ch + 1;",
                            "int a = base + i * (rs.ch + 1) + j, b = a + rs.ch;", None, "topology"),
    "mirror_sign": ("return mk<R>(x, (u < 0 ? -y : y), zc + zoff);", "return mk<R>(x, (u < 0 ? y : y), zc + zoff);",
                    None, "mirror"),
    # the configurations' scenario preset: tests/test_oracle_config_preset.py
    "cfg_speed_scale": ("9.2f + 2.0f * lane_pm1(seed, 0)", "9.2f + 2.5f * lane_pm1(seed, 0)", CFG_FN, "", CFG),
    "cfg_pitch_nominal": ("0.10f + 5.0f * deg * lane_normal(seed, 32)", "5.0f * deg * lane_normal(seed, 32)", CFG_FN,
                          "", CFG),
    "cfg_rates_in_degrees": ("30.0f * deg * lane_normal(seed, 40)", "30.0f * lane_normal(seed, 40)", CFG_FN, "", CFG),
    "cfg_gust_unscaled": ("lane_u01(seed, 23) * (1.0f / 0.69191763f)", "lane_u01(seed, 23)", CFG_FN, "", CFG),
    "cfg_wind_sd": ("cf.wind_sd * lane_normal(seed, 46)", "0.5f * cf.wind_sd * lane_normal(seed, 46)", CFG_FN, "",
                    CFG),
    "cfg_grip_lane": ("0.022f * lane_pm1(seed, 13)", "0.022f * lane_pm1(seed, 12)", CFG_FN, "", CFG),
    "cfg_cf_factor_dropped": ("d[S_CF] *= 0.90f + 0.25f * lane_u01(seed, 20);", "", CFG_FN, "", CFG),
    "cfg_warp_one_sided": ("cf.warp_max * lane_pm1(seed, 21)", "cf.warp_max * lane_u01(seed, 21)", CFG_FN, "", CFG),
    "cfg_shared_lane": ("d[S_FYAW] = 5.0f * deg * lane_normal(seed, 50);", "d[S_FYAW] = 5.0f * deg * lane_normal(seed, 34);",
                        CFG_FN, "", CFG),
    "box_muller_radius": ("std::sqrt(-2.0 * std::log(u1))", "std::sqrt(-std::log(u1))", BM_FN, "", CFG),
    "box_muller_angle_lane": ("std::cos(6.283185307179586 * u2)", "std::cos(6.283185307179586 * u1)", BM_FN, "", CFG),
    # pose recording (sim.cu:117-128): tests/test_oracle_physics.py
    "rec_index_ceil": ("int rec_i = step0 / rec_every;", "int rec_i = (step0 + rec_every - 1) / rec_every;", None,
                       "pose_recording", PHYS),
    "rec_alive_inverted": ("r[7] = s.done ? 0.0f : 1.0f;", "r[7] = s.done ? 1.0f : 0.0f;", None, "pose_recording", PHYS),
    "rec_one_step_late": ("if (rec && st % rec_every == 0 && rec_i < rec_cap)",
                          "if (rec && (st + 1) % rec_every == 0 && rec_i < rec_cap)", None, "pose_recording", PHYS),
}


def mutate(src, old, new, marker):
    start = src.index(marker) if marker else 0
    pos = src.index(old, start)
    return src[:pos] + new + src[pos + len(old):]


def build(src_text, out):
    cpp = out + ".cpp"
    with open(cpp, "w") as f:
        f.write(src_text)
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                        "-pthread", cpp, "-o", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def run_pins(so, kexpr, path=PINS):
    env = dict(os.environ, PLANES_ORACLE_SO=so)
    cmd = [sys.executable, "-m", "pytest", path, "-q", "-x", "-p", "no:cacheprovider"] + (["-k", kexpr] if kexpr else [])
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-2000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_planted_mistakes_are_caught(tmp_path):
    src = open(SRC).read()
    jobs = {"control": (src, "attitude or grip or clamp or grand or elite or identical_parents or tournament or "
                             "refinement or topology or mirror", PINS),
            "control_cfg": (src, "", CFG),
            "control_phys": (src, "pose_recording", PHYS)}
    for name, (old, new, marker, kexpr, *path) in MUTANTS.items():
        m = mutate(src, old, new, marker)
        assert m != src, name
        jobs[name] = (m, kexpr, path[0] if path else PINS)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        builds = {n: ex.submit(build, t, str(tmp_path / f"{n}.so")) for n, (t, _, _) in jobs.items()}
        runs = {n: ex.submit(lambda n=n: run_pins(builds[n].result(), jobs[n][1], jobs[n][2])) for n in jobs}
        res = {n: f.result() for n, f in runs.items()}
    for control in ("control", "control_cfg", "control_phys"):
        rc, out = res.pop(control)
        assert rc == 0, out
    survivors = [n for n, (rc, _) in res.items() if rc != 1]
    assert not survivors, {n: res[n][1] for n in survivors}
