"""Closed forms and invariants that pin the oracle's flight model where the reference
ships no vectors (its sim kernel is CUDA-only): FP32 vs FP64 convergence, ballistic
motion at zero air density, angular-momentum conservation without torque, normal-flip
invariance of the panel force law, flat-plate special cases, robust scores by hand."""
import numpy as np
import pytest

import oracle_lib as O

PRESETS = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_presets.npz"),
                  allow_pickle=False)["genomes"]


def _groups(a, b):
    out = []
    for sl in (slice(0, 3), slice(3, 7), slice(7, 10), slice(10, 13)):
        out.append(np.max(np.linalg.norm(a[:, sl] - b[:, sl], axis=1) /
                          np.maximum(np.linalg.norm(b[:, sl], axis=1), 1e-6)))
    return out


def test_fp32_matches_fp64_over_half_second():
    gs = np.concatenate([PRESETS, O.random_genomes(27, 5)])
    worst = 0.0
    for g in gs:
        for seed in (1, 2):
            a = O.trajectory(g, seed, 300, 32).astype(np.float64)
            b = O.trajectory(g, seed, 300, 64)
            n = min(len(a), len(b))
            if n < 300:
                continue
            worst = max(worst, max(_groups(a[:n], b[:n])))
    assert worst < 1e-4, worst


def test_ballistic_at_zero_density():
    g = PRESETS[0]
    sc = O.scenario(7).copy()
    sc[8] = 0.0  # air density -> no aerodynamic force
    tr = O.trajectory(g, 7, 400, 64, scenario_override=sc)
    st0, _, _, _ = O.launch(g, 7, prec=64)
    dt = float(np.float32(1.0) / np.float32(600.0))
    n = np.arange(1, len(tr) + 1, dtype=np.float64)
    vz = st0[9] - n * 9.81 * float(np.float32(9.81)) / 9.81 * dt
    z = st0[2] + n * st0[9] * dt - float(np.float32(9.81)) * dt * dt * n * (n + 1) / 2
    x = st0[0] + n * st0[7] * dt
    assert np.allclose(tr[:, 2], z, rtol=0, atol=1e-9)
    assert np.allclose(tr[:, 0], x, rtol=0, atol=1e-9)
    assert np.allclose(tr[:, 9], vz, rtol=0, atol=1e-9)


def test_angular_momentum_conserved_without_torque():
    """rho = 0: only the gyroscopic term acts; world-frame L = R I w is conserved to O(dt)."""
    g = PRESETS[3]
    sc = O.scenario(11).copy()
    sc[8] = 0.0
    sc[5:8] = [9.0, -7.0, 5.0]  # strong release rates
    tr = O.trajectory(g, 11, 600, 64, scenario_override=sc)
    _, mp, _, _ = O.launch(g, 11, prec=64, )
    # printed inertia with the overridden scenario (material/infill/bias unchanged)
    _, mp, _, _ = O.launch(g, 11, prec=64)
    I = np.array([[mp[5], mp[8], mp[9]], [mp[8], mp[6], mp[10]], [mp[9], mp[10], mp[7]]])

    def L(row):
        w, x, y, z = row[3:7]
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        return R @ (I @ row[10:13])

    L0 = L(tr[0])
    drift = max(np.linalg.norm(L(r) - L0) / np.linalg.norm(L0) for r in tr)
    # explicit Euler drifts O(dt) (~4e-3 here); a wrong-sign gyroscopic term would rotate L at
    # 2|w| ~ 26 rad/s and give an O(1) change
    assert drift < 1e-2, drift


def test_force_law_invariant_under_normal_flip():
    for g in O.random_genomes(6, 9):
        a = O.trajectory(g, 3, 300, 32)
        b = O.trajectory(g, 3, 300, 32, flip_normals=True)
        assert np.array_equal(a, b)


def _plate(alpha, rho=1.2, cf=0.02, V=10.0, A=0.01):
    c = np.zeros((1, 3))
    n = np.array([[0.0, 0.0, 1.0]])
    vel = np.array([-V * np.cos(alpha), 0.0, V * np.sin(alpha)])  # relative flow = vel - wind at the CG
    return O.aero(c, [A], n, np.zeros(3), [1, 0, 0, 0], -vel * 0 + vel, np.zeros(3), np.zeros(3), rho, cf)


def test_flat_plate_special_cases():
    rho, V, A, cf = 1.2, 10.0, 0.01, 0.02
    qd = 0.5 * rho * V * V * A
    # edge-on: pure skin friction along the flow
    F, T = _plate(0.0, rho, cf, V, A)
    assert np.allclose(F, [qd * cf, 0, 0], atol=1e-15) and np.allclose(T, 0)
    # face-on: normal coefficient 2 (2 s|s| at s = 1), no friction
    F, _ = _plate(np.pi / 2, rho, cf, V, A)
    assert np.allclose(F, [0, 0, -2 * qd], rtol=1e-12, atol=1e-15)
    # small angle: normal-force slope 1.9 per radian
    a = 1e-4
    F, _ = _plate(a, rho, 0.0, V, A)
    assert abs(-F[2] / (qd * a) - 1.9) < 1e-3


def test_panel_torque_is_offset_cross_force():
    c = np.array([[0.1, 0.05, -0.02]])
    n = np.array([[0.0, 0.6, 0.8]])
    vel = np.array([8.0, 1.0, -2.0])
    F, T = O.aero(c, [0.02], n, np.zeros(3), [1, 0, 0, 0], vel, np.zeros(3), np.zeros(3), 1.2, 0.03)
    assert np.allclose(T, np.cross(c[0], F), rtol=1e-12, atol=1e-18)


def test_robust_score_hand_vectors():
    # mean 2.5, tail (4+3)//4 = 1 -> CVaR = worst = 1
    assert np.isclose(O.robust_scores([4, 1, 3, 2], 1, 4)[0], 0.65 * 2.5 + 0.25 * 1 + 0.10 * 1)
    # 8 samples: tail 2
    s = np.array([5, -100, 7, 3, 9, 1, 2, 8], np.float32)
    expect = 0.65 * s.mean() + 0.25 * (-100 + 1) / 2 + 0.10 * -100
    assert np.isclose(O.robust_scores(s, 1, 8)[0], expect, rtol=1e-6)
    # 128-scenario holdout: tail 32
    h = np.arange(128, dtype=np.float32)
    assert np.isclose(O.robust_scores(h, 1, 128)[0], 0.65 * 63.5 + 0.25 * 15.5 + 0.0, rtol=1e-6)


def test_unprintable_and_nan_paths():
    g = PRESETS[0].copy()
    g[0] = 0.25
    g[19] = 0.65  # wide tail pushes the span beyond the print volume
    out, printable, nt, nv = O.mass(g)
    if not printable:
        fit, steps, _ = O.fly(g[None], [0], [1], prec=32)
        assert fit[0] == -100.0 and steps[0] == 0


def test_triangle_counts_closed_form():
    lo, hi = O.bounds()
    rng = np.random.default_rng(0)
    for g in (lo + rng.random((400, 35), dtype=np.float32) * (hi - lo)).astype(np.float32):
        lay = g[26]
        k = 1 if lay < 0.25 else (2 if lay < 0.5 else (3 if lay < 0.75 else 2))
        bi = g[13] > np.float32(0.02)
        V, T = O.mesh(g)
        assert len(T) == 174 + 2 * k + (164 if bi else 0)
        assert len(V) == 119 + 4 * k + (110 if bi else 0)
        assert T.min() >= 0 and T.max() < len(V)


def test_mesh_mirror_symmetry_and_geometry_test_properties():
    """tests/geometry_test.cpp:18-36 properties on the 5 presets."""
    for i, g in enumerate(PRESETS):
        V, T = O.mesh(g)
        grid = V[:102].reshape(17, 6, 3)
        assert np.allclose(grid[::-1, :, 0], grid[:, :, 0], atol=1e-6)
        assert np.allclose(grid[::-1, :, 1], -grid[:, :, 1], atol=1e-6)
        assert np.allclose(grid[::-1, :, 2], grid[:, :, 2], atol=1e-6)
        out, printable, nt, nv = O.mass(g)
        assert 0 < nv <= 512 and 0 < nt <= 900 and printable
        assert 0.005 < out[0] < 0.100 and np.all(out[5:8] > 0)
        pm, sh = O.printed(out, O.scenario(1000 + i))
        assert pm[0] > out[0] and pm[1] > 0 and pm[11] > 0 and pm[15] > 0 and pm[19] > 0
        assert np.linalg.norm(sh) < 0.02


def test_champion_stl_topology():
    """The reference's champion STL (out/shape-robust-run, 500 facets) is a triple-fin
    monoplane shell: 2 x 180 offset triangles + 2 x 70 boundary-edge quads."""
    g = PRESETS[4].copy()
    g[13] = 0.0   # monoplane
    g[26] = 0.6   # triple fin
    V, T = O.mesh(g)
    assert len(T) == 180
    from collections import Counter
    edges = Counter()
    for t in T:
        for k in range(3):
            a, b = int(t[k]), int(t[(k + 1) % 3])
            edges[(min(a, b), max(a, b))] += 1
    boundary = sum(1 for c in edges.values() if c == 1)
    assert 2 * len(T) + 2 * boundary == 500


def test_pose_recording_selects_the_reference_steps():
    """sim.cu:117-128: before step s a frame {pos, q, alive} is written when s % rec_every == 0,
    at index rec_i, which restarts at step0 / rec_every at every 600-step chunk (sim.cu:117; so
    with rec_every = 7 chunk 1's first frame, step 602, overwrites chunk 0's last, step 595) and
    grows by one per frame; after the flight ends the frames keep its final pose with alive = 0.
    The recorder's frames are compared with the oracle's separate trajectory loop (the state
    after every step) at the steps that rule selects."""
    steps, cap = 1800, 400
    for g, seed in ((PRESETS[1], 42), (PRESETS[0], 3)):
        tr = O.trajectory(g, seed, steps, 32, max_steps=steps)
        st0 = O.launch(g, seed, prec=32)[0]
        n = len(tr)
        for r in (6, 7):
            _, rec = O.record(g, seed, max_steps=steps, rec_every=r, rec_cap=cap)
            frame_step = {}
            for step0 in range(0, steps, 600):
                i = step0 // r
                for s in range(step0, min(step0 + 600, steps)):
                    if s % r == 0 and i < cap:
                        frame_step[i] = s
                        i += 1
            for i, s in frame_step.items():
                state = st0 if s == 0 else tr[min(s, n) - 1]
                assert np.array_equal(rec[i, :7], state[:7].astype(np.float32)), (seed, r, i, s)
                assert rec[i, 7] == (1.0 if s < n else 0.0), (seed, r, i, s)
            assert np.all(rec[max(frame_step) + 1:] == 0)  # slots no step selected stay untouched
    # the cases cover both flags: one flight ends inside the window, and frames past it are dead
    ends = [len(O.trajectory(g, s, steps, 32, max_steps=steps)) for g, s in ((PRESETS[1], 42), (PRESETS[0], 3))]
    assert min(ends) < steps
