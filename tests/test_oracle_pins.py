"""Pins of the oracle functions that the reference's own files leave unchecked: the seeded
launch (sim.cu:70-109), GA breeding (sim.cu:472-490, with grand() sim.cu:318-321 and the
mt19937 engine of sim.cu:398) and the mesh at resolutions other than CH=5, SP=8
(common.h:177-192 with CH/SP as parameters).  Each check comes from the mathematics or a
published definition, not from the oracle's own formulas:

- launch: elementary rotation matrices and quaternion rotation by the Hamilton product,
  the velocity direction from the pitch/yaw angles, the impulse law recomputed with
  numpy's inverse and cross product, the 35 rad/s clamp, the showcase grid;
- breeding: statistics of the offspring of populations built so that every operator's
  effect can be read off (identical parents; two parent types of unequal fitness), against
  the stated probabilities and distributions; the engine against the C++ standard's check
  value and numpy's independent MT19937; grand() against generate_canonical's definition;
- mesh: closed-form vertex/triangle counts and deck topology at any (ch, sp), mirror
  symmetry, and refinement: the (2ch, 2sp) mesh holds the (ch, sp) mesh's vertices bit for
  bit at the shared (u, xc) nodes.

A planted slip in the oracle (a sign in qaxis(y, -pitch), grip - shift, the cross product's
order, R against its transpose, the blend range, the mutation rate or scale, the anneal,
the tournament size or direction, the immigrant rate, a grid index) fails one of these.
"""
import numpy as np
import pytest
from scipy import stats

import oracle_lib as O

LO, HI = O.bounds()
RANGE = (HI.astype(np.float64) - LO.astype(np.float64))
PRESETS = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "ref_presets.npz"),
                  allow_pickle=False)["genomes"]

# scenario descriptor fields (common.h:377-387 order + base_z)
SPEED, PITCH, YAW, ROLL, HEIGHT, RR, PR, YR = range(8)
FPITCH, FYAW, GX, GY, GZ = 13, 14, 15, 16, 17
BX, BY, BZ = 20, 21, 22


# ---------------------------------------------------------------- launch (sim.cu:70-109)

def rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def hamilton(a, b):
    w1, v1, w2, v2 = a[0], np.asarray(a[1:]), b[0], np.asarray(b[1:])
    return np.concatenate([[w1 * w2 - v1 @ v2], w1 * v2 + w2 * v1 + np.cross(v1, v2)])


def rotate(q, v):
    """q (w, x, y, z) applied to v as q (0, v) q*."""
    qc = np.array([q[0], -q[1], -q[2], -q[3]])
    return hamilton(hamilton(q, np.concatenate([[0.0], v])), qc)[1:]


def rotate_inv(q, v):
    return rotate(np.array([q[0], -q[1], -q[2], -q[3]]), v)


def inertia(mp):
    return np.array([[mp[5], mp[8], mp[9]], [mp[8], mp[6], mp[10]], [mp[9], mp[10], mp[7]]], np.float64)


def override(seed, **fields):
    sc = O.scenario(seed).copy()
    for k, v in fields.items():
        sc[int(k[1:])] = v
    return sc


def test_launch_attitude_and_velocity_closed_forms():
    """q = qaxis(z, yaw) qaxis(y, -pitch) qaxis(x, roll) (sim.cu:86-89) is the rotation
    Rz(yaw) Ry(-pitch) Rx(roll): its body axes are that matrix's columns, so the nose points
    along (cos p cos y, cos p sin y, sin p); the velocity is speed along the force direction
    (pitch + force_pitch, yaw + force_yaw) (sim.cu:93-95); the position is (0, 0, height)."""
    g = PRESETS[2]
    cases = [dict(f1=0.35, f2=-0.4, f3=0.6, f13=0.1, f14=-0.2), dict(f1=-0.2, f2=0.9, f3=-1.1, f13=-0.05, f14=0.3),
             dict(f1=0.0, f2=0.0, f3=0.0, f13=0.0, f14=0.0)]
    seeds = [3, 17, 101, 5003]
    scs = [override(7, **c) for c in cases] + [O.scenario(s) for s in seeds]
    for sc in scs:
        st, mp, sh, nt = O.launch(g, 7, prec=64, scenario_override=sc)
        p, y, r = (float(np.float64(sc[k])) for k in (PITCH, YAW, ROLL))
        R = rz(y) @ ry(-p) @ rx(r)
        q = st[3:7]
        assert abs(np.linalg.norm(q) - 1) < 1e-14
        for k in range(3):
            assert np.allclose(rotate(q, np.eye(3)[k]), R[:, k], rtol=0, atol=1e-13), k
        nose = np.array([np.cos(p) * np.cos(y), np.cos(p) * np.sin(y), np.sin(p)])
        assert np.allclose(rotate(q, [1.0, 0, 0]), nose, rtol=0, atol=1e-13)
        th, ps = p + float(sc[FPITCH]), y + float(sc[FYAW])
        v = float(sc[SPEED]) * np.array([np.cos(th) * np.cos(ps), np.cos(th) * np.sin(ps), np.sin(th)])
        assert np.allclose(st[7:10], v, rtol=0, atol=1e-12)
        assert abs(np.linalg.norm(st[7:10]) - float(sc[SPEED])) < 1e-12
        assert np.array_equal(st[0:3], [0.0, 0.0, float(sc[HEIGHT])])


def test_launch_spin_without_and_with_grip_offset():
    """wb = release rates + I^-1 (0.12 (grip - cg_shift) x R^T (m v)), |wb| <= 35
    (sim.cu:96-107).  With no print bias the CG does not shift; a zero grip then leaves the
    release rates exactly, and a grip offset gives the impulse law recomputed here with numpy's
    inverse of the printed inertia and the inverse rotation by the conjugate quaternion."""
    g = PRESETS[0]
    base = dict(f1=0.25, f2=0.5, f3=-0.7, f13=0.05, f14=-0.1, f5=0.3, f6=-0.2, f7=0.5, f20=0.0, f21=0.0, f22=0.0)
    sc = override(11, **base, f15=0.0, f16=0.0, f17=0.0)
    st, mp, sh, nt = O.launch(g, 11, prec=64, scenario_override=sc)
    assert np.array_equal(sh, [0.0, 0.0, 0.0])
    assert np.array_equal(st[10:13], sc[[RR, PR, YR]].astype(np.float64))
    Iinv = np.linalg.inv(inertia(mp))
    for grip in ([0.02, 0.0, 0.0], [0.0, -0.015, 0.0], [0.0, 0.0, 0.008], [0.012, 0.01, -0.006]):
        sc = override(11, **base, f15=grip[0], f16=grip[1], f17=grip[2])
        st, mp2, sh, _ = O.launch(g, 11, prec=64, scenario_override=sc)
        assert np.array_equal(mp2, mp)
        gr = sc[[GX, GY, GZ]].astype(np.float64)
        Jb = rotate_inv(st[3:7], mp[0] * st[7:10])
        dw = Iinv @ (float(np.float32(0.12)) * np.cross(gr, Jb))
        want = sc[[RR, PR, YR]].astype(np.float64) + dw
        assert np.linalg.norm(want) < 35
        assert np.allclose(st[10:13], want, rtol=1e-9, atol=1e-12), (grip, st[10:13], want)
        assert np.linalg.norm(dw) > 0.05  # the offset's contribution is far above the tolerance
    # with a print bias the CG shifts and the lever arm is grip - shift
    sc = override(11, **{**base, "f20": 0.015, "f21": -0.008, "f22": 0.004}, f15=0.01, f16=0.0, f17=0.0)
    st, mp3, sh, _ = O.launch(g, 11, prec=64, scenario_override=sc)
    assert np.linalg.norm(sh) > 1e-3
    Jb = rotate_inv(st[3:7], mp3[0] * st[7:10])
    dw = np.linalg.inv(inertia(mp3)) @ (float(np.float32(0.12)) * np.cross(sc[[GX, GY, GZ]].astype(np.float64) - sh, Jb))
    assert np.allclose(st[10:13], sc[[RR, PR, YR]].astype(np.float64) + dw, rtol=1e-9, atol=1e-12)


def test_launch_spin_clamp_and_showcase_grid():
    """|wb| > 35 rad/s is scaled back to 35 along its direction (sim.cu:107); the showcase grid
    puts plane p at (-(p // 25) 1.4 s, (p % 25 - 12) s, height + 0.25 sin(12.9898 p)) (sim.cu:71-78)."""
    g = PRESETS[1]
    sc = override(5, f5=30.0, f6=-25.0, f7=20.0, f15=0.0, f16=0.0, f17=0.0, f20=0.0, f21=0.0, f22=0.0)
    st, _, _, _ = O.launch(g, 5, prec=64, scenario_override=sc)
    rates = sc[[RR, PR, YR]].astype(np.float64)
    assert abs(np.linalg.norm(st[10:13]) - 35.0) < 1e-12
    assert np.allclose(st[10:13], rates * 35.0 / np.linalg.norm(rates), rtol=0, atol=1e-12)
    for p in (0, 1, 12, 13, 24, 25, 49, 77, 310):
        s = 0.6
        st, _, _, _ = O.launch(g, 5, prec=64, plane_index=p, spread=s)
        h = float(O.scenario(5)[HEIGHT])
        arg = float(np.float32(p) * np.float32(12.9898))  # the reference forms the angle in float
        want = [-(p // 25) * s * 1.4, (p % 25 - 12) * s, h + 0.25 * np.sin(arg)]
        assert np.allclose(st[0:3], want, rtol=0, atol=1e-6), (p, st[0:3], want)
    st, _, _, _ = O.launch(g, 5, prec=64, plane_index=40, spread=0.0)
    assert np.array_equal(st[0:2], [0.0, 0.0])


# ---------------------------------------------------------------- GA engine and grand()

def test_mt19937_engine_standard_check_value_and_independent_implementation():
    """The C++ standard fixes mt19937's 10000th output of a default-seeded engine (5489) at
    4123659995 ([rand.predef]); numpy's MT19937 with the classic seeding is an independent
    implementation of the same engine."""
    assert int(O.mt19937(5489, 10000)[-1]) == 4123659995
    bg = np.random.MT19937(0)
    bg._legacy_seeding(42)
    assert np.array_equal(O.mt19937(42, 5000), bg.random_raw(5000).astype(np.uint32))


def canonical_f32(x):
    """generate_canonical<float, 24> of one 32-bit draw (libstdc++): float(x) / 2^32, values
    that round to 1 replaced by the largest float below 1."""
    u = x.astype(np.float32) * np.float32(2.0 ** -32)
    return np.where(u >= 1, np.nextafter(np.float32(1), np.float32(0)), u).astype(np.float32)


def test_random_genomes_follow_grand_rule():
    """grand() (sim.cu:318-321): gene i = clamp(lo_i + u (hi_i - lo_i)) with u one canonical float
    per gene, in gene order, from mt19937(seed)."""
    n = 300
    bg = np.random.MT19937(0)
    bg._legacy_seeding(42)
    u = canonical_f32(bg.random_raw(n * 35)).reshape(n, 35)
    lo, hi = LO.astype(np.float32), HI.astype(np.float32)
    want = np.minimum(hi, np.maximum(lo, lo + u * (hi - lo)))
    assert np.array_equal(O.random_genomes(n, 42), want)


# ---------------------------------------------------------------- breeding (sim.cu:472-490)

POP = 6400
ELITE = POP // 64


def breed(gs, acc, gen, gens, seed=42):
    idx = O.ga_rank(acc)
    return O.ga_breed(gs, acc, idx, gen, gens, seed=seed), idx


def near(a, b):
    """equal to within the rounding of a float blend of two equal values"""
    return np.abs(a - b) <= 4 * np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)


def test_breeding_elite_rows_copied_exactly():
    """the first pop/64 children are the ranked elite, bit for bit (sim.cu:474)"""
    rng = np.random.default_rng(3)
    gs = (LO + rng.random((1000, 35)) * RANGE).astype(np.float32)
    acc = rng.normal(5, 2, 1000).astype(np.float32)
    nxt, idx = breed(gs, acc, 3, 10)
    assert np.array_equal(nxt[:1000 // 64], gs[idx[:1000 // 64]])
    assert not np.array_equal(nxt[1000 // 64], gs[idx[1000 // 64]])
    assert np.all(nxt >= LO) and np.all(nxt <= HI)


@pytest.mark.parametrize("gen,gens", [(0, 10), (6, 10)])
def test_breeding_immigrants_and_mutation_from_identical_parents(gen, gens):
    """All parents equal G: a bred child equals G except where mutated, so the mutation rate
    (0.25 per gene), the mutation scale (N(0, 0.08 range anneal), anneal = 1 - 0.6 gen/gens) and
    the clamp to the bounds can be read off; immigrants (5 %) are rows unlike G in most genes,
    with genes uniform in the bounds (grand())."""
    G = (LO + 0.5 * RANGE).astype(np.float32)
    G[0] = np.float32(HI[0] - 0.03 * RANGE[0])  # near the upper bound: mutations clamp there
    G[5] = np.float32(LO[5] + 0.03 * RANGE[5])  # near the lower bound
    gs = np.tile(G, (POP, 1))
    acc = np.linspace(10, 1, POP).astype(np.float32)
    nxt, _ = breed(gs, acc, gen, gens)
    kids = nxt[ELITE:].astype(np.float64)
    same = near(kids, G.astype(np.float64))
    off = ~same
    imm = off.sum(1) > 20  # Binomial(35, 0.25) >= 21 has probability < 1e-5
    n = len(kids)
    rate = imm.mean()
    assert abs(rate - 0.05) < 4 * np.sqrt(0.05 * 0.95 / n), rate
    uim = (kids[imm] - LO) / RANGE
    assert stats.kstest(uim.ravel(), "uniform").pvalue > 1e-3
    bred = kids[~imm]
    mutated = off[~imm]
    m = mutated.mean()
    nb = mutated.size
    assert abs(m - 0.25) < 4 * np.sqrt(0.25 * 0.75 / nb), m
    anneal = 1.0 - 0.6 * gen / gens
    sig = 0.08 * RANGE * anneal
    z = (bred - G) / sig
    interior = [j for j in range(35) if j not in (0, 5)]
    zi = z[:, interior][mutated[:, interior]]
    assert stats.kstest(zi, "norm").pvalue > 1e-3
    assert abs(zi.std() - 1.0) < 0.02
    # clamps: gene 0 sits 0.03 range below its upper bound, gene 5 above its lower bound
    assert np.all(bred <= HI) and np.all(bred >= LO)
    for j, bound, lim in ((0, HI[0], (HI[0] - G[0]) / sig[0]), (5, LO[5], (LO[5] - G[5]) / sig[5])):
        col = bred[mutated[:, j], j]
        at = np.mean(col == np.float64(bound))
        want = stats.norm.sf(abs(float(lim)))
        assert abs(at - want) < 4 * np.sqrt(want * (1 - want) / len(col)) + 2e-3, (j, at, want)


def mix_cdf(x, lo, hi, s, w):
    """CDF of a mixture: weight 1-w U(lo, hi), weight w U(lo, hi) + N(0, s)."""
    G = lambda t: t * stats.norm.cdf(t) + stats.norm.pdf(t)
    conv = s / (hi - lo) * (G((x - lo) / s) - G((x - hi) / s))
    return (1 - w) * np.clip((x - lo) / (hi - lo), 0, 1) + w * conv


def test_breeding_blend_and_tournament_of_four():
    """Two parent types: A (fitness 10, a quarter of the population, genes at 0.3 of the range)
    and B (fitness 5, at 0.7).  A tournament of four (first contestant through the ranking, three
    by raw index; the fitter wins, sim.cu:478-482) picks an A parent with probability
    1 - 0.75^4; a child of two A (two B) parents equals A (B) except where mutated.  A child of
    one of each has genes a A + (1 - a) B with a ~ U(-0.1, 1.1) per gene (sim.cu:485-486), plus
    the mutation on a quarter of them: the blend weights read off those genes follow that
    mixture."""
    gen, gens = 4, 10
    A = (LO + 0.3 * RANGE).astype(np.float32)
    B = (LO + 0.7 * RANGE).astype(np.float32)
    is_a = (np.arange(POP) % 4) == 1
    gs = np.where(is_a[:, None], A, B).astype(np.float32)
    acc = np.where(is_a, 10.0, 5.0).astype(np.float32)
    nxt, idx = breed(gs, acc, gen, gens)
    assert np.all(is_a[idx[:POP // 4]])
    kids = nxt[ELITE:].astype(np.float64)
    eqa = near(kids, A.astype(np.float64)).sum(1)
    eqb = near(kids, B.astype(np.float64)).sum(1)
    # immigrant genes are uniform in the bounds, so only ~68 % of them lie within 0.35 |A - B|
    # of the A-B interval (Binomial(35, 0.68): 24 +- 3 genes), where a bred child's genes lie
    # unless a mutation moves them (>= 32 of 35 but for ~0.1 % of children)
    between = ((kids - A) / (B - A.astype(np.float64)))
    inside = ((between > -0.35) & (between < 1.35)).sum(1)
    imm = (inside <= 31) & (eqa < 12) & (eqb < 12)
    aa, bb = (eqa >= 12) & ~imm, (eqb >= 12) & ~imm
    ab = ~(imm | aa | bb)
    nb = len(kids) - imm.sum()
    p = 1 - 0.75 ** 4
    for obs, want in ((aa.sum(), p * p), (bb.sum(), (1 - p) ** 2), (ab.sum(), 2 * p * (1 - p))):
        assert abs(obs / nb - want) < 4 * np.sqrt(want * (1 - want) / nb), (obs / nb, want)
    assert abs(imm.mean() - 0.05) < 4 * np.sqrt(0.05 * 0.95 / len(kids))
    a = ((kids[ab] - B) / (A.astype(np.float64) - B)).ravel()
    s = 0.08 * (1 - 0.6 * gen / gens) / 0.4  # the mutation in units of |A - B| = 0.4 range
    ks = stats.kstest(a, lambda x: mix_cdf(x, -0.1, 1.1, s, 0.25))
    assert ks.pvalue > 1e-3, ks
    # a blend range of [0, 1], or weights drawn once per child, would fail the test above;
    # direct checks of the weights' support and spread:
    inner = a[(a > -0.1) & (a < 1.1)]
    assert abs(inner.mean() - 0.5) < 0.01 and abs(len(inner) / len(a) - mix_cdf(1.1, -0.1, 1.1, s, 0.25)
                                                   + mix_cdf(-0.1, -0.1, 1.1, s, 0.25)) < 0.01


# ---------------------------------------------------------------- mesh at any (ch, sp)

RES = [(2, 3), (3, 4), (5, 8), (6, 10), (8, 13), (10, 16), (12, 24)]


def layout_fins(g):
    lay = g[26]
    return 1 if lay < 0.25 else (2 if lay < 0.5 else (3 if lay < 0.75 else 2))


def family_genomes():
    """every fin layout, mono- and biplane"""
    out = []
    for k, lay in enumerate((0.1, 0.3, 0.6, 0.9)):
        for gap in (0.0, 0.04):
            g = PRESETS[k % 5].copy()
            g[26], g[13] = lay, gap
            out.append(g)
    return out


@pytest.mark.parametrize("ch,sp", RES)
def test_mesh_counts_and_deck_topology_at_any_resolution(ch, sp):
    """Per deck (2sp+1)(ch+1) vertices and 2 (2sp) ch triangles (a, b, a+1), (b, b+1, a+1) with
    b = a + ch + 1 over the (u, xc) grid (common.h:177-192); tail 9 / 8, keel 8 / 6, 4 / 2 per
    fin, struts 8 / 4 on a biplane -- none of which depend on (ch, sp)."""
    for g in family_genomes():
        k, bi = layout_fins(g), g[13] > np.float32(0.02)
        decks = 2 if bi else 1
        V, T = O.mesh(g, ch=ch, sp=sp)
        nd_v, nd_t = (2 * sp + 1) * (ch + 1), 2 * 2 * sp * ch
        assert len(V) == decks * nd_v + 9 + 4 * k + 8 + (8 if bi else 0)
        assert len(T) == decks * nd_t + 8 + 2 * k + 6 + (4 if bi else 0)
        want = []
        for i in range(2 * sp):
            for j in range(ch):
                a = i * (ch + 1) + j
                b = a + ch + 1
                want += [(a, b, a + 1), (b, b + 1, a + 1)]
        want = np.array(want)
        assert np.array_equal(T[:nd_t], want)
        if bi:
            off = nd_v + 9 + 4 * k
            assert np.array_equal(T[nd_t + 8 + 2 * k: 2 * nd_t + 8 + 2 * k], want + off)
        assert T.min() >= 0 and T.max() < len(V)


@pytest.mark.parametrize("ch,sp", RES)
def test_mesh_mirror_symmetry_at_any_resolution(ch, sp):
    """The deck grid is mirror-symmetric about y = 0 bit for bit: station i and 2sp - i have
    |u| equal, the same x and z, and opposite y (common.h:171)."""
    for g in family_genomes():
        V, _ = O.mesh(g, ch=ch, sp=sp)
        grid = V[:(2 * sp + 1) * (ch + 1)].reshape(2 * sp + 1, ch + 1, 3)
        assert np.array_equal(grid[::-1, :, 0], grid[:, :, 0])
        assert np.array_equal(grid[::-1, :, 1], -grid[:, :, 1])
        assert np.array_equal(grid[::-1, :, 2], grid[:, :, 2])
        assert np.all(grid[sp + 1:, :, 1] > 0) and np.all(grid[sp, :, 1] == 0)


@pytest.mark.parametrize("ch,sp", [(2, 3), (3, 4), (5, 8), (6, 10)])
def test_mesh_refinement_keeps_shared_nodes_bit_for_bit(ch, sp):
    """u_i = (i - sp)/sp and xc_j = j/ch are correctly rounded quotients, so the (2ch, 2sp) grid
    node (2i, 2j) is the same (u, xc) as node (i, j) of the (ch, sp) grid and carries the same
    vertex bit for bit; every part outside the decks is unchanged."""
    for g in family_genomes():
        bi = g[13] > np.float32(0.02)
        V1, T1 = O.mesh(g, ch=ch, sp=sp)
        V2, T2 = O.mesh(g, ch=2 * ch, sp=2 * sp)
        n1, n2 = (2 * sp + 1) * (ch + 1), (4 * sp + 1) * (2 * ch + 1)
        g1 = V1[:n1].reshape(2 * sp + 1, ch + 1, 3)
        g2 = V2[:n2].reshape(4 * sp + 1, 2 * ch + 1, 3)
        assert np.array_equal(g2[::2, ::2], g1)
        k = layout_fins(g)
        rest1, rest2 = V1[n1:n1 + 9 + 4 * k], V2[n2:n2 + 9 + 4 * k]
        assert np.array_equal(rest1, rest2)
        if bi:
            u1 = V1[n1 + 9 + 4 * k: 2 * n1 + 9 + 4 * k].reshape(2 * sp + 1, ch + 1, 3)
            u2 = V2[n2 + 9 + 4 * k: 2 * n2 + 9 + 4 * k].reshape(4 * sp + 1, 2 * ch + 1, 3)
            assert np.array_equal(u2[::2, ::2], u1)
        d = 2 if bi else 1
        assert np.array_equal(V1[d * n1 + 9 + 4 * k:], V2[d * n2 + 9 + 4 * k:])  # keel, struts
        nt1, nt2 = d * 4 * sp * ch, d * 4 * (2 * sp) * (2 * ch)
        # non-deck triangles: the same, shifted by the vertex-count difference
        t1 = T1[4 * sp * ch: 4 * sp * ch + 8 + 2 * k] - n1
        t2 = T2[16 * sp * ch: 16 * sp * ch + 8 + 2 * k] - n2
        assert np.array_equal(t1, t2)
        assert np.array_equal(T1[nt1 + 8 + 2 * k:] - (d * n1), T2[nt2 + 8 + 2 * k:] - (d * n2))
