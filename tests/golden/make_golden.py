"""Writes the golden fixtures under tests/golden/ from the reference's own code.

Run in the container that holds the reference checkout, after `make -C oracle ref`
has compiled oracle/_ref/libplanes_ref.so from the reference's src/common.h,
src/designs.cu and src/stl.cu (the file set of its `make host-test` recipe).
The fixtures are data only: inputs and the reference's outputs.

  python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402  (only for random_genomes = mt19937 + grand rule)

F32P, I32P, IP = O.F32P, O.I32P, O.IP


def main():
    ref = C.CDLL(O.REF_SO)
    ref.ref_build_mesh.argtypes = [F32P, F32P, I32P, IP, IP]
    ref.ref_mass_props.argtypes = [F32P, F32P]
    ref.ref_scenario.argtypes = [C.c_int, F32P]
    ref.ref_printed.argtypes = [F32P, C.c_int, F32P, F32P]
    ref.ref_wind.argtypes = [C.c_float] * 4 + [C.c_int, F32P]
    ref.ref_preset.argtypes = [C.c_int, F32P, C.c_char_p, C.c_int]
    ref.ref_wing_pt.argtypes = [F32P, C.c_float, C.c_float, C.c_float, C.c_float, F32P]
    sz = [C.c_int() for _ in range(4)]
    ref.ref_sizes(*[C.byref(x) for x in sz])

    names, presets = [], []
    for i in range(ref.ref_n_presets()):
        g = np.zeros(35, np.float32)
        nm = C.create_string_buffer(64)
        ref.ref_preset(i, g, nm, 64)
        names.append(nm.value.decode())
        presets.append(g)
    presets = np.array(presets, np.float32)
    np.savez_compressed(os.path.join(HERE, "ref_presets.npz"), genomes=presets, names=np.array(names),
                        sizes=np.array([x.value for x in sz], np.int32))

    # genomes: 5 presets + draws of the reference's grand() rule from mt19937(0)
    rnd = O.random_genomes(507, 0)
    gs = np.concatenate([presets, rnd]).astype(np.float32)
    n = len(gs)
    mass = np.zeros((n, 20), np.float32)
    nt = np.zeros(n, np.int32)
    nv = np.zeros(n, np.int32)
    fits = np.zeros(n, np.int32)
    Vs, Ts = [], []
    for i, g in enumerate(gs):
        V = np.zeros(512 * 3, np.float32)
        T = np.zeros(900 * 3, np.int32)
        a, b = C.c_int(), C.c_int()
        fits[i] = ref.ref_build_mesh(g, V, T, C.byref(a), C.byref(b))
        nv[i], nt[i] = a.value, b.value
        ref.ref_mass_props(g, mass[i])
        if i < 24:
            Vs.append(V[:3 * a.value].reshape(-1, 3))
            Ts.append(T[:3 * b.value].reshape(-1, 3))
    np.savez_compressed(os.path.join(HERE, "ref_mass.npz"), genomes=gs, mass=mass, nt=nt, nv=nv, fits=fits)
    np.savez_compressed(os.path.join(HERE, "ref_mesh.npz"), genomes=gs[:24],
                        V=np.concatenate(Vs), T=np.concatenate(Ts), nv=nv[:24], nt=nt[:24])

    # wing_pt mirror samples (geometry_test.cpp:19-23 grid) for the presets
    wp = np.zeros((5, 2, 9, 6, 3), np.float32)
    for i, g in enumerate(presets):
        for s in range(9):
            for j in range(6):
                u, xc = np.float32(s / 8), np.float32(j / 5)
                ref.ref_wing_pt(g, -u, xc, 0, 1, wp[i, 0, s, j])
                ref.ref_wing_pt(g, u, xc, 0, 1, wp[i, 1, s, j])
    np.savez_compressed(os.path.join(HERE, "ref_wing_pt.npz"), wp=wp)

    # scenarios: small, negative, GA/holdout/audit seeds and int32-overflow seeds
    seeds = list(range(-40, 600)) + [100003 + p * 7919 for p in range(128)] + \
        [9001 + i * 7919 for i in range(0, 4096, 37)] + [730001, 930001, 2930001, 1900001] + \
        [58040099, 58040100, 58040101, 790000000, 2 ** 31 - 1, -2 ** 31, 123456789, -987654321]
    seeds = np.array(seeds, np.int32)
    sc = np.zeros((len(seeds), 25), np.float32)
    for i, s in enumerate(seeds):
        ref.ref_scenario(int(s), sc[i])
    np.savez_compressed(os.path.join(HERE, "ref_scenario.npz"), seeds=seeds, desc=sc)

    # printed mass props: genome i x seed
    pi = np.arange(0, 64, dtype=np.int32)
    ps = (1000 + 13 * pi).astype(np.int32)
    pout = np.zeros((64, 20), np.float32)
    psh = np.zeros((64, 3), np.float32)
    for k in range(64):
        ref.ref_printed(mass[pi[k]], int(ps[k]), pout[k], psh[k])
    np.savez_compressed(os.path.join(HERE, "ref_printed.npz"), gi=pi, seeds=ps, out=pout, shift=psh)

    # wind field samples
    rng = np.random.default_rng(7)
    m = 512
    P = (rng.normal(size=(m, 3)) * np.array([8, 3, 1.5])).astype(np.float32)
    t = rng.uniform(0, 20, m).astype(np.float32)
    ws = rng.integers(-1000, 3_000_000, m).astype(np.int32)
    W = np.zeros((m, 3), np.float32)
    for k in range(m):
        ref.ref_wind(P[k, 0], P[k, 1], P[k, 2], t[k], int(ws[k]), W[k])
    np.savez_compressed(os.path.join(HERE, "ref_wind.npz"), p=P, t=t, seeds=ws, w=W)
    print("golden fixtures written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
