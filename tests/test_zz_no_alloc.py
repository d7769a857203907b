"""pr_eval_noalloc, the allocation-free variant of pr_prepare + pr_scenarios + pr_fly (layout 0):
through the C ABI, two calls in a row -- the second with a smaller input, so that it runs on
the scratch array the first left behind -- each bit-identical to the same evaluation on a
fresh context of the allocating path; sizes beyond the array's capacity are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def PA():
    import planes_amd
    print("device:", planes_amd.device_name(), flush=True)
    return planes_amd


def noalloc(PA, genomes, seeds, params):
    n, k = len(genomes), len(seeds)
    dg = PA.DeviceBuffer(4 * 35 * n)
    dg.upload(np.ascontiguousarray(genomes, np.float32))
    ds = PA.DeviceBuffer(4 * k)
    ds.upload(np.ascontiguousarray(seeds, np.int32))
    df = PA.DeviceBuffer(4 * n * k)
    dst = PA.DeviceBuffer(4 * n * k)
    print(f"pr_eval_noalloc: {n} genomes x {k} seeds", flush=True)
    rc = PA.LIB.pr_eval_noalloc(dg.ptr, n, ds.ptr, k, C.byref(params), df.ptr, dst.ptr, None)
    assert rc == 0, PA.LIB.pr_last_error()
    assert PA.LIB.pr_device_sync() == 0
    fit = df.download(np.float32, n * k)
    steps = dst.download(np.int32, n * k)
    for b in (dg, ds, df, dst):
        b.free()
    return fit, steps


def allocating(PA, genomes, seeds, params):
    pl = PA.Planes(len(genomes), len(seeds), len(genomes) * len(seeds))
    fit, steps = pl.evaluate(genomes, seeds, params)
    pl.close()
    return np.asarray(fit).reshape(-1), np.asarray(steps).reshape(-1)


def test_noalloc_matches_allocating_path_on_two_calls(PA):
    params = PA.SimParams.reference(max_steps=3000)
    cases = [(O.random_genomes(1500, 71), np.arange(1, 11, dtype=np.int32)),
             (O.random_genomes(300, 72), np.array([5, 9, 2 ** 31 - 3], np.int32))]
    for gs, seeds in cases:
        a, sa = noalloc(PA, gs, seeds, params)
        b, sb = allocating(PA, gs, seeds, params)
        assert np.array_equal(a, b) and np.array_equal(sa, sb)
        assert (a == -100).sum() > 0 and np.all(sa[a != -100] > 0)


def test_noalloc_refuses_sizes_beyond_its_array(PA):
    params = PA.SimParams.reference(max_steps=600)
    dg = PA.DeviceBuffer(4 * 35 * 2049)
    ds = PA.DeviceBuffer(4 * 17)
    df = PA.DeviceBuffer(4 * 2049 * 17)
    assert PA.LIB.pr_eval_noalloc(dg.ptr, 2049, ds.ptr, 1, C.byref(params), df.ptr, None, None) == -1
    assert b"2048" in PA.LIB.pr_last_error()
    assert PA.LIB.pr_eval_noalloc(dg.ptr, 10, ds.ptr, 17, C.byref(params), df.ptr, None, None) == -1
    assert PA.LIB.pr_eval_noalloc(dg.ptr, 0, ds.ptr, 1, C.byref(params), df.ptr, None, None) == -1
