"""The allocation-free batch getter (stg_get_batch_tet_counts_noalloc) against
the allocating one (stg_get_batch_tet_counts), through the C ABI: two calls
in a row on different simulations and species, the second with fewer tets,
then the size limit."""
import ctypes as C

import numpy as np
import pytest

from steps_amd import _native as N
from steps_amd import workloads as W
from steps_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _noalloc(eng, spec, idx, real=0):
    lib = N.load()
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.full(len(idx), -1.0)
    N.check(lib.stg_get_batch_tet_counts_noalloc(eng.h, real, spec, N.ptr(idx, C.c_int64), len(idx),
                                                 N.ptr(out, C.c_double)), eng.h)
    return out


def test_noalloc_batch_getter_matches_the_allocating_one():
    w = W.four_species_cube(n=12, side=12e-6, seed=2, end_time=0.02)
    print(f"case noalloc getter, {w.n_tets} tets", flush=True)
    eng = Engine(w.fm, [w.mt], [4])
    for s in range(len(w.fm.specs)):
        eng.set_tet_counts(s, w.counts[s])
    eng.run(w.end_time)
    small = W.ab_binding_cube(n=3, side=3e-6, n_a=900, n_b=700, seed=1)
    eng2 = Engine(small.fm, [small.mt], [5])
    for s in range(3):
        eng2.set_tet_counts(s, small.counts[s])
    eng2.run(0.05)
    rng = np.random.default_rng(0)
    big = rng.integers(0, w.n_tets, size=50_000)
    few = rng.integers(0, small.n_tets, size=77)
    for e, spec, idx in ((eng, 2, big), (eng2, 1, few)):
        ref = e.batch_tet_counts(spec, idx)
        got = _noalloc(e, spec, idx)
        assert np.array_equal(got, ref)
        assert np.array_equal(got, e.get_tet_counts(spec)[idx].astype(float))
    assert len(_noalloc(eng, 0, np.zeros(0, dtype=np.int64))) == 0
    with pytest.raises(N.ArgErr, match="at most 65536"):
        _noalloc(eng, 0, np.zeros(65537, dtype=np.int64))
    with pytest.raises(N.ArgErr, match="out of range"):
        _noalloc(eng2, 0, np.array([small.n_tets]))
