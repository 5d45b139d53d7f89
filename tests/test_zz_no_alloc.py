"""ogp_conj_find_noalloc (the allocation-free threshold search) against ogp_conj_find: identical
records on two calls in a row, the second on a smaller search, and the limits reported as errors
with the count."""
import ctypes

import numpy as np
import pytest

import opengpop_amd as ogp
from opengpop_amd import _capi as C
from opengpop_amd import scenarios

pytestmark = pytest.mark.gpu


def _same(x, y):
    assert len(x) == len(y)
    for f in ("a", "b", "step", "distance", "time", "pos_a", "pos_b"):
        assert np.array_equal(getattr(x, f), getattr(y, f)), f


def test_noalloc_find_equals_find_on_two_calls():
    d = scenarios.config4(n_random=360, n_pairs=40, n_setA=80, days=0.25)
    with ogp.Context(degree=20) as c:
        ca = ogp.ClosestApproach(c, d["step"])
        ca.initialize(d["r"], d["v"], d["sets"])
        ca.propagate(d["steps"])
        ref = ca.find_closest(threshold=10e3)
        bytes_before = c.device_bytes
        got = ca.find_closest_noalloc(10e3)
        assert c.device_bytes == bytes_before
        assert len(ref) >= 15
        _same(got, ref)
        # second call: a smaller search (fewer objects, fewer steps, three sets); the array
        # still holds the first call's candidates and counter
        keep = np.r_[0:40, 80:200, 260:440]
        sets = np.repeat(np.arange(3, dtype=np.uint32), [40, 120, 180])
        ca.initialize(d["r"][keep], d["v"][keep], sets)
        ca.propagate(200)
        got2 = ca.find_closest_noalloc(10e3)
        ref2 = ca.find_closest(threshold=10e3)
        assert 0 < len(ref2) < len(ref)
        _same(got2, ref2)
        # too small a capacity: an error, with the count
        cnt = C._I64()
        L = ogp.lib()
        z = np.zeros(8)
        u = np.zeros(8, np.uint32)
        k = np.zeros(8, np.int32)
        rc = L.ogp_conj_find_noalloc(c.h, 10e3, 1, u.ctypes.data, u.ctypes.data, k.ctypes.data,
                                     z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                                     ctypes.byref(cnt))
        assert rc == C.ERR_ARGUMENT and cnt.value == len(ref2)
        # more candidates than the array holds (four sets of 110: 72,600 pairs, all below a
        # threshold of 10^6 km): an error, with the count, whatever the capacity
        ca.initialize(d["r"], d["v"], np.repeat(np.arange(4, dtype=np.uint32), 110))
        ca.propagate(20)
        rc = L.ogp_conj_find_noalloc(c.h, 1e9, 8, u.ctypes.data, u.ctypes.data, k.ctypes.data,
                                     z.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                                     ctypes.byref(cnt))
        assert rc == C.ERR_ARGUMENT and cnt.value == 72_600
        assert "65536" in L.ogp_last_error(c.h).decode()
