"""des_worst_quality_noalloc (partials in a static device array, no
allocation) against des_worst_quality (partials in the context's buffers),
through the C-ABI: two calls in a row, the second on a smaller model, each
compared exactly (quality and host element index), and once more after the
smaller model has written the shared array."""
import ctypes as C

import pytest

import scenarios as SC

pytestmark = pytest.mark.gpu


def _noalloc(g):
    q, w = C.c_double(), C.c_int()
    rc = g.lib.des_worst_quality_noalloc(g.ctx, C.byref(q), C.byref(w))
    assert rc == 0, g.lib.des_last_error(g.ctx).decode()
    return q.value, w.value


def test_noalloc_quality_equals_allocating_entry_point():
    from dynearthsol_amd.lib import DesModel
    big = DesModel(SC.build("evp3", nelem=6000))
    small = DesModel(SC.build("evp2", nelem=1500))
    big.run(7)
    small.run(5)
    print("models:", big.E, small.E, "elements", flush=True)
    a = _noalloc(big)
    assert a == big.worst_quality(), a
    b = _noalloc(small)
    assert b == small.worst_quality(), b
    assert _noalloc(big) == big.worst_quality()
    small.run(3)
    assert _noalloc(small) == small.worst_quality()
    big.close()
    small.close()

