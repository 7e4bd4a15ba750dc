"""Consecutive runs (des_run / des_batch_run with nothing in between) carry
on as one run: the next run's first step completes the last one's deferred
tail -- volume swap, mass, shape functions, Jaumann rotation -- inside its own
passes instead of in extra kernels between the runs.  Runs of 7 + 11 + 7 steps
must therefore end bitwise where one run of 25 steps ends, in FP64 and in
mixed precision, one pass per launch and in the persistent launch, for single
models and batches; in FP64 also where the same runs end when a download
between them completes the tail with the per-function kernels (the
reference's own order: update_mesh, then rotate_stress).  A change of
precision completes the tail first."""
import numpy as np
import pytest

import scenarios as SC
from dynearthsol_amd import abi

pytestmark = pytest.mark.gpu

NAMES = ["evp3", "ep3", "psevp2", "bc3", "mat2", "fields3", "dynmat3", "maxwell3", "slant3"]
CHUNKS = [7, 11, 7]            # steps 10 and 20 update dt inside the second and third run


def _same(a, b, tag):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), (tag, k, float(np.max(np.abs(a[k] - b[k]))))


def _single(m, runs, precision, mode=0, read_between=False):
    from dynearthsol_amd.lib import DesModel
    g = DesModel(m)
    g.set_launch_mode(mode)
    for i, n in enumerate(runs):
        p = precision[i] if isinstance(precision, list) else precision
        g.run(n, p)
        if read_between and i + 1 < len(runs):
            g.download()                 # completes the deferred tail
    out = g.download(), g.scalars()
    g.close()
    return out


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name", NAMES)
def test_chunked_runs_equal_one_run_fp64(name, mode):
    m = SC.build(name)
    print(name, m.nelem, "elements, launch mode", mode, flush=True)
    one, s1 = _single(m, [sum(CHUNKS)], abi.DES_FP64, mode)
    cont, s2 = _single(m, CHUNKS, abi.DES_FP64, mode)
    flushed, s3 = _single(m, CHUNKS, abi.DES_FP64, mode, read_between=True)
    _same(one, cont, (name, "continued"))
    _same(one, flushed, (name, "tail completed between the runs"))
    assert s1 == s2 == s3


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name", ["evp3", "mat3", "maxwell2", "fields3", "bc3"])
def test_chunked_runs_equal_one_run_mixed(name, mode):
    m = SC.build(name)
    print(name, m.nelem, "elements, mixed, launch mode", mode, flush=True)
    one, s1 = _single(m, [sum(CHUNKS)], abi.DES_MIXED, mode)
    cont, s2 = _single(m, CHUNKS, abi.DES_MIXED, mode)
    _same(one, cont, (name, "continued"))
    assert s1 == s2


@pytest.mark.parametrize("name", ["evp3", "maxwell2"])
def test_precision_change_completes_the_tail(name):
    """FP64 -> mixed -> FP64 runs back to back are the same runs with the
    deferred tail completed between them."""
    m = SC.build(name)
    print(name, m.nelem, "elements, FP64 / mixed / FP64", flush=True)
    prec = [abi.DES_FP64, abi.DES_MIXED, abi.DES_FP64]
    a, sa = _single(m, CHUNKS, prec)
    b, sb = _single(m, CHUNKS, prec, read_between=True)
    _same(a, b, name)
    assert sa == sb


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("precision", [abi.DES_FP64, abi.DES_MIXED])
def test_chunked_batch_runs_equal_one_run(precision, mode):
    from dynearthsol_amd.lib import DesBatch
    names = ["evp3", "maxwell3", "evp2", "maxwell2", "viscous2"]
    models = [SC.build(n) for n in names]
    models.append(SC.build("evp3", nelem=900))
    print("batch of", [m.nelem for m in models], "precision", precision, "mode", mode, flush=True)
    out = []
    for runs in ([sum(CHUNKS)], CHUNKS):
        bt = DesBatch(models)
        bt.set_launch_mode(mode)
        for n in runs:
            bt.run(n, precision)
        bt.sync()
        out.append([(bt.download(k), bt.scalars(k)) for k in range(len(models))])
        bt.close()
    for k in range(len(models)):
        _same(out[0][k][0], out[1][k][0], k)
        assert out[0][k][1] == out[1][k][1], k
