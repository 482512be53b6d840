"""The allocation-free extension entry point (bsx_extend_batch_device_noalloc)
against bsx_extend_batch_device on the same device buffers: two calls in a row,
the second with a smaller batch (the shared scratch array keeps the first
call's contents), every output field and the cell counts bit-exact, under the
plain and the learned policy; a batch above the stated maximum is rejected
before any work is queued."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
EXT = ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w", "end_bonus")
FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off", "cells")


def _tiled(b, reps):
    """reps copies of a batch, each copy's h0 shifted so the copies differ."""
    from bsbolt_amd.tasks import ExtTasks
    nq, nt = len(b.query), len(b.target)
    return ExtTasks(np.tile(b.query, reps), np.concatenate([b.q_off + np.uint64(r * nq) for r in range(reps)]),
                    np.tile(b.qlen, reps), np.tile(b.target, reps),
                    np.concatenate([b.t_off + np.uint64(r * nt) for r in range(reps)]), np.tile(b.tlen, reps),
                    np.concatenate([np.maximum(b.h0 + 3 * r - 9, 1) for r in range(reps)]),
                    np.tile(b.w, reps), np.tile(b.end_bonus, reps))


def _resident(engine, b):
    dev = {k: engine.to_device(np.ascontiguousarray(getattr(b, k))) for k in EXT}
    n = len(b)

    def outs():
        r = {k: engine.empty((n,), np.int32) for k in FIELDS[:-1]}
        r["cells"] = engine.empty((n,), np.int64)
        return r
    return dev, outs(), outs()


def _ptrs(dev, res, n):
    p = {k: v.ptr for k, v in dev.items()}
    p.update({k: v.ptr for k, v in res.items()})
    p["n"] = n
    return p


@pytest.mark.parametrize("learned", [False, True])
def test_noalloc_equals_allocating_entry_point(engine, learned):
    from bsbolt_amd import band, bsx
    from bsbolt_amd.tasks import random_ext_tasks
    assert bsx.BSX_NOALLOC_MAX_TASKS == 262144
    big = _tiled(random_ext_tasks(bsx.BSX_NOALLOC_MAX_TASKS // 8, seed=71, qmax=160, h0max=140,
                                  w_choices=(16, 100, 200), n_frac=0.001), 8)
    small = random_ext_tasks(30000, seed=72, qmax=700, h0max=200, w_choices=(8, 100))
    if learned:
        engine.set_band_model(band.load_model())
        engine.set_band_policy(bsx.BAND_LEARNED)
    try:
        for call, b in enumerate((big, small)):  # second call smaller than the first
            n = len(b)
            print(f"noalloc call {call}: {n} tasks, learned={learned}", flush=True)
            dev, ra, rb = _resident(engine, b)
            engine.extend_device(_ptrs(dev, ra, n))
            engine.sync()
            engine.extend_device(_ptrs(dev, rb, n), noalloc=True)
            engine.sync()
            for k in FIELDS:
                assert np.array_equal(ra[k].numpy(), rb[k].numpy()), f"call {call}: {k}"
            if call == 1:  # and the oracle on the smaller batch
                want, _ = O.extend2_batch(b, nthreads=16)
                for k in FIELDS:
                    assert np.array_equal(rb[k].numpy(), want[k]), f"oracle: {k}"
            del dev, ra, rb
    finally:
        engine.set_band_policy(bsx.BAND_PLAIN)


def test_noalloc_rejects_batches_above_its_maximum(engine):
    from bsbolt_amd import bsx
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(16, seed=73)
    dev, ra, _ = _resident(engine, b)
    p = _ptrs(dev, ra, bsx.BSX_NOALLOC_MAX_TASKS + 1)  # checked before any work is queued
    with pytest.raises(RuntimeError, match="at most"):
        engine.extend_device(p, noalloc=True)
    lib = bsx.lib()
    assert b"at most 262144 tasks" in lib.bsx_last_error(ctypes.c_void_p(engine.ctx))
