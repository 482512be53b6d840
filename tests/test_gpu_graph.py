"""Graph-captured batch submission: a captured extension / CIGAR batch is
replayed on refilled device buffers, with eager work on the same context
between replays, and every replay must equal the oracle for the data it ran on."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
EXT = ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w", "end_bonus")
FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off", "cells")


def _ext_buffers(engine, batches):
    n = len(batches[0])
    caps = {k: max(getattr(b, k).nbytes for b in batches) for k in EXT}
    dev = {k: engine.empty((caps[k],), np.uint8) for k in EXT}
    res = {k: engine.empty((n,), np.int32) for k in FIELDS[:-1]}
    res["cells"] = engine.empty((n,), np.int64)
    ptrs = {k: v.ptr for k, v in dev.items()}
    ptrs.update({k: v.ptr for k, v in res.items()})
    ptrs["n"] = n
    return dev, res, ptrs


@pytest.mark.parametrize("learned", [False, True])
def test_extend_graph_replays(engine, learned):
    from bsbolt_amd import band, bsx
    from bsbolt_amd.tasks import random_ext_tasks
    batches = [random_ext_tasks(20000, seed=60 + s, qmax=150, h0max=120, w_choices=(16, 100), n_frac=0.001)
               for s in range(3)]
    dev, res, ptrs = _ext_buffers(engine, batches)
    if learned:
        engine.set_band_model(band.load_model())
        engine.set_band_policy(bsx.BAND_LEARNED)
    try:
        for k in EXT:
            dev[k].upload(getattr(batches[0], k))
        g = engine.capture_extend(ptrs)
        for rep, b in enumerate(batches + [batches[0]]):
            for k in EXT:
                dev[k].upload(getattr(b, k))
            # eager work on the context between replays, with a larger batch
            other = random_ext_tasks(30000 + 1000 * rep, seed=90 + rep, qmax=200)
            eager = engine.extend(other)
            g.launch()
            engine.sync()
            want, _ = O.extend2_batch(b, nthreads=16)
            for k in FIELDS:
                assert np.array_equal(res[k].numpy(), want[k]), f"replay {rep}: {k}"
            ow, _ = O.extend2_batch(other, nthreads=16)
            assert np.array_equal(eager["score"], ow["score"])
        g.close()
    finally:
        engine.set_band_policy(bsx.BAND_PLAIN)


def test_global_graph_replays(engine):
    from bsbolt_amd.tasks import GlbTasks
    rng = np.random.default_rng(3)
    batches = []
    for s in range(3):
        pairs, ws = [], []
        for _ in range(4000):
            ql = int(rng.integers(20, 120))
            q = rng.integers(0, 4, ql).astype(np.uint8)
            t = q.copy()
            for _ in range(int(rng.integers(0, 3))):
                p = int(rng.integers(0, len(t)))
                t = np.delete(t, p) if rng.random() < 0.5 else np.insert(t, p, int(rng.integers(0, 4)))
            t = np.where(rng.random(len(t)) < 0.03, rng.integers(0, 4, len(t)), t).astype(np.uint8)
            pairs.append((q, t))
            ws.append(O.global_band(ql, len(t), 100))
        batches.append(GlbTasks.from_pairs(pairs, ws))
    n = 4000
    slot = 256
    fields = ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")
    caps = {k: max(getattr(b, k).nbytes for b in batches) for k in fields}
    dev = {k: engine.empty((caps[k],), np.uint8) for k in fields}
    score = engine.empty((n,), np.int32)
    ncig = engine.empty((n,), np.int32)
    cig = engine.empty((n * slot,), np.uint32)
    ptrs = {k: v.ptr for k, v in dev.items()}
    ptrs.update(n=n, score=score.ptr, n_cigar=ncig.ptr, cigar=cig.ptr, cigar_slot=slot)
    for k in fields:
        dev[k].upload(getattr(batches[0], k))
    g = engine.capture_global(ptrs)
    for rep, b in enumerate(batches + [batches[1]]):
        for k in fields:
            dev[k].upload(getattr(b, k))
        engine.global_(batches[(rep + 1) % 3])  # eager work between replays
        g.launch()
        engine.sync()
        ws_, wn, wc, _ = O.global2_batch(b, slot)
        assert np.array_equal(score.numpy(), ws_), f"replay {rep}: score"
        assert np.array_equal(ncig.numpy(), wn), f"replay {rep}: n_cigar"
        got = cig.numpy().reshape(n, slot)
        for i in range(n):
            assert np.array_equal(got[i, :wn[i]], wc[i, :wn[i]]), f"replay {rep}: task {i}"
    g.close()


def test_stream_global_in_graph_batches(engine):
    """A CIGAR batch streamed through a captured graph in fixed-size chunks
    (plus an eager remainder) equals the one-call result and the oracle."""
    from bsbolt_amd.tasks import GlbTasks
    rng = np.random.default_rng(8)
    pairs, ws = [], []
    for _ in range(10500):
        ql = int(rng.integers(40, 101))
        q = rng.integers(0, 4, ql).astype(np.uint8)
        t = np.where(rng.random(ql) < 0.04, rng.integers(0, 4, ql), q).astype(np.uint8)
        if rng.random() < 0.3:
            t = np.insert(t, int(rng.integers(0, ql)), rng.integers(0, 4, int(rng.integers(1, 4)))).astype(np.uint8)
        pairs.append((q, t))
        ws.append(O.global_band(ql, len(t), 100))
    tk = GlbTasks.from_pairs(pairs, ws)
    s, nc, cg = engine.stream_global(tk, batch=4096)
    s1, nc1, cg1 = engine.global_(tk, slot=cg.shape[1])
    assert np.array_equal(s, s1) and np.array_equal(nc, nc1) and np.array_equal(cg, cg1)
    ws_, wn, wc, _ = O.global2_batch(tk, cg.shape[1])
    assert np.array_equal(s, ws_) and np.array_equal(nc, wn)


def test_stream_cfg4_cigar_pass_full_graph_batch(engine):
    """The RRBS workload's CIGAR pass (every try of mem_reg2aln's band doubling,
    built in rounds on the GPU) streamed through hipGraph-captured batches of
    262,144 tasks: at least one full batch plus an eager remainder, equal to the
    oracle on every task (scores, CIGAR counts, CIGARs)."""
    from bsbolt_amd.tasks import GlbTasks
    from bsbolt_amd.workload import CONFIGS, Workload, build_glb_batch
    wl = Workload(CONFIGS["cfg4"].scaled(0.05))
    try:
        wl.build_ext_batch(lambda t: engine.extend(t, cells=False))
        gt = build_glb_batch(wl, lambda t: engine.global_(t))[0]
    finally:
        wl.close()
    m = len(gt)
    reps = -(-300000 // m)  # enough copies for one full 262,144-task graph batch and a remainder
    qn, tn = len(gt.query), len(gt.target)
    tk = GlbTasks(np.tile(gt.query, reps), np.concatenate([gt.q_off + np.uint64(r * qn) for r in range(reps)]),
                  np.tile(gt.qlen, reps), np.tile(gt.target, reps),
                  np.concatenate([gt.t_off + np.uint64(r * tn) for r in range(reps)]), np.tile(gt.tlen, reps),
                  np.tile(gt.w, reps))
    n = len(tk)
    assert n >= 262144 and n % 262144 != 0
    print(f"cfg4 CIGAR pass: {m} tasks x {reps} = {n}, graph batches of 262144", flush=True)
    s, nc, cg = engine.stream_global(tk, batch=262144)
    ws_, wn, wc, _ = O.global2_batch(tk, cg.shape[1], nthreads=16)
    assert np.array_equal(s, ws_) and np.array_equal(nc, wn)
    bad = [i for i in range(n) if not np.array_equal(cg[i, :nc[i]], wc[i, :wn[i]])]
    assert not bad, f"{len(bad)} CIGARs differ, first {bad[:5]}"
