"""GPU parity of the global/CIGAR kernel against the reference's golden outputs
and the CPU oracle (scores and CIGARs bit-exact)."""
import numpy as np
import pytest

from conftest import GLB_GOLDEN, golden_glb
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def scoring_of(sc):
    from bsbolt_amd.bsx import Scoring
    return Scoring.make(mat=sc["mat"], o_del=sc["o_del"], e_del=sc["e_del"], o_ins=sc["o_ins"],
                        e_ins=sc["e_ins"])


@pytest.mark.parametrize("name", GLB_GOLDEN)
def test_global_matches_reference_golden(engine, name):
    t, sc, rs, rn, rc = golden_glb(name)
    score, ncig, cig = engine.global_(t, scoring_of(sc), slot=rc.shape[1])
    assert np.array_equal(score, rs)
    assert np.array_equal(ncig, rn)
    for i in range(len(t)):
        assert np.array_equal(cig[i, :ncig[i]], rc[i, :rn[i]]), f"task {i}"


def test_global_score_only(engine):
    t, sc, rs, rn, rc = golden_glb("glb_random_bsbolt.npz")
    score, ncig, cig = engine.global_(t, scoring_of(sc), cigar=False)
    assert np.array_equal(score, rs)


def test_global_random_vs_oracle(engine):
    from bsbolt_amd.tasks import GlbTasks
    rng = np.random.default_rng(44)
    pairs, ws = [], []
    for _ in range(20000):
        ql = int(rng.integers(1, 260))
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = list(np.where(rng.random(ql) < 0.05, rng.integers(0, 4, ql), q % 4))
        for _ in range(int(rng.integers(0, 4))):
            p = int(rng.integers(0, len(t) + 1)); L = int(rng.integers(1, 12))
            if rng.random() < 0.5:
                t[p:p] = list(rng.integers(0, 4, L))
            else:
                del t[p:p + L]
        t = np.array(t or [1], np.uint8)
        pairs.append((q, t)); ws.append(O.global_band(ql, len(t), 100) + int(rng.integers(0, 3)))
    tk = GlbTasks.from_pairs(pairs, ws)
    slot = int((tk.qlen + tk.tlen).max())
    score, ncig, cig = engine.global_(tk, slot=slot)
    os_, on, oc, _ = O.global2_batch(tk, slot)
    assert np.array_equal(score, os_)
    assert np.array_equal(ncig, on)
    for i in range(len(tk)):
        assert np.array_equal(cig[i, :ncig[i]], oc[i, :on[i]]), f"task {i}"


def test_global_rejects_out_of_band_tasks(engine):
    from bsbolt_amd.bsx import BsxError, BSX_EINVAL
    from bsbolt_amd.tasks import GlbTasks
    q = np.array([0, 1, 2, 3] * 5, np.uint8)
    pairs = [(q, q), (q, np.concatenate([q, q])), (q, q[:18])]
    tk = GlbTasks.from_pairs(pairs, [5, 5, 5])   # task 1: |tlen-qlen| = 20 > w
    (score, ncig, cig), rc = engine.global_(tk, raise_on_error=False)
    assert rc == BSX_EINVAL
    assert score[0] == 20 and ncig[0] == 1 and cig[0, 0] == (20 << 4)
    assert score[1] == 0  # untouched


def test_global_cigar_slot_overflow(engine):
    from bsbolt_amd.bsx import BSX_ERANGE
    from bsbolt_amd.tasks import GlbTasks
    q = np.array([0, 1, 2, 3] * 10, np.uint8)
    t = np.concatenate([q[:10], q[12:25], [0, 0, 0], q[25:]]).astype(np.uint8)
    tk = GlbTasks.from_pairs([(q, t)], 10)
    (score, ncig, cig), rc = engine.global_(tk, slot=2, raise_on_error=False)
    assert rc == BSX_ERANGE and ncig[0] == -1


def test_global_long_queries_narrow_and_wide_bands(engine):
    """Queries longer than any group (the lanes hold the band, not the query),
    w = 0, and bands wider than both sequences (taken as the full matrix).  The session engine keeps its side streams on, so the tasks with
    bands of 25-64 columns also run on the side stream's own slot region with
    this enlarged slot size (its offset scales with the slot)."""
    from bsbolt_amd.tasks import GlbTasks
    rng = np.random.default_rng(45)
    pairs, ws = [], []
    for k in range(300):
        ql = int(rng.integers(2100, 5000)) if k < 200 else int(rng.integers(1, 300))
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = list(np.where(rng.random(ql) < 0.05, rng.integers(0, 4, ql), q % 4))
        if k % 3:
            for _ in range(int(rng.integers(1, 4))):
                p = int(rng.integers(0, len(t) + 1)); L = int(rng.integers(1, 6))
                if rng.random() < 0.5:
                    t[p:p] = list(rng.integers(0, 4, L))
                else:
                    del t[p:p + L]
        t = np.array(t or [1], np.uint8)
        d = abs(len(t) - ql)
        w = (d + int(rng.integers(0, 40))) if k < 200 else [d, 10**6, 2**28][k % 3]
        pairs.append((q, t)); ws.append(w)
    tk = GlbTasks.from_pairs(pairs, ws)
    slot = int((tk.qlen + tk.tlen).max())
    engine.set_global_zslot(int((np.minimum(tk.qlen.astype(np.int64), 2 * tk.w.astype(np.int64) + 1)
                                 * tk.tlen).max()) + 64)
    try:
        score, ncig, cig = engine.global_(tk, slot=slot)
    finally:
        engine.set_global_zslot(1 << 16)
    os_, on, oc, _ = O.global2_batch(tk, slot)
    assert np.array_equal(score, os_)
    assert np.array_equal(ncig, on)
    for i in range(len(tk)):
        assert np.array_equal(cig[i, :ncig[i]], oc[i, :on[i]]), f"task {i}"


def test_global_wide_bands_walk_paths():
    """Bands wider than the walk's LDS window (n_col > 1021: every lane loads
    the direction byte) and bands of one window row per refill (n_col 513-1024),
    on an engine of their own: the direction scratch grows to #CU * 128 slots of
    ~1.1 MB for this test only."""
    from bsbolt_amd.bsx import Engine
    from bsbolt_amd.tasks import GlbTasks
    rng = np.random.default_rng(46)
    pairs, ws = [], []
    for ql, w in [(1030, 520)] * 4 + [(800, 300)] * 4:
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = list(np.where(rng.random(ql) < 0.05, rng.integers(0, 4, ql), q % 4))
        for _ in range(3):
            p = int(rng.integers(0, len(t) + 1)); L = int(rng.integers(1, 4))
            if rng.random() < 0.5:
                t[p:p] = list(rng.integers(0, 4, L))
            else:
                del t[p:p + L]
        pairs.append((q, np.array(t, np.uint8))); ws.append(w)
    tk = GlbTasks.from_pairs(pairs, ws)
    slot = int((tk.qlen + tk.tlen).max())
    need = int((np.minimum(tk.qlen.astype(np.int64), 2 * tk.w.astype(np.int64) + 1) * tk.tlen).max())
    e = Engine(0)
    try:
        e.L.bsx_set_side_streams(e.ctx, 0)  # no side-stream slot region: #CU * 128 slots
        e.set_global_zslot(need + 64)
        score, ncig, cig = e.global_(tk, slot=slot)
    finally:
        e.close()
    os_, on, oc, _ = O.global2_batch(tk, slot)
    assert np.array_equal(score, os_)
    assert np.array_equal(ncig, on)
    for i in range(len(tk)):
        assert np.array_equal(cig[i, :ncig[i]], oc[i, :on[i]]), f"task {i}"


def test_scalar_ksw_global2_drop_in(engine):
    """bsx_ksw_global2 through the C symbol with the reference's exact
    signature (ksw.h:84): score, n_cigar and a malloc'd CIGAR the caller frees."""
    import ctypes
    from bsbolt_amd import bsx
    L = bsx.lib()
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    rng = np.random.default_rng(84)
    mat = O.fill_scmat(1, 4)
    for _ in range(40):
        ql = int(rng.integers(1, 150))
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = (q % 4).copy()
        if ql > 10:  # an indel and a substitution
            k = int(rng.integers(1, ql - 1))
            t = np.concatenate([t[:k], rng.integers(0, 4, int(rng.integers(1, 4))).astype(np.uint8), t[k:]])
            t[0] = (t[0] + 1) % 4
        w = max(abs(len(t) - ql) + 3, int(rng.integers(1, 40)))
        want_s, want_c = O.global2(q, t, w)
        want_n = len(want_c)
        nc = ctypes.c_int(-7)
        cp = ctypes.POINTER(ctypes.c_uint32)()
        s = L.bsx_ksw_global2(ql, q.ctypes.data, len(t), t.ctypes.data, 5, mat.ctypes.data, 6, 1, 6, 1, w,
                              ctypes.byref(nc), ctypes.byref(cp))
        assert (s, nc.value) == (want_s, want_n)
        assert [cp[k] for k in range(nc.value)] == list(want_c)
        libc.free(ctypes.cast(cp, ctypes.c_void_p))
        # score only (n_cigar / cigar NULL, ksw.c:510-513)
        assert L.bsx_ksw_global2(ql, q.ctypes.data, len(t), t.ctypes.data, 5, mat.ctypes.data, 6, 1, 6, 1, w,
                                 None, None) == want_s


def test_host_call_sizes_the_direction_slot(engine):
    """Tasks whose band x target exceeds the default 64 KB direction slot
    (500-base reads, bands of ~241 columns): the host-memory call and the scalar
    drop-in size the slot for their own call; the context's slot is unchanged."""
    import ctypes
    from bsbolt_amd import bsx
    from bsbolt_amd.tasks import GlbTasks
    rng = np.random.default_rng(47)
    pairs, ws = [], []
    for _ in range(64):
        ql = int(rng.integers(450, 520))
        q = rng.integers(0, 4, ql).astype(np.uint8)
        t = list(np.where(rng.random(ql) < 0.03, rng.integers(0, 4, ql), q))
        p = int(rng.integers(0, ql)); t[p:p] = list(rng.integers(0, 4, int(rng.integers(1, 6))))
        t = np.array(t, np.uint8)
        pairs.append((q, t)); ws.append(abs(len(t) - ql) + 117)
    tk = GlbTasks.from_pairs(pairs, ws)
    need = np.minimum(tk.qlen.astype(np.int64), 2 * tk.w.astype(np.int64) + 1) * tk.tlen
    assert need.max() > 65536
    slot = int((tk.qlen + tk.tlen).max())
    score, ncig, cig = engine.global_(tk, slot=slot)
    os_, on, oc, _ = O.global2_batch(tk, slot)
    assert np.array_equal(score, os_) and np.array_equal(ncig, on)
    for i in range(len(tk)):
        assert np.array_equal(cig[i, :ncig[i]], oc[i, :on[i]]), f"task {i}"
    L = bsx.lib()
    mat = O.fill_scmat(1, 4)
    q, t = pairs[0]
    nc = ctypes.c_int(0)
    cp = ctypes.POINTER(ctypes.c_uint32)()
    s = L.bsx_ksw_global2(len(q), q.ctypes.data, len(t), t.ctypes.data, 5, mat.ctypes.data, 6, 1, 6, 1, ws[0],
                          ctypes.byref(nc), ctypes.byref(cp))
    assert (s, nc.value, [cp[k] for k in range(nc.value)]) == (int(os_[0]), int(on[0]), list(oc[0, :on[0]]))
    ctypes.CDLL(None).free(ctypes.cast(cp, ctypes.c_void_p))
