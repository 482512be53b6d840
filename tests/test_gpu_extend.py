"""GPU parity of the extension kernels (libbsx.so through the C ABI) against
the reference's golden outputs and against the CPU oracle.  Bit-exact."""
import ctypes

import numpy as np
import pytest

from conftest import EXT_GOLDEN, golden_ext
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")


def scoring_of(sc):
    from bsbolt_amd.bsx import Scoring
    return Scoring.make(mat=sc["mat"], o_del=sc["o_del"], e_del=sc["e_del"], o_ins=sc["o_ins"],
                        e_ins=sc["e_ins"], zdrop=sc["zdrop"])


def assert_same(got, want, label, fields=FIELDS):
    for k in fields:
        bad = np.nonzero(np.asarray(got[k]) != np.asarray(want[k]))[0]
        assert len(bad) == 0, f"{label}: {k} differs at {len(bad)} tasks, first {bad[:8]}"


@pytest.mark.parametrize("name", EXT_GOLDEN)
def test_extend_matches_reference_golden(engine, name):
    t, sc, ref = golden_ext(name)
    got = engine.extend(t, scoring_of(sc))
    assert_same(got, ref, name)


def test_extend_cells_match_oracle_on_golden(engine):
    t, sc, ref = golden_ext("ext_random_bsbolt.npz")
    got = engine.extend(t, scoring_of(sc))
    want, _ = O.extend2_batch(t, nthreads=8, scoring=sc)
    assert np.array_equal(got["cells"], want["cells"])


@pytest.mark.parametrize("seed,qmax,n", [(101, 150, 60000), (102, 260, 30000), (103, 2048, 1500)])
def test_extend_random_vs_oracle(engine, seed, qmax, n):
    from bsbolt_amd.tasks import random_ext_tasks
    t = random_ext_tasks(n, seed=seed, qmax=qmax, h0max=200, w_choices=(8, 16, 32, 64, 100, 200),
                         eb_choices=(0, 5, 30), n_frac=0.002)
    got = engine.extend(t)
    want, _ = O.extend2_batch(t, nthreads=16)
    assert_same(got, want, f"seed {seed}", FIELDS + ("cells",))


def test_extend_every_bin_boundary(engine):
    from bsbolt_amd.tasks import ExtTasks
    rng = np.random.default_rng(7)
    pairs, h0 = [], []
    for ql in (1, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 513,
               1023, 1024, 1025, 2047, 2048):
        for _ in range(3):
            q = rng.integers(0, 4, ql).astype(np.uint8)
            t = np.where(rng.random(ql) < 0.03, rng.integers(0, 4, ql), q).astype(np.uint8)
            pairs.append((q, np.concatenate([t, rng.integers(0, 4, int(rng.integers(0, 300)))]).astype(np.uint8)))
            h0.append(int(rng.integers(1, 300)))
    tk = ExtTasks.from_pairs(pairs, h0, 100, 30)
    got = engine.extend(tk)
    want, _ = O.extend2_batch(tk, nthreads=16)
    assert_same(got, want, "bins", FIELDS + ("cells",))


def test_extend_device_path(engine):
    from bsbolt_amd.tasks import random_ext_tasks
    t = random_ext_tasks(5000, seed=21)
    keep, ptrs = engine.tasks_to_device(t)
    out = {k: engine.to_device(np.full(len(t), -7, np.int32)) for k in FIELDS}
    cells = engine.to_device(np.zeros(len(t), np.int64))
    ptrs.update({k: v.ptr for k, v in out.items()})
    ptrs["cells"] = cells.ptr
    ptrs["n"] = len(t)
    engine.extend_device(ptrs)
    engine.sync()
    got = {k: v.numpy() for k, v in out.items()}
    got["cells"] = cells.numpy()
    want, _ = O.extend2_batch(t, nthreads=16)
    assert_same(got, want, "device path", FIELDS + ("cells",))


def test_extend_rejects_bad_tasks_and_leaves_their_slots(engine):
    from bsbolt_amd.bsx import BsxError, BSX_EINVAL
    from bsbolt_amd.tasks import ExtTasks
    rng = np.random.default_rng(3)
    good = [(rng.integers(0, 4, 30).astype(np.uint8), rng.integers(0, 4, 60).astype(np.uint8)) for _ in range(6)]
    pairs = good[:2] + [(np.zeros(0, np.uint8), good[2][1])] + good[3:4] + \
        [(rng.integers(0, 4, 3000).astype(np.uint8), good[4][1])] + good[5:]
    h0 = [20, 20, 20, 0, 20, 20]  # task 2: qlen 0, task 3: h0 0, task 4: qlen > 2048
    t = ExtTasks.from_pairs(pairs, h0, 100, 30)
    out = {k: np.full(len(t), -12345, np.int32) for k in FIELDS}
    with pytest.raises(BsxError) as ei:
        engine.extend(t, out=out, cells=False)
    assert ei.value.code == BSX_EINVAL and "task 2" in str(ei.value)
    for k in FIELDS:
        assert (out[k][[2, 3, 4]] == -12345).all()
    for i in (0, 1, 5):
        r = O.extend2(t.query_of(i), t.target_of(i), 20)
        assert tuple(int(out[k][i]) for k in FIELDS) == r[:6]


def test_scalar_ksw_extend2_drop_in(engine):
    from bsbolt_amd import bsx
    rng = np.random.default_rng(8)
    mat = O.fill_scmat(1, 4)
    for _ in range(20):
        ql = int(rng.integers(1, 120)); q = rng.integers(0, 5, ql).astype(np.uint8)
        t = np.concatenate([q % 4, rng.integers(0, 4, 40)]).astype(np.uint8)
        h0 = int(rng.integers(1, 90))
        want = O.extend2(q, t, h0, 100, 30, 100)[:6]
        assert bsx.ksw_extend2(ql, q, len(t), t, 5, mat, 6, 1, 6, 1, 100, 30, 100, h0) == want
        # the C symbol with the reference's exact signature
        o = (ctypes.c_int * 5)()
        s = bsx.lib().bsx_ksw_extend2(ql, q.ctypes.data, len(t), t.ctypes.data, 5, mat.ctypes.data,
                                      6, 1, 6, 1, 100, 30, 100, h0, *[ctypes.addressof(o) + 4 * k for k in range(5)])
        assert (s, *list(o)) == want


def test_scalar_drop_in_from_short_lived_threads(engine):
    """The reference's kt_for starts fresh worker threads for every batch of
    reads (kthread.c:49-58); the drop-in's contexts must stay bounded by the
    peak number of concurrent callers, not grow with the threads."""
    import threading
    from bsbolt_amd import bsx
    L = bsx.lib()
    mat = O.fill_scmat(1, 4)
    rng = np.random.default_rng(81)
    cases = []
    for _ in range(64):
        ql = int(rng.integers(1, 100)); q = rng.integers(0, 5, ql).astype(np.uint8)
        t = np.concatenate([q % 4, rng.integers(0, 4, 30)]).astype(np.uint8)
        h0 = int(rng.integers(1, 60))
        cases.append((q, t, h0, O.extend2(q, t, h0, 100, 30, 100)[:6]))
    bad = []

    def worker(k):
        q, t, h0, want = cases[k]
        o = (ctypes.c_int * 5)()
        s = L.bsx_ksw_extend2(len(q), q.ctypes.data, len(t), t.ctypes.data, 5, mat.ctypes.data,
                              6, 1, 6, 1, 100, 30, 100, h0, *[ctypes.addressof(o) + 4 * j for j in range(5)])
        if (s, *list(o)) != want:
            bad.append(k)

    conc = 4
    for wave in range(0, len(cases), conc):  # 16 batches of 4 fresh threads each
        th = [threading.Thread(target=worker, args=(k,)) for k in range(wave, wave + conc)]
        for x in th:
            x.start()
        for x in th:
            x.join()
    assert not bad, bad
    assert 1 <= L.bsx_dropin_contexts() <= conc + 1  # +1: the single-thread test before may have made one


def test_extend_is_deterministic(engine):
    from bsbolt_amd.tasks import random_ext_tasks
    t = random_ext_tasks(20000, seed=33)
    a = engine.extend(t)
    b = engine.extend(t)
    for k in a:
        assert np.array_equal(a[k], b[k])
