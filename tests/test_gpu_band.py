"""GPU checks of the learned-band form against the CPU oracle: features and
labels bit-exact, narrow-band extension results identical to the reference's,
the MLP forward/backward/Adam kernels within stated tolerances, and what the
trained predictor gains over running without it."""
import numpy as np
import pytest

from conftest import workload_tasks
from oracle import band_mlp as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")


def _both(wl_small):
    from bsbolt_amd.workload import concat_ext
    return concat_ext(list(wl_small))


def _eligible(b):
    """Tasks a fixed-class run routes to the frame kernel: queries of at most 256 bases."""
    return b.qlen <= B.QMAX


def _policy(engine, policy, cls=0):
    from bsbolt_amd import bsx
    engine.set_band_policy(policy, cls)


def test_features_match_oracle(engine, wl_small):
    b = _both(wl_small)
    got = engine.band_features(b)
    want = B.features(b)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} tasks differ, first {bad[:5]}"


def test_features_random_tasks_with_n(engine):
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(4000, seed=31, qmax=300, h0max=300, w_choices=(1, 16, 100), eb_choices=(0, 30),
                         n_frac=0.02)
    assert np.array_equal(engine.band_features(b), B.features(b))


def test_labels_match_oracle(engine, wl_small):
    b = _both(wl_small)
    got = engine.band_labels(b)
    want = B.labels(b)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"{len(bad)} labels differ, first {bad[:5]}: gpu {got[bad[:5]]} cpu {want[bad[:5]]}"
    assert (want < 4).mean() > 0.3


@pytest.mark.parametrize("cls", [0, 1, 2, 3])
def test_fixed_class_equals_reference(engine, wl_small, cls):
    from bsbolt_amd import bsx
    b = _both(wl_small)
    want, _ = O.extend2_batch(b, nthreads=16)
    try:
        engine.set_band_policy(bsx.BAND_FIXED, cls)
        got = engine.extend(b)
        stats = engine.band_stats()
    finally:
        engine.set_band_policy(bsx.BAND_PLAIN, 0)
    for k in FIELDS + ("cells",):
        bad = np.nonzero(got[k] != want[k])[0]
        assert len(bad) == 0, f"class {cls}: {k} differs at {len(bad)} tasks, first {bad[:5]}"
    wf = B.FRAMES[cls]
    el = _eligible(b)
    assert stats[cls, 0] == int(el.sum())
    # the GPU frame run falls back exactly where the CPU statement rejects
    idx = np.nonzero(el)[0]
    rej = 0
    for i in idx:
        r = B.narrow_extend(b.query_of(i), b.target_of(i), int(b.h0[i]), wf, w=int(b.w[i]),
                            end_bonus=int(b.end_bonus[i]))[0]
        rej += r != 1
    assert stats[cls, 1] == rej
    assert stats[4, 0] == len(b) - int(el.sum())


@pytest.mark.parametrize("cls", [1, 2, 3])
def test_fixed_class_random_scorings(engine, cls):
    """Other gap costs (general e, asymmetric o/e), other matrices, N bases."""
    from bsbolt_amd import bsx
    from bsbolt_amd.bsx import Scoring
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(6000, seed=40 + cls, qmax=256, h0max=80, w_choices=(16, 40, 100, 200),
                         eb_choices=(0, 5, 30), n_frac=0.01)
    for sc in (dict(o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100),
               dict(o_del=5, e_del=2, o_ins=3, e_ins=1, zdrop=50),
               dict(o_del=4, e_del=3, o_ins=7, e_ins=2, zdrop=0)):
        want, _ = O.extend2_batch(b, nthreads=16, scoring=sc)
        try:
            engine.set_band_policy(bsx.BAND_FIXED, cls)
            got = engine.extend(b, Scoring.make(**sc))
        finally:
            engine.set_band_policy(bsx.BAND_PLAIN, 0)
        for k in FIELDS + ("cells",):
            assert np.array_equal(got[k], want[k]), (sc, k)


@pytest.mark.parametrize("name", ["ext_random_bsbolt.npz", "ext_edge_bsbolt.npz", "ext_brute_bsbolt.npz",
                                  "ext_random_nozdrop.npz", "ext_random_asym.npz", "ext_random_matrix.npz"])
def test_fixed_class_on_reference_golden(engine, name):
    """Frame runs (24 and 80 columns) with their fallback reproduce the reference's
    own outputs, including other scoring matrices and gap costs."""
    from bsbolt_amd import bsx
    from conftest import golden_ext
    from test_gpu_extend import scoring_of
    t, sc, ref = golden_ext(name)
    for cls in (0, 2):
        try:
            engine.set_band_policy(bsx.BAND_FIXED, cls)
            got = engine.extend(t, scoring_of(sc))
        finally:
            engine.set_band_policy(bsx.BAND_PLAIN)
        for k in FIELDS:
            bad = np.nonzero(got[k] != ref[k])[0]
            assert len(bad) == 0, f"{name} class {cls}: {k} differs at {len(bad)} tasks, first {bad[:5]}"


def _feats_labels(engine, b):
    return engine.band_features(b), engine.band_labels(b)


def test_forward_matches_oracle(engine, wl_small):
    b = _both(wl_small)
    X = engine.band_features(b)
    p = B.init_params(7)
    p[-5:] = [0.2, -0.1, 0.0, 0.3, -0.2]
    engine.set_band_model(p)
    lg, cl = engine.band_forward(X)
    want = B.forward(p, X)
    # tolerance: FP32 accumulation order inside MFMA and BF16 re-rounding of
    # hidden activations that land near a rounding boundary
    assert np.abs(lg - want).max() < 2e-2 * max(1.0, np.abs(want).max())
    srt = np.sort(want, axis=1)
    clear = srt[:, -1] - srt[:, -2] > 5e-2
    assert np.array_equal(cl[clear], np.argmax(want, axis=1)[clear])


def test_grad_matches_oracle(engine, wl_small):
    b = _both(wl_small)
    X = engine.band_features(b)
    rng = np.random.default_rng(2)
    y = rng.integers(0, 5, len(b)).astype(np.uint8)
    p = B.init_params(8)
    engine.set_band_model(p)
    g, loss = engine.band_grad(X, y)
    gw, lw = B.grad(p, X, y)
    assert abs(loss - lw) < 1e-3 * max(1.0, lw)
    rel = np.linalg.norm(g - gw) / np.linalg.norm(gw)
    assert rel < 1e-2, rel
    # each parameter block separately
    o = 0
    for name, shape in B.SHAPES:
        k = int(np.prod(shape))
        gb, wb = g[o:o + k], gw[o:o + k]
        assert np.linalg.norm(gb - wb) <= 2e-2 * np.linalg.norm(wb) + 1e-6, name
        o += k


def test_grad_ragged_tile(engine):
    rng = np.random.default_rng(4)
    X = B.bf16_bits(rng.normal(0, 1, (37, 32)).astype(np.float32))
    X[:, 24:] = 0
    y = rng.integers(0, 5, 37).astype(np.uint8)
    p = B.init_params(9)
    engine.set_band_model(p)
    g, loss = engine.band_grad(X, y)
    gw, lw = B.grad(p, X, y)
    assert abs(loss - lw) < 1e-3 * max(1.0, lw)
    assert np.linalg.norm(g - gw) / np.linalg.norm(gw) < 1e-2


def test_adam_step_matches_oracle(engine):
    rng = np.random.default_rng(5)
    X = B.bf16_bits(rng.normal(0, 1, (512, 32)).astype(np.float32))
    X[:, 24:] = 0
    y = rng.integers(0, 5, 512).astype(np.uint8)
    p0 = B.init_params(10)
    engine.set_band_model(p0)
    g, _ = engine.band_grad(X, y)
    Xd, yd = engine.to_device(X), engine.to_device(y)
    engine.band_train_device(Xd.ptr, yd.ptr, 512, 512, 1, 1e-3)
    engine.sync()
    p1 = engine.get_band_model()
    want, _, _ = B.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, 1e-3)
    assert np.abs(p1 - want).max() < 1e-6


def test_fused_predictor_matches_features_and_forward(engine, wl_small):
    """The learned policy's fused pass (features into LDS, MFMA tiles with the
    weights staged in LDS) gives the classes of the separate features and
    forward kernels, task for task (the same BF16 operands and MFMA order)."""
    b = _both(wl_small)
    p = B.init_params(5)
    p[-5:] = [0.4, 0.1, -0.2, 0.3, 0.0]
    engine.set_band_model(p)
    X = engine.band_features(b)
    _, cl = engine.band_forward(X)
    fused = engine.band_predict(b)
    assert np.array_equal(fused, cl)
    assert len(set(fused.tolist())) > 1


def test_learned_policy_results_equal_plain(engine, wl_small):
    from bsbolt_amd import bsx
    b = _both(wl_small)
    want, _ = O.extend2_batch(b, nthreads=16)
    p = B.init_params(11)
    p[-5:] = [1.0, 2.0, 1.5, 0.5, 0.0]  # bias the untrained net towards narrow classes
    engine.set_band_model(p)
    try:
        engine.set_band_policy(bsx.BAND_LEARNED)
        got = engine.extend(b)
        stats = engine.band_stats()
    finally:
        engine.set_band_policy(bsx.BAND_PLAIN)
    for k in FIELDS + ("cells",):
        assert np.array_equal(got[k], want[k]), k
    assert stats[:, 0].sum() == len(b)


def test_trained_predictor_gains(engine):
    """Train on one simulation, evaluate on another (same read recipe, disjoint
    genome and read seeds): the trained net routes more tasks to frames that
    hold their whole run, and fewer to frames that fall back, than the
    untrained net; and against the method without a network (the 24-column
    frame tried on every task, plain fallback) it cuts the frame runs that fall
    back while keeping most of the runs that hold (counts, not times)."""
    from bsbolt_amd import bsx
    train = workload_tasks("cfg2", 0.03)
    test = workload_tasks("cfg1", 0.01)
    Xtr, ytr = _feats_labels(engine, train)
    perm = np.random.default_rng(0).permutation(len(train))
    Xtr, ytr = np.ascontiguousarray(Xtr[perm]), np.ascontiguousarray(ytr[perm])
    p0 = B.init_params(12)
    engine.set_band_model(p0)

    def routed(params):
        engine.set_band_model(params)
        try:
            engine.set_band_policy(bsx.BAND_LEARNED)
            engine.extend(test, cells=False)
            return engine.band_stats()
        finally:
            engine.set_band_policy(bsx.BAND_PLAIN)

    s0 = routed(p0)
    engine.set_band_model(p0)
    Xd, yd = engine.to_device(Xtr), engine.to_device(ytr)
    steps = 600
    losses = engine.empty((steps,), np.float32)
    engine.band_train_device(Xd.ptr, yd.ptr, len(ytr), 2048, steps, 3e-3, losses.ptr)
    engine.sync()
    ls = losses.numpy()
    assert ls[-50:].mean() < 0.7 * ls[:10].mean(), (ls[:10].mean(), ls[-50:].mean())
    p1 = engine.get_band_model()
    s1 = routed(p1)
    acc0 = (s0[:4, 0] - s0[:4, 1]).sum()
    acc1 = (s1[:4, 0] - s1[:4, 1]).sum()
    rej0, rej1 = s0[:4, 1].sum(), s1[:4, 1].sum()
    # the method without a network: the 24-column frame tried on every task
    engine.set_band_policy(bsx.BAND_FIXED, 0)
    try:
        engine.extend(test, cells=False)
        sf = engine.band_stats()
    finally:
        engine.set_band_policy(bsx.BAND_PLAIN)
    accf, rejf = (sf[:4, 0] - sf[:4, 1]).sum(), sf[:4, 1].sum()
    print(f"untrained: narrow accepted {acc0} rejected {rej0}; trained: accepted {acc1} rejected {rej1}; "
          f"no predictor (24-column frame on every task): accepted {accf} rejected {rejf}; of {len(test)}")
    assert acc1 > 0.1 * len(test)
    assert rej1 < 0.5 * acc1
    # what the trained network gains over the method without it: far fewer frame
    # runs that fall back (work done twice), while keeping the frame runs that hold
    assert rej1 < 0.8 * rejf, (rej1, rejf)
    assert acc1 > 0.8 * accf, (acc1, accf)
    assert acc1 - rej1 > acc0 - rej0
