"""CPU checks of the learned-band oracle (oracle/band_oracle.c, oracle/band_mlp.py):
the band-following frame run against the compiled reference, the task features
against hand-computed values, and the MLP backward pass against finite
differences."""
import numpy as np
import pytest

from oracle import band_mlp as B
from oracle import oracle as O

FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")


def _reference(b):
    """Outputs of the compiled reference where it is built, else the oracle; cells from the oracle."""
    want, _ = O.extend2_batch(b, nthreads=8)
    if O.ref_available():
        ref, _ = O.extend2_batch(b, nthreads=8, use_ref=True, want_cells=False)
        for k in FIELDS:
            assert np.array_equal(ref[k], want[k])
    return want


def _check_sound(b, wf):
    want = _reference(b)
    acc = 0
    whys = np.zeros(3, int)
    for i in range(len(b)):
        r, out, cells, why = B.narrow_extend(b.query_of(i), b.target_of(i), int(b.h0[i]), wf, w=int(b.w[i]),
                                             end_bonus=int(b.end_bonus[i]))
        # -2: a column entered the frame without its first-row state (the argument the GPU frame relies on)
        assert r in (0, 1), f"task {i} wf {wf}: r = {r}"
        if r == 1:
            acc += 1
            got = dict(zip(FIELDS, out))
            for k in FIELDS:
                assert got[k] == want[k][i], f"task {i} wf {wf}: {k} {got[k]} != {want[k][i]}"
            assert cells == want["cells"][i], f"task {i} wf {wf}: cells"
        else:
            whys[why] += 1
    return acc, whys


_WHYS = {}


@pytest.mark.parametrize("wf", B.FRAMES)
def test_frame_run_sound_on_workloads(wl_small, wf):
    tot_acc = 0
    tot_why = np.zeros(3, int)
    for b in wl_small:
        acc, whys = _check_sound(b, wf)
        tot_acc += acc
        tot_why += whys
    n = sum(len(b) for b in wl_small)
    assert tot_acc > 0.3 * n, f"wf {wf}: accepted only {tot_acc} of {n}"
    _WHYS[wf] = tot_why


def test_frame_every_rejection_reason_occurs(wl_small):
    if len(_WHYS) < len(B.FRAMES):
        for wf in B.FRAMES:
            _WHYS[wf] = sum(_check_sound(b, wf)[1] for b in wl_small)
    tot = sum(_WHYS.values())
    assert all(tot[1:] > 0), _WHYS


@pytest.mark.parametrize("seed", [11, 12])
def test_frame_run_sound_on_random_tasks(seed):
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(3000, seed=seed, qmax=200, h0max=120, w_choices=(16, 32, 64, 100, 200),
                         eb_choices=(0, 5, 30), n_frac=0.002)
    for wf in B.FRAMES:
        _check_sound(b, wf)


@pytest.mark.parametrize("seed", [21, 22])
def test_frame_run_sound_on_random_scorings(seed):
    """General gap extensions and asymmetric gap costs (the first-row insertion
    values then fall by e_ins per column)."""
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(1500, seed=seed, qmax=200, h0max=120, w_choices=(16, 40, 100),
                         eb_choices=(0, 5, 30), n_frac=0.01)
    for sc in (dict(o_del=5, e_del=2, o_ins=3, e_ins=1, zdrop=50), dict(o_del=4, e_del=3, o_ins=7, e_ins=2, zdrop=0)):
        want, _ = O.extend2_batch(b, nthreads=8, scoring=sc)
        for i in range(len(b)):
            for wf in (24, 80):
                r, out, cells, why = B.narrow_extend(b.query_of(i), b.target_of(i), int(b.h0[i]), wf, w=int(b.w[i]),
                                                     end_bonus=int(b.end_bonus[i]), zdrop=sc["zdrop"],
                                                     o_del=sc["o_del"], e_del=sc["e_del"], o_ins=sc["o_ins"],
                                                     e_ins=sc["e_ins"])
                assert r in (0, 1)
                if r == 1:
                    assert out == tuple(int(want[k][i]) for k in FIELDS) and cells == want["cells"][i], (sc, i, wf)


def test_frame_wider_than_query_is_the_reference_run():
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(500, seed=5, qmax=120, h0max=100, w_choices=(8, 16, 32, 200), eb_choices=(0, 30))
    want = _reference(b)
    for i in range(len(b)):
        r, out, cells, why = B.narrow_extend(b.query_of(i), b.target_of(i), int(b.h0[i]), 144, w=int(b.w[i]),
                                             end_bonus=int(b.end_bonus[i]))
        assert r == 1
        assert out == tuple(int(want[k][i]) for k in FIELDS) and cells == want["cells"][i]


def test_frame_rejections_crafted():
    rng = np.random.default_rng(3)
    q = rng.integers(0, 4, 60).astype(np.uint8)
    # h0 = 60: the first-row insertion values h0 - 7 - (j-1) are still positive at
    # column 24, where the band of row 0 (the whole query) leaves a 24-column frame
    r, out, cells, why = B.narrow_extend(q, q, 60, 24)
    assert r == 0 and why == 2
    # ... and reach 0 at column 54, inside an 80-column frame; the exact match
    # then keeps its band inside the frame
    r, out, cells, why = B.narrow_extend(q, q, 60, 80)
    assert r == 1 and out[0] == 60 + 60
    # h0 = 5 leaves no first-row values; a 12-base exact match fits 24 columns
    r, out, cells, why = B.narrow_extend(q[:12], q[:12], 5, 24)
    assert r == 1 and out[0] == 5 + 12 and cells == int(O.extend2_batch(_pairs([(q[:12], q[:12])], 5),
                                                                          nthreads=1)[0]["cells"][0])
    # a 60-base exact match from h0 = 5 scores 5 + i after row i: its insertion arm
    # H(i, i+d) = 5 + i - 6 - d reaches column ~2i while the deletion arm keeps the
    # band start near i/2, so the band outgrows 24 columns with non-zero cells past it
    r, out, cells, why = B.narrow_extend(q, q, 5, 24)
    assert r == 0 and why == 1


def _pairs(pairs, h0):
    from bsbolt_amd.tasks import ExtTasks
    return ExtTasks.from_pairs(pairs, h0=[h0] * len(pairs), w=100, end_bonus=30)


def test_labels_are_smallest_accepted_class(wl_small):
    b = wl_small[0]
    idx = np.arange(0, len(b), 7)
    s = b.subset(idx)
    lab = B.labels(s)
    for i in range(len(s)):
        acc = [B.narrow_extend(s.query_of(i), s.target_of(i), int(s.h0[i]), wf, w=int(s.w[i]),
                               end_bonus=int(s.end_bonus[i]))[0] == 1 for wf in B.FRAMES]
        want = next((c for c in range(4) if acc[c]), 4) if s.qlen[i] <= B.QMAX else 4
        assert lab[i] == want


# ---------------- features ----------------

def test_features_closed_form():
    from bsbolt_amd.tasks import ExtTasks
    rng = np.random.default_rng(1)
    q = rng.integers(0, 4, 20).astype(np.uint8)
    t_same = q.copy()
    t_shift = np.concatenate([[2], q]).astype(np.uint8)  # query p == target p+1
    qn = q.copy()
    qn[5] = 4
    b = ExtTasks.from_pairs([(q, t_same), (q, t_shift), (qn, t_same)], h0=[30, 30, 19], w=100, end_bonus=30)
    f = B.features_f32(b)
    K = min(20, B.SCAN)  # bases covered by the ungapped scans
    # task 0: exact match of 20 bases
    assert f[0, 0] == 20 / 128 and f[0, 1] == 20 / 256 and f[0, 2] == 30 / 64
    assert f[0, 3] == min(100, 20 + 30 - 6 + 1) / 128  # clamped band (ksw.c:399-407)
    assert f[0, 4] == 30 / 32 and f[0, 5] == (30 - 7) / 64 and f[0, 6] == (30 - 7) / 64
    assert f[0, 7] == (30 - 7 - 20) / 64
    assert f[0, 8] == K / 64 and f[0, 9] == K / 64 and f[0, 10] == K / 128 and f[0, 11] == K / 64
    assert f[0, 12] == K / 128 and f[0, 13] == min(K, 16) / 16 and f[0, 18] == 0 and f[0, 19] == 0
    qk = q[:K]
    assert f[0, 20] == (qk == 3).sum() / 128 and f[0, 23] == 0
    assert f[0, 21] == ((qk == 1) | (qk == 2)).sum() / 128
    assert f[0, 22] == ((q[:K] == 1) & (q[1:K + 1] == 2)).sum() / 32
    # task 1: the +1 diagonal matches everywhere
    assert f[1, 14] == K / 64
    # task 2: one N in the query at position 5 (mat[*][4] = -1)
    assert f[2, 23] == 1 / 8 and f[2, 12] == 5 / 128 and f[2, 8] == (K - 1) / 64
    assert f[2, 11] == (K - 1 - 1) / 64 and f[2, 9] == (K - 2) / 64


def test_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 0.0, 3.0e-3], np.float32)
    bits = B.bf16_bits(x)
    assert list(bits[:5]) == [0x3F80, 0x3F80, 0x3F82, 0xC020, 0x0000]
    assert abs(B.bf16_val(bits[5:])[0] - 3e-3) < 3e-3 * 2 ** -8


def test_features_stored_as_bf16(wl_small):
    b = wl_small[1]
    s = b.subset(np.arange(0, 400))
    fb = B.features(s)
    f = B.features_f32(s)
    assert fb.shape == (400, 32) and not fb[:, 24:].any()
    assert np.allclose(B.bf16_val(fb[:, :24]), f, rtol=2 ** -8, atol=0)


# ---------------- MLP ----------------

def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 24)).astype(np.float32)
    y = rng.integers(0, 5, n)
    return X, y


def test_grad64_matches_finite_differences():
    X, y = _data(40, 0)
    p = B.init_params(3).astype(np.float64)
    p[B.NPARAMS - 5:] = np.array([0.1, -0.2, 0.05, 0.0, 0.3])
    g = B.grad64(p, X, y)
    rng = np.random.default_rng(9)
    idx = np.concatenate([rng.choice(B.NPARAMS, 60, replace=False), np.arange(B.NPARAMS - 5, B.NPARAMS)])
    eps = 1e-6
    for k in idx:
        d = np.zeros(B.NPARAMS)
        d[k] = eps
        fd = (B.loss64(p + d, X, y) - B.loss64(p - d, X, y)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_mixed_precision_grad_close_to_exact():
    X, y = _data(256, 1)
    p = B.init_params(4)
    g, loss = B.grad(p, X, y)
    g64 = B.grad64(p, X, y)
    assert abs(loss - B.loss64(p, B.bf16(X), y)) < 2e-2
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 5e-2


def test_adam_first_step_is_sign_step():
    p = np.array([0.5, -0.25, 1.0], np.float32)
    g = np.array([0.3, -2.0, 1e-3], np.float32)
    m = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    p1, m1, v1 = B.adam(p, m, v, g, 1, 0.01)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-6)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)


def test_adam_constant_gradient_many_steps():
    """With a constant gradient the bias-corrected moments are exactly g and g^2
    at every step t, so each step moves by lr * g / (|g| + eps): after T steps
    p = p0 - T * lr * g / (|g| + eps).  A bias correction with the wrong step
    count (e.g. t stuck at 1, or t off by one) breaks this from step 2 on."""
    rng = np.random.default_rng(21)
    p0 = rng.normal(0, 1, 300).astype(np.float32)
    g = rng.normal(0, 1, 300).astype(np.float32)
    g[np.abs(g) < 1e-3] = 1e-3
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    lr, T = 1e-3, 25
    for t in range(1, T + 1):
        p, m, v = B.adam(p, m, v, g, t, lr)
    want = p0 - T * lr * g / (np.abs(g) + 1e-8)
    assert np.abs(p - want).max() < 2e-5
    # a bias correction stuck at step 1 drifts away from the closed form
    q, m2, v2 = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for t in range(1, T + 1):
        q, m2, v2 = B.adam(q, m2, v2, g, 1, lr)
    assert np.abs(q - want).max() > 1e-3
