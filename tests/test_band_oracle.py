"""CPU checks of the learned-band oracle (oracle/band_oracle.c, oracle/band_mlp.py):
the narrow-band certificate against the compiled reference, the task features
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


def _check_sound(b, wn):
    want = _reference(b)
    acc = 0
    whys = np.zeros(4, int)
    for i in range(len(b)):
        r, out, cells, why = B.narrow_extend(b.query_of(i), b.target_of(i), int(b.h0[i]), wn, w=int(b.w[i]),
                                             end_bonus=int(b.end_bonus[i]))
        if r == 1:
            acc += 1
            got = dict(zip(FIELDS, out))
            for k in FIELDS:
                assert got[k] == want[k][i], f"task {i} wn {wn}: {k} {got[k]} != {want[k][i]}"
            assert cells == want["cells"][i], f"task {i} wn {wn}: cells"
        else:
            whys[why] += 1
    return acc, whys


_WHYS = {}


@pytest.mark.parametrize("wn", [8, 16, 32, 64])
def test_certificate_sound_on_workloads(wl_small, wn):
    tot_acc = 0
    tot_why = np.zeros(4, int)
    for b in wl_small:
        acc, whys = _check_sound(b, wn)
        tot_acc += acc
        tot_why += whys
    n = sum(len(b) for b in wl_small)
    assert tot_acc > 0.05 * n, f"wn {wn}: accepted only {tot_acc} of {n}"
    _WHYS[wn] = tot_why


def test_certificate_every_rejection_reason_occurs(wl_small):
    if len(_WHYS) < 4:
        for wn in (8, 16, 32, 64):
            _WHYS[wn] = sum(_check_sound(b, wn)[1] for b in wl_small)
    tot = sum(_WHYS.values())
    assert all(tot[1:] > 0), _WHYS


@pytest.mark.parametrize("seed", [11, 12])
def test_certificate_sound_on_random_tasks(seed):
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(3000, seed=seed, qmax=200, h0max=120, w_choices=(16, 32, 64, 100, 200),
                         eb_choices=(0, 5, 30), n_frac=0.002)
    for wn in (8, 16, 32, 64):
        _check_sound(b, wn)


def test_certificate_full_band_is_the_reference_run():
    from bsbolt_amd.tasks import random_ext_tasks
    b = random_ext_tasks(500, seed=5, qmax=120, h0max=100, w_choices=(8, 16, 32), eb_choices=(0, 30))
    want = _reference(b)
    for i in range(len(b)):
        r, out, cells, why = B.narrow_extend(b.query_of(i), b.target_of(i), int(b.h0[i]), 64, w=int(b.w[i]),
                                             end_bonus=int(b.end_bonus[i]))
        assert r == 1
        assert out == tuple(int(want[k][i]) for k in FIELDS) and cells == want["cells"][i]


def test_certificate_ramp_rejection():
    # h0 = 60: the first-row insertion values h0-7-(j-1) are still positive past column 9
    rng = np.random.default_rng(3)
    q = rng.integers(0, 4, 50).astype(np.uint8)
    r, out, cells, why = B.narrow_extend(q, q, 60, 8)
    assert r == 0 and why == 2
    # h0 = 5 leaves no ramp, and a 12-base exact match keeps every cell at
    # distance 8 from the diagonal at 0 (its deletion/insertion arms lose 6+1 per base)
    r, out, cells, why = B.narrow_extend(q[:12], q[:12], 5, 8)
    assert r == 1 and out[0] == 5 + 12
    # with h0 = 14 the same match scores 26 and its insertion arm (H - 7 - d at
    # d columns right of the diagonal) leaves the band at the end of a row
    r, out, cells, why = B.narrow_extend(q[:12], q[:12], 14, 8)
    assert r == 0 and why == 3


def test_labels_are_smallest_accepted_class(wl_small):
    b = wl_small[0]
    idx = np.arange(0, len(b), 7)
    s = b.subset(idx)
    lab = B.labels(s)
    for i in range(len(s)):
        acc = [B.narrow_extend(s.query_of(i), s.target_of(i), int(s.h0[i]), wn, w=int(s.w[i]),
                               end_bonus=int(s.end_bonus[i]))[0] == 1 for wn in B.WIDTHS]
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
