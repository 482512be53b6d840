"""Pins the CPU oracle (oracle/ksw_oracle.c) against the reference's own outputs
(golden fixtures written from the compiled bsbolt/External/BWA/ksw.c), against
closed forms and, where the reference sources are present, against the compiled
reference directly."""
import itertools

import numpy as np
import pytest

from conftest import EXT_GOLDEN, GLB_GOLDEN, golden_ext, golden_glb
from oracle import oracle as O

FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")


@pytest.mark.parametrize("name", EXT_GOLDEN)
def test_oracle_extend_matches_reference_golden(name):
    t, sc, ref = golden_ext(name)
    out, _ = O.extend2_batch(t, nthreads=4, scoring=sc)
    for k in FIELDS:
        bad = np.nonzero(out[k] != ref[k])[0]
        assert len(bad) == 0, f"{name}: {k} differs at tasks {bad[:10]}"


@pytest.mark.parametrize("name", GLB_GOLDEN)
def test_oracle_global_matches_reference_golden(name):
    t, sc, rs, rn, rc = golden_glb(name)
    score, ncig, cig, _ = O.global2_batch(t, rc.shape[1], scoring=sc)
    assert np.array_equal(score, rs)
    assert np.array_equal(ncig, rn)
    for i in range(len(t)):
        assert np.array_equal(cig[i, :ncig[i]], rc[i, :rn[i]]), f"task {i}"


def test_fill_scmat_closed_form():
    m = O.fill_scmat(1, 4).reshape(5, 5)
    # bwa.c:169-178: a on the ACGT diagonal, -b off it, -1 for N
    assert (np.diag(m)[:4] == 1).all()
    assert m[4, 4] == -1 and (m[4] == -1).all() and (m[:, 4] == -1).all()
    off = m[:4, :4][~np.eye(4, dtype=bool)]
    assert (off == -4).all()


@pytest.mark.parametrize("qlen", [0, 1, 5, 6, 7, 30, 150, 205, 206, 400])
def test_cal_max_gap_closed_form(qlen):
    # BSBolt defaults: a=1, o=6, e=1, w=100 -> min(max(qlen-5, 1), 200)
    assert O.cal_max_gap(qlen) == min(max(qlen - 5, 1), 200)
    # e=2, o=3, a=2: trunc((2q-3)/2 + 1) = trunc(q - 0.5) = q - 1 for q >= 1
    assert O.cal_max_gap(qlen, a=2, o_del=3, e_del=2, o_ins=3, e_ins=2, w=50) == min(max(qlen - 1, 1), 100)


def test_global_band_closed_form():
    # bwa.c:231-241 with a=1,o=6,e=1: gap = max(((l+1)>>1) - 5, 1)
    for lq, rl in [(100, 100), (100, 103), (151, 140), (20, 21), (3, 9)]:
        gap = max(((lq + 1) >> 1) - 5, 1)
        d = abs(rl - lq)
        w = min((gap + d + 1) >> 1, 100)
        assert O.global_band(lq, rl, 100) == max(w, d + 3)


def test_get_seq_forward_reverse_and_bridge():
    rng = np.random.default_rng(0)
    seq = rng.integers(0, 4, 37).astype(np.uint8)
    pac = np.zeros(37 // 4 + 1, np.uint8)
    for i, c in enumerate(seq):                       # bntseq.c:236 _set_pac
        pac[i >> 2] |= c << ((~i & 3) << 1)
    L = len(seq)
    assert np.array_equal(O.get_seq(L, pac, 3, 20), seq[3:20])
    rc = (3 - seq)[::-1]                               # reverse-complement strand at [L, 2L)
    assert np.array_equal(O.get_seq(L, pac, L + 5, L + 30), rc[5:30])
    assert np.array_equal(O.get_seq(L, pac, 20, 3), seq[3:20])       # swapped ends
    assert len(O.get_seq(L, pac, L - 3, L + 3)) == 0                 # bridges the strands
    assert np.array_equal(O.get_seq(L, pac, -4, 5), seq[:5])         # clamped


def test_alphabet_and_bisulfite_conversion():
    assert [O.nt4(c) for c in "ACGTacgtN-X"] == [0, 1, 2, 3, 0, 1, 2, 3, 4, 5, 4]
    assert O.bs_convert("ACGTCCGA", 0) == "ATGTTTGA"   # C->T
    assert O.bs_convert("ACGTCCGA", 1) == "ACATCCAA"   # G->A


@pytest.mark.parametrize("ql,h0", [(1, 1), (10, 5), (57, 20), (150, 100)])
def test_extend_exact_match_closed_form(ql, h0):
    q = np.random.default_rng(ql).integers(0, 4, ql).astype(np.uint8)
    s, qle, tle, gtle, gs, mo, cells = O.extend2(q, q, h0)
    assert (s, qle, tle, gtle, gs, mo) == (h0 + ql, ql, ql, ql, h0 + ql, 0)


def test_extend_empty_target_closed_form():
    q = np.array([0, 1, 2], np.uint8)
    assert O.extend2(q, np.zeros(0, np.uint8), 17)[:6] == (17, 0, 0, 0, -1, 0)


def test_extend_all_mismatch_first_row_ramp():
    # h0=10, 1-base query, target of mismatches: row 0 cell = max(G[0]+s, E, F) with G[0]=h0
    q = np.array([0], np.uint8)
    t = np.array([1, 1, 1], np.uint8)
    s, qle, tle, gtle, gs, mo, cells = O.extend2(q, t, 10)
    # row 0: H(0,0) = 10-4 = 6 < h0, so the local score stays h0, but column qlen-1 is
    # reached with 6 (gscore 6, gtle 1).  Row 1: M = H(0,-1)+s = (10-7)-4 < 0, E = F = 0,
    # so the row maximum is 0 and the loop ends (ksw.c:454).
    assert (s, qle, tle) == (10, 0, 0)
    assert gs == 6 and gtle == 1


def test_extend_zdrop_row():
    # 20 matches then mismatches: with zdrop=10 the loop stops once max - m - gap term > 10
    rng = np.random.default_rng(5)
    q = rng.integers(0, 4, 60).astype(np.uint8)
    t = np.concatenate([q[:20], (q[20:] + 1) % 4]).astype(np.uint8)
    full = O.extend2(q, t, 30, zdrop=0)
    cut = O.extend2(q, t, 30, zdrop=10)
    assert full[:3] == cut[:3] == (50, 20, 20)
    assert cut[6] < full[6]   # fewer cells evaluated after the z-drop exit


def test_extend_invariants_random():
    rng = np.random.default_rng(7)
    for _ in range(500):
        ql = int(rng.integers(1, 80)); tl = int(rng.integers(0, 160))
        q = rng.integers(0, 5, ql).astype(np.uint8); t = rng.integers(0, 4, tl).astype(np.uint8)
        h0 = int(rng.integers(1, 100)); w = int(rng.integers(0, 120))
        s, qle, tle, gtle, gs, mo, cells = O.extend2(q, t, h0, w)
        assert s >= h0 and 0 <= qle <= ql and 0 <= tle <= tl and gs <= s
        assert (qle == 0 and tle == 0) == (s == h0)
        assert 0 <= mo and cells <= ql * tl


def test_extend_cells_closed_form_full_band():
    # exact match with h0 large: no cell is ever zero, so the band is only the
    # diagonal strip, cells(row i) = min(i+w+1, qlen) - max(i-w, 0)
    ql, w = 12, 3
    q = np.arange(ql, dtype=np.uint8) % 4
    s, *_, cells = O.extend2(q, q, 200, w)
    assert cells == sum(min(i + w + 1, ql) - max(i - w, 0) for i in range(ql))


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference (oracle/_ref) not present")
def test_oracle_vs_compiled_reference_brute_force_tiny():
    """All query strings of length <= 4 over ACGTN against all targets of length
    <= 4 over ACGT, two h0 values, default scoring."""
    qs = [np.array(q, np.uint8) for L in range(1, 5) for q in itertools.product(range(5), repeat=L)]
    ts = [np.array(t, np.uint8) for L in range(0, 5) for t in itertools.product(range(4), repeat=L)]
    rng = np.random.default_rng(3)
    sel_q = [qs[i] for i in rng.choice(len(qs), 250, replace=False)]
    for h0 in (3, 9):
        for q in sel_q:
            for t in ts:
                a = O.extend2(q, t, h0)[:6]
                b = O.extend2(q, t, h0, use_ref=True)[:6]
                assert a == b, (q, t, h0, a, b)


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference (oracle/_ref) not present")
def test_oracle_vs_compiled_reference_random_params():
    rng = np.random.default_rng(99)
    for _ in range(3000):
        ql = int(rng.integers(1, 130)); q = rng.integers(0, 5, ql).astype(np.uint8)
        t = np.concatenate([np.where(rng.random(ql) < 0.9, q % 4, rng.integers(0, 4, ql)),
                            rng.integers(0, 4, int(rng.integers(0, 80)))]).astype(np.uint8)
        kw = dict(mat=O.fill_scmat(int(rng.integers(1, 4)), int(rng.integers(1, 7))),
                  o_del=int(rng.integers(0, 9)), e_del=int(rng.integers(1, 4)),
                  o_ins=int(rng.integers(0, 9)), e_ins=int(rng.integers(1, 4)))
        args = (q, t, int(rng.integers(1, 250)), int(rng.integers(0, 210)), int(rng.integers(0, 40)),
                int(rng.choice([0, 15, 100])))
        assert O.extend2(*args, **kw)[:6] == O.extend2(*args, **kw, use_ref=True)[:6]


@pytest.mark.skipif(not O.ref_available(), reason="compiled reference (oracle/_ref) not present")
def test_global2_batch_threads_and_compiled_reference():
    """The threaded batch driver gives the same scores and CIGARs on 1 and 8
    threads, through the restatement and through the compiled reference's
    ksw_global2, over 9000 tasks (more than two 4096-task chunks) with indels
    and bands from 0 to 40."""
    from bsbolt_amd.tasks import GlbTasks
    rng = np.random.default_rng(5)
    pairs, ws = [], []
    for _ in range(9000):
        ql = int(rng.integers(1, 120))
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = list(np.where(rng.random(ql) < 0.05, rng.integers(0, 4, ql), q % 4))
        for _ in range(int(rng.integers(0, 3))):
            p = int(rng.integers(0, len(t) + 1)); L = int(rng.integers(1, 4))
            if rng.random() < 0.5:
                t[p:p] = list(rng.integers(0, 4, L))
            else:
                del t[p:p + L]
        t = np.array(t or [1], np.uint8)
        pairs.append((q, t)); ws.append(abs(len(t) - ql) + int(rng.integers(0, 40)))
    tk = GlbTasks.from_pairs(pairs, ws)
    slot = int((tk.qlen + tk.tlen).max())
    base = O.global2_batch(tk, slot)
    for kw in (dict(nthreads=8), dict(use_ref=True), dict(nthreads=8, use_ref=True)):
        got = O.global2_batch(tk, slot, **kw)
        for k in range(3):
            assert np.array_equal(base[k], got[k]), kw
