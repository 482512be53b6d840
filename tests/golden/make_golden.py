"""Writes the golden fixtures of tests/golden/*.npz from the reference's own
ksw_extend2 / ksw_global2 (bsbolt/External/BWA/ksw.c), compiled by
oracle/Makefile into oracle/_ref/libkswref.so.  Run in the container that holds
the reference sources:  make -C oracle && python tests/golden/make_golden.py

Each fixture stores the task inputs, the scoring and the reference's outputs;
the inputs are seeded synthetic tasks (see the generators below) and, for the
realistic set, tasks produced by bsbolt_amd.workload (mem_chain2aln-style flanks
of simulated bisulfite reads).
"""
import itertools
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from bsbolt_amd.tasks import ExtTasks, GlbTasks, random_ext_tasks  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

SCORINGS = {
    # BSBolt Align defaults (Parser.py:47-81): -A1 -B4 -O6,6 -E1,1 -d100, end bonus -L30,30
    "bsbolt": dict(a=1, b=4, o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100),
    "nozdrop": dict(a=1, b=4, o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=0),
    "asym": dict(a=2, b=5, o_del=5, e_del=2, o_ins=3, e_ins=1, zdrop=40),
}


def mat_for(sc):
    return O.fill_scmat(sc["a"], sc["b"])


def run_ref_ext(t: ExtTasks, sc, mat):
    outs = {k: np.zeros(len(t), np.int32) for k in ("score", "qle", "tle", "gtle", "gscore", "max_off")}
    for i in range(len(t)):
        r = O.extend2(t.query_of(i), t.target_of(i), int(t.h0[i]), int(t.w[i]), int(t.end_bonus[i]),
                      sc["zdrop"], mat=mat, o_del=sc["o_del"], e_del=sc["e_del"], o_ins=sc["o_ins"],
                      e_ins=sc["e_ins"], use_ref=True)
        for k, v in zip(outs, r[:6]):
            outs[k][i] = v
    return outs


def save_ext(name, t: ExtTasks, sc, mat, note):
    outs = run_ref_ext(t, sc, mat)
    np.savez_compressed(os.path.join(OUT, name), **t.arrays(), **{"ref_" + k: v for k, v in outs.items()},
                        mat=mat, o_del=sc["o_del"], e_del=sc["e_del"], o_ins=sc["o_ins"], e_ins=sc["e_ins"],
                        zdrop=sc["zdrop"], note=np.array(note))
    print(name, len(t))


def edge_tasks(rng):
    """Hand-built corner cases: tiny lengths, empty target, N bases in both
    sequences, exact matches, all-mismatch rows, narrow/zero bands, h0 at the
    edges of the first-row/first-column ramps, long insertion/deletion tails."""
    pairs, h0, w, eb = [], [], [], []

    def add(q, t, hh, ww=100, e=30):
        pairs.append((np.array(q, np.uint8), np.array(t, np.uint8)))
        h0.append(hh); w.append(ww); eb.append(e)
    for ql in (1, 2, 3, 17, 64, 65, 128, 129, 200, 256, 257, 511, 700):
        q = rng.integers(0, 4, ql)
        for hh in (1, 6, 7, 8, 30, 120):
            add(q, q, hh)                                   # exact match
            add(q, [], hh)                                  # empty target
            add(q, (q + 1) % 4, hh)                         # all mismatch
            add(q, np.concatenate([q, rng.integers(0, 4, ql)]), hh, 100, 0)
        add(q, q, 1, 0)                                     # zero band
        add(q, q, 50, 1)
        add(np.where(rng.random(ql) < 0.2, 4, q), q, 40)     # N in query
        tq = q.copy(); tq[rng.random(ql) < 0.2] = 4
        add(q, tq, 40)                                      # N in target
        if ql > 10:
            add(q, np.concatenate([q[:ql // 2], rng.integers(0, 4, 9), q[ql // 2:]]), 25)   # insertion in target
            add(q, np.concatenate([q[:ql // 3], q[ql // 3 + 7:]]), 25)                     # deletion
            add(q, np.concatenate([q[:5], (q[5:] + 2) % 4]), 200)                          # z-drop after 5
            add(q, np.concatenate([q[:ql // 2], (q[ql // 2:] + 1) % 4]), 12, 8)            # narrow band
    return ExtTasks.from_pairs(pairs, h0, w, eb)


def brute_tasks():
    """Every query of length 1..3 over {A,C,G,T,N} against every target of
    length 0..3 over {A,C,G,T}, at h0 in {1, 7, 12} (w=100, end_bonus=30)."""
    qs = [q for L in (1, 2, 3) for q in itertools.product(range(5), repeat=L)]
    ts = [t for L in (0, 1, 2, 3) for t in itertools.product(range(4), repeat=L)]
    pairs, h0 = [], []
    for hh in (1, 7, 12):
        for q in qs:
            for t in ts:
                pairs.append((np.array(q, np.uint8), np.array(t, np.uint8)))
                h0.append(hh)
    return ExtTasks.from_pairs(pairs, h0, 100, 30)


def global_tasks(rng, n):
    pairs, ws = [], []
    for _ in range(n):
        ql = int(rng.integers(1, 160))
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = [int(x) % 4 for x in q]
        for k in range(len(t)):
            if rng.random() < 0.04:
                t[k] = int(rng.integers(0, 4))
        for _ in range(int(rng.integers(0, 3))):
            p = int(rng.integers(0, len(t) + 1)); L = int(rng.integers(1, 8))
            if rng.random() < 0.5:
                t[p:p] = list(rng.integers(0, 4, L))
            else:
                del t[p:p + L]
        if not t:
            t = [0]
        t = np.array(t, np.uint8)
        # band as bwa_gen_cigar2 chooses it (bwa.c:231-241), sometimes wider
        w = O.global_band(ql, len(t), 100)
        if rng.random() < 0.3:
            w += int(rng.integers(0, 30))
        pairs.append((q, t)); ws.append(w)
    return GlbTasks.from_pairs(pairs, ws)


def save_glb(name, t: GlbTasks, sc, mat):
    n = len(t)
    slot = int((t.qlen + t.tlen).max())
    score = np.zeros(n, np.int32); ncig = np.zeros(n, np.int32); cig = np.zeros((n, slot), np.uint32)
    for i in range(n):
        q = t.query[int(t.q_off[i]):int(t.q_off[i]) + int(t.qlen[i])]
        tt = t.target[int(t.t_off[i]):int(t.t_off[i]) + int(t.tlen[i])]
        s, cg = O.global2(q, tt, int(t.w[i]), mat=mat, o_del=sc["o_del"], e_del=sc["e_del"],
                          o_ins=sc["o_ins"], e_ins=sc["e_ins"], use_ref=True)
        score[i] = s; ncig[i] = len(cg); cig[i, :len(cg)] = cg
    np.savez_compressed(os.path.join(OUT, name), **{k: getattr(t, k) for k in
                                                   ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")},
                        ref_score=score, ref_n_cigar=ncig, ref_cigar=cig, mat=mat, o_del=sc["o_del"],
                        e_del=sc["e_del"], o_ins=sc["o_ins"], e_ins=sc["e_ins"])
    print(name, n)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libkswref.so missing: run `make -C oracle` where the reference sources are")
    rng = np.random.default_rng(20261202)
    sc = SCORINGS["bsbolt"]; mat = mat_for(sc)
    save_ext("ext_random_bsbolt.npz",
             random_ext_tasks(6000, seed=11, qmax=260, h0max=300, w_choices=(1, 2, 5, 8, 16, 32, 64, 100, 200),
                              eb_choices=(0, 5, 30), n_frac=0.01),
             sc, mat, "random flank-shaped tasks, BSBolt default scoring")
    save_ext("ext_edge_bsbolt.npz", edge_tasks(rng), sc, mat, "hand-built corner cases")
    save_ext("ext_brute_bsbolt.npz", brute_tasks(), sc, mat, "exhaustive tiny strings")
    sc2 = SCORINGS["nozdrop"]
    save_ext("ext_random_nozdrop.npz", random_ext_tasks(3000, seed=12, qmax=200, h0max=200,
                                                        w_choices=(8, 32, 100), eb_choices=(0, 30)),
             sc2, mat_for(sc2), "z-drop disabled")
    sc3 = SCORINGS["asym"]
    save_ext("ext_random_asym.npz", random_ext_tasks(3000, seed=13, qmax=200, h0max=200,
                                                     w_choices=(4, 16, 50, 100), eb_choices=(0, 10)),
             sc3, mat_for(sc3), "asymmetric gap costs, a=2 b=5")
    # a matrix that is not of the bwa_fill_scmat form (negative and positive N entries)
    sc4 = dict(SCORINGS["bsbolt"]); m4 = rng.integers(-6, 4, 25).astype(np.int8); m4[[0, 6, 12, 18]] = [3, 2, 3, 2]
    save_ext("ext_random_matrix.npz", random_ext_tasks(2000, seed=14, qmax=150, h0max=120, n_frac=0.05),
             sc4, m4, "irregular 5x5 matrix")
    save_glb("glb_random_bsbolt.npz", global_tasks(rng, 4000), sc, mat)
    save_glb("glb_random_asym.npz", global_tasks(rng, 1500), sc3, mat_for(sc3))


if __name__ == "__main__":
    main()
