"""Writes tests/golden/md_bsbolt.npz from the reference's own bwa_gen_cigar2
(bsbolt/External/BWA/bwa.c:199-351, compiled with the sources it links by
oracle/Makefile into oracle/_ref/libbwaref.so).  Run in the container that
holds the reference sources:  make -C oracle && python tests/golden/make_golden_md.py

Inputs: a 20 kb synthetic bisulfite index (tests/md_cases.py) and regions on
the four strands plus the boundary cases of the context window; each region
gets a partly unmethylated read with substitutions and indels.  Stored: the
genome (the test rebuilds pac/opac from it), the regions, both reads, and the
reference's score, CIGAR, NM, MD:Z, XB:Z and methylation counts.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import md_cases as M  # noqa: E402
from oracle import oracle as O  # noqa: E402

GLEN, SEED, N = 20000, 7, 700


def main():
    g, pac, opac, l_pac, crick_l = M.index(GLEN, SEED)
    rng = np.random.default_rng(11)
    mat = O.fill_scmat()
    rows = []
    for rb, re in M.regions(rng, N, GLEN, l_pac):
        q, oq = M.reads_for(rng, l_pac, opac, crick_l, rb, re)
        r = O.ref_gen_cigar2(mat, 6, 1, 6, 1, 100, l_pac, pac, opac, q, rb, re, crick_l, oq)
        if r is None:
            continue
        rows.append((rb, re, q, oq, r))
    cat = lambda xs, dt: np.concatenate([np.asarray(x, dt) for x in xs]) if xs else np.zeros(0, dt)
    offs = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.int64)
    qs = [r[2] for r in rows]
    oqs = [r[3] for r in rows]
    cgs = [r[4]["cigar"] for r in rows]
    mds = [(r[4]["md"] or "").encode() for r in rows]
    xbs = [(r[4]["xb"] or "").encode() for r in rows]
    np.savez_compressed(
        os.path.join(HERE, "md_bsbolt.npz"), genome=g, glen=GLEN, l_pac=l_pac, crick_l=crick_l,
        rb=np.array([r[0] for r in rows], np.int64), re=np.array([r[1] for r in rows], np.int64),
        query=cat(qs, np.uint8), q_off=offs(qs), oquery=cat(oqs, np.uint8), oq_off=offs(oqs),
        cigar=cat(cgs, np.uint32), cg_off=offs(cgs),
        md=np.frombuffer(b"".join(mds), np.uint8), md_off=offs(mds),
        xb=np.frombuffer(b"".join(xbs), np.uint8), xb_off=offs(xbs),
        has_md=np.array([r[4]["md"] is not None for r in rows]),
        nm=np.array([r[4]["NM"] for r in rows], np.int32), score=np.array([r[4]["score"] for r in rows], np.int32),
        meth=np.array([r[4]["meth"] for r in rows], np.int32),
        note=np.array("bwa_gen_cigar2 of the compiled reference, -A1 -B4 -O6,6 -E1,1, w=100"))
    print("md_bsbolt.npz", len(rows), "cases,", int(sum(r[4]["md"] is None for r in rows)), "without MD")


if __name__ == "__main__":
    main()
