"""Synthetic inputs for the NM / MD:Z / XB:Z layer (bwa_gen_cigar2's tail,
bsbolt/External/BWA/bwa.c:250-343): a small bisulfite index and regions on all
four strands of the doubled reference, each with a read as sequenced (partly
unmethylated, with substitutions and indels) and the converted read the
aligner saw.  Used by tests/golden/make_golden_md.py and the GPU tests."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
from bsbolt_amd.workload import bs_pac, make_genome  # noqa: E402
from oracle import oracle as O  # noqa: E402


def pack2(codes):
    """2-bit forward-only layout of the .pac files (first base in the top bits),
    l/4+1 bytes as bwa.c:430-431 reads them."""
    codes = np.asarray(codes, np.uint8)
    out = np.zeros(len(codes) // 4 + 1, np.uint8)
    idx = np.arange(len(codes))
    np.bitwise_or.at(out, idx >> 2, ((codes & 3) << ((~idx & 3) << 1)).astype(np.uint8))
    return out


def index(glen, seed):
    """(genome, pac, opac, l_pac, crick_l): pac = [C->T copy] ++ [G->A copy]
    (bntseq.c:262-269), opac = the unconverted genome twice (bwtindex.c:303-304)."""
    g = make_genome(glen, seed)
    pac = np.concatenate([bs_pac(g), np.zeros(1, np.uint8)])
    opac = pack2(np.concatenate([g, g]))
    return g, pac, opac, 2 * glen, glen


def regions(rng, n, glen, l_pac):
    """Regions [rb, re) on the four strands, plus the boundary cases of the
    window logic: the reference start, rb == l_pac, a window that bridges the
    strands, the last bases before the reference end."""
    out = []
    for _ in range(n):
        L = int(rng.integers(12, 170))
        q = int(rng.integers(4))
        rb = q * glen + int(rng.integers(0, glen - L))
        out.append((rb, rb + L))
    extra = [(0, 40), (1, 60), (2, 33), (l_pac, l_pac + 50), (l_pac + 1, l_pac + 30), (l_pac + 2, l_pac + 80),
             (l_pac - 45, l_pac - 1), (l_pac - 60, l_pac), (glen - 30, glen + 30), (glen, glen + 40),
             (l_pac + glen - 20, l_pac + glen + 20), (2 * l_pac - 90, 2 * l_pac - 3)]
    return out + extra


def reads_for(rng, l_pac, opac, crick_l, rb, re, p_unmeth=0.7, p_sub=0.03, p_indel=0.015):
    """(query, oquery) for region [rb, re): the reference bases of that strand,
    with the C (G on the Crick copy) of non-CpG context converted with
    p_unmeth, then substitutions and 1-3 base indels; query is oquery after the
    in-silico conversion matching the pac copy the region lies on."""
    ref = O.get_seq(l_pac, opac, rb, re).astype(np.uint8)
    is_crick = (rb >= crick_l and rb < l_pac) or rb >= crick_l + l_pac
    rbase, mbase = (2, 0) if is_crick else (1, 3)
    seq = ref.copy()
    conv = (seq == rbase) & (rng.random(len(seq)) < p_unmeth)
    seq[conv] = mbase
    out = []
    k = 0
    while k < len(seq):
        u = rng.random()
        if u < p_sub:
            out.append((int(seq[k]) + 1 + int(rng.integers(3))) & 3)
        elif u < p_sub + p_indel / 2:
            k += int(rng.integers(1, 4)) - 1
        elif u < p_sub + p_indel:
            out.append(int(seq[k]))
            out.extend(int(x) for x in rng.integers(0, 4, int(rng.integers(1, 4))))
        else:
            out.append(int(seq[k]))
        k += 1
    oq = np.array(out if out else [0], np.uint8)
    ct = (rb < crick_l) or (l_pac <= rb < l_pac + crick_l)  # regions whose pac copy reads C->T
    q = oq.copy()
    if ct:
        q[q == 1] = 3
    else:
        q[q == 2] = 0
    return q, oq


def cigar_stage(mat, l_pac, pac, query, rb, re, w_=100, o_del=6, e_del=1, o_ins=6, e_ins=1):
    """The CIGAR bwa_gen_cigar2 computes for a region (bwa.c:211-248): query and
    reference window reversed when rb >= l_pac, band from orc_global_band,
    ksw_global2 through the oracle.  None where the reference returns no CIGAR."""
    if rb >= re or (rb < l_pac < re) or len(query) == 0:
        return None
    rseq = O.get_seq(l_pac, pac, rb, re)
    if len(rseq) != re - rb:
        return None
    q = np.asarray(query, np.uint8)
    if rb >= l_pac:
        q, rseq = q[::-1].copy(), rseq[::-1].copy()
    w = O.global_band(len(q), len(rseq), w_, int(mat[0]), o_del, e_del, o_ins, e_ins)
    _, cig = O.global2(q, rseq, w, mat=mat, o_del=o_del, e_del=e_del, o_ins=o_ins, e_ins=e_ins)
    return np.asarray(cig, np.uint32)
