"""Workload generator: synthetic bisulfite genomes and reads, turned into the
ksw_extend2 / ksw_global2 task batches that BSBolt's aligner would issue.

Python front end of libbsxwl.so (bsbolt_amd/csrc/workload.cpp).  The
extension rounds of mem_chain2aln (bsbolt/External/BWA/bwamem.c:636-790) need
the results of earlier rounds (right-flank h0 = left score, w<<1 retries, seed
containment), so `build_ext_batch` takes an `extend` callable: in the product it
is the GPU engine (bsx.Engine.extend).

Named workloads (sizes in bases; seeds fixed so every run sees the same batch):
  cfg0  2 Mb genome,  20k single-end 100 bp reads
  cfg1  50 Mb genome, 1M paired-end 2x150 pairs, insert ~ N(350, 50)
  cfg2  training set for the band predictor: cfg1 recipe, disjoint seeds
  cfg3  100 Mb genome, 500k single-end 250 bp reads, 1% subs, 0.5% indels (geometric, <= 30)
  cfg4  100 Mb genome, RRBS: 2M single-end reads from MspI fragments, 40-100 bp, 20% duplicates
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from .tasks import ExtTasks, GlbTasks, RESULT_FIELDS

HERE = os.path.dirname(os.path.abspath(__file__))
WL_PATH = os.path.join(HERE, "libbsxwl.so")

# BSBolt Align defaults (bsbolt/Utils/Parser.py:40-106 -> bwa mem argv,
# bsbolt/Utils/Launcher.py:77-108): -A1 -B4 -O6,6 -E1,1 -L30,30 -w100 -d100 -k19
DEFAULT_OPT = dict(a=1, b=4, o_del=6, e_del=1, o_ins=6, e_ins=1, pen_clip5=30, pen_clip3=30, w=100,
                   zdrop=100, min_seed_len=19)


@dataclass
class SimConfig:
    genome_len: int
    genome_seed: int
    n_frag: int
    read_seed: int
    read_len: int = 100
    paired: bool = False
    insert_mean: float = 350.0
    insert_sd: float = 50.0
    p_cpg: float = 0.8
    p_ch: float = 0.01
    sub: float = 0.005
    indel: float = 0.0005
    geometric_indels: bool = False
    trim_min: int = 0
    rrbs: bool = False
    dup: float = 0.0
    p_shadow: float = 0.7        # reads that also extend one shadow chain
    p_shadow_len: float = 0.25   # shadow seed length = 19 + Geometric(p) - 1
    p_at: float = 0.295          # 41% GC
    cpg_frac: float = 0.01       # CpG share of dinucleotides
    opt: dict = field(default_factory=lambda: dict(DEFAULT_OPT))

    def scaled(self, frac):
        """Same recipe with a fraction of the reads (and genome >= 1 Mb)."""
        import dataclasses
        return dataclasses.replace(self, n_frag=max(1, int(self.n_frag * frac)),
                                   genome_len=max(1_000_000, int(self.genome_len * min(1.0, frac * 4))))


CONFIGS = {
    "cfg0": SimConfig(2_000_000, 0, 20_000, 100, read_len=100),
    "cfg1": SimConfig(50_000_000, 1, 1_000_000, 101, read_len=150, paired=True),
    "cfg2": SimConfig(50_000_000, 2, 500_000, 102, read_len=150, paired=True),
    "cfg3": SimConfig(100_000_000, 3, 500_000, 103, read_len=250, sub=0.01, indel=0.005,
                      geometric_indels=True),
    "cfg4": SimConfig(100_000_000, 4, 2_000_000, 104, read_len=100, trim_min=40, rrbs=True, dup=0.2,
                      p_cpg=0.7),
}

_wl = None
_P = C.c_void_p


def wl():
    global _wl
    if _wl is None:
        if not os.path.exists(WL_PATH):
            raise RuntimeError(f"libbsxwl.so not built ({WL_PATH}); run __graft_entry__.build()")
        L = C.CDLL(WL_PATH)
        L.wl_genome.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, _P]
        L.wl_bs_pac.argtypes = [_P, C.c_int64, _P]
        L.wl_create.restype = _P
        L.wl_create.argtypes = [_P, _P, C.c_int64]
        L.wl_destroy.argtypes = [_P]
        L.wl_simulate.restype = C.c_int64
        L.wl_simulate.argtypes = [_P, _P, C.c_uint64, C.c_int64, _P, _P, _P, _P, _P, _P, C.c_int64, _P]
        L.wl_read_len.restype = C.c_int32
        L.wl_read_len.argtypes = [_P, C.c_int64]
        L.wl_read_seq.argtypes = [_P, C.c_int64, _P]
        L.wl_read_oseq.argtypes = [_P, C.c_int64, _P]
        L.wl_chains_init.restype = C.c_int64
        L.wl_chains_init.argtypes = [_P, C.c_int64, _P, _P, _P, _P, _P]
        L.wl_round_emit.restype = C.c_int64
        L.wl_round_emit.argtypes = [_P]
        L.wl_round_sizes.argtypes = [_P, _P, _P]
        L.wl_round_tasks.argtypes = [_P] + [_P] * 11
        L.wl_round_consume.argtypes = [_P] + [_P] * 6
        L.wl_n_calls.restype = C.c_int64
        L.wl_n_calls.argtypes = [_P]
        L.wl_regions.restype = C.c_int64
        L.wl_regions.argtypes = [_P, C.c_int64, _P, C.c_int64]
        L.wl_n_reads.restype = C.c_int64
        L.wl_n_reads.argtypes = [_P]
        L.wl_get_seq.restype = C.c_int64
        L.wl_get_seq.argtypes = [_P, C.c_int64, C.c_int64, _P]
        L.wl_assess_conversion.restype = C.c_int
        L.wl_assess_conversion.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_float]
        L.wl_bseq_records.restype = C.c_int
        L.wl_bseq_records.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_float, _P]
        _wl = L
    return _wl


def _p(a):
    return a.ctypes.data


def make_genome(length, seed, p_at=0.295, cpg=0.01):
    g = np.zeros(length, np.uint8)
    wl().wl_genome(seed, length, p_at, cpg, _p(g))
    return g


def bs_pac(genome):
    """Forward-only 2-bit pac of [genome C->T] ++ [genome G->A]; l_pac = 2*len."""
    pac = np.zeros((2 * len(genome) + 3) // 4, np.uint8)
    wl().wl_bs_pac(_p(genome), len(genome), _p(pac))
    return pac


class Workload:
    """Genome + bisulfite index + simulated reads + their seeds."""

    def __init__(self, cfg: SimConfig):
        self.cfg = cfg
        L = wl()
        self.genome = make_genome(cfg.genome_len, cfg.genome_seed, cfg.p_at, cfg.cpg_frac)
        self.pac = bs_pac(self.genome)
        o = cfg.opt
        self.opt = np.array([o[k] for k in ("a", "b", "o_del", "e_del", "o_ins", "e_ins", "pen_clip5",
                                            "pen_clip3", "w", "zdrop", "min_seed_len")], np.int32)
        self.ctx = L.wl_create(_p(self.opt), _p(self.pac), cfg.genome_len)
        simcfg = np.array([cfg.read_len, int(cfg.paired), cfg.insert_mean, cfg.insert_sd, cfg.p_cpg, cfg.p_ch,
                           cfg.sub, cfg.indel, int(cfg.geometric_indels), cfg.trim_min, int(cfg.rrbs), cfg.dup, cfg.p_shadow, cfg.p_shadow_len],
                          np.float64)
        cap = cfg.n_frag * (2 if cfg.paired else 1) * (cfg.read_len // max(1, o["min_seed_len"]) + 2)
        self.seed_read = np.zeros(cap, np.int32)
        self.seed_chain = np.zeros(cap, np.int32)
        self.seed_rbeg = np.zeros(cap, np.int64)
        self.seed_qbeg = np.zeros(cap, np.int32)
        self.seed_len = np.zeros(cap, np.int32)
        ns = C.c_int64(0)
        self.n_reads = L.wl_simulate(self.ctx, _p(self.genome), cfg.read_seed, cfg.n_frag, _p(simcfg),
                                     _p(self.seed_read), _p(self.seed_chain), _p(self.seed_rbeg), _p(self.seed_qbeg),
                                     _p(self.seed_len), cap, C.byref(ns))
        if self.n_reads < 0:
            raise ValueError("simulation failed (no RRBS fragments?)")
        n = ns.value
        self.seed_read, self.seed_rbeg = self.seed_read[:n], self.seed_rbeg[:n]
        self.seed_chain = self.seed_chain[:n]
        self.seed_qbeg, self.seed_len = self.seed_qbeg[:n], self.seed_len[:n]
        self.l_pac = 2 * cfg.genome_len

    def close(self):
        if getattr(self, "ctx", None):
            wl().wl_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def read(self, i):
        n = wl().wl_read_len(self.ctx, i)
        out = np.zeros(n, np.uint8)
        wl().wl_read_seq(self.ctx, i, _p(out))
        return out

    def oread(self, i):
        """Read i as sequenced (before the aligner's in-silico conversion)."""
        n = wl().wl_read_len(self.ctx, i)
        out = np.zeros(n, np.uint8)
        wl().wl_read_oseq(self.ctx, i, _p(out))
        return out

    def opac(self):
        """The unconverted reference in the .pac 2-bit layout, both copies
        (bwtindex.c:303-304 writes it as .opac): l_pac/4+1 bytes."""
        return pack2(np.concatenate([self.genome, self.genome]))

    def get_seq(self, beg, end):
        out = np.zeros(abs(end - beg) + 1, np.uint8)
        n = wl().wl_get_seq(self.ctx, beg, end, _p(out))
        return out[:n]

    def regions(self, i, cap=64):
        out = np.zeros((cap, 9), np.int64)
        n = wl().wl_regions(self.ctx, i, _p(out), cap)
        return out[:min(n, cap)]

    def build_ext_batch(self, extend, max_rounds=64, on_round=None, seeds=None):
        """Runs mem_chain2aln over every read in rounds; `extend(ExtTasks) ->
        dict of result arrays`.  Returns (all tasks issued as one ExtTasks,
        their results, per-task side 0=left/1=right, per-task round); the read
        of each task is left in self.task_read.  `seeds` = (read, chain, rbeg,
        qbeg, len) arrays replaces the simulator's seeds (e.g. chains taken
        from the reference aligner's own trace); a seed's score is its length,
        as mem_chain sets it for SMEM seeds."""
        L = wl()
        sr, sc, sb, sq, sl = (self.seed_read, self.seed_chain, self.seed_rbeg, self.seed_qbeg, self.seed_len) \
            if seeds is None else (np.ascontiguousarray(seeds[0], np.int32), np.ascontiguousarray(seeds[1], np.int32),
                                   np.ascontiguousarray(seeds[2], np.int64), np.ascontiguousarray(seeds[3], np.int32),
                                   np.ascontiguousarray(seeds[4], np.int32))
        nch = L.wl_chains_init(self.ctx, len(sr), _p(sr), _p(sc), _p(sb), _p(sq), _p(sl))
        # read of each chain: chains are consecutive runs of equal (read, chain) keys
        key_change = np.ones(len(sr), bool)
        key_change[1:] = (sr[1:] != sr[:-1]) | (sc[1:] != sc[:-1])
        chain_read = sr[key_change]
        assert len(chain_read) == nch
        parts, results, sides, rounds, treads = [], [], [], [], []
        for r in range(max_rounds):
            n = L.wl_round_emit(self.ctx)
            if n == 0:
                break
            qt, tt = C.c_int64(0), C.c_int64(0)
            L.wl_round_sizes(self.ctx, C.byref(qt), C.byref(tt))
            q = np.zeros(qt.value + 1, np.uint8)
            t = np.zeros(tt.value + 1, np.uint8)
            qo = np.zeros(n, np.uint64); to = np.zeros(n, np.uint64)
            ql, tl, h0, w, eb, cid, side = [np.zeros(n, np.int32) for _ in range(7)]
            L.wl_round_tasks(self.ctx, _p(q), _p(qo), _p(ql), _p(t), _p(to), _p(tl), _p(h0), _p(w), _p(eb),
                             _p(cid), _p(side))
            tasks = ExtTasks(q, qo, ql, t, to, tl, h0, w, eb)
            res = extend(tasks)
            L.wl_round_consume(self.ctx, *[_p(np.ascontiguousarray(res[k], np.int32)) for k in RESULT_FIELDS])
            parts.append(tasks); results.append({k: np.asarray(res[k]) for k in RESULT_FIELDS})
            sides.append(side); rounds.append(np.full(n, r, np.int32)); treads.append(chain_read[cid])
            if on_round:
                on_round(r, n)
        else:
            raise RuntimeError("mem_chain2aln rounds did not converge")
        batch = concat_ext(parts)
        self.task_read = np.concatenate(treads) if treads else np.zeros(0, np.int32)
        res = {k: np.concatenate([x[k] for x in results]) if results else np.zeros(0, np.int32)
               for k in RESULT_FIELDS}
        return batch, res, (np.concatenate(sides) if sides else np.zeros(0, np.int32)), \
            (np.concatenate(rounds) if rounds else np.zeros(0, np.int32))


def infer_bw(l1, l2, score, a, q, r):
    """Band that can hold the gaps of a local hit of `score` (bwamem.c:796-803),
    vectorised; C's (int) cast of the double quotient truncates toward zero."""
    l1, l2, score = (np.asarray(x, np.int64) for x in (l1, l2, score))
    w = np.trunc((np.minimum(l1, l2) * a - score - q) / r + 2.0).astype(np.int64)
    w = np.maximum(w, np.abs(l1 - l2))
    return np.where((l1 == l2) & (l1 * a - score < ((q + r - a) << 1)), 0, w)


def cigar_band(w_, l_query, rlen, mat0=1, o_del=6, e_del=1, o_ins=6, e_ins=1):
    """Global band of bwa_gen_cigar2 (bwa.c:231-241), vectorised."""
    l_query, rlen, w_ = (np.asarray(x, np.int64) for x in (l_query, rlen, w_))
    half = (l_query + 1) >> 1
    max_ins = np.trunc((half * mat0 - o_ins) / e_ins + 1.0).astype(np.int64)
    max_del = np.trunc((half * mat0 - o_del) / e_del + 1.0).astype(np.int64)
    max_gap = np.maximum(np.maximum(max_ins, max_del), 1)
    d = np.abs(rlen - l_query)
    w = (max_gap + d + 1) >> 1
    w = np.minimum(w, w_)
    return np.maximum(w, d + 3)


def build_glb_batch(wl, global_fn, min_score=0, mat=None, opt=None, primary_only=True, return_regions=False):
    """CIGAR pass over every alignment region of a workload (see cigar_tries),
    in rounds.  primary_only keeps each read's best region, the one its SAM
    record comes from.  Returns what cigar_tries returns; with return_regions
    also a dict of per-region arrays read, qb, qe, rb, re."""
    regs = []
    for i in range(wl.n_reads):
        rs_ = wl.regions(i)
        if primary_only and len(rs_):  # the SAM record of a read comes from its best region
            rs_ = rs_[np.argmax(rs_[:, 4]):][:1]
        for r in rs_:
            qb, qe, rb, re, score, truesc, w = (int(x) for x in r[:7])
            if score < min_score or qe <= qb or rb >= re:
                continue
            regs.append((i, qb, qe, rb, re, truesc, w))
    out = cigar_tries(regs, wl.read, wl.get_seq, wl.l_pac, global_fn, mat=mat, opt=opt)
    if return_regions:
        R = np.array([r for r in regs], np.int64).reshape(-1, 7)
        reg = {"read": R[:, 0].copy()}
        for k, c in (("qb", 1), ("qe", 2), ("rb", 3), ("re", 4)):
            reg[k] = R[:, c].copy()
        out = out + (reg,)
    return out


def cigar_tries(regs, read_fn, get_seq, l_pac, global_fn, mat=None, opt=None, trace=None):
    """The CIGAR tries of mem_reg2aln (bwamem.c:1197-1233) and bwa_gen_cigar2
    (bwa.c:199-248) for a list of regions (read, qb, qe, rb, re, truesc, w),
    run in rounds: band w2 = max of the inferred deletion/insertion bands
    (capped by the region's band when above -w), then up to 3 tries doubling w2
    (at most 4w) while the global score stays below truesc - a and changes.
    Each try reverses query and reference when rb >= l_pac (leftmost indels)
    and runs ksw_global2 at the band bwa_gen_cigar2 derives from w2; a region
    with w2 == 0 and equal lengths is scored ungapped without DP (bwa.c:222-230).

    read_fn(i) gives read i's converted codes, get_seq(rb, re) the reference
    window; `global_fn(GlbTasks) -> (score, n_cigar, cigar[n, slot])` runs one
    round.  Returns (all issued tasks as one GlbTasks, their scores, n_cigar,
    CIGAR lists, region index per task, round per task, final score per
    region).  A `trace` list receives (region, try, w2 passed to
    bwa_gen_cigar2, score) for every try, in the order of the tries."""
    o = dict(DEFAULT_OPT) if opt is None else opt
    mat = fill_matrix(o["a"], o["b"]) if mat is None else np.asarray(mat, np.int8).reshape(25)
    nreg = len(regs)
    R = np.array([r[1:] for r in regs], np.int64).reshape(-1, 6)
    qlen = R[:, 1] - R[:, 0]
    rlen = R[:, 3] - R[:, 2]
    truesc = R[:, 4]
    tmp = infer_bw(qlen, rlen, truesc, o["a"], o["o_del"], o["e_del"])
    w2 = np.maximum(infer_bw(qlen, rlen, truesc, o["a"], o["o_ins"], o["e_ins"]), tmp)
    w2 = np.where(w2 > o["w"], np.minimum(w2, R[:, 5]), w2)
    last = np.full(nreg, -(1 << 30), np.int64)
    tries = np.zeros(nreg, np.int64)
    final = np.full(nreg, -(1 << 30), np.int64)
    active = np.ones(nreg, bool)
    seqs = {}
    parts, scores, ncig, cigs, owner, rnd = [], [], [], [], [], []
    wmax = o["w"] << 2
    for rr in range(3):
        idx = np.nonzero(active)[0]
        if len(idx) == 0:
            break
        w2[idx] = np.minimum(w2[idx], wmax)
        pairs, pw, pidx, sc_round = [], [], [], np.zeros(len(idx), np.int64)
        for k, g in enumerate(idx):
            i, qb, qe, rb, re = regs[g][:5]
            if g not in seqs:
                qs = read_fn(i)[qb:qe].copy()
                rs = get_seq(rb, re)
                if len(rs) != re - rb or (rb < l_pac < re):
                    seqs[g] = None
                else:
                    if rb >= l_pac:  # leftmost indels on the reverse strand (bwa.c:216-221)
                        qs, rs = qs[::-1].copy(), rs[::-1].copy()
                    seqs[g] = (qs, rs)
            if seqs[g] is None:
                sc_round[k] = -(1 << 30)
                continue
            qs, rs = seqs[g]
            if len(qs) == len(rs) and w2[g] == 0:  # ungapped (bwa.c:222-230)
                sc_round[k] = int(mat[rs.astype(np.int64) * 5 + qs].sum())
                continue
            pairs.append((qs, rs))
            pw.append(int(cigar_band(w2[g], len(qs), len(rs), mat[0], o["o_del"], o["e_del"], o["o_ins"],
                                     o["e_ins"])))
            pidx.append(k)
        if pairs:
            t = GlbTasks.from_pairs(pairs, pw)
            s_, nc, cg = global_fn(t)
            for j, k in enumerate(pidx):
                sc_round[k] = int(s_[j])
            parts.append(t)
            scores.append(np.asarray(s_))
            ncig.append(np.asarray(nc))
            cigs.extend([list(cg[j, :nc[j]]) if nc[j] > 0 else [] for j in range(len(pairs))])
            owner.append(idx[np.asarray(pidx, np.int64)])
            rnd.append(np.full(len(pairs), rr, np.int32))
        # do-while tail of mem_reg2aln (bwamem.c:1229-1233)
        if trace is not None:
            trace.extend((int(g), rr, int(w2[g]), int(sc_round[k])) for k, g in enumerate(idx))
        final[idx] = sc_round
        stop = (sc_round == last[idx]) | (w2[idx] == wmax)
        cont = ~stop
        last[idx] = np.where(cont, sc_round, last[idx])
        w2[idx] = np.where(cont, w2[idx] << 1, w2[idx])
        tries[idx] += 1
        cont &= (tries[idx] < 3) & (sc_round < truesc[idx] - o["a"])
        active[:] = False
        active[idx[cont]] = True
    batch = concat_glb(parts)
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)
    return (batch, cat(scores, np.int32), cat(ncig, np.int32), cigs, cat(owner, np.int64), cat(rnd, np.int32),
            final)


def md_inputs(wl, regions, owner, cigars):
    """Inputs of the NM / MD:Z / XB:Z pass over the regions of a CIGAR pass
    (build_glb_batch with return_regions): one task per region with its final
    CIGAR -- the last try's, or l_query M for a region scored without DP
    (bwa.c:222-230).  Returns (region index, rb, re, oquery, oq_off, l_query,
    n_cigar, cigar[n, slot]).  The read goes in as mem_reg2aln passes it to
    bwa_gen_cigar2 (bwamem.c:1228): from the read's first base, with
    l_query = qe - qb -- the oquery argument is not offset by qb, unlike the
    converted query."""
    last = {}
    for t, g in enumerate(owner):
        last[int(g)] = t
    idx, cg = [], []
    for g in range(len(regions["rb"])):
        rb, re = int(regions["rb"][g]), int(regions["re"][g])
        lq = int(regions["qe"][g] - regions["qb"][g])
        if g in last:
            idx.append(g); cg.append(np.asarray(cigars[last[g]], np.uint32))
        elif lq == re - rb and not (rb < wl.l_pac < re) and re <= 2 * wl.l_pac:
            idx.append(g); cg.append(np.array([lq << 4], np.uint32))
    idx = np.array(idx, np.int64)
    rb = regions["rb"][idx].astype(np.int64)
    re = regions["re"][idx].astype(np.int64)
    l_query = (regions["qe"] - regions["qb"])[idx].astype(np.int32)
    reads, parts = {}, []
    for t, g in enumerate(idx):
        i = int(regions["read"][g])
        if i not in reads:
            reads[i] = wl.oread(i)
        parts.append(reads[i][:int(l_query[t])])
    n = len(idx)
    oq_off = np.zeros(n, np.uint64)
    if n:
        oq_off[1:] = np.cumsum([len(p) for p in parts])[:-1]
    slot = max(1, max((len(c) for c in cg), default=1))
    cig = np.zeros((n, slot), np.uint32)
    for t, c in enumerate(cg):
        cig[t, :len(c)] = c
    return (idx, rb, re, np.concatenate(parts).astype(np.uint8) if parts else np.zeros(1, np.uint8), oq_off,
            l_query, np.array([len(c) for c in cg], np.int32), cig)


def assess_conversion(seq1, seq2=None, substitution_proportion=0.1):
    """Conversion type of an undirectional read or pair (bs_helpers.cpp:41-62):
    0 = C->T, 1 = G->A, 2 = align both patterns.  Sequences as ASCII strings."""
    s1 = seq1.encode()
    s2 = seq2.encode() if seq2 is not None else s1
    return wl().wl_assess_conversion(s1, s2, 1 if seq2 is not None else 0, substitution_proportion)


def bseq_records(seq1, seq2=None, undirectional=False, substitution_proportion=0.1):
    """The read records bwa mem makes for a read or pair (bwa.c:73-133), as
    (first, read group, substitution pattern) triples: pattern 0 converts
    C->T, 1 converts G->A (bs_helpers.cpp:18-31)."""
    out = np.zeros(12, np.int32)
    n = wl().wl_bseq_records(seq1.encode(), seq2.encode() if seq2 is not None else None, int(undirectional),
                             substitution_proportion, _p(out))
    return [tuple(int(x) for x in out[3 * k:3 * k + 3]) for k in range(n)]


def pack2(codes):
    """2-bit forward-only layout of the .pac files (first base in the top bits),
    len/4+1 bytes as bwa.c:430-431 reads them."""
    codes = np.asarray(codes, np.uint8)
    out = np.zeros(len(codes) // 4 + 1, np.uint8)
    idx = np.arange(len(codes))
    np.bitwise_or.at(out, idx >> 2, ((codes & 3) << ((~idx & 3) << 1)).astype(np.uint8))
    return out


def fill_matrix(a=1, b=4):
    """bwa_fill_scmat (bwa.c:169-178)."""
    m = np.full((5, 5), -b, np.int8)
    np.fill_diagonal(m, a)
    m[4, :] = -1
    m[:, 4] = -1
    return m.reshape(25)


def concat_glb(parts):
    if not parts:
        z = np.zeros(0, np.int32)
        return GlbTasks(np.zeros(0, np.uint8), np.zeros(0, np.uint64), z, np.zeros(0, np.uint8),
                        np.zeros(0, np.uint64), z, z)
    qbase = np.cumsum([0] + [len(p.query) for p in parts[:-1]], dtype=np.uint64)
    tbase = np.cumsum([0] + [len(p.target) for p in parts[:-1]], dtype=np.uint64)
    return GlbTasks(np.concatenate([p.query for p in parts]),
                    np.concatenate([p.q_off + qb for p, qb in zip(parts, qbase)]),
                    np.concatenate([p.qlen for p in parts]),
                    np.concatenate([p.target for p in parts]),
                    np.concatenate([p.t_off + tb for p, tb in zip(parts, tbase)]),
                    np.concatenate([p.tlen for p in parts]),
                    np.concatenate([p.w for p in parts]))


def concat_ext(parts):
    if not parts:
        z = np.zeros(0, np.int32)
        return ExtTasks(np.zeros(0, np.uint8), np.zeros(0, np.uint64), z, np.zeros(0, np.uint8),
                        np.zeros(0, np.uint64), z, z, z, z)
    qbase = np.cumsum([0] + [len(p.query) for p in parts[:-1]], dtype=np.uint64)
    tbase = np.cumsum([0] + [len(p.target) for p in parts[:-1]], dtype=np.uint64)
    return ExtTasks(np.concatenate([p.query for p in parts]),
                    np.concatenate([p.q_off + qb for p, qb in zip(parts, qbase)]),
                    np.concatenate([p.qlen for p in parts]),
                    np.concatenate([p.target for p in parts]),
                    np.concatenate([p.t_off + tb for p, tb in zip(parts, tbase)]),
                    np.concatenate([p.tlen for p in parts]),
                    np.concatenate([p.h0 for p in parts]),
                    np.concatenate([p.w for p in parts]),
                    np.concatenate([p.end_bonus for p in parts]))
