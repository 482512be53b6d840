"""Writes tests/golden/cfg0_bwamem.npz: the cfg0 workload run through the
reference aligner itself.

Steps (all in a temporary directory; nothing of the reference is kept):
 1. the cfg0 genome (2 Mb, one contig) and its 20,000 single-end 100 bp
    directional bisulfite reads from the project's simulator
    (bsbolt_amd/workload.py), written as FASTA / FASTQ (reads as sequenced);
 2. `bwa index -a bwtsw` and `bwa mem` with BSBolt Align's default argv
    (bsbolt/Utils/Launcher.py:77-108, defaults bsbolt/Utils/Parser.py:31-115),
    one thread, verbosity 4, using oracle/_ref/bwa: the reference's own
    sources compiled by oracle/Makefile, with the call recorder oracle/bwa_dump.c
    linked around ksw_extend2 / ksw_global2 / bwa_gen_cigar2;
 3. the recorder's calls and the aligner's own verbosity-4 trace (chains,
    extension lines and regions of mem_chain2aln, bwamem.c:1171,736,764,776;
    mem_reg2aln's band and tries, bwamem.c:1222,1229) are split per read and
    stored as arrays.

The fixture holds data only: every ksw_extend2 call of every read (inputs and
the six outputs); for the first N_TRACE reads their chains and regions; for
every read its final regions, its CIGAR tries (bwa_gen_cigar2 inputs and
outputs, including the NM, MD:Z and XB:Z text) and its SAM records (flag, POS,
CIGAR); the same for a second run on N_EDGE edited copies of the reads (random
ends, deletions near an end), whose regions mem_reg2aln clips and squeezes;
and SHA-256 digests of the index's .pac / .opac.

Run:  python tests/golden/make_golden_cfg0.py   (needs oracle/_ref/bwa: `make -C oracle bwa`)
"""
import hashlib
import os
import re
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
BWA = os.path.join(ROOT, "oracle", "_ref", "bwa")
OUT = os.path.join(HERE, "cfg0_bwamem.npz")
N_TRACE = 4000   # reads whose chains / regions are kept
N_EDGE = 2000    # edited reads of the second run (clipping and deletion squeeze)
# BSBolt Align's bwa mem argv at its defaults (Launcher.py:77-108; Parser.py:31-115), one thread
BSB_ARGV = ["-Y", "-A", "1", "-B", "4", "-D", "0.5", "-E", "1,1", "-L", "30,30", "-T", "10", "-U", "17",
            "-W", "0", "-c", "500", "-d", "100", "-k", "19", "-m", "50", "-r", "1.5", "-t", "1", "-w", "100",
            "-y", "20", "-O", "6,6", "-h", "100,200", "-e", "0.1", "-l", "0.5", "-n", "5", "-Z", "0.95"]
ACGTN = np.frombuffer(b"ACGTN", np.uint8)


def write_inputs(d, wl):
    with open(os.path.join(d, "ref.fa"), "w") as f:
        f.write(">chr1\n")
        s = ACGTN[wl.genome].tobytes().decode()
        for i in range(0, len(s), 80):
            f.write(s[i:i + 80] + "\n")
    with open(os.path.join(d, "reads.fq"), "w") as f:
        for i in range(wl.n_reads):
            r = ACGTN[wl.oread(i)].tobytes().decode()
            f.write(f"@r{i}\n{r}\n+\n{'I' * len(r)}\n")


def parse_dump(path):
    """Records of the call recorder (layout in oracle/bwa_dump.c)."""
    b = open(path, "rb").read()
    p, E, G, Cg = 0, [], [], []
    i32 = lambda k: struct.unpack_from(f"<{k}i", b, p)
    while p < len(b):
        tag = b[p:p + 1]
        p += 1
        if tag == b"E":
            qlen, tlen, m, od, ed, oi, ei, w, eb, zd, h0 = i32(11); p += 44
            mat = np.frombuffer(b, np.int8, m * m, p); p += m * m
            q = np.frombuffer(b, np.uint8, qlen, p); p += qlen
            t = np.frombuffer(b, np.uint8, tlen, p); p += tlen
            out = i32(6); p += 24
            E.append(dict(qlen=qlen, tlen=tlen, gap=(od, ed, oi, ei), w=w, end_bonus=eb, zdrop=zd, h0=h0,
                          mat=mat.copy(), q=q.copy(), t=t.copy(), out=out))
        elif tag == b"G":
            qlen, tlen, m, od, ed, oi, ei, w = i32(8); p += 32
            p += m * m
            q = np.frombuffer(b, np.uint8, qlen, p); p += qlen
            t = np.frombuffer(b, np.uint8, tlen, p); p += tlen
            score, nc = i32(2); p += 8
            cig = np.frombuffer(b, np.uint32, max(nc, 0), p).copy(); p += 4 * max(nc, 0)
            G.append(dict(qlen=qlen, tlen=tlen, w=w, q=q.copy(), t=t.copy(), score=score, n_cigar=nc, cigar=cig))
        elif tag == b"C":
            w_, lq = i32(2); p += 8
            l_pac, rb, re_, crick_l = struct.unpack_from("<4q", b, p); p += 32
            want, = i32(1); p += 4
            q = np.frombuffer(b, np.uint8, max(lq, 0), p).copy(); p += max(lq, 0)
            score, nc, nm = i32(3); p += 12
            cig = np.frombuffer(b, np.uint32, nc, p).copy(); p += 4 * nc
            tl, = i32(1); p += 4
            tail = b[p:p + tl].decode(); p += tl
            meth = i32(4); p += 16
            Cg.append(dict(w_=w_, l_query=lq, l_pac=l_pac, rb=rb, re=re_, crick_l=crick_l, want=want, q=q,
                           score=score, n_cigar=nc, nm=nm, cigar=cig, tail=tail, meth=meth, n_g=len(G)))
        else:
            raise ValueError(f"bad record tag {tag!r} at {p - 1}")
    return E, G, Cg


RE_READ = re.compile(r"^=====> Processing read 'r(\d+)' <=====")
RE_FIN = re.compile(r"^=====> Finalizing read 'r(\d+)' <=====")
RE_CHAIN = re.compile(r"^\* Found CHAIN\((\d+)\): n=(\d+); weight=(\d+)\t(.*)$")
RE_SEED = re.compile(r"^(-?\d+);(\d+);(\d+),(\d+)\(")
RE_EXT = re.compile(r"^\*\*\* (Left|Right) extension: prev_score=(-?\d+); score=(-?\d+); bandwidth=(\d+); "
                    r"max_off_diagonal_dist=(-?\d+)")
RE_REG = re.compile(r"^\*\*\* Added alignment region: \[(\d+),(\d+)\) <=> \[(\d+),(\d+)\); score=(-?\d+); "
                    r"\{left,right\}_bandwidth=\{(\d+),(\d+)\}")
RE_FINAL_REG = re.compile(r"^\*\* (-?\d+), \[(\d+),(\d+)\) <=> \[(\d+),(\d+)\)$")
RE_BAND = re.compile(r"^\* Band width: inferred=(-?\d+), cmd_opt=(\d+), alnreg=(\d+)")
RE_TRY = re.compile(r"^\* Final alignment: w2=(-?\d+), global_sc=(-?\d+), local_sc=(-?\d+)")


def parse_trace(path, n_reads):
    """Per-read chains, extension lines, regions and mem_reg2aln tries, in order."""
    chains = [[] for _ in range(n_reads)]
    ext = [[] for _ in range(n_reads)]
    regs = [[] for _ in range(n_reads)]
    final = [[] for _ in range(n_reads)]
    reg2aln = [[] for _ in range(n_reads)]   # per mem_reg2aln call: (inferred, cmd, alnreg, [tries])
    sam = []                                 # (read, flag, rname, pos, cigar) of every SAM record
    cur, fin = -1, -1
    with open(path, errors="replace") as f:
        for line in f:
            line = line.rstrip("\n")
            if line.startswith("@"):
                continue  # SAM header
            if line and line[0] == "r" and "\t" in line:  # SAM record
                f_ = line.split("\t")
                sam.append((int(f_[0][1:]), int(f_[1]), f_[2], int(f_[3]), f_[5]))
                continue
            m = RE_READ.match(line)
            if m:
                cur, fin = int(m.group(1)), -1
                continue
            m = RE_FIN.match(line)
            if m:
                fin, cur = int(m.group(1)), -1
                continue
            if cur >= 0:
                m = RE_CHAIN.match(line)
                if m:
                    seeds = []
                    for tok in m.group(4).split("\t"):
                        s = RE_SEED.match(tok)
                        seeds.append(tuple(int(x) for x in s.groups()))  # score, len, qbeg, rbeg
                    assert len(seeds) == int(m.group(2))
                    chains[cur].append(seeds)
                    continue
                m = RE_EXT.match(line)
                if m:
                    ext[cur].append((0 if m.group(1) == "Left" else 1,) + tuple(int(x) for x in m.groups()[1:]))
                    continue
                m = RE_REG.match(line)
                if m:
                    regs[cur].append(tuple(int(x) for x in m.groups()))
                    continue
                m = RE_FINAL_REG.match(line)
                if m:
                    final[cur].append(tuple(int(x) for x in m.groups()))
                    continue
            if fin >= 0:
                m = RE_BAND.match(line)
                if m:
                    reg2aln[fin].append([tuple(int(x) for x in m.groups()), []])
                    continue
                m = RE_TRY.match(line)
                if m:
                    reg2aln[fin][-1][1].append(tuple(int(x) for x in m.groups()))
    return chains, ext, regs, final, reg2aln, sam


def edge_reads(wl, n):
    """n reads as sequenced, edited so that mem_reg2aln clips and squeezes:
    random 5' bases, random 3' bases, both, or a 1-3 base deletion 2-6 bases
    from one end (seeded; built from the first n cfg0 reads)."""
    rng = np.random.default_rng(20240)
    out = []
    for k in range(n):
        r = wl.oread(k).copy()
        kind = k % 4
        if kind in (0, 2):
            m = int(rng.integers(12, 36)); r[:m] = rng.integers(0, 4, m)
        if kind in (1, 2):
            m = int(rng.integers(12, 36)); r[len(r) - m:] = rng.integers(0, 4, m)
        if kind == 3:
            d, p = int(rng.integers(1, 4)), int(rng.integers(2, 7))
            p = p if rng.random() < 0.5 else len(r) - p - d
            r = np.concatenate([r[:p], r[p + d:]])
        out.append(r.astype(np.uint8))
    return out


def tries_and_sam(Cg, reg2aln, final, sam, n_reads):
    """mem_reg2aln's tries matched to the recorder's bwa_gen_cigar2 calls, the
    final regions and the SAM records, as fixture arrays."""
    want = [c for c in Cg if c["want"]]
    calls = [(i, blk) for i in range(n_reads) for blk in reg2aln[i]]
    assert sum(len(blk[1]) for _, blk in calls) == len(want), (sum(len(b[1]) for _, b in calls), len(want))
    rows, cg_words, tails, k = [], [], [], 0
    for ci, (i, (band, tries)) in enumerate(calls):
        for ti, (w2, gsc, lsc) in enumerate(tries):
            c = want[k]; k += 1
            assert c["w_"] == w2 and c["score"] == gsc, (i, ti)
            rows.append((i, ci, ti, band[0], band[1], band[2], w2, gsc, lsc, c["l_query"], c["rb"], c["re"],
                         c["n_cigar"], c["nm"], *c["meth"]))
            cg_words.append(c["cigar"])
            tails.append(c["tail"])
    F = [(i,) + r for i in range(n_reads) for r in final[i]]
    return {
        "final_region": np.array(F, np.int64).reshape(-1, 6),  # read, score, qb, qe, rb, re
        "reg2aln": np.array(rows, np.int64).reshape(-1, 18),
        "reg2aln_cigar": np.concatenate(cg_words).astype(np.uint32) if cg_words else np.zeros(0, np.uint32),
        "reg2aln_tail": np.frombuffer("\n".join(tails).encode(), np.uint8).copy(),
        # SAM records (mem_reg2aln's squeeze and clipping, bwamem.c:1237-1263, end up here)
        "sam": np.array([(r, fl, p) for r, fl, _, p, _ in sam], np.int64).reshape(-1, 3),  # read, flag, pos
        "sam_cigar": np.frombuffer("\n".join(c for *_, c in sam).encode(), np.uint8).copy(),
    }


def main():
    from bsbolt_amd.workload import CONFIGS, Workload, bs_pac, pack2
    if not os.path.exists(BWA):
        raise SystemExit(f"{BWA} not built: run `make -C oracle bwa` (needs /root/reference)")
    cfg = CONFIGS["cfg0"]
    wl = Workload(cfg)
    n_reads = wl.n_reads
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, wl)
        fa = os.path.join(d, "ref.fa")
        subprocess.run([BWA, "index", "-a", "bwtsw", fa], check=True, capture_output=True)
        pac = open(fa + ".pac", "rb").read()
        opac = open(fa + ".opac", "rb").read()
        l_pac = 2 * cfg.genome_len
        nb = (l_pac + 3) // 4   # data bytes (the file then holds the l_pac % 4 trailer)
        ours_pac = bs_pac(wl.genome)[:nb].tobytes()
        ours_opac = pack2(np.concatenate([wl.genome, wl.genome]))[:nb].tobytes()
        print(f"index: .pac {'==' if pac[:nb] == ours_pac else '!='} wl_bs_pac, "
              f".opac {'==' if opac[:nb] == ours_opac else '!='} pack2(genome x2)", flush=True)
        env = dict(os.environ, BSX_DUMP=os.path.join(d, "dump.bin"))
        with open(os.path.join(d, "trace.txt"), "w") as out:
            subprocess.run([BWA, "mem", "-v", "4", *BSB_ARGV, fa, os.path.join(d, "reads.fq")], check=True,
                           stdout=out, stderr=subprocess.DEVNULL, env=env)
        E, G, Cg = parse_dump(os.path.join(d, "dump.bin"))
        chains, ext, regs, final, reg2aln, sam = parse_trace(os.path.join(d, "trace.txt"), n_reads)
        # second run: edited reads whose regions get clipped / squeezed by mem_reg2aln
        er = edge_reads(wl, N_EDGE)
        with open(os.path.join(d, "edge.fq"), "w") as f:
            for i, r in enumerate(er):
                f.write(f"@r{i}\n{ACGTN[r].tobytes().decode()}\n+\n{'I' * len(r)}\n")
        env["BSX_DUMP"] = os.path.join(d, "edge.bin")
        with open(os.path.join(d, "edge.txt"), "w") as out:
            subprocess.run([BWA, "mem", "-v", "4", *BSB_ARGV, fa, os.path.join(d, "edge.fq")], check=True,
                           stdout=out, stderr=subprocess.DEVNULL, env=env)
        _, _, eCg = parse_dump(os.path.join(d, "edge.bin"))
        _, _, _, efinal, ereg2aln, esam = parse_trace(os.path.join(d, "edge.txt"), N_EDGE)
    # ---- extension calls, split per read by the trace's extension lines ----
    per_read = [len(x) for x in ext]
    assert sum(per_read) == len(E), (sum(per_read), len(E))
    e_read = np.repeat(np.arange(n_reads, dtype=np.int32), per_read)
    for k, (e, rd) in enumerate(zip(E, e_read)):
        side, prev, score, bw, mo = ext[rd][k - int(np.searchsorted(e_read, rd))]
        assert e["out"][0] == score and e["w"] == bw and e["out"][5] == mo, k
    mat = E[0]["mat"]
    assert all(np.array_equal(e["mat"], mat) and e["gap"] == E[0]["gap"] and e["zdrop"] == E[0]["zdrop"]
               for e in E)
    ql = np.array([e["qlen"] for e in E], np.int32)
    tl = np.array([e["tlen"] for e in E], np.int32)
    q_off = np.zeros(len(E), np.uint64); q_off[1:] = np.cumsum(ql[:-1])
    t_off = np.zeros(len(E), np.uint64); t_off[1:] = np.cumsum(tl[:-1])
    fx = dict(
        query=np.concatenate([e["q"] for e in E]), q_off=q_off, qlen=ql,
        target=np.concatenate([e["t"] for e in E]), t_off=t_off, tlen=tl,
        h0=np.array([e["h0"] for e in E], np.int32), w=np.array([e["w"] for e in E], np.int32),
        end_bonus=np.array([e["end_bonus"] for e in E], np.int32),
        mat=mat, o_del=np.int32(E[0]["gap"][0]), e_del=np.int32(E[0]["gap"][1]), o_ins=np.int32(E[0]["gap"][2]),
        e_ins=np.int32(E[0]["gap"][3]), zdrop=np.int32(E[0]["zdrop"]),
        e_read=e_read, e_side=np.array([x[0] for r in ext for x in r], np.int8),
    )
    for j, k in enumerate(("score", "qle", "tle", "gtle", "gscore", "max_off")):
        fx["ref_" + k] = np.array([e["out"][j] for e in E], np.int32)
    # ---- producer trace of the first N_TRACE reads ----
    nt = min(N_TRACE, n_reads)
    sr, sc, sk, sl, sq, sb = [], [], [], [], [], []
    for i in range(nt):
        for c, seeds in enumerate(chains[i]):
            for (score, ln, qb, rb) in seeds:
                sr.append(i); sc.append(c); sk.append(score); sl.append(ln); sq.append(qb); sb.append(rb)
    fx.update(n_trace=np.int32(nt), seed_read=np.array(sr, np.int32), seed_chain=np.array(sc, np.int32),
              seed_score=np.array(sk, np.int32), seed_len=np.array(sl, np.int32),
              seed_qbeg=np.array(sq, np.int32), seed_rbeg=np.array(sb, np.int64))
    R = [(i,) + r for i in range(nt) for r in regs[i]]
    fx["region"] = np.array(R, np.int64).reshape(-1, 8)        # read, qb, qe, rb, re, score, lw, rw
    fx.update(tries_and_sam(Cg, reg2aln, final, sam, n_reads))
    for k, v in tries_and_sam(eCg, ereg2aln, efinal, esam, N_EDGE).items():
        fx["edge_" + k] = v
    fx["edge_reads"] = np.concatenate(er)
    fx["edge_len"] = np.array([len(r) for r in er], np.int32)
    fx["n_reads"] = np.int32(n_reads)
    fx["pac_sha256"] = np.frombuffer(hashlib.sha256(pac[:nb]).digest(), np.uint8).copy()
    fx["opac_sha256"] = np.frombuffer(hashlib.sha256(opac[:nb]).digest(), np.uint8).copy()
    fx["oread_sha256"] = np.frombuffer(hashlib.sha256(np.concatenate(
        [wl.oread(i) for i in range(n_reads)]).tobytes()).digest(), np.uint8).copy()
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT}: {len(E)} ksw_extend2 calls of {n_reads} reads; {len(fx['seed_read'])} seeds, "
          f"{len(R)} regions of the first {nt} reads; {len(fx['reg2aln'])} CIGAR tries, {len(sam)} SAM records; "
          f"edited-read run: {len(fx['edge_reg2aln'])} tries, {len(fx['edge_sam'])} SAM records; "
          f"{len(G)} ksw_global2 calls; "
          f"{os.path.getsize(OUT) / 1e6:.2f} MB")
    wl.close()


if __name__ == "__main__":
    main()
