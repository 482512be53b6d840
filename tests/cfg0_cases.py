"""The cfg0 batch as the reference aligner issued it (tests/golden/cfg0_bwamem.npz,
written by tests/golden/make_golden_cfg0.py from bsbolt/External/BWA compiled
here), and replays of its chains and regions through this project's task
producer, shared by the CPU and GPU tests."""
import numpy as np

from conftest import load_golden

FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")
_FX = None


def fixture():
    global _FX
    if _FX is None:
        _FX = load_golden("cfg0_bwamem.npz")
    return _FX


def ext_tasks():
    """Every ksw_extend2 call of bwa mem on cfg0: (ExtTasks, scoring dict, reference outputs)."""
    from bsbolt_amd.tasks import ExtTasks, EXT_FIELDS
    z = fixture()
    t = ExtTasks(*[z[k] for k in EXT_FIELDS])
    sc = dict(mat=z["mat"], o_del=int(z["o_del"]), e_del=int(z["e_del"]), o_ins=int(z["o_ins"]),
              e_ins=int(z["e_ins"]), zdrop=int(z["zdrop"]))
    return t, sc, {k: z["ref_" + k] for k in FIELDS}


def workload():
    from bsbolt_amd.workload import CONFIGS, Workload
    return Workload(CONFIGS["cfg0"])


def replay_chains(wl, extend):
    """The reference's chains of the traced reads through the producer's
    mem_chain2aln rounds.  Returns (tasks in the reference's call order, their
    results, read of each task, regions per read as (qb, qe, rb, re, score, w))."""
    z = fixture()
    nt = int(z["n_trace"])
    batch, res, side, rnd = wl.build_ext_batch(
        extend, seeds=(z["seed_read"], z["seed_chain"], z["seed_rbeg"], z["seed_qbeg"], z["seed_len"]))
    order = np.lexsort((rnd, wl.task_read))  # per read, round order = the reference's sequential order
    regions = {i: [tuple(int(x) for x in r[[0, 1, 2, 3, 4, 6]]) for r in wl.regions(i)] for i in range(nt)}
    return batch, res, wl.task_read[order], order, regions


def reference_calls_of_traced_reads():
    """Indices (into the fixture's call list) of the calls of the traced reads."""
    z = fixture()
    return np.nonzero(z["e_read"] < int(z["n_trace"]))[0]


def reg2aln_regions(run=""):
    """Each mem_reg2aln call of a reference run ("" = the cfg0 reads, "edge_" =
    the edited reads) as a region (read, qb, qe, rb, re, truesc, w) -- qb/qe
    from the read's final region with that window -- with the reference's try:
    w2, global score, CIGAR, NM and MD/XB text."""
    z = fixture()
    R = z[run + "reg2aln"]
    fin = z[run + "final_region"]
    tails = bytes(z[run + "reg2aln_tail"]).decode().split("\n")
    cig_off = np.concatenate([[0], np.cumsum(R[:, 12])])
    out = []
    for k, row in enumerate(R):
        i, ci, ti, inf, cmd, alnreg, w2, gsc, lsc, lq, rb, re, nc, nm = (int(x) for x in row[:14])
        m = fin[(fin[:, 0] == i) & (fin[:, 4] == rb) & (fin[:, 5] == re) & (fin[:, 3] - fin[:, 2] == lq)]
        assert len(m) >= 1, f"no final region of read {i} for window [{rb},{re})"
        qb, qe = int(m[0, 2]), int(m[0, 3])
        md, _, xb = tails[k].partition("\tXB:Z:")
        out.append(dict(region=(i, qb, qe, rb, re, lsc, alnreg), w2=w2, score=gsc, n_cigar=nc, nm=nm,
                        cigar=z[run + "reg2aln_cigar"][cig_off[k]:cig_off[k] + nc], md=md, xb=xb,
                        meth=[int(x) for x in row[14:18]], inferred=inf))
    return out


def check_replay(batch, res, task_read, order, regions):
    """The replayed producer issues exactly the reference's calls, and ends
    with exactly its regions."""
    z = fixture()
    idx = reference_calls_of_traced_reads()
    assert len(order) == len(idx), f"producer issued {len(order)} calls, the reference {len(idx)}"
    assert np.array_equal(task_read, z["e_read"][idx])
    for k in ("qlen", "tlen", "h0", "w", "end_bonus"):
        got = getattr(batch, k)[order]
        assert np.array_equal(got, z[k][idx]), f"task field {k} differs from the reference's calls"
    for j, (a, b) in enumerate(zip(order, idx)):
        qo, to = int(z["q_off"][b]), int(z["t_off"][b])
        assert np.array_equal(batch.query_of(a), z["query"][qo:qo + int(z["qlen"][b])]), f"call {j}: query"
        assert np.array_equal(batch.target_of(a), z["target"][to:to + int(z["tlen"][b])]), f"call {j}: target"
    for k in FIELDS:
        assert np.array_equal(np.asarray(res[k])[order], z["ref_" + k][idx]), f"result {k}"
    reg = z["region"]
    for i in range(int(z["n_trace"])):
        want = [(int(r[1]), int(r[2]), int(r[3]), int(r[4]), int(r[5]), max(int(r[6]), int(r[7])))
                for r in reg[reg[:, 0] == i]]
        assert regions[i] == want, f"read {i}: regions {regions[i]} != reference {want}"


# BSBolt Align's bwa mem argv at its defaults (bsbolt/Utils/Launcher.py:77-108;
# bsbolt/Utils/Parser.py:31-115), one thread
BSB_ARGV = ["-Y", "-A", "1", "-B", "4", "-D", "0.5", "-E", "1,1", "-L", "30,30", "-T", "10", "-U", "17",
            "-W", "0", "-c", "500", "-d", "100", "-k", "19", "-m", "50", "-r", "1.5", "-t", "1", "-w", "100",
            "-y", "20", "-O", "6,6", "-h", "100,200", "-e", "0.1", "-l", "0.5", "-n", "5", "-Z", "0.95"]


def write_fasta_fastq(d, wl, n_reads=None):
    """The workload's genome as a one-contig FASTA and its reads as sequenced
    as FASTQ (names r0, r1, ...); returns (fasta path, fastq path)."""
    import os
    acgtn = np.frombuffer(b"ACGTN", np.uint8)
    fa, fq = os.path.join(d, "ref.fa"), os.path.join(d, "reads.fq")
    with open(fa, "w") as f:
        f.write(">chr1\n")
        s = acgtn[wl.genome].tobytes().decode()
        for i in range(0, len(s), 80):
            f.write(s[i:i + 80] + "\n")
    with open(fq, "w") as f:
        for i in range(wl.n_reads if n_reads is None else min(n_reads, wl.n_reads)):
            r = acgtn[wl.oread(i)].tobytes().decode()
            f.write(f"@r{i}\n{r}\n+\n{'I' * len(r)}\n")
    return fa, fq


def sam_records(run=""):
    """A reference run's SAM records: list of (read, flag, pos, CIGAR string)."""
    z = fixture()
    cig = bytes(z[run + "sam_cigar"]).decode().split("\n")
    return [(int(r), int(f), int(p), c) for (r, f, p), c in zip(z[run + "sam"], cig)]


def read_lengths(run=""):
    z = fixture()
    return np.full(int(z["n_reads"]), 100, np.int32) if run == "" else z["edge_len"]


def reg2aln_vs_sam(run=""):
    """The mem_reg2aln calls of a run paired with the SAM records they became:
    for every read with as many calls as records (XA hits add calls without
    records), in order.  Returns (cases, records)."""
    cases = reg2aln_regions(run)
    recs = sam_records(run)
    by_c, by_r = {}, {}
    for c in cases:
        by_c.setdefault(c["region"][0], []).append(c)
    for r in recs:
        if not r[1] & 4:
            by_r.setdefault(r[0], []).append(r)
    out_c, out_r = [], []
    for i, cs in by_c.items():
        rs = by_r.get(i, [])
        if len(cs) == len(rs):
            out_c += cs
            out_r += rs
    return out_c, out_r


def cigar_str(words, clip="S"):
    """BAM CIGAR words as SAM text, clip ops (3) printed as `clip`."""
    return "".join(f"{w >> 4}{'MID' [w & 0xf] if (w & 0xf) < 3 else clip}" for w in words)
