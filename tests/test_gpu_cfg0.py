"""cfg0 on the GPU against the reference aligner itself.

The fixture tests/golden/cfg0_bwamem.npz is what bsbolt/External/BWA, compiled
from its own sources, did on the cfg0 workload with BSBolt's default argv:
every ksw_extend2 call and its six outputs, and for 4,000 reads the chains,
regions and mem_reg2aln CIGAR tries of its own trace.  Here the HIP kernels
run those calls (plain form, learned form, every frame class), the producer's
mem_chain2aln rounds run the reference's chains with the GPU extending, and
the CIGAR tries run on the GPU's ksw_global2 and NM/MD/XB kernels.  Bit-exact."""
import numpy as np
import pytest

import cfg0_cases as K

pytestmark = pytest.mark.gpu


def _scoring(sc):
    from bsbolt_amd.bsx import Scoring
    return Scoring.make(mat=sc["mat"], o_del=sc["o_del"], e_del=sc["e_del"], o_ins=sc["o_ins"],
                        e_ins=sc["e_ins"], zdrop=sc["zdrop"])


@pytest.mark.parametrize("form", ["plain", "learned", "f24", "f40", "f80", "f144"])
def test_gpu_equals_every_reference_call(engine, form):
    from bsbolt_amd import band, bsx
    t, sc, ref = K.ext_tasks()
    print(f"cfg0 reference calls: {len(t)} tasks, form {form}", flush=True)
    if form == "learned":
        engine.set_band_model(band.load_model())
        engine.set_band_policy(bsx.BAND_LEARNED)
    elif form != "plain":
        engine.set_band_policy(bsx.BAND_FIXED, ("f24", "f40", "f80", "f144").index(form))
    try:
        got = engine.extend(t, _scoring(sc))
    finally:
        engine.set_band_policy(bsx.BAND_PLAIN)
    for k in K.FIELDS:
        bad = np.nonzero(got[k] != ref[k])[0]
        assert len(bad) == 0, f"{form}: {k} differs at {len(bad)} calls, first {bad[:8]}"


def test_gpu_producer_replays_reference_chains(engine):
    wl = K.workload()
    try:
        out = K.replay_chains(wl, lambda t: engine.extend(t, cells=False))
        K.check_replay(*out)
    finally:
        wl.close()


def test_gpu_cigar_tries_replay_reference_reg2aln(engine):
    from bsbolt_amd.workload import cigar_tries, pack2
    cases = K.reg2aln_regions()
    wl = K.workload()
    try:
        trace = []
        out = cigar_tries([c["region"] for c in cases], wl.read, wl.get_seq, wl.l_pac,
                          lambda t: engine.global_(t), trace=trace)
        batch, scores, ncig, cigs, owner, rnd, final = out
        assert [(w2, sc) for _, _, w2, sc in trace] == [(c["w2"], c["score"]) for c in cases]
        last = {int(g): t for t, g in enumerate(owner)}
        n = len(cases)
        reg = [c["region"] for c in cases]
        ours = [cigs[last[g]] if g in last else [(reg[g][2] - reg[g][1]) << 4] for g in range(n)]
        for g, c in enumerate(cases):
            assert np.array_equal(np.asarray(ours[g], np.uint32), c["cigar"]), f"region {g}: CIGAR"
        # NM, MD:Z and XB:Z of those CIGARs on the GPU
        oq = [wl.oread(r[0])[:r[2] - r[1]] for r in reg]
        oq_off = np.zeros(n, np.uint64)
        oq_off[1:] = np.cumsum([len(x) for x in oq])[:-1]
        slot = max(len(x) for x in ours)
        cig = np.zeros((n, slot), np.uint32)
        for g, x in enumerate(ours):
            cig[g, :len(x)] = x
        opac = pack2(np.concatenate([wl.genome, wl.genome]))
        got = engine.md(opac, wl.l_pac, wl.cfg.genome_len, np.array([r[3] for r in reg], np.int64),
                        np.array([r[4] for r in reg], np.int64), np.concatenate(oq), oq_off,
                        np.array([len(x) for x in oq], np.int32), np.array([len(x) for x in ours], np.int32), cig)
        for g, c in enumerate(cases):
            assert (int(got["nm"][g]), got["md"][g], got["xb"][g], [int(x) for x in got["meth"][g]]) == \
                (c["nm"], c["md"], c["xb"], c["meth"]), f"region {g}: NM/MD/XB"
    finally:
        wl.close()


@pytest.mark.parametrize("run", ["", "edge_"])
def test_gpu_reg2aln_cigar_equals_reference_sam(engine, run):
    """bsx_reg2aln_cigar_device on the reference's bwa_gen_cigar2 CIGARs gives
    the CIGAR, POS and strand of its SAM records (deletion squeeze, clipping)."""
    cases, recs = K.reg2aln_vs_sam(run)
    lens = K.read_lengths(run)
    n = len(cases)
    slot = max(len(c["cigar"]) for c in cases)
    cig = np.zeros((n, slot), np.uint32)
    for g, c in enumerate(cases):
        cig[g, :len(c["cigar"])] = c["cigar"]
    reg = np.array([c["region"][:5] for c in cases], np.int64)
    out, nout, pos, rev = engine.reg2aln_cigar(4_000_000, reg[:, 3], reg[:, 4], reg[:, 1], reg[:, 2],
                                               lens[reg[:, 0]], [len(c["cigar"]) for c in cases], cig)
    for g, r in enumerate(recs):
        assert (K.cigar_str(out[g, :nout[g]]), int(pos[g]) % 2_000_000 + 1, bool(rev[g])) == \
            (r[3], r[2], bool(r[1] & 16)), (run, g)
    # the squeeze branch on crafted CIGARs, against the oracle
    from oracle import oracle as O
    w = lambda k, op: k << 4 | op
    crafted = [([w(3, 2), w(50, 0)], 100, 153, 0, 50, 50), ([w(50, 0), w(2, 2)], 100, 152, 0, 50, 50),
               ([w(1, 2), w(50, 0), w(2, 2)], 100, 153, 0, 50, 50), ([w(40, 0), w(3, 2)], 1500, 1543, 5, 45, 60),
               ([w(2, 2), w(40, 0)], 1500, 1542, 0, 40, 40)]
    cg = np.zeros((len(crafted), 3), np.uint32)
    for g, x in enumerate(crafted):
        cg[g, :len(x[0])] = x[0]
    a = np.array([x[1:] for x in crafted], np.int64)
    out, nout, pos, rev = engine.reg2aln_cigar(1000, a[:, 0], a[:, 1], a[:, 2], a[:, 3], a[:, 4],
                                               [len(x[0]) for x in crafted], cg)
    for g, x in enumerate(crafted):
        k, words, p, rv = O.reg2aln_cigar(1000, *x[1:], np.array(x[0], np.uint32))
        assert (list(out[g, :nout[g]]), int(pos[g]), int(rev[g])) == (words, p, rv), g
