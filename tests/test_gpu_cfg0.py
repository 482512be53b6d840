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
