"""The CIGAR pass of the RRBS (MspI fragments, 40-100 bp) and indel-heavy
250 bp workloads on the GPU: extension rounds, then mem_reg2aln's band
inference and retries with ksw_global2 on the GPU.  Every global task's score
and CIGAR must equal the oracle's, and the whole pipeline must end in the same
per-region scores as the same pipeline run on the CPU oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cpu_global(t):
    slot = int((t.qlen + t.tlen).max())
    s, nc, cg, _ = O.global2_batch(t, slot)
    return s, nc, cg


@pytest.mark.parametrize("cfg,scale", [("cfg4", 0.02), ("cfg3", 0.01)])
def test_cigar_pass_matches_oracle(engine, cfg, scale):
    from bsbolt_amd.workload import CONFIGS, Workload, build_glb_batch
    w = Workload(CONFIGS[cfg].scaled(scale))
    try:
        w.build_ext_batch(lambda t: engine.extend(t, cells=False))
        batch, s, nc, cigs, owner, rnd, final = build_glb_batch(w, lambda t: engine.global_(t), min_score=10)
    finally:
        w.close()
    assert len(batch) > 200
    slot = int((batch.qlen + batch.tlen).max())
    ws, wn, wc, _ = O.global2_batch(batch, slot)
    assert np.array_equal(s, ws)
    assert np.array_equal(nc, wn)
    for k in range(len(batch)):
        assert cigs[k] == list(wc[k, :wn[k]]), f"task {k}: CIGAR differs"
    # the same pipeline on the CPU oracle reaches the same regions and scores
    w2 = Workload(CONFIGS[cfg].scaled(scale))
    try:
        w2.build_ext_batch(lambda t: O.extend2_batch(t, nthreads=16)[0])
        b2, s2, nc2, cigs2, owner2, rnd2, final2 = build_glb_batch(w2, _cpu_global, min_score=10)
    finally:
        w2.close()
    assert np.array_equal(final, final2)
    assert np.array_equal(owner, owner2) and np.array_equal(rnd, rnd2)
