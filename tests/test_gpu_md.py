"""GPU NM / MD:Z / XB:Z kernel (bsx_md_batch_device; bwa_gen_cigar2's tail,
bsbolt/External/BWA/bwa.c:250-343) against the reference's outputs in the
committed fixture and against the oracle (orc_gen_md) on generated cases."""
import os

import numpy as np
import pytest

import md_cases as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "md_bsbolt.npz")


@pytest.fixture(scope="module")
def eng():
    from bsbolt_amd import bsx
    e = bsx.Engine(0)
    yield e
    e.close()


def _pack_cigars(cgs):
    slot = max(1, max(len(c) for c in cgs))
    out = np.zeros((len(cgs), slot), np.uint32)
    for i, c in enumerate(cgs):
        out[i, :len(c)] = c
    return out, np.array([len(c) for c in cgs], np.int32)


def _run(eng, opac, l_pac, crick_l, rbs, res, oqs, cgs, **kw):
    cig, nc = _pack_cigars(cgs)
    oq_off = np.concatenate([[0], np.cumsum([len(x) for x in oqs])[:-1]]).astype(np.uint64)
    return eng.md(opac, l_pac, crick_l, np.asarray(rbs, np.int64), np.asarray(res, np.int64),
                  np.concatenate(oqs).astype(np.uint8), oq_off, np.array([len(x) for x in oqs], np.int32),
                  nc, cig, **kw)


def test_md_kernel_matches_reference_fixture(eng):
    d = np.load(FIX)
    g = d["genome"]
    opac = M.pack2(np.concatenate([g, g]))
    n = len(d["rb"])
    sl = lambda k, o, i: d[k][d[o][i]:d[o][i + 1]]
    oqs = [sl("oquery", "oq_off", i) for i in range(n)]
    cgs = [sl("cigar", "cg_off", i) for i in range(n)]
    print(f"md fixture: {n} cases", flush=True)
    out = _run(eng, opac, int(d["l_pac"]), int(d["crick_l"]), d["rb"], d["re"], oqs, cgs)
    for i in range(n):
        if not d["has_md"][i]:
            assert out["status"][i] == 1 and out["nm"][i] == -1, i
            continue
        assert out["status"][i] == 0, i
        assert out["md"][i] == bytes(sl("md", "md_off", i)).decode(), i
        assert out["xb"][i] == bytes(sl("xb", "xb_off", i)).decode(), i
        assert out["nm"][i] == d["nm"][i], i
        assert list(out["meth"][i]) == list(d["meth"][i]), i


def test_md_kernel_matches_oracle_generated(eng):
    g, pac, opac, l_pac, crick_l = M.index(200000, 31)
    rng = np.random.default_rng(17)
    mat = O.fill_scmat()
    rbs, res, oqs, cgs = [], [], [], []
    for rb, re in M.regions(rng, 6000, 200000, l_pac):
        q, oq = M.reads_for(rng, l_pac, opac, crick_l, rb, re, p_unmeth=float(rng.random()),
                            p_sub=0.04, p_indel=0.02)
        cg = M.cigar_stage(mat, l_pac, pac, q, rb, re)
        if cg is None:
            continue
        rbs.append(rb); res.append(re); oqs.append(oq); cgs.append(cg)
    print(f"md generated: {len(rbs)} cases", flush=True)
    out = _run(eng, opac, l_pac, crick_l, rbs, res, oqs, cgs)
    st_seen = set()
    for i in range(len(rbs)):
        st, md, xb, nm, meth = O.gen_md(l_pac, opac, rbs[i], res[i], crick_l, oqs[i], cgs[i])
        st_seen.add(st)
        assert out["status"][i] == st, i
        assert (out["md"][i], out["xb"][i], int(out["nm"][i]), list(out["meth"][i])) == (md, xb, nm, meth), i
    assert 0 in st_seen and 1 in st_seen


def test_md_kernel_slot_overflow(eng):
    """A slot too small for its string: status -1, empty strings, NM -1; the
    other tasks of the batch are computed."""
    g, pac, opac, l_pac, crick_l = M.index(5000, 4)
    rng = np.random.default_rng(2)
    mat = O.fill_scmat()
    rbs, res, oqs, cgs = [], [], [], []
    for rb in (200, 900, 3000):
        q, oq = M.reads_for(rng, l_pac, opac, crick_l, rb, rb + 120, p_sub=0.08)
        rbs.append(rb); res.append(rb + 120); oqs.append(oq); cgs.append(M.cigar_stage(mat, l_pac, pac, q, rb, rb + 120))
    full = _run(eng, opac, l_pac, crick_l, rbs, res, oqs, cgs)
    cap = max(len(m) for m in full["md"]) - 1
    small = _run(eng, opac, l_pac, crick_l, rbs, res, oqs, cgs, md_slot=cap)
    for i in range(3):
        if len(full["md"][i]) > cap:
            assert small["status"][i] == -1 and small["nm"][i] == -1 and small["md"][i] == ""
        else:
            assert small["status"][i] == 0 and small["md"][i] == full["md"][i]
    assert -1 in set(small["status"].tolist())


def test_md_kernel_on_workload_cigar_pass(eng):
    """The CIGAR pass of a simulated RRBS workload (ksw_global2 on the GPU with
    mem_reg2aln's band inference and retries), then NM/MD/XB of every region's
    final CIGAR on the GPU against the oracle."""
    from bsbolt_amd.workload import CONFIGS, Workload, build_glb_batch, md_inputs
    cfg = CONFIGS["cfg4"].scaled(0.004)
    wl = Workload(cfg)
    try:
        wl.build_ext_batch(lambda t: eng.extend(t, cells=False))
        out = build_glb_batch(wl, lambda t: eng.global_(t), return_regions=True)
        reg, owner, cigs = out[7], out[4], out[3]
        idx, rb, re, oq, oq_off, lq, nc, cig = md_inputs(wl, reg, owner, cigs)
        opac = wl.opac()
        print(f"md workload: {len(idx)} regions", flush=True)
        got = eng.md(opac, wl.l_pac, cfg.genome_len, rb, re, oq, oq_off, lq, nc, cig)
        for i in range(len(idx)):
            o = int(oq_off[i])
            st, md, xb, nm, meth = O.gen_md(wl.l_pac, opac, int(rb[i]), int(re[i]), cfg.genome_len,
                                            oq[o:o + int(lq[i])], cig[i, :nc[i]])
            assert got["status"][i] == st, i
            assert (got["md"][i], got["xb"][i], int(got["nm"][i]), list(got["meth"][i])) == (md, xb, nm, meth), i
        assert len(idx) > 1000
    finally:
        wl.close()


def test_md_kernel_rejects_like_the_oracle(eng):
    """Inconsistent tasks (CIGAR longer than the read, a soft clip, an empty,
    reversed, negative or out-of-range region, more ops than the slot holds)
    get the oracle's status -3 and no strings; a read with N bases and the valid
    task beside them are computed."""
    g, pac, opac, l_pac, crick_l = M.index(4000, 3)
    oq = O.get_seq(l_pac, opac, 100, 150).astype(np.uint8)
    oqn = oq.copy()
    oqn[[3, 17, 40]] = 4
    ok = np.array([50 << 4], np.uint32)
    cases = [(100, 150, oq, ok), (100, 150, oqn, ok), (100, 150, oq, np.array([51 << 4], np.uint32)),
             (100, 150, oq, np.array([(2 << 4) | 4, 48 << 4], np.uint32)), (150, 150, oq, ok), (160, 150, oq, ok),
             (-1, 49, oq, ok), (2 * l_pac - 20, 2 * l_pac + 30, oq, ok)]
    out = _run(eng, opac, l_pac, crick_l, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases],
               [c[3] for c in cases])
    for i, (rb, re, q, cg) in enumerate(cases):
        st, md, xb, nm, meth = O.gen_md(l_pac, opac, rb, re, crick_l, q, cg)
        assert out["status"][i] == st, i
        assert (out["md"][i], out["xb"][i], int(out["nm"][i]), list(out["meth"][i])) == (md, xb, nm, meth), i
    assert list(out["status"]) == [0, 0, -3, -3, -3, -3, -3, -3]
    # more CIGAR ops than the slot holds
    cig = np.zeros((1, 1), np.uint32)
    cig[0, 0] = 50 << 4
    r = eng.md(opac, l_pac, crick_l, np.array([100], np.int64), np.array([150], np.int64), oq,
               np.zeros(1, np.uint64), np.array([50], np.int32), np.array([2], np.int32), cig)
    assert r["status"][0] == -3 and r["nm"][0] == -1
