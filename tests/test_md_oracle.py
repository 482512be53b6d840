"""The NM / MD:Z / XB:Z restatement (oracle/ksw_oracle.c orc_gen_md, following
bsbolt/External/BWA/bwa.c:180-197,250-343) against the reference's own
bwa_gen_cigar2: the committed fixture (tests/golden/md_bsbolt.npz, written by
tests/golden/make_golden_md.py from the compiled reference) and, where the
compiled reference is present, fresh cases; plus closed-form checks."""
import os

import numpy as np
import pytest

import md_cases as M
from oracle import oracle as O

FIX = os.path.join(os.path.dirname(__file__), "golden", "md_bsbolt.npz")


def _fixture():
    d = np.load(FIX)
    g = d["genome"]
    opac = M.pack2(np.concatenate([g, g]))
    return d, opac


def _sl(d, k, i):
    o = d[k + "_off"] if k + "_off" in d else d[{"query": "q_off", "oquery": "oq_off", "cigar": "cg_off"}[k]]
    return d[k][o[i]:o[i + 1]]


def test_orc_gen_md_matches_reference_fixture():
    d, opac = _fixture()
    l_pac, crick_l = int(d["l_pac"]), int(d["crick_l"])
    n = len(d["rb"])
    assert n > 700
    kinds = set()
    for i in range(n):
        st, md, xb, nm, meth = O.gen_md(l_pac, opac, int(d["rb"][i]), int(d["re"][i]), crick_l,
                                        _sl(d, "oquery", i), _sl(d, "cigar", i))
        if not d["has_md"][i]:
            assert st == 1 and nm == -1, i
            kinds.add("bridge")
            continue
        assert st == 0, i
        assert md == bytes(_sl(d, "md", i)).decode(), i
        assert xb == bytes(_sl(d, "xb", i)).decode(), i
        assert nm == int(d["nm"][i]), i
        assert meth == [int(x) for x in d["meth"][i]], i
        kinds.add("X" in xb or "x" in xb)
    assert "bridge" in kinds and True in kinds


def test_cigar_stage_matches_reference_fixture():
    d, opac = _fixture()
    g = d["genome"]
    _, pac, _, l_pac, _ = M.index(len(g), 7)
    mat = O.fill_scmat()
    for i in range(0, len(d["rb"]), 7):
        cg = M.cigar_stage(mat, l_pac, pac, _sl(d, "query", i), int(d["rb"][i]), int(d["re"][i]))
        assert np.array_equal(cg, _sl(d, "cigar", i)), i


def test_exact_read_closed_form():
    """A read equal to its region (no conversion, no mutation) on each strand:
    MD is the region length, NM 0, and XB holds one methylated context letter
    per C (G on the Crick copy) of the region, no other letters."""
    g, pac, opac, l_pac, crick_l = M.index(4000, 3)
    for rb in (100, crick_l + 250, l_pac + 300, l_pac + crick_l + 400):
        re = rb + 80
        oq = O.get_seq(l_pac, opac, rb, re).astype(np.uint8)
        is_crick = (crick_l <= rb < l_pac) or rb >= crick_l + l_pac
        rbase = 2 if is_crick else 1
        cig = np.array([80 << 4], np.uint32)
        st, md, xb, nm, meth = O.gen_md(l_pac, opac, rb, re, crick_l, oq, cig)
        assert st == 0 and md == "80" and nm == 0
        letters = [c for c in xb if c.isalpha()]
        assert len(letters) == int((oq == rbase).sum()) == sum(meth[0::2])
        assert all(c.isupper() for c in letters) and meth[1] == meth[3] == 0


@pytest.mark.skipif(not os.path.exists(O.BWAREF_PATH), reason="compiled reference (oracle/_ref) absent")
def test_orc_gen_md_matches_live_reference():
    g, pac, opac, l_pac, crick_l = M.index(30000, 21)
    rng = np.random.default_rng(5)
    mat = O.fill_scmat()
    seen = 0
    for rb, re in M.regions(rng, 250, 30000, l_pac):
        q, oq = M.reads_for(rng, l_pac, opac, crick_l, rb, re, p_unmeth=0.4, p_sub=0.05, p_indel=0.03)
        r = O.ref_gen_cigar2(mat, 6, 1, 6, 1, 100, l_pac, pac, opac, q, rb, re, crick_l, oq)
        if r is None:
            continue
        st, md, xb, nm, meth = O.gen_md(l_pac, opac, rb, re, crick_l, oq, r["cigar"])
        if r["md"] is None:
            assert st == 1
            continue
        if st == -2:  # the reference reads past its clamped window: only at the reference ends
            assert rb < 2 or re > 2 * l_pac - 3, (rb, re)
            continue
        assert (st, md, xb, nm, meth) == (0, r["md"], r["xb"], r["NM"], r["meth"]), (rb, re)
        seen += 1
    assert seen > 200


def test_orc_gen_md_rejects_inconsistent_inputs():
    """Regions the reference makes no CIGAR for and CIGARs that do not fit the
    read get status -3 and no strings (the reference would read past its
    buffers there)."""
    g, pac, opac, l_pac, crick_l = M.index(4000, 3)
    oq = O.get_seq(l_pac, opac, 100, 150).astype(np.uint8)
    ok = np.array([50 << 4], np.uint32)
    assert O.gen_md(l_pac, opac, 100, 150, crick_l, oq, ok)[0] == 0
    for rb, re, cig in ((100, 150, np.array([51 << 4], np.uint32)),          # consumes 51 of 50 read bases
                        (100, 150, np.array([(2 << 4) | 4, 48 << 4], np.uint32)),  # soft clip: not a ksw_global2 op
                        (150, 150, ok), (160, 150, ok), (-1, 49, ok),         # empty, reversed, negative region
                        (2 * l_pac - 20, 2 * l_pac + 30, ok)):                 # past the doubled reference
        st, md, xb, nm, meth = O.gen_md(l_pac, opac, rb, re, crick_l, oq, cig)
        assert (st, md, xb, nm) == (-3, "", "", -1), (rb, re)


def test_reads_with_n_bases():
    """N (code 4) in the read: a mismatch against the reference base in MD."""
    g, pac, opac, l_pac, crick_l = M.index(4000, 3)
    rb, re = 500, 560
    oq = O.get_seq(l_pac, opac, rb, re).astype(np.uint8)
    oq[10] = 4
    st, md, xb, nm, meth = O.gen_md(l_pac, opac, rb, re, crick_l, oq, np.array([60 << 4], np.uint32))
    assert st == 0 and nm == 1
    ref = O.get_seq(l_pac, opac, rb, re)
    assert md == f"10{'ACGT'[ref[10]]}49"
