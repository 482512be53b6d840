"""ctypes front end of the CPU oracle (oracle/ksw_oracle.c) and of the compiled
reference kernel (oracle/_ref/libkswref.so, built from the reference's own
bsbolt/External/BWA/ksw.c where that source is present).

Test infrastructure only: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libkswref.so")
BWAREF_PATH = os.path.join(HERE, "_ref", "libbwaref.so")

_lib = None
_ref = None
_bwaref = None

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.orc_extend2.restype = _I
        L.orc_extend2.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I,
                                  _P, _P, _P, _P, _P, _P]
        L.orc_global2.restype = _I
        L.orc_global2.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _P, _P]
        L.orc_extend2_batch.restype = C.c_double
        L.orc_extend2_batch.argtypes = [_P, _P, _I, _I, _I, _I, _I, _I64, _P, _P, _P, _P, _P, _P,
                                        _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I]
        L.orc_global2_batch.restype = C.c_double
        L.orc_global2_batch.argtypes = [_P, _P, _I, _I, _I, _I, _I64, _P, _P, _P, _P, _P, _P, _P,
                                        _P, _P, _P, _I64, _I]
        L.orc_cal_max_gap.restype = _I
        L.orc_cal_max_gap.argtypes = [_I] * 7
        L.orc_global_band.restype = _I
        L.orc_global_band.argtypes = [_I] * 8
        L.orc_get_seq.restype = _I64
        L.orc_get_seq.argtypes = [_I64, _P, _I64, _I64, _P]
        L.orc_fill_scmat.restype = None
        L.orc_fill_scmat.argtypes = [_I, _I, _P]
        L.orc_free.argtypes = [_P]
        L.orc_nt4.restype = _I
        L.orc_nt4.argtypes = [_I]
        L.orc_bs_convert.argtypes = [C.c_char_p, _I, _I]
        L.orc_gen_md.restype = _I
        L.orc_gen_md.argtypes = [_I64, _P, _I64, _I64, _I64, _I, _P, _I, _P, _P, _I, _P, _P, _I, _P, _P, _P]
        L.orc_gen_md_batch.restype = C.c_double
        L.orc_gen_md_batch.argtypes = [_I64, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _I64, _P, _I64, _P, _P,
                                       _I64, _P, _P, _P, _P, _I]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own ksw.c, compiled by oracle/Makefile (container only)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"compiled reference missing: {REF_PATH}")
        R = C.CDLL(REF_PATH)
        R.ksw_extend2.restype = _I
        R.ksw_extend2.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _I, _I, _I,
                                  _P, _P, _P, _P, _P]
        R.ksw_global2.restype = _I
        R.ksw_global2.argtypes = [_I, _P, _I, _P, _I, _P, _I, _I, _I, _I, _I, _P, _P]
        _ref = R
    return _ref


def bwaref():
    """The reference's bwa.c (bwa_gen_cigar2) with the sources it links, compiled by
    oracle/Makefile (container only)."""
    global _bwaref
    if _bwaref is None:
        if not os.path.exists(BWAREF_PATH):
            raise RuntimeError(f"compiled reference missing: {BWAREF_PATH}")
        R = C.CDLL(BWAREF_PATH)
        R.bwa_gen_cigar2.restype = C.POINTER(C.c_uint32)
        R.bwa_gen_cigar2.argtypes = [_P, _I, _I, _I, _I, _I, _I64, _P, _P, _I, _P, _I64, _I64, _P, _P, _P, _I64, _P,
                                     _P, _P, _P, _P]
        _bwaref = R
    return _bwaref


def ref_gen_cigar2(mat, o_del, e_del, o_ins, e_ins, w, l_pac, pac, opac, query, rb, re, crick_l, oquery):
    """One bwa_gen_cigar2 call of the compiled reference (bwa.c:199-351).  Returns
    None when it returns no CIGAR, else dict(score, cigar [uint32], NM, md, xb,
    meth [cg_meth, cg_unmeth, ch_meth, ch_unmeth]); md and xb are None when the
    reference appended no strings (its context window bridged the strands)."""
    R = bwaref()
    q = np.array(query, np.uint8)
    oq = np.array(oquery, np.uint8)
    mat = np.ascontiguousarray(mat, np.int8)
    sc, nc, nm = C.c_int(0), C.c_int(0), C.c_int(0)
    m = [C.c_int(0) for _ in range(4)]
    p = R.bwa_gen_cigar2(mat.ctypes.data, o_del, e_del, o_ins, e_ins, w, l_pac, pac.ctypes.data, opac.ctypes.data,
                         len(q), q.ctypes.data, rb, re, C.byref(sc), C.byref(nc), C.byref(nm), crick_l,
                         oq.ctypes.data, *[C.byref(x) for x in m])
    if not p:
        return None
    n = nc.value
    cig = np.array([p[k] for k in range(n)], np.uint32)
    md = xb = None
    if nm.value >= 0:
        raw = C.string_at(C.addressof(p.contents) + 4 * n)
        md_b, _, rest = raw.partition(b"\t")
        md, xb = md_b.decode(), rest[len(b"XB:Z:"):].decode()
    lib().orc_free(C.cast(p, _P))
    return dict(score=sc.value, cigar=cig, NM=nm.value, md=md, xb=xb, meth=[x.value for x in m])


def gen_md(l_pac, opac, rb, re, crick_l, oquery, cigar, md_cap=4096, xb_cap=4096):
    """orc_gen_md for one task: (status, md, xb, NM, meth[4])."""
    oq = np.ascontiguousarray(oquery, np.uint8)
    cg = np.ascontiguousarray(cigar, np.uint32)
    md = C.create_string_buffer(md_cap)
    xb = C.create_string_buffer(xb_cap)
    ml, xl, nm = C.c_int(0), C.c_int(0), C.c_int(0)
    cnt = np.zeros(4, np.int32)
    r = lib().orc_gen_md(l_pac, opac.ctypes.data, rb, re, crick_l, len(oq), oq.ctypes.data, len(cg), cg.ctypes.data,
                         md, md_cap, C.byref(ml), xb, xb_cap, C.byref(xl), C.byref(nm), cnt.ctypes.data)
    return r, md.raw[:ml.value].decode(), xb.raw[:xl.value].decode(), nm.value, [int(x) for x in cnt]


def gen_md_batch(l_pac, opac, crick_l, rb, re, oquery, oq_off, l_query, n_cigar, cigar, md_slot, xb_slot,
                 nthreads=1):
    """orc_gen_md over a batch (cigar [n, slot]) on nthreads host threads.
    Returns (dict of md / xb (lists of str), nm, meth [n, 4], status; seconds)."""
    n = len(rb)
    arr = lambda x, dt: np.ascontiguousarray(x, dt)
    rb, re, oq_off = arr(rb, np.int64), arr(re, np.int64), arr(oq_off, np.uint64)
    l_query, n_cigar, cigar = arr(l_query, np.int32), arr(n_cigar, np.int32), arr(cigar, np.uint32).reshape(n, -1)
    opac, oquery = arr(opac, np.uint8), arr(oquery, np.uint8)
    md = np.zeros((n, md_slot), np.uint8)
    xb = np.zeros((n, xb_slot), np.uint8)
    ml, xl, nm, st = (np.zeros(n, np.int32) for _ in range(4))
    meth = np.zeros((n, 4), np.int32)
    dt = lib().orc_gen_md_batch(n, l_pac, opac.ctypes.data, crick_l, rb.ctypes.data, re.ctypes.data,
                                oquery.ctypes.data, oq_off.ctypes.data, l_query.ctypes.data, n_cigar.ctypes.data,
                                cigar.ctypes.data, cigar.shape[1], md.ctypes.data, md_slot, ml.ctypes.data,
                                xb.ctypes.data, xb_slot, xl.ctypes.data, nm.ctypes.data, meth.ctypes.data,
                                st.ctypes.data, nthreads)
    return {"md": [bytes(md[i, :ml[i]]).decode() for i in range(n)],
            "xb": [bytes(xb[i, :xl[i]]).decode() for i in range(n)], "nm": nm, "meth": meth, "status": st}, dt


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def fill_scmat(a=1, b=4):
    m = np.zeros(25, np.int8)
    lib().orc_fill_scmat(a, b, _ptr(m))
    return m


def extend2(query, target, h0, w=100, end_bonus=30, zdrop=100, mat=None,
            o_del=6, e_del=1, o_ins=6, e_ins=1, use_ref=False):
    """One extension.  Returns (score, qle, tle, gtle, gscore, max_off, cells);
    cells is None for the compiled reference (it does not count them)."""
    q = _u8(query)
    t = _u8(target)
    mat = fill_scmat() if mat is None else np.ascontiguousarray(mat, np.int8)
    out = np.zeros(5, np.int32)
    o = [out[k:k + 1].ctypes.data_as(_P) for k in range(5)]
    if use_ref:
        s = ref().ksw_extend2(len(q), _ptr(q), len(t), _ptr(t), 5, _ptr(mat), o_del, e_del,
                              o_ins, e_ins, w, end_bonus, zdrop, h0, *o)
        return (int(s), *[int(x) for x in out], None)
    cells = np.zeros(1, np.int64)
    s = lib().orc_extend2(len(q), _ptr(q), len(t), _ptr(t), 5, _ptr(mat), o_del, e_del,
                          o_ins, e_ins, w, end_bonus, zdrop, h0, *o, _ptr(cells))
    return (int(s), *[int(x) for x in out], int(cells[0]))


def global2(query, target, w, mat=None, o_del=6, e_del=1, o_ins=6, e_ins=1,
            cigar=True, use_ref=False):
    """One global alignment.  Returns (score, [cigar ints] or None)."""
    q = _u8(query)
    t = _u8(target)
    mat = fill_scmat() if mat is None else np.ascontiguousarray(mat, np.int8)
    nc = C.c_int(0)
    cp = C.POINTER(C.c_uint32)()
    if use_ref:
        fn, free = ref().ksw_global2, C.CDLL(None).free
    else:
        fn, free = lib().orc_global2, lib().orc_free
    if cigar:
        s = fn(len(q), _ptr(q), len(t), _ptr(t), 5, _ptr(mat), o_del, e_del, o_ins, e_ins, w,
               C.byref(nc), C.byref(cp))
        cg = [int(cp[k]) for k in range(nc.value)]
        if cp:
            free(C.cast(cp, _P))
        return int(s), cg
    s = fn(len(q), _ptr(q), len(t), _ptr(t), 5, _ptr(mat), o_del, e_del, o_ins, e_ins, w, None, None)
    return int(s), None


def extend2_batch(b, nthreads=1, use_ref=False, scoring=None, want_cells=True):
    """Run a task batch (attributes query, q_off, qlen, target, t_off, tlen, h0,
    w, end_bonus) on `nthreads` host threads.  Returns (dict of outputs, seconds)."""
    sc = scoring or {}
    mat = np.ascontiguousarray(sc.get("mat", fill_scmat()), np.int8)
    n = len(b.qlen)
    outs = {k: np.zeros(n, np.int32) for k in ("score", "qle", "tle", "gtle", "gscore", "max_off")}
    cells = np.zeros(n, np.int64) if (want_cells and not use_ref) else None
    fn = C.cast(ref().ksw_extend2, _P) if use_ref else None
    arrs = [np.ascontiguousarray(getattr(b, k)) for k in
            ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w", "end_bonus")]
    assert arrs[1].dtype == np.uint64 and arrs[4].dtype == np.uint64
    assert all(arrs[k].dtype == np.int32 for k in (2, 5, 6, 7, 8))
    dt = lib().orc_extend2_batch(
        fn, _ptr(mat), sc.get("o_del", 6), sc.get("e_del", 1), sc.get("o_ins", 6), sc.get("e_ins", 1),
        sc.get("zdrop", 100), n, *[_ptr(a) for a in arrs],
        *[_ptr(outs[k]) for k in ("score", "qle", "tle", "gtle", "gscore", "max_off")],
        _ptr(cells), nthreads)
    if cells is not None:
        outs["cells"] = cells
    return outs, dt


def global2_batch(b, slot, scoring=None, nthreads=1, use_ref=False):
    """ksw_global2 with CIGARs over a task batch (attributes query, q_off, qlen,
    target, t_off, tlen, w) on `nthreads` host threads: this restatement, or the
    compiled reference with use_ref.  Returns (score, n_cigar, cigar[n, slot],
    seconds)."""
    sc = scoring or {}
    mat = np.ascontiguousarray(sc.get("mat", fill_scmat()), np.int8)
    n = len(b.qlen)
    score = np.zeros(n, np.int32)
    ncig = np.zeros(n, np.int32)
    cig = np.zeros(n * slot, np.uint32)
    fn = C.cast(ref().ksw_global2, _P) if use_ref else None
    arrs = [np.ascontiguousarray(getattr(b, k)) for k in
            ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")]
    dt = lib().orc_global2_batch(fn, _ptr(mat), sc.get("o_del", 6), sc.get("e_del", 1), sc.get("o_ins", 6),
                                 sc.get("e_ins", 1), n, *[_ptr(a) for a in arrs], _ptr(score), _ptr(ncig),
                                 _ptr(cig), slot, nthreads)
    return score, ncig, cig.reshape(n, slot), dt


def cal_max_gap(qlen, a=1, o_del=6, e_del=1, o_ins=6, e_ins=1, w=100):
    return lib().orc_cal_max_gap(a, o_del, e_del, o_ins, e_ins, w, qlen)


def global_band(l_query, rlen, w_=100, mat0=1, o_del=6, e_del=1, o_ins=6, e_ins=1):
    return lib().orc_global_band(mat0, o_del, e_del, o_ins, e_ins, w_, l_query, rlen)


def get_seq(l_pac, pac, beg, end):
    pac = _u8(pac)
    out = np.zeros(max(0, abs(end - beg)) + 1, np.uint8)
    n = lib().orc_get_seq(l_pac, _ptr(pac), beg, end, _ptr(out))
    return out[:n]


def nt4(ch):
    return lib().orc_nt4(ord(ch))


def bs_convert(seq, pattern):
    buf = C.create_string_buffer(seq.encode())
    lib().orc_bs_convert(buf, pattern, len(seq))
    return buf.value.decode()


def reg2aln_cigar(l_pac, rb, re, qb, qe, l_seq, cigar):
    """orc_reg2aln_cigar for one region: (n_ops or -1/-2, [SAM CIGAR words], pos, is_rev)."""
    L = lib()
    c = np.ascontiguousarray(cigar, np.uint32)
    out = np.zeros(len(c) + 2, np.uint32)
    pos, rev = C.c_int64(0), C.c_int(0)
    fn = L.orc_reg2aln_cigar
    fn.restype = C.c_int
    fn.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P]
    k = fn(l_pac, rb, re, qb, qe, l_seq, len(c), _ptr(c), _ptr(out), len(out), C.byref(pos), C.byref(rev))
    return k, [int(x) for x in out[:max(k, 0)]], pos.value, rev.value
