"""Python host of libbsx.so (ctypes), the role bsbolt.Align plays for the
reference aligner, with the operator interface of the reference kernels:

  ksw_extend2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins,
              w, end_bonus, zdrop, h0) -> (score, qle, tle, gtle, gscore, max_off)
      bsbolt/External/BWA/ksw.h:108, ksw.c:380-479
  ksw_global2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w)
      -> (score, n_cigar, cigar)           ksw.h:84, ksw.c:504-606

plus the batched forms the GPU is built for (Engine.extend / Engine.global_).
Every call runs on the GPU through the C ABI; there is no CPU fallback: if the
library or a GPU is missing, loading fails with an error.
"""
import ctypes as C
import os

import numpy as np

from .tasks import ExtTasks, GlbTasks, RESULT_FIELDS

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsx.so")

BSX_OK, BSX_EINVAL, BSX_EDEVICE, BSX_ENOMEM, BSX_ERANGE = 0, -1, -2, -3, -4
BSX_MAX_QLEN = 2048
BSX_NOALLOC_MAX_TASKS = 262144
TUNE_FIELDS = 4
_ERRNAME = {BSX_EINVAL: "BSX_EINVAL", BSX_EDEVICE: "BSX_EDEVICE", BSX_ENOMEM: "BSX_ENOMEM",
            BSX_ERANGE: "BSX_ERANGE"}

# Every symbol include/bsx.h declares.
EXPORTS = ("bsx_create", "bsx_destroy", "bsx_last_error", "bsx_set_stream", "bsx_sync", "bsx_version",
           "bsx_default_scoring", "bsx_extend_batch", "bsx_extend_batch_device", "bsx_global_batch",
           "bsx_global_batch_device", "bsx_set_global_zslot", "bsx_ksw_extend2", "bsx_dev_alloc",
           "bsx_dev_free", "bsx_memcpy", "bsx_timer_start", "bsx_timer_stop", "bsx_set_band_policy",
           "bsx_band_stats", "bsx_band_model_set", "bsx_band_model_get", "bsx_band_features_device",
           "bsx_band_labels_device", "bsx_band_forward_device", "bsx_band_grad_device",
           "bsx_band_train_device", "bsx_set_kernel_timing", "bsx_kernel_times", "bsx_ext_bin_limits",
           "bsx_graph_capture_extend", "bsx_graph_capture_global", "bsx_graph_launch", "bsx_graph_destroy",
           "bsx_set_side_streams", "bsx_md_batch_device", "bsx_band_predict_device", "bsx_dropin_contexts",
           "bsx_extend_batch_device_noalloc", "bsx_ksw_global2", "bsx_reg2aln_cigar_device",
           "bsx_set_tuning", "bsx_get_tuning")
TIMED_SLOTS = 16
# lanes per task and columns per lane of the plain form's bins (bsx_internal.h kExtShape);
# one lane per task is the lane-per-task kernel
EXT_SHAPES = ((4, 4), (1, 32), (1, 48), (1, 64), (16, 6), (16, 8), (32, 6), (32, 8), (64, 8), (64, 16), (64, 32))

# learned-band form (include/bsx.h)
BAND_PLAIN, BAND_LEARNED, BAND_FIXED = 0, 1, 2
BAND_CLASSES = 5              # classes 0..3 = band-following frames of 24/40/80/144 columns, 4 = plain
BAND_FRAMES = (24, 40, 80, 144)
BAND_FEATURES = 24
BAND_FEATURE_STRIDE = 32
BAND_PARAMS = 6085


class BsxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAME.get(code, code)}: {msg}")
        self.code = code


class Scoring(C.Structure):
    _fields_ = [("mat", C.c_int8 * 25), ("o_del", C.c_int32), ("e_del", C.c_int32),
                ("o_ins", C.c_int32), ("e_ins", C.c_int32), ("zdrop", C.c_int32)]

    @classmethod
    def make(cls, mat=None, o_del=6, e_del=1, o_ins=6, e_ins=1, zdrop=100, a=1, b=4):
        s = cls()
        m = fill_scmat(a, b) if mat is None else np.asarray(mat, np.int8).reshape(25)
        for k in range(25):
            s.mat[k] = int(m[k])
        s.o_del, s.e_del, s.o_ins, s.e_ins, s.zdrop = o_del, e_del, o_ins, e_ins, zdrop
        return s

    def as_dict(self):
        return dict(mat=np.array(list(self.mat), np.int8), o_del=self.o_del, e_del=self.e_del,
                    o_ins=self.o_ins, e_ins=self.e_ins, zdrop=self.zdrop)


class _ExtTasks(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in
                                     ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w",
                                      "end_bonus")]


class _ExtResults(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in RESULT_FIELDS + ("cells",)]


class _GlbTasks(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in
                                     ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")]


class _GlbResults(C.Structure):
    _fields_ = [("score", C.c_void_p), ("n_cigar", C.c_void_p), ("cigar", C.c_void_p),
                ("cigar_slot", C.c_int64)]


class _MdTasks(C.Structure):
    _fields_ = [("n", C.c_int64), ("l_pac", C.c_int64), ("crick_l", C.c_int64), ("opac", C.c_void_p),
                ("rb", C.c_void_p), ("re", C.c_void_p), ("oquery", C.c_void_p), ("oq_off", C.c_void_p),
                ("l_query", C.c_void_p), ("n_cigar", C.c_void_p), ("cigar", C.c_void_p), ("cigar_slot", C.c_int64)]


class _Reg2alnTasks(C.Structure):
    _fields_ = [("n", C.c_int64), ("l_pac", C.c_int64), ("rb", C.c_void_p), ("re", C.c_void_p),
                ("qb", C.c_void_p), ("qe", C.c_void_p), ("l_seq", C.c_void_p), ("n_cigar", C.c_void_p),
                ("cigar", C.c_void_p), ("cigar_slot", C.c_int64)]


class _Reg2alnResults(C.Structure):
    _fields_ = [("cigar", C.c_void_p), ("cigar_slot", C.c_int64), ("n_cigar", C.c_void_p), ("pos", C.c_void_p),
                ("is_rev", C.c_void_p)]


class _MdResults(C.Structure):
    _fields_ = [("md", C.c_void_p), ("md_slot", C.c_int64), ("md_len", C.c_void_p), ("xb", C.c_void_p),
                ("xb_slot", C.c_int64), ("xb_len", C.c_void_p), ("nm", C.c_void_p), ("meth", C.c_void_p),
                ("status", C.c_void_p)]


_lib = None


def lib():
    """Load libbsx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libbsx.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.bsx_create.restype = C.c_void_p
        L.bsx_create.argtypes = [C.c_int]
        L.bsx_destroy.argtypes = [C.c_void_p]
        L.bsx_last_error.restype = C.c_char_p
        L.bsx_last_error.argtypes = [C.c_void_p]
        L.bsx_set_stream.argtypes = [C.c_void_p, C.c_void_p]
        L.bsx_sync.argtypes = [C.c_void_p]
        L.bsx_version.restype = C.c_char_p
        L.bsx_default_scoring.argtypes = [C.POINTER(Scoring)]
        for f, tk, rs in (("bsx_extend_batch", _ExtTasks, _ExtResults),
                          ("bsx_extend_batch_device", _ExtTasks, _ExtResults),
                          ("bsx_extend_batch_device_noalloc", _ExtTasks, _ExtResults),
                          ("bsx_global_batch", _GlbTasks, _GlbResults),
                          ("bsx_global_batch_device", _GlbTasks, _GlbResults)):
            getattr(L, f).argtypes = [C.c_void_p, C.POINTER(Scoring), C.POINTER(tk), C.POINTER(rs)]
            getattr(L, f).restype = C.c_int
        L.bsx_set_global_zslot.argtypes = [C.c_void_p, C.c_int64]
        L.bsx_dev_alloc.restype = C.c_void_p
        L.bsx_dev_alloc.argtypes = [C.c_void_p, C.c_size_t]
        L.bsx_dev_free.argtypes = [C.c_void_p, C.c_void_p]
        L.bsx_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.bsx_timer_start.argtypes = [C.c_void_p]
        L.bsx_timer_stop.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        P = C.c_void_p
        L.bsx_set_band_policy.argtypes = [P, C.c_int, C.c_int]
        L.bsx_band_stats.argtypes = [P, P]
        L.bsx_band_model_set.argtypes = [P, P, C.c_int64]
        L.bsx_band_model_get.argtypes = [P, P, C.c_int64]
        L.bsx_band_features_device.argtypes = [P, C.POINTER(Scoring), C.POINTER(_ExtTasks), P]
        L.bsx_band_labels_device.argtypes = [P, C.POINTER(Scoring), C.POINTER(_ExtTasks), P]
        L.bsx_band_predict_device.argtypes = [P, C.POINTER(Scoring), C.POINTER(_ExtTasks), P]
        L.bsx_band_predict_device.restype = C.c_int
        L.bsx_band_forward_device.argtypes = [P, P, C.c_int64, P, P]
        L.bsx_band_grad_device.argtypes = [P, P, P, C.c_int64, P]
        L.bsx_band_train_device.argtypes = [P, P, P, C.c_int64, C.c_int64, C.c_int, C.c_float, P]
        L.bsx_graph_capture_extend.argtypes = [P, C.POINTER(Scoring), C.POINTER(_ExtTasks), C.POINTER(_ExtResults),
                                               C.POINTER(C.c_void_p)]
        L.bsx_graph_capture_global.argtypes = [P, C.POINTER(Scoring), C.POINTER(_GlbTasks), C.POINTER(_GlbResults),
                                               C.POINTER(C.c_void_p)]
        L.bsx_graph_capture_extend.restype = C.c_int
        L.bsx_graph_capture_global.restype = C.c_int
        L.bsx_graph_launch.argtypes = [P, P]
        L.bsx_graph_launch.restype = C.c_int
        L.bsx_graph_destroy.argtypes = [P]
        L.bsx_graph_destroy.restype = None
        L.bsx_set_kernel_timing.argtypes = [P, C.c_int]
        L.bsx_kernel_times.argtypes = [P, P, C.c_int, P]
        L.bsx_ext_bin_limits.argtypes = [P, C.c_int]
        L.bsx_set_side_streams.argtypes = [P, C.c_int]
        L.bsx_md_batch_device.argtypes = [P, C.POINTER(_MdTasks), C.POINTER(_MdResults)]
        L.bsx_md_batch_device.restype = C.c_int
        L.bsx_reg2aln_cigar_device.argtypes = [P, C.POINTER(_Reg2alnTasks), C.POINTER(_Reg2alnResults)]
        L.bsx_reg2aln_cigar_device.restype = C.c_int
        L.bsx_set_tuning.argtypes = [P, P, C.c_int]
        L.bsx_set_tuning.restype = C.c_int
        L.bsx_get_tuning.argtypes = [P, P, C.c_int, C.c_char_p, C.c_int]
        L.bsx_get_tuning.restype = C.c_int
        for f in ("bsx_set_kernel_timing", "bsx_kernel_times", "bsx_ext_bin_limits", "bsx_set_side_streams"):
            getattr(L, f).restype = C.c_int
        for f in ("bsx_set_band_policy", "bsx_band_stats", "bsx_band_model_set", "bsx_band_model_get",
                  "bsx_band_features_device", "bsx_band_labels_device", "bsx_band_forward_device",
                  "bsx_band_grad_device", "bsx_band_train_device"):
            getattr(L, f).restype = C.c_int
        L.bsx_ksw_extend2.restype = C.c_int
        L.bsx_ksw_extend2.argtypes = [C.c_int, P, C.c_int, P, C.c_int, P] + [C.c_int] * 8 + [P] * 5
        L.bsx_ksw_global2.restype = C.c_int
        L.bsx_ksw_global2.argtypes = [C.c_int, P, C.c_int, P, C.c_int, P] + [C.c_int] * 5 + [P, P]
        L.bsx_dropin_contexts.restype = C.c_int64
        L.bsx_dropin_contexts.argtypes = []
        _lib = L
    return _lib


def ext_bin_limits():
    """Largest query length of each plain-kernel bin."""
    n = lib().bsx_ext_bin_limits(None, 0)
    out = np.zeros(n, np.int32)
    lib().bsx_ext_bin_limits(out.ctypes.data, n)
    return out


def fill_scmat(a=1, b=4):
    """bwa_fill_scmat (bsbolt/External/BWA/bwa.c:169-178)."""
    m = np.full((5, 5), -b, np.int8)
    np.fill_diagonal(m, a)
    m[4, :] = -1
    m[:, 4] = -1
    return m.reshape(25)


def _p(a):
    return None if a is None else a.ctypes.data


class Engine:
    """A libbsx context on one GPU (one HIP stream)."""

    def __init__(self, device=0):
        self.L = lib()
        self.ctx = self.L.bsx_create(device)
        if not self.ctx:
            raise RuntimeError(f"bsx_create({device}) failed: no usable GPU")

    def close(self):
        if self.ctx:
            self.L.bsx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != BSX_OK:
            raise BsxError(rc, self.L.bsx_last_error(self.ctx).decode())

    def last_error(self):
        return self.L.bsx_last_error(self.ctx).decode()

    def set_stream(self, stream_handle):
        self._check(self.L.bsx_set_stream(self.ctx, stream_handle))

    def sync(self):
        self._check(self.L.bsx_sync(self.ctx))

    def set_global_zslot(self, nbytes):
        self.L.bsx_set_global_zslot(self.ctx, int(nbytes))

    # ---- device memory ----
    def to_device(self, arr):
        """Copies a numpy array into new device memory; returns a DevArray."""
        arr = np.ascontiguousarray(arr)
        d = DevArray(self, arr.nbytes, arr.dtype, arr.shape)
        if arr.nbytes:
            self._check(self.L.bsx_memcpy(self.ctx, d.ptr, arr.ctypes.data, arr.nbytes, 1))
        return d

    def empty(self, shape, dtype):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return DevArray(self, n, np.dtype(dtype), shape)

    def timer_start(self):
        self._check(self.L.bsx_timer_start(self.ctx))

    def timer_stop(self):
        ms = C.c_float(0)
        self._check(self.L.bsx_timer_stop(self.ctx, C.byref(ms)))
        return ms.value

    def tasks_to_device(self, tasks):
        """Uploads an ExtTasks batch; returns (dict of DevArrays, dict of addresses)."""
        keep = {k: self.to_device(getattr(tasks, k)) for k in
                ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w", "end_bonus")}
        return keep, {k: v.ptr for k, v in keep.items()}

    # ---- extension ----
    def extend(self, tasks: ExtTasks, scoring: Scoring = None, cells=True, out=None, raise_on_error=True):
        """Host-memory batch of ksw_extend2 calls -> dict of int32 arrays
        (score, qle, tle, gtle, gscore, max_off) and int64 cells."""
        sc = scoring or Scoring.make()
        n = len(tasks)
        res = out if out is not None else {k: np.zeros(n, np.int32) for k in RESULT_FIELDS}
        if cells and "cells" not in res:
            res["cells"] = np.zeros(n, np.int64)
        tk = _ExtTasks(n, *[_p(getattr(tasks, k)) for k in
                            ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w", "end_bonus")])
        rs = _ExtResults(*[_p(res.get(k)) for k in RESULT_FIELDS + ("cells",)])
        rc = self.L.bsx_extend_batch(self.ctx, C.byref(sc), C.byref(tk), C.byref(rs))
        if raise_on_error:
            self._check(rc)
        return res if raise_on_error else (res, rc)

    def extend_device(self, dev, scoring: Scoring = None, noalloc: bool = False):
        """Device-memory batch: `dev` maps the task/result field names to
        device addresses (e.g. torch tensors' data_ptr()).  Asynchronous; call
        sync() to wait and collect precondition errors.  noalloc=True calls the
        allocation-free entry point (at most BSX_NOALLOC_MAX_TASKS tasks)."""
        sc = scoring or Scoring.make()
        tk = _ExtTasks(dev["n"], *[dev.get(k) for k in
                                   ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w",
                                    "end_bonus")])
        rs = _ExtResults(*[dev.get(k) for k in RESULT_FIELDS + ("cells",)])
        f = self.L.bsx_extend_batch_device_noalloc if noalloc else self.L.bsx_extend_batch_device
        self._check(f(self.ctx, C.byref(sc), C.byref(tk), C.byref(rs)))

    # ---- graph-captured submission ----
    def capture_extend(self, dev, scoring: Scoring = None):
        """Graph of one device-memory extension batch (pointers as extend_device)."""
        sc = scoring or Scoring.make()
        tk = self._ext_tasks_dev(dev)
        rs = _ExtResults(*[dev.get(k) for k in RESULT_FIELDS + ("cells",)])
        h = C.c_void_p()
        self._check(self.L.bsx_graph_capture_extend(self.ctx, C.byref(sc), C.byref(tk), C.byref(rs), C.byref(h)))
        return Graph(self, h.value)

    def capture_global(self, dev, scoring: Scoring = None):
        sc = scoring or Scoring.make()
        tk = _GlbTasks(dev["n"], *[dev.get(k) for k in
                                   ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")])
        rs = _GlbResults(dev.get("score"), dev.get("n_cigar"), dev.get("cigar"), dev.get("cigar_slot", 0))
        h = C.c_void_p()
        self._check(self.L.bsx_graph_capture_global(self.ctx, C.byref(sc), C.byref(tk), C.byref(rs), C.byref(h)))
        return Graph(self, h.value)

    def stream_global(self, tasks: GlbTasks, batch=262144, slot=None, scoring: Scoring = None):
        """CIGARs of a host batch streamed through one captured graph per
        `batch` tasks (the last, shorter chunk runs eagerly).  Returns
        (score, n_cigar, cigar[n, slot]) like global_."""
        sc = scoring or Scoring.make()
        n = len(tasks)
        if slot is None:
            slot = int((tasks.qlen + tasks.tlen).max()) if n else 1
        score = np.zeros(n, np.int32)
        ncig = np.zeros(n, np.int32)
        cig = np.zeros((n, slot), np.uint32)
        full = (n // batch) * batch
        if full:
            from .tasks import GlbTasks as _G
            chunks = [slice(a, a + batch) for a in range(0, full, batch)]
            subs = []
            for sl in chunks:
                q0, q1 = int(tasks.q_off[sl.start]), int(tasks.q_off[sl.stop - 1]) + int(tasks.qlen[sl.stop - 1])
                t0, t1 = int(tasks.t_off[sl.start]), int(tasks.t_off[sl.stop - 1]) + int(tasks.tlen[sl.stop - 1])
                subs.append(_G(tasks.query[q0:q1], tasks.q_off[sl] - q0, tasks.qlen[sl], tasks.target[t0:t1],
                               tasks.t_off[sl] - t0, tasks.tlen[sl], tasks.w[sl]))
            fields = ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")
            dev = {k: self.empty((max(getattr(b, k).nbytes for b in subs),), np.uint8) for k in fields}
            ds, dn = self.empty((batch,), np.int32), self.empty((batch,), np.int32)
            dc = self.empty((batch * slot,), np.uint32)
            ptrs = {k: v.ptr for k, v in dev.items()}
            ptrs.update(n=batch, score=ds.ptr, n_cigar=dn.ptr, cigar=dc.ptr, cigar_slot=slot)
            for k in fields:
                dev[k].upload(getattr(subs[0], k))
            g = self.capture_global(ptrs, sc)
            try:
                for sl, b in zip(chunks, subs):
                    for k in fields:
                        dev[k].upload(getattr(b, k))
                    g.launch()
                    self.sync()
                    score[sl], ncig[sl] = ds.numpy(), dn.numpy()
                    # words past a task's n_cigar are left as the device buffer held them
                    blk = dc.numpy().reshape(batch, slot)
                    cig[sl] = np.where(np.arange(slot)[None, :] < ncig[sl][:, None], blk, 0)
            finally:
                g.close()
        if full < n:
            rest = _glb_subset(tasks, full, n)
            s_, nc_, cg_ = self.global_(rest, sc, cigar=True, slot=slot)
            score[full:], ncig[full:], cig[full:] = s_, nc_, cg_
        return score, ncig, cig

    # ---- per-kernel timing ----
    def set_side_streams(self, enable):
        """Spread independent launches over side streams (default) or run them in order."""
        self._check(self.L.bsx_set_side_streams(self.ctx, 1 if enable else 0))

    def set_kernel_timing(self, max_calls):
        self._check(self.L.bsx_set_kernel_timing(self.ctx, int(max_calls)))

    def kernel_times(self):
        """(ms summed per slot, number of calls recorded); slots: plain bins, narrow classes, predictor."""
        ms = np.zeros(TIMED_SLOTS, np.float32)
        calls = C.c_int(0)
        self._check(self.L.bsx_kernel_times(self.ctx, ms.ctypes.data, TIMED_SLOTS, C.byref(calls)))
        return ms, calls.value

    # ---- learned-band form ----
    def set_band_policy(self, policy, band_class=0):
        self._check(self.L.bsx_set_band_policy(self.ctx, int(policy), int(band_class)))

    def band_stats(self):
        """(routed, rejected) per class of the last learned/fixed extension call."""
        out = np.zeros(2 * BAND_CLASSES, np.int64)
        self._check(self.L.bsx_band_stats(self.ctx, out.ctypes.data))
        return out.reshape(BAND_CLASSES, 2)

    def set_band_model(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(-1)
        self._check(self.L.bsx_band_model_set(self.ctx, p.ctypes.data, p.size))

    def get_band_model(self):
        p = np.zeros(BAND_PARAMS, np.float32)
        self._check(self.L.bsx_band_model_get(self.ctx, p.ctypes.data, p.size))
        return p

    def _ext_tasks_dev(self, dev):
        return _ExtTasks(dev["n"], *[dev.get(k) for k in
                                     ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w",
                                      "end_bonus")])

    def band_features_device(self, dev, feats_ptr, scoring: Scoring = None):
        sc = scoring or Scoring.make()
        self._check(self.L.bsx_band_features_device(self.ctx, C.byref(sc), C.byref(self._ext_tasks_dev(dev)),
                                                     feats_ptr))

    def band_labels_device(self, dev, labels_ptr, scoring: Scoring = None):
        sc = scoring or Scoring.make()
        self._check(self.L.bsx_band_labels_device(self.ctx, C.byref(sc), C.byref(self._ext_tasks_dev(dev)),
                                                   labels_ptr))

    def band_features(self, tasks: ExtTasks, scoring: Scoring = None):
        """Features of a host batch -> uint16 [n, 32] (BF16 bit patterns)."""
        keep, dev = self.tasks_to_device(tasks)
        dev["n"] = len(tasks)
        f = self.empty((len(tasks), BAND_FEATURE_STRIDE), np.uint16)
        self.band_features_device(dev, f.ptr, scoring)
        self.sync()
        return f.numpy()

    def band_labels(self, tasks: ExtTasks, scoring: Scoring = None):
        """Smallest frame class whose band-following run holds each task (4: none)."""
        keep, dev = self.tasks_to_device(tasks)
        dev["n"] = len(tasks)
        lab = self.empty((len(tasks),), np.uint8)
        self.band_labels_device(dev, lab.ptr, scoring)
        self.sync()
        return lab.numpy()

    def band_predict(self, tasks: ExtTasks, scoring: Scoring = None):
        """Classes the learned policy's fused predictor gives a host batch."""
        sc = scoring or Scoring.make()
        keep, dev = self.tasks_to_device(tasks)
        dev["n"] = len(tasks)
        cl = self.empty((len(tasks),), np.uint8)
        self._check(self.L.bsx_band_predict_device(self.ctx, C.byref(sc), C.byref(self._ext_tasks_dev(dev)), cl.ptr))
        self.sync()
        return cl.numpy()

    def band_forward(self, feats):
        """MLP forward of host features uint16 [n, 32] -> (logits [n, 5], class [n])."""
        feats = np.ascontiguousarray(feats, np.uint16)
        n = feats.shape[0]
        fd = self.to_device(feats)
        lg = self.empty((n, BAND_CLASSES), np.float32)
        cl = self.empty((n,), np.uint8)
        self._check(self.L.bsx_band_forward_device(self.ctx, fd.ptr, n, lg.ptr, cl.ptr))
        self.sync()
        return lg.numpy(), cl.numpy()

    def band_grad(self, feats, labels):
        """Mean cross-entropy gradient (6085) and mean loss over host samples."""
        feats = np.ascontiguousarray(feats, np.uint16)
        labels = np.ascontiguousarray(labels, np.uint8)
        fd, ld = self.to_device(feats), self.to_device(labels)
        g = self.empty((BAND_PARAMS + 1,), np.float32)
        self._check(self.L.bsx_band_grad_device(self.ctx, fd.ptr, ld.ptr, feats.shape[0], g.ptr))
        self.sync()
        out = g.numpy()
        return out[:BAND_PARAMS], float(out[BAND_PARAMS])

    def band_train_device(self, feats_ptr, labels_ptr, n, batch, steps, lr, losses_ptr=None):
        self._check(self.L.bsx_band_train_device(self.ctx, feats_ptr, labels_ptr, int(n), int(batch), int(steps),
                                                 float(lr), losses_ptr))

    # ---- global ----
    def global_(self, tasks: GlbTasks, scoring: Scoring = None, cigar=True, slot=None,
                raise_on_error=True):
        """Host-memory batch of ksw_global2 calls -> (score, n_cigar, cigar[n, slot])."""
        sc = scoring or Scoring.make()
        n = len(tasks)
        if slot is None:
            slot = int((tasks.qlen + tasks.tlen).max()) if n else 1
        score = np.zeros(n, np.int32)
        ncig = np.zeros(n, np.int32)
        cig = np.zeros(n * slot, np.uint32) if cigar else None
        tk = _GlbTasks(n, *[_p(getattr(tasks, k)) for k in
                            ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")])
        rs = _GlbResults(_p(score), _p(ncig), _p(cig), slot if cigar else 0)
        rc = self.L.bsx_global_batch(self.ctx, C.byref(sc), C.byref(tk), C.byref(rs))
        out = (score, ncig, cig.reshape(n, slot) if cigar else None)
        if raise_on_error:
            self._check(rc)
            return out
        return out, rc

    def global_device(self, dev, scoring: Scoring = None):
        sc = scoring or Scoring.make()
        tk = _GlbTasks(dev["n"], *[dev.get(k) for k in
                                   ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")])
        rs = _GlbResults(dev.get("score"), dev.get("n_cigar"), dev.get("cigar"), dev.get("cigar_slot", 0))
        self._check(self.L.bsx_global_batch_device(self.ctx, C.byref(sc), C.byref(tk), C.byref(rs)))


    # ---- NM / MD:Z / XB:Z (bwa_gen_cigar2's tail, bwa.c:250-343) ----
    def md_device(self, dev):
        """Device-memory form: `dev` maps the bsx_md_tasks / bsx_md_results field
        names to device addresses (and n, l_pac, crick_l and the slot sizes to ints)."""
        tk = _MdTasks(dev["n"], dev["l_pac"], dev["crick_l"], *[dev[k] for k in
                      ("opac", "rb", "re", "oquery", "oq_off", "l_query", "n_cigar", "cigar")], dev["cigar_slot"])
        rs = _MdResults(dev["md"], dev["md_slot"], dev["md_len"], dev["xb"], dev["xb_slot"], dev["xb_len"],
                        dev["nm"], dev["meth"], dev["status"])
        self._check(self.L.bsx_md_batch_device(self.ctx, C.byref(tk), C.byref(rs)))

    def get_tuning(self):
        """(table [tpb_plain, tpb_narrow, shape_var_65_96, shape_var_97_128],
        source "default" | "cache" | "caller", device key)."""
        t = np.zeros(TUNE_FIELDS, np.int32)
        key = C.create_string_buffer(64)
        src = self.L.bsx_get_tuning(self.ctx, t.ctypes.data, TUNE_FIELDS, key, 64)
        if src < 0:
            self._check(src)
        return [int(x) for x in t], ("default", "cache", "caller")[src], key.value.decode()

    def set_tuning(self, table=None):
        """Sets the tuning table (None: the defaults)."""
        if table is None:
            self._check(self.L.bsx_set_tuning(self.ctx, None, 0))
        else:
            t = np.ascontiguousarray(table, np.int32)
            self._check(self.L.bsx_set_tuning(self.ctx, t.ctypes.data, len(t)))

    def reg2aln_cigar(self, l_pac, rb, re, qb, qe, l_seq, n_cigar, cigar):
        """SAM CIGAR and position of regions from their bwa_gen_cigar2 CIGARs
        (mem_reg2aln's deletion squeeze and clipping): host arrays in, host
        arrays out (cigar [n, slot+2], n_cigar, pos, is_rev)."""
        return _reg2aln_cigar(self, l_pac, rb, re, qb, qe, l_seq, n_cigar, cigar)

    def md(self, opac, l_pac, crick_l, rb, re, oquery, oq_off, l_query, n_cigar, cigar, md_slot=None, xb_slot=None):
        """Host arrays in, host arrays out: NM, MD:Z and XB:Z of each task's
        CIGAR (cigar is [n, slot] uint32 as global_ returns it).  Returns a dict
        with md / xb (lists of str), nm, meth [n, 4], status."""
        n = len(rb)
        if md_slot is None:
            md_slot = int(4 * (np.max(l_query) if n else 1) + 4 * cigar.shape[1] + 16)
        if xb_slot is None:
            xb_slot = int(4 * (np.max(l_query) if n else 1) + 16)
        host = {"opac": np.ascontiguousarray(opac, np.uint8), "rb": np.ascontiguousarray(rb, np.int64),
                "re": np.ascontiguousarray(re, np.int64), "oquery": np.ascontiguousarray(oquery, np.uint8),
                "oq_off": np.ascontiguousarray(oq_off, np.uint64), "l_query": np.ascontiguousarray(l_query, np.int32),
                "n_cigar": np.ascontiguousarray(n_cigar, np.int32),
                "cigar": np.ascontiguousarray(cigar, np.uint32).reshape(n, -1)}
        keep = {k: self.to_device(v) for k, v in host.items()}
        out = {"md": self.empty((n, md_slot), np.uint8), "md_len": self.empty((n,), np.int32),
               "xb": self.empty((n, xb_slot), np.uint8), "xb_len": self.empty((n,), np.int32),
               "nm": self.empty((n,), np.int32), "meth": self.empty((n, 4), np.int32),
               "status": self.empty((n,), np.int32)}
        dev = {k: v.ptr for k, v in {**keep, **out}.items()}
        dev.update(n=n, l_pac=int(l_pac), crick_l=int(crick_l), cigar_slot=int(host["cigar"].shape[1]),
                   md_slot=md_slot, xb_slot=xb_slot)
        self.md_device(dev)
        self.sync()
        h = {k: v.numpy() for k, v in out.items()}
        return {"md": [bytes(h["md"][i, :h["md_len"][i]]).decode() for i in range(n)],
                "xb": [bytes(h["xb"][i, :h["xb_len"][i]]).decode() for i in range(n)],
                "nm": h["nm"], "meth": h["meth"], "status": h["status"]}


def _reg2aln_cigar(eng, l_pac, rb, re, qb, qe, l_seq, n_cigar, cigar):
    n = len(rb)
    cigar = np.ascontiguousarray(cigar, np.uint32).reshape(n, -1)
    host = {"rb": np.ascontiguousarray(rb, np.int64), "re": np.ascontiguousarray(re, np.int64),
            "qb": np.ascontiguousarray(qb, np.int32), "qe": np.ascontiguousarray(qe, np.int32),
            "l_seq": np.ascontiguousarray(l_seq, np.int32), "n_cigar": np.ascontiguousarray(n_cigar, np.int32),
            "cigar": cigar}
    keep = {k: eng.to_device(v) for k, v in host.items()}
    slot = cigar.shape[1] + 2
    out = {"cigar": eng.empty((n, slot), np.uint32), "n_cigar": eng.empty((n,), np.int32),
           "pos": eng.empty((n,), np.int64), "is_rev": eng.empty((n,), np.int32)}
    tk = _Reg2alnTasks(n, int(l_pac), *[keep[k].ptr for k in ("rb", "re", "qb", "qe", "l_seq", "n_cigar", "cigar")],
                       cigar.shape[1])
    rs = _Reg2alnResults(out["cigar"].ptr, slot, out["n_cigar"].ptr, out["pos"].ptr, out["is_rev"].ptr)
    eng._check(eng.L.bsx_reg2aln_cigar_device(eng.ctx, C.byref(tk), C.byref(rs)))
    eng.sync()
    h = {k: v.numpy() for k, v in out.items()}
    return h["cigar"], h["n_cigar"], h["pos"], h["is_rev"]


def _glb_subset(t, a, b):
    q0, t0 = int(t.q_off[a]), int(t.t_off[a])
    q1 = int(t.q_off[b - 1]) + int(t.qlen[b - 1])
    t1 = int(t.t_off[b - 1]) + int(t.tlen[b - 1])
    return GlbTasks(t.query[q0:q1], t.q_off[a:b] - q0, t.qlen[a:b], t.target[t0:t1], t.t_off[a:b] - t0,
                    t.tlen[a:b], t.w[a:b])


class Graph:
    """A captured batch call; launch() replays it on the context's stream."""

    def __init__(self, eng, handle):
        self.eng, self.h = eng, handle

    def launch(self):
        self.eng._check(self.eng.L.bsx_graph_launch(self.eng.ctx, self.h))

    def close(self):
        if self.h:
            self.eng.L.bsx_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DevArray:
    """Device memory owned through a libbsx context (no other HIP runtime is
    needed in the process)."""

    def __init__(self, eng, nbytes, dtype, shape):
        self.eng, self.nbytes, self.dtype, self.shape = eng, nbytes, np.dtype(dtype), shape
        self.ptr = eng.L.bsx_dev_alloc(eng.ctx, max(1, nbytes))
        if not self.ptr:
            raise BsxError(BSX_ENOMEM, eng.last_error())

    def upload(self, arr):
        """Copies a host array into the start of this buffer (synchronous)."""
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        if arr.nbytes:
            self.eng._check(self.eng.L.bsx_memcpy(self.eng.ctx, self.ptr, arr.ctypes.data, arr.nbytes, 1))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            self.eng._check(self.eng.L.bsx_memcpy(self.eng.ctx, out.ctypes.data, self.ptr, self.nbytes, 2))
        return out

    def free(self):
        if self.ptr and self.eng.ctx:
            self.eng.L.bsx_dev_free(self.eng.ctx, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_default = None


def _engine():
    global _default
    if _default is None:
        _default = Engine(0)
    return _default


def ksw_extend2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0):
    """One ksw_extend2 call on the GPU (argument meaning as ksw.h:84-108).
    Returns (score, qle, tle, gtle, gscore, max_off).  Where the reference
    asserts (h0 <= 0, ksw.c:385) this raises BsxError(BSX_EINVAL)."""
    if m != 5:
        raise BsxError(BSX_EINVAL, "m must be 5")
    q = np.ascontiguousarray(np.asarray(query, np.uint8)[:qlen])
    t = np.ascontiguousarray(np.asarray(target, np.uint8)[:tlen])
    tk = ExtTasks.from_pairs([(q, t)], h0, w, end_bonus)
    sc = Scoring.make(mat=mat, o_del=o_del, e_del=e_del, o_ins=o_ins, e_ins=e_ins, zdrop=zdrop)
    r = _engine().extend(tk, sc, cells=False)
    return tuple(int(r[k][0]) for k in RESULT_FIELDS)


def ksw_global2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w):
    """One ksw_global2 call on the GPU.  Returns (score, n_cigar, [cigar uint32])."""
    if m != 5:
        raise BsxError(BSX_EINVAL, "m must be 5")
    q = np.asarray(query, np.uint8)[:qlen]
    t = np.asarray(target, np.uint8)[:tlen]
    tk = GlbTasks.from_pairs([(q, t)], w)
    sc = Scoring.make(mat=mat, o_del=o_del, e_del=e_del, o_ins=o_ins, e_ins=e_ins)
    s, nc, cg = _engine().global_(tk, sc, cigar=True)
    n = int(nc[0])
    return int(s[0]), n, [int(x) for x in cg[0, :n]]
