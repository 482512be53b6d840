"""Task batches in the structure-of-arrays layout of include/bsx.h.

An extension task is the argument list of one ksw_extend2 call
(bsbolt/External/BWA/ksw.h:108): query codes, target codes, h0, band w and
end_bonus.  A batch concatenates the sequences and keeps per-task offsets and
lengths, so it can be handed to the HIP kernels, to the CPU oracle or written
to disk unchanged.
"""
from dataclasses import dataclass, field

import numpy as np

EXT_FIELDS = ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w", "end_bonus")
EXT_DTYPES = dict(query=np.uint8, q_off=np.uint64, qlen=np.int32, target=np.uint8, t_off=np.uint64,
                  tlen=np.int32, h0=np.int32, w=np.int32, end_bonus=np.int32)
GLB_FIELDS = ("query", "q_off", "qlen", "target", "t_off", "tlen", "w")
RESULT_FIELDS = ("score", "qle", "tle", "gtle", "gscore", "max_off")


@dataclass
class ExtTasks:
    """A batch of ksw_extend2 calls."""
    query: np.ndarray
    q_off: np.ndarray
    qlen: np.ndarray
    target: np.ndarray
    t_off: np.ndarray
    tlen: np.ndarray
    h0: np.ndarray
    w: np.ndarray
    end_bonus: np.ndarray
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        for k in EXT_FIELDS:
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=EXT_DTYPES[k]))

    def __len__(self):
        return len(self.qlen)

    @classmethod
    def from_pairs(cls, pairs, h0, w=100, end_bonus=30):
        """pairs: list of (query codes, target codes); h0/w/end_bonus scalars or sequences."""
        n = len(pairs)
        ql = np.array([len(q) for q, _ in pairs], np.int32)
        tl = np.array([len(t) for _, t in pairs], np.int32)
        qo = np.zeros(n, np.uint64)
        to = np.zeros(n, np.uint64)
        if n:
            qo[1:] = np.cumsum(ql[:-1], dtype=np.uint64)
            to[1:] = np.cumsum(tl[:-1], dtype=np.uint64)
        q = np.concatenate([np.asarray(p[0], np.uint8) for p in pairs]) if n else np.zeros(0, np.uint8)
        t = np.concatenate([np.asarray(p[1], np.uint8) for p in pairs]) if n else np.zeros(0, np.uint8)
        b = lambda v: np.broadcast_to(np.asarray(v, np.int32), (n,)).copy()
        return cls(q, qo, ql, t, to, tl, b(h0), b(w), b(end_bonus))

    def subset(self, idx):
        """A compacted batch holding tasks idx (in that order)."""
        idx = np.asarray(idx, np.int64)
        pairs = [(self.query_of(i), self.target_of(i)) for i in idx]
        return ExtTasks.from_pairs(pairs, self.h0[idx], self.w[idx], self.end_bonus[idx])

    def query_of(self, i):
        o = int(self.q_off[i])
        return self.query[o:o + int(self.qlen[i])]

    def target_of(self, i):
        o = int(self.t_off[i])
        return self.target[o:o + int(self.tlen[i])]

    def arrays(self):
        return {k: getattr(self, k) for k in EXT_FIELDS}

    def save(self, path, **extra):
        np.savez_compressed(path, **self.arrays(), **extra)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        t = cls(*[z[k] for k in EXT_FIELDS])
        t.meta = {k: z[k] for k in z.files if k not in EXT_FIELDS}
        return t


@dataclass
class GlbTasks:
    """A batch of ksw_global2 calls."""
    query: np.ndarray
    q_off: np.ndarray
    qlen: np.ndarray
    target: np.ndarray
    t_off: np.ndarray
    tlen: np.ndarray
    w: np.ndarray

    def __post_init__(self):
        for k in GLB_FIELDS:
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=EXT_DTYPES[k]))

    def __len__(self):
        return len(self.qlen)

    @classmethod
    def from_pairs(cls, pairs, w):
        e = ExtTasks.from_pairs(pairs, 1, w, 0)
        return cls(e.query, e.q_off, e.qlen, e.target, e.t_off, e.tlen, e.w)


def random_ext_tasks(n, seed=0, qmax=150, h0max=150, w_choices=(100,), eb_choices=(30,),
                     sub=0.05, indel=0.3, n_frac=0.0, tail_max=None):
    """Seeded synthetic extension tasks shaped like seed flanks: the target is
    the query with substitutions / short indels, followed by unrelated bases
    (the window beyond the alignment, cal_max_gap-sized, bwamem.c:625-632)."""
    rng = np.random.default_rng(seed)
    pairs, h0s, ws, ebs = [], [], [], []
    for _ in range(n):
        ql = int(rng.integers(1, qmax + 1))
        q = rng.integers(0, 4, ql).astype(np.uint8)
        if n_frac > 0:
            q[rng.random(ql) < n_frac] = 4
        t = list(np.where(q == 4, rng.integers(0, 4, ql), q).astype(np.uint8))
        for k in range(len(t)):
            if rng.random() < sub:
                t[k] = int(rng.integers(0, 4))
        if rng.random() < indel:
            p = int(rng.integers(0, len(t) + 1))
            L = int(rng.integers(1, 6))
            if rng.random() < 0.5:
                t[p:p] = list(rng.integers(0, 4, L))
            else:
                del t[p:p + L]
        extra = max(1, ql - 5) if tail_max is None else int(rng.integers(0, tail_max + 1))
        t = np.array(t + list(rng.integers(0, 4, extra)), np.uint8)
        pairs.append((q, t))
        h0s.append(int(rng.integers(1, h0max + 1)))
        ws.append(int(rng.choice(w_choices)))
        ebs.append(int(rng.choice(eb_choices)))
    return ExtTasks.from_pairs(pairs, h0s, ws, ebs)
