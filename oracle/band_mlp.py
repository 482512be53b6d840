"""CPU oracle of the learned-band predictor (test infrastructure only: imported
by tests/ and the smoke check as the checker, never by the product path).

  features(batch)        the 24 task features (oracle/band_oracle.c), stored as
                         BF16 bit patterns [n, 32] like the GPU
  forward(params, X)     the MLP 24 -> 64 -> 64 -> 5 with the GPU's mixed
                         precision: BF16 inputs/weights/hidden activations,
                         FP32 accumulation and biases
  grad(params, X, y)     mean softmax cross-entropy and its gradient, with the
                         same BF16 rounding of the matrix operands
  grad64(params, X, y)   the same loss in float64 without rounding (the form the
                         finite-difference checks pin)
  adam(...)              one Adam step (betas 0.9/0.999, eps 1e-8)
  labels(batch)          smallest frame class whose band-following run is
                         accepted (oracle/band_oracle.c)

Parameter vector: W1[64][24], b1[64], W2[64][64], b2[64], W3[5][64], b3[5].
"""
import ctypes as C

import numpy as np

from . import oracle as _O

NF, NH, NC = 24, 64, 5
STRIDE = 32
SHAPES = (("W1", (NH, NF)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NC, NH)), ("b3", (NC,)))
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
assert NPARAMS == 6085
SCAN = 16
QMAX = 256
FRAMES = (24, 40, 80, 144)  # columns of the frame classes 0..3


def unpack(p):
    out, o = {}, 0
    for k, s in SHAPES:
        n = int(np.prod(s))
        out[k] = np.asarray(p[o:o + n]).reshape(s)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[k], np.float64).reshape(-1) for k, _ in SHAPES])


def init_params(seed=0):
    """He-style initialisation (uniform), zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for k, s in SHAPES:
        if k.startswith("b"):
            d[k] = np.zeros(s)
        else:
            lim = np.sqrt(6.0 / s[1])
            d[k] = rng.uniform(-lim, lim, s)
    return pack(d).astype(np.float32)


def bf16_bits(x):
    """float32 -> BF16 bit pattern, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_val(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(x):
    return bf16_val(bf16_bits(x))


def _lib():
    L = _O.lib()
    if not getattr(L, "_band_ready", False):
        P, I, I64 = C.c_void_p, C.c_int, C.c_int64
        L.orc_narrow_extend.restype = I
        L.orc_narrow_extend.argtypes = [I, P, I, P, P] + [I] * 9 + [P, P, P]
        L.orc_band_labels.restype = None
        L.orc_band_labels.argtypes = [I64, P, I, I, I, I, I] + [P] * 9 + [I, P]
        L.orc_band_features_batch.restype = None
        L.orc_band_features_batch.argtypes = [I64, P, I, I, I, I] + [P] * 9 + [I, P]
        L._band_ready = True
    return L


def _arrs(b):
    return [np.ascontiguousarray(getattr(b, k)) for k in
            ("query", "q_off", "qlen", "target", "t_off", "tlen", "h0", "w", "end_bonus")]


def _sc(scoring):
    sc = scoring or {}
    mat = np.ascontiguousarray(sc.get("mat", _O.fill_scmat()), np.int8)
    return mat, sc.get("o_del", 6), sc.get("e_del", 1), sc.get("o_ins", 6), sc.get("e_ins", 1), sc.get("zdrop", 100)


def features_f32(b, scoring=None):
    mat, od, ed, oi, ei, _ = _sc(scoring)
    n = len(b.qlen)
    f = np.zeros((n, NF), np.float32)
    arrs = _arrs(b)
    _lib().orc_band_features_batch(n, mat.ctypes.data, od, ed, oi, ei, *[a.ctypes.data for a in arrs], SCAN,
                                   f.ctypes.data)
    return f


def features(b, scoring=None):
    """BF16 bit patterns [n, 32] (columns 24..31 zero), as the GPU stores them."""
    f = features_f32(b, scoring)
    out = np.zeros((len(f), STRIDE), np.uint16)
    out[:, :NF] = bf16_bits(f)
    return out


def labels(b, scoring=None):
    mat, od, ed, oi, ei, zd = _sc(scoring)
    n = len(b.qlen)
    lab = np.zeros(n, np.uint8)
    arrs = _arrs(b)
    _lib().orc_band_labels(n, mat.ctypes.data, od, ed, oi, ei, zd, *[a.ctypes.data for a in arrs], QMAX,
                           lab.ctypes.data)
    return lab


def narrow_extend(query, target, h0, wf, w=100, end_bonus=30, zdrop=100, mat=None, o_del=6, e_del=1, o_ins=6,
                  e_ins=1):
    """(accepted, outputs(6), cells, why) of one run in a band-following frame of wf columns."""
    q = np.ascontiguousarray(query, np.uint8)
    t = np.ascontiguousarray(target, np.uint8)
    mat = _O.fill_scmat() if mat is None else np.ascontiguousarray(mat, np.int8)
    out = np.zeros(6, np.int32)
    cl = C.c_int64(0)
    why = C.c_int(0)
    r = _lib().orc_narrow_extend(len(q), q.ctypes.data, len(t), t.ctypes.data, mat.ctypes.data, o_del, e_del, o_ins,
                                 e_ins, w, end_bonus, zdrop, h0, wf, out.ctypes.data, C.byref(cl), C.byref(why))
    return r, tuple(int(x) for x in out), int(cl.value), int(why.value)


def _xval(X):
    X = np.asarray(X)
    if X.dtype == np.uint16:
        X = bf16_val(X)[:, :NF]
    return np.asarray(X, np.float32)[:, :NF]


def forward(params, X, return_hidden=False):
    """Logits [n, 5] with BF16 operands and FP32 accumulation."""
    d = unpack(np.asarray(params, np.float32))
    x = bf16(_xval(X))
    W1, W2, W3 = (bf16(d[k].astype(np.float32)) for k in ("W1", "W2", "W3"))
    z1 = x.astype(np.float64) @ W1.T.astype(np.float64) + d["b1"]
    h1 = np.maximum(z1, 0).astype(np.float32)
    h1b = bf16(h1)
    z2 = h1b.astype(np.float64) @ W2.T.astype(np.float64) + d["b2"]
    h2 = np.maximum(z2, 0).astype(np.float32)
    h2b = bf16(h2)
    lg = (h2b.astype(np.float64) @ W3.T.astype(np.float64) + d["b3"]).astype(np.float32)
    if return_hidden:
        return lg, dict(x=x, h1=h1, h1b=h1b, h2=h2, h2b=h2b, W1=W1, W2=W2, W3=W3)
    return lg


def predict(params, X):
    return np.argmax(forward(params, X), axis=1).astype(np.uint8)


def grad(params, X, y):
    """(mean gradient [6085] float64, mean loss) with the GPU's BF16 rounding of
    every matrix operand (activations, deltas, weights)."""
    y = np.asarray(y, np.int64)
    n = len(y)
    lg, h = forward(params, X, return_hidden=True)
    z = lg.astype(np.float64)
    z = z - z.max(axis=1, keepdims=True)
    e = np.exp(z)
    p = e / e.sum(axis=1, keepdims=True)
    loss = float(-np.log(np.maximum(p[np.arange(n), y], 1e-30)).mean())
    dl = p.copy()
    dl[np.arange(n), y] -= 1.0
    dlb = bf16(dl.astype(np.float32)).astype(np.float64)
    dh2 = dlb @ h["W3"].astype(np.float64)
    dz2 = np.where(h["h2"] > 0, dh2, 0.0)
    dz2b = bf16(dz2.astype(np.float32)).astype(np.float64)
    dh1 = dz2b @ h["W2"].astype(np.float64)
    dz1 = np.where(h["h1"] > 0, dh1, 0.0)
    dz1b = bf16(dz1.astype(np.float32)).astype(np.float64)
    g = dict(W3=dlb.T @ h["h2b"].astype(np.float64), b3=dl.sum(0),
             W2=dz2b.T @ h["h1b"].astype(np.float64), b2=dz2.sum(0),
             W1=dz1b.T @ h["x"].astype(np.float64), b1=dz1.sum(0))
    return pack(g) / n, loss


def loss64(params, X, y):
    d = unpack(np.asarray(params, np.float64))
    x = _xval(X).astype(np.float64)
    h1 = np.maximum(x @ d["W1"].T + d["b1"], 0)
    h2 = np.maximum(h1 @ d["W2"].T + d["b2"], 0)
    z = h2 @ d["W3"].T + d["b3"]
    z = z - z.max(axis=1, keepdims=True)
    lp = z - np.log(np.exp(z).sum(axis=1, keepdims=True))
    return float(-lp[np.arange(len(y)), np.asarray(y, np.int64)].mean())


def grad64(params, X, y):
    """Exact float64 backward pass of loss64 (pinned by finite differences)."""
    y = np.asarray(y, np.int64)
    n = len(y)
    d = unpack(np.asarray(params, np.float64))
    x = _xval(X).astype(np.float64)
    z1 = x @ d["W1"].T + d["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ d["W2"].T + d["b2"]
    h2 = np.maximum(z2, 0)
    z = h2 @ d["W3"].T + d["b3"]
    z = z - z.max(axis=1, keepdims=True)
    p = np.exp(z) / np.exp(z).sum(axis=1, keepdims=True)
    dl = p.copy()
    dl[np.arange(n), y] -= 1.0
    dz2 = np.where(z2 > 0, dl @ d["W3"], 0.0)
    dz1 = np.where(z1 > 0, dz2 @ d["W2"], 0.0)
    g = dict(W3=dl.T @ h2, b3=dl.sum(0), W2=dz2.T @ h1, b2=dz2.sum(0), W1=dz1.T @ x, b1=dz1.sum(0))
    return pack(g) / n


def adam(params, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in FP32; returns (params, m, v)."""
    p = np.asarray(params, np.float32)
    g = np.asarray(g, np.float32)
    m = (np.float32(b1) * m + np.float32(1 - b1) * g).astype(np.float32)
    v = (np.float32(b2) * v + np.float32(1 - b2) * g * g).astype(np.float32)
    c1 = np.float32(1 - b1 ** step)
    c2 = np.float32(1 - b2 ** step)
    p = (p - np.float32(lr) * (m / c1) / (np.sqrt(v / c2) + np.float32(eps))).astype(np.float32)
    return p, m, v
