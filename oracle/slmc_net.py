"""ORACLE — test infrastructure only.

FP64 numpy restatement of the self-learning effective model the library evaluates on the GPU
(no counterpart in the reference, which has no SLMC). Only tests/ may import this module.

Field tensor: site i of the L1 x L2 honeycomb lattice is unit cell i // 2 = x + L1 y and
sublattice i % 2; channel 4 * sublattice + (s - 1) is the one-hot value of s at (slice t, y, x).
Network: h1 = tanh(conv(X, W1) + b1), h2 = tanh(conv(h1, W2) + b2), o = conv(h2, W3),
E = Σ o, with 3 x 3 x 3 kernels (tap k = 9 (dt+1) + 3 (dy+1) + (dx+1)) that are periodic in
y and x and zero-padded in t. Loss = Var_b(E_b + log W_b).

Pins: the backward pass is checked against central finite differences of the forward pass
(tests/test_slmc_oracle.py); the forward pass against a direct triple loop on a tiny lattice
and against its translation equivariance (E is invariant under lattice translations).
"""
from __future__ import annotations

import numpy as np

CIN = 8


class Geometry:
    def __init__(self, L1: int, L2: int, nt: int):
        self.X, self.Y, self.T = L1, L2, nt
        self.P = nt * L1 * L2
        self.n = 2 * L1 * L2
        q = np.arange(self.P)
        x, y, t = q % L1, (q // L1) % L2, q // (L1 * L2)
        self.pos = np.zeros((27, self.P), dtype=np.int64)
        self.valid = np.zeros((27, self.P), dtype=bool)
        for k in range(27):
            dt, dy, dx = k // 9 - 1, (k // 3) % 3 - 1, k % 3 - 1
            tt = t + dt
            self.valid[k] = (tt >= 0) & (tt < nt)
            self.pos[k] = (np.clip(tt, 0, nt - 1) * L2 + (y + dy) % L2) * L1 + (x + dx) % L1


def param_layout(C: int):
    """Offsets of W1 [C,8,27], b1 [C], W2 [C,C,27], b2 [C], W3 [1,C,27] in the flat vector."""
    sizes = [("W1", C * CIN * 27), ("b1", C), ("W2", C * C * 27), ("b2", C), ("W3", C * 27)]
    out, off = {}, 0
    for name, n in sizes:
        out[name] = (off, n)
        off += n
    return out, off


def unpack(p, C: int):
    lay, _ = param_layout(C)
    g = lambda k: p[lay[k][0]:lay[k][0] + lay[k][1]]
    return (g("W1").reshape(C, CIN, 27), g("b1"), g("W2").reshape(C, C, 27), g("b2"),
            g("W3").reshape(1, C, 27))


def bf16_weights(p, C: int):
    """The parameters the GPU kernels evaluate: W1, W2, W3 rounded to BF16 (round to nearest
    even of their FP32 values), biases as FP32; returned as float64."""
    q = np.asarray(p, dtype=np.float32).copy()
    lay, _ = param_layout(C)
    for name in ("W1", "W2", "W3"):
        off, n = lay[name]
        u = q[off:off + n].view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        q[off:off + n] = u.astype(np.uint32).view(np.float32)
    return q.astype(np.float64)


def embed(geo: Geometry, fields):
    """fields: [B, ns, nt] values 1..4 -> X [B, 8, P]."""
    fields = np.asarray(fields)
    B = fields.shape[0]
    X = np.zeros((B, CIN, geo.P))
    for i in range(geo.n):
        sub, cell = i % 2, i // 2
        for t in range(geo.T):
            q = t * geo.X * geo.Y + cell
            X[np.arange(B), sub * 4 + fields[:, i, t].astype(np.int64) - 1, q] = 1.0
    return X


def _gather(geo, a):
    """a [B, c, P] -> [B, c, 27, P] with zero padding in t."""
    return a[:, :, geo.pos] * geo.valid[None, None]


def forward(geo: Geometry, p, C: int, fields):
    W1, b1, W2, b2, W3 = unpack(np.asarray(p, dtype=np.float64), C)
    X = embed(geo, fields)
    G0 = _gather(geo, X)
    h1 = np.tanh(np.einsum("oik,bikq->boq", W1, G0) + b1[None, :, None])
    G1 = _gather(geo, h1)
    h2 = np.tanh(np.einsum("oik,bikq->boq", W2, G1) + b2[None, :, None])
    G2 = _gather(geo, h2)
    o = np.einsum("oik,bikq->boq", W3, G2)[:, 0]
    cache = dict(G0=G0, h1=h1, G1=G1, h2=h2, G2=G2)
    return o.sum(axis=1), cache


def energy(geo, p, C, fields):
    return forward(geo, p, C, fields)[0]


def _scatter(geo, dG):
    """adjoint of _gather: dG [B, c, 27, P] -> [B, c, P]."""
    B, c = dG.shape[:2]
    out = np.zeros((B, c, geo.P))
    for k in range(27):
        v = dG[:, :, k, :] * geo.valid[k][None, None]
        np.add.at(out, (slice(None), slice(None), geo.pos[k]), v)
    return out


def loss_grad(geo: Geometry, p, C: int, fields, logw):
    """loss = mean_b (E_b + logw_b - mean)^2 and its gradient (flat, the layout of param_layout)."""
    p = np.asarray(p, dtype=np.float64)
    W1, b1, W2, b2, W3 = unpack(p, C)
    E, c = forward(geo, p, C, fields)
    B = len(E)
    d = E + np.asarray(logw) - np.mean(E + np.asarray(logw))
    loss = np.mean(d * d)
    gE = 2.0 * d / B
    do = np.broadcast_to(gE[:, None, None], (B, 1, geo.P))
    gW3 = np.einsum("boq,bikq->oik", do, c["G2"])
    dh2 = _scatter(geo, np.einsum("oik,boq->bikq", W3, do))
    dpre2 = dh2 * (1.0 - c["h2"] ** 2)
    gW2 = np.einsum("boq,bikq->oik", dpre2, c["G1"])
    gb2 = dpre2.sum(axis=(0, 2))
    dh1 = _scatter(geo, np.einsum("oik,boq->bikq", W2, dpre2))
    dpre1 = dh1 * (1.0 - c["h1"] ** 2)
    gW1 = np.einsum("boq,bikq->oik", dpre1, c["G0"])
    gb1 = dpre1.sum(axis=(0, 2))
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel()])
    return loss, grad


def energy_loops(geo: Geometry, p, C: int, field):
    """E of one configuration by direct loops (tiny lattices only)."""
    W1, b1, W2, b2, W3 = unpack(np.asarray(p, dtype=np.float64), C)
    X, Y, T = geo.X, geo.Y, geo.T

    def taps(t, y, x):
        for dt in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    tt = t + dt
                    if 0 <= tt < T:
                        yield 9 * (dt + 1) + 3 * (dy + 1) + dx + 1, tt, (y + dy) % Y, (x + dx) % X

    h1 = np.zeros((C, T, Y, X))
    for co in range(C):
        for t in range(T):
            for y in range(Y):
                for x in range(X):
                    s = b1[co]
                    for k, tt, yy, xx in taps(t, y, x):
                        for sub in (0, 1):
                            v = int(field[2 * (xx + X * yy) + sub, tt])
                            s += W1[co, 4 * sub + v - 1, k]
                    h1[co, t, y, x] = np.tanh(s)

    def conv(a, W, bias, act):
        cout, cin = W.shape[:2]
        out = np.zeros((cout, T, Y, X))
        for co in range(cout):
            for t in range(T):
                for y in range(Y):
                    for x in range(X):
                        s = bias[co] if bias is not None else 0.0
                        for k, tt, yy, xx in taps(t, y, x):
                            s += W[co, :, k] @ a[:, tt, yy, xx]
                        out[co, t, y, x] = np.tanh(s) if act else s
        return out

    h2 = conv(h1, W2, b2, True)
    o = conv(h2, W3, None, False)
    return o.sum()
