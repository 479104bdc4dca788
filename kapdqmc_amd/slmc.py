"""Self-learning Monte Carlo for the t-U model (no counterpart in the reference).

``EffectiveModel`` wraps the library's convolutional effective model of the HS field
(kdq_net_*): E(s) of a batch of field arrays, the loss Var_b(E + log W) and its gradient, Adam
steps, and the inner chain on exp(-E) alone. ``DQMCBatch.slmc_move`` (kdq_slmc_move) proposes
a new field for every chain by running that inner chain from the current field and accepts it
with min(1, W(s')/W(s) e^{E(s') - E(s)}) using the exact FP64 weights, so the chains keep
sampling W exactly whatever the network has learned.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _stack(fields, ns, nt):
    """[B] arrays (ns, nt) -> contiguous uint8 [B, nt, ns] (each s[i + t ns], as the library reads it)."""
    arr = np.stack([np.asarray(f, dtype=np.uint8).reshape(ns, nt, order="F").T for f in fields])
    return np.ascontiguousarray(arr)


class EffectiveModel:
    """Translation-equivariant conv net E(s) on an L1 x L2 honeycomb lattice with nt slices.

    channels: hidden channels C; max_batch: the most configurations (or chains) one call takes.
    Parameters are FP32 masters, laid out as [W1 (C, 8, 27), b1 (C), W2 (C, C, 27), b2 (C),
    W3 (C, 27)]; the GPU kernels evaluate the network with W1, W2, W3 rounded to BF16 (nearest
    even) and FP32 biases, activations and accumulation.
    """

    def __init__(self, L1: int, L2: int, nt: int, channels: int = 8, seed: int = 0, max_batch: int = 64):
        self.L = L.load()
        h = C.c_void_p()
        L.check_net(self.L.kdq_net_create(L1, L2, nt, channels, seed, max_batch, C.byref(h)))
        self.h = h
        self.L1, self.L2, self.nt, self.channels, self.max_batch = L1, L2, nt, channels, max_batch
        self.ns = 2 * L1 * L2

    @classmethod
    def for_model(cls, model, **kw):
        """Network for a t-U model struct's lattice (``model.site``) and slice count."""
        return cls(model.site[0], model.site[1], model.Nt, **kw)

    def close(self):
        if getattr(self, "h", None):
            self.L.kdq_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        L.check_net(rc, self.h)

    @property
    def num_params(self) -> int:
        return int(self.L.kdq_net_num_params(self.h))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, dtype=np.float32)
        self._chk(self.L.kdq_net_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {a.size}")
        self._chk(self.L.kdq_net_set_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float))))

    def _fields(self, fields):
        fields = list(fields)
        if not 1 <= len(fields) <= self.max_batch:
            raise ValueError(f"batch of {len(fields)} configurations; the network takes 1..{self.max_batch}")
        return _stack(fields, self.ns, self.nt)

    def energy(self, fields) -> np.ndarray:
        """E(s) of each field array (shape (ns, nt), values 1..4)."""
        buf = self._fields(fields)
        E = np.zeros(len(buf))
        self._chk(self.L.kdq_net_energy(self.h, buf.ctypes.data_as(C.POINTER(C.c_uint8)), len(buf),
                                        E.ctypes.data_as(C.POINTER(C.c_double))))
        return E

    def loss_grad(self, fields, logw):
        """(Var_b(E_b + logw_b), gradient as FP32 [num_params])."""
        buf = self._fields(fields)
        lw = np.ascontiguousarray(logw, dtype=np.float64)
        loss = C.c_double()
        g = np.zeros(self.num_params, dtype=np.float32)
        self._chk(self.L.kdq_net_loss_grad(self.h, buf.ctypes.data_as(C.POINTER(C.c_uint8)),
                                           lw.ctypes.data_as(C.POINTER(C.c_double)), len(buf), C.byref(loss),
                                           g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def train_step(self, fields, logw, lr: float = 1e-3) -> float:
        """One Adam step on Var_b(E_b + logw_b); returns the loss before the step."""
        buf = self._fields(fields)
        lw = np.ascontiguousarray(logw, dtype=np.float64)
        loss = C.c_double()
        self._chk(self.L.kdq_net_train_step(self.h, buf.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            lw.ctypes.data_as(C.POINTER(C.c_double)), len(buf), float(lr),
                                            C.byref(loss)))
        return loss.value

    def sample(self, fields, nsteps: int, seed: int = 0):
        """Run the inner chain on exp(-E) from each configuration; returns (final fields,
        their E, accepted single-field changes per configuration)."""
        buf = self._fields(fields)
        E = np.zeros(len(buf))
        acc = np.zeros(len(buf), dtype=np.int32)
        self._chk(self.L.kdq_net_sample(self.h, buf.ctypes.data_as(C.POINTER(C.c_uint8)), len(buf), int(nsteps),
                                        int(seed), E.ctypes.data_as(C.POINTER(C.c_double)),
                                        acc.ctypes.data_as(C.POINTER(C.c_int32))))
        out = [np.asfortranarray(b.T) for b in buf]
        return out, E, acc


def train(batch, net: EffectiveModel, rounds: int, sweeps_per_round: int = 1, lr: float = 1e-3,
          epochs: int = 20, history: int = 0):
    """Train ``net`` on the (field, log W) pairs a DQMCBatch produces: each round runs
    ``sweeps_per_round`` local sweeps of every chain, adds the chains' fields and exact log W
    to a buffer (the last ``history`` rounds, all if 0), and takes ``epochs`` Adam steps on
    minibatches of it. Returns the loss of each round's last step."""
    buf_f, buf_w, losses = [], [], []
    rng = np.random.default_rng(0)
    for _ in range(rounds):
        batch.phy_update(sweeps_per_round, False)
        lw = batch.log_weight()
        for c in range(batch.n):
            buf_f.append(batch.get_fields(c))
            buf_w.append(lw[c])
        if history:
            keep = history * batch.n
            buf_f, buf_w = buf_f[-keep:], buf_w[-keep:]
        loss = float("nan")
        for _ in range(epochs):
            m = min(net.max_batch, len(buf_f))
            idx = rng.choice(len(buf_f), size=m, replace=False)
            loss = net.train_step([buf_f[i] for i in idx], np.array([buf_w[i] for i in idx]), lr)
        losses.append(loss)
    return losses
