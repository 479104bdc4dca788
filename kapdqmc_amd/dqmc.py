"""Reference-shaped entry points over the HIP library.

``Initial_s`` and ``phy_update`` keep the signatures and the argument meaning of the
reference's generics (src/KAPDQMC.jl:12-16, src/tU/GreenMatrix.jl:1-8,
src/tU/phy_update.jl:3-152): ``phy_update`` mutates and returns ``s`` and appends one CSV
row per sweep when ``record`` is true. ``DQMCBatch`` is the handle-level API (a batch of
independent chains advancing together on one GPU).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .rng import initial_fields


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


KIND_CODE = {"tU": L.MODEL_TU, "tV": L.MODEL_TV, "tUV": L.MODEL_TUV}
NOBS = {"tU": 6, "tV": 10, "tUV": 7}


def field_shape(m):
    """Shape of the reference's field array: t-U (Ns, Nt), t-V (Ns/2, 3, Nt), t-U-V (Ns z, Nt)."""
    if m.kind == "tU":
        return (m.Ns, m.Nt)
    if m.kind == "tV":
        return (m.Ns // 2, m.nnidx.shape[1], m.Nt)
    return (m.Ns * m.nnidx.shape[1], m.Nt)


def elements_per_slice(m) -> int:
    """Field elements of one slice: t-U Ns, t-V 3 Ns/2, t-U-V Ns z."""
    return int(np.prod(field_shape(m)[:-1]))


def replay_from_visiting_order(m, choice, u):
    """Arranges proposal draws given in the reference's visiting order into the replay layout
    of kdq_set_replay, [sweeps, 2, Nt, n_elem].

    The visiting order of one sweep (one `sampler` draw and one `rand(rng)` per proposal, in
    that order): the forward pass lt = 1..Nt, then the backward pass lt = Nt..1
    (src/tU/phy_update.jl:53,101); within a slice the field elements in order, except t-V,
    whose forward pass visits the bond families 3, 2, 1 and its backward pass 1, 2, 3, each
    over its bonds i = 1..Ns/2 (src/tV/phy_update.jl:64,131). `choice` is the index of the
    sampled value among the other three in increasing order (0, 1, 2)."""
    E, nt = elements_per_slice(m), m.Nt
    c = np.asarray(choice, dtype=np.uint8).ravel()
    uu = np.asarray(u, dtype=np.float64).ravel()
    per = 2 * nt * E
    if c.size != uu.size or c.size % per:
        raise ValueError(f"draws must hold whole sweeps of {per} proposals")
    ns = c.size // per
    order = np.arange(E)
    if m.kind == "tV":
        nb = m.Ns // 2
        fwd = np.concatenate([np.arange(nb) + j * nb for j in (2, 1, 0)])
        bwd = np.concatenate([np.arange(nb) + j * nb for j in (0, 1, 2)])
    else:
        fwd = bwd = order
    oc = np.zeros((ns, 2, nt, E), dtype=np.uint8)
    ou = np.zeros((ns, 2, nt, E))
    vc, vu = c.reshape(ns, 2, nt, E), uu.reshape(ns, 2, nt, E)
    for d, perm in ((0, fwd), (1, bwd)):
        lts = np.arange(nt) if d == 0 else np.arange(nt)[::-1]     # visiting position -> lt - 1
        oc[:, d, lts[:, None], perm[None, :]] = vc[:, d]
        ou[:, d, lts[:, None], perm[None, :]] = vu[:, d]
    return oc, ou


def _descriptors(models, seeds, chains, delay, sweep0, precision=L.PREC_FP64):
    """kdq_model_desc / kdq_run_desc arrays for a list of model structs (plus the numpy
    arrays they point into, which must outlive the call)."""
    n = len(models)
    kind = getattr(models[0], "kind", "tU")
    keep = []
    descs = (L.ModelDesc * n)()
    runs = (L.RunDesc * n)()
    for c, m in enumerate(models):
        if getattr(m, "kind", "tU") != kind:
            raise ValueError("all chains of a batch must share the model kind")
        f = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
        arrs = dict(eta=f(m.η), gamma=f(m.γ), K=f(np.real(m.K)), eK=f(np.real(m.eK)),
                    eKinv=f(np.real(m.eKinv)), halfeK=f(np.real(m.HalfeK)),
                    halfeKinv=f(np.real(m.HalfeKinv)), Pt=f(np.real(m.Pt)))
        if kind != "tUV":
            arrs["alpha"] = f(m.α)
        nodes = np.ascontiguousarray(m.nodes, dtype=np.int64)
        keep += list(arrs.values()) + [nodes]
        d = descs[c]
        d.kind, d.ns, d.nt, d.n_nodes = KIND_CODE[kind], m.Ns, m.Nt, len(nodes)
        if kind != "tU":
            # model.nnidx in column-major (linear) order as 1-based (x, y) pairs
            bonds = np.ascontiguousarray(np.transpose(m.nnidx, (1, 0, 2)).reshape(-1, 2), dtype=np.int32)
            keep.append(bonds)
            d.z = m.nnidx.shape[1]
            d.bonds = bonds.ctypes.data_as(C.POINTER(C.c_int32))
        if kind == "tUV":
            d.a, d.is_complex, d.U, d.V = m.a, int(m.Type is complex), m.U, m.V
        d.nodes = nodes.ctypes.data_as(C.POINTER(C.c_int64))
        d.lattice = L.LATTICE_CODE[m.Lattice]
        d.site[0], d.site[1] = m.site
        d.ht = m.Ht
        for k, v in arrs.items():
            setattr(d, k, _dp(v))
        mats = dict(K=m.K, eK=m.eK, eKinv=m.eKinv, halfeK=m.HalfeK, halfeKinv=m.HalfeKinv, Pt=m.Pt)
        if any(np.iscomplexobj(v) and np.any(np.imag(v) != 0) for v in mats.values()):
            if kind != "tU":
                raise ValueError("complex model matrices are supported for the t-U model only")
            for k, v in mats.items():           # complex hopping (flux != 0): imaginary parts too
                im = f(np.imag(v))
                keep.append(im)
                setattr(d, k + "_im", _dp(im))
        r = runs[c]
        r.seed, r.chain, r.sweep0, r.precision, r.delay = seeds[c], chains[c], sweep0, precision, delay
    return descs, runs, keep


class DQMCBatch:
    """A batch of chains of one model kind (t-U, t-V or t-U-V) on the current GPU.

    models: list of model structs of one kind; seeds: list of int. Chains may differ in Ns,
    couplings and Nt: chains of one Nt share launches, and the Nt groups sweep concurrently
    (per-slice calls — wrap, update_layer, measure — need one Nt).
    delay: local-update block size k in sites (t-U, ≤ 64) or field elements (t-V / t-U-V,
    ≤ 32); 1 = one Sherman-Morrison update per flip.
    """

    def __init__(self, models, seeds, chains=None, delay: int = 32, precision: str = "fp64",
                 sweep0: int = 0):
        self.L = L.load()
        if not isinstance(models, (list, tuple)):
            models = [models]
        n = len(models)
        seeds = list(seeds) if isinstance(seeds, (list, tuple)) else [seeds] * n
        chains = list(range(n)) if chains is None else list(chains)
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.models = list(models)
        self.kind = getattr(models[0], "kind", "tU")
        self.nobs = NOBS[self.kind]
        self.precision = precision
        descs, runs, self._keep = _descriptors(self.models, seeds, chains, delay, sweep0,
                                               L.PREC_MIXED if precision == "mixed" else L.PREC_FP64)
        h = C.c_void_p()
        L.check(self.L.kdq_create(descs, runs, n, C.byref(h)))
        self.h = h
        self.n = n
        self.seeds, self.chains, self.delay = seeds, chains, delay

    def close(self):
        if getattr(self, "h", None):
            self.L.kdq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        L.check(rc, self.h)

    # fields ---------------------------------------------------------------
    def set_fields(self, c: int, s):
        m = self.models[c]
        a = np.asfortranarray(np.asarray(s, dtype=np.uint8))
        if a.shape != field_shape(m):
            raise ValueError(f"s must have shape {field_shape(m)}")
        self._chk(self.L.kdq_set_fields(self.h, c, a.ctypes.data_as(C.POINTER(C.c_uint8))))

    def get_fields(self, c: int) -> np.ndarray:
        m = self.models[c]
        a = np.zeros(field_shape(m), dtype=np.uint8, order="F")
        self._chk(self.L.kdq_get_fields(self.h, c, a.ctypes.data_as(C.POINTER(C.c_uint8))))
        return a

    def initial_s(self):
        self._chk(self.L.kdq_initial_s(self.h))

    # sweeps ---------------------------------------------------------------
    def phy_update(self, sweeps: int, record: bool = True) -> np.ndarray:
        rows = np.zeros((self.n, sweeps, self.nobs))
        self._chk(self.L.kdq_phy_update(self.h, sweeps, int(record), _dp(rows)))
        return rows

    # fine-grained ---------------------------------------------------------
    def get_G(self, c: int) -> np.ndarray:
        n = self.models[c].Ns
        buf = np.zeros(2 * n * n)
        self._chk(self.L.kdq_get_G(self.h, c, _dp(buf)))
        return buf.view(np.complex128).reshape(n, n).T.copy()

    def G4(self, c: int, tau1: int, tau2: int):
        """G(τ1), G(τ2), G(τ1,τ2), G(τ2,τ1) of chain c from its current fields, as
        G4(model, s, τ1, τ2) returns them (src/tU/GreenMatrix.jl:126-226; t-U)."""
        n = self.models[c].Ns
        bufs = [np.zeros(2 * n * n) for _ in range(4)]
        self._chk(self.L.kdq_G4(self.h, c, int(tau1), int(tau2), *[_dp(b) for b in bufs]))
        return [b.view(np.complex128).reshape(n, n).T.copy() for b in bufs]

    def set_G(self, c: int, G):
        n = self.models[c].Ns
        buf = np.ascontiguousarray(np.asarray(G, dtype=np.complex128).T).reshape(-1).view(np.float64).copy()
        self._chk(self.L.kdq_set_G(self.h, c, _dp(buf)))

    def wrap(self, lt: int, forward: bool):
        self._chk(self.L.kdq_wrap(self.h, lt, int(forward)))

    def update_layer(self, lt: int, direction: int, sweep: int):
        self._chk(self.L.kdq_update_layer(self.h, lt, direction, sweep))

    def init_chain(self):
        self._chk(self.L.kdq_init_chain(self.h))

    def measure(self, lt: int) -> np.ndarray:
        out = np.zeros((self.n, self.nobs))
        self._chk(self.L.kdq_measure(self.h, lt, _dp(out)))
        return out

    def set_replay(self, c: int, choice, u):
        """Replayed draws for chain c (kdq_set_replay): arrays [sweeps, 2, Nt, n_elem] (see
        replay_from_visiting_order), used by the next `sweeps` sweeps instead of the Philox
        stream; empty arrays return the chain to Philox."""
        ch = np.ascontiguousarray(choice, dtype=np.uint8).ravel()
        uu = np.ascontiguousarray(u, dtype=np.float64).ravel()
        per = 2 * self.models[c].Nt * elements_per_slice(self.models[c])
        if ch.size != uu.size or ch.size % per:
            raise ValueError("replay arrays must hold whole sweeps")
        self._chk(self.L.kdq_set_replay(self.h, c, ch.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(uu),
                                        ch.size // per))

    def health(self) -> np.ndarray:
        """[chains, 4]: max |G_wrapped - G_recomputed| at the stabilisation nodes since the last
        call, non-finite G, singular overlap, lost rank (kdq_get_health)."""
        out = np.zeros((self.n, 4))
        self._chk(self.L.kdq_get_health(self.h, _dp(out)))
        return out

    def check_G(self) -> np.ndarray:
        """[chains, 2]: max |G² - G| and |tr G - (N - Np)| of the current G (kdq_check_G)."""
        out = np.zeros((self.n, 2))
        self._chk(self.L.kdq_check_G(self.h, _dp(out)))
        return out

    def measure_noalloc(self, lt: int) -> np.ndarray:
        """measure() through kdq_measure_noalloc: the working matrices live in the library's
        per-device static array (at most 32 MiB for the whole batch), nothing is allocated."""
        out = np.zeros((self.n, self.nobs))
        self._chk(self.L.kdq_measure_noalloc(self.h, lt, _dp(out)))
        return out

    def profile(self, families=()):
        """Enable HIP-event timing of the named kernel families (see _lib.PROF_FAMILIES)."""
        mask = 0
        for f in families:
            mask |= 1 << L.PROF_FAMILIES.index(f)
        self._chk(self.L.kdq_profile(self.h, mask))

    def profile_read(self) -> dict:
        """{family: (total device ms, launches)} since the last profile() call."""
        ms = np.zeros(len(L.PROF_FAMILIES))
        n = np.zeros(len(L.PROF_FAMILIES), dtype=np.int64)
        self._chk(self.L.kdq_profile_read(self.h, _dp(ms), n.ctypes.data_as(C.POINTER(C.c_int64))))
        return {f: (float(ms[i]), int(n[i])) for i, f in enumerate(L.PROF_FAMILIES)}

    def counters(self) -> np.ndarray:
        out = np.zeros((self.n, 2))
        self._chk(self.L.kdq_get_counters(self.h, _dp(out)))
        return out

    def near_ties(self) -> np.ndarray:
        """Per chain: decisions so far with |uniform draw − acceptance probability| < 1e-12,
        the only ones a rounding difference between two FP64 implementations may flip."""
        out = np.zeros(self.n)
        self._chk(self.L.kdq_get_near_ties(self.h, _dp(out)))
        return out

    def set_option(self, name: str, value: int):
        """kdq_set_option: "graph" (0/1), "delay" (t-U block size), "gemm_tile" (0, 32, 64)."""
        self._chk(self.L.kdq_set_option(self.h, L.OPTIONS[name], int(value)))
        if name == "delay":
            self.delay = int(value)

    def precision_error(self) -> np.ndarray:
        """max |G_FP32 - G_FP64| per chain over the stabilisations since the last call
        (mixed precision; zeros for FP64 batches)."""
        out = np.zeros(self.n)
        self._chk(self.L.kdq_get_precision_error(self.h, _dp(out)))
        return out

    def log_weight(self) -> np.ndarray:
        """log W(s) = 2 log|det(P† B(2Θ,0) P)| + Σ log γ(s) of every chain's current fields."""
        out = np.zeros(self.n)
        self._chk(self.L.kdq_log_weight(self.h, _dp(out)))
        return out

    def logdet(self) -> np.ndarray:
        """log|det(U_L U_R)| of the overlap factorised by the last get_G! of each chain."""
        out = np.zeros(self.n)
        self._chk(self.L.kdq_get_logdet(self.h, _dp(out)))
        return out

    def slmc_move(self, net, nsteps: int) -> np.ndarray:
        """One self-learning global move per chain (t-U): ``nsteps`` inner-chain steps of
        ``net`` (a kapdqmc_amd.slmc.EffectiveModel) from the current fields, accepted with
        min(1, W(s')/W(s) e^{E(s')-E(s)}). Returns 1.0 / 0.0 per chain."""
        out = np.zeros(self.n)
        self._chk(self.L.kdq_slmc_move(self.h, net.h, int(nsteps), _dp(out)))
        return out


def Initial_s(model, seed: int = 0, chain: int = 0) -> np.ndarray:
    """Initial_s (src/tU/GreenMatrix.jl:1-8, src/tV/GreenMatrix.jl:1-10,
    src/tUV/GreenMatrix.jl:1-8): UInt8 field array of the model's shape, i.i.d. uniform over
    {1..4}, drawn from the library's Philox stream instead of a MersenneTwister."""
    shape = field_shape(model)
    n_elem = int(np.prod(shape[:-1]))
    flat = initial_fields(seed, chain, n_elem, model.Nt)          # [n_elem, Nt]
    return np.asfortranarray(flat.reshape(shape, order="F"))


def _lattice_tag(lattice: str, kind: str = "tU") -> str:
    if kind == "tU":
        tags = {"SQUARE90": "□90", "SQUARE45": "□45", "HoneyComb60": "HC", "HoneyComb120": "HC120"}
    else:
        tags = {"SQUARE": "□", "HoneyComb60": "HC", "HoneyComb120": "HC120"}
    if lattice not in tags:
        raise ValueError(f"Lattice: {lattice} is not allowed !")
    return tags[lattice]


def phy_update(path: str, model, s: np.ndarray, Sweeps: int, record: bool = False, *, seed: int = 0,
               chain: int = 0, delay: int = 32, sweep0: int = 0) -> np.ndarray:
    """phy_update on the GPU (t-U: src/tU/phy_update.jl:3-152, t-V: src/tV/phy_update.jl:1-188,
    t-U-V: src/tUV/phy_update.jl:1-186). Mutates ``s`` in place, returns it, and appends one
    row of observables per sweep to the reference's CSV file name when ``record`` is true."""
    b = DQMCBatch([model], [seed], [chain], delay=delay, sweep0=sweep0)
    try:
        b.set_fields(0, s)
        rows = b.phy_update(Sweeps, record)[0]
        s[...] = b.get_fields(0)
    finally:
        b.close()
    if record:
        kind = getattr(model, "kind", "tU")
        name = _lattice_tag(model.Lattice, kind)
        if kind == "tU" and model.Θquench == 0.0:
            fn = f"tUphy{name}_t{model.Ht}U{model.Hu1}size{model.site}Δt{model.Δt}Θ{model.Θrelax}BS{model.BatchSize}.csv"
        elif kind == "tU":
            fn = (f"tUphy{name}_t{model.Ht}U{model.Hu1}_{model.Hu2}size{model.site}Δt{model.Δt}"
                  f"Θ{model.Θrelax}_{model.Θquench}BS{model.BatchSize}.csv")
        elif kind == "tV" and len(np.unique(model.α)) == 1:
            fn = f"tVphy{name}_t{model.Ht}V{model.Hv1}size{model.site}Δt{model.Δt}Θ{model.Θrelax}BS{model.BatchSize}.csv"
        elif kind == "tV":
            fn = (f"tVphy{name}_t{model.Ht}V{model.Hv1}_{model.Hv2}size{model.site}Δt{model.Δt}"
                  f"Θ{model.Θrelax}_{model.Θquench}BS{model.BatchSize}.csv")
        else:
            fn = f"tUVphy{name}_t{model.Ht}U{model.U}V{model.V}size{model.site}Δt{model.Δt}Θ{model.Θ}BS{model.BatchSize}.csv"
        with open(os.path.join(path, fn), "a") as fh:
            for r in rows:
                fh.write(",".join(repr(float(x)) for x in r) + "\n")
    return s


class SCEEBatch:
    """Swap-replica Rényi-2 chains (ctrl_SCEEicr) of the t-U (src/tU/SCEE.jl:1-262), t-V
    (src/tV/SCEE.jl:1-431) or t-U-V (src/tUV/SCEE.jl:1-339) model.

    indexA / indexB: the reference's 1-based site lists; lam, nlam: λ and Nλ. Each chain
    holds two replica fields (ss[1], ss[2]) in the model's field layout; replica r draws from
    Philox stream 1 + r. delay: t-U sites per decision block (1..16); t-V bonds per delayed
    block (8, or 4; below 4 element by element with in-place updates); t-U-V decides element
    by element.
    """

    def __init__(self, models, indexA, indexB, lam: float, nlam: int, seeds, chains=None,
                 delay: int = 8, sweep0: int = 0, _dop_alpha=None):
        self.L = L.load()
        if not isinstance(models, (list, tuple)):
            models = [models]
        n = len(models)
        seeds = list(seeds) if isinstance(seeds, (list, tuple)) else [seeds] * n
        chains = list(range(n)) if chains is None else list(chains)
        self.models = list(models)
        descs, runs, self._keep = _descriptors(self.models, seeds, chains, delay, sweep0)
        regs = (L.SceeDesc * n)()
        for c in range(n):
            ia = np.ascontiguousarray(indexA[c] if _is_nested(indexA) else indexA, dtype=np.int64)
            ib = np.ascontiguousarray(indexB[c] if _is_nested(indexB) else indexB, dtype=np.int64)
            self._keep += [ia, ib]
            regs[c].nA, regs[c].nB = len(ia), len(ib)
            regs[c].indexA = ia.ctypes.data_as(C.POINTER(C.c_int64))
            regs[c].indexB = ib.ctypes.data_as(C.POINTER(C.c_int64))
            regs[c].lam, regs[c].nlam = lam, nlam
        h = C.c_void_p()
        if _dop_alpha is None:
            L.check(self.L.kdq_scee_create(descs, runs, regs, n, C.byref(h)))
        else:
            L.check(self.L.kdq_dop_create(descs, runs, regs, float(_dop_alpha), n, C.byref(h)))
        self.h, self.n = h, n

    def close(self):
        if getattr(self, "h", None):
            self.L.kdq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        L.check(rc, self.h)

    def set_fields(self, c: int, replica: int, s):
        a = np.asfortranarray(np.asarray(s, dtype=np.uint8))
        if a.shape != field_shape(self.models[c]):
            raise ValueError(f"s must have shape {field_shape(self.models[c])}")
        self._chk(self.L.kdq_scee_set_fields(self.h, c, replica, a.ctypes.data_as(C.POINTER(C.c_uint8))))

    def get_fields(self, c: int, replica: int) -> np.ndarray:
        a = np.zeros(field_shape(self.models[c]), dtype=np.uint8, order="F")
        self._chk(self.L.kdq_scee_get_fields(self.h, c, replica, a.ctypes.data_as(C.POINTER(C.c_uint8))))
        return a

    def init(self):
        self._chk(self.L.kdq_scee_init(self.h))

    def set_option(self, name: str, value: int):
        self._chk(self.L.kdq_set_option(self.h, L.OPTIONS[name], int(value)))

    def sweeps(self, n: int) -> np.ndarray:
        """n sweeps of every chain (set-up included); returns [nchains, n] values of O."""
        O = np.zeros((self.n, n))
        self._chk(self.L.kdq_scee_sweeps(self.h, n, _dp(O)))
        return O

    def G4(self, c: int, replica: int):
        ns = self.models[c].Ns
        bufs = [np.zeros(2 * ns * ns) for _ in range(4)]   # interleaved, column-major
        self._chk(self.L.kdq_scee_get_G4(self.h, c, replica, *[_dp(x) for x in bufs]))
        return [(x[0::2] + 1j * x[1::2]).reshape((ns, ns), order="F") for x in bufs]   # Gt, G0, Gt0, G0t

    profile = DQMCBatch.profile
    profile_read = DQMCBatch.profile_read

    def near_ties(self) -> np.ndarray:
        """[chains, replicas]: decisions so far with |uniform draw − acceptance probability|
        < 1e-12."""
        nrep = 1 if isinstance(self, DOPBatch) else 2
        out = np.zeros(nrep * self.n)
        self._chk(self.L.kdq_get_near_ties(self.h, _dp(out)))
        return out.reshape(nrep, self.n).T.copy()

    def detg(self) -> np.ndarray:
        out = np.zeros((self.n, 2))
        self._chk(self.L.kdq_scee_detg(self.h, _dp(out)))
        return out


class DOPBatch(SCEEBatch):
    """Disorder operator of the t-U model (ctrl_SCDOPicr, src/tU/disorder_operate.jl:1-247)
    for a batch of chains: one field per chain, sampled with W · |det X_A|^{2λ} |det X_B|^{2(1-λ)},
    X = e^{iα}I + (1 - e^{iα})G(Θ)[region] (DOPBuffer(alpha, index), get_ABX!). sweeps()
    returns per sweep the mean over slices of (det X_A / det X_B)^(1/Nλ); detg() gives
    |det X_A|², |det X_B|². The field draws from Philox stream 1."""

    def __init__(self, models, alpha: float, indexA, indexB, lam: float, nlam: int, seeds, chains=None,
                 delay: int = 8, sweep0: int = 0):
        super().__init__(models, indexA, indexB, lam, nlam, seeds, chains, delay, sweep0, _dop_alpha=alpha)
        self.alpha = alpha

    def set_fields(self, c: int, s):
        super().set_fields(c, 0, s)

    def get_fields(self, c: int) -> np.ndarray:
        return super().get_fields(c, 0)

    def G4(self, c: int):
        return super().G4(c, 0)


def ctrl_SCDOPicr(path: str, model, alpha: float, indexA, indexB, Sweeps: int, λ: float, Nλ: int, s,
                  record: bool = False, *, seed: int = 0, chain: int = 0, delay: int = 8, sweep0: int = 0):
    """ctrl_SCDOPicr on the GPU (src/tU/disorder_operate.jl:1-200): mutates the field s in
    place, returns it, and appends the row [λ, O_1, …, O_Sweeps] to the reference's CSV file
    name (:26-30) when ``record`` is true."""
    b = DOPBatch([model], alpha, indexA, indexB, λ, Nλ, [seed], [chain], delay=delay, sweep0=sweep0)
    try:
        b.set_fields(0, s)
        O = b.sweeps(Sweeps)[0]
        s[...] = b.get_fields(0)
    finally:
        b.close()
    if record:
        name = _lattice_tag(model.Lattice, "tU")
        if len(np.unique(model.α)) == 1:
            fn = (f"tUSCDOP{name}_t{model.Ht}U{model.Hu1}size{model.site}Δt{model.Δt}Θ{model.Θrelax}"
                  f"N{Nλ}BS{model.BatchSize}.csv")
        else:
            fn = (f"tUSCDOP{name}_t{model.Ht}U{model.Hu1}_{model.Hu2}size{model.site}Δt{model.Δt}"
                  f"Θ{model.Θrelax}_{model.Θquench}N{Nλ}BS{model.BatchSize}.csv")
        with open(os.path.join(path, fn), "a") as fh:
            fh.write(",".join(repr(float(x)) for x in [λ, *O]) + "\n")
    return s


def _is_nested(x) -> bool:
    return len(x) > 0 and hasattr(x[0], "__len__")


def ctrl_SCEEicr(path: str, model, indexA, indexB, Sweeps: int, λ: float, Nλ: int, ss, record: bool = False,
                 *, seed: int = 0, chain: int = 0, delay: int = 8, sweep0: int = 0):
    """ctrl_SCEEicr on the GPU (t-U src/tU/SCEE.jl:1-262, t-V src/tV/SCEE.jl:1-431, t-U-V
    src/tUV/SCEE.jl:1-339): mutates the two replica fields ss[0], ss[1] in place, returns ss,
    and appends the row [λ, O_1, …, O_Sweeps] to the reference's CSV file name when
    ``record`` is true."""
    b = SCEEBatch([model], indexA, indexB, λ, Nλ, [seed], [chain], delay=delay, sweep0=sweep0)
    try:
        b.set_fields(0, 0, ss[0])
        b.set_fields(0, 1, ss[1])
        O = b.sweeps(Sweeps)[0]
        ss[0][...] = b.get_fields(0, 0)
        ss[1][...] = b.get_fields(0, 1)
    finally:
        b.close()
    if record:
        kind = getattr(model, "kind", "tU")
        name = _lattice_tag(model.Lattice, kind)
        if kind == "tU" and model.Θquench == 0.0:
            fn = (f"tUSCEE{name}_t{model.Ht}U{model.Hu1}size{model.site}Δt{model.Δt}Θ{model.Θrelax}"
                  f"N{Nλ}BS{model.BatchSize}.csv")
        elif kind == "tU":
            fn = (f"tUSCEE{name}_t{model.Ht}U{model.Hu1}_{model.Hu2}size{model.site}Δt{model.Δt}"
                  f"Θ{model.Θrelax}_{model.Θquench}N{Nλ}BS{model.BatchSize}.csv")
        elif kind == "tV" and len(np.unique(model.α)) == 1:   # src/tV/SCEE.jl:25-29
            fn = (f"tVSCEE{name}_t{model.Ht}V{model.Hv1}size{model.site}Δt{model.Δt}Θ{model.Θrelax}"
                  f"N{Nλ}BS{model.BatchSize}.csv")
        elif kind == "tV":
            fn = (f"tVSCEE{name}_t{model.Ht}V{model.Hv1}_{model.Hv2}size{model.site}Δt{model.Δt}"
                  f"Θ{model.Θrelax}_{model.Θquench}N{Nλ}BS{model.BatchSize}.csv")
        else:   # src/tUV/SCEE.jl:25
            fn = (f"tUVSCEE{name}_t{model.Ht}U{model.U}V{model.V}size{model.site}Δt{model.Δt}Θ{model.Θ}"
                  f"N{Nλ}BS{model.BatchSize}.csv")
        with open(os.path.join(path, fn), "a") as fh:
            fh.write(",".join(repr(float(x)) for x in [λ, *O]) + "\n")
    return ss
