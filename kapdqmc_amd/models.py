"""Host-side model constructors mirroring the reference's parameter structs.

``tU_Hubbard_Para`` restates ``src/tU/model.jl:39-96``, ``tV_Hubbard_Para`` restates
``src/tV/model.jl:30-122`` and ``tUV_Hubbard_Para`` restates ``src/tUV/model.jl:30-118``.
They run once on the host (two symmetric eigen-decompositions, as the reference's
``LAPACK.syevd!`` calls) and produce the constants the GPU library consumes.
Keyword names and meanings follow the reference exactly.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

from .geometry import nnK_Matrix, nnidx_F

GAMMA = np.array([1 + math.sqrt(6) / 3, 1 + math.sqrt(6) / 3, 1 - math.sqrt(6) / 3, 1 - math.sqrt(6) / 3])
ETA = np.array([math.sqrt(2 * (3 - math.sqrt(6))), -math.sqrt(2 * (3 - math.sqrt(6))),
                math.sqrt(2 * (3 + math.sqrt(6))), -math.sqrt(2 * (3 + math.sqrt(6)))])


def _nodes(Nt: int, BatchSize: int) -> np.ndarray:
    """Stabilisation nodes (src/tU/model.jl:81-85; identical in tV/tUV)."""
    h = Nt // 2
    if h % BatchSize == 0:
        return np.arange(0, Nt + 1, BatchSize, dtype=np.int64)
    left = list(range(h - BatchSize, 0, -BatchSize))[::-1]
    right = list(range(h, Nt + 1, BatchSize))
    return np.array([0] + left + right + [Nt], dtype=np.int64)


def _quench_profile(H1, H2, Theta_relax, Theta_quench, dt, Nt):
    """Piecewise-linear coupling profile (src/tU/model.jl:41-52)."""
    if Theta_quench > 0.0 and abs(H1 - H2) > 0:
        nq = round(Theta_quench / dt)
        t = np.arange(1, nq + 1) / nq              # LinRange(H1, H2, nq + 1)[2:end] as Julia's lerp
        lin = (1.0 - t) * H1 + t * H2
        nr = round(Theta_relax / dt)
        prof = np.concatenate([np.full(nr, H1), lin, lin[::-1], np.full(nr, H1)])
    else:
        if not (H1 == H2 and Theta_quench < 1e-7):
            raise AssertionError("For Θquench=0, the two couplings must be equal")
        prof = H1 * np.ones(Nt)
    if np.linalg.norm(prof[::-1] - prof) >= 1e-10:
        raise AssertionError("coupling profile is not symmetric!")
    if len(prof) != Nt:
        raise AssertionError("Length of coupling profile does not match Nt!")
    return prof


def _exp_sym(A: np.ndarray, dt: float):
    """exp(∓Δτ A / 2), exp(∓Δτ A) of a real symmetric or complex Hermitian A (syevd!)."""
    E, V = np.linalg.eigh(A)
    f = lambda c: (V * np.exp(c * E)) @ V.conj().T
    return f(-dt / 2), f(-dt), f(dt / 2), f(dt)


@dataclass
class tU_Hubbard_Para_:
    """Field-for-field mirror of ``tU_Hubbard_Para_`` (src/tU/model.jl:14-37)."""
    Lattice: str
    Ht: float
    Hu1: float
    Hu2: float
    site: list
    Θrelax: float
    Θquench: float
    Ns: int
    Nt: int
    K: np.ndarray
    BatchSize: int
    Δt: float
    α: np.ndarray
    γ: np.ndarray
    η: np.ndarray
    Pt: np.ndarray
    HalfeK: np.ndarray
    eK: np.ndarray
    HalfeKinv: np.ndarray
    eKinv: np.ndarray
    nodes: np.ndarray
    kind: str = field(default="tU")


def tU_Hubbard_Para(*, Ht, Hu1, Hu2, Δt, Θrelax, Θquench, Lattice: str, site, BatchSize,
                    Initial: str = "H0", flux: float = 0.0) -> tU_Hubbard_Para_:
    """t-U Hubbard model constants (src/tU/model.jl:39-96).

    The trial state is the lowest Ns/2 eigenvectors of K + 1e-5·diag(-1,+1,...) (the
    reference's staggered potential, model.jl:68-76), or the "V" product state. flux != 0
    gives complex Hermitian K, eK, eKinv, HalfeK, HalfeKinv and a complex Pt (the reference
    stores them ComplexF64 in any case, model.jl:24-34).
    """
    site = list(site)
    Nt = round(2 * (Θrelax + Θquench) / Δt)
    Hu = _quench_profile(Hu1, Hu2, Θrelax, Θquench, Δt, Nt)
    alpha = np.sqrt(Δt * Hu / 2)
    K = nnK_Matrix(Lattice, site, flux=flux)
    Ns = K.shape[0]
    HalfeK, eK, HalfeKinv, eKinv = _exp_sym(Ht * K, Δt)
    if np.linalg.norm(eK @ eKinv - np.eye(Ns)) >= 1e-10:
        raise AssertionError("eK*eKinv does not equal identity!")
    Pt = np.zeros((Ns, Ns // 2), dtype=K.dtype)   # ComplexF64 with complex hopping (model.jl:68)
    if Initial == "H0":
        KK = K + 1e-5 * np.diag(np.tile([-1.0, 1.0], Ns // 2))
        _, V = np.linalg.eigh(KK)
        Pt[:] = V[:, : Ns // 2]
    elif Initial == "V":
        for i in range(1, Ns // 2 + 1):
            Pt[i * 2 - 1, i - 1] = 1.0
    else:
        raise ValueError(f"Initial must be 'H0' or 'V', got {Initial!r}")
    return tU_Hubbard_Para_(Lattice, float(Ht), float(Hu1), float(Hu2), site, float(Θrelax),
                            float(Θquench), Ns, Nt, K, int(BatchSize), float(Δt), alpha,
                            GAMMA.copy(), ETA.copy(), Pt, HalfeK, eK, HalfeKinv, eKinv,
                            _nodes(Nt, BatchSize))


@dataclass
class tV_Hubbard_Para_:
    """Mirror of ``tV_Hubbard_Para_`` (src/tV/model.jl:3-28); UV is replaced by the bond
    list (the 2x2 rotations are applied per bond instead of as dense N x N products)."""
    Lattice: str
    Ht: float
    Hv1: float
    Hv2: float
    site: list
    Θrelax: float
    Θquench: float
    Ns: int
    Nt: int
    K: np.ndarray
    BatchSize: int
    Δt: float
    α: np.ndarray
    γ: np.ndarray
    η: np.ndarray
    Pt: np.ndarray
    HalfeK: np.ndarray
    eK: np.ndarray
    HalfeKinv: np.ndarray
    eKinv: np.ndarray
    nnidx: np.ndarray
    nodes: np.ndarray
    kind: str = field(default="tV")

    def UV(self) -> np.ndarray:
        """Dense bond-rotation matrices UV[:, :, j] (src/tV/model.jl:103-110)."""
        Ns = self.Ns
        UV = np.zeros((Ns, Ns, self.nnidx.shape[1]))
        h = 2 ** 0.5 / 2
        for j in range(self.nnidx.shape[1]):
            for i in range(self.nnidx.shape[0]):
                x, y = self.nnidx[i, j] - 1
                UV[x, x, j] = UV[x, y, j] = UV[y, x, j] = -h
                UV[y, y, j] = h
        return UV


def tV_Hubbard_Para(*, Ht, Hv1, Hv2, Δt, Θrelax, Θquench, Lattice: str, site, BatchSize,
                    Initial: str = "H0") -> tV_Hubbard_Para_:
    """Spinless t-V model constants (src/tV/model.jl:30-122)."""
    site = list(site)
    if "HoneyComb" not in Lattice:
        raise ValueError("tV supports honeycomb lattices only")
    K = nnK_Matrix(Lattice, site)
    Ns = K.shape[0]
    HalfeK, eK, HalfeKinv, eKinv = _exp_sym(-Ht * K, Δt)
    Pt = np.zeros((Ns, Ns // 2))
    if Initial == "H0":
        KK = K + 1e-3 * np.diag(np.tile([-1.0, 1.0], Ns // 2))
        _, V = np.linalg.eigh(KK)
        Pt[:] = V[:, Ns // 2:]
    elif Initial == "V":
        for i in range(1, Ns // 2 + 1):
            Pt[i * 2 - 2, i - 1] = 1.0
    else:
        raise ValueError(f"Initial must be 'H0' or 'V', got {Initial!r}")
    Pt = HalfeKinv @ Pt
    Nt = round(2 * (Θrelax + Θquench) / Δt)
    Hv = _quench_profile(Hv1, Hv2, Θrelax, Θquench, Δt, Nt)
    alpha = np.sqrt(Δt * Hv / 2)
    return tV_Hubbard_Para_(Lattice, float(Ht), float(Hv1), float(Hv2), site, float(Θrelax),
                            float(Θquench), Ns, Nt, K, int(BatchSize), float(Δt), alpha,
                            GAMMA.copy(), ETA.copy(), Pt, HalfeK, eK, HalfeKinv, eKinv,
                            nnidx_F(Lattice, site), _nodes(Nt, BatchSize))


@dataclass
class tUV_Hubbard_Para_:
    """Mirror of ``tUV_Hubbard_Para_`` (src/tUV/model.jl:3-27)."""
    Lattice: str
    Ht: float
    U: float
    V: float
    a: float
    Type: type
    site: list
    Θ: float
    Ns: int
    Nt: int
    K: np.ndarray
    BatchSize: int
    Δt: float
    γ: np.ndarray
    η: np.ndarray
    Pt: np.ndarray
    HalfeK: np.ndarray
    eK: np.ndarray
    HalfeKinv: np.ndarray
    eKinv: np.ndarray
    nnidx: np.ndarray      # [Ns, z, 2] directed bonds, 1-based
    nodes: np.ndarray
    kind: str = field(default="tUV")


def tUV_Hubbard_Para(*, Ht, Hu, Hv, Lattice: str, site, Δt, Θ, BatchSize,
                     Initial: str = "H0") -> tUV_Hubbard_Para_:
    """t-U-V model constants (src/tUV/model.jl:30-118)."""
    site = list(site)
    if "HoneyComb" not in Lattice:
        raise ValueError("tUV supports honeycomb lattices only")
    K = nnK_Matrix(Lattice, site)
    Ns = K.shape[0]
    HalfeK, eK, HalfeKinv, eKinv = _exp_sym(-Ht * K, Δt)
    Pt = np.zeros((Ns, Ns // 2))
    if Initial == "H0":
        KK = K + 1e-5 * np.diag(np.tile([-1.0, 1.0], Ns // 2))
        _, Vv = np.linalg.eigh(KK)
        Pt[:] = Vv[:, Ns // 2:]
    elif Initial == "V":
        for i in range(1, Ns // 2 + 1):
            Pt[i * 2 - 2, i - 1] = 1.0
    else:
        raise ValueError(f"Initial must be 'H0' or 'V', got {Initial!r}")
    Pt = HalfeKinv @ Pt
    Nt = 2 * math.ceil(Θ / Δt)
    z = 3.0
    if not Hu ** 2 > 4 * z ** 2 * Hv ** 2:
        raise AssertionError("For stability, require U^2 > 4 z^2 V^2")
    Type = complex if Hu > 0 else float
    if Hu > 0:
        g = Hu / z / 2 + math.sqrt(Hu ** 2 / z ** 2 / 4 - Hv ** 2)
    else:
        g = Hu / z / 2 - math.sqrt(Hu ** 2 / z ** 2 / 4 - Hv ** 2)
    a = 0.0 if g == 0 else Hv / g
    eta = math.sqrt(abs(Δt * g)) * np.array([math.sqrt(3 - math.sqrt(6)), -math.sqrt(3 - math.sqrt(6)),
                                             math.sqrt(3 + math.sqrt(6)), -math.sqrt(3 + math.sqrt(6))])
    tmp = nnidx_F(Lattice, site)
    nnidx = np.zeros((Ns, tmp.shape[1], 2), dtype=np.int64)
    for i in range(tmp.shape[0]):
        for j in range(tmp.shape[1]):
            nnidx[2 * i, j] = tmp[i, j]
            nnidx[2 * i + 1, j] = tmp[i, j][::-1]
    return tUV_Hubbard_Para_(Lattice, float(Ht), float(Hu), float(Hv), a, Type, site, float(Θ),
                             Ns, int(Nt), K, int(BatchSize), float(Δt), GAMMA.copy(), eta, Pt,
                             HalfeK, eK, HalfeKinv, eKinv, nnidx, _nodes(int(Nt), BatchSize))


def native_model(kind: str, **kw):
    """The same constructors run by the library's C++ builder (``kdq_model_build``): the
    keyword arguments of ``tU_Hubbard_Para`` / ``tV_Hubbard_Para`` / ``tUV_Hubbard_Para``,
    the same struct back. What a C or C++ caller without Python gets; Python callers may use
    either. The eigenvectors come from the library's own eigensolver, so ``Pt`` equals the
    numpy-built one up to a rotation inside its span (the same projector)."""
    import ctypes as C

    from . import _lib as L

    lib = L.load()
    p = L.ModelParams()
    p.kind = {"tU": L.MODEL_TU, "tV": L.MODEL_TV, "tUV": L.MODEL_TUV}[kind]
    p.lattice = L.LATTICE_CODE[kw["Lattice"]]
    p.site[0], p.site[1] = kw["site"]
    p.batch_size = kw["BatchSize"]
    p.initial = L.INITIAL_CODE[kw.get("Initial", "H0")]
    p.ht, p.dt = kw["Ht"], kw["Δt"]
    if kind == "tUV":
        p.c1, p.c2, p.theta_relax = kw["Hu"], kw["Hv"], kw["Θ"]
    else:
        c = "Hu" if kind == "tU" else "Hv"
        p.c1, p.c2 = kw[c + "1"], kw[c + "2"]
        p.theta_relax, p.theta_quench = kw["Θrelax"], kw["Θquench"]
    h = C.c_void_p()
    d = L.ModelDesc()
    L.check(lib.kdq_model_build(C.byref(p), C.byref(h), C.byref(d)))
    try:
        ns, nt = d.ns, d.nt
        mat = lambda ptr, cols=None: np.ctypeslib.as_array(ptr, shape=((cols or ns) * ns,)).reshape(
            (cols or ns, ns)).T.copy()
        vec = lambda ptr, n: np.ctypeslib.as_array(ptr, shape=(n,)).copy()
        K, eK, eKinv = mat(d.K), mat(d.eK), mat(d.eKinv)
        hK, hKinv, Pt = mat(d.halfeK), mat(d.halfeKinv), mat(d.Pt, ns // 2)
        nodes = vec(d.nodes, d.n_nodes)
        eta, gamma = vec(d.eta, 4), vec(d.gamma, 4)
        site = list(kw["site"])
        if kind == "tU":
            return tU_Hubbard_Para_(kw["Lattice"], float(p.ht), float(p.c1), float(p.c2), site,
                                    float(p.theta_relax), float(p.theta_quench), ns, nt, K, int(p.batch_size),
                                    float(p.dt), vec(d.alpha, nt), gamma, eta, Pt, hK, eK, hKinv, eKinv, nodes)
        bonds = vec(d.bonds, 2 * 3 * (ns // 2 if kind == "tV" else ns)).astype(np.int64)
        if kind == "tV":
            nnidx = bonds.reshape(3, ns // 2, 2).transpose(1, 0, 2).copy()
            return tV_Hubbard_Para_(kw["Lattice"], float(p.ht), float(p.c1), float(p.c2), site,
                                    float(p.theta_relax), float(p.theta_quench), ns, nt, K, int(p.batch_size),
                                    float(p.dt), vec(d.alpha, nt), gamma, eta, Pt, hK, eK, hKinv, eKinv, nnidx,
                                    nodes)
        nnidx = bonds.reshape(3, ns, 2).transpose(1, 0, 2).copy()
        return tUV_Hubbard_Para_(kw["Lattice"], float(p.ht), float(p.c1), float(p.c2), d.a,
                                 complex if d.is_complex else float, site, float(p.theta_relax), ns, nt, K,
                                 int(p.batch_size), float(p.dt), gamma, eta, Pt, hK, eK, hKinv, eKinv, nnidx,
                                 nodes)
    finally:
        lib.kdq_model_free(h)
