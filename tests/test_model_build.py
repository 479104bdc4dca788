"""The library's host model builder (kdq_model_build, C++, no GPU work) against the Python
host constructors, the oracle's geometry and Jacobi eigensolver, and facts that fix the
result independently of either eigensolver: e^{-Δτ H} by scaled-and-squared Padé
(scipy.linalg.expm), eK·eKinv = I, HalfeK² = eK, the trial state as the projector onto
the lowest (t-U) or highest (t-V, t-U-V) half of the staggered K. Runs on the CPU."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg

from kapdqmc_amd.models import native_model, tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

CTORS = {"tU": tU_Hubbard_Para, "tV": tV_Hubbard_Para, "tUV": tUV_Hubbard_Para}
MATS = ("K", "eK", "eKinv", "HalfeK", "HalfeKinv")


def proj(P):
    return P @ P.T


def check_same(kind, kw, tol=1e-13):
    a, b = native_model(kind, **kw), CTORS[kind](**kw)
    assert (a.Ns, a.Nt, a.BatchSize) == (b.Ns, b.Nt, b.BatchSize)
    for f in MATS:
        assert np.abs(getattr(a, f) - getattr(b, f)).max() < tol, f
    assert np.array_equal(a.K, b.K)
    assert np.array_equal(a.nodes, b.nodes)
    assert np.array_equal(a.η, b.η) and np.array_equal(a.γ, b.γ)
    if kind != "tUV":
        assert np.array_equal(a.α, b.α)
    if kind != "tU":
        assert np.array_equal(a.nnidx, b.nnidx)
    if kind == "tUV":
        assert a.a == b.a and a.Type is b.Type
    assert np.abs(proj(a.Pt) - proj(b.Pt)).max() < 1e-9
    return a


@pytest.mark.parametrize("lattice", ["HoneyComb120", "HoneyComb60", "SQUARE90", "SQUARE45"])
@pytest.mark.parametrize("initial", ["H0", "V"])
def test_tu_equals_host_constructor(lattice, initial):
    # a quench profile (Hu1 != Hu2) and an odd node layout (Nt/2 not a multiple of BatchSize)
    check_same("tU", dict(Ht=1.0, Hu1=3.0, Hu2=5.0, Δt=0.1, Θrelax=1.0, Θquench=0.5, Lattice=lattice,
                          site=[4, 3], BatchSize=3, Initial=initial))


@pytest.mark.parametrize("h1,h2", [(0.0, 3.8), (1.3, 3.8)])
def test_quench_profile_is_julia_linrange(h1, h2):
    """A step count that does not divide the range (the reference's quench test Hu 0 -> 3.8,
    here over Θquench = 0.7, 7 slices): the profile is Julia's LinRange element for element,
    (1 - t) h1 + t h2 with t = i / nq, in both constructors; numpy.linspace differs from it in
    the last bit at some i, so the equality below would catch it."""
    kw = dict(Ht=1.0, Hu1=h1, Hu2=h2, Δt=0.1, Θrelax=0.4, Θquench=0.7, Lattice="HoneyComb120",
              site=[3, 3], BatchSize=10, Initial="H0")
    m = check_same("tU", kw)
    t = np.arange(1, 8) / 7
    lin = np.array([(1.0 - x) * h1 + x * h2 for x in t])   # Julia: T((1-t)*a + t*b)
    prof = np.concatenate([np.full(4, h1), lin, lin[::-1], np.full(4, h1)])
    assert np.array_equal(m.α, np.sqrt(0.1 * prof / 2))
    if h1 == 0.0:
        assert (np.linspace(h1, h2, 8)[1:] != lin).any()


@pytest.mark.parametrize("lattice", ["HoneyComb120", "HoneyComb60"])
@pytest.mark.parametrize("initial", ["H0", "V"])
def test_tv_and_tuv_equal_host_constructors(lattice, initial):
    check_same("tV", dict(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.05, Θrelax=1.0, Θquench=0.0, Lattice=lattice,
                          site=[3, 3], BatchSize=4, Initial=initial))
    for U in (4.0, -4.0):
        check_same("tUV", dict(Ht=1.0, Hu=U, Hv=0.5, Lattice=lattice, site=[3, 3], Δt=0.1, Θ=1.0,
                               BatchSize=4, Initial=initial))


def test_against_oracle_geometry_and_eigensolver(oracle_mod):
    m = native_model("tU", Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                     Lattice="HoneyComb120", site=[6, 4], BatchSize=5)
    assert np.array_equal(m.K, oracle_mod.nnK("HoneyComb120", [6, 4]))
    KK = m.K + 1e-5 * np.diag(np.tile([-1.0, 1.0], m.Ns // 2))
    _, V = oracle_mod.eigh(KK)
    assert np.abs(proj(V[:, : m.Ns // 2]) - proj(m.Pt)).max() < 1e-8


@pytest.mark.parametrize("kind,lattice,L", [("tU", "HoneyComb120", 6), ("tU", "SQUARE90", 5),
                                            ("tV", "HoneyComb60", 4)])
def test_propagators_against_pade(kind, lattice, L):
    kw = dict(Ht=1.3, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice=lattice, site=[L, L], BatchSize=5)
    kw.update(Hu1=4.0, Hu2=4.0) if kind == "tU" else kw.update(Hv1=1.0, Hv2=1.0)
    m = native_model(kind, **kw)
    H = (1.3 if kind == "tU" else -1.3) * m.K
    n = m.Ns
    assert np.abs(m.eK - scipy.linalg.expm(-0.1 * H)).max() < 1e-13
    assert np.abs(m.HalfeKinv - scipy.linalg.expm(0.05 * H)).max() < 1e-13
    assert np.abs(m.eK @ m.eKinv - np.eye(n)).max() < 1e-13
    assert np.abs(m.HalfeK @ m.HalfeK - m.eK).max() < 1e-13
    # trial state: orthonormal columns spanning an invariant subspace of the staggered K
    mu = 1e-3 if kind == "tV" else 1e-5
    KK = m.K + mu * np.diag(np.tile([-1.0, 1.0], n // 2))
    P = m.Pt if kind == "tU" else m.HalfeK @ m.Pt
    assert np.abs(P.T @ P - np.eye(n // 2)).max() < 1e-12
    R = KK @ P - P @ (P.T @ KK @ P)
    assert np.abs(R).max() < 1e-11
    w = np.linalg.eigvalsh(KK)
    e = np.sort(np.linalg.eigvalsh(P.T @ KK @ P))
    ref = w[: n // 2] if kind == "tU" else w[n // 2:]
    assert np.abs(e - ref).max() < 1e-11


def test_large_lattice():
    # 15 x 15 honeycomb, N = 450 (the largest single-chain lattice of the workloads)
    kw = dict(Ht=1.0, Hu1=5.0, Hu2=5.0, Δt=0.1, Θrelax=12.0, Θquench=0.0, Lattice="HoneyComb120",
              site=[15, 15], BatchSize=10)
    m = check_same("tU", kw, tol=1e-12)
    assert m.Ns == 450 and m.Nt == 240


def test_reference_asserts_become_errors():
    from kapdqmc_amd._lib import KdqError
    with pytest.raises(KdqError, match="U\\^2 > 4 z\\^2 V\\^2"):
        native_model("tUV", Ht=1.0, Hu=1.0, Hv=0.5, Lattice="HoneyComb120", site=[3, 3], Δt=0.1, Θ=1.0,
                     BatchSize=5)
    with pytest.raises(KdqError, match="must be equal"):
        native_model("tU", Ht=1.0, Hu1=3.0, Hu2=4.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                     Lattice="HoneyComb120", site=[3, 3], BatchSize=5)
    with pytest.raises(KdqError, match="honeycomb"):
        native_model("tV", Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="SQUARE90",
                     site=[3, 3], BatchSize=5)


def test_null_and_bad_arguments():
    from kapdqmc_amd import _lib as L
    lib = L.load()
    h, d = C.c_void_p(), L.ModelDesc()
    assert lib.kdq_model_build(None, C.byref(h), C.byref(d)) == L.KDQ_EARG
    p = L.ModelParams()
    p.kind, p.lattice, p.batch_size, p.dt, p.ht = L.MODEL_TU, 7, 5, 0.1, 1.0
    p.site[0] = p.site[1] = 3
    assert lib.kdq_model_build(C.byref(p), C.byref(h), C.byref(d)) == L.KDQ_EARG
    assert b"lattice" in lib.kdq_last_error(None)
    p.lattice, p.site[0] = 0, 0
    assert lib.kdq_model_build(C.byref(p), C.byref(h), C.byref(d)) == L.KDQ_EARG
    assert lib.kdq_model_free(None) == L.KDQ_OK


C_PROGRAM = r"""
#include <stdio.h>
#include "kapdqmc.h"
int main(void) {
    kdq_model_params p = {0};
    p.kind = KDQ_MODEL_TU; p.lattice = KDQ_LATTICE_HC120; p.site[0] = 3; p.site[1] = 3;
    p.batch_size = 5; p.initial = KDQ_INITIAL_H0; p.ht = 1.0; p.c1 = p.c2 = 4.0; p.dt = 0.1;
    p.theta_relax = 1.0; p.theta_quench = 0.0;
    kdq_model m; kdq_model_desc d;
    int rc = kdq_model_build(&p, &m, &d);
    if (rc != KDQ_OK) { printf("error %d %s\n", rc, kdq_last_error(NULL)); return 1; }
    printf("%d %d %d %.17g %.17g\n", d.ns, d.nt, d.n_nodes, d.eK[0], d.alpha[0]);
    return kdq_model_free(m);
}
"""


def test_header_is_c_and_builder_runs_from_c(tmp_path):
    # the header compiles as C99 and a C program builds a model through the library
    import os
    import subprocess
    from kapdqmc_amd import _lib
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "m.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "m"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(root, "include"), str(src), "-o",
                    str(exe), "-L", libdir, "-l:libkapdqmc_hip.so", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                        site=[3, 3], BatchSize=5)
    assert [int(x) for x in out[:3]] == [m.Ns, m.Nt, len(m.nodes)]
    assert abs(float(out[3]) - m.eK[0, 0]) < 1e-14 and float(out[4]) == m.α[0]
