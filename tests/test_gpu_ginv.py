"""The register factorisations of the stabilisation against the blocked kernels, on identical
batches: the overlap solve of get_G! by the Gauss-Jordan inverse and a GEMM (option "ginv")
against the blocked LU with its triangular solves, and each Cholesky-QR pass by the register
factor-inverse and a GEMM (option "qinv") against the blocked Cholesky and the in-place
triangular solve; both default for Np <= 144. G within 1e-11 relative (the paths round
differently), fields and measurement rows as in the other rounding-level comparisons. Cases
cover every kernel instantiation (Np = 9, 16, 36, 81, 100, 144), a batch of uneven Np in one
handle, t-V and t-U-V sweeps, mixed precision, the split node (second pass beside get_G!) and
log W through the summed factor diagonals; the oracle comparisons of the default path are the
sweep parity tests."""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def tu(L, U=4.0, theta=1.0, bs=5):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=bs)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def sweep(models, ginv, sweeps=2, delay=16, precision="fp64", opt="ginv", extra=()):
    from kapdqmc_amd.dqmc import DQMCBatch
    b = DQMCBatch(models, list(range(9, 9 + len(models))), list(range(len(models))), delay=delay,
                  precision=precision)
    b.set_option(opt, ginv)
    for k, v in extra:
        b.set_option(k, v)
    b.initial_s()
    rows = b.phy_update(sweeps, True)
    out = ([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows)
    b.close()
    return out


def agree(a, b, gtol=1e-11, rtol=1e-9, atol=1e-11):
    fa, ga, ra = a
    fb, gb, rb = b
    for c in range(len(fa)):
        assert np.array_equal(fa[c], fb[c]), c
        assert rel(ga[c], gb[c]) < gtol, (c, rel(ga[c], gb[c]))
    assert np.allclose(ra, rb, rtol=rtol, atol=atol), np.abs(ra - rb).max()


@pytest.mark.parametrize("opt", ["ginv", "qinv"])
@pytest.mark.parametrize("L", [3, 4, 6, 9, 10, 12])
def test_tu_register_factor_equals_blocked(L, opt):
    print(f"case: register vs blocked {opt}, t-U L={L} (Np={L * L})", flush=True)
    models = [tu(L), tu(L, U=2.0)]
    ns = 1 if L >= 10 else 2
    agree(sweep(models, 1, sweeps=ns, opt=opt), sweep(models, 0, sweeps=ns, opt=opt))


@pytest.mark.parametrize("split", [0, 1])
def test_tu_both_register_paths_against_blocked(split):
    print(f"case: register vs blocked factorisations together, qsplit={split}", flush=True)
    models = [tu(6), tu(4, U=2.0)]
    ex = (("qinv", 1), ("qsplit", split))
    ex0 = (("qinv", 0), ("qsplit", split))
    agree(sweep(models, 1, extra=ex), sweep(models, 0, extra=ex0))


def test_log_weight_register_qr():
    # log W sums log L_jj of both Cholesky-QR passes over the chain: the same with either path
    from kapdqmc_amd.dqmc import DQMCBatch
    print("case: log W with register vs blocked Cholesky-QR", flush=True)
    models = [tu(4), tu(6, U=2.0)]
    out = []
    for q in (1, 0):
        b = DQMCBatch(models, [3, 4], [0, 1], delay=16)
        b.set_option("qinv", q)
        b.initial_s()
        out.append(b.log_weight())
        b.close()
    assert np.allclose(out[0], out[1], rtol=1e-11, atol=1e-9), (out[0], out[1])


@pytest.mark.parametrize("opt", ["ginv", "qinv"])
def test_uneven_batch(opt):
    print(f"case: register vs blocked {opt}, uneven batch L = 3/4/6", flush=True)
    models = [tu(3), tu(6, U=6.0), tu(4, U=2.0)]
    agree(sweep(models, 1, opt=opt), sweep(models, 0, opt=opt))


@pytest.mark.parametrize("opt", ["ginv", "qinv"])
@pytest.mark.parametrize("name", ["tV", "tUV"])
def test_r2_models(name, opt):
    print(f"case: register vs blocked {opt}, {name}", flush=True)
    if name == "tV":
        models = [tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                                  Lattice="HoneyComb120", site=[L, L], BatchSize=5) for L in (3, 4)]
    else:
        models = [tUV_Hubbard_Para(Ht=1.0, Hu=U, Hv=0.5, Lattice="HoneyComb120", site=[4, 4], Δt=0.1, Θ=1.0,
                                   BatchSize=5) for U in (4.0, -4.0)]
    agree(sweep(models, 1, delay=8, opt=opt), sweep(models, 0, delay=8, opt=opt))


@pytest.mark.parametrize("opt", ["ginv", "qinv"])
def test_mixed_precision(opt):
    print(f"case: register vs blocked {opt}, mixed precision", flush=True)
    models = [tu(4), tu(6, U=2.0)]
    agree(sweep(models, 1, precision="mixed", opt=opt), sweep(models, 0, precision="mixed", opt=opt), gtol=1e-5,
          rtol=1e-4, atol=1e-6)
