"""Pins of the numpy restatement of the self-learning effective model (oracle/slmc_net.py):
forward pass vs direct loops, backward pass vs central finite differences, and invariance of
E under lattice translations. The first test plants a dropped term in the oracle's backward
pass and checks that the finite-difference comparison catches it."""
import numpy as np

from oracle import slmc_net as sn


def _setup(L1=2, L2=3, nt=4, C=3, B=3, seed=1):
    geo = sn.Geometry(L1, L2, nt)
    _, npar = sn.param_layout(C)
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, npar)
    f = rng.integers(1, 5, size=(B, geo.n, geo.T))
    lw = rng.normal(0, 1, B)
    return geo, C, p, f, lw, rng


def _fd_err(geo, C, p, f, lw, rng, grad_fn, n=25, eps=1e-6):
    _, g = grad_fn(geo, p, C, f, lw)
    err = 0.0
    for i in rng.choice(len(p), n, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (sn.loss_grad(geo, pp, C, f, lw)[0] - sn.loss_grad(geo, pm, C, f, lw)[0]) / (2 * eps)
        err = max(err, abs(fd - g[i]) / (abs(fd) + 1e-6))
    return err


def test_backward_matches_finite_differences():
    geo, C, p, f, lw, rng = _setup()
    assert _fd_err(geo, C, p, f, lw, rng, sn.loss_grad) < 1e-5

    # planted mistake: drop the tanh derivative of layer 2 -> the check must fail
    def broken(geo, p, C, f, lw):
        loss, g = sn.loss_grad(geo, p, C, f, lw)
        W1, b1, W2, b2, W3 = sn.unpack(p, C)
        E, c = sn.forward(geo, p, C, f)
        B = len(E)
        d = E + lw - np.mean(E + lw)
        do = np.broadcast_to((2.0 * d / B)[:, None, None], (B, 1, geo.P))
        dh2 = sn._scatter(geo, np.einsum("oik,boq->bikq", W3, do))
        gb2_wrong = dh2.sum(axis=(0, 2))
        lay, _ = sn.param_layout(C)
        g = g.copy()
        g[lay["b2"][0]:lay["b2"][0] + C] = gb2_wrong
        return loss, g

    assert _fd_err(geo, C, p, f, lw, np.random.default_rng(5), broken, n=len(p)) > 1e-3


def test_forward_matches_loops():
    geo, C, p, f, lw, rng = _setup(L1=3, L2=2, nt=3)
    E = sn.energy(geo, p, C, f)
    El = [sn.energy_loops(geo, p, C, f[b]) for b in range(len(f))]
    assert np.allclose(E, El, rtol=1e-12, atol=1e-12)


def test_energy_translation_invariant():
    geo, C, p, f, lw, rng = _setup(L1=3, L2=4, nt=5, B=2)
    E = sn.energy(geo, p, C, f)
    # translate every field by one cell in x and two in y
    X, Y = geo.X, geo.Y
    g = np.empty_like(f)
    for i in range(geo.n):
        sub, cell = i % 2, i // 2
        x, y = cell % X, cell // X
        j = 2 * (((x + 1) % X) + X * ((y + 2) % Y)) + sub
        g[:, j, :] = f[:, i, :]
    assert np.allclose(sn.energy(geo, p, C, g), E, rtol=1e-12)


def test_bf16_rounding_matches_torch():
    """bf16_weights rounds W1, W2, W3 as torch's bfloat16 conversion does (nearest even) and
    leaves the biases alone."""
    import torch
    C = 3
    _, n = sn.param_layout(C)
    rng = np.random.default_rng(11)
    p = rng.normal(0, 0.5, n).astype(np.float32)
    p[:5] = [1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 2 ** -8 + 2 ** -16, 0.0]   # ties and small values
    q = sn.bf16_weights(p, C)
    ref = torch.from_numpy(p).to(torch.bfloat16).float().double().numpy()
    lay, _ = sn.param_layout(C)
    for name in ("W1", "W2", "W3"):
        off, k = lay[name]
        assert np.array_equal(q[off:off + k], ref[off:off + k])
    for name in ("b1", "b2"):
        off, k = lay[name]
        assert np.array_equal(q[off:off + k], p[off:off + k].astype(np.float64))
    assert q[0] == 1.0 and q[1] == 1 + 2 ** -6   # ties to even
