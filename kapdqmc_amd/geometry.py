"""Lattice geometry for the host-side model constructors.

Restates the parts of ``src/public/Geometry.jl`` that the t-U, t-V and t-U-V models use.
Indices are 1-based, exactly as in the reference, so that field layouts and area lists
cross the boundary unchanged.
"""
from __future__ import annotations

import numpy as np

LATTICES = ("HoneyComb120", "HoneyComb60", "SQUARE90", "SQUARE45")


def _mod1(a: int, n: int) -> int:
    r = a % n
    return r if r != 0 else n


def xy_i(site, x: int, y: int) -> int:
    """Even 1-based index of unit cell (x, y) (Geometry.jl:8-13)."""
    return 2 * (x + (y - 1) * site[0])


def i_xy(site, i: int):
    """Unit-cell coordinates of 1-based index i (Geometry.jl:18-21)."""
    j = -(-i // 2)
    return _mod1(j, site[0]), -(-j // site[0])


def nn2idx(lattice: str, site, idx: int):
    """1-based nearest-neighbour list of site idx (Geometry.jl:74-134)."""
    x, y = i_xy(site, idx)
    L1, L2 = site
    if lattice == "SQUARE90":
        if idx % 2 == 1:
            return [idx + 1, xy_i(site, _mod1(x - 1, L1), _mod1(y - 1, L2)),
                    xy_i(site, x, _mod1(y - 1, L2)), xy_i(site, _mod1(x - 1, L1), y)]
        return [xy_i(site, _mod1(x + 1, L1), y) - 1, xy_i(site, x, _mod1(y + 1, L2)) - 1,
                idx - 1, xy_i(site, _mod1(x + 1, L1), _mod1(y + 1, L2)) - 1]
    if lattice == "SQUARE45":
        if idx % 2 == 1:
            return [xy_i(site, _mod1(x - 1, L1), y), idx + 1,
                    xy_i(site, _mod1(x - 1, L1), _mod1(y + 1, L2)), xy_i(site, x, _mod1(y - 1, L2))]
        return [xy_i(site, x, _mod1(y + 1, L2)) - 1, xy_i(site, _mod1(x + 1, L1), _mod1(y - 1, L2)) - 1,
                idx - 1, xy_i(site, _mod1(x + 1, L1), _mod1(y, L2)) - 1]
    if lattice == "HoneyComb120":
        if idx % 2 == 1:
            return [idx + 1, xy_i(site, _mod1(x + 1, L1), y), xy_i(site, x, _mod1(y - 1, L2))]
        return [idx - 1, xy_i(site, x, _mod1(y + 1, L2)) - 1, xy_i(site, _mod1(x - 1, L1), y) - 1]
    if lattice == "HoneyComb60":
        if idx % 2 == 1:
            return [idx + 1, xy_i(site, _mod1(x + 1, L1), _mod1(y - 1, L2)),
                    xy_i(site, x, _mod1(y - 1, L2))]
        return [idx - 1, xy_i(site, x, _mod1(y + 1, L2)) - 1,
                xy_i(site, _mod1(x - 1, L1), _mod1(y + 1, L2)) - 1]
    raise ValueError(f"Lattice: {lattice} is not allowed !")


def nnK_Matrix(lattice: str, site, t=(1.0, 1.0, 1.0), flux: float = 0.0) -> np.ndarray:
    """Nearest-neighbour hopping matrix (Geometry.jl:178-207): real for flux = 0; complex
    otherwise, with the square lattices' bonds carrying e^{±i flux/4} (the first two
    neighbours e^{+i flux/4}, the last two e^{-i flux/4}) and the honeycomb's the real t."""
    ns = 2 * site[0] * site[1]
    flux1, flux2 = np.exp(1j * flux / 4), np.exp(-1j * flux / 4)
    K = np.zeros((ns, ns)) if flux == 0 else np.zeros((ns, ns), dtype=np.complex128)
    for i in range(1, ns + 1):
        nn = nn2idx(lattice, site, i)
        if lattice.startswith("SQUARE"):
            for j, nj in enumerate(nn):
                v = flux1 if j < 2 else flux2
                K[i - 1, nj - 1] = v if flux != 0 else 1.0
        else:
            for j, nj in enumerate(nn):
                K[i - 1, nj - 1] = t[j]
    if np.linalg.norm(K - K.conj().T) >= 1e-8:
        raise AssertionError("K is not Hermitian!")
    return K


def nnidx_F(lattice: str, site):
    """Bond list of the honeycomb lattice (Geometry.jl:24-66): array [ns/2, 3, 2] of 1-based
    (A-site, neighbour) pairs, bond family j along the second axis."""
    if "HoneyComb" not in lattice:
        raise ValueError("nnidx_F: only honeycomb lattices are supported")
    ns = 2 * site[0] * site[1]
    out = np.zeros((ns // 2, 3, 2), dtype=np.int64)
    for count, i in enumerate(range(1, ns + 1, 2)):
        for j, nj in enumerate(nn2idx(lattice, site, i)):
            out[count, j] = (i, nj)
    return out


def area_index(lattice: str, site, area):
    """Site list of a rectangular area of unit cells (Geometry.jl:209-307, the rectangle
    branch used by the honeycomb and SQUARE90 lattices)."""
    (x0, y0), (x1, y1) = area
    if lattice not in ("HoneyComb120", "HoneyComb60", "SQUARE90"):
        raise NotImplementedError(f"area_index for {lattice}")
    if x0 < 0:
        raise NotImplementedError("zigzag/bearded areas are not supported")
    index = []
    for lx in range(x0, x1 + 1):
        for ly in range(y0, y1 + 1):
            a = xy_i(site, lx, ly) - 1
            index += [a, a + 1]
    return np.array(index, dtype=np.int64)
