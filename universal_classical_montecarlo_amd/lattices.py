"""Synthetic inputs for the benchmarks, built with numpy: pair tables of the
periodic simple-cubic lattice (the form the library recognises as the cubic
stencil: every row holds the six periodic neighbours), the plaquette cells of a
periodic square lattice, the 4-spin plaquette term tabulated for
ucmc_ising_cells_create, and random +-1 / unit-vector fields."""
from __future__ import annotations

import numpy as np


def cubic_csr(L):
    """CSR rows (+x, +y, +z, -x, -y, -z) of the periodic L^3 lattice, site
    id (z L + y) L + x (row 0 starts 1, L: the library's cubic detection reads
    L there)."""
    i = np.arange(L ** 3, dtype=np.int64)
    x, y, z = i % L, (i // L) % L, i // (L * L)
    def site(xx, yy, zz):
        return ((zz % L) * L + (yy % L)) * L + (xx % L)
    nb = np.stack([site(x + 1, y, z), site(x, y + 1, z), site(x, y, z + 1), site(x - 1, y, z),
                   site(x, y - 1, z), site(x, y, z - 1)], axis=1)
    off = (np.arange(L ** 3 + 1, dtype=np.int64) * 6).astype(np.uint32)
    return off, np.ascontiguousarray(nb.reshape(-1).astype(np.uint32))


def square_plaquettes(Lx, Ly):
    """Cells (x, y), (x+1, y), (x+1, y+1), (x, y+1) of the periodic Lx x Ly
    lattice, site id y Lx + x, one per site."""
    i = np.arange(Lx * Ly, dtype=np.int64)
    x, y = i % Lx, i // Lx
    xp, yp = (x + 1) % Lx, (y + 1) % Ly
    return np.ascontiguousarray(np.stack([y * Lx + x, y * Lx + xp, yp * Lx + xp, yp * Lx + x],
                                         axis=1).astype(np.uint32))


def four_spin_term(K=1.0):
    """E = -K s0 s1 s2 s3 tabulated as ucmc_ising_cells wants it: energy[bits]
    and delta[which, bits] = energy[bits ^ (1 << which)] - energy[bits] (bit k
    set when link k holds -1)."""
    bits = np.arange(16)
    parity = np.array([bin(b).count("1") & 1 for b in bits])
    energy = np.where(parity, K, -K).astype(np.float64)
    delta = np.array([[energy[b ^ (1 << w)] - energy[b] for b in bits] for w in range(4)], np.float64)
    return delta, energy


def random_ising(n, seed):
    return np.where(np.random.default_rng(seed).random(n) < 0.5, -1, 1).astype(np.int8)


def random_unit_vectors(n, N, seed):
    x = np.random.default_rng(seed).normal(size=(n, N))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return np.ascontiguousarray(x.reshape(-1))
