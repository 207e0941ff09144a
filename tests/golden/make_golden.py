"""Writes tests/golden/reference_vectors.json from the compiled reference
(oracle/_ref/libucmc_ref.so, built by `make -C oracle ref` from the reference's
own headers).  Every value is produced by the reference's code: its Pcg32, its
builders, its models' energy_diff_at / energy_at, its proposals, its drivers,
and (for the keyed replays) its per-site update applied in colour order.

Run in the container that holds the reference sources:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import (checkerboard, cubic_edges, ref, square_edges)  # noqa: E402


def ref_spins(n, seed, stream):
    u = np.zeros(n)
    ref().ref_pcg32_block(seed, stream, n, u)
    return np.where(u < 0.5, -1, 1).astype(np.int8)


def ref_unit(N, n, seed, stream):
    out = np.zeros(n * N)
    ref().ref_sphere(N, seed, stream, n, out)
    return out


def main():
    r = ref()
    g = {}

    # F1: Pcg32 streams (rng.hpp)
    g["pcg32"] = []
    for seed, stream, n in [(42, 54, 6), (2026, 1, 8), (0, 0, 4), (7, 3, 5), (2**63 + 5, 2**40, 4)]:
        u = np.zeros(n)
        r.ref_pcg32_block(seed, stream, n, u)
        g["pcg32"].append({"seed": seed, "stream": stream,
                           "u32": [int(round(x * 2**32)) for x in u]})

    # F5: table layouts and builder errors (lattice/*.hpp)
    g["square"] = []
    for Lx, Ly in [(4, 4), (2, 2), (1, 4), (1, 1), (3, 5), (2, 6)]:
        total = r.ref_square_lattice(Lx, Ly, None, None)
        off = np.zeros(Lx * Ly + 1, np.uint32)
        b = np.zeros(max(total, 1), np.uint32)
        r.ref_square_lattice(Lx, Ly, off, b)
        g["square"].append({"Lx": Lx, "Ly": Ly, "offsets": off.tolist(), "bonds": b[:total].tolist()})
    g["square_errors"] = [{"Lx": 0, "Ly": 3, "code": int(r.ref_square_lattice(0, 3, None, None))},
                          {"Lx": 3, "Ly": 0, "code": int(r.ref_square_lattice(3, 0, None, None))}]
    g["pair_graph"] = []
    for n, ei, ej, sym in [(2, [0, 0], [1, 1], 1), (3, [], [], 1), (2, [0], [1], 0), (2, [0], [0], 1),
                          (2, [0], [2], 1), (4, [0, 1, 2, 3, 0], [1, 2, 3, 0, 2], 1)]:
        ei_a = np.array(ei, np.uint32)
        ej_a = np.array(ej, np.uint32)
        off = np.zeros(n + 1, np.uint32)
        b = np.zeros(max(2 * len(ei), 1), np.uint32)
        code = r.ref_pair_graph(n, len(ei), ei_a, ej_a, sym, off, b)
        g["pair_graph"].append({"n": n, "ei": ei, "ej": ej, "symmetric": sym, "result": int(code),
                                "offsets": off.tolist() if code >= 0 else None,
                                "bonds": b[:max(code, 0)].tolist() if code >= 0 else None})
    L = 2
    nc = 6 * L ** 4
    cells = np.zeros(4 * nc, np.uint32)
    inc_off = np.zeros(4 * L ** 4 + 1, np.uint64)
    inc_cell = np.zeros(4 * nc, np.uint32)
    inc_which = np.zeros(4 * nc, np.uint8)
    r.ref_hypercubic(L, cells, inc_off, inc_cell, inc_which)
    g["hypercubic2"] = {"cells": cells.tolist(), "inc_off": inc_off.tolist(),
                        "inc_cell": inc_cell.tolist(), "inc_which": inc_which.tolist()}

    # F2: Ising energy_diff_at / energy_at on a random 8x8 (example/ising.hpp)
    ei, ej = square_edges(8, 8)
    s = ref_spins(64, 3, 3)
    dE = np.zeros(64, np.int32)
    ea = np.zeros(64, np.int32)
    r.ref_ising_eval(64, len(ei), ei, ej, s, 1, dE, ea)
    g["ising_eval"] = {"L": 8, "J": 1, "spins": s.tolist(), "dE_flip": dE.tolist(), "energy_at": ea.tolist()}

    # F3: proposals (dof/*.hpp)
    out = np.zeros(20 * 3)
    r.ref_local_rotation(3, 0.5, np.array([1.0, 0.0, 0.0]), 99, 3, 20, out)
    g["local_rotation3"] = {"delta": 0.5, "seed": 99, "stream": 3, "cur": [1.0, 0.0, 0.0], "out": out.tolist()}
    out = np.zeros(10 * 2)
    r.ref_local_rotation(2, 1.0, np.array([0.0, 1.0]), 5, 2, 10, out)
    g["local_rotation2"] = {"delta": 1.0, "seed": 5, "stream": 2, "cur": [0.0, 1.0], "out": out.tolist()}
    g["sphere"] = {}
    for N, cnt in [(2, 8), (3, 8), (4, 4)]:
        out = np.zeros(cnt * N)
        r.ref_sphere(N, 10 + N, 1, cnt, out)
        g["sphere"][str(N)] = {"seed": 10 + N, "stream": 1, "out": out.tolist()}
    out = np.zeros(10)
    r.ref_u1_shift(1.0, 0.3, 4, 4, 10, out)
    g["u1_shift"] = {"width": 1.0, "cur": 0.3, "seed": 4, "stream": 4, "out": out.tolist()}
    ident3 = np.zeros(18)
    ident3[[0, 8, 16]] = 1.0
    out = np.zeros(3 * 18)
    r.ref_sun_shift(3, 0.2, ident3, 12, 2, 3, out)
    g["su3_shift"] = {"width": 0.2, "seed": 12, "stream": 2, "out": out.tolist()}
    ident2 = np.zeros(8)
    ident2[[0, 6]] = 1.0
    out = np.zeros(3 * 8)
    r.ref_sun_shift(2, 0.7, ident2, 13, 2, 3, out)
    g["su2_shift"] = {"width": 0.7, "seed": 13, "stream": 2, "out": out.tolist()}
    a = np.array([0.0, 0.31, 0.2, -0.1, 0.05, 0.4, -0.2, -0.1, 0.0, -0.7, 0.3, 0.2, -0.05, 0.4,
                  -0.3, 0.2, 0.0, 0.39])
    out = np.zeros(18)
    r.ref_sun_expm(3, a, out)
    g["su3_expm"] = {"a": a.tolist(), "out": out.tolist()}

    # F4: float dE (example/vector_spin.hpp, u1_gauge.hpp, sun_gauge.hpp)
    ei, ej = square_edges(4, 4)
    sp = ref_unit(3, 16, 8, 5)
    props = ref_unit(3, 16, 8, 6)
    dE = np.zeros(16)
    ea = np.zeros(16)
    r.ref_vec_eval(3, 16, len(ei), ei, ej, sp, 1.0, props, dE, ea)
    g["vec3_eval"] = {"L": 4, "J": 1.0, "spins": sp.tolist(), "props": props.tolist(),
                      "dE": dE.tolist(), "energy_at": ea.tolist()}
    for L in (2, 3):
        n = 4 * L ** 4
        th = np.zeros(n)
        r.ref_pcg32_block(3, 4, n, th)
        th = (th * 2 - 1) * 3.0
        cand = np.zeros(n)
        r.ref_pcg32_block(5, 4, n, cand)
        cand = (cand * 2 - 1) * 3.0
        dE = np.zeros(n)
        mp = np.zeros(1)
        r.ref_u1_eval(L, 1.3, th, cand, dE, mp)
        g[f"u1_eval_L{L}"] = {"L": L, "beta": 1.3, "theta": th.tolist(), "cand": cand.tolist(),
                              "dE": dE.tolist(), "mean_plaquette": float(mp[0])}
    L = 2
    n = 4 * L ** 4
    links = np.zeros(n * 18)
    cands = np.zeros(n * 18)
    for i in range(n):
        o = np.zeros(18)
        r.ref_sun_shift(3, 0.9, ident3, 100 + i, 1, 1, o)
        links[i * 18:(i + 1) * 18] = o
        r.ref_sun_shift(3, 0.4, o, 500 + i, 1, 1, o)
        cands[i * 18:(i + 1) * 18] = o
    dE = np.zeros(n)
    mp = np.zeros(1)
    te = np.zeros(1)
    r.ref_sun_eval(3, L, 1.3, links, cands, dE, mp, te)
    g["su3_eval"] = {"L": 2, "beta": 1.3, "links": links.tolist(), "cands": cands.tolist(),
                     "dE": dE.tolist(), "mean_plaquette": float(mp[0]), "total_energy": float(te[0])}

    # F6: keyed colour-order replays through the reference's own per-site update
    ei, ej = square_edges(16, 16)
    s = ref_spins(256, 21, 0)
    acc = np.zeros(5, np.uint64)
    col = checkerboard((16, 16))
    s2 = s.copy()
    r.ref_ising_keyed(256, len(ei), ei, ej, s2, 1, 0.44, col, 2, 1234, 7, 5, acc)
    g["ising_keyed"] = {"L": 16, "beta": 0.44, "J": 1, "run_seed": 1234, "sweep0": 7, "spins": s.tolist(),
                        "final": s2.tolist(), "accepts": [int(x) for x in acc]}
    ei, ej = cubic_edges(4)
    sp = ref_unit(3, 64, 9, 1)
    col = checkerboard((4, 4, 4))
    sp2 = sp.copy()
    acc = np.zeros(2, np.uint64)
    tr = np.zeros(2 * 64)
    r.ref_vec_keyed(3, 64, len(ei), ei, ej, sp2, 1.0, 0.693, 1, 0.5, col, 2, 77, 0, 2, acc, tr)
    g["vec3_keyed"] = {"L": 4, "beta": 0.693, "delta": 0.5, "run_seed": 77, "spins": sp.tolist(),
                       "final": sp2.tolist(), "accepts": [int(x) for x in acc], "trace_dE": tr.tolist()}
    L = 2
    n = 4 * L ** 4
    th = np.zeros(n)
    r.ref_pcg32_block(8, 1, n, th)
    th = (th * 2 - 1) * 3.0
    th2 = th.copy()
    acc = np.zeros(2, np.uint64)
    r.ref_u1_keyed(L, th2, 0.9, 1.0, 1.0, 5, 0, 2, acc, None)
    g["u1_keyed"] = {"L": 2, "model_beta": 0.9, "driver_beta": 1.0, "width": 1.0, "run_seed": 5,
                     "theta": th.tolist(), "final": th2.tolist(), "accepts": [int(x) for x in acc]}
    l2 = links.copy()
    acc = np.zeros(1, np.uint64)
    r.ref_sun_keyed(3, L, l2, 5.7, 1.0, 0.2, 4, 3, 0, 1, acc)
    g["su3_keyed"] = {"L": 2, "model_beta": 5.7, "driver_beta": 1.0, "width": 0.2, "hits": 4,
                      "run_seed": 3, "links": links.tolist(), "final": l2.tolist(),
                      "accepts": [int(x) for x in acc]}

    # F7: the reference's own chains (test/model_check.cpp:171-222, u1_check, su_check)
    ei, ej = square_edges(4, 4)
    s = np.ones(16, np.int8)
    e = np.zeros(4000)
    r.ref_ising_metropolis_seq(16, len(ei), ei, ej, s, 1, 0.3, 11, 1, 2000, 4000, e, None, None)
    g["chain_ising_metropolis"] = {"mean_e_per_site": float(e.mean() / 16), "final": s.tolist(),
                                   "e_head": e[:10].tolist()}
    s = np.ones(16, np.int8)
    e = np.zeros(20000)
    sz = np.zeros(20000, np.uint64)
    mc = np.zeros(1)
    r.ref_ising_wolff_seq(16, len(ei), ei, ej, s, 1, 0.3, 12, 1, 2000, 20000, e, sz, mc)
    g["chain_ising_wolff"] = {"mean_e_per_site": float(e.mean() / 16), "mean_cluster": float(mc[0]),
                              "final": s.tolist(), "sizes_head": [int(x) for x in sz[:20]]}
    sp = np.zeros(48)
    sp[2::3] = 1.0
    e = np.zeros(4000)
    r.ref_vec_metropolis_seq(3, 16, len(ei), ei, ej, sp, 1.0, 0.3, 0, 0.0, 13, 1, 2000, 4000, e)
    g["chain_heis_metropolis"] = {"mean_e_per_site": float(e.mean() / 16), "final": sp.tolist()}
    sp = np.zeros(48)
    sp[2::3] = 1.0
    e = np.zeros(20000)
    sz = np.zeros(20000, np.uint64)
    r.ref_vec_wolff_seq(3, 16, len(ei), ei, ej, sp, 1.0, 0.3, 14, 1, 2000, 20000, e, sz)
    g["chain_heis_wolff"] = {"mean_e_per_site": float(e.mean() / 16), "final": sp.tolist()}
    L = 2
    th = np.zeros(4 * L ** 4)
    pl = np.zeros(50)
    r.ref_u1_metropolis_seq(L, th, 0.5, 0.5, 0.7, 60, 2, 20, 50, pl)
    g["chain_u1"] = {"L": 2, "beta": 0.5, "width": 0.7, "seed": 60, "plaquettes": pl.tolist()}
    lk = np.zeros(4 * L ** 4 * 18)
    for i in range(4 * L ** 4):
        lk[i * 18:(i + 1) * 18] = ident3
    pl = np.zeros(10)
    r.ref_sun_metropolis_seq(3, L, lk, 4.0, 4.0, 0.3, 90, 2, 5, 10, pl)
    g["chain_su3"] = {"L": 2, "beta": 4.0, "width": 0.3, "seed": 90, "plaquettes": pl.tolist()}

    path = os.path.join(HERE, "reference_vectors.json")
    with open(path, "w") as f:
        json.dump(g, f)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
