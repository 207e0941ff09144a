"""Headline benchmark: 2D Ising L=4096 Metropolis (checkerboard, both colours
fused in one pass) on one MI355X, site-updates/s and fraction of the HBM
roofline, with the reference's own sequential CPU driver timed beside it.

One "step" is one full Metropolis sweep (every site updated once) over the
resident int8 lattice.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "site-updates/s/GPU + %HBM roofline; SLMC vs plain indep. samples/s (τ_int)"
HBM_PEAK_GBS = 8000.0
BYTES_PER_SITE = 2.0  # int8 read + int8 write per site-update


def traffic_from_profile():
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), if the summary has been committed."""
    p = os.path.join(ROOT, "profiles", "ising_fused_4096_traffic.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(L, seconds_target=12.0):
    """The reference's Metropolis<Ising, Flip, Pcg32>::sweep (compiled from its
    own headers into oracle/_ref) on one host core, over a bounded number of
    sweeps of the same L x L lattice.  Falls back to the oracle restatement."""
    from oracle_lib import REF_SO, oracle, ref, square_edges, csr_square
    from universal_classical_montecarlo_amd import lattices
    n = L * L
    spins = lattices.random_ising(n, 7)
    kind = "reference"
    sweeps = 1
    if os.path.exists(REF_SO):
        r = ref()
        ei, ej = square_edges(L, L)
        def run(k):
            s = spins.copy()
            r.ref_ising_metropolis_seq(n, len(ei), ei, ej, s, 1, 0.44, 2026, 1, k, 0, None, None, None)
    else:
        kind = "port"
        o = oracle()
        off, bonds = csr_square(L, L)
        def run(k):
            s = spins.copy()
            o.orc_ising_metropolis_seq(n, off, bonds, s, 1, 0.44, 2026, 1, k, 0, None, None, None)
    t0 = time.perf_counter()
    run(1)
    t1 = time.perf_counter() - t0
    sweeps = max(1, int(seconds_target / max(t1, 1e-6)))
    t0 = time.perf_counter()
    run(sweeps)
    dt = time.perf_counter() - t0
    return {"value": n * sweeps / dt, "unit": "site-updates/s", "cores": 1, "kind": kind,
            "sample": f"{sweeps} sequential sweeps of the {L}x{L} lattice at beta=0.44 "
                      f"(Metropolis<Ising,Flip,Pcg32>, one thread, {platform.processor() or 'host CPU'})",
            "seconds": dt}


def cpu_baseline_all_cores(L, threads=None, seconds_target=6.0):
    """The same compiled reference chain on every host core of this GPU's share
    (at most 16): independent chains, one per thread, each on its own copy of
    an L x L lattice with its own stream, aggregated (the ctypes call releases
    the GIL).  Reported beside the one-core baseline."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle_lib import REF_SO, ref, square_edges
    from universal_classical_montecarlo_amd import lattices
    if not os.path.exists(REF_SO):
        return None
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    n = L * L
    r = ref()
    ei, ej = square_edges(L, L)
    spins = lattices.random_ising(n, 7)
    def run(args):
        stream, k = args
        s = spins.copy()
        r.ref_ising_metropolis_seq(n, len(ei), ei, ej, s, 1, 0.44, 2026, stream, k, 0, None, None, None)
    t0 = time.perf_counter()
    run((1, 1))
    sweeps = max(1, int(seconds_target / max(time.perf_counter() - t0, 1e-6)))
    with ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(run, [(1 + t, sweeps) for t in range(threads)]))
        dt = time.perf_counter() - t0
    return {"value": threads * n * sweeps / dt, "unit": "site-updates/s", "cores": threads,
            "kind": "reference",
            "sample": f"{threads} independent chains x {sweeps} sweeps of the {L}x{L} lattice, "
                      "one thread each"}


def expm_products(width, samples=200000, seed=1):
    """Mean number of 3x3 complex products in the device expm of an SUShift<3>
    generator (Horner terms m plus squarings s, with the kernel's rule: halve
    the inf-norm n to <= 0.5, m = 4..8 by n), over generators drawn as the
    proposal draws them."""
    rng = np.random.default_rng(seed)
    d = width * (2 * rng.random((samples, 3)) - 1)
    d -= d.mean(axis=1, keepdims=True)
    ab = width * (2 * rng.random((samples, 3, 2)) - 1)
    mod = np.hypot(ab[..., 0], ab[..., 1])   # |x_01|, |x_02|, |x_12|
    rows = np.stack([np.abs(d[:, 0]) + mod[:, 0] + mod[:, 1],
                     np.abs(d[:, 1]) + mod[:, 0] + mod[:, 2],
                     np.abs(d[:, 2]) + mod[:, 1] + mod[:, 2]], axis=1)
    n = rows.max(axis=1)
    s = np.zeros(samples)
    while (n > 0.5).any():
        big = n > 0.5
        n = np.where(big, n * 0.5, n)
        s += big
    m = np.select([n <= 0.08, n <= 0.16, n <= 0.28, n <= 0.43], [4, 5, 6, 7], 8)
    return float((m + s).mean())


def suite(quick=False):
    """Throughput of the other workloads (one GPU): Heisenberg 128^3, Wolff,
    SU(3) 16^4 / 24^4, the 256-replica XY batch and SLMC, each bounded
    to a few seconds; reported beside the headline, not as `value`."""
    from universal_classical_montecarlo_amd import abi, lattices
    out = {}
    # C2b: 3D Heisenberg 128^3, LocalRotation(0.5), beta = 0.693, 2-colour
    L = 64 if quick else 128
    off, bonds = lattices.cubic_csr(L)
    v = abi.System.vector(3, off, bonds, 1.0)
    v.upload(lattices.random_unit_vectors(L ** 3, 3, 5))
    v.metropolis(0.693, 5, 1, proposal=abi.PROP_ROTATION, width=0.5)
    ms, n = v.bench(0.693, 50, 2, proposal=abi.PROP_ROTATION, width=0.5)
    rate = L ** 3 * 50 / (ms * 1e-3)
    out["heisenberg_128_metropolis"] = {"site_updates_per_s": rate,
                                        "hbm_frac_24B": 24 * rate / 1e9 / HBM_PEAK_GBS}
    # Wolff reflection clusters (label propagation at this size): 2000 Metropolis
    # sweeps bring the random start close to equilibrium at beta = 0.693 (~T_c),
    # then 500 untimed flips
    v.metropolis(0.693, 2000, 4, proposal=abi.PROP_ROTATION, width=0.5)
    v.wolff(0.693, 500, 3)
    t0 = time.perf_counter()
    sizes, _ = v.wolff(0.693, 500, 3, flip0=500)
    dt = time.perf_counter() - t0
    out["heisenberg_128_wolff"] = {"flipped_sites_per_s": float(sizes.sum()) / dt,
                                   "flips_per_s": 500 / dt, "mean_cluster": float(sizes.mean())}
    v.close()
    # Wolff on the 2D Ising lattice of the headline (tiled label path), beta = 0.44
    L = 1024 if quick else 4096
    w = abi.System.ising_square(L, L, 1)   # the reference's initial state, all +1
    w.metropolis(0.44, 200, 5)
    w.wolff(0.44, 50, 6)
    t0 = time.perf_counter()
    sizes, _ = w.wolff(0.44, 100, 6, flip0=50)
    dt = time.perf_counter() - t0
    out[f"ising_{L}_wolff"] = {"flipped_sites_per_s": float(sizes.sum()) / dt, "flips_per_s": 100 / dt,
                               "mean_cluster": float(sizes.mean())}
    w.close()
    # C3: SU(3) 24^4 (16^4 quick), beta = 5.7 set once (model), SUShift 0.2, 4 hits
    for L in ((16,) if quick else (16, 24)):
        g = abi.System.sun_gauge(3, L, 5.7)
        g.metropolis(1.0, 2, 1, proposal=abi.PROP_SU_SHIFT, width=0.2, hits=4)
        ms, n = g.bench(1.0, 10, 2, proposal=abi.PROP_SU_SHIFT, width=0.2, hits=4,
                        reunitarize_every=10)
        links = 4 * L ** 4
        rate = links * 10 / (ms * 1e-3)
        o = g.observe()
        flop = 2466 + 4 * (198 + 36 + 198 * expm_products(0.2))
        out[f"su3_{L}^4_4hit"] = {"link_updates_per_s": rate, "plaquette": o.plaquette,
                                 "flop_per_link": flop, "tflops": rate * flop / 1e12,
                                 "fp32_fma_peak_frac": rate * flop / 78.6e12}
        g.close()
    # C4: 256 XY replicas under parallel tempering
    Ls, ladders, betas = [], [], []
    grid = 0.8 + 0.6 * np.arange(64) / 63
    for ld, (Lr, count) in enumerate([(32, 64), (64, 64), (128, 64), (256, 32), (512, 32)]):
        bs = grid if count == 64 else grid[::2]
        Ls += [Lr] * count
        ladders += [ld] * count
        betas += list(bs)
    pt = abi.System.pt(Ls, ladders, betas)
    tile, _ = pt.pt_autotune(1.0)
    pt.pt_run(1.0, 10, 1, use_graph=True)
    t0 = time.perf_counter()
    pt.pt_run(1.0, 100, 2, sweep0=10, use_graph=True)
    dt = time.perf_counter() - t0
    acc, sacc, stry = pt.pt_stats()
    out["xy_pt_256_replicas"] = {"site_updates_per_s": pt.sites * 100 / dt, "tile": tile,
                                 "hbm_frac_8B": 8 * pt.sites * 100 / dt / 1e9 / HBM_PEAK_GBS,
                                 "swap_acceptance": float(sacc.sum() / max(stry.sum(), 1))}
    pt.close()
    # A user cell term through the generic table-driven path: the 4-spin plaquette
    # term E = -K s0 s1 s2 s3 (K = 1) as CellModel<IsingDof, Term, CellComplex<4>>
    L = 256 if quick else 1024
    delta, energy = lattices.four_spin_term(1.0)
    c = abi.System.ising_cells(L * L, lattices.square_plaquettes(L, L), delta, energy)
    c.metropolis(0.5, 5, 1)
    ms, n = c.bench(0.5, 50, 2)
    out[f"ising_cells_plaquette_{L}"] = {"site_updates_per_s": L * L * 50 / (ms * 1e-3),
                                         "colours": int(c.info.ncolours)}
    c.close()
    # C5: SLMC vs plain GPU Metropolis, independent samples / s
    from universal_classical_montecarlo_amd import slmc
    out["slmc"] = slmc.compare(L=64 if quick else 256, nconfigs=4000 if quick else 20000,
                               plain_sweeps=20000 if quick else 200000,
                               slmc_steps=2000 if quick else 100000)
    return out


def profile_others():
    """Each of the other configurations at full size for a few steps, for
    rocprofv3 --kernel-trace (kernel durations of K2-K9)."""
    from universal_classical_montecarlo_amd import abi, lattices, slmc
    L = 128
    off, bonds = lattices.cubic_csr(L)
    v = abi.System.vector(3, off, bonds, 1.0)
    v.upload(lattices.random_unit_vectors(L ** 3, 3, 5))
    v.metropolis(0.693, 20, 1, proposal=abi.PROP_ROTATION, width=0.5)
    v.wolff(0.693, 50, 3)
    v.close()
    w = abi.System.ising_square(4096, 4096, 1)
    w.upload(lattices.random_ising(4096 * 4096, 7))
    w.metropolis(0.44, 20, 5)
    w.wolff(0.44, 20, 6)
    w.close()
    g = abi.System.sun_gauge(3, 24, 5.7)
    g.metropolis(1.0, 2, 1, proposal=abi.PROP_SU_SHIFT, width=0.2, hits=4)
    g.reunitarize()
    g.close()
    Ls, ladders, betas = [], [], []
    grid = 0.8 + 0.6 * np.arange(64) / 63
    for ld, (Lr, count) in enumerate([(32, 64), (64, 64), (128, 64), (256, 32), (512, 32)]):
        Ls += [Lr] * count
        ladders += [ld] * count
        betas += list(grid if count == 64 else grid[::2])
    pt = abi.System.pt(Ls, ladders, betas)
    pt.pt_run(1.0, 20, 1, tile=64, use_graph=True)
    pt.close()
    f = abi.System.plaquette_ising(256, 1.0, 0.2)
    _, J, _, _ = slmc.fit_effective_model(f, 0.40, burn=500, nconfigs=200, every=1, run_seed=3)
    f.slmc_run(J, 0.40, 200, 4, record=False)
    f.metropolis(0.40, 500, 5)
    f.close()
    print(json.dumps({"profile": "other_configs"}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--L", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-suite", action="store_true")
    ap.add_argument("--quick-suite", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only the headline kernel and a same-size copy kernel, for rocprofv3 --pmc")
    ap.add_argument("--profile-others", action="store_true",
                    help="the other configurations for a few steps each, for rocprofv3 --kernel-trace")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X, and the lattice "
              "of this workload is one GPU's share; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from universal_classical_montecarlo_amd import abi, lattices
    L = args.L
    if args.profile_others:
        profile_others()
        return
    if args.profile:
        # 50 fused sweeps, then 50 copies of the same byte count (read L^2 + write L^2)
        s = abi.System.ising_square(L, L, 1)
        s.upload(lattices.random_ising(L * L, 7))
        s.metropolis(0.44, 50, 2026)
        s.close()
        print(json.dumps({"profile": "ising_fused", "L": L,
                          "copy_GBps": abi.measure_copy(L * L, 50)}), flush=True)
        return
    sysm = abi.System.ising_square(L, L, 1)
    assert sysm.info.schedule == abi.SCHED_FUSED
    sysm.upload(lattices.random_ising(L * L, 7))
    beta = 0.44
    # warmup (untimed)
    sysm.metropolis(beta, args.warmup, 2026, sweep0=0)
    # timed region: exactly `steps` sweeps, HIP events on the library's stream
    ms, launches = sysm.bench(beta, args.steps, 2026, repeats=1)
    sites = L * L
    kernel_s = ms * 1e-3 / launches
    value = sites * args.steps / (ms * 1e-3)
    achieved = BYTES_PER_SITE * sites / kernel_s / 1e9
    obs = sysm.observe()
    sysm.close()
    # the measured ceiling: a plain device copy of the same bytes (L^2 read + L^2 written)
    copy_gbs = abi.measure_copy(sites, 50)

    # out-of-cache size (16384^2 = 268 MB int8) for the HBM-bound regime
    ooc = None
    if L == 4096:
        L2 = 16384
        s2 = abi.System.ising_square(L2, L2, 1)
        s2.upload(np.where(np.random.default_rng(7).random(L2 * L2) < 0.5, -1, 1).astype(np.int8))
        s2.metropolis(beta, 5, 11)
        ms2, n2 = s2.bench(beta, 100, 11)
        gbs2 = BYTES_PER_SITE * L2 * L2 / (ms2 * 1e-3 / n2) / 1e9
        s2.close()
        copy2 = abi.measure_copy(L2 * L2, 20)
        ooc = {"L": L2, "site_updates_per_s": L2 * L2 * 100 / (ms2 * 1e-3), "achieved_GBps": gbs2,
               "hbm_frac": gbs2 / HBM_PEAK_GBS, "copy_GBps": copy2, "frac_of_copy": gbs2 / copy2}

    cpu = None if args.no_cpu else cpu_baseline(L)
    cpu_all = None if args.no_cpu else cpu_baseline_all_cores(1024)
    extra = None if args.no_suite else suite(quick=args.quick_suite)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": f"2D Ising square periodic L={L}, J=1, beta={beta}, Flip proposal, "
                               "checkerboard Metropolis with both colours fused (2 B/site)",
                   "L": L, "beta": beta, "sites": sites, "energy_per_site_end": obs.energy / sites},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(),
                     "kernel": "ising_fused", "kernel_avg_us": kernel_s * 1e6,
                     "bytes_per_site": BYTES_PER_SITE, "copy_GBps": copy_gbs,
                     "frac_of_copy": achieved / copy_gbs},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in
                                                  ("value", "unit", "cores", "kind", "sample")},
        "cpu_baseline_all_cores": cpu_all,
        "out_of_cache": ooc,
        "other_configs": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
