"""Tabulated Ising cell terms on the GPU (CellModel<IsingDof, Term, List>,
model/cell.hpp:16-99): the device chain against the oracle's keyed replay, bit
for bit (fields and accept counts), for cell arities 1 to 4, random cell lists
(sites may repeat inside a cell) and random tables; the arity-2 IsingTerm case
against the Ising model itself; observables and traced dE against the oracle."""
import numpy as np
import pytest

from universal_classical_montecarlo_amd import abi
from oracle_lib import (cell_values, cells_colouring, checkerboard, csr_square, delta_from_energy,
                        ising_pair_term, oracle, pair_cells, random_spins,
                        square_plaquettes)

pytestmark = pytest.mark.gpu


def _device_run(s, spins, beta, seed, sweep0, nsweeps):
    s.upload(spins)
    ser = s.metropolis(beta, nsweeps, seed, sweep0=sweep0,
                       flags=abi.RECORD_ACCEPTS | abi.RECORD_ENERGY | abi.RECORD_ORDER, record=True)
    return s.download(), np.array([o.accepted for o in ser], np.uint64), ser


def _oracle_run(n, cells, A, delta, spins, beta, col, ncol, seed, sweep0, nsweeps):
    s = spins.copy()
    acc = np.zeros(nsweeps, np.uint64)
    assert oracle().orc_ising_cells_metropolis(n, len(cells), A, cells, delta.reshape(-1), s, beta, col,
                                               ncol, seed, sweep0, nsweeps, acc) == 0
    return s, acc


@pytest.mark.parametrize("A,n,ncells,physical", [(1, 300, 200, True), (2, 500, 900, True), (3, 400, 500, False),
                                                 (4, 600, 700, True), (4, 257, 300, False)])
def test_cells_random_tables_bit_exact(gpu, A, n, ncells, physical):
    print(f"cells A={A} n={n} cells={ncells} physical={physical}", flush=True)
    rng = np.random.default_rng(100 * A + n)
    cells = rng.integers(0, n, size=(ncells, A)).astype(np.uint32)
    energy = np.round(rng.normal(size=1 << A), 3)
    # a term whose delta is its energy difference, or an arbitrary table (the
    # device must reproduce whatever delta the term reports)
    delta = delta_from_energy(energy, A) if physical else np.round(rng.normal(size=(A, 1 << A)), 3)
    s = abi.System.ising_cells(n, cells, delta, energy)
    col = s.colouring()
    want_col, ncol = cells_colouring(n, cells)
    assert np.array_equal(col, want_col)
    spins = random_spins(n, 7 + A, 1)
    got, acc, ser = _device_run(s, spins, 0.6, 21, 4, 5)
    want, wacc = _oracle_run(n, cells, A, delta, spins, 0.6, want_col, ncol, 21, 4, 5)
    assert np.array_equal(got, want)
    assert np.array_equal(acc, wacc)
    e = oracle().orc_ising_cells_energy(len(cells), A, cells, energy, want)
    assert abs(ser[-1].energy - e) <= 1e-9 * (1 + abs(e))
    assert ser[-1].order[0] == float(want.astype(np.int64).sum())
    s.close()


@pytest.mark.parametrize("Lx,Ly,beta,J", [(16, 16, 0.44, 1), (12, 20, 0.3, -1)])
def test_cells_ising_term_equals_ising(gpu, Lx, Ly, beta, J):
    """IsingTerm over the bonds as arity-2 cells: the Ising chain itself."""
    off, b = csr_square(Lx, Ly)
    n = Lx * Ly
    delta, energy = ising_pair_term(J)
    c = abi.System.ising_cells(n, pair_cells(off, b), delta, energy)
    assert np.array_equal(c.colouring(), checkerboard((Lx, Ly)))
    spins = random_spins(n, 40 + Lx, 0)
    got, acc, _ = _device_run(c, spins, beta, 9, 0, 6)
    want = spins.copy()
    wacc = np.zeros(6, np.uint64)
    oracle().orc_ising_metropolis_keyed(n, off, b, want, J, beta, checkerboard((Lx, Ly)), 2, 9, 0, 6, wacc)
    assert np.array_equal(got, want) and np.array_equal(acc, wacc)
    c.close()


def test_cells_plaquette_trace_is_finite_difference(gpu):
    """Traced device dE = E(site flipped) - E for the sites of the first colour
    class (the later classes see its updates), for a 4-spin plaquette term and a
    3-site term on the same lattice, as two cell systems."""
    L = 16
    n = L * L
    for A, cells, K in ((4, square_plaquettes(L, L), 0.7),
                        (3, square_plaquettes(L, L)[:, :3].copy(), -1.3)):
        energy = np.array([-K * np.prod(cell_values(b, A)) for b in range(1 << A)])
        delta = delta_from_energy(energy, A)
        s = abi.System.ising_cells(n, cells, delta, energy)
        spins = random_spins(n, 3, A)
        s.upload(spins)
        dE, acc = s.trace(0.5, 13)
        o = oracle()
        e0 = o.orc_ising_cells_energy(len(cells), A, cells, energy, spins)
        first = np.flatnonzero(s.colouring() == 0)
        assert len(first) > 0
        for i in first:
            f = spins.copy()
            f[i] = -f[i]
            want = o.orc_ising_cells_energy(len(cells), A, cells, energy, f) - e0
            assert abs(dE[i] - want) <= 1e-5 * (1 + abs(want)), (A, i)
        s.close()


def test_cells_errors(gpu):
    cells = np.array([[0, 1], [1, 2]], np.uint32)
    d, e = ising_pair_term(1)
    with pytest.raises(abi.UcmcError) as ex:
        abi.System.ising_cells(2, cells, d, e)             # site 2 of 2
    assert ex.value.code == abi.UCMC_E_OUT_OF_RANGE
    with pytest.raises(abi.UcmcError) as ex:
        abi.System.ising_cells(3, np.zeros((2, 5), np.uint32), np.zeros(5 << 5), np.zeros(32))
    assert ex.value.code == abi.UCMC_E_UNSUPPORTED
    bad = e.copy()
    bad[1] = np.nan
    with pytest.raises(abi.UcmcError) as ex:
        abi.System.ising_cells(3, cells, d, bad)
    assert ex.value.code == abi.UCMC_E_INVALID_ARGUMENT
    s = abi.System.ising_cells(3, cells, d, e)
    with pytest.raises(abi.UcmcError):
        s.metropolis(0.5, 1, 1, proposal=abi.PROP_SPHERE)   # Flip only
    s.close()
