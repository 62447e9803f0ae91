"""Checkpoint / restart around the library's output in the reference's
SCF-state records: oracle/scfstate.py restates xc_hamiltonian_store /
xc_hamiltonian_recover (ParaGauss modules/xc_hamiltonian.f90:2407-2634) over
readwriteblocked_module's blocked file format, and the tests write real
files with it.

  * the packed V_xc and E_xc of a build survive store / recover bit for bit,
    closed and open shell;
  * the library's state written after those records and loaded into a new
    builder of the same geometry lets the next incremental builds give the
    bits of the uninterrupted SCF;
  * recover's spin conversions: open-shell records into a closed-shell run
    average the two spins, closed-shell records into an open-shell run copy
    the one spin into both.
"""
import os

import numpy as np
import pytest

from oracle import scfstate
from pgxc import XCBuilder, basis, grid, systems
from pgxc.xc import BF16X3, FP64, PBE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    els, xyz = systems.h2o()
    b = basis.seeded_basis(els, xyz, seed=1, shell_set="small")
    g = grid.molecular_grid(els, xyz, nrad=40, nang=302)
    C, occ = systems.random_orbitals(b.nbas, 5, seed=2)
    return b, g, C, occ


def write_tape(path, ham, exc, n_spin, state):
    t = scfstate.BlockedTape(str(path), "w")
    scfstate.store(t, ham, exc, n_spin)
    scfstate.put_state(t, state)
    t.close()
    # whole blocks of 128 doubles, each one sequential record
    assert os.path.getsize(path) % (8 * scfstate.LENBLK + 8) == 0


def read_tape(path, n_spin, npk):
    t = scfstate.BlockedTape(str(path), "r")
    ham, exc, init = scfstate.recover(t, n_spin, npk)
    state = scfstate.get_state(t)
    t.close()
    return ham, exc, init, state


def same(a, b):
    return np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]


def test_closed_shell_restart_continues_bitwise(case, tmp_path):
    b, g, C, occ = case
    npk = b.nbas * (b.nbas + 1) // 2
    chain = [systems.perturbed_orbitals(C, 1e-4, seed=30 + k) for k in range(2)]
    tape = tmp_path / "scfstate"
    with XCBuilder(b, g) as xc:  # the uninterrupted SCF
        xc.set_orbitals(C, occ)
        ham0, e0, q0 = xc.build(PBE, FP64)
        write_tape(tape, ham0, e0, 1, xc.save_state())
        ref = []
        for Ck in chain:
            xc.set_orbitals(Ck, occ, incremental=True)
            ref.append(xc.build(PBE, BF16X3))
    ham, exc, init, state = read_tape(tape, 1, npk)
    assert np.array_equal(ham, ham0) and exc == e0 and not init
    with XCBuilder(b, g) as xc:  # the restarted SCF
        xc.load_state(state)
        for Ck, r in zip(chain, ref):
            xc.set_orbitals(Ck, occ, incremental=True)
            assert same(xc.build(PBE, BF16X3), r)


def test_open_shell_records_and_spin_conversions(case, tmp_path):
    b, g, C, occ = case
    npk = b.nbas * (b.nbas + 1) // 2
    Cb = systems.perturbed_orbitals(C, 1e-2, seed=21)
    Cs = np.stack([C, Cb], axis=2)
    occs = np.stack([occ, np.concatenate([occ[:-1], [0.0]])], axis=1)
    tape = tmp_path / "scfstate_open"
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(Cs, occs)
        ham2, e2, _ = xc.build(PBE, FP64)
        write_tape(tape, ham2, e2, 2, xc.save_state())
        C1 = np.stack([systems.perturbed_orbitals(Cs[:, :, s], 1e-4, seed=40 + s) for s in range(2)],
                      axis=2)
        xc.set_orbitals(C1, occs, incremental=True)
        ref = xc.build(PBE, FP64)
    # the same spin count: bit for bit, and the state continues the SCF
    ham, exc, init, state = read_tape(tape, 2, npk)
    assert np.array_equal(ham, ham2) and exc == e2 and not init
    with XCBuilder(b, g) as xc:
        xc.load_state(state)
        xc.set_orbitals(C1, occs, incremental=True)
        assert same(xc.build(PBE, FP64), ref)
    # polarized records into a restricted run: the mean of the two spins
    ham, exc, init, _ = read_tape(tape, 1, npk)
    assert np.array_equal(ham, (ham2[0::2] + ham2[1::2]) * 0.5) and exc == e2
    # restricted records into a polarized run: one spin copied into both
    tape1 = tmp_path / "scfstate_closed"
    with XCBuilder(b, g) as xc:
        xc.set_orbitals(C, occ)
        ham1, e1, _ = xc.build(PBE, FP64)
        write_tape(tape1, ham1, e1, 1, xc.save_state())
    ham, exc, init, _ = read_tape(tape1, 2, npk)
    assert np.array_equal(ham[0::2], ham1) and np.array_equal(ham[1::2], ham1)
    assert exc == e1 and init
