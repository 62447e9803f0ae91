"""CPU restatement of ParaGauss's SCF-state records of the XC Hamiltonian.

TEST INFRASTRUCTURE ONLY: tests/ use it to round-trip the library's packed
V_xc and E_xc through the reference's checkpoint records; the product never
imports it.

  BlockedTape   readwriteblocked_module's file format
                (modules/readwriteblocked_module.f90:476-540 write, :177-276
                read, :720-770 stopwrite): a stream of doubles cut into
                blocks of lenblk doubles, each block one Fortran unformatted
                sequential record (4-byte length markers before and after);
                the last block is written whole (its tail holds what the
                buffer held before).  lenblk = machineparameters_rw_buffer,
                default 128 (modules/machineparameters_module.f90:68).
  store         xc_hamiltonian_store (modules/xc_hamiltonian.f90:2407-2486),
                numeric exchange: [1], [n_spin, xc_length],
                ham(1:xc_length:n_spin), ham(2:xc_length:2) if n_spin > 1,
                [exc_int].
  recover       xc_hamiltonian_recover (:2488-2634): the same records, with
                the spin conversions: polarized data into a restricted run
                average the spins, restricted data into a polarized run copy
                the one spin into both.
"""
import struct

import numpy as np

LENBLK = 128


class BlockedTape:
    """Write: put() items, then close(); read: get(n) items."""

    def __init__(self, path, mode, lenblk=LENBLK):
        self.path, self.mode, self.lenblk = path, mode, lenblk
        self.buf = np.zeros(lenblk)
        self.fill = 0  # write: items in buf; read: items left in buf
        self.pos = 0
        self.f = open(path, "wb" if mode == "w" else "rb")

    def _write_block(self):
        data = self.buf.astype("<f8").tobytes()
        n = struct.pack("<i", len(data))
        self.f.write(n + data + n)

    def put(self, items):
        items = np.asarray(items, dtype=np.float64).ravel()
        i = 0
        while i < len(items):
            take = min(self.lenblk - self.fill, len(items) - i)
            self.buf[self.fill:self.fill + take] = items[i:i + take]
            self.fill += take
            i += take
            if self.fill == self.lenblk:  # a full buffer goes out as one record
                self._write_block()
                self.fill = 0

    def close(self):
        if self.mode == "w" and self.fill:
            self._write_block()
        self.f.close()

    def _read_block(self):
        head = self.f.read(4)
        if len(head) < 4:
            raise EOFError("end of file " + self.path)
        (nb,) = struct.unpack("<i", head)
        data = self.f.read(nb)
        (tail,) = struct.unpack("<i", self.f.read(4))
        if tail != nb or nb != 8 * self.lenblk:
            raise ValueError("not a blocked record of %d doubles" % self.lenblk)
        self.buf = np.frombuffer(data, dtype="<f8").copy()
        self.pos, self.fill = 0, self.lenblk

    def get(self, n):
        out = np.empty(n)
        i = 0
        while i < n:
            if self.fill == 0:
                self._read_block()
            take = min(self.fill, n - i)
            out[i:i + take] = self.buf[self.pos:self.pos + take]
            self.pos += take
            self.fill -= take
            i += take
        return out


def store(tape, ham, exc, n_spin):
    """ham: packed V_xc of xc_length = n_spin * n(n+1)/2, spins interleaved."""
    ham = np.asarray(ham, dtype=np.float64)
    xc_length = ham.size
    tape.put([1.0])
    tape.put([float(n_spin), float(xc_length)])
    tape.put(ham[0::n_spin])
    if n_spin > 1:
        tape.put(ham[1::2])
    tape.put([exc])


def recover(tape, n_spin, npk):
    """(ham of n_spin * npk, exc, mat_initialized) for a run with n_spin spins."""
    xc_length = n_spin * npk
    if tape.get(1)[0] == 0.0:
        return np.zeros(xc_length), None, True
    spin_stored, _ = (int(v) for v in tape.get(2))
    ham = np.zeros(xc_length)
    ham[0::n_spin] = tape.get(xc_length // n_spin)
    initialized = False
    if spin_stored > 1:
        if n_spin > 1:
            ham[1::2] = tape.get(xc_length // 2)
        else:
            ham = (ham + tape.get(xc_length)) * 0.5
    elif n_spin > 1:
        ham[1::2] = ham[0::2]
        initialized = True
    exc = tape.get(1)[0]
    return ham, exc, initialized


def put_state(tape, blob):
    """The library's incremental state after the XC records: [nbytes], then
    the bytes as doubles (zero-padded to whole doubles), as a Fortran caller
    hands them to readwriteblocked_write."""
    blob = bytes(blob)
    pad = (-len(blob)) % 8
    tape.put([float(len(blob))])
    tape.put(np.frombuffer(blob + b"\0" * pad, dtype="<f8"))


def get_state(tape):
    nbytes = int(tape.get(1)[0])
    words = tape.get((nbytes + 7) // 8)
    return words.astype("<f8").tobytes()[:nbytes]
