"""The restatement of the SCF-state records (oracle/scfstate.py) without a GPU:
its blocked tape is what Fortran unformatted sequential I/O writes and reads
(checked with flang both ways), the records come in xc_hamiltonian_store's
order (modules/xc_hamiltonian.f90:2478-2484: flag, n_spin and xc_length, the
spins, exc_int), and recover converts between spin counts as the reference
does (:2563-2626)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import scfstate

FORTRAN = """
program tape
  implicit none
  real(8) :: buf(128), tot
  integer :: r, i
  open(10, file='from_fortran.bin', form='unformatted', access='sequential', status='replace')
  do r = 1, 3
     buf = [(r * 1000.0d0 + i, i = 1, 128)]
     write(10) buf
  end do
  close(10)
  open(11, file='from_python.bin', form='unformatted', access='sequential', status='old')
  tot = 0
  do r = 1, 2
     read(11) buf
     tot = tot + sum(buf)
  end do
  close(11)
  write(*, '(es24.16)') tot
end program tape
"""


def test_blocked_tape_is_fortran_sequential_records(tmp_path):
    flang = shutil.which("flang")
    if flang is None:
        pytest.skip("flang not available")
    (tmp_path / "tape.f90").write_text(FORTRAN)
    t = scfstate.BlockedTape(str(tmp_path / "from_python.bin"), "w")
    items = np.arange(1, 201, dtype=np.float64)  # 200 items: one full block and a partial one
    t.put(items[:77])
    t.put(items[77:])
    t.close()
    subprocess.run([flang, "-o", str(tmp_path / "tape"), str(tmp_path / "tape.f90")], check=True,
                   cwd=tmp_path)
    out = subprocess.run([str(tmp_path / "tape")], capture_output=True, text=True, check=True,
                         cwd=tmp_path)
    # the second block holds items 129..200, and after them what the buffer
    # held before (items 73..128), as the reference's stopwrite writes it
    assert float(out.stdout) == items.sum() + items[72:128].sum()
    r = scfstate.BlockedTape(str(tmp_path / "from_fortran.bin"), "r")
    got = np.concatenate([r.get(5), r.get(300), r.get(79)])
    want = np.concatenate([rr * 1000.0 + np.arange(1, 129) for rr in (1, 2, 3)])
    assert np.array_equal(got, want)
    with pytest.raises(EOFError):
        r.get(1)


def test_record_order_and_spin_conversions(tmp_path):
    npk = 10
    rng = np.random.default_rng(1)
    ham2 = rng.normal(size=2 * npk)  # alpha / beta interleaved
    t = scfstate.BlockedTape(str(tmp_path / "t"), "w")
    scfstate.store(t, ham2, -7.25, 2)
    scfstate.put_state(t, b"\x01\x02\x03")
    t.close()
    r = scfstate.BlockedTape(str(tmp_path / "t"), "r")
    stream = r.get(1 + 2 + 2 * npk + 1 + 1 + 1)
    assert stream[0] == 1.0 and stream[1] == 2.0 and stream[2] == 2 * npk
    assert np.array_equal(stream[3:3 + npk], ham2[0::2])
    assert np.array_equal(stream[3 + npk:3 + 2 * npk], ham2[1::2])
    assert stream[3 + 2 * npk] == -7.25 and stream[4 + 2 * npk] == 3.0
    for n_spin, want, init in ((2, ham2, False), (1, (ham2[0::2] + ham2[1::2]) * 0.5, False)):
        r = scfstate.BlockedTape(str(tmp_path / "t"), "r")
        ham, exc, got_init = scfstate.recover(r, n_spin, npk)
        assert np.array_equal(ham, want) and exc == -7.25 and got_init == init
        assert scfstate.get_state(r) == b"\x01\x02\x03"
    ham1 = rng.normal(size=npk)
    t = scfstate.BlockedTape(str(tmp_path / "t1"), "w")
    scfstate.store(t, ham1, 1.5, 1)
    t.close()
    ham, exc, init = scfstate.recover(scfstate.BlockedTape(str(tmp_path / "t1"), "r"), 2, npk)
    assert np.array_equal(ham[0::2], ham1) and np.array_equal(ham[1::2], ham1) and init
    assert os.path.getsize(tmp_path / "t1") == 8 * 128 + 8
