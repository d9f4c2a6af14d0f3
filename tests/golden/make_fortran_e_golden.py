"""Writes tests/golden/fortran_e_format.npz: values and their Fortran '(3E15.7)' and
'(7E16.8)' editing as this image's Fortran compiler (flang) writes them -- the formats of the
reference's Capsule and center_velocity files (main.f90:463,467).  Fixture generator for
tests/test_formats.py; run it from the repository root:

    python tests/golden/make_fortran_e_golden.py

Negative zero is left out: the reference's compiler (Intel Fortran, default nominus0) writes it
without the sign, the Fortran 2003 default used by flang keeps it.
"""
import os
import subprocess
import tempfile

import numpy as np

FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
SRC = """program efmt
  implicit none
  integer(8) :: n
  real(8), allocatable :: v(:)
  open(10, file='v.bin', access='stream', form='unformatted', status='old')
  read(10) n
  allocate(v(n))
  read(10) v
  close(10)
  open(11, file='e157.txt', status='replace')
  write(11, '(3E15.7)') v
  close(11)
  open(12, file='e168.txt', status='replace')
  write(12, '(7E16.8)') v
  close(12)
end program efmt
"""


def values():
    rng = np.random.default_rng(2024)
    mag = 10.0 ** rng.uniform(-120, 120, 600)
    v = mag * rng.choice([-1.0, 1.0], 600)
    edge = [0.0, 1.0, -1.0, 0.5, 9.99999995, 9.99999994, 9.999999949, 99999999.5, 1e-99,
            9.9999999e-100, 1e100, -1e100, 1e99, 1234568.5, 1234567.5, 6249.999999999999,
            0.0016, 1e-4, 2.5, 1.0 / 3.0, 2.0 / 3.0, 5e-324, 1.7976931348623157e308]
    # three columns per line for '(3E15.7)' and seven for '(7E16.8)': 21 per block of lines
    v = np.concatenate([np.array(edge), v, rng.uniform(-10, 10, 200)])
    pad = (-v.size) % 21
    return np.concatenate([v, rng.uniform(0, 1, pad)])


def main():
    v = values()
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "efmt.f90"), "w") as f:
            f.write(SRC)
        with open(os.path.join(d, "v.bin"), "wb") as f:
            f.write(np.int64(v.size).tobytes() + v.astype("<f8").tobytes())
        subprocess.run([FLANG, "-O0", "efmt.f90", "-o", "efmt"], cwd=d, check=True)
        subprocess.run(["./efmt"], cwd=d, check=True)
        e157 = open(os.path.join(d, "e157.txt")).read()
        e168 = open(os.path.join(d, "e168.txt")).read()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran_e_format.npz")
    np.savez_compressed(out, values=v, e157=np.array(e157), e168=np.array(e168))
    print("wrote", out, v.size, "values")


if __name__ == "__main__":
    main()
