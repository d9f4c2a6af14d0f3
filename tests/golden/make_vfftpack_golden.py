"""Writes tests/golden/vfftpack_ref.npz: the reference's own vendored VFFTPACK (libs/fft/fft.f,
built unchanged by oracle/Makefile into oracle/_ref/libref.so) applied forward (vrfftf) and
backward (vrfftb) to seeded random sequences of the lengths the flow solver uses, one sequence
per call as solver1d makes them (solver.f90:61-66,139-145).  Run where /root/reference exists:

    make -C oracle ref && python tests/golden/make_vfftpack_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle  # noqa: E402

LENGTHS = (8, 12, 16, 17, 32, 64, 128)


def main():
    if _oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libref.so not built (needs the reference tree)")
    rng = np.random.default_rng(20261203)
    out = {}
    for n in LENGTHS:
        x = rng.standard_normal((3, n))
        out[f"x{n}"] = x
        out[f"f{n}"] = np.stack([_oracle.ref_vrfft(r) for r in x])
        out[f"b{n}"] = np.stack([_oracle.ref_vrfft(r, inverse=True) for r in x])
    np.savez(os.path.join(HERE, "vfftpack_ref.npz"), lengths=np.array(LENGTHS), **out)
    print("wrote", ", ".join(str(n) for n in LENGTHS))


if __name__ == "__main__":
    main()
