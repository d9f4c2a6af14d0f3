"""Writes tests/golden/kernel_ref.npz: the reference's own compiled `kernel` (kernel.f90:7-25,
built unchanged by oracle/Makefile into oracle/_ref/libref.so) evaluated on a fixed set of
arguments, plus its `kerneltest` sum (kernel.f90:43-63).  Run where /root/reference exists:

    make -C oracle ref && python tests/golden/make_kernel_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle  # noqa: E402


def arguments():
    rng = np.random.default_rng(20261202)
    grid = np.arange(-2.0, 2.0 + 1e-12, 1.0 / 256.0)
    edges = np.array([0.5, 1.5, -0.5, -1.5, 0.0, 1.0, -1.0])
    near = np.concatenate([np.nextafter(edges, np.inf), np.nextafter(edges, -np.inf)])
    rand = rng.uniform(-1.75, 1.75, 4096)
    return np.concatenate([grid, edges, near, rand])


def main():
    r = _oracle.ref_lib()
    if r is None:
        raise SystemExit("oracle/_ref/libref.so not built (needs the reference tree)")
    xs = arguments()
    vals = np.array([r.ref_kernel(float(x)) for x in xs])
    np.savez(os.path.join(HERE, "kernel_ref.npz"), r=xs, kernel=vals,
             kerneltest=np.array([r.ref_kerneltest()]))
    print("wrote", len(xs), "values; kerneltest =", repr(r.ref_kerneltest()))


if __name__ == "__main__":
    main()
