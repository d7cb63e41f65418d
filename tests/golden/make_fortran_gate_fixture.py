"""Writes tests/golden/fortran_visco_test.npz from the reference's fixture
tests/data/fortran_visco_test.txt (visccoefs.out of the original Fortran
sl3d visco_test at lmax 32 on prem.simp.vm1: rows `output l m Re(U) Im(U)
Re(phi) Im(phi)`, SI, output k at t = k kyr): the rows of outputs 1..5 as
arrays (output, l, m as int16; U and phi as complex128). Data only; run in
this container, where the reference is present:

    python3 tests/golden/make_fortran_gate_fixture.py
"""

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCE = "/root/reference/tests/data/fortran_visco_test.txt"


def main():
    rows = np.loadtxt(SOURCE)
    rows = rows[rows[:, 0] >= 1]
    np.savez_compressed(
        os.path.join(HERE, "fortran_visco_test.npz"),
        output=rows[:, 0].astype(np.int16), l=rows[:, 1].astype(np.int16),
        m=rows[:, 2].astype(np.int16), u=rows[:, 3] + 1j * rows[:, 4],
        phi=rows[:, 5] + 1j * rows[:, 6])


if __name__ == "__main__":
    main()
