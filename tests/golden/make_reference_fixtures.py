"""Convert the two matrix data files the reference ships into small numpy fixtures.

Run in the build container (the reference tree is not present on GPU machines):
    python tests/golden/make_reference_fixtures.py /root/reference
  mob.csv        -- PETSc ASCII dense dump of a 121 x 121 grand mobility (Nm=33, Nc=2, beta=0.1,
                    the test_sim.py system with mm_HI=False)
  src/saddle.bin -- PETSc binary Mat (classid 1211216) of a 214 x 214 saddle matrix (Nm=30, Nc=2)
Both are parsed as numbers only.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_petsc_ascii_dense(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            try:
                rows.append([float(v) for v in parts])
            except ValueError:
                continue   # header lines ("Mat Object", "type: seqdense")
    return np.array(rows)


def read_petsc_binary_mat(path):
    raw = open(path, "rb").read()
    head = np.frombuffer(raw[:16], ">i4")
    assert head[0] == 1211216, "not a PETSc Mat"
    m, n, nz = int(head[1]), int(head[2]), int(head[3])
    off = 16
    rowlen = np.frombuffer(raw[off:off + 4 * m], ">i4"); off += 4 * m
    cols = np.frombuffer(raw[off:off + 4 * nz], ">i4"); off += 4 * nz
    vals = np.frombuffer(raw[off:off + 8 * nz], ">f8")
    A = np.zeros((m, n))
    k = 0
    for i in range(m):
        for _ in range(rowlen[i]):
            A[i, cols[k]] = vals[k]
            k += 1
    return A


def main(ref):
    M = read_petsc_ascii_dense(os.path.join(ref, "mob.csv"))
    A = read_petsc_binary_mat(os.path.join(ref, "src", "saddle.bin"))
    np.savez_compressed(os.path.join(HERE, "ref_mob_csv.npz"), M=M)
    np.savez_compressed(os.path.join(HERE, "ref_saddle_bin.npz"), A=A)
    print(M.shape, A.shape)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
