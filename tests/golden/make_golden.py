"""Writes the golden fixtures under tests/golden/ from data files the reference's
own tests hold (run where the reference's files are; the GPU tests read only the fixtures).

testMatrix_upper.npz: src/maspack/solvers/testMatrix.mat (PardisoSolverTest.java:
608-688), a 2529x2529 symmetric matrix given as "[ (i j v) ... ]" followed by a
vector "[ ... ]"; stored as the 0-based upper CSR plus that vector and the SHA-256
of the source text.
"""
import hashlib
import os
import re

import numpy as np
import scipy.sparse as sp

REF = "/root/reference/src/maspack/solvers/testMatrix.mat"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    raw = open(REF, "rb").read()
    txt = raw.decode()
    a, rest = txt.split("]", 1)
    trip = re.findall(r"\(\s*(\d+)\s+(\d+)\s+([-+0-9.eE]+)\s*\)", a)
    I = np.array([int(t[0]) for t in trip])
    J = np.array([int(t[1]) for t in trip])
    V = np.array([float(t[2]) for t in trip])
    n = max(I.max(), J.max()) + 1
    vec = np.array([float(x) for x in rest.replace("[", " ").replace("]", " ").split()])
    A = sp.csr_matrix((V, (I, J)), shape=(n, n))
    U = sp.csr_matrix(sp.triu(A))
    U.sort_indices()
    np.savez_compressed(os.path.join(HERE, "testMatrix_upper.npz"),
                        rowOffs=U.indptr.astype(np.int32), colIdxs=U.indices.astype(np.int32),
                        vals=U.data, trailing_vector=vec, n_entries_full=np.int64(len(V)),
                        sha256=np.frombuffer(hashlib.sha256(raw).digest(), np.uint8))
    print("n=%d entries=%d upper=%d vector=%d" % (n, len(V), U.nnz, len(vec)))


if __name__ == "__main__":
    main()
