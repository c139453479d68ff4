"""Writes tests/golden/fixture_spai.npz: the oracle's static SPAI M (MGS, the
reference's operation order) and |I_k| per column for the reference's two
matrix fixtures (tests/golden/matrices/*.mtx.gz, copied unchanged from the
reference tree's matrices/), plus the oracle's BiCGSTAB iteration spread with
the reference driver's solve settings (b = 1, x0 = 1, absolute TOL = 1e-7,
MAX_ITER = 10000; src/spai_multi_dcu.cpp:1920-1934).  The oracle is pinned by
tests/test_oracle.py; these vectors spare the GPU tests its 2 x ~150 s run on
circuit_2.

Run: python tests/golden/make_fixture_spai.py   (after __graft_entry__.build())
"""
import gzip
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import dcu_amd as D  # noqa: E402
from oracle import binding as O  # noqa: E402


def read_fixture(name):
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, name)
        with gzip.open(os.path.join(HERE, "matrices", name + ".gz"), "rb") as f, open(path, "wb") as g:
            shutil.copyfileobj(f, g)
        return D.Problem.read_mtx(path).system(0)


def main():
    out = {}
    for name in ("fill_3elt.mtx", "circuit_2.mtx"):
        key = name.split(".")[0]
        A = read_fixture(name)
        m = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
        n1, n2, t, e = O.column_sets(A.n, A.colptr, A.rowidx)
        out[f"{key}_m"] = m
        out[f"{key}_n1"] = n1
        if np.all(np.isfinite(m)):
            Ac = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
            Mc = O.csc2csr(A.n, A.colptr, A.rowidx, m)[:3]
            b = np.ones(A.n)
            spread = O.iteration_spread(Ac, Mc, b, 1e-7, 10000, threads=(1, 2, 4, 8))
            # the reference driver starts from x0 = 1
            its = []
            before = O.num_threads()
            for th in (1, 2, 4, 8):
                O.set_threads(th)
                its.append(O.bicgstab(Ac, Mc, b, np.ones(A.n), 1e-7, 10000)[1])
            O.set_threads(before)
            out[f"{key}_iters_x0zero"] = np.array(spread)
            out[f"{key}_iters_x0ones"] = np.array(its)
        print(key, A.n, A.nnz, "M finite", np.all(np.isfinite(m)), flush=True)
    np.savez_compressed(os.path.join(HERE, "fixture_spai.npz"), **out)


if __name__ == "__main__":
    main()
