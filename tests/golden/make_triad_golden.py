"""Writes tests/golden/triads_32.npz from the reference's triad-list files
(3e-9/full_yNN.dat, the lists nonlinRead reads, start.f90:2983-3069) for
files 1, 6 and 9 of its 32x32x153 case.

Every reference list is a subset of the threshold-0 list that
rnl_multiblock_nofft_amd.triads.full_lists(32, 32) generates, so a file is
stored as one bit per generated row (kept or not) plus the file's row count
per type; the lists are rebuilt from the generator.  Row order within a list
is the generator's (it only affects the rounding of the sums).

Run in the container that has /root/reference:
    python tests/golden/make_triad_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from rnl_multiblock_nofft_amd import triads  # noqa: E402

FILES = (1, 6, 9)


def main():
    full = triads.full_lists(32, 32)
    index = {t: {tuple(r): i for i, r in enumerate(full[t].tolist())} for t in range(1, 10)}
    out = {}
    for q in FILES:
        nx, nz, j, lists = triads.read_list("/root/reference/3e-9/full_y%02d.dat" % q)
        assert (nx, nz, j) == (15, 15, q)
        for t in range(1, 10):
            m = np.zeros(len(full[t]), bool)
            for r in lists[t].tolist():
                m[index[t][tuple(r)]] = True
            out["f%d_t%d" % (q, t)] = np.packbits(m)
            out["f%d_n%d" % (q, t)] = np.array(len(lists[t]))
    np.savez_compressed(os.path.join(HERE, "triads_32.npz"), **out)


def load():
    """{file index: {type: int8 (len, 4)}} from the stored bit masks."""
    full = triads.full_lists(32, 32)
    g = np.load(os.path.join(HERE, "triads_32.npz"))
    res = {}
    for q in FILES:
        lists = {}
        for t in range(1, 10):
            m = np.unpackbits(g["f%d_t%d" % (q, t)])[: len(full[t])].astype(bool)
            lists[t] = full[t][m]
            assert len(lists[t]) == int(g["f%d_n%d" % (q, t)])
        res[q] = lists
    return res


if __name__ == "__main__":
    main()
