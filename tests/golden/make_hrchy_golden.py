"""Writes hrchy_ref.npz: the reference's own hierarchy tables as a fixture.

TEST INFRASTRUCTURE.  Runs the reference's src/hrchy_space.cc, compiled from
the reference sources by oracle/Makefile into oracle/_ref/libhrchy_ref.so
(alloc_hrchy_space, hrchy_space.cc:137-231, with and without an hrchy_filter,
hrchy_space.h:29-36), and stores the breadth-first index table n and the
ptr_p1 / ptr_m1 neighbour tables for every case below.  The CPU tests check
the oracle's enumeration against these tables wherever the reference is
absent (and the GPU tests check the library against the oracle).

Usage (in the container that holds the reference): python tests/golden/make_hrchy_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

# (n_dim K, max_depth L) of the unfiltered cases
CASES = [(1, 0), (1, 5), (2, 10), (3, 4), (4, 3), (5, 5), (8, 3), (21, 2), (16, 3), (7, 4)]
# filters (index ndarray, depth) -> bool, and the (K, L) each is applied at
FILTERS = {
    "max occupation 2": lambda n, d: n.max() <= 2,
    "weighted depth": lambda n, d: int((n * (np.arange(n.size) + 1)).sum()) <= 7,
    # rejects the intermediate index (1, 0, ...): its canonical descendants vanish too
    "prune through an intermediate": lambda n, d: not (n[0] == 1 and n[1:].sum() == 0),
}
FILTER_CASES = [(3, 5), (4, 4), (6, 3)]


def key(K, L, name=None):
    return "K%d_L%d" % (K, L) + ("" if name is None else "_" + name.replace(" ", "_"))


def main():
    from oracle import oracle as O
    out = {}
    for K, L in CASES:
        ref = O.ref_hrchy(K, L)
        if ref is None:
            raise SystemExit("oracle/_ref not built: run make -C oracle with the reference present")
        for tag, a in zip(("n", "p1", "m1"), ref):
            out[key(K, L) + "_" + tag] = a
    for name, fn in FILTERS.items():
        for K, L in FILTER_CASES:
            ref = O.ref_hrchy_filtered(K, L, fn)
            for tag, a in zip(("n", "p1", "m1"), ref):
                out[key(K, L, name) + "_" + tag] = a
    path = os.path.join(HERE, "hrchy_ref.npz")
    np.savez_compressed(path, **out)
    print("wrote %s (%d arrays, %d bytes)" % (path, len(out), os.path.getsize(path)))


if __name__ == "__main__":
    main()
