"""Write golden vectors from the reference's own code compiled in place
(oracle/_ref/libmctref.so, built by `make -C oracle ref` from /root/reference).

Only data is written: seeded inputs and the reference's outputs.  Run from the
repo root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def kdtree_golden():
    """kdtree2 1-NN (src/kdtree2.f90) on random and lattice (tie-heavy) sites."""
    rng = np.random.default_rng(20261202)
    sets = []
    for t in range(12):
        S = int(rng.integers(2, 420))
        pts = rng.uniform(0, 20, (S, 3))
        q = rng.uniform(-1, 21, (256, 3))
        sets.append((pts, q))
    for t in range(12):
        g = rng.integers(0, 6, (int(rng.integers(14, 160)), 3)).astype(float) * 2.0
        g = np.unique(g, axis=0)
        rng.shuffle(g)
        q = rng.integers(0, 11, (256, 3)).astype(float)
        sets.append((g, q))
    arrays = {}
    for i, (pts, q) in enumerate(sets):
        arrays["pts_%d" % i] = pts
        arrays["q_%d" % i] = q
        arrays["idx_%d" % i] = po.ref_kdtree_nearest(pts, q)
    np.savez_compressed(os.path.join(OUT, "kdtree2_nearest.npz"), n=len(sets), **arrays)


if __name__ == "__main__":
    if po.ref() is None:
        sys.exit("oracle/_ref/libmctref.so is not built (make -C oracle ref)")
    kdtree_golden()
    print("wrote golden vectors to", OUT)
