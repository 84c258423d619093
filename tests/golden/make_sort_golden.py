"""Writes tests/golden/sort_golden.npz: the inputs of the reference's only
hot-path test (framework/tests/test_gpu.py:17-61) re-created with the Python
standard library, plus the expected Morton bucket of every cell.

Inputs: N=10,000 cells, grid (50, 50, 50), voxel size 2.0, random.seed(42),
three random.uniform(0, dim * 2.0 * 0.99) draws per cell in x, y, z order
(test_gpu.py:33-37). The expected bucket is computed here with a bit-by-bit
Morton interleave of int(x / 2.0) (test_gpu.py:40), independent of the
oracle's mask-and-shift closed form.
"""
import os
import random

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def morton_bits(x, y, z):
    k = 0
    for b in range(10):
        k |= ((x >> b) & 1) << (3 * b)
        k |= ((y >> b) & 1) << (3 * b + 1)
        k |= ((z >> b) & 1) << (3 * b + 2)
    return k


def main():
    n, gdim, vs = 10_000, (50, 50, 50), 2.0
    random.seed(42)
    pos = np.empty((n, 3), dtype=np.float64)
    keys = np.empty(n, dtype=np.uint32)
    for i in range(n):
        p = [random.uniform(0, gdim[j] * vs * 0.99) for j in range(3)]
        pos[i] = p
        keys[i] = morton_bits(*(int(x / vs) for x in p))
    np.savez_compressed(os.path.join(HERE, "sort_golden.npz"), pos=pos, keys=keys,
                        dims=np.array(gdim, dtype=np.int32), voxel=np.array([vs]))


if __name__ == "__main__":
    main()
