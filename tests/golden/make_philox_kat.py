"""Writes tests/golden/philox_kat.json: Philox4x32-10 vectors from two
independent sources, used to pin the oracle's Philox.

1. Random123's published known-answer vectors (kat_vectors, Salmon et al.).
2. rocRAND's philox4x32_10_engine::ten_rounds, compiled here for the host
   from /opt/rocm/include/rocrand/rocrand_philox4x32_10.h with hipcc, on
   seeded random counters and keys.
"""
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RANDOM123 = [
    # ctr[4], key[2], out[4]
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]

SRC = r'''
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 run(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};
int main() {
    Probe p;
    unsigned c0, c1, c2, c3, k0, k1;
    while (scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 o = p.run(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
    }
    return 0;
}
'''


def rocrand_vectors(n=64, seed=7):
    rng = np.random.default_rng(seed)
    ins = rng.integers(0, 2**32, size=(n, 6), dtype=np.uint64)
    ins[0] = 0
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.cpp")
        exe = os.path.join(d, "probe")
        open(src, "w").write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", src, "-o", exe],
                       check=True)
        txt = "\n".join(" ".join(str(int(v)) for v in row) for row in ins) + "\n"
        out = subprocess.run([exe], input=txt, capture_output=True, text=True, check=True).stdout.split("\n")
    vecs = []
    for row, line in zip(ins, out):
        o = [int(v) for v in line.split()]
        vecs.append(([int(v) for v in row[:4]], [int(v) for v in row[4:]], o))
    return vecs


def main():
    data = {"random123": RANDOM123, "rocrand": rocrand_vectors()}
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(data, f, indent=0)


if __name__ == "__main__":
    main()
