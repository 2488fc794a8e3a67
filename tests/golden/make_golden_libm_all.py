"""Writes golden_libm_all.npz: per-chunk digests of the correctly rounded sinf, cosf and tanf over ALL 2^32 float bit
patterns, computed with the reference's own correctly rounded functions (cuda/common/libm_cr.cuh, compiled in place into
oracle/_ref/libref_libm.so by `make -C oracle ref`).

Chunk c holds the 2^24 patterns c * 2^24 ... (c + 1) * 2^24 - 1; its digest is the sum (mod 2^64) over the chunk of
mix64(x << 32 | f(x)), mix64 being SplitMix64's finaliser. The GPU test computes the same sums on the device
(tests/devlibm/devlibm.hip) and compares all 768; a differing chunk is then compared pattern by pattern.

    python tests/golden/make_golden_libm_all.py      (about a minute per function on 8 cores)
"""
import multiprocessing as mp
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CHUNK_BITS = 24
FUNCS = ("sin", "cos", "tan")


def mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def chunk_digest(x_bits, out_bits):
    v = mix64((np.asarray(x_bits, np.uint64) << np.uint64(32)) | np.asarray(out_bits, np.uint64))
    return np.uint64(int(v.sum(dtype=np.uint64)))  # numpy's uint64 sum wraps mod 2^64


def _work(args):
    fn, c = args
    import oracle as orc
    x = np.arange(c << CHUNK_BITS, (c + 1) << CHUNK_BITS, dtype=np.uint64).astype(np.uint32)
    return fn, c, chunk_digest(x, orc.RefLibm().unary(fn, x))


def main():
    out = {name: np.zeros(1 << (32 - CHUNK_BITS), np.uint64) for name in FUNCS}
    jobs = [(fn, c) for fn in range(3) for c in range(1 << (32 - CHUNK_BITS))]
    with mp.Pool(min(8, os.cpu_count() or 1)) as pool:
        for fn, c, d in pool.imap_unordered(_work, jobs, chunksize=4):
            out[FUNCS[fn]][c] = d
    path = os.path.join(HERE, "golden_libm_all.npz")
    np.savez_compressed(path, **{k + "_chunks": v for k, v in out.items()})
    print("wrote", path)


if __name__ == "__main__":
    main()
