"""Writes tests/golden/xxh3_vectors.json: XXH3_64bits_withSeed over 8-byte little-endian inputs,
computed with the `xxhash` library (libxxhash), independent of this repository's hash code.

The reference hashes a 64-bit edge index with XXH3_64bits_withSeed (include/types.h:8-9,
include/bucket.h:54-63): seeds are the sketch seed (checksum) and seed + 5*c (column c).
Run: python tests/golden/make_golden.py
"""
import json
import os
import random
import struct

import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = random.Random(20240607)
    seeds = [0, 1, 42, 47, 52, 42 + 5 * 24, 0xFFFFFFFF, 0x100000000, 0xFFFFFFFFFFFFFFFF, 0x8000000000000000,
             0x0123456789ABCDEF]
    seeds += [rng.getrandbits(64) for _ in range(8)]
    inputs = [0, 1, 2, 0xFFFFFFFF, 0x100000000, 0xFFFFFFFFFFFFFFFF, (3 << 32) | 1000, (1 << 32) | 2]
    inputs += [(rng.randrange(1 << 18) << 32) | rng.randrange(1 << 18) for _ in range(12)]
    inputs += [rng.getrandbits(64) for _ in range(6)]
    vecs = []
    for s in seeds:
        for x in inputs:
            vecs.append([str(x), str(s), str(xxhash.xxh3_64_intdigest(struct.pack("<Q", x), seed=s))])
    with open(os.path.join(HERE, "xxh3_vectors.json"), "w") as f:
        json.dump({"library": "xxhash", "library_version": xxhash.VERSION, "xxhash_version": xxhash.XXHASH_VERSION,
                   "function": "XXH3_64bits_withSeed(8-byte LE input, seed)", "vectors": vecs}, f, indent=0)
    print(len(vecs), "vectors")


if __name__ == "__main__":
    main()
