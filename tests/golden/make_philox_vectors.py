"""Writes tests/golden/philox_openrand.json: output words of the reference's
own Philox generator (third_party/openrand/philox.h, compiled where it lies
into oracle/_ref/libfj_ref.so by `make -C oracle ref`).

Each case is one generator openrand::Philox(seed, ctr, global_seed, ctr1):
its first 6 rand<uint32_t>() words and the FP32 bit patterns of the first 6
rand<float>() draws of a second generator built with the same arguments.
The cases cover the argument arrangement of the scanner-glare stream
(ctr = x, ctr1 = y, global_seed = medium) at small, large and negative
(two's-complement) coordinates, and extreme seeds.

Run in a container that holds the reference: python tests/golden/make_philox_vectors.py
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libfj_ref.so")
COUNT = 6


def cases():
    rng = np.random.default_rng(20111112)
    out = [(0, 0, 0, 0), (0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF),
           (0x0123456789ABCDEF, 17, 1, 4), (0xCBF29CE484222325, 3839, 0, 2159),
           (0xCBF29CE484222325, 0xFFFFFFFF, 1, 0xFFFFFFFE)]  # x = -1, y = -2
    for _ in range(59):
        seed = int(rng.integers(0, 2 ** 63)) * 2 + int(rng.integers(0, 2))
        out.append((seed, int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 2)), int(rng.integers(0, 2 ** 32))))
    return out


def main():
    lib = C.CDLL(REF_LIB)
    lib.fjr_philox_draws.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
    rows = []
    for seed, ctr, global_seed, ctr1 in cases():
        words = (C.c_uint32 * COUNT)()
        floats = (C.c_float * COUNT)()
        lib.fjr_philox_draws(seed, ctr, global_seed, ctr1, COUNT, words, floats)
        fbits = np.frombuffer(bytes(floats), dtype=np.uint32)
        rows.append({"seed": str(seed), "ctr": ctr, "global_seed": global_seed, "ctr1": ctr1,
                     "words": [int(w) for w in words], "float_bits": [int(b) for b in fbits]})
    path = os.path.join(HERE, "philox_openrand.json")
    with open(path, "w") as f:
        json.dump({"source": "third_party/openrand/philox.h (openrand::Philox), rand<uint32_t>() and rand<float>()",
                   "cases": rows}, f, indent=0)
    print(f"wrote {len(rows)} cases to {path}")


if __name__ == "__main__":
    main()
