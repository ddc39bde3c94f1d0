"""Writes tests/golden/threefry_vectors.json: inputs and outputs of the reference's own
Threefry-2x32-20 (the vendored Random123 header under the reference tree, built by
oracle/Makefile into oracle/_ref/libref_threefry.so).  Run in the build container:
    make -C oracle ref && python tests/golden/make_threefry_vectors.py
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle  # noqa: E402


def main():
    assert oracle.ref_lib() is not None, "build oracle/_ref first (make -C oracle ref)"
    rng = random.Random(20261203)
    inputs = [(0, 0, 0, 0), (0xFFFFFFFF,) * 4, (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
              (1, 0xF00DCAFE, 0, 0xDECAFBAD)]
    # reference-shaped draws: counter {n, 0xf00dcafe}, key {cell, 0xdecafbad + seed}
    for _ in range(64):
        inputs.append((rng.randrange(1, 200), 0xF00DCAFE, rng.randrange(0, 1 << 20),
                       (0xDECAFBAD + rng.randrange(-2**31, 2**31)) & 0xFFFFFFFF))
    for _ in range(64):
        inputs.append(tuple(rng.getrandbits(32) for _ in range(4)))
    out = [list(oracle.threefry(*v, which="ref")) for v in inputs]
    with open(os.path.join(HERE, "threefry_vectors.json"), "w") as f:
        json.dump({"source": "exec/src/gpgmp/src/cl_include/Random123/threefry.h threefry2x32 (20 rounds)",
                   "fields": "ctr0 ctr1 key0 key1 -> out0 out1",
                   "inputs": [list(v) for v in inputs], "outputs": out}, f)
    print("wrote", len(inputs), "vectors")


if __name__ == "__main__":
    main()
