"""Writes schnakenberg256_it6000.npz: the state of the 256x256 Schnakenberg run (seed0 5,
initial Poisson(200)/(75) counts of configs.schnakenberg_state) after 6000 host-loop
iterations on the oracle -- the Turing regime, where hundreds of SSA events per cell and
iteration occur in the spots.  Parity tests resume the loop from it on the GPU and on the
oracle (both skip the first 6000 iterations of the schedule and the seeds).

    python tests/golden/make_turing_snapshot.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

ITERATIONS = 6000
SEED0 = 5


def main():
    from inchman_amd import configs
    from oracle import oracle

    m = configs.schnakenberg_model(256)
    st = configs.schnakenberg_state(256)
    out, t = oracle.run(m, st, ITERATIONS, SEED0)
    np.savez_compressed(os.path.join(HERE, "schnakenberg256_it6000.npz"), state=out.astype(np.int32),
                        iterations=np.int64(ITERATIONS), seed0=np.int64(SEED0), sim_time=np.float32(t))
    print("sim time %.4f, U max %d, V max %d" % (t, out[0].max(), out[1].max()))


if __name__ == "__main__":
    main()
