"""Write the regression fixtures of tests/golden/ from the C oracle.

The reference holds no golden vectors for this path (no .npz/.npy fixtures
for the event/sparse/jitconn operators exist in its tree), so these
fixtures freeze the oracle's outputs after it has been pinned by the
closed-form, period and cross-restatement checks in tests/test_oracle_*.py.

    python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import np_ref, orc  # noqa: E402


def main():
    np.savez_compressed(os.path.join(HERE, "lfsr88_streams.npz"),
                        **{f"seed{s}": orc.lfsr88_stream(s, 64) for s in (0, 1, 42, 1234)})
    # one-hot jitconn columns: M[:, c] for a few columns of small shapes
    out = {}
    for shape in ((100, 200), (2, 1000), (1000, 2)):
        for p in (0.01, 0.1, 0.5):
            for layout in (0, 1):
                for od in (0, 1):
                    r, c, w = orc.jitconn_enumerate(1, od, layout, shape[0], shape[1], 0.0, 1.0,
                                                    np_ref.conn_len(p), 1234)
                    key = f"{shape[0]}x{shape[1]}_p{p}_l{layout}_o{od}"
                    out[key + "_rows"] = r
                    out[key + "_cols"] = c
                    out[key + "_w"] = w
    np.savez_compressed(os.path.join(HERE, "jitconn_uniform_conn.npz"), **out)


if __name__ == "__main__":
    main()
