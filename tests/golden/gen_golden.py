"""Write the regression fixtures of tests/golden/ from the C oracle.

The reference holds no golden vectors for this path (no .npz/.npy fixtures
for the event/sparse/jitconn operators exist in its tree), so these
fixtures freeze the oracle's outputs after it has been pinned by the
closed-form, period and cross-restatement checks in tests/test_oracle_*.py.

    python tests/golden/gen_golden.py
"""
import os
import sys
import zlib

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
    # a 200 x 100 event csrmv (ragged seeded CSR, heterogeneous f32 weights),
    # both orientations, from the serial oracle loops
    from tests.helpers import bernoulli, coba4000_inputs, random_csr
    indices, indptr = random_csr(200, 100, 0.1, seed=2024)
    data = np.random.default_rng(5).normal(0.5, 0.1, indices.size).astype(np.float32)
    ev_t, ev_g = bernoulli(200, 0.3, 6), bernoulli(100, 0.3, 7)
    np.savez_compressed(os.path.join(HERE, "event_csrmv_200x100.npz"), indices=indices, indptr=indptr, data=data,
                        events_transpose=ev_t, events_gather=ev_g,
                        y_transpose=orc.event_csrmv(data, indices, indptr, ev_t, (200, 100), True),
                        y_gather=orc.event_csrmv(data, indices, indptr, ev_g, (200, 100), False))
    # the 4000-neuron COBA network's first 100 steps (serial oracle run)
    _, idx, ptr, V0 = coba4000_inputs(0)
    spikes, raster = orc.coba_net_run(3200, 800, ptr, idx, 0.6, 6.7, 20.0, V0, 100,
                                      (-60.0, -50.0, -60.0, 20.0, 5.0, 1.0))
    np.savez_compressed(os.path.join(HERE, "coba4000_raster100.npz"), raster_bits=np.packbits(raster, axis=1),
                        spikes=np.int64(spikes), csr_crc=np.uint32(zlib.crc32(idx.tobytes() + ptr.tobytes())),
                        V0=V0)


if __name__ == "__main__":
    main()
