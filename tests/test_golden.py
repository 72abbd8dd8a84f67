"""The committed fixtures of tests/golden/ against the oracle -- CPU only.

The reference holds no golden vectors for this path; these fixtures freeze
the pinned oracle (gen_golden.py) so that a later change of the oracle, of
the seeded input generators, or of numpy's streams shows up here.  The
event csrmv fixture is also checked against its FP64 closed form
(events @ dense(A), as event/tests/test_event_csrmv.py:66-68 does)."""
import os
import zlib

import numpy as np

from oracle import orc
from tests.helpers import coba4000_inputs, dense_of

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_event_csrmv_fixture():
    d = np.load(os.path.join(HERE, "event_csrmv_200x100.npz"))
    idx, ptr, w = d["indices"], d["indptr"], d["data"]
    y_t = orc.event_csrmv(w, idx, ptr, d["events_transpose"], (200, 100), True)
    y_g = orc.event_csrmv(w, idx, ptr, d["events_gather"], (200, 100), False)
    np.testing.assert_array_equal(y_t, d["y_transpose"])
    np.testing.assert_array_equal(y_g, d["y_gather"])
    D = dense_of(w, idx, ptr, (200, 100))
    np.testing.assert_allclose(d["y_transpose"], d["events_transpose"].astype(np.float64) @ D, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(d["y_gather"], D @ d["events_gather"].astype(np.float64), rtol=1e-5, atol=1e-5)


def test_coba4000_raster_fixture():
    d = np.load(os.path.join(HERE, "coba4000_raster100.npz"))
    _, idx, ptr, V0 = coba4000_inputs(0)
    assert zlib.crc32(idx.tobytes() + ptr.tobytes()) == int(d["csr_crc"])
    np.testing.assert_array_equal(V0, d["V0"])
    spikes, raster = orc.coba_net_run(3200, 800, ptr, idx, 0.6, 6.7, 20.0, V0, 100,
                                      (-60.0, -50.0, -60.0, 20.0, 5.0, 1.0))
    assert spikes == int(d["spikes"]) > 0
    np.testing.assert_array_equal(raster, np.unpackbits(d["raster_bits"], axis=1)[:, :4000])
