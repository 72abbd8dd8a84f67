"""Host-side logic of the batched network path (no GPU)."""
import numpy as np

from brainpy_amd.dyn.batched import max_in_degree


def test_max_in_degree_counts_duplicates_per_population():
    # 3 E neurons, 2 I neurons; rows list local post ids
    rows = [[1, 1, 4], [1, 3], [0],          # E rows: neuron 1 gets 3 E synapses
            [1, 1, 1, 1], [7, -1, 2]]        # I rows: neuron 1 gets 4 I synapses; 7 and -1 are skipped
    indptr = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    indices = np.concatenate([np.asarray(r, np.int32) for r in rows])
    assert max_in_degree(indices, indptr, 3, 2) == 4
    # E population only
    assert max_in_degree(indices[:6], indptr[:4], 3, 0) == 3
    # no synapses
    assert max_in_degree(np.zeros(0, np.int32), np.zeros(6, np.int32), 3, 2) == 0
