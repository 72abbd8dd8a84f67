"""Seeded synthetic inputs shared by the tests and the benchmark."""
import numpy as np

from brainpy_amd.connect import fixed_prob_csr  # noqa: F401


def random_csr(m, n, density, seed, dtype=np.int32):
    """Ragged CSR with Bernoulli(density) entries (variable row lengths,
    empty rows included)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mask = rng.random((m, n)) < density
    rows, cols = np.nonzero(mask)
    indptr = np.zeros(m + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=m), out=indptr[1:])
    return cols.astype(dtype), indptr.astype(dtype)


def dense_of(data, indices, indptr, shape):
    m, n = shape
    D = np.zeros((m, n), np.float64)
    rows = np.repeat(np.arange(m), np.diff(indptr.astype(np.int64)))
    w = np.broadcast_to(np.asarray(data, np.float64), indices.shape) if np.size(data) == 1 else np.asarray(data, np.float64)
    np.add.at(D, (rows, indices.astype(np.int64)), w)
    return D


def bernoulli(n, rate, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.random(n) < rate


def f32_to_bf16_bits(x):
    """Round-to-nearest-even float32 -> bfloat16 bit patterns (uint16)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r


def bf16_bits_to_f32(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def coba4000_inputs(seed=0):
    """The 4000-neuron E/I COBA network of EICOBANet(seed=seed) on the host:
    the four FixedProb p=0.02 projections in EICOBANet's order (E2I, I2I,
    I2E, E2E; seed * 7919 + k + 1), their combined per-neuron CSR, and the
    initial potentials U[-60, -50] from PCG64(seed)."""
    from brainpy_amd.dyn.batched import combined_csr
    size = {"E": 3200, "I": 800}
    csrs = {}
    for k, (name, pre, post) in enumerate((("E2I", "E", "I"), ("I2I", "I", "I"), ("I2E", "I", "E"),
                                           ("E2E", "E", "E"))):
        csrs[name] = fixed_prob_csr(size[pre], size[post], 0.02, seed * 7919 + k + 1)
    V0 = np.random.Generator(np.random.PCG64(seed)).uniform(-60.0, -50.0, 4000).astype(np.float32)
    idx, ptr = combined_csr(csrs, 3200, 800)
    return csrs, idx, ptr, V0
