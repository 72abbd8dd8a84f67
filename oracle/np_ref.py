"""Independent numpy restatement of LFSR88 and of the just-in-time
connectivity generators (test infrastructure only).

Written separately from ``bpx_oracle.c`` (vectorised over streams instead of
looping per stream) so the two restatements cross-check each other.

* LFSR88: ``brainpy/_src/math/tifunc.py:47-180``.
* row layout (Taichi CPU kernels): ``jitconn/_event_matvec_taichi.py:38-85``.
* lane32 layout (Taichi GPU kernels): ``jitconn/_event_matvec_taichi.py:92-154``.
"""
import numpy as np

U32 = np.uint32
_M0, _M1, _M2 = U32(0xFFFFFFFE), U32(0xFFFFFFF8), U32(0xFFFFFFF0)


def key(seed):
    """lfsr88_key: (seed+1, seed+7, seed+15, 0) in uint32 (tifunc.py:67)."""
    s = np.asarray(seed, dtype=np.uint64).astype(U32)
    with np.errstate(over="ignore"):
        return [s + U32(1), s + U32(7), s + U32(15), np.zeros_like(s)]


def next_key(k):
    """lfsr88_next_key (tifunc.py:80-86) on arrays of states."""
    k0, k1, k2 = k[0], k[1], k[2]
    b = ((k0 << U32(13)) ^ k0) >> U32(19)
    s1 = ((k0 & _M0) << U32(12)) ^ b
    b = ((k1 << U32(2)) ^ k1) >> U32(25)
    s2 = ((k1 & _M1) << U32(4)) ^ b
    b = ((k2 << U32(3)) ^ k2) >> U32(11)
    s3 = ((k2 & _M2) << U32(17)) ^ b
    return [s1, s2, s3, b]


def bits(k):
    return k[0] ^ k[1] ^ k[2]


def stream(seed, count):
    k = key(np.asarray([seed]))
    out = np.empty(count, U32)
    for i in range(count):
        k = next_key(k)
        out[i] = bits(k)[0]
    return out


def randint(k, low, high):
    """lfsr88_random_integers (tifunc.py:136-147) -> (new key, int64 values)."""
    k = next_key(k)
    d = U32(high + 1 - low)
    return k, (bits(k) % d).astype(np.int64) + low


def uniform(k, low, high):
    """lfsr88_uniform (tifunc.py:156-168) in float32, mul then add."""
    k = next_key(k)
    r = bits(k).astype(np.float32) * np.float32(2.3283064365386963e-10)
    return k, (r * np.float32(np.float32(high) - np.float32(low))) + np.float32(low)


def normal(k, mu, sigma):
    """lfsr88_normal (tifunc.py:89-133), float32 Box-Muller sine branch."""
    k = next_key(k)
    u1 = bits(k).astype(np.float32) * np.float32(2.3283064365386963e-10)
    k = next_key(k)
    u2 = bits(k).astype(np.float32) * np.float32(2.3283064365386963e-10)
    u1 = np.maximum(u1, np.float32(1e-10))
    mag = np.sqrt(np.float32(-2.0) * np.log(u1))
    z = mag * np.sin(np.float32(6.283185307179586) * u2)
    return k, np.float32(mu) + np.float32(sigma) * z


def _draw(k, dist, w0, w1):
    if dist == "uniform":
        return uniform(k, w0, w1)
    if dist == "normal":
        return normal(k, w0, w1)
    return k, None


def enumerate_row(dist, outdim_parallel, num_row, num_col, w0, w1, clen, seed):
    """All (row, col, w) of the row layout, as sorted-by-(row,col) arrays."""
    outer, inner = (num_row, num_col) if outdim_parallel else (num_col, num_row)
    a = np.arange(outer, dtype=np.int64)
    k = key((np.uint64(seed) + a.astype(np.uint64)) & np.uint64(0xFFFFFFFF))
    k, b = randint(k, 0, clen - 1)
    R, C, W = [], [], []
    alive = b < inner
    while alive.any():
        k, w = _draw(k, dist, w0, w1)
        idx = np.nonzero(alive)[0]
        R.append(a[idx])
        C.append(b[idx])
        W.append(np.full(idx.size, np.float32(w0)) if w is None else w[idx])
        k, inc = randint(k, 1, clen)
        b = np.where(alive, b + inc, b)
        alive = alive & (b < inner)
    return _pack(R, C, W, outdim_parallel)


def enumerate_lane32(dist, outdim_parallel, num_row, num_col, w0, w1, clen, seed):
    outer, inner = (num_row, num_col) if outdim_parallel else (num_col, num_row)
    step = np.int64(max((num_row + 1) >> 5, 1))
    a = np.repeat(np.arange(outer, dtype=np.int64), 32)
    lane = np.tile(np.arange(32, dtype=np.int64), outer)
    k = key((np.uint64(seed) + (a * 32 + lane).astype(np.uint64)) & np.uint64(0xFFFFFFFF))
    # u32 arithmetic of the GPU kernel, carried in int64 modulo 2^32
    start = (step * lane - 1) & 0xFFFFFFFF
    end = np.minimum((start + step) & 0xFFFFFFFF, inner)
    k, inc = randint(k, 1, clen)
    b = (start + inc) & 0xFFFFFFFF
    R, C, W = [], [], []
    alive = b < end
    while alive.any():
        k, w = _draw(k, dist, w0, w1)
        idx = np.nonzero(alive)[0]
        R.append(a[idx])
        C.append(b[idx])
        W.append(np.full(idx.size, np.float32(w0)) if w is None else w[idx])
        k, inc = randint(k, 1, clen)
        b = np.where(alive, b + inc, b)
        alive = alive & (b < end)
    return _pack(R, C, W, outdim_parallel)


def _pack(R, C, W, outdim_parallel):
    if not R:
        z = np.zeros(0, np.int64)
        return z, z, np.zeros(0, np.float32)
    a = np.concatenate(R)
    b = np.concatenate(C)
    w = np.concatenate(W).astype(np.float32)
    rows, cols = (a, b) if outdim_parallel else (b, a)
    order = np.lexsort((cols, rows))
    return rows[order], cols[order], w[order]


def dense(dist, outdim_parallel, layout, num_row, num_col, w0, w1, clen, seed):
    """Dense FP64 matrix M (num_row, num_col) of the generated connectivity."""
    f = enumerate_lane32 if layout == "lane32" else enumerate_row
    r, c, w = f(dist, outdim_parallel, num_row, num_col, w0, w1, clen, seed)
    M = np.zeros((num_row, num_col), np.float64)
    np.add.at(M, (r, c), w.astype(np.float64))
    return M


def conn_len(conn_prob):
    """clen = ceil(1/p)*2-1 as the reference computes it: 1/p in Python
    float, rounded to float32 by jnp.ceil, then int32
    (jitconn/_event_matvec_taichi.py:401-402)."""
    return int(np.ceil(np.float32(1.0 / conn_prob)) * 2 - 1)
