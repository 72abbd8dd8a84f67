"""CPU oracle for the event-driven synaptic hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the reference the GPU results are
compared with (or, in the CPU baseline, timed as the CPU path).  The product
path (``brainpy_amd``) never imports it and has no CPU fallback.

* ``oracle.orc``   -- ctypes binding of ``liborc.so`` (the C restatement in
  ``bpx_oracle.c``, built by ``oracle/Makefile``).
* ``oracle.np_ref`` -- an independent, vectorised numpy restatement of LFSR88
  and of both just-in-time connectivity layouts, used to cross-check the C
  restatement on small shapes.
* ``oracle.coo_ref`` -- numpy restatement of the COO products (coomv,
  seg_matmul).
* ``oracle.bsr_ref`` -- numpy restatement of the block-sparse products
  (cusparse_bcsr_matvec, bcsrmm).

Parity unpinned: the normal-distribution weights of jitconn (log/sqrt/sin
are device/libm math, pinned to a tolerance only) and the brainpylib
``gpu_jit_*`` connectivity (its RNG is not in the reference tree; the
Taichi twins define the connectivity restated here).
"""
