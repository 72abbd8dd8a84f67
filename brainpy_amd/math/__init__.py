"""``brainpy.math`` operator namespaces of the event-driven synaptic path:
``event``, ``sparse`` and ``jitconn`` (mirrors of brainpy/math/event.py,
sparse.py and jitconn.py)."""
from . import event, jitconn, sparse

__all__ = ["event", "sparse", "jitconn"]
