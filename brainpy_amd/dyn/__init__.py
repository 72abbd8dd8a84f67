"""Network-level driver of the event-driven path: the E/I COBA LIF network
step (``EICOBANet``)."""
from .coba import EICOBANet, expon_coef

__all__ = ["EICOBANet", "expon_coef"]
