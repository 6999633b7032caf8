"""MI355X-native P3M long-range Coulomb solver (drop-in for the k-space part
of ESPResSo's P3M / P3MGPU actors)."""
from . import electrostatics  # noqa: F401
from .batch import P3MBatch, P3MServer, tune_model  # noqa: F401

__all__ = ["electrostatics", "P3MBatch", "P3MServer", "tune_model"]
