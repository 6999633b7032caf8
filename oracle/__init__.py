"""TEST INFRASTRUCTURE ONLY: the FP64 CPU oracle (see oracle.py)."""
