"""TEST INFRASTRUCTURE ONLY: CPU checkers for the PO / ray-trace path (see bind.py)."""
