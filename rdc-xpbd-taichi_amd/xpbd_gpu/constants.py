"""Shared numerical tolerance of the solver (xpbd_gpu/constants.py:3)."""

EPSILON = 1e-12
