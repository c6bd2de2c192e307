"""Tracking constants of the reference (cuslines/cuda_python/_globals.py:1-8)."""

EXCESS_ALLOC_FACT = 2
MAX_SLINES_PER_SEED = 10
MAX_SLINE_LEN = 501
PMF_THRESHOLD_P = 0.05
REAL_SIZE = 4
THR_X_BL = 64
THR_X_SL = 32
NORM_EPS = 1e-12
