"""Tracking constants of the reference (cuslines/cuda_python/_globals.py:1-8)
that the host side uses: the streamline slot length, the float size of the
output buffers and the launch shape of the reference's kernels."""

MAX_SLINE_LEN = 501
REAL_SIZE = 4
THR_X_BL = 64
THR_X_SL = 32
