"""The host check behind the MFMA bootstrap kernel's low-rank form
(cuslines_hat_lowrank: H = U U^T of rank r, R its leveraged centred residual
matrix, boot_utils.py:86-94), on the CPU: the getter matrices of the order-6
and order-8 bases have it (rank = number of basis functions); a perturbed R,
a hat matrix whose R was built for another H, and a matrix that is not a
projection do not."""

import ctypes as C

import numpy as np
import pytest


def _rank(H, R):
    from gpustreamlines_amd import _lib

    L = _lib.load()
    H = np.ascontiguousarray(H, dtype=np.float32)
    R = np.ascontiguousarray(R, dtype=np.float32)
    r = C.c_int32(-1)
    assert L.cuslines_hat_lowrank(H.ctypes.data, R.ctypes.data, len(H), C.byref(r)) == 0
    return r.value


def _bvecs(n, seed):
    v = np.random.default_rng(seed).normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("order,n,rank", [(6, 160, 28), (8, 90, 45), (4, 30, 15)])
def test_getter_matrices_have_the_form(order, n, rank):
    from gpustreamlines_amd import shm

    H, R = shm.hat_and_lcr(_bvecs(n, order), sh_order_max=order)
    assert _rank(H, R) == rank


def test_other_matrices_do_not():
    from gpustreamlines_amd import shm

    H, R = shm.hat_and_lcr(_bvecs(160, 1), sh_order_max=6)
    bad = np.array(R, dtype=np.float32)
    bad[3, 5] += 1e-3
    assert _rank(H, bad) == 0
    H2, R2 = shm.hat_and_lcr(_bvecs(160, 2), sh_order_max=6)
    assert _rank(H, R2) == 0
    M = np.random.default_rng(3).normal(size=(160, 160)) * 0.01
    assert _rank(M @ M.T, shm.lcr_matrix(M @ M.T)) == 0
    assert _rank(np.eye(40), shm.lcr_matrix(np.eye(40) * 0.5)) == 0
