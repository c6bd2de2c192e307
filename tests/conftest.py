import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


SMALL_SHAPE = (32, 36, 30)


@pytest.fixture(scope="session")
def small_workloads():
    """Small synthetic workloads of every getter, built once per session."""
    from gpustreamlines_amd import workloads as W

    return {k: W.build(k, shape=SMALL_SHAPE) for k in ("prob", "ptt", "csa", "opdt")}


def packed_points(res):
    """Valid points of every streamline of an oracle result, concatenated."""
    sl, lens = res["slines"], res["sline_len"]
    if len(lens) == 0:
        return np.zeros((0, 3), dtype=np.float32)
    return np.concatenate([sl[i, :n] for i, n in enumerate(lens)])
