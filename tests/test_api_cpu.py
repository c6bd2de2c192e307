"""Host-side API behaviour that needs no GPU: argument validation in the
order the reference performs it, getter data preparation, output
containers."""

import numpy as np
import pytest

from gpustreamlines_amd import direction_getters as D
from gpustreamlines_amd.array_sequence import StreamlineArray


def test_boot_model_type_validation():
    with pytest.raises(ValueError):
        D.BootDirectionGetter("dti", 1, np.eye(2), np.eye(2), np.zeros((1, 2)), np.zeros((1, 2)),
                              np.zeros((3, 1)), np.array([1, 0, 0]))


def test_boot_from_gradient_table():
    from gpustreamlines_amd.sphere import single_shell_gtab, symmetric_sphere

    v, _ = symmetric_sphere(724)
    g = single_shell_gtab()
    for make in (D.BootDirectionGetter.from_dipy_csa, D.BootDirectionGetter.from_dipy_opdt):
        dg = make(g, v)
        assert dg.H.shape == (160, 160) and dg.delta_nr == 28
        assert dg.sampling_matrix.shape == (724, 28)
        assert dg.b0s_mask.sum() == 10


def test_tight_fit_tiling():
    assert D.calculate_tight_fit(724) == (16, 8, 8)
    assert D.calculate_tight_fit(1) == (1, 1, 1)
    assert D.calculate_tight_fit(100) == (8, 4, 4)


def test_ptt_prepare_data():
    from gpustreamlines_amd.sphere import symmetric_sphere

    v, _ = symmetric_sphere(724)
    rng = np.random.default_rng(0)
    data = rng.normal(size=(4, 5, 6, 724)).astype(np.float32)
    stop = rng.random((4, 5, 6)).astype(np.float32)
    dg = D.PttDirectionGetter(odf_lut_res=16)
    out = dg.prepare_data(data, stop, 0.3, v, False)
    assert out.min() >= 0
    assert np.all(out[stop < 0.3] == 0)
    mx = out.max(axis=3)
    assert np.allclose(mx[mx > 0], 1.0)
    assert (dg.log2_x, dg.log2_y, dg.log2_z) == (4, 3, 3)
    lut = dg.sphere_vertices_lut_h
    assert lut.shape == (16, 16, 16)
    g = np.linspace(-1, 1, 16)
    for i, j, k in [(0, 0, 0), (15, 3, 7), (8, 8, 8)]:
        p = np.array([g[i], g[j], g[k]])
        assert lut[i, j, k] == np.argmin(np.linalg.norm(v - p, axis=1))
    assert np.array_equal(data.shape, out.shape)


def test_tracker_rejects_mismatched_sphere():
    from gpustreamlines_amd import GPUTracker, ProbDirectionGetter

    with pytest.raises(ValueError):
        GPUTracker(ProbDirectionGetter(), np.zeros((2, 2, 2, 5), np.float32), np.zeros((2, 2, 2), np.float32),
                   0.1, np.zeros((4, 3), np.float32), np.zeros((1, 2), np.int32))


def test_tracker_rejects_more_gpus_than_present():
    """Asking for more GPUs than the machine has raises RuntimeError before
    any device memory is touched (cu_tractography.py:161-164 raises the same
    type); without any GPU the failed device count is that error too."""
    from gpustreamlines_amd import GPUTracker, ProbDirectionGetter

    v = np.eye(3, dtype=np.float32)
    with pytest.raises(RuntimeError):
        GPUTracker(ProbDirectionGetter(), np.zeros((2, 2, 2, 3), np.float32), np.zeros((2, 2, 2), np.float32),
                   0.1, v, np.array([[0, 1], [1, 2]], np.int32), ngpus=10**6)


def test_streamline_array():
    a = StreamlineArray([np.zeros((3, 3)), np.ones((2, 3))])
    assert len(a) == 2 and a[1].shape == (2, 3) and a._data.shape == (5, 3)
    b = StreamlineArray.from_packed(a._data, a._lengths)
    assert [x.shape for x in b] == [(3, 3), (2, 3)]


def test_ptt_prepare_on_device_defers_to_the_gpu():
    """prepare_on_device: prepare_data keeps the raw ODF (the GPU clips,
    masks and normalises it) and only chooses the tiling, like the host
    recipe does."""
    from gpustreamlines_amd.sphere import symmetric_sphere

    v, _ = symmetric_sphere(724)
    rng = np.random.default_rng(1)
    data = rng.normal(size=(3, 4, 5, 724)).astype(np.float32)
    stop = rng.random((3, 4, 5)).astype(np.float32)
    dev = D.PttDirectionGetter(odf_lut_res=16, prepare_on_device=True)
    out = dev.prepare_data(data, stop, 0.3, v, False)
    assert np.array_equal(out, data)
    host = D.PttDirectionGetter(odf_lut_res=16)
    host.prepare_data(data, stop, 0.3, v, False)
    assert (dev.log2_x, dev.log2_y, dev.log2_z) == (host.log2_x, host.log2_y, host.log2_z)
    assert dev.sphere_vertices_lut_h is None
    s, thr, symm = dev._device_inputs
    assert np.array_equal(s, stop) and thr == 0.3 and symm is False
