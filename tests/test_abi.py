"""The C ABI library loads on a host without a GPU and exports every entry
point include/cuslines_hip.h declares (no compute calls here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cuslines_hip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^int (cuslines_\w+)\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("cuslines_tracker_create", "cuslines_phase1_async", "cuslines_phase2_async",
                 "cuslines_scan_async", "cuslines_compact_async", "cuslines_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from gpustreamlines_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libcuslines_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    L.cuslines_abi_version.restype = ctypes.c_int
    assert L.cuslines_abi_version() == _lib.ABI_VERSION


def test_params_struct_layout_matches_header():
    """Field order/size of cuslines_params as ctypes sees it == the C struct
    (checked through a compiled probe of offsetof)."""
    import subprocess
    import tempfile

    from gpustreamlines_amd._lib import cuslines_params

    fields = [f for f, _ in cuslines_params._fields_]
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "%s"\nint main(){\n' % HEADER
    src += 'printf("%zu\\n", sizeof(cuslines_params));\n'
    for f in fields:
        src += 'printf("%%zu\\n", offsetof(cuslines_params, %s));\n' % f
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        exe = os.path.join(d, "p")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-o", exe, c], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(cuslines_params)
    assert out[1:] == [getattr(cuslines_params, f).offset for f in fields]


def test_errors_are_reported_without_gpu():
    from gpustreamlines_amd import _lib

    L = _lib.load()
    h = ctypes.c_void_p()
    bad = _lib.params_struct(dict(dimx=0, dimy=1, dimz=1, dimt=1, model_type=2, nverts=1, nedges=1))
    assert L.cuslines_tracker_create(ctypes.byref(bad), 0, ctypes.byref(h)) == 1  # CUSLINES_EINVAL
    assert "positive" in _lib.last_error()


def test_training_batch_limit_is_checked_without_gpu():
    from gpustreamlines_amd import _lib

    L = _lib.load()
    nbytes = ctypes.c_size_t()
    L.cuslines_mlp_workspace_bytes.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_size_t)]
    assert L.cuslines_mlp_workspace_bytes(8_388_480, ctypes.byref(nbytes)) == 0
    # 65535 chunks of 128 pairs, each pair's activations plus the partials
    assert nbytes.value == 4 * (8_388_480 * (4 * 128 + 2 * 24 + 1) + 65535 * 128 * 129)
    assert L.cuslines_mlp_workspace_bytes(8_388_481, ctypes.byref(nbytes)) == 1  # CUSLINES_EINVAL
    assert "8,388,480" in _lib.last_error()
    assert L.cuslines_mlp_workspace_bytes(0, ctypes.byref(nbytes)) == 1


def test_chunk_entry_points_check_arguments_without_gpu():
    """The bounded-chunk and graph entry points refuse null handles before
    any HIP call; destroying a null graph is a no-op."""
    from gpustreamlines_amd import _lib

    L = _lib.load()
    nbytes = ctypes.c_size_t()
    assert L.cuslines_chunk_workspace_bytes(None, 10, 10, ctypes.byref(nbytes)) == 1  # CUSLINES_EINVAL
    args = [None, None, 10, 0, None, None, None, 10, None, None, None, None, None, None, None, 0]
    assert L.cuslines_chunk_async(*args) == 1
    assert "null tracker" in _lib.last_error()
    g = ctypes.c_void_p()
    assert L.cuslines_chunk_graph_create(*args, ctypes.byref(g)) == 1
    assert g.value is None
    assert L.cuslines_graph_launch(None) == 1
    assert L.cuslines_graph_destroy(None) == 0
