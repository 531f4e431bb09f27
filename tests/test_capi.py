"""The C-ABI library loads on a host without a GPU and exports every entry
point include/mmnorm_amd.h declares (no compute call is made)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mmnorm_amd.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|size_t|const char \*)\s*(mmn_\w+)\(", src, re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    for want in ("mmn_cuda_image", "mmn_cuda_sn_est", "mmn_bp_forward",
                 "mmn_bp_forward_workspace_size", "mmn_sn_from_moment",
                 "mmn_bp_backward", "mmn_bp_backward_workspace_size", "mmn_rsdf_bfs",
                 "mmn_rsdf_bfs_workspace_size", "mmn_li_threshold",
                 "mmn_li_threshold_workspace_size", "mmn_ccl6", "mmn_ccl6_workspace_size",
                 "mmn_label_sizes", "mmn_abi_version", "mmn_last_error", "mmn_bp_profile",
                 "mmn_bp_profile_workspace_size", "mmn_bp_profile_backward",
                 "mmn_bp_profile_backward_workspace_size"):
        assert want in names


def test_library_exports_every_declared_symbol():
    from mmnorm_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmmnorm_hip.so not built")
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name
    assert lib.mmn_abi_version() == 1


def test_workspace_queries_and_argument_errors_without_gpu():
    from mmnorm_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmmnorm_hip.so not built")
    lib = _lib.load()
    assert lib.mmn_rsdf_bfs_workspace_size(4, 4, 4) > 4 * 4 * 4 * 12
    assert lib.mmn_rsdf_bfs_workspace_size(0, 4, 4) == 0
    assert lib.mmn_bp_backward_workspace_size(16, 32, 100) > 16 * 32 * 8
    assert lib.mmn_bp_forward_workspace_size(1000, 3200, 1) == 32 * 1000 * 32
    assert lib.mmn_bp_forward_workspace_size(1000, 20, 0) == 0
    # Without a device the plan assumes 256 CUs x 8 resident workgroups of 512
    # voxels.  128x128x96 voxels, 3600 rows, with the moment: 3072 voxel
    # blocks, fewer than 8 rounds of 2048, so every block is split into the 6
    # row segments that give 8 rounds, with 32-byte partials per voxel.
    assert lib.mmn_bp_forward_workspace_size(128 * 128 * 96, 3600, 1) == 6 * 128 * 128 * 96 * 32
    assert lib.mmn_bp_forward_workspace_size(128 * 128 * 96, 3600, 0) == 6 * 128 * 128 * 96 * 8
    # 256x256x128 voxels: 16384 blocks are 8 rounds already, no split, no workspace
    assert lib.mmn_bp_forward_workspace_size(256 * 256 * 128, 10000, 1) == 0
    # the 550x600x250 real-data grid with 16 rows: too few rows to split
    assert lib.mmn_bp_forward_workspace_size(550 * 600 * 250, 16, 1) == 0
    # argument checks run before any HIP call
    with pytest.raises(ValueError, match="NULL"):
        _lib.call("mmn_bp_forward", None, None, None, None, None, None, 0, None, None, None,
                  77e9, 1e6, 1, 1, 1, 1, 1, 1, 0, 2, 32, None, 0, None)
    with pytest.raises(ValueError, match="variant"):
        _lib.call("mmn_bp_forward", 8, None, 8, 8, 8, None, 0, 8, 8, 8, 77e9, 1e6, 1, 1, 1,
                  1, 1, 1, 0, 7, 32, None, 0, None)
    with pytest.raises(ValueError, match="K must be"):
        _lib.call("mmn_bp_forward", 8, None, 8, 8, 8, None, 0, 8, 8, 8, 77e9, 1e6, 1, 1, 1,
                  1, 1, 1, 0, 2, 12, None, 0, None)
    with pytest.raises(ValueError, match="K must be 16 or 32"):
        _lib.call("mmn_bp_backward", 8, 8, 8, 1.0, 8, 8, 8, None, 0, 8, 8, 8, 77e9, 1e6,
                  1, 1, 1, 1, 1, 1, 0, 48, None, 0, None)
    with pytest.raises(ValueError, match="workspace"):
        _lib.call("mmn_bp_backward", 8, 8, 8, 1.0, 8, 8, 8, None, 0, 8, 8, 8, 77e9, 1e6,
                  1, 1, 1, 1, 1, 1, 0, 32, None, 0, None)
    with pytest.raises(ValueError, match="workspace"):
        _lib.call("mmn_rsdf_bfs", 8, 8, 8, 8, 8, 1, 1e-3, 4, 4, 4, None, 0, None)
    # the no-allocation BFS refuses a grid beyond its 128 MiB scratch before any HIP call
    with pytest.raises(ValueError, match="scratch"):
        _lib.call("mmn_rsdf_bfs_noalloc", 8, 8, 8, 8, 8, 1, 1e-3, 2048, 2048, 4, None)
    assert lib.mmn_ccl6_workspace_size(4, 4, 4) >= 4 * 4 * 4 * 4
    assert lib.mmn_li_threshold_workspace_size(100) >= 800
    with pytest.raises(ValueError, match="workspace"):
        _lib.call("mmn_ccl6", 8, None, 8, 4, 4, 4, None, 0, None)
    with pytest.raises(ValueError, match="exponent"):
        _lib.call("mmn_li_threshold", 8, None, None, None, 8, 10, -1.0, 10, None, 0, None)
    with pytest.raises(ValueError, match="NULL"):
        _lib.call("mmn_label_sizes", None, 8, 10, 3, None)
    with pytest.raises(ValueError, match="sizes"):
        _lib.call("mmn_cuda_image", 8, 8, 8, 8, 8, 8, 8, 1.0, 1.0, 1.0, 0, 1, 1, 1, 1, 1, 0,
                  0, None)
    # range-profile pair: workspace sizes from the bounding box, argument checks
    bb = (ctypes.c_double * 6)(0.0, 0.0, 0.0, 0.1, 0.1, 0.05)
    bad = (ctypes.c_double * 6)(0.0, 0.0, 0.0, -0.1, 0.1, 0.05)
    pb, pbad = ctypes.addressof(bb), ctypes.addressof(bad)
    # M = ceil(2 |diag| / h) + 6 table points per row, h = c / (64 (F-1) df):
    # tables of 8 + 32 bytes per point at least
    h = 2.99792458e8 / (64 * 63 * 62.5e6)
    M = int(np.ceil(2 * np.sqrt(0.1 ** 2 + 0.1 ** 2 + 0.05 ** 2) / h)) + 6
    assert lib.mmn_bp_profile_workspace_size(1000, 100, 64, 77e9, 62.5e6, pb, 1) >= 100 * M * 40
    assert lib.mmn_bp_profile_workspace_size(1000, 100, 64, 77e9, 62.5e6, pbad, 1) == 0
    assert lib.mmn_bp_profile_workspace_size(1000, 100, 64, float("nan"), 62.5e6, pb, 1) == 0
    assert lib.mmn_bp_profile_backward_workspace_size(1000, 100, 64, 77e9, 62.5e6, pb) > 0
    assert lib.mmn_bp_profile_backward_workspace_size(0, 100, 64, 77e9, 62.5e6, pb) == 0
    with pytest.raises(ValueError, match="NULL"):
        _lib.call("mmn_bp_profile", 8, None, 8, 8, 8, None, 0, 8, 8, 8, 77e9, 62.5e6, 1, 1, 1,
                  1, 1, 64, 0, None, None, 0, None)
    with pytest.raises(ValueError, match="bounding box"):
        _lib.call("mmn_bp_profile", 8, None, 8, 8, 8, None, 0, 8, 8, 8, 77e9, 62.5e6, 1, 1, 1,
                  1, 1, 64, 0, pbad, None, 0, None)
    with pytest.raises(ValueError, match="workspace"):
        _lib.call("mmn_bp_profile", 8, None, 8, 8, 8, None, 0, 8, 8, 8, 77e9, 62.5e6, 1, 1, 1,
                  1, 1, 64, 0, pb, None, 0, None)
    with pytest.raises(ValueError, match="finite"):
        _lib.call("mmn_bp_profile_backward", 8, 8, 8, 1.0, 8, 8, 8, None, 0, 8, 8, 8,
                  float("inf"), 62.5e6, 1, 1, 1, 1, 1, 64, 0, pb, None, 0, None)
    with pytest.raises(ValueError, match="workspace"):
        _lib.call("mmn_bp_profile_backward", 8, 8, 8, 1.0, 8, 8, 8, None, 0, 8, 8, 8, 77e9,
                  62.5e6, 1, 1, 1, 1, 1, 64, 0, pb, None, 0, None)
    ctypes.get_errno()
