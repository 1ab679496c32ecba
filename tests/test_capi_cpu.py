"""The C-ABI library loads and exports every symbol include/mctomo_hip.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mctomo_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", src, flags=re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ("fastMarching3D", "mct_voronoi_to_grid", "mct_body_create", "mct_body_likelihood",
              "mct_body_likelihood_batch_dev", "mct_destroy", "mct_sync", "mct_version"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import mctomo_amd
    path = mctomo_amd.lib_path()
    if not os.path.exists(path):
        pytest.fail("libmctomo_hip.so is not built (python -m mctomo_amd._build)")
    L = ctypes.CDLL(path)
    for n in declared_functions():
        assert hasattr(L, n), n
    assert mctomo_amd.version().startswith("mctomo-hip")


def test_create_without_gpu_reports_hip_error_not_crash():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import mctomo_amd
    import numpy as np
    g = mctomo_amd.Grid(5, 5, 5, 0, 0, 0, 1, 1, 1)
    with pytest.raises(mctomo_amd.MCTError):
        mctomo_amd.BodyLikelihood(g, np.zeros((1, 3)), np.zeros((1, 3, 1)), np.ones((1, 2, 1)), 1)
