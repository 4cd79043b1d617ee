"""The C-ABI library loads on the CPU and exports every symbol that
include/dream_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dream_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dream_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_all_exports():
    from dream_amd import _lib
    assert sorted(_lib.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    import dream_amd
    L = dream_amd.lib()
    for name in _declared():
        assert hasattr(L, name), name
    # Symbols are plain C (unmangled) in the shared object.
    raw = ctypes.CDLL(dream_amd.LIB_PATH)
    for name in _declared():
        assert getattr(raw, name) is not None


def test_cpu_only_calls():
    import dream_amd
    L = dream_amd.lib()
    assert L.dream_hip_version().startswith(b"dream_hip")
    assert L.dream_hip_device_count() >= 0
    L.dream_hip_abort(2)
    assert L.dream_hip_is_running() == 0
    L.dream_hip_set_running()
    assert L.dream_hip_is_running() == 1
    assert L.dream_hip_das_setup_has_changed(None, 1, 10, 10, 2, 2, 0) == 1


def test_api_checks_before_device_use():
    """Argument checks of the reference's gateways run before any GPU call."""
    import numpy as np
    from dream_amd import das, dream_arr_rect
    with pytest.raises(RuntimeError, match="Error in dream_arr_rect"):
        dream_arr_rect.dream_arr_rect(np.zeros((4, 2)), [1, 2], np.zeros((3, 3)), [1, 1, 1, 10],
                                      0.0, [1, 1500, 0], "off", 0.0, "off", [0, 0], "off", None,
                                      0.0, "stop")
    with pytest.raises(RuntimeError, match="Error in dream_arr_rect"):
        dream_arr_rect.dream_arr_rect(np.zeros((4, 3)), [1, 2], np.zeros((3, 3)), [1, 1, 1, 10],
                                      0.0, [1, 1500, 0], "bogus", 0.0, "off", [0, 0], "off",
                                      None, 0.0, "stop")
    with pytest.raises(RuntimeError, match="Argument 2"):
        das.das(np.zeros((10, 4)), np.zeros((4, 2)), None, np.zeros((5, 3)), 0.02, 0.0, 1500.0,
                "saft", "ignore", "gpu")
    with pytest.raises(RuntimeError, match="Error in das"):
        das.das(np.zeros((10, 4)), np.zeros((4, 3)), None, np.zeros((5, 3)), 0.02, [0.0, 1.0],
                1500.0, "saft", "ignore", "gpu")
