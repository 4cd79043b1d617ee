"""dream_amd -- MI355X (gfx950) implementation of DREAM's two data-parallel hot paths.

* ``dream_amd.dream_arr_rect.dream_arr_rect`` / ``dream_amd.dreamrect.dreamrect``:
  discrete-representation spatial impulse responses (drop-ins for the
  reference's Python modules of the same names).
* ``dream_amd.das.das`` / ``dream_amd.das_f.das``: delay-and-sum in float64 /
  float32 (drop-ins for the reference's ``das`` / ``das_f``).
* ``dream_amd.torch_ops``: device-tensor API -- FP32 data with FP64 delays,
  linear interpolation, apodization, batched frames, the differentiable DAS
  (``DASFunction``) and batched uneven SIR plans.

All compute runs in ``libdream_hip.so`` (HIP kernels for gfx950, C ABI in
``include/dream_hip.h``); there is no CPU fallback.
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__version__ = "0.1.0"
