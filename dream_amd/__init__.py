"""dream_amd -- MI355X (gfx950) implementation of DREAM's data-parallel hot paths.

* ``dream_amd.dream_arr_rect`` / ``dreamrect`` / ``dream_arr_circ`` / ``dreamcirc`` /
  ``dream_arr_cylind`` / ``dreamcylind`` / ``dream_arr_annu``: discrete-representation
  spatial impulse responses of rectangular, circular and cylindrically curved elements,
  their arrays and annular arrays, with or without attenuation (drop-ins for the
  reference's Python modules).
* ``dream_amd.das`` / ``das_f``: delay-and-sum in float64 / float32 (drop-ins for the
  reference's ``das`` / ``das_f``); ``dream_amd.das_uni``: DAS over implicit uniform grids.
* ``dream_amd.fftconv_p`` / ``sum_fftconv``: the data-synthesis convolutions.
* ``dream_amd.torch_ops``: device-tensor API -- FP32 data with FP64 delays, linear
  interpolation, apodization, batched frames, the differentiable DAS (``DASFunction``),
  uneven SIR and DAS batches in one persistent launch, and on-device convolution.

All compute runs in ``libdream_hip.so`` (HIP kernels for gfx950, C ABI in
``include/dream_hip.h``); there is no CPU fallback.
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__version__ = "0.2.0"
