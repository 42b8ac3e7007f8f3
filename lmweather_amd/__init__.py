"""lmweather_amd -- the MI355X (gfx950) dynamical core behind lmweather's wx.sphere / wx.plane operators.

The compute path is liblmw_hip.so (hand-written HIP kernels, C ABI in include/lmw.h); this package is the host-side
mirror of the reference classes lmweather::Sphere and lmweather::Plane over that ABI, plus ensembles, the
initial-condition builders of the benchmark configurations and the learned coarse-step corrector.
"""
from ._lib import (BF16X3, BF16X6, FP32, FP64, DeviceError, PlaneParams, SphereParams)  # noqa: F401
from .sphere import Ensemble, Sphere, device_info, fft, gauss_legendre  # noqa: F401
from .plane import Plane  # noqa: F401
