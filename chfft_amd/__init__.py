"""chfft_amd — chfft's transform engine on MI355X (gfx950) HIP kernels.

The public names mirror chfft 0.3.4 (lib.rs:21-25). All compute runs in
libchfft_hip.so through the C ABI of include/chfft_hip.h.
"""
from ._lib import build, lib, LIB_PATH, ChfftError  # noqa: F401
from .api import (  # noqa: F401
    CFft1D, CFft2D, Dct1D, DctType, Mdct1D, MultiFFT, Plan, RFft1D, sine_window, vorbis_window,
)
