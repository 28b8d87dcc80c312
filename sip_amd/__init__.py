"""MI355X (gfx950) batched FFT engine: drop-in for the SiP FFT plan/exec path.

The product is libasdfft.so (C ABI in include/fft_api.h); this package builds it
(sip_amd.build) and exposes a Python/torch host side (sip_amd.fft).
"""
from .fft import (Istft, Plan, RaggedPlan, Server, asd_fft_c2c, asd_fft_c2c_sep, asd_fft_c2r, asd_fft_r2c,  # noqa: F401
                  fft_shift, ifft_shift)
