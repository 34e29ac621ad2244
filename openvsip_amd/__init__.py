"""openvsip_amd -- MI355X (gfx950) backend for OpenVSIP's fast-convolution path.

The product is libovxhip.so (HIP kernels + C-ABI, include/ovxhip.h).  This
package holds its sources (csrc/), the Python ctypes binding used by the
tests and the benchmark (api.py), and seeded workload generators.
"""
from ._lib import OvxError, OvxInvalid, OvxLibraryMissing, OvxUnsupported, LIB_PATH  # noqa: F401
from .api import (COL, FFT_FWD, FFT_INV, ROW, Batch, Convolution, Correlation, Fastconv, Fft,  # noqa: F401
                  Fftm, Replica, init, variants)
from .api import CorrReference, Server, corr_request, fastconv_request, fftm_request  # noqa: F401
from .api import FftmReal, Storage, autotune, autotune_batch, tuned_batch_variant, tuned_variant, vmmul  # noqa: F401
