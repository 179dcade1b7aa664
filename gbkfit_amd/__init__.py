"""gbkfit_amd: an MI355X (gfx950) driver for gbkfit's model-evaluation hot
path -- smdisk/mcdisk evaluation into a spectral cube, supersample
downscaling, PSF/LSF FFT convolution, moment maps and the residual / chi^2
objective -- behind gbkfit's own driver/backend interface.

Device work runs in hand-written HIP kernels (libgbkhip.so, C ABI in
include/gbkhip.h); torch-ROCm tensors only own device buffers.
"""
__version__ = '0.1.0'

from . import _native  # noqa: F401
