"""MI355X-native scattering-device hot path of Sassena (coherent "all" and
incoherent "self" intermediate scattering functions with orientational
averaging and FFT time autocorrelation).  The compute path is
libsassena_amd.so (hand-written gfx950 HIP behind the C ABI of
include/sassena_amd.h); this package is its Python host side."""
from .device import ScatterDevice, pad_length, version  # noqa: F401
from ._lib import SxError  # noqa: F401
from .batch import Batch  # noqa: F401
