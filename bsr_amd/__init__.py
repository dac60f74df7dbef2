"""MI355X-native radial core of bsr_mat: B-spline Slater integrals R^k in
band storage and their contraction with the angular-coefficient lists into
the channel blocks of H and S.  The compute path is the HIP library
libbsr_gpu.so (C ABI in include/bsr_gpu.h); this package is its host side."""
from .core import Context  # noqa: F401
from .lib import BsrGpuError  # noqa: F401
