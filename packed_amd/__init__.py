"""packed_amd -- MI355X (gfx950) FFT and overlap-save cross-correlation for the packed
split-complex layout, behind a C ABI (include/pcx_gpu.h, libpcx_gpu.so)."""
from ._lib import PcxError  # noqa: F401
from .fft import fft_queue, fft_unit, fxcorr_unit, md_storage, pidx, real_size, vector  # noqa: F401
from .vexpr import conj, repack, store  # noqa: F401

__all__ = ["fft_unit", "fxcorr_unit", "fft_queue", "vector", "md_storage", "pidx", "real_size", "PcxError", "conj", "store",
           "repack"]
