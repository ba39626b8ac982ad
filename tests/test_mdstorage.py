"""CPU check of the batch-storage layout: the slice distance of a two-axis md::dynamic_storage
(calc_storage_size, mdstorage.hpp:879-885), at values worked by hand."""
import pytest


@pytest.mark.parametrize("inner,pack,alignment,stride", [
    (1001, 8, 8, 2016),   # lcm(16, 8) = 16: 2002 -> 2016
    (5, 1, 8, 16),        # lcm(2, 8) = 8: 10 -> 16
    (17, 16, 8, 64),      # lcm(32, 8) = 32: 34 -> 64
    (3, 4, 4, 8),         # lcm(8, 4) = 8: 6 -> 8
    (33, 2, 32, 96),      # lcm(4, 32) = 32: 66 -> 96
    (3000, 1, 64, 6016),  # lcm(2, 64) = 64: 6000 -> 6016
    (4096, 8, 8, 8192),   # no padding
])
def test_row_stride(inner, pack, alignment, stride):
    from packed_amd.fft import md_row_stride

    assert md_row_stride(inner, pack, alignment) == stride
