"""GPU check of pcx_fft_batch_noalloc (the allocation-free batched FFT, whose four-step
intermediate is a fixed 128 MiB device array per GPU) against pcx_fft_batch through the C ABI:
bit-identical outputs on consecutive calls, the second with a smaller batch (the array keeps the
first call's contents), forward and inverse, single-pass and four-step sizes, and a batch larger
than the array (several chunks)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(lib, plan, inverse, n, batch, seed):
    import torch

    from packed_amd._lib import check

    st = torch.cuda.current_stream().cuda_stream
    src = torch.empty(2 * n * batch, dtype=torch.float32, device="cuda")
    check(lib.pcx_generate(src.data_ptr(), n, n * batch, seed, 1, 1.0, st))
    want = torch.full_like(src, float("nan"))
    got = torch.full_like(src, float("nan"))
    check(lib.pcx_fft_batch(plan, inverse, 1, want.data_ptr(), n, 2 * n, src.data_ptr(), n, 2 * n, n, batch, st))
    check(lib.pcx_fft_batch_noalloc(plan, inverse, 1, got.data_ptr(), n, 2 * n, src.data_ptr(), n, 2 * n, n, batch,
                                    st))
    torch.cuda.synchronize()
    return got, want


@pytest.mark.parametrize("lg,batches", [(10, (300, 7)), (16, (512, 3)), (20, (130, 2)), (24, (1, 1))])
def test_noalloc_matches_batch(lg, batches):
    import ctypes as C

    import torch

    from packed_amd._lib import check, lib

    L = lib()
    n = 1 << lg
    h = C.c_void_p()
    check(L.pcx_plan_create(n, 2048, 0, C.byref(h)))
    try:
        for inverse in (0, 1):
            for k, batch in enumerate(batches):  # the second call is smaller than the first
                got, want = _pair(L, h, inverse, n, batch, 0x5EED0A00 + 16 * lg + 4 * inverse + k)
                assert not torch.isnan(want).any()
                assert torch.equal(got, want), (lg, inverse, batch)
    finally:
        L.pcx_plan_destroy(h)


def test_noalloc_validates_like_batch():
    import ctypes as C

    import torch

    from packed_amd._lib import ERR_INVALID_ARGUMENT, check, lib

    L = lib()
    h = C.c_void_p()
    check(L.pcx_plan_create(1 << 16, 2048, 0, C.byref(h)))
    buf = torch.zeros(1 << 18, dtype=torch.float32, device="cuda")
    try:
        # batch distance smaller than one vector
        rc = L.pcx_fft_batch_noalloc(h, 0, 1, buf.data_ptr(), 1 << 16, 100, buf.data_ptr(), 1 << 16, 100, 1 << 16, 2,
                                     None)
        assert rc == ERR_INVALID_ARGUMENT
        assert L.pcx_fft_batch_noalloc(None, 0, 1, None, 8, 0, None, 8, 0, 8, 1, None) == ERR_INVALID_ARGUMENT
    finally:
        L.pcx_plan_destroy(h)
