"""GPU parity of the uneven-batch work queue (pcx_queue_*): transforms of every power-of-two size
from 2 to 2^20, mixed directions, at arbitrary offsets, each checked against the FP64 oracle;
eager execution, hipGraph replays (with eager work between them) and the autotuned schedule."""
import numpy as np
import pytest

from helpers import fft_tol

pytestmark = pytest.mark.gpu


def _jobs(rng, sizes):
    jobs, off = [], 0
    for lg in sizes:
        inv = int(rng.integers(0, 2))
        n = 1 << lg
        off += int(rng.integers(0, 3)) * 8  # gaps between vectors
        jobs.append((lg, inv, off, off))
        off += 2 * n
    return jobs, off


def _fill(orc, jobs, total):
    buf = np.zeros(total, np.float32)
    data = []
    for k, (lg, inv, so, _) in enumerate(jobs):
        n = 1 << lg
        z = orc.gen_normal(0x5EED0004 + k, n)
        buf[so:so + n] = z.real
        buf[so + n:so + 2 * n] = z.imag
        data.append(z.astype(np.complex128))
    return buf, data


def _check(orc, jobs, data, out, normalized=True):
    worst = 0.0
    for (lg, inv, _, do), z in zip(jobs, data):
        n = 1 << lg
        got = out[do:do + n].astype(np.float64) + 1j * out[do + n:do + 2 * n]
        want = orc.fft64(z, inverse=bool(inv))
        if inv and normalized:
            want = want / n
        err = orc.rel_l2(got, want)
        assert err < fft_tol(n), (lg, inv, err)
        worst = max(worst, err / fft_tol(n))
    return worst


@pytest.fixture(scope="module")
def pa():
    import packed_amd

    return packed_amd


def _case(orc, seed):
    rng = np.random.default_rng(seed)
    sizes = [int(v) for v in rng.integers(1, 13, 200)] + list(range(1, 21)) + [13, 14, 15, 16, 16, 18]
    rng.shuffle(sizes)
    jobs, total = _jobs(rng, sizes)
    buf, data = _fill(orc, jobs, total)
    return jobs, total, buf, data


def test_queue_eager(orc, pa):
    import torch

    jobs, total, buf, data = _case(orc, 1)
    src = torch.from_numpy(buf).cuda()
    dst = torch.zeros_like(src)
    q = pa.fft_queue(jobs)
    q.exec(dst.data_ptr(), src.data_ptr())
    torch.cuda.synchronize()
    _check(orc, jobs, data, dst.cpu().numpy())
    f, b = q.work()
    assert b == sum(16 * (1 << lg) for lg, *_ in jobs)


def test_queue_in_place_unnormalized(orc, pa):
    import torch

    jobs, total, buf, data = _case(orc, 2)
    x = torch.from_numpy(buf).cuda()
    q = pa.fft_queue(jobs, normalized=False)
    q.exec(x.data_ptr(), x.data_ptr())
    torch.cuda.synchronize()
    _check(orc, jobs, data, x.cpu().numpy(), normalized=False)


def test_queue_graph_replays(orc, pa):
    """Captured pass replayed three times with eager work in between; every replay checked."""
    import torch

    jobs, total, buf, data = _case(orc, 3)
    src = torch.from_numpy(buf).cuda()
    dst = torch.zeros_like(src)
    q = pa.fft_queue(jobs)
    s = torch.cuda.Stream()
    q.capture(dst.data_ptr(), src.data_ptr(), s)
    for rep in range(3):
        dst.zero_()
        torch.cuda.synchronize()
        q.launch(s)
        s.synchronize()
        _check(orc, jobs, data, dst.cpu().numpy())
        junk = torch.randn(1 << 20, device="cuda") * 2  # eager work between replays
        torch.cuda.synchronize()
        del junk


def test_queue_autotune_and_cache(orc, pa, tmp_path):
    import torch

    jobs, total, buf, data = _case(orc, 4)
    src = torch.from_numpy(buf).cuda()
    dst = torch.zeros_like(src)
    cache = str(tmp_path / "tune.txt")
    q = pa.fft_queue(jobs)
    q.autotune(dst.data_ptr(), src.data_ptr(), cache)
    dst.zero_()
    q.exec(dst.data_ptr(), src.data_ptr())
    torch.cuda.synchronize()
    _check(orc, jobs, data, dst.cpu().numpy())
    text = open(cache).read()
    assert "queue_grid " in text
    q2 = pa.fft_queue(jobs)
    q2.autotune(dst.data_ptr(), src.data_ptr(), cache)  # reads the cache
    dst.zero_()
    q2.exec(dst.data_ptr(), src.data_ptr())
    torch.cuda.synchronize()
    _check(orc, jobs, data, dst.cpu().numpy())


def test_capture_then_autotune_then_capture(orc, pa):
    """Autotuning after a capture drops the captured graph (its kernels, grids and twiddle tables
    may all change); launching it fails until the queue is captured again, and the new graph
    gives the oracle's results.  Autotuning refuses dst == src (it overwrites dst many times)."""
    import torch

    from packed_amd import PcxError

    jobs, total, buf, data = _case(orc, 5)
    src = torch.from_numpy(buf).cuda()
    dst = torch.zeros_like(src)
    q = pa.fft_queue(jobs)
    s = torch.cuda.Stream()
    q.capture(dst.data_ptr(), src.data_ptr(), s)
    with pytest.raises(PcxError):
        q.autotune(src.data_ptr(), src.data_ptr())
    q.autotune(dst.data_ptr(), src.data_ptr())
    with pytest.raises(PcxError):
        q.launch(s)
    q.capture(dst.data_ptr(), src.data_ptr(), s)
    dst.zero_()
    torch.cuda.synchronize()
    q.launch(s)
    s.synchronize()
    _check(orc, jobs, data, dst.cpu().numpy())


def test_queue_above_persistent_range(orc, pa):
    """Sizes above 2^20 run as grouped launches after the persistent one, in the same pass."""
    import torch

    rng = np.random.default_rng(6)
    jobs, total = _jobs(rng, [21, 5, 16, 21, 20, 12])
    buf, data = _fill(orc, jobs, total)
    src = torch.from_numpy(buf).cuda()
    dst = torch.zeros_like(src)
    q = pa.fft_queue(jobs)
    q.exec(dst.data_ptr(), src.data_ptr())
    torch.cuda.synchronize()
    _check(orc, jobs, data, dst.cpu().numpy())
