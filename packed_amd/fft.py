"""Python mirror of the reference's FFT-unit and cross-correlation interfaces over the C ABI.

``vector``      pcxo::vector<T, P> (vector.hpp:26-382): n complex values in packs of P reals
                then P imaginaries, ceil(n/P)*2P elements, zero-filled -- here resident in device
                memory (a torch float32 / float64 tensor on the GPU; torch is only the allocator).
                The default pack is 32 / sizeof(T) (types.hpp:70-72): 8 for float, 4 for double.
``fft_unit``    pcxo::fft_unit<T, Order> (fft.hpp:565-915), T = float or double: size(),
                operator() (forward, in or out of place, zero-extended source), ifft<Normalized>(),
                fft_raw / ifft_raw on raw device pointers.
``fxcorr_unit`` pcxo::fxcorr_unit<float> (fxcorr.hpp:44-147): construct from the kernel g and an
                FFT size, call on a signal (in place), or out of place.

Errors raise ``PcxError`` (a ValueError, the counterpart of std::invalid_argument) carrying the
reference's messages.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import ORDER, PRECISION, PcxError, check, lib

__all__ = ["vector", "md_storage", "fft_unit", "fxcorr_unit", "fft_queue", "PcxError", "real_size", "pidx"]


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("packed_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


def real_size(n: int, pack: int) -> int:
    """vector.hpp:376-378: floats occupied by n complex values in pack P."""
    return ((n + pack - 1) // pack) * pack * 2


def pidx(i, pack: int):
    """element_access.hpp:8-11: float offset of the real part of complex i."""
    i = np.asarray(i, dtype=np.int64)
    return i + (i // pack) * pack


def _stream_ptr(stream) -> int | None:
    if stream is None:
        torch = _torch()
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


DTYPES = ("float32", "float64")


def default_pack(dtype: str = "float32") -> int:
    """types.hpp:70-72 default_pack_size<T> = 32 / sizeof(T)."""
    return 8 if dtype == "float32" else 4


class vector:
    """Device-resident packed complex vector (pcxo::vector<T, P> layout), T = float or double."""

    def __init__(self, n: int, pack: int | None = None, data=None, dtype: str | None = None):
        if data is not None and dtype is None:
            dtype = str(data.dtype).replace("torch.", "")
        dtype = dtype or "float32"
        if dtype not in DTYPES:
            raise PcxError(_lib.ERR_INVALID_ARGUMENT, f"element type {dtype!r} is not float32 or float64")
        pack = default_pack(dtype) if pack is None else pack
        if pack < 1 or pack & (pack - 1):
            raise PcxError(_lib.ERR_INVALID_PACK, f"pack size (which is {pack}) is not a power of two")
        torch = _torch()
        self.n = int(n)
        self.pack = int(pack)
        self.dtype = dtype
        if data is None:
            data = torch.zeros(real_size(self.n, self.pack), dtype=getattr(torch, dtype), device="cuda")
        self.data = data

    def size(self) -> int:
        return self.n

    def __len__(self) -> int:
        return self.n

    def data_ptr(self) -> int:
        return self.data.data_ptr()

    @classmethod
    def from_complex(cls, z, pack: int | None = None, dtype: str = "float32") -> "vector":
        torch = _torch()
        z = np.asarray(z)
        pack = default_pack(dtype) if pack is None else pack
        buf = np.zeros(real_size(len(z), pack), np.dtype(dtype))
        idx = pidx(np.arange(len(z)), pack)
        buf[idx] = np.real(z)
        buf[idx + pack] = np.imag(z)
        return cls(len(z), pack, torch.from_numpy(buf).cuda(), dtype)

    # vector_arithm.hpp:450-495: lazy expressions, evaluated on the GPU by store()
    def _expr(self):
        from .vexpr import Leaf

        return Leaf(self)

    def __add__(self, o):
        return self._expr() + o

    def __radd__(self, o):
        return o + self._expr()

    def __sub__(self, o):
        return self._expr() - o

    def __rsub__(self, o):
        return o - self._expr()

    def __mul__(self, o):
        return self._expr() * o

    def __rmul__(self, o):
        return o * self._expr()

    def __truediv__(self, o):
        return self._expr() / o

    def __rtruediv__(self, o):
        return o / self._expr()

    def to_complex(self, n: int | None = None, pack: int | None = None) -> np.ndarray:
        """Complex values read in `pack` layout (default: this vector's pack)."""
        n = self.n if n is None else n
        pack = self.pack if pack is None else pack
        buf = self.data.cpu().numpy()
        idx = pidx(np.arange(n), pack)
        return buf[idx].astype(np.float64) + 1j * buf[idx + pack].astype(np.float64)


def md_row_stride(inner: int, pack: int, alignment: int) -> int:
    """Reals between slices of a two-axis md::dynamic_storage (calc_storage_size,
    mdstorage.hpp:879-885): ceil(2*inner / A) * A with A = lcm(2*pack, alignment)."""
    import math

    a = math.lcm(2 * pack, alignment)
    return (2 * inner + a - 1) // a * a


class md_storage:
    """Two-axis md::dynamic_storage<T, basis, P, Alignment> (mdstorage.hpp:1478-1499) in device
    memory: `outer` slices of `inner` packed complex values, row_stride reals apart, with
    row_stride = ceil(2*inner / A)*A, A = lcm(2P, Alignment) (calc_storage_size,
    mdstorage.hpp:879-885); zero-filled.  storage[i] is slice i as a `vector` view (no copy).
    Extents are in layout order: the inner (contiguous) axis first."""

    def __init__(self, inner: int, outer: int, pack: int | None = None, alignment: int | None = None,
                 dtype: str = "float32"):
        if dtype not in DTYPES:
            raise PcxError(_lib.ERR_INVALID_ARGUMENT, f"element type {dtype!r} is not float32 or float64")
        pack = default_pack(dtype) if pack is None else pack
        alignment = default_pack(dtype) if alignment is None else alignment
        for what, v in (("pack", pack), ("alignment", alignment)):
            if v < 1 or v & (v - 1):
                raise PcxError(_lib.ERR_INVALID_PACK, f"{what} size (which is {v}) is not a power of two")
        torch = _torch()
        self.inner, self.outer, self.pack, self.dtype = int(inner), int(outer), int(pack), dtype
        self.row_stride = md_row_stride(self.inner, self.pack, int(alignment))
        self.data = torch.zeros(self.row_stride * self.outer, dtype=getattr(torch, dtype), device="cuda")

    def size(self) -> int:
        """Extent of the outer axis (mdstorage.hpp:1466-1468)."""
        return self.outer

    def __len__(self) -> int:
        return self.outer

    def data_ptr(self) -> int:
        return self.data.data_ptr()

    def __getitem__(self, i: int) -> vector:
        if not 0 <= i < self.outer:
            raise IndexError(i)
        view = self.data[i * self.row_stride: i * self.row_stride + real_size(self.inner, self.pack)]
        return vector(self.inner, self.pack, view, self.dtype)


class fft_unit:
    """fft.hpp:565-915 fft_unit<T, Order> on the GPU; dtype "float64" makes fft_unit<double, Order>."""

    def __init__(self, fft_size: int, sub_size: int = 2048, order: str = "normal", precision: str = "fp32",
                 dtype: str = "float32"):
        if order not in ORDER:
            raise PcxError(_lib.ERR_INVALID_ARGUMENT, f"unknown fft_order {order!r}")
        if dtype not in DTYPES:
            raise PcxError(_lib.ERR_INVALID_ARGUMENT, f"element type {dtype!r} is not float32 or float64")
        import ctypes as C

        h = C.c_void_p()
        create = lib().pcx_plan_create_f64 if dtype == "float64" else lib().pcx_plan_create
        check(create(int(fft_size), int(sub_size), ORDER[order], C.byref(h)))
        self._h = h
        self.order = order
        self.dtype = dtype
        f64 = dtype == "float64"
        self._fft = lib().pcx_fft_f64 if f64 else lib().pcx_fft
        self._ifft = lib().pcx_ifft_f64 if f64 else lib().pcx_ifft
        self._batch = lib().pcx_fft_batch_f64 if f64 else lib().pcx_fft_batch
        if precision != "fp32":
            check(lib().pcx_plan_set_precision(h, PRECISION[precision]))
        self.precision = precision

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.pcx_plan_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def size(self) -> int:
        return int(lib().pcx_plan_size(self._h))

    # ---------------------------------------------------------------- launch-shape tuning
    def set_shape(self, n2_log2: int = 0, cols: int = 0, rows: int = 0) -> None:
        """Four-step split 2^n2_log2 column length, `cols` columns per column tile, `rows` rows per
        row workgroup (0 = any; pcx_plan_set_shape)."""
        check(lib().pcx_plan_set_shape(self._h, int(n2_log2), int(cols), int(rows)))

    def shape(self):
        import ctypes as C

        a, c, g = C.c_int(), C.c_int(), C.c_int()
        check(lib().pcx_plan_shape(self._h, C.byref(a), C.byref(c), C.byref(g)))
        return a.value, c.value, g.value

    def set_grid(self, grid_col: int, grid_row: int) -> None:
        check(lib().pcx_plan_set_grid(self._h, int(grid_col), int(grid_row)))

    def set_chunk(self, nbytes: int) -> None:
        check(lib().pcx_plan_set_chunk(self._h, int(nbytes)))

    def set_one_pass(self, enable: bool) -> None:
        check(lib().pcx_plan_set_one_pass(self._h, int(bool(enable))))

    def set_three_pass(self, enable: bool) -> None:
        """Natural order, 2^21 .. 2^24 (FP64: 2^21 .. 2^23): three passes (default) or the four-step."""
        check(lib().pcx_plan_set_three_pass(self._h, int(bool(enable))))

    # ---------------------------------------------------------------- raw pointer forms
    def fft_raw(self, dst: int, src: int | None = None, dst_pack: int = 1, src_pack: int | None = None,
                src_size: int | None = None, stream=None) -> None:
        """fft_raw<PD,PS>(dst[, src[, src_size]]) (fft.hpp:624-687) on device pointers."""
        src = dst if src is None else src
        src_pack = dst_pack if src_pack is None else src_pack
        src_size = self.size() if src_size is None else src_size
        check(self._fft(self._h, dst, dst_pack, src, src_pack, src_size, _stream_ptr(stream)))

    def ifft_raw(self, dst: int, src: int | None = None, normalized: bool = True, dst_pack: int = 1,
                 src_pack: int | None = None, stream=None) -> None:
        """ifft_raw<Normalized,PD,PS>(dst[, src]) (fft.hpp:689-721) on device pointers."""
        src = dst if src is None else src
        src_pack = dst_pack if src_pack is None else src_pack
        check(self._ifft(self._h, int(normalized), dst, dst_pack, src, src_pack, _stream_ptr(stream)))

    def batch(self, dst: int, src: int, batch: int, dst_pack: int, dst_dist: int, src_pack: int,
              src_dist: int, inverse: bool = False, normalized: bool = True, src_size: int | None = None,
              stream=None) -> None:
        src_size = self.size() if src_size is None else src_size
        check(self._batch(self._h, int(inverse), int(normalized), dst, dst_pack, dst_dist, src, src_pack,
                          src_dist, src_size, batch, _stream_ptr(stream)))

    # ---------------------------------------------------------------- vector forms
    def _check_type(self, *vs: vector) -> None:
        for v in vs:
            if v.dtype != self.dtype:
                raise PcxError(_lib.ERR_INVALID_ARGUMENT,
                               f"vector element type {v.dtype} does not match the unit's {self.dtype}")

    def _check_dst(self, v: vector, what: str) -> None:
        if v.size() != self.size():
            raise PcxError(_lib.ERR_SIZE_MISMATCH, f"{what} size (which is {v.size()}) is not equal to fft size "
                                                   f"(which is {self.size()})")

    def across(self, dst_vecs, src_vecs, inverse: bool = False, normalized: bool = True, stream=None) -> None:
        """FFT across a range of vectors (fft_unit_par, fft.hpp:2168-2572, in natural order):
        dst_vecs[j][k] = sum_i src_vecs[i][k] W^(i j) for every element k."""
        torch = _torch()
        n = self.size()
        if len(dst_vecs) != n or len(src_vecs) != n:
            raise PcxError(_lib.ERR_SIZE_MISMATCH,
                           f"vector count (which is {len(src_vecs)}) is not equal to fft size (which is {n})")
        length = src_vecs[0].size()
        if any(v.size() != length for v in list(src_vecs) + list(dst_vecs)):
            raise PcxError(_lib.ERR_SIZE_MISMATCH, "vectors differ in size")
        if len({v.pack for v in src_vecs}) != 1 or len({v.pack for v in dst_vecs}) != 1:
            raise PcxError(_lib.ERR_INVALID_PACK, "vectors of one range share a pack size")
        spk, dpk = src_vecs[0].pack, dst_vecs[0].pack
        if spk != dpk:
            # the kernel reads element k of every source while other workgroups write element k of
            # the destinations at other packed offsets: no destination may overlap a source
            # (fft.hpp:783 refuses an in-place repack)
            esz = 8 if self.dtype == "float64" else 4
            s0 = np.array([v.data_ptr() for v in src_vecs], np.int64)
            d0 = np.array([v.data_ptr() for v in dst_vecs], np.int64)
            s1 = s0 + esz * real_size(length, spk)
            d1 = d0 + esz * real_size(length, dpk)
            if np.any((d0[:, None] < s1[None, :]) & (s0[None, :] < d1[:, None])):
                raise PcxError(_lib.ERR_INPLACE_REPACK,
                               f"cannot perform in-place repack from source pack size (which is {spk}) to "
                               f"destination pack size (which is {dpk})")
        sp = torch.tensor([v.data_ptr() for v in src_vecs], dtype=torch.int64, device="cuda")
        dp = torch.tensor([v.data_ptr() for v in dst_vecs], dtype=torch.int64, device="cuda")
        fn = lib().pcx_fft_across_f64 if self.dtype == "float64" else lib().pcx_fft_across
        check(fn(self._h, int(inverse), int(normalized), dp.data_ptr(), dst_vecs[0].pack, sp.data_ptr(),
                 src_vecs[0].pack, length, _stream_ptr(stream)))
        self._across_tables = (sp, dp)  # kept alive until the next call (asynchronous launch)

    def _storage_call(self, dst: md_storage, src: md_storage | None, inverse: bool, normalized: bool,
                      stream) -> None:
        """Every slice of a batch storage in one launch (the rows of a two-axis md::dynamic_storage)."""
        src = dst if src is None else src
        self._check_type(dst, src)
        if dst.inner != self.size():
            raise PcxError(_lib.ERR_SIZE_MISMATCH, f"destination size (which is {dst.inner}) is not equal to fft "
                                                   f"size (which is {self.size()})")
        if src.inner > self.size() or (inverse and src.inner != self.size()):
            raise PcxError(_lib.ERR_SIZE_MISMATCH, f"source size (which is {src.inner}) does not fit fft size "
                                                   f"(which is {self.size()})")
        if dst.outer != src.outer:
            raise PcxError(_lib.ERR_SIZE_MISMATCH, f"destination slice count (which is {dst.outer}) is not equal to "
                                                   f"source slice count (which is {src.outer})")
        self.batch(dst.data_ptr(), src.data_ptr(), dst.outer, dst.pack, dst.row_stride, src.pack, src.row_stride,
                   inverse=inverse, normalized=normalized, src_size=src.inner, stream=stream)

    def __call__(self, dst: vector, src: vector | None = None, stream=None) -> None:
        """operator()(v) in place, or operator()(dst, src) with src zero-extended (fft.hpp:728-820).
        With md_storage arguments every slice is transformed in one launch."""
        if isinstance(dst, md_storage):
            self._storage_call(dst, src, False, True, stream)
            return
        self._check_type(dst, *([] if src is None else [src]))
        if src is None:
            if dst.size() != self.size():
                raise PcxError(_lib.ERR_SIZE_MISMATCH, f"input size (which is {dst.size()}) is not equal to fft "
                                                       f"size (which is {self.size()})")
            self.fft_raw(dst.data_ptr(), dst.data_ptr(), dst.pack, dst.pack, self.size(), stream)
            return
        self._check_dst(dst, "destination")
        if src.size() > self.size():
            raise PcxError(_lib.ERR_SIZE_MISMATCH, f"source size (which is {src.size()}) is bigger than fft size "
                                                   f"(which is {self.size()})")
        self.fft_raw(dst.data_ptr(), src.data_ptr(), dst.pack, src.pack, src.size(), stream)

    def ifft(self, dst: vector, src: vector | None = None, normalized: bool = True, stream=None) -> None:
        """ifft<Normalized>(v) / ifft<Normalized>(dst, src) (fft.hpp:828-915); md_storage: every slice."""
        if isinstance(dst, md_storage):
            self._storage_call(dst, src, True, normalized, stream)
            return
        self._check_type(dst, *([] if src is None else [src]))
        if src is None:
            if dst.size() != self.size():
                raise PcxError(_lib.ERR_SIZE_MISMATCH, f"input size (which is {dst.size()} is not equal to fft "
                                                       f"size (which is {self.size()})")
            src = dst
        else:
            self._check_dst(dst, "destination")
            if src.size() != self.size():
                raise PcxError(_lib.ERR_SIZE_MISMATCH, f"source size (which is {src.size()}) is not equal to fft "
                                                       f"size (which is {self.size()})")
        self.ifft_raw(dst.data_ptr(), src.data_ptr(), normalized, dst.pack, src.pack, stream)


class fxcorr_unit:
    """fxcorr.hpp:44-147 fxcorr_unit<float> on the GPU: out[i] = sum_m f[i+m] * conj(g[m])."""

    def __init__(self, g: vector, fft_size: int, stream=None, precision: str = "fp32"):
        import ctypes as C

        h = C.c_void_p()
        check(lib().pcx_fxcorr_create(g.data_ptr(), g.pack, g.size(), int(fft_size), _stream_ptr(stream),
                                      C.byref(h)))
        self._h = h
        if precision != "fp32":
            check(lib().pcx_fxcorr_set_precision(h, PRECISION[precision]))
        self.fft_size = int(fft_size)
        self.g_size = g.size()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.pcx_fxcorr_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def step(self, pack: int = 8) -> int:
        return int(lib().pcx_fxcorr_step(self._h, pack))

    def __call__(self, f: vector, out: vector | None = None, stream=None) -> None:
        """operator()(vector&) in place (fxcorr.hpp:102-136), or into `out` (the out-of-place overload).
        With md_storage arguments every slice is a channel: one launch, one shared kernel spectrum."""
        if isinstance(f, md_storage):
            dst = f if out is None else out
            if not isinstance(dst, md_storage) or dst.outer != f.outer or dst.inner < f.inner:
                raise PcxError(_lib.ERR_SIZE_MISMATCH, "destination storage must have the source's slice count "
                                                       "and at least its slice size")
            if f.dtype != "float32" or dst.dtype != "float32":
                raise PcxError(_lib.ERR_UNSUPPORTED, "cross-correlation is built for float vectors")
            check(lib().pcx_fxcorr_exec_batch(self._h, dst.data_ptr(), dst.pack, dst.row_stride, f.data_ptr(), f.pack,
                                              f.row_stride, f.inner, f.outer, _stream_ptr(stream)))
            return
        dst = f if out is None else out
        if dst.size() < f.size():
            raise PcxError(_lib.ERR_SIZE_MISMATCH, f"destination size (which is {dst.size()}) is smaller than "
                                                   f"the signal (which is {f.size()})")
        check(lib().pcx_fxcorr_exec(self._h, dst.data_ptr(), dst.pack, f.data_ptr(), f.pack, f.size(),
                                    _stream_ptr(stream)))

    def batch(self, dst: int, src: int, length: int, channels: int, pack: int, dist: int, stream=None,
              workspace: int | None = None, workspace_bytes: int = 0) -> None:
        """Multi-channel correlation on raw device pointers: channel c at c*dist floats.  With a
        workspace (device pointer, >= workspace_min() bytes) the unit's scratch comes from it (the
        reference's TmpFactory_ slot) and the unit allocates nothing of its own."""
        if workspace is None:
            check(lib().pcx_fxcorr_exec_batch(self._h, dst, pack, dist, src, pack, dist, length, channels,
                                              _stream_ptr(stream)))
        else:
            check(lib().pcx_fxcorr_exec_batch_ws(self._h, dst, pack, dist, src, pack, dist, length, channels,
                                                 workspace, int(workspace_bytes), _stream_ptr(stream)))

    def workspace_min(self) -> int:
        return int(lib().pcx_fxcorr_workspace_min(self._h))


class fft_queue:
    """Work queue of uneven power-of-two transforms (pcx_queue_*): jobs = [(log2n, inverse,
    src_offset, dst_offset)], each vector in split layout (re[N] then im[N]) at its float offset."""

    def __init__(self, jobs, normalized: bool = True):
        import ctypes as C

        arr = (_lib.Job * len(jobs))()
        for k, (lg, inv, so, do) in enumerate(jobs):
            arr[k].log2n, arr[k].inverse, arr[k].reserved = int(lg), int(inv), 0
            arr[k].src_offset, arr[k].dst_offset = int(so), int(do)
        h = C.c_void_p()
        check(lib().pcx_queue_create(arr, len(jobs), int(normalized), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.pcx_queue_destroy(h)
            self._h = None

    def exec(self, dst: int, src: int, stream=None) -> None:
        check(lib().pcx_queue_exec(self._h, dst, src, _stream_ptr(stream)))

    def autotune(self, dst: int, src: int, cache_path: str | None = None, stream=None) -> None:
        check(lib().pcx_queue_autotune(self._h, dst, src, _stream_ptr(stream),
                                       cache_path.encode() if cache_path else None))

    def capture(self, dst: int, src: int, stream) -> None:
        check(lib().pcx_queue_capture(self._h, dst, src, _stream_ptr(stream)))

    def launch(self, stream) -> None:
        check(lib().pcx_queue_launch(self._h, _stream_ptr(stream)))

    def work(self):
        import ctypes as C

        f, b = C.c_double(), C.c_double()
        check(lib().pcx_queue_work(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value
