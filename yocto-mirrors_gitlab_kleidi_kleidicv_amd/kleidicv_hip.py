"""Python host binding of libkleidicv_hip.so (the C ABI in include/kleidicv_hip).

Thin ctypes layer used by the tests, the benchmark and the serving loop.
Images are interleaved HWC uint8 arrays: numpy arrays (host memory, staged by
the library) or torch CUDA tensors (device memory, used in place).  Every
call goes to the HIP library; there is no CPU fallback, and loading fails
loudly when the library has not been built.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkleidicv_hip.so")

OK, NOT_IMPLEMENTED, NULL_POINTER, RANGE, ALLOCATION, ALIGNMENT = range(6)
ERROR_NAMES = ["OK", "NOT_IMPLEMENTED", "NULL_POINTER", "RANGE", "ALLOCATION",
               "ALIGNMENT", "CONTEXT_MISMATCH"]
BORDER = {"CONSTANT": 0, "REPLICATE": 1, "REFLECT": 2, "WRAP": 3,
          "REVERSE": 4, "TRANSPARENT": 5, "NONE": 6}
INTERP = {"NEAREST": 0, "LINEAR": 1}
NV12_TO_RGB, NV12_TO_BGR, NV12_TO_RGBA, NV12_TO_BGRA = 0x02, 0x22, 0x42, 0x62
NV21_TO_RGB, NV21_TO_BGR, NV21_TO_RGBA, NV21_TO_BGRA = 0x12, 0x32, 0x52, 0x72


class KleidiCVError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s returned %s" % (what, ERROR_NAMES[code] if code < len(ERROR_NAMES) else code))
        self.code = code


_lib = None


def lib():
    """The loaded C library; raises when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libkleidicv_hip.so is not built: run "
                               "python yocto-mirrors_gitlab_kleidi_kleidicv_amd/build.py")
        # PyTorch ships its own copy of the HIP runtime under the same soname
        # (libamdhip64.so.7).  Loading torch first makes this library bind to
        # that copy, so one runtime serves both and torch tensors, streams
        # and events are valid handles here.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = C.CDLL(LIB_PATH)
        _lib.kleidicv_hip_default_stream.restype = C.c_void_p
    return _lib


def data_function(name, *argtypes):
    """Callable behind a KLEIDICV_API_DECLARATION function-pointer variable."""
    ptr = C.c_void_p.in_dll(lib(), name)
    return C.CFUNCTYPE(C.c_int, *argtypes)(ptr.value)


# Argument types of the function-pointer data symbols (kleidicv.h:2715, 2782).
_P, _S = C.c_void_p, C.c_size_t
REMAP_S16_ARGTYPES = [_P, _S, _S, _S, _P, _S, _S, _S, _S, _P, _S, _P, _S, C.c_int, _P]
REMAP_F32_ARGTYPES = [_P, _S, _S, _S, _P, _S, _S, _S, _S, _P, _S, _P, _S, C.c_int, C.c_int, _P]
YUV_SP_ARGTYPES = [_P, _S, _P, _S, _P, _S, _S, _S, C.c_int]


# ---------------------------------------------------------------------------
# helpers
def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(x):
    if x is None:
        return C.c_void_p(0)
    if _is_torch(x):
        return C.c_void_p(x.data_ptr())
    return C.c_void_p(x.ctypes.data)


def _shape(x):
    s = tuple(x.shape)
    return s if len(s) == 3 else (s[0], s[1], 1)


def _stride(x):
    """Row stride in bytes of a (h, w[, c]) array whose rows may be padded."""
    if _is_torch(x):
        return x.stride(0) * x.element_size()
    return x.strides[0]


def _empty_like_shape(like, shape, dtype=np.uint8):
    if _is_torch(like):
        import torch
        return torch.empty(shape, dtype=torch.uint8, device=like.device)
    return np.empty(shape, dtype=dtype)


def _sz(*vals):
    return [C.c_size_t(int(v)) for v in vals]


def _check(code, what):
    if code != OK:
        raise KleidiCVError(code, what)


def _stream_arg(stream):
    if stream is None:
        return C.c_void_p(0)
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


# ---------------------------------------------------------------------------
# Drop-in API (host or device buffers, synchronous) and device API (stream)
def gaussian_blur(src, kernel_size, sigma=0.0, border="REVERSE", dst=None,
                  stream=None, device_api=False):
    h, w, c = _shape(src)
    dst = _empty_like_shape(src, tuple(src.shape)) if dst is None else dst
    args = [_ptr(src), *_sz(_stride(src)), _ptr(dst), *_sz(_stride(dst), w, h, c,
            kernel_size, kernel_size), C.c_float(sigma), C.c_float(sigma),
            C.c_int(BORDER[border])]
    if device_api:
        _check(lib().kleidicv_hip_gaussian_blur_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_gaussian_blur_u8")
    else:
        _check(lib().kleidicv_gaussian_blur_u8(*args), "kleidicv_gaussian_blur_u8")
    return dst


def separable_filter_2d(src, kernel_x, kernel_y, border="REPLICATE", dst=None,
                        stream=None, device_api=False):
    h, w, c = _shape(src)
    dst = _empty_like_shape(src, tuple(src.shape)) if dst is None else dst
    kx = np.ascontiguousarray(kernel_x, dtype=np.uint8)
    ky = np.ascontiguousarray(kernel_y, dtype=np.uint8)
    args = [_ptr(src), *_sz(_stride(src)), _ptr(dst), *_sz(_stride(dst), w, h, c),
            _ptr(kx), *_sz(len(kx)), _ptr(ky), *_sz(len(ky)), C.c_int(BORDER[border])]
    if device_api:
        _check(lib().kleidicv_hip_separable_filter_2d_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_separable_filter_2d_u8")
    else:
        _check(lib().kleidicv_separable_filter_2d_u8(*args),
               "kleidicv_separable_filter_2d_u8")
    return dst


def blur_and_downsample(src, border="REVERSE", dst=None, stream=None, device_api=False):
    """cv::pyrDown: binomial 5x5 at even rows and columns, ((w+1)/2, (h+1)/2)."""
    h, w, c = _shape(src)
    shape = ((h + 1) // 2, (w + 1) // 2) + ((c,) if len(src.shape) == 3 else ())
    dst = _empty_like_shape(src, shape) if dst is None else dst
    args = [_ptr(src), *_sz(_stride(src), w, h), _ptr(dst), *_sz(_stride(dst), c),
            C.c_int(BORDER[border])]
    if device_api:
        _check(lib().kleidicv_hip_blur_and_downsample_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_blur_and_downsample_u8")
    else:
        _check(lib().kleidicv_blur_and_downsample_u8(*args),
               "kleidicv_blur_and_downsample_u8")
    return dst


def median_blur(src, kernel_size, border="REPLICATE", dst=None, stream=None,
                device_api=False):
    h, w, c = _shape(src)
    dst = _empty_like_shape(src, tuple(src.shape)) if dst is None else dst
    args = [_ptr(src), *_sz(_stride(src)), _ptr(dst), *_sz(_stride(dst), w, h, c,
            kernel_size, kernel_size), C.c_int(BORDER[border])]
    if device_api:
        _check(lib().kleidicv_hip_median_blur_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_median_blur_u8")
    else:
        _check(lib().kleidicv_median_blur_u8(*args), "kleidicv_median_blur_u8")
    return dst


def resize_linear(src, dst_width, dst_height, dst=None, stream=None,
                  device_api=False):
    h, w, c = _shape(src)
    if dst is None:
        shape = (dst_height, dst_width, c) if len(src.shape) == 3 else (dst_height, dst_width)
        dst = _empty_like_shape(src, shape)
    args = [_ptr(src), *_sz(_stride(src), w, h), _ptr(dst),
            *_sz(_stride(dst), dst_width, dst_height, c)]
    if device_api:
        _check(lib().kleidicv_hip_resize_linear_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_resize_linear_u8")
    else:
        _check(lib().kleidicv_resize_linear_u8(*args), "kleidicv_resize_linear_u8")
    return dst


def remap_s16point5(src, mapxy, mapfrac, border="CONSTANT", border_value=(0, 0, 0, 0),
                    dst=None, stream=None, device_api=False):
    h, w, c = _shape(src)
    dh, dw = mapfrac.shape[0], mapfrac.shape[1]
    if dst is None:
        dst = _empty_like_shape(src, (dh, dw, c) if len(src.shape) == 3 else (dh, dw))
    bv = np.array(border_value, dtype=np.uint8)
    args = [_ptr(src), *_sz(_stride(src), w, h), _ptr(dst),
            *_sz(_stride(dst), dw, dh, c), _ptr(mapxy), *_sz(_stride(mapxy)),
            _ptr(mapfrac), *_sz(_stride(mapfrac)), C.c_int(BORDER[border]), _ptr(bv)]
    if device_api:
        _check(lib().kleidicv_hip_remap_s16point5_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_remap_s16point5_u8")
    else:
        fn = data_function("kleidicv_remap_s16point5_u8", *REMAP_S16_ARGTYPES)
        _check(fn(*args), "kleidicv_remap_s16point5_u8")
    return dst


def remap_f32(src, mapx, mapy, interpolation="LINEAR", border="REPLICATE",
              border_value=(0, 0, 0, 0), dst=None, stream=None, device_api=False):
    h, w, c = _shape(src)
    dh, dw = mapx.shape[0], mapx.shape[1]
    if dst is None:
        dst = _empty_like_shape(src, (dh, dw, c) if len(src.shape) == 3 else (dh, dw))
    bv = np.array(border_value, dtype=np.uint8)
    args = [_ptr(src), *_sz(_stride(src), w, h), _ptr(dst), *_sz(_stride(dst), dw, dh, c),
            _ptr(mapx), *_sz(_stride(mapx)), _ptr(mapy), *_sz(_stride(mapy)),
            C.c_int(INTERP[interpolation]), C.c_int(BORDER[border]), _ptr(bv)]
    if device_api:
        _check(lib().kleidicv_hip_remap_f32_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_remap_f32_u8")
    else:
        fn = data_function("kleidicv_remap_f32_u8", *REMAP_F32_ARGTYPES)
        _check(fn(*args), "kleidicv_remap_f32_u8")
    return dst


def warp_perspective(src, transform, dst_width, dst_height, interpolation="LINEAR",
                     border="REPLICATE", border_value=(0, 0, 0, 0), dst=None,
                     stream=None, device_api=False):
    h, w, c = _shape(src)
    if dst is None:
        dst = _empty_like_shape(src, (dst_height, dst_width, c) if len(src.shape) == 3
                                else (dst_height, dst_width))
    t = np.ascontiguousarray(transform, dtype=np.float32).reshape(9)
    bv = np.array(border_value, dtype=np.uint8)
    args = [_ptr(src), *_sz(_stride(src), w, h), _ptr(dst),
            *_sz(_stride(dst), dst_width, dst_height), _ptr(t), *_sz(c),
            C.c_int(INTERP[interpolation]), C.c_int(BORDER[border]), _ptr(bv)]
    if device_api:
        _check(lib().kleidicv_hip_warp_perspective_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_warp_perspective_u8")
    else:
        _check(lib().kleidicv_warp_perspective_u8(*args), "kleidicv_warp_perspective_u8")
    return dst


def yuv_sp_to_rgb(y, uv, color_format=NV12_TO_RGB, dst=None, stream=None,
                  device_api=False):
    h, w = y.shape[0], y.shape[1]
    oc = 4 if color_format & 0x40 else 3
    dst = _empty_like_shape(y, (h, w, oc)) if dst is None else dst
    args = [_ptr(y), *_sz(_stride(y)), _ptr(uv), *_sz(_stride(uv)), _ptr(dst),
            *_sz(_stride(dst), w, h), C.c_int(color_format)]
    if device_api:
        _check(lib().kleidicv_hip_yuv_semiplanar_to_rgb_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_yuv_semiplanar_to_rgb_u8")
    else:
        fn = data_function("kleidicv_yuv_semiplanar_to_rgb_u8", *YUV_SP_ARGTYPES)
        _check(fn(*args), "kleidicv_yuv_semiplanar_to_rgb_u8")
    return dst


def yuv_to_rgb(src, width, height, color_format, dst=None, stream=None, device_api=False):
    """One-buffer YUV (planar 4:2:0, 4:2:2, 4:4:4) to RGB/BGR/RGBA/BGRA.
    `src` holds the source rows (2-D, rows may be padded)."""
    oc = 4 if color_format & 0x40 else 3
    dst = _empty_like_shape(src, (height, width, oc)) if dst is None else dst
    args = [_ptr(src), *_sz(_stride(src)), _ptr(dst), *_sz(_stride(dst), width, height),
            C.c_int(color_format)]
    if device_api:
        _check(lib().kleidicv_hip_yuv_to_rgb_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_yuv_to_rgb_u8")
    else:
        _check(lib().kleidicv_yuv_to_rgb_u8(*args), "kleidicv_yuv_to_rgb_u8")
    return dst


def rgb_to_yuv(src, color_format, dst=None, stream=None, device_api=False):
    """RGB/BGR(A) (h, w, 3|4) to one-buffer YUV: planar 4:2:0 ((3h+1)/2 rows
    of w bytes), 4:2:2 (h rows of 2w) or 4:4:4 (h rows of 3w)."""
    h, w, _ = _shape(src)
    base = color_format & 0x0F
    shape = {0x01: (h, 3 * w), 0x04: (h, 2 * w)}.get(base, ((3 * h + 1) // 2, w))
    dst = _empty_like_shape(src, shape) if dst is None else dst
    args = [_ptr(src), *_sz(_stride(src)), _ptr(dst), *_sz(_stride(dst), w, h),
            C.c_int(color_format)]
    if device_api:
        _check(lib().kleidicv_hip_rgb_to_yuv_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_rgb_to_yuv_u8")
    else:
        _check(lib().kleidicv_rgb_to_yuv_u8(*args), "kleidicv_rgb_to_yuv_u8")
    return dst


def rgb_to_yuv_semiplanar(src, color_format, y=None, uv=None, stream=None, device_api=False):
    """RGB/BGR(A) (h, w, 3|4) to NV12/NV21: luma (h, w) and chroma
    ((h+1)/2, 2*((w+1)/2))."""
    h, w, _ = _shape(src)
    y = _empty_like_shape(src, (h, w)) if y is None else y
    uv = _empty_like_shape(src, ((h + 1) // 2, 2 * ((w + 1) // 2))) if uv is None else uv
    args = [_ptr(src), *_sz(_stride(src)), _ptr(y), *_sz(_stride(y)), _ptr(uv),
            *_sz(_stride(uv), w, h), C.c_int(color_format)]
    if device_api:
        _check(lib().kleidicv_hip_rgb_to_yuv_semiplanar_u8(*args, _stream_arg(stream)),
               "kleidicv_hip_rgb_to_yuv_semiplanar_u8")
    else:
        _check(lib().kleidicv_rgb_to_yuv_semiplanar_u8(*args),
               "kleidicv_rgb_to_yuv_semiplanar_u8")
    return y, uv


def fail_device_allocation(after):
    """Test hook: the library's device allocation that follows `after`
    successful ones fails with out-of-memory (once); negative disarms."""
    lib().kcv_test_fail_device_allocation(C.c_long(int(after)))


def release_staging():
    """Test hook: the next drop-in call allocates its device staging anew
    (the library otherwise keeps one staging buffer for those calls)."""
    lib().kcv_test_release_staging()


def synchronize(stream=None):
    _check(lib().kleidicv_hip_synchronize(_stream_arg(stream)), "kleidicv_hip_synchronize")


# ---------------------------------------------------------------------------
# Batched execution (kleidicv_hip_batch_*): device tensors only.
OP_GAUSSIAN_BLUR, OP_SEPARABLE_FILTER_2D, OP_MEDIAN_BLUR, OP_RESIZE_LINEAR = 1, 2, 3, 4
OP_REMAP_S16POINT5, OP_REMAP_F32, OP_WARP_PERSPECTIVE, OP_YUV_SEMIPLANAR_TO_RGB = 5, 6, 7, 8


class OpDesc(C.Structure):
    """kleidicv_hip_op_desc."""
    _fields_ = [
        ("op", C.c_int), ("stage", C.c_uint),
        ("src", C.c_void_p), ("src_stride", C.c_size_t), ("src_width", C.c_size_t),
        ("src_height", C.c_size_t),
        ("dst", C.c_void_p), ("dst_stride", C.c_size_t), ("dst_width", C.c_size_t),
        ("dst_height", C.c_size_t),
        ("channels", C.c_size_t), ("kernel_size", C.c_size_t), ("sigma", C.c_float),
        ("kernel_x", C.c_uint8 * 7), ("kernel_y", C.c_uint8 * 7),
        ("border_type", C.c_int), ("interpolation", C.c_int),
        ("border_value", C.c_uint8 * 4),
        ("aux0", C.c_void_p), ("aux0_stride", C.c_size_t),
        ("aux1", C.c_void_p), ("aux1_stride", C.c_size_t),
        ("transformation", C.c_float * 9), ("color_format", C.c_int),
    ]


class Tuning(C.Structure):
    """kleidicv_hip_batch_tuning."""
    _fields_ = [("blocks_per_cu", C.c_uint), ("resize_band_rows", C.c_uint),
                ("xcd_order", C.c_uint), ("use_graph", C.c_uint)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        return cls(**{k: int(d[k]) for k, _ in cls._fields_})


def _desc(op, src, dst, stage, **kw):
    d = OpDesc()
    d.op = op
    d.stage = stage
    sh, sw, sc = _shape(src)
    d.src, d.src_stride, d.src_width, d.src_height = _ptr(src).value, _stride(src), sw, sh
    if isinstance(dst, tuple):
        # an output shape only: the serving engine takes the buffer from its ring
        d.dst, d.dst_stride, d.dst_width, d.dst_height = None, 0, dst[1], dst[0]
    else:
        dh, dw, _ = _shape(dst)
        d.dst, d.dst_stride, d.dst_width, d.dst_height = _ptr(dst).value, _stride(dst), dw, dh
    d.channels = sc
    for k, v in kw.items():
        setattr(d, k, v)
    return d


def op_gaussian_blur(src, dst, kernel_size, sigma=0.0, border="REVERSE", stage=0):
    return _desc(OP_GAUSSIAN_BLUR, src, dst, stage, kernel_size=kernel_size, sigma=sigma,
                 border_type=BORDER[border])


def op_separable_filter_2d(src, dst, kernel_x, kernel_y, border="REPLICATE", stage=0):
    d = _desc(OP_SEPARABLE_FILTER_2D, src, dst, stage, kernel_size=len(kernel_x),
              border_type=BORDER[border])
    for i, (a, b) in enumerate(zip(kernel_x, kernel_y)):
        d.kernel_x[i], d.kernel_y[i] = int(a), int(b)
    return d


def op_median_blur(src, dst, kernel_size, border="REPLICATE", stage=0):
    return _desc(OP_MEDIAN_BLUR, src, dst, stage, kernel_size=kernel_size,
                 border_type=BORDER[border])


def op_resize_linear(src, dst, stage=0):
    return _desc(OP_RESIZE_LINEAR, src, dst, stage)


def _bv(d, border_value):
    for i, v in enumerate(tuple(border_value)[:4]):
        d.border_value[i] = int(v)


def op_remap_s16point5(src, dst, mapxy, mapfrac, border="CONSTANT", border_value=(0, 0, 0, 0),
                       stage=0):
    d = _desc(OP_REMAP_S16POINT5, src, dst, stage, border_type=BORDER[border],
              aux0=_ptr(mapxy).value, aux0_stride=_stride(mapxy),
              aux1=_ptr(mapfrac).value, aux1_stride=_stride(mapfrac))
    _bv(d, border_value)
    return d


def op_remap_f32(src, dst, mapx, mapy, interpolation="LINEAR", border="REPLICATE",
                 border_value=(0, 0, 0, 0), stage=0):
    d = _desc(OP_REMAP_F32, src, dst, stage, interpolation=INTERP[interpolation],
              border_type=BORDER[border], aux0=_ptr(mapx).value, aux0_stride=_stride(mapx),
              aux1=_ptr(mapy).value, aux1_stride=_stride(mapy))
    _bv(d, border_value)
    return d


def op_warp_perspective(src, dst, transform, interpolation="LINEAR", border="REPLICATE",
                        border_value=(0, 0, 0, 0), stage=0):
    d = _desc(OP_WARP_PERSPECTIVE, src, dst, stage, interpolation=INTERP[interpolation],
              border_type=BORDER[border])
    for i, v in enumerate(np.asarray(transform, dtype=np.float32).reshape(9)):
        d.transformation[i] = float(v)
    _bv(d, border_value)
    return d


def op_yuv_sp_to_rgb(y, uv, dst, color_format=NV12_TO_RGB, stage=0):
    d = _desc(OP_YUV_SEMIPLANAR_TO_RGB, y, dst, stage, aux0=_ptr(uv).value,
              aux0_stride=_stride(uv), color_format=color_format)
    d.channels = 1
    return d


class Batch:
    """A validated batch of ops over device tensors, run as one persistent
    kernel launch per (stage, kernel family, variant) segment.

    The tensors referenced by the descriptors must stay alive (and in place)
    for as long as the batch is used; `keep` holds references for that.
    """

    def __init__(self, descs, tuning=None, keep=()):
        L = lib()
        L.kleidicv_hip_batch_work_items.restype = C.c_size_t
        arr = (OpDesc * len(descs))(*descs)
        self._keep = tuple(keep)
        self._h = C.c_void_p(0)
        failed = C.c_size_t(0)
        t = C.byref(tuning) if tuning is not None else None
        code = L.kleidicv_hip_batch_create(arr, C.c_size_t(len(descs)), t, C.byref(self._h),
                                           C.byref(failed))
        if code != OK:
            err = KleidiCVError(code, "kleidicv_hip_batch_create (op %d)" % failed.value)
            err.failed_index = failed.value
            raise err

    def run(self, stream=None):
        _check(lib().kleidicv_hip_batch_run(self._h, _stream_arg(stream)),
               "kleidicv_hip_batch_run")

    def tune(self, stream=None, iterations=5):
        best = Tuning()
        ms = C.c_float(0)
        _check(lib().kleidicv_hip_batch_tune(self._h, _stream_arg(stream), C.c_uint(iterations),
                                             C.byref(best), C.byref(ms)),
               "kleidicv_hip_batch_tune")
        return best, ms.value

    @property
    def tuning(self):
        t = Tuning()
        _check(lib().kleidicv_hip_batch_get_tuning(self._h, C.byref(t)),
               "kleidicv_hip_batch_get_tuning")
        return t

    @tuning.setter
    def tuning(self, t):
        _check(lib().kleidicv_hip_batch_set_tuning(self._h, C.byref(t)),
               "kleidicv_hip_batch_set_tuning")

    @property
    def work_items(self):
        return int(lib().kleidicv_hip_batch_work_items(self._h))

    def close(self):
        if self._h:
            lib().kleidicv_hip_batch_destroy(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_ops(descs, stream=None):
    """Each descriptor through its single-image device entry point, in order
    on `stream` (kleidicv_hip_run_ops): the unbatched counterpart of a batch."""
    arr = (OpDesc * len(descs))(*descs)
    failed = C.c_size_t(0)
    code = lib().kleidicv_hip_run_ops(arr, C.c_size_t(len(descs)), _stream_arg(stream),
                                      C.byref(failed))
    if code != OK:
        err = KleidiCVError(code, "kleidicv_hip_run_ops (op %d)" % failed.value)
        err.failed_index = failed.value
        raise err


def batch_workspace_size(descs, tuning=None):
    """Device bytes kleidicv_hip_batch_launch needs for these descriptors."""
    arr = (OpDesc * len(descs))(*descs)
    n = C.c_size_t(0)
    failed = C.c_size_t(0)
    code = lib().kleidicv_hip_batch_workspace_size(
        arr, C.c_size_t(len(descs)), C.byref(tuning) if tuning is not None else None,
        C.byref(n), C.byref(failed))
    if code != OK:
        err = KleidiCVError(code, "kleidicv_hip_batch_workspace_size (op %d)" % failed.value)
        err.failed_index = failed.value
        raise err
    return n.value


def batch_launch(descs, workspace, stream=None, tuning=None):
    """One-shot batch into a caller-provided device workspace (a uint8 tensor).

    Returns 0 when launched; when the workspace is too small nothing is
    launched and the required byte count is returned instead.
    """
    arr = (OpDesc * len(descs))(*descs)
    failed = C.c_size_t(0)
    need = C.c_size_t(0)
    nbytes = workspace.numel() if workspace is not None else 0
    code = lib().kleidicv_hip_batch_launch(
        arr, C.c_size_t(len(descs)), C.byref(tuning) if tuning is not None else None,
        _ptr(workspace), C.c_size_t(nbytes), _stream_arg(stream), C.byref(need),
        C.byref(failed))
    if code == RANGE and need.value > nbytes:
        return need.value
    if code != OK:
        err = KleidiCVError(code, "kleidicv_hip_batch_launch (op %d)" % failed.value)
        err.failed_index = failed.value
        raise err
    return 0
