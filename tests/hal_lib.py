"""ctypes bindings of the OpenCV HAL entry points' C twins
(include/kleidicv_hip/kleidicv_hal.h), with OpenCV's encodings."""

import ctypes as C

import kleidicv_hip as K

OK, NOT_IMPLEMENTED, UNKNOWN = 0, 1, -1
CV_8U, CV_16U, CV_32F = 0, 2, 5
BORDER = {"CONSTANT": 0, "REPLICATE": 1, "REFLECT": 2, "WRAP": 3, "REFLECT_101": 4,
          "TRANSPARENT": 5, "ISOLATED": 16}
INTER_NEAREST, INTER_LINEAR, INTER_CUBIC, INTER_AREA = 0, 1, 2, 3

_P, _S, _I, _D, _B = C.c_void_p, C.c_size_t, C.c_int, C.c_double, C.c_bool
ARGTYPES = {
    "bgr_to_yuv422": [_P, _S, _P, _S, _I, _I, _I, _B, _I, _I],
    "bgr_to_yuv420_p": [_P, _S, _P, _S, _I, _I, _I, _B, _I],
    "bgr_to_yuv420_sp": [_P, _S, _P, _S, _P, _S, _I, _I, _I, _B, _I],
    "bgr_to_yuv": [_P, _S, _P, _S, _I, _I, _I, _I, _B, _B],
    "yuv422_to_bgr": [_P, _S, _P, _S, _I, _I, _I, _B, _I, _I],
    "yuv_to_bgr_sp": [_P, _S, _P, _S, _I, _I, _I, _B, _I],
    "yuv_to_bgr_sp_ex": [_P, _S, _P, _S, _P, _S, _I, _I, _I, _B, _I],
    "yuv_to_bgr_p": [_P, _S, _P, _S, _I, _I, _I, _B, _I],
    "yuv_to_bgr": [_P, _S, _P, _S, _I, _I, _I, _I, _B, _B],
    "separable_filter_2d": [_P, _S, _I, _P, _S, _I, _I, _I, _I, _I, _I, _I, _P, _I, _P, _I,
                            _I, _I, _I, _D, _I],
    "gaussian_blur_binomial": [_P, _S, _P, _S, _I, _I, _I, _I, _S, _S, _S, _S, _S, _I],
    "gaussian_blur": [_P, _S, _P, _S, _I, _I, _I, _I, _S, _S, _S, _S, _S, _S, _D, _D, _I],
    "resize": [_I, _P, _S, _I, _I, _P, _S, _I, _I, _D, _D, _I],
    "medianBlur": [_P, _S, _P, _S, _I, _I, _I, _I, _I],
    "pyrdown": [_P, _S, _I, _I, _P, _S, _I, _I, _I, _I, _I],
    "remap_s16": [_I, _P, _S, _I, _I, _P, _S, _I, _I, _P, _S, _P, _S, _I, _I, _P],
    "remap_f32": [_I, _P, _S, _I, _I, _P, _S, _I, _I, _P, _S, _P, _S, _I, _I, _P],
    "warp_perspective": [_I, _P, _S, _I, _I, _P, _S, _I, _I, _P, _I, _I, _P],
}

# Itanium-mangled names of the C++ entry points (namespace kleidicv::hal),
# as an OpenCV build patched for the reference adapter references them.
MANGLED = {
    "bgr_to_yuv422": "_ZN8kleidicv3hal13bgr_to_yuv422EPKhmPhmiiibii",
    "bgr_to_yuv420_p": "_ZN8kleidicv3hal15bgr_to_yuv420_pEPKhmPhmiiibi",
    "bgr_to_yuv420_sp": "_ZN8kleidicv3hal16bgr_to_yuv420_spEPKhmPhmS3_miiibi",
    "bgr_to_yuv": "_ZN8kleidicv3hal10bgr_to_yuvEPKhmPhmiiiibb",
    "yuv422_to_bgr": "_ZN8kleidicv3hal13yuv422_to_bgrEPKhmPhmiiibii",
    "yuv_to_bgr_sp": "_ZN8kleidicv3hal13yuv_to_bgr_spEPKhmPhmiiibi",
    "yuv_to_bgr_sp_ex": "_ZN8kleidicv3hal16yuv_to_bgr_sp_exEPKhmS2_mPhmiiibi",
    "yuv_to_bgr_p": "_ZN8kleidicv3hal12yuv_to_bgr_pEPKhmPhmiiibi",
    "yuv_to_bgr": "_ZN8kleidicv3hal10yuv_to_bgrEPKhmPhmiiiibb",
    "separable_filter_2d": "_ZN8kleidicv3hal19separable_filter_2dEPKhmiPhmiiiiiiiS2_iS2_iiiidi",
    "gaussian_blur_binomial": "_ZN8kleidicv3hal22gaussian_blur_binomialEPKhmPhmiiiimmmmmi",
    "gaussian_blur": "_ZN8kleidicv3hal13gaussian_blurEPKhmPhmiiiimmmmmmddi",
    "resize": "_ZN8kleidicv3hal6resizeEiPKhmiiPhmiiddi",
    "medianBlur": "_ZN8kleidicv3hal10medianBlurEPKhmPhmiiiii",
    "pyrdown": "_ZN8kleidicv3hal7pyrdownEPKhmiiPhmiiiii",
    "remap_s16": "_ZN8kleidicv3hal9remap_s16EiPKhmiiPhmiiPsmPtmiiPKd",
    "remap_f32": "_ZN8kleidicv3hal9remap_f32EiPKhmiiPhmiiPKfmS5_miiPKd",
    "warp_perspective": "_ZN8kleidicv3hal16warp_perspectiveEiPKhmiiPhmiiPKdiiS5_",
}


def fn(name):
    f = getattr(K.lib(), "kleidicv_hip_hal_" + name)
    f.argtypes = ARGTYPES[name]
    f.restype = C.c_int
    return f


def cv_type(depth, channels):
    return depth + ((channels - 1) << 3)
