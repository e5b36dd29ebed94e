"""ctypes binding of the CPU FP64 oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (imfit_amd) never
imports it.

`OracleModel` exposes the reference ModelObject's setup and evaluation methods
(core/model_object.h:38-347) under the same names, so a test can drive the oracle
and the GPU library through one call sequence.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        for name, args in [
            ("orc_set_threads", [vp, C.c_int]),
            ("orc_set_conv_mode", [vp, C.c_int]),
            ("orc_add_function", [vp, C.c_char_p, C.c_int]),
            ("orc_set_subsampling", [vp, C.c_int]),
            ("orc_get_nparams", [vp]),
            ("orc_add_psf", [vp, _dp, C.c_int, C.c_int, C.c_int]),
            ("orc_setup_model_image", [vp, C.c_int, C.c_int]),
            ("orc_add_image_data", [vp, _dp, C.c_int, C.c_int]),
            ("orc_add_image_characteristics", [vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double]),
            ("orc_add_error_vector", [vp, _dp, C.c_long, C.c_int]),
            ("orc_add_mask_vector", [vp, _dp, C.c_long, C.c_int]),
            ("orc_add_oversampled_psf", [vp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int]),
            ("orc_use_cash_statistic", [vp]),
            ("orc_use_poisson_mlr", [vp]),
            ("orc_use_model_errors", [vp]),
            ("orc_final_setup", [vp]),
            ("orc_create_model_image", [vp, _dp]),
            ("orc_get_model_image", [vp, _dp]),
            ("orc_get_expanded_model_image", [vp, _dp]),
            ("orc_get_weights", [vp, _dp]),
            ("orc_compute_deviates", [vp, _dp, _dp]),
            ("orc_fit_statistic", [vp, _dp, _dp]),
            ("orc_convolve", [_dp, C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int]),
            ("orc_downsample_and_replace", [_dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_int,
                                            C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.orc_get_nvalid.argtypes = [vp]
        L.orc_get_nvalid.restype = C.c_long
        L.orc_func_value.argtypes = [C.c_char_p, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_func_value.restype = C.c_double
        L.orc_conv_seconds.argtypes = [C.c_int]
        L.orc_conv_seconds.restype = C.c_double
        L.orc_calculate_bn.argtypes = [C.c_double]
        L.orc_calculate_bn.restype = C.c_double
        L.orc_bessel_k01.argtypes = [C.c_double, _dp, _dp]
        L.orc_bessel_k01.restype = None
        L.orc_enorm.argtypes = [C.c_long, _dp]
        L.orc_enorm.restype = C.c_double
        L.orc_expdisk3d_qags.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                         _dp, C.POINTER(C.c_int)]
        L.orc_expdisk3d_qags.restype = C.c_double
        L.orc_qags_test.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.POINTER(C.c_int)]
        L.orc_qags_test.restype = C.c_double
        L.orc_set_psf_interpolation.argtypes = [vp, C.c_int]
        L.orc_set_psf_interpolation.restype = C.c_int
        L.orc_psf_interp_value.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_psf_interp_value.restype = C.c_double
        L.orc_lanczos.argtypes = [C.c_double, C.c_int]
        L.orc_lanczos.restype = C.c_double
        L.orc_spline_knot_derivatives.argtypes = [_dp, C.c_int, _dp]
        L.orc_spline_knot_derivatives.restype = None
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own mp_enorm / DownsampleAndReplace / Calculate_bn, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libimfitref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R._Z8mp_enormiPKd.argtypes = [C.c_int, _dp]
        R._Z8mp_enormiPKd.restype = C.c_double
        R._Z12Calculate_bnd.argtypes = [C.c_double]
        R._Z12Calculate_bnd.restype = C.c_double
        R._Z20DownsampleAndReplacePKdiiiiPdiiiiiiii.argtypes = [
            _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
            C.c_int, C.c_int, C.c_int, C.c_int]
        R._Z20DownsampleAndReplacePKdiiiiPdiiiiiiii.restype = None
        _REF = R
    return _REF


class OracleModel:
    """CPU restatement of ModelObject, driven by the same call sequence as the GPU one."""

    def __init__(self, threads=1, conv_mode=0):
        self._L = lib()
        self.h = self._L.orc_create()
        self._L.orc_set_threads(self.h, threads)
        self._L.orc_set_conv_mode(self.h, conv_mode)
        self.ncols = self.nrows = 0

    def __del__(self):
        try:
            self._L.orc_destroy(self.h)
        except Exception:
            pass

    def _chk(self, st, what):
        if st < 0:
            raise RuntimeError("oracle %s failed (%d)" % (what, st))
        return st

    def AddFunction(self, name, new_set=False):
        return self._chk(self._L.orc_add_function(self.h, name.encode(), int(new_set)), "AddFunction")

    def SetSubsampling(self, on):
        self._L.orc_set_subsampling(self.h, int(on))

    def GetNParams(self):
        return self._L.orc_get_nparams(self.h)

    def AddPSFVector(self, psf, normalize=True):
        psf = np.ascontiguousarray(psf, dtype=np.float64)
        return self._chk(self._L.orc_add_psf(self.h, _ptr(psf), psf.shape[1], psf.shape[0], int(normalize)),
                         "AddPSFVector")

    def SetupModelImage(self, ncols, nrows):
        self.ncols, self.nrows = ncols, nrows
        return self._chk(self._L.orc_setup_model_image(self.h, ncols, nrows), "SetupModelImage")

    def AddImageDataVector(self, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        self.nrows, self.ncols = data.shape
        return self._chk(self._L.orc_add_image_data(self.h, _ptr(data), self.ncols, self.nrows), "AddImageData")

    def AddImageCharacteristics(self, gain, read_noise, exp_time=1.0, n_combined=1, original_sky=0.0):
        return self._L.orc_add_image_characteristics(self.h, gain, read_noise, exp_time, n_combined, original_sky)

    def AddErrorVector(self, err, input_type=100):
        err = np.ascontiguousarray(err, dtype=np.float64)
        return self._chk(self._L.orc_add_error_vector(self.h, _ptr(err), err.size, input_type), "AddErrorVector")

    def AddMaskVector(self, mask, input_type=10):
        mask = np.ascontiguousarray(mask, dtype=np.float64)
        return self._chk(self._L.orc_add_mask_vector(self.h, _ptr(mask), mask.size, input_type), "AddMaskVector")

    def AddOversampledPsf(self, psf, scale, x1, x2, y1, y2, normalize=True):
        psf = np.ascontiguousarray(psf, dtype=np.float64)
        return self._chk(self._L.orc_add_oversampled_psf(self.h, _ptr(psf), psf.shape[1], psf.shape[0], scale,
                                                         x1, x2, y1, y2, int(normalize)), "AddOversampledPsf")

    def SetupPsfInterpolation(self, method=1):
        """PointSource interpolation: 1 bicubic, 2 lanczos2, 3 lanczos3 (status as the
        reference's: -1 no PSF, -2 PSF smaller than 4 x 4)."""
        return self._L.orc_set_psf_interpolation(self.h, int(method))

    def UseCashStatistic(self):
        self._L.orc_use_cash_statistic(self.h)

    def UsePoissonMLR(self):
        self._L.orc_use_poisson_mlr(self.h)

    def UseModelErrors(self):
        self._L.orc_use_model_errors(self.h)

    def FinalSetupForFitting(self):
        return self._L.orc_final_setup(self.h)

    def GetNValidPixels(self):
        return self._L.orc_get_nvalid(self.h)

    def CreateModelImage(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._L.orc_create_model_image(self.h, _ptr(p))

    def GetModelImageVector(self):
        out = np.empty((self.nrows, self.ncols))
        self._L.orc_get_model_image(self.h, _ptr(out))
        return out

    def GetWeightImageVector(self):
        out = np.empty((self.nrows, self.ncols))
        self._L.orc_get_weights(self.h, _ptr(out))
        return out

    def ComputeDeviates(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        out = np.empty(self.nrows * self.ncols)
        self._L.orc_compute_deviates(self.h, _ptr(p), _ptr(out))
        return out

    def GetFitStatistic(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        out = np.zeros(1)
        self._L.orc_fit_statistic(self.h, _ptr(p), _ptr(out))
        return float(out[0])


def func_value(name, params, x0, y0, x, y, subsampling=False):
    p = np.ascontiguousarray(params, dtype=np.float64)
    return lib().orc_func_value(name.encode(), _ptr(p), x0, y0, x, y, int(subsampling))


def convolve(image, psf, normalize=True, mode=0):
    img = np.array(image, dtype=np.float64, order="C")
    psf = np.ascontiguousarray(psf, dtype=np.float64)
    lib().orc_convolve(_ptr(img), img.shape[1], img.shape[0], _ptr(psf), psf.shape[1], psf.shape[0],
                       int(normalize), mode)
    return img


def conv_seconds(reset=False):
    """Wall time the oracle spent in FFT convolutions since the last reset."""
    return lib().orc_conv_seconds(int(reset))


def enorm(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_enorm(x.size, _ptr(x))


def calculate_bn(n):
    return lib().orc_calculate_bn(n)


def bessel_k01(x):
    k0, k1 = np.zeros(1), np.zeros(1)
    lib().orc_bessel_k01(float(x), _ptr(k0), _ptr(k1))
    return float(k0[0]), float(k1[0])


def expdisk3d_qags(params6, x0, y0, x, y, epsrel=1e-6):
    p = np.ascontiguousarray(params6, dtype=np.float64)
    err = np.zeros(1)
    nev = C.c_int(0)
    v = lib().orc_expdisk3d_qags(_ptr(p), x0, y0, x, y, epsrel, _ptr(err), C.byref(nev))
    return v, float(err[0]), nev.value


def qags_test(which, a, b, epsrel=1e-6):
    err = np.zeros(1)
    nev = C.c_int(0)
    v = lib().orc_qags_test(which, a, b, epsrel, _ptr(err), C.byref(nev))
    return v, float(err[0]), nev.value


def psf_interp_value(method, psf, x, y):
    """PsfInterpolator value at offset (x, y) from the PSF centre (1 bicubic, 2 lanczos2, 3 lanczos3)."""
    p = np.ascontiguousarray(psf, dtype=np.float64)
    return lib().orc_psf_interp_value(int(method), _ptr(p), p.shape[1], p.shape[0], float(x), float(y))


def lanczos(x, n):
    return lib().orc_lanczos(float(x), int(n))


def spline_knot_derivatives(y):
    """Knot derivatives of the natural cubic spline through y at unit spacing."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    d = np.empty_like(y)
    lib().orc_spline_knot_derivatives(_ptr(y), y.size, _ptr(d))
    return d
