"""imfit_amd -- MI355X model-image library for imfit's ModelObject hot path.

`ModelObjectGPU` is a thin ctypes binding of the C ABI in include/imfit_gpu.h
(libimfit_gpu.so, built in-tree by imfit_amd/build.py).  Its method names are
the reference ModelObject's (core/model_object.h:38-347), so callers written
against ModelObject -- and the parity tests -- read the same.  There is no CPU
fallback: if the HIP library is missing this module raises on import.
"""
import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimfit_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "imfit_gpu.h")

if not os.path.exists(LIB_PATH):
    raise ImportError("imfit_amd: %s is missing; build it with `python imfit_amd/build.py`" % LIB_PATH)

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

WEIGHTS_ARE_SIGMAS, WEIGHTS_ARE_VARIANCES, WEIGHTS_ARE_WEIGHTS = 100, 110, 120
MASK_ZERO_IS_GOOD, MASK_ZERO_IS_BAD = 10, 20
PRECISION_FP64, PRECISION_FP32 = 0, 1
GRADIENT_ADJOINT, GRADIENT_TANGENT = 0, 1
INTERP_BICUBIC, INTERP_LANCZOS2, INTERP_LANCZOS3 = 1, 2, 3
INTERPOLATION = {"bicubic": INTERP_BICUBIC, "lanczos2": INTERP_LANCZOS2, "lanczos3": INTERP_LANCZOS3}

_SIGS = {
    "ifg_create": ([C.c_int], _vp),
    "ifg_destroy": ([_vp], None),
    "ifg_last_error": ([_vp], C.c_char_p),
    "ifg_set_precision": ([_vp, C.c_int], C.c_int),
    "ifg_add_function": ([_vp, C.c_char_p, C.c_int], C.c_int),
    "ifg_set_subsampling": ([_vp, C.c_int], C.c_int),
    "ifg_set_psf_interpolation": ([_vp, C.c_int], C.c_int),
    "ifg_get_nparams": ([_vp], C.c_int),
    "ifg_add_psf": ([_vp, _dp, C.c_int, C.c_int, C.c_int], C.c_int),
    "ifg_setup_model_image": ([_vp, C.c_int, C.c_int], C.c_int),
    "ifg_add_image_data": ([_vp, _dp, C.c_int, C.c_int], C.c_int),
    "ifg_add_image_characteristics": ([_vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double], C.c_int),
    "ifg_add_error_vector": ([_vp, _dp, C.c_long, C.c_int], C.c_int),
    "ifg_add_mask_vector": ([_vp, _dp, C.c_long, C.c_int], C.c_int),
    "ifg_add_oversampled_psf": ([_vp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
                                C.c_int),
    "ifg_use_cash_statistic": ([_vp], C.c_int),
    "ifg_use_poisson_mlr": ([_vp], C.c_int),
    "ifg_use_model_errors": ([_vp], C.c_int),
    "ifg_final_setup": ([_vp], C.c_int),
    "ifg_which_fit_statistic": ([_vp], C.c_int),
    "ifg_get_ndata": ([_vp], C.c_long),
    "ifg_get_nvalid": ([_vp], C.c_long),
    "ifg_create_model_image": ([_vp, _dp], C.c_int),
    "ifg_get_model_image": ([_vp, _dp], C.c_int),
    "ifg_compute_deviates": ([_vp, _dp, _dp], C.c_int),
    "ifg_fit_statistic": ([_vp, _dp, _dp], C.c_int),
    "ifg_fit_statistic_batch": ([_vp, _dp, C.c_int, _dp], C.c_int),
    "ifg_jacobian": ([_vp, _dp, C.c_int, C.POINTER(C.c_int), _dp, _dp], C.c_int),
    "ifg_normal_equations": ([_vp, _dp, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp], C.c_int),
    "ifg_fit_statistic_gradient": ([_vp, _dp, _dp, _dp], C.c_int),
    "ifg_set_gradient_method": ([_vp, C.c_int], C.c_int),
    "ifg_deviates_vjp": ([_vp, _dp, _dp, _dp, _dp], C.c_int),
    "ifg_model_image_vjp": ([_vp, _dp, _dp, _dp], C.c_int),
    "ifg_set_kernel_timing": ([_vp, C.c_int], C.c_int),
    "ifg_set_graph_mode": ([_vp, C.c_int], C.c_int),
    "ifg_get_kernel_timing": ([_vp, C.c_int, C.c_char_p, C.c_int, _dp, C.POINTER(C.c_long)], C.c_int),
}

class ParamInfo(C.Structure):
    """ifg_param_info: mpfit's fixed / limited / limits per parameter."""
    _fields_ = [("fixed", C.c_int), ("limited", C.c_int * 2), ("limits", C.c_double * 2)]


class LMConfig(C.Structure):
    _fields_ = [("ftol", C.c_double), ("xtol", C.c_double), ("gtol", C.c_double), ("maxiter", C.c_int),
                ("maxfev", C.c_int)]


class LMResult(C.Structure):
    _fields_ = [("bestnorm", C.c_double), ("orignorm", C.c_double), ("status", C.c_int), ("niter", C.c_int),
                ("nfev", C.c_int), ("nfree", C.c_int), ("npegged", C.c_int)]


class DEConfig(C.Structure):
    """ifg_de_config: imfit's Differential Evolution settings."""
    _fields_ = [("ftol", C.c_double), ("seed", C.c_ulong), ("max_generations", C.c_int),
                ("pop_per_param", C.c_int), ("mode", C.c_int)]


class DEResult(C.Structure):
    _fields_ = [("best", C.c_double), ("nfev", C.c_long), ("status", C.c_int), ("generations", C.c_int),
                ("population", C.c_int)]


DE_GENERATIONAL, DE_SEQUENTIAL = 0, 1
DE_CONVERGED, DE_MAXGEN = 1, 5
_SIGS["ifg_fit_de"] = ([_vp, _dp, C.POINTER(ParamInfo), C.POINTER(DEConfig), C.POINTER(DEResult)], C.c_int)
_SIGS["ifg_fit_lm"] = ([_vp, _dp, C.POINTER(ParamInfo), C.POINTER(LMConfig), C.c_int, C.POINTER(LMResult)],
                       C.c_int)
_SIGS["ifg_fit_lm_normal"] = ([_vp, _dp, C.POINTER(ParamInfo), C.POINTER(LMConfig), C.POINTER(LMResult)], C.c_int)
_SIGS["ifg_fit_lm_normal_fd"] = ([_vp, _dp, C.POINTER(ParamInfo), C.POINTER(LMConfig), C.POINTER(LMResult)],
                                 C.c_int)

_SIGS["ifg_batch_create"] = ([C.c_int], _vp)
_SIGS["ifg_batch_destroy"] = ([_vp], None)
_SIGS["ifg_batch_last_error"] = ([_vp], C.c_char_p)
_SIGS["ifg_batch_add"] = ([_vp, _vp], C.c_int)
_SIGS["ifg_batch_size"] = ([_vp], C.c_int)
_SIGS["ifg_batch_fit_statistic"] = ([_vp, _dp, _dp], C.c_int)
_SIGS["ifg_batch_normal_equations"] = ([_vp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, _dp], C.c_int)
_SIGS["ifg_batch_fit_lm"] = ([_vp, _dp, C.POINTER(ParamInfo), C.POINTER(LMConfig), C.POINTER(LMResult)], C.c_int)
_SIGS["ifg_batch_set_graph_mode"] = ([_vp, C.c_int], C.c_int)
_SIGS["ifg_batch_set_lm_resident"] = ([_vp, C.c_int], C.c_int)
_SIGS["ifg_autotune"] = ([_vp, _dp, C.c_int], C.c_int)
_SIGS["ifg_get_launch_shape"] = ([_vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int)
_SIGS["ifg_init_genrand"] = ([C.c_ulong], None)
_SIGS["ifg_use_bootstrap"] = ([_vp], C.c_int)
_SIGS["ifg_make_bootstrap_sample"] = ([_vp], C.c_int)
_SIGS["ifg_set_bootstrap_indices"] = ([_vp, C.POINTER(C.c_long), C.c_long], C.c_int)
_SIGS["ifg_get_bootstrap_indices"] = ([_vp, C.POINTER(C.c_long)], C.c_int)
_SIGS["ifg_stop_bootstrap"] = ([_vp], C.c_int)
_SIGS["ifg_set_launch_shape"] = ([_vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int)
_SIGS["ifg_batch_autotune"] = ([_vp, _dp, C.POINTER(C.c_int), _dp], C.c_int)
_SIGS["ifg_fit_statistic_noalloc"] = ([_vp, _dp, _dp], C.c_int)

_LIB = C.CDLL(LIB_PATH)
for _name, (_args, _res) in _SIGS.items():
    _f = getattr(_LIB, _name)
    _f.argtypes = _args
    _f.restype = _res


def lib():
    return _LIB


def header_symbols():
    """Names of the entry points declared in include/imfit_gpu.h."""
    with open(HEADER) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(ifg_[a-z0-9_]+)\s*\(", txt)))


def _ptr(a):
    return a.ctypes.data_as(_dp)


# The per-evaluation entry points once more, through a second handle whose pointer
# arguments are plain addresses: building ctypes pointer objects costs several
# microseconds a call, a few percent of a whole evaluation.
_FAST = C.CDLL(LIB_PATH)
_FAST.ifg_fit_statistic.argtypes = [_vp, _vp, _vp]
_FAST.ifg_fit_statistic.restype = C.c_int


def _f64(a):
    """a as a C-contiguous float64 array (no copy when it already is one)."""
    if type(a) is np.ndarray and a.dtype == np.float64 and a.flags.c_contiguous:
        return a
    return np.ascontiguousarray(a, dtype=np.float64)


def _addr(a):
    return a.__array_interface__["data"][0]


class IfgError(RuntimeError):
    pass


def init_genrand(seed):
    """Seed the process-wide MT19937 that bootstrap samples are drawn from (imfit --seed)."""
    _LIB.ifg_init_genrand(int(seed))


_DEVFN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, _dp, C.c_long, _dp)
_LIB.ifg_mpfit.argtypes = [C.c_long, C.c_int, _dp, C.c_void_p, C.c_void_p, _DEVFN, C.c_void_p, C.c_void_p]
_LIB.ifg_mpfit.restype = C.c_int


def mpfit(deviates, params, m, limits=None, fixed=None, ftol=1e-8, xtol=0.0, maxiter=1000):
    """mpfit (ifg_mpfit: MINPACK LMDIF as imfit runs it, forward-difference Jacobian) on
    `deviates(p) -> m values` from the start `params`, within limits [(lo, hi) or None]
    and fixed flags.  Returns (params, LMResult).  Exceptions raised by `deviates` end
    the fit and are re-raised."""
    x = np.array(params, dtype=np.float64)
    n = x.size
    errors = []

    def cb(ctx, npar, p, mm, out):
        try:
            v = np.asarray(deviates(np.ctypeslib.as_array(p, shape=(npar,)).copy()), dtype=np.float64)
            if v.size != mm:
                raise ValueError("deviates returned %d values, expected %d" % (v.size, mm))
            np.ctypeslib.as_array(out, shape=(mm,))[:] = v
            return 0
        except Exception as e:   # noqa: BLE001 -- handed back to the caller below
            errors.append(e)
            return -1

    fn = _DEVFN(cb)
    pinfo = ModelObjectGPU._pinfo(limits, fixed, n)
    cfg = LMConfig(ftol, xtol, 0.0, maxiter, 0)
    res = LMResult()
    st = _LIB.ifg_mpfit(int(m), n, _ptr(x), C.cast(pinfo, C.c_void_p) if pinfo is not None else None,
                        C.cast(C.pointer(cfg), C.c_void_p), fn, None, C.cast(C.pointer(res), C.c_void_p))
    if errors:
        raise errors[0]
    if st < 0 and st > -16:
        raise IfgError("mpfit failed (%d)" % st)
    return x, res


class ModelObjectGPU:
    """ModelObject on one MI355X: setup and evaluation through the C ABI."""

    def __init__(self, device=0, precision=PRECISION_FP64):
        self.h = _LIB.ifg_create(device)
        if not self.h:
            raise IfgError("ifg_create failed")
        self._chk(_LIB.ifg_set_precision(self.h, precision), "SetPrecision")
        self.ncols = self.nrows = 0
        self._nparams = 0          # GetNParams(), kept current by AddFunction
        self._stat = np.zeros(1)   # GetFitStatistic's result slot
        self._stat_addr = _addr(self._stat)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            _LIB.ifg_destroy(h)
            self.h = None

    def _chk(self, st, what):
        if st < 0:
            raise IfgError("%s failed (%d): %s" % (what, st, _LIB.ifg_last_error(self.h).decode()))
        return st

    # ---- setup (same call order as core/setup_model_object.cpp:58-153)
    def AddFunction(self, name, new_set=False):
        st = self._chk(_LIB.ifg_add_function(self.h, name.encode(), int(new_set)), "AddFunction")
        self._nparams = _LIB.ifg_get_nparams(self.h)
        return st

    def SetSubsampling(self, on):
        return self._chk(_LIB.ifg_set_subsampling(self.h, int(on)), "SetSubsampling")

    def SetupPsfInterpolation(self, method=INTERP_BICUBIC):
        """Point-source PSF interpolation: INTERP_BICUBIC, INTERP_LANCZOS2 or INTERP_LANCZOS3
        (ModelObject::SetupPsfInterpolation, core/model_object.cpp:343-380)."""
        return self._chk(_LIB.ifg_set_psf_interpolation(self.h, int(method)), "SetupPsfInterpolation")

    def GetNParams(self):
        return _LIB.ifg_get_nparams(self.h)

    def _vec(self, params):
        """params as a C-contiguous float64 vector of GetNParams() values: the library
        reads exactly that many doubles, so a shorter vector is refused here."""
        p = _f64(params)
        if p.ndim != 1 or p.shape[0] != self._nparams:
            raise ValueError("expected %d parameters, got shape %s" % (self._nparams, p.shape))
        return p

    def AddPSFVector(self, psf, normalize=True):
        psf = np.ascontiguousarray(psf, dtype=np.float64)
        return self._chk(_LIB.ifg_add_psf(self.h, _ptr(psf), psf.shape[1], psf.shape[0], int(normalize)),
                         "AddPSFVector")

    def SetupModelImage(self, ncols, nrows):
        self.ncols, self.nrows = ncols, nrows
        return self._chk(_LIB.ifg_setup_model_image(self.h, ncols, nrows), "SetupModelImage")

    def AddImageDataVector(self, data):
        data = np.ascontiguousarray(data, dtype=np.float64)
        self.nrows, self.ncols = data.shape
        return self._chk(_LIB.ifg_add_image_data(self.h, _ptr(data), self.ncols, self.nrows), "AddImageDataVector")

    def AddImageCharacteristics(self, gain, read_noise, exp_time=1.0, n_combined=1, original_sky=0.0):
        return self._chk(_LIB.ifg_add_image_characteristics(self.h, gain, read_noise, exp_time, n_combined,
                                                            original_sky), "AddImageCharacteristics")

    def AddErrorVector(self, err, input_type=WEIGHTS_ARE_SIGMAS):
        err = np.ascontiguousarray(err, dtype=np.float64)
        return self._chk(_LIB.ifg_add_error_vector(self.h, _ptr(err), err.size, input_type), "AddErrorVector")

    def AddMaskVector(self, mask, input_type=MASK_ZERO_IS_GOOD):
        mask = np.ascontiguousarray(mask, dtype=np.float64)
        return self._chk(_LIB.ifg_add_mask_vector(self.h, _ptr(mask), mask.size, input_type), "AddMaskVector")

    def AddOversampledPsf(self, psf, scale, x1, x2, y1, y2, normalize=True):
        psf = np.ascontiguousarray(psf, dtype=np.float64)
        return self._chk(_LIB.ifg_add_oversampled_psf(self.h, _ptr(psf), psf.shape[1], psf.shape[0], scale, x1, x2,
                                                      y1, y2, int(normalize)), "AddOversampledPsf")

    def UseCashStatistic(self):
        return self._chk(_LIB.ifg_use_cash_statistic(self.h), "UseCashStatistic")

    def UsePoissonMLR(self):
        return self._chk(_LIB.ifg_use_poisson_mlr(self.h), "UsePoissonMLR")

    def UseModelErrors(self):
        return self._chk(_LIB.ifg_use_model_errors(self.h), "UseModelErrors")

    def FinalSetupForFitting(self):
        return self._chk(_LIB.ifg_final_setup(self.h), "FinalSetupForFitting")

    def WhichFitStatistic(self):
        return _LIB.ifg_which_fit_statistic(self.h)

    def GetNDataValues(self):
        return _LIB.ifg_get_ndata(self.h)

    def GetNValidPixels(self):
        return _LIB.ifg_get_nvalid(self.h)

    # ---- bootstrap resampling (core/model_object.cpp:2036-2087)
    def UseBootstrap(self):
        return self._chk(_LIB.ifg_use_bootstrap(self.h), "UseBootstrap")

    def MakeBootstrapSample(self):
        return self._chk(_LIB.ifg_make_bootstrap_sample(self.h), "MakeBootstrapSample")

    def SetBootstrapIndices(self, indices):
        idx = np.ascontiguousarray(indices, dtype=np.int64)
        return self._chk(_LIB.ifg_set_bootstrap_indices(self.h, idx.ctypes.data_as(C.POINTER(C.c_long)), idx.size),
                         "SetBootstrapIndices")

    def GetBootstrapIndices(self):
        out = np.empty(self.GetNValidPixels(), dtype=np.int64)
        self._chk(_LIB.ifg_get_bootstrap_indices(self.h, out.ctypes.data_as(C.POINTER(C.c_long))),
                  "GetBootstrapIndices")
        return out

    def StopBootstrap(self):
        return self._chk(_LIB.ifg_stop_bootstrap(self.h), "StopBootstrap")

    # ---- hot path
    def CreateModelImage(self, params):
        p = self._vec(params)
        self._chk(_LIB.ifg_create_model_image(self.h, _ptr(p)), "CreateModelImage")

    def GetModelImageVector(self):
        out = np.empty((self.nrows, self.ncols))
        self._chk(_LIB.ifg_get_model_image(self.h, _ptr(out)), "GetModelImageVector")
        return out

    def ComputeDeviates(self, params):
        p = self._vec(params)
        out = np.empty(self.nrows * self.ncols)
        self._chk(_LIB.ifg_compute_deviates(self.h, _ptr(p), _ptr(out)), "ComputeDeviates")
        return out

    def GetFitStatistic(self, params):
        p = _f64(params)
        if p.ndim != 1 or p.shape[0] != self._nparams:
            self._vec(p)   # raises
        st = _FAST.ifg_fit_statistic(self.h, _addr(p), self._stat_addr)
        if st < 0:
            self._chk(st, "GetFitStatistic")
        return float(self._stat[0])

    def GetFitStatisticNoAlloc(self, params):
        """GetFitStatistic through ifg_fit_statistic_noalloc: no device or pinned
        allocation in the call (per-call memory from the library's fixed per-GPU work
        area).  The handle needs one ordinary evaluation first."""
        p = self._vec(params)
        out = np.zeros(1)
        self._chk(_LIB.ifg_fit_statistic_noalloc(self.h, _ptr(p), _ptr(out)), "GetFitStatisticNoAlloc")
        return float(out[0])

    def GetFitStatisticBatch(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        if p.ndim != 2 or p.shape[1] != self._nparams:
            raise ValueError("params must be (n_vectors, %d)" % self._nparams)
        out = np.zeros(p.shape[0])
        self._chk(_LIB.ifg_fit_statistic_batch(self.h, _ptr(p), p.shape[0], _ptr(out)), "GetFitStatisticBatch")
        return out

    # ---- derivatives
    def _ifree(self, ifree):
        if ifree is None:
            ifree = np.arange(self.GetNParams())
        return np.ascontiguousarray(ifree, dtype=np.int32)

    def Jacobian(self, params, ifree=None):
        """(deviates, J) with J[:, j] = d(deviate)/d(param ifree[j]) (nData x nFree)."""
        p = self._vec(params)
        fr = self._ifree(ifree)
        n = self.nrows * self.ncols
        fvec = np.empty(n)
        fjac = np.empty((fr.size, n))
        self._chk(_LIB.ifg_jacobian(self.h, _ptr(p), fr.size, fr.ctypes.data_as(C.POINTER(C.c_int)), _ptr(fvec),
                                    _ptr(fjac)), "Jacobian")
        return fvec, fjac.T

    def NormalEquations(self, params, ifree=None):
        """(J^T J, J^T r, statistic) reduced on the GPU."""
        p = self._vec(params)
        fr = self._ifree(ifree)
        jtj = np.empty((fr.size, fr.size))
        jtr = np.empty(fr.size)
        stat = np.zeros(1)
        self._chk(_LIB.ifg_normal_equations(self.h, _ptr(p), fr.size, fr.ctypes.data_as(C.POINTER(C.c_int)),
                                            _ptr(jtj), _ptr(jtr), _ptr(stat)), "NormalEquations")
        return jtj, jtr, float(stat[0])

    def GetFitStatisticGradient(self, params):
        """(gradient of GetFitStatistic w.r.t. every parameter, statistic)."""
        p = self._vec(params)
        g = np.empty(self.GetNParams())
        stat = np.zeros(1)
        self._chk(_LIB.ifg_fit_statistic_gradient(self.h, _ptr(p), _ptr(g), _ptr(stat)), "GetFitStatisticGradient")
        return g, float(stat[0])

    def SetGradientMethod(self, method):
        """GRADIENT_ADJOINT (reverse mode, the default) or GRADIENT_TANGENT (one
        derivative plane per parameter) for GetFitStatisticGradient."""
        self._chk(_LIB.ifg_set_gradient_method(self.h, int(method)), "SetGradientMethod")

    def _outputs(self, u):
        w = np.ascontiguousarray(u, dtype=np.float64).ravel()
        if w.size != self.nrows * self.ncols:
            raise ValueError("expected %d output weights, got %d" % (self.nrows * self.ncols, w.size))
        return w

    def DeviatesVJP(self, params, u):
        """(J^T u for the deviates' Jacobian J, deviates) by the adjoint pass."""
        p = self._vec(params)
        w = self._outputs(u)
        g = np.empty(self.GetNParams())
        fvec = np.empty(self.nrows * self.ncols)
        self._chk(_LIB.ifg_deviates_vjp(self.h, _ptr(p), _ptr(w), _ptr(g), _ptr(fvec)), "DeviatesVJP")
        return g, fvec

    def ModelImageVJP(self, params, u):
        """sum_i u_i d(model_i)/d(params) over the nData model pixels (adjoint pass)."""
        p = self._vec(params)
        w = self._outputs(u)
        g = np.empty(self.GetNParams())
        self._chk(_LIB.ifg_model_image_vjp(self.h, _ptr(p), _ptr(w), _ptr(g)), "ModelImageVJP")
        return g

    # ---- Levenberg-Marquardt fits
    @staticmethod
    def _pinfo(limits, fixed, n):
        if limits is None and fixed is None:
            return None
        arr = (ParamInfo * n)()
        for i in range(n):
            if fixed is not None and fixed[i]:
                arr[i].fixed = 1
            if limits is not None and limits[i] is not None:
                lo, hi = limits[i]
                if lo is not None:
                    arr[i].limited[0], arr[i].limits[0] = 1, lo
                if hi is not None:
                    arr[i].limited[1], arr[i].limits[1] = 1, hi
        return arr

    def FitLM(self, params, limits=None, fixed=None, ftol=1e-8, xtol=0.0, maxiter=1000, jacobian=None,
              method="mpfit"):
        """Levenberg-Marquardt fit.  method "mpfit": MINPACK LMDIF as imfit runs it, with
        jacobian "fd" (default; one GPU evaluation per free parameter) or "analytic";
        method "normal": LM on GPU-reduced normal equations, with jacobian "analytic"
        (default) or "fd" (forward differences from one batched launch).  Returns
        (params, LMResult)."""
        if jacobian is None:
            jacobian = "analytic" if method == "normal" else "fd"
        x = np.array(self._vec(params), dtype=np.float64)
        pinfo = self._pinfo(limits, fixed, x.size)
        cfg = LMConfig(ftol, xtol, 0.0, maxiter, 0)
        res = LMResult()
        if method == "normal" and jacobian == "fd":
            st = _LIB.ifg_fit_lm_normal_fd(self.h, _ptr(x), pinfo, C.byref(cfg), C.byref(res))
        elif method == "normal":
            st = _LIB.ifg_fit_lm_normal(self.h, _ptr(x), pinfo, C.byref(cfg), C.byref(res))
        else:
            mode = 1 if jacobian == "analytic" else 0
            st = _LIB.ifg_fit_lm(self.h, _ptr(x), pinfo, C.byref(cfg), mode, C.byref(res))
        if st < 0 and st > -16:
            self._chk(st, "FitLM")
        return x, res

    def FitDE(self, params, limits, fixed=None, ftol=1e-8, seed=1, max_generations=600,
              mode=DE_GENERATIONAL, pop_per_param=8):
        """imfit's Differential Evolution (DiffEvolnFit): every free parameter needs both
        limits.  mode DE_GENERATIONAL evaluates each generation's trials in one batched
        launch; DE_SEQUENTIAL is DESolver's own loop (one evaluation per trial).
        Returns (best params, DEResult)."""
        x = np.array(self._vec(params), dtype=np.float64)
        pinfo = self._pinfo(limits, fixed, x.size)
        cfg = DEConfig(ftol, seed, max_generations, pop_per_param, mode)
        res = DEResult()
        st = _LIB.ifg_fit_de(self.h, _ptr(x), pinfo, C.byref(cfg), C.byref(res))
        if st < 0:
            self._chk(st, "FitDE")
        return x, res

    # ---- instrumentation
    def SetKernelTiming(self, on):
        return self._chk(_LIB.ifg_set_kernel_timing(self.h, int(on)), "SetKernelTiming")

    def SetGraphMode(self, on):
        return self._chk(_LIB.ifg_set_graph_mode(self.h, int(on)), "SetGraphMode")

    # ---- launch-shape autotuning
    def Autotune(self, params):
        """Time GetFitStatistic (one vector) or GetFitStatisticBatch (rows of a 2-D array)
        at params under nearby launch shapes and keep the fastest; returns the number of
        timed trials (0: the shapes came from this process's record for the geometry)."""
        p = np.ascontiguousarray(np.atleast_2d(np.asarray(params, dtype=np.float64)))
        if p.shape[1] != self.GetNParams():
            raise ValueError("params must have GetNParams() columns")
        return self._chk(_LIB.ifg_autotune(self.h, _ptr(p), p.shape[0]), "Autotune")

    def SetLaunchShape(self, grid, cols, rows, population=False):
        """Apply a launch shape (as GetLaunchShape reports it) to grid 0 (main) or a region."""
        return self._chk(_LIB.ifg_set_launch_shape(self.h, grid, int(population), cols, rows), "SetLaunchShape")

    def GetLaunchShape(self, grid=0, population=False):
        """(columns per workgroup of the column pass, rows per workgroup of the row passes)
        of grid 0 (main image) or 1.. (oversampled regions)."""
        c, r = C.c_int(0), C.c_int(0)
        self._chk(_LIB.ifg_get_launch_shape(self.h, grid, int(population), C.byref(c), C.byref(r)), "GetLaunchShape")
        return c.value, r.value

    def GetKernelTiming(self):
        """{kernel name: (total ms, launches)} recorded with HIP events on the handle's stream."""
        out = {}
        i = 0
        while True:
            name = C.create_string_buffer(64)
            ms = np.zeros(1)
            n = C.c_long(0)
            cnt = _LIB.ifg_get_kernel_timing(self.h, i, name, 64, _ptr(ms), C.byref(n))
            if cnt <= i:
                break
            out[name.value.decode()] = (float(ms[0]), int(n.value))
            i += 1
        return out


class BatchGPU:
    """Many set-up ModelObjectGPU models (e.g. cutouts of different sizes) evaluated
    together in merged launches (the ifg_batch_* entry points).  Parameters are given
    per model (a list of vectors); the models must share precision, statistic and
    sub-pixel switch, and are kept alive by the batch."""

    def __init__(self, models, device=0):
        self.h = _LIB.ifg_batch_create(device)
        if not self.h:
            raise IfgError("ifg_batch_create failed")
        self.models = list(models)
        for m in self.models:
            self._chk(_LIB.ifg_batch_add(self.h, m.h), "ifg_batch_add")
        self.nparams = [m.GetNParams() for m in self.models]

    def __del__(self):
        if getattr(self, "h", None):
            _LIB.ifg_batch_destroy(self.h)
            self.h = None

    def _chk(self, st, what):
        if st < 0:
            raise IfgError("%s failed (%d): %s" % (what, st, _LIB.ifg_batch_last_error(self.h).decode()))
        return st

    def _concat(self, params):
        if len(params) != len(self.models):
            raise ValueError("one parameter vector per model")
        for p, n in zip(params, self.nparams):
            if len(p) != n:
                raise ValueError("parameter vector of the wrong length")
        return np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64) for p in params]))

    def _split(self, x):
        out, o = [], 0
        for n in self.nparams:
            out.append(x[o:o + n].copy())
            o += n
        return out

    def GetFitStatistic(self, params):
        x = self._concat(params)
        out = np.zeros(len(self.models))
        self._chk(_LIB.ifg_batch_fit_statistic(self.h, _ptr(x), _ptr(out)), "GetFitStatistic")
        return out

    def NormalEquations(self, params, ifree=None):
        """Per model (J^T J as a full matrix, J^T r); and the statistics."""
        x = self._concat(params)
        if ifree is None:
            ifree = [np.arange(n) for n in self.nparams]
        nf = np.array([len(f) for f in ifree], dtype=np.int32)
        fr = np.ascontiguousarray(np.concatenate([np.asarray(f) for f in ifree]).astype(np.int32))
        total = int(sum(k * (k + 1) // 2 + k for k in nf))
        normal = np.zeros(total)
        stat = np.zeros(len(self.models))
        self._chk(_LIB.ifg_batch_normal_equations(self.h, _ptr(x), nf.ctypes.data_as(C.POINTER(C.c_int)),
                                                  fr.ctypes.data_as(C.POINTER(C.c_int)), _ptr(normal), _ptr(stat)),
                  "NormalEquations")
        out, o = [], 0
        for k in nf:
            jtj = np.zeros((k, k))
            iu = np.triu_indices(k)
            jtj[iu] = normal[o:o + k * (k + 1) // 2]
            jtj = jtj + np.triu(jtj, 1).T
            o += k * (k + 1) // 2
            out.append((jtj, normal[o:o + k].copy()))
            o += k
        return out, stat

    def FitLM(self, params, limits=None, fixed=None, ftol=1e-8, xtol=0.0, maxiter=1000):
        """The normal-equation LM of every model at once.  limits / fixed: per model
        (lists as for ModelObjectGPU.FitLM) or None.  Returns (params per model,
        LMResult per model)."""
        x = self._concat(params)
        pinfo = None
        if limits is not None or fixed is not None:
            pinfo = (ParamInfo * x.size)()
            o = 0
            for k, n in enumerate(self.nparams):
                part = ModelObjectGPU._pinfo(None if limits is None else limits[k],
                                             None if fixed is None else fixed[k], n)
                if part is not None:
                    for i in range(n):
                        pinfo[o + i] = part[i]
                o += n
        cfg = LMConfig(ftol, xtol, 0.0, maxiter, 0)
        res = (LMResult * len(self.models))()
        self._chk(_LIB.ifg_batch_fit_lm(self.h, _ptr(x), pinfo, C.byref(cfg), res), "FitLM")
        return self._split(x), list(res)

    def SetGraphMode(self, on):
        return self._chk(_LIB.ifg_batch_set_graph_mode(self.h, int(on)), "SetGraphMode")

    def SetLMResident(self, on):
        """FitLM's driver: False (default) runs the host-driven rounds (the LM step on
        the host, one synchronisation per round); True keeps the fits resident on the
        GPU (staging, evaluation and LM steps on the device, one host wait per 8
        rounds)."""
        return self._chk(_LIB.ifg_batch_set_lm_resident(self.h, int(on)), "SetLMResident")

    def Autotune(self, params):
        """One column-strip width and one rows-per-workgroup multiplier for every model,
        timed on GetFitStatistic at params; returns (trials, width, multiplier)."""
        x = self._concat(params)
        c, f = C.c_int(0), np.zeros(1)
        n = self._chk(_LIB.ifg_batch_autotune(self.h, _ptr(x), C.byref(c), _ptr(f)), "Autotune")
        return n, c.value, float(f[0])
