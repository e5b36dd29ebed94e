"""One model fitted to several images at once (multimfit), on the GPU.

Follows ModelObjectMultImage (core/model_object_multimage.cpp:300-445) and its
parameter bookkeeping (core/param_holder.cpp:150-272,
core/paramvector_processing.cpp:42-104):

  * global parameters: five image-description parameters (PIXEL_SCALE, IMAGE_PA,
    FLUX_SCALE, X0, Y0) for each image after the first (the reference image), then
    the model's parameters as a configuration file lists them;
  * image 0 takes the model parameters unchanged; image k > 0 takes its own X0, Y0
    for the first function set, and for every later set the reference-image offset
    from the first set, rotated by IMAGE_PA and scaled by PIXEL_SCALE
    (CalculateOffset_X0Y0);
  * each function then sees the image's parameters as its Setup applies them
    (e.g. func_sersic.cpp:113-117): PA - IMAGE_PA, intensities x FLUX_SCALE,
    lengths x PIXEL_SCALE (BrokenExponential's alpha / PIXEL_SCALE);
  * the statistic is the sum of the images' statistics, the deviates their
    concatenation; every image's evaluation runs in one batched launch.

Image-info files (core/imageparams_file_parser.cpp): blocks opened by IMAGE_START,
with DATA, ERROR, MASK, PSF, GAIN, READNOISE, EXPTIME, NCOMBINED, ORIGINAL_SKY,
WEIGHT, the flags MASK_ZERO_IS_BAD / ERRORS_ARE_VARIANCES / ERRORS_ARE_WEIGHTS,
the image-description parameters with limits or "fixed" (ignored for the first
image), oversampled PSFs (OVERSAMPLED_PSF, OVERSAMPLE_SCALE, OVERSAMPLED_REGION,
:499-516), and a block's own functions after FUNCTIONS_START (:536-580): a function
section of the configuration-file form whose functions exist in that image only,
in their own function sets after the global ones, with their parameters taken as
given (no image-description scaling, core/model_object_multimage.cpp: only global
functions see SetImageParameters) and appended to the global vector image by
image (core/param_holder.cpp:236-246).

Fits: mpfit on the images' concatenated deviates (`FitMpfit`, multimfit's L-M:
the same mpfit code as ifg_fit_lm, driven through ifg_mpfit), or
Levenberg-Marquardt on the normal equations (`FitLM`), where each image's J^T J
and J^T r come from the GPU (analytic Jacobians) and are pushed through the
derivatives of the parameter assembly.  Bootstrap resampling draws every image's
sample in image order from the one MT19937 stream (ModelObjectMultImage::
MakeBootstrapSample, core/model_object_multimage.cpp:749-760).
"""
import math
import os

import numpy as np

from . import (MASK_ZERO_IS_BAD, MASK_ZERO_IS_GOOD, PRECISION_FP64, WEIGHTS_ARE_SIGMAS, WEIGHTS_ARE_VARIANCES,
               WEIGHTS_ARE_WEIGHTS, BatchGPU, ModelObjectGPU, init_genrand, mpfit)
from .config import (ChopComment, ConfigError, ParamLimit, ParseFunctionSection, ReadConfigFile, SplitString,
                     TrimWhitespace, strtod)
from .fits import DetermineImageOffset, read_image

N_IMAGE_PARAMS = 5
IMAGE_PARAM_NAMES = ("PIXEL_SCALE", "IMAGE_PA", "FLUX_SCALE", "X0", "Y0")
DEG2RAD = 0.017453292519943295

# Own-parameter roles under the image-description parameters: (index of PA or None,
# intensity indices, length indices, inverse-length indices).
ROLES = {
    "Sersic": (0, (3,), (4,), ()),
    "Sersic_GenEllipse": (0, (4,), (5,), ()),
    "Exponential": (0, (2,), (3,), ()),
    "BrokenExponential": (0, (2,), (3, 4, 5), (6,)),
    "Gaussian": (0, (2,), (3,), ()),
    "Moffat": (0, (2,), (3,), ()),
    "FlatSky": (None, (0,), (), ()),
    "ExponentialDisk3D": (0, (2,), (3, 5), ()),
    "EdgeOnDisk": (0, (1,), (2, 4), ()),
    "PointSource": (None, (0,), (), ()),          # func_pointsource.cpp:185-191
    "PointSourceRot": (0, (1,), (), ()),          # func_pointsource-rot.cpp:175-182
}
NPARAMS = {"Sersic": 5, "Sersic_GenEllipse": 6, "Exponential": 4, "BrokenExponential": 7, "Gaussian": 4,
           "Moffat": 5, "FlatSky": 1, "ExponentialDisk3D": 6, "EdgeOnDisk": 5, "PointSource": 1,
           "PointSourceRot": 2}


class ImageInfo:
    def __init__(self):
        self.data = self.error = self.mask = self.psf = None
        self.gain, self.readnoise, self.exptime, self.ncombined, self.original_sky = 1.0, 0.0, 1.0, 1, 0.0
        self.weight = 1.0
        self.mask_type, self.error_type = MASK_ZERO_IS_GOOD, WEIGHTS_ARE_SIGMAS
        self.params = [1.0, 0.0, 1.0, 0.0, 0.0]     # PIXEL_SCALE IMAGE_PA FLUX_SCALE X0 Y0
        self.limits = [ParamLimit() for _ in range(N_IMAGE_PARAMS)]
        self.overpsf, self.overpsf_scale, self.overpsf_region = [], [], []
        self.functions = None   # the block's own function section (FunctionSection), if any


def read_image_info(path, resolve=None):
    """ImageInfo per IMAGE_START block of an image-info file.  resolve(name) maps the
    file names it lists (default: as given)."""
    resolve = resolve or (lambda name: name)
    with open(path) as f:
        lines = f.read().split("\n")
    infos = []
    func_lines = None   # lines of the current block's function section, with their numbers
    for k, raw in enumerate(lines, start=1):
        line = TrimWhitespace(raw)
        if not line or line[0] == "#":
            continue
        if "IMAGE_START" in line:
            if func_lines is not None:
                _finish_functions(infos[-1], func_lines)
                func_lines = None
            infos.append(ImageInfo())
            continue
        if not infos:
            raise ConfigError(0, k, "image-info line before the first IMAGE_START")
        if func_lines is not None:
            func_lines.append((line, k))
            continue
        if "FUNCTIONS_START" in line:   # the rest of the block is a function section
            func_lines = []
            continue
        info = infos[-1]
        pieces = SplitString(ChopComment(line))
        key = pieces[0]
        if key == "MASK_ZERO_IS_BAD":
            info.mask_type = MASK_ZERO_IS_BAD
        elif key == "MASK_ZERO_IS_GOOD":
            info.mask_type = MASK_ZERO_IS_GOOD
        elif key == "ERRORS_ARE_VARIANCES":
            info.error_type = WEIGHTS_ARE_VARIANCES
        elif key == "ERRORS_ARE_WEIGHTS":
            info.error_type = WEIGHTS_ARE_WEIGHTS
        elif key == "ERRORS_ARE_SIGMAS":
            info.error_type = WEIGHTS_ARE_SIGMAS
        elif len(pieces) < 2:
            raise ConfigError(0, k, "image-info line %d needs a value" % k)
        elif key in ("DATA", "ERROR", "MASK", "PSF"):
            setattr(info, key.lower(), resolve(pieces[1]))
        elif key in ("GAIN", "READNOISE", "EXPTIME", "ORIGINAL_SKY", "WEIGHT"):
            attr = {"GAIN": "gain", "READNOISE": "readnoise", "EXPTIME": "exptime",
                    "ORIGINAL_SKY": "original_sky", "WEIGHT": "weight"}[key]
            setattr(info, attr, strtod(pieces[1])[0])
        elif key == "NCOMBINED":
            info.ncombined = int(strtod(pieces[1])[0])
        elif key in IMAGE_PARAM_NAMES:
            j = IMAGE_PARAM_NAMES.index(key)
            info.params[j] = strtod(pieces[1])[0]
            lim = ParamLimit()
            if len(pieces) > 2:
                if pieces[2] == "fixed":
                    lim.fixed = 1
                else:
                    lo_hi = SplitString(pieces[2], ",")
                    if len(lo_hi) != 2:
                        raise ConfigError(0, k, "bad limits on image-info line %d" % k)
                    lim.limited = [1, 1]
                    lim.limits = [strtod(lo_hi[0])[0], strtod(lo_hi[1])[0]]
            info.limits[j] = lim
        elif key == "OVERSAMPLED_PSF":
            info.overpsf.append(resolve(pieces[1]))
        elif key == "OVERSAMPLE_SCALE":
            info.overpsf_scale.append(int(strtod(pieces[1])[0]))
        elif key == "OVERSAMPLED_REGION":
            info.overpsf_region.append(pieces[1])
    if func_lines is not None:
        _finish_functions(infos[-1], func_lines)
    if not infos:
        raise ConfigError(0, -1, "no IMAGE_START block in %r" % path)
    return infos


def _finish_functions(info, func_lines):
    """Parse a block's FUNCTIONS_START section (core/imageparams_file_parser.cpp:620-650)."""
    lines = [ln for ln, _ in func_lines]
    numbers = [k for _, k in func_lines]
    info.functions = ParseFunctionSection(lines, True, numbers, checkParameterLimits=True)


class ModelObjectMultImageGPU:
    """A configuration file's model on several images (one GPU ModelObject each).

    stat: "chi2" (data-based errors, or the images' ERROR files), "model" (errors from
    the model, imfit --model-errors), "cash" or "mlr"."""

    def __init__(self, config, infos, stat="chi2", precision=PRECISION_FP64, device=0):
        self.cfg = config if hasattr(config, "function_names") else ReadConfigFile(config)
        self.infos = infos
        self.n_images = len(infos)
        self.stat = stat
        self.models, self.offsets = [], []
        for info in infos:
            m = ModelObjectGPU(device, precision)
            for name, new_set in self.cfg.functions():
                m.AddFunction(name, new_set)
            if info.functions is not None:   # the image's own functions, new function sets after the global ones
                starts = set(info.functions.fset_starts)
                for j, name in enumerate(info.functions.function_names):
                    m.AddFunction(name, j in starts)
            if info.psf is not None:
                m.AddPSFVector(read_image(info.psf))
            data = read_image(info.data)
            m.AddImageDataVector(data)
            m.AddImageCharacteristics(info.gain, info.readnoise, info.exptime, info.ncombined, info.original_sky)
            off = DetermineImageOffset(info.data)
            self._add_oversampled(m, info, off)
            if info.mask is not None:
                m.AddMaskVector(read_image(info.mask), info.mask_type)
            if stat == "cash":
                m.UseCashStatistic()
            elif stat == "mlr":
                m.UsePoissonMLR()
            elif info.error is not None:
                m.AddErrorVector(read_image(info.error), info.error_type)
            elif stat == "model":
                m.UseModelErrors()
            m.FinalSetupForFitting()
            self.models.append(m)
            self.offsets.append(off)
        self.batch = BatchGPU(self.models) if self.n_images > 1 else None
        self.n_image_params = N_IMAGE_PARAMS * (self.n_images - 1)
        self.n_model = len(self.cfg.parameters)
        self.n_local = [0 if info.functions is None else len(info.functions.parameters) for info in infos]
        # function layout: (name, first own parameter in the model vector)
        self._funcs, off = [], 0
        for k, (name, new_set) in enumerate(self.cfg.functions()):
            if new_set:
                off += 2
            self._funcs.append((name, off, new_set))
            off += NPARAMS[name]
        self._set_xy = [(i, j) for i, j in self.cfg.xy_index]

    @staticmethod
    def _add_oversampled(m, info, off):
        """OVERSAMPLED_PSF / OVERSAMPLE_SCALE / OVERSAMPLED_REGION of one image: one PSF and
        scale for every region, or one per region (core/multimfit_main.cpp, as imfit's
        command-line options); regions in the image file's pixels, moved into an image
        section's frame."""
        files, scales, regions = info.overpsf, info.overpsf_scale, info.overpsf_region
        if not files and not regions:
            return
        if not files or not regions or len(scales) != len(files):
            raise ConfigError(0, -1, "oversampled PSF, scale and region must be given together")
        if len(files) > 1 and len(files) != len(regions):
            raise ConfigError(0, -1, "one oversampled PSF, or one per region")
        for k, region in enumerate(regions):
            q = 0 if len(files) == 1 else k
            x, y = SplitString(region, ",")
            x1, x2 = [int(v) for v in SplitString(x, ":")]
            y1, y2 = [int(v) for v in SplitString(y, ":")]
            m.AddOversampledPsf(read_image(files[q]), int(scales[q]), x1 - off[0], x2 - off[0], y1 - off[1],
                                y2 - off[1])

    # ---- parameters
    def GetNParams(self):
        return self.n_image_params + self.n_model + sum(self.n_local)

    def _local_start(self, k):
        """Index of image k's first own (per-image function) parameter in the global vector."""
        return self.n_image_params + self.n_model + sum(self.n_local[:k])

    def initial_parameters(self):
        """The global vector of the files' values, X0/Y0 moved into the images' frames
        (image sections, core/utilities_multimfit.cpp:14-63): image-description
        parameters, the model's, then each image's own functions' parameters."""
        p = [v for info in self.infos[1:] for v in info.params] + list(self.cfg.parameters)
        for info in self.infos:
            if info.functions is not None:
                p += list(info.functions.parameters)
        return self._to_data_frame(np.array(p, dtype=np.float64))

    def limits_and_fixed(self):
        lims, fixed = [], []
        for info in self.infos[1:]:
            for lim in info.limits:
                lims.append((lim.limits[0], lim.limits[1]) if lim.limited[0] and lim.limited[1] else None)
                fixed.append(lim.fixed == 1)
        lims += self.cfg.limit_pairs()
        fixed += list(self.cfg.fixed())
        for info in self.infos:   # core/multimfit_main.cpp:585-612
            if info.functions is None:
                continue
            for lim in info.functions.limits:
                lims.append((lim.limits[0], lim.limits[1]) if lim.limited[0] and lim.limited[1] else None)
                fixed.append(lim.fixed == 1)
        shift = self._frame_shift()
        out = [None if l is None else (l[0] - shift[k], l[1] - shift[k]) for k, l in enumerate(lims)]
        return out, np.array(fixed, dtype=bool)

    def _frame_shift(self):
        s = np.zeros(self.GetNParams())
        for k in range(1, self.n_images):
            s[N_IMAGE_PARAMS * k - 2] = self.offsets[k][0]
            s[N_IMAGE_PARAMS * k - 1] = self.offsets[k][1]
        for i, j in self._set_xy:
            s[self.n_image_params + i] = self.offsets[0][0]
            s[self.n_image_params + j] = self.offsets[0][1]
        for k, info in enumerate(self.infos):   # own functions: in their image's frame
            if info.functions is None:
                continue
            b = self._local_start(k)
            for i, j in info.functions.xy_index:
                s[b + i] = self.offsets[k][0]
                s[b + j] = self.offsets[k][1]
        return s

    def _to_data_frame(self, p):
        return p - self._frame_shift()

    def to_file_frame(self, p):
        return np.asarray(p, dtype=np.float64) + self._frame_shift()

    def params_for_image(self, p, k):
        """The parameter vector image k's ModelObject evaluates, and its derivative
        with respect to the global vector (n_model x n_global)."""
        p = np.asarray(p, dtype=np.float64)
        nI, nM, nG = self.n_image_params, self.n_model, self.GetNParams()
        M = p[nI:nI + nM].copy()
        A = np.zeros((nM, nG))
        A[:, nI:nI + nM] = np.eye(nM)
        out, A = self._global_for_image(p, k, M, A)
        nL = self.n_local[k]
        if nL == 0:
            return out, A
        b = self._local_start(k)   # the image's own parameters, as given (param_holder.cpp:236-246)
        L = np.zeros((nL, nG))
        L[:, b:b + nL] = np.eye(nL)
        return np.concatenate([out, p[b:b + nL]]), np.vstack([A, L])

    def _global_for_image(self, p, k, M, A):
        nI, nM = self.n_image_params, self.n_model
        if k == 0:
            return M, A
        base = N_IMAGE_PARAMS * (k - 1)
        s, rot, fsc, x0i, y0i = p[base:base + N_IMAGE_PARAMS]
        iS, iR, iF, iX, iY = range(base, base + N_IMAGE_PARAMS)
        out = M.copy()
        X00, Y00 = M[0], M[1]
        out[0], out[1] = x0i, y0i
        A[0, :], A[1, :] = 0.0, 0.0
        A[0, iX], A[1, iY] = 1.0, 1.0
        c, sn = math.cos(DEG2RAD * rot), math.sin(DEG2RAD * rot)
        for i, j in self._set_xy[1:]:   # CalculateOffset_X0Y0 (paramvector_processing.cpp:80-104)
            dx, dy = M[i] - X00, M[j] - Y00
            ex, ey = dx * c + dy * sn, -dx * sn + dy * c
            out[i], out[j] = x0i + s * ex, y0i + s * ey
            A[i, :], A[j, :] = 0.0, 0.0
            A[i, iX], A[j, iY] = 1.0, 1.0
            A[i, iS], A[j, iS] = ex, ey
            A[i, iR], A[j, iR] = s * DEG2RAD * (-dx * sn + dy * c), s * DEG2RAD * (-dx * c - dy * sn)
            A[i, nI + i], A[i, nI + j], A[i, nI + 0], A[i, nI + 1] = s * c, s * sn, -s * c, -s * sn
            A[j, nI + i], A[j, nI + j], A[j, nI + 0], A[j, nI + 1] = -s * sn, s * c, s * sn, -s * c
        for name, off, _ in self._funcs:   # FunctionObject::Setup with the image's parameters
            pa, ints, lens, invs = ROLES[name]
            if pa is not None:
                out[off + pa] = M[off + pa] - rot
                A[off + pa, iR] = -1.0
            for q in ints:
                out[off + q] = M[off + q] * fsc
                A[off + q, nI + off + q], A[off + q, iF] = fsc, M[off + q]
            for q in lens:
                out[off + q] = M[off + q] * s
                A[off + q, nI + off + q], A[off + q, iS] = s, M[off + q]
            for q in invs:
                out[off + q] = M[off + q] / s
                A[off + q, nI + off + q], A[off + q, iS] = 1.0 / s, -M[off + q] / (s * s)
        return out, A

    # ---- evaluation
    def GetFitStatistic(self, p):
        """Sum of the images' statistics (core/model_object_multimage.cpp:363-408)."""
        vecs = [self.params_for_image(p, k)[0] for k in range(self.n_images)]
        if self.batch is None:
            return self.models[0].GetFitStatistic(vecs[0])
        return float(np.sum(self.batch.GetFitStatistic(vecs)))

    def GetFitStatistics(self, p):
        vecs = [self.params_for_image(p, k)[0] for k in range(self.n_images)]
        return np.array([m.GetFitStatistic(v) for m, v in zip(self.models, vecs)])

    def ComputeDeviates(self, p):
        """The images' deviates one after another (core/model_object_multimage.cpp:411-438)."""
        return np.concatenate([m.ComputeDeviates(self.params_for_image(p, k)[0])
                               for k, m in enumerate(self.models)])

    def NormalEquations(self, p, free):
        """J^T J, J^T r and the statistic of the global parameters `free`, from each
        image's GPU normal equations over all its parameters (one batched launch)."""
        vecs, As = zip(*[self.params_for_image(p, k) for k in range(self.n_images)])
        if self.batch is None:
            res = [self.models[0].NormalEquations(vecs[0])]
        else:
            pairs, stats = self.batch.NormalEquations(list(vecs))
            res = [(a, b, st) for (a, b), st in zip(pairs, stats)]
        n = len(free)
        jtj, jtr, stat = np.zeros((n, n)), np.zeros(n), 0.0
        for (a, b, st), A in zip(res, As):
            Af = A[:, free]
            jtj += Af.T @ a @ Af
            jtr += Af.T @ b
            stat += st
        return jtj, jtr, stat

    def FitMpfit(self, p0=None, ftol=1e-8, maxiter=1000):
        """multimfit's L-M: mpfit on the images' concatenated deviates, forward-difference
        Jacobian, within the files' limits and fixed flags (ftol 1e-8 as multimfit's
        default).  Returns (parameters in the files' frame, statistic, LMResult)."""
        p = self.initial_parameters() if p0 is None else np.array(p0, dtype=np.float64)
        lims, fixed = self.limits_and_fixed()
        m = sum(int(mo.GetNDataValues()) for mo in self.models)
        x, res = mpfit(self.ComputeDeviates, p, m, limits=lims, fixed=fixed, ftol=ftol, maxiter=maxiter)
        return self.to_file_frame(x), self.GetFitStatistic(x), res

    def Bootstrap(self, n, seed, best, ftol=1e-8):
        """n bootstrap fits from the best-fit vector `best` (data frame): seed the MT19937,
        switch every image to bootstrap samples, then per iteration draw each image's
        sample in image order and refit with mpfit (core/bootstrap_errors.cpp:160-250).
        Returns the successful fits' parameters in the files' frame, one row each."""
        init_genrand(seed)
        for mo in self.models:
            mo.UseBootstrap()
        lims, fixed = self.limits_and_fixed()
        m = sum(int(mo.GetNDataValues()) for mo in self.models)
        rows = []
        try:
            for _ in range(n):
                for mo in self.models:
                    mo.MakeBootstrapSample()
                x, res = mpfit(self.ComputeDeviates, best, m, limits=lims, fixed=fixed, ftol=ftol)
                if res.status > 0:
                    rows.append(self.to_file_frame(x))
        finally:
            for mo in self.models:
                mo.StopBootstrap()
        return np.array(rows)

    def FitLM(self, p0=None, ftol=1e-8, maxiter=500):
        """Levenberg-Marquardt on the summed statistic within the files' limits and fixed
        flags.  Returns (parameters in the files' frame, statistic, iterations)."""
        p = self.initial_parameters() if p0 is None else np.array(p0, dtype=np.float64)
        lims, fixed = self.limits_and_fixed()
        free = np.flatnonzero(~fixed)
        lo = np.array([-np.inf if lims[i] is None else lims[i][0] for i in free])
        hi = np.array([np.inf if lims[i] is None else lims[i][1] for i in free])
        lam = 1e-3
        jtj, jtr, stat = self.NormalEquations(p, free)
        it = 0
        for it in range(1, maxiter + 1):
            d = np.diag(jtj).copy()
            d[d <= 0] = 1.0
            improved = False
            while lam < 1e16:
                try:
                    step = np.linalg.solve(jtj + lam * np.diag(d), -jtr)
                except np.linalg.LinAlgError:
                    lam *= 10.0
                    continue
                q = p.copy()
                q[free] = np.clip(p[free] + step, lo, hi)
                s_new = self.GetFitStatistic(q)
                if np.isfinite(s_new) and s_new <= stat:
                    improved = True
                    break
                lam *= 10.0
            if not improved:
                break
            rel = (stat - s_new) / max(stat, 1e-300)
            p = q
            lam = max(lam / 10.0, 1e-12)
            jtj, jtr, stat = self.NormalEquations(p, free)
            if rel < ftol:
                break
        return self.to_file_frame(p), stat, it
