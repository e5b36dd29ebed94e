"""GPU model objects from imfit's inputs: a configuration file and FITS images.

`imfit_setup` does what imfit does between reading its command line and starting
the solver (core/imfit_main.cpp:182-356 and core/setup_model_object.cpp:36-153),
with the GPU ModelObject in place of the CPU one:

  1. read the configuration file with limits, apply its GAIN / READNOISE / EXPTIME /
     NCOMBINED / ORIGINAL_SKY unless given as arguments (imfit_main.cpp:184-191);
  2. read the data image; a section "name.fits[x1:x2,y1:y2]" gives X0/Y0 offsets
     (imfit_main.cpp:200-211); read mask and noise images, which must match the
     data's size (core/getimages.cpp:35-110), and the PSF;
  3. PSF, data and image characteristics, oversampled PSFs with their regions
     shifted by the offsets (core/psf_oversampling_info.cpp:194-204), mask, then the
     statistic and the noise image (setup_model_object.cpp:58-153);
  4. the functions of the file (imfit_main.cpp:253-266), whose parameter count must
     match the file's values;
  5. FinalSetupForFitting, then X0/Y0 values and limits are moved into the
     section's frame (imfit_main.cpp:279-356).

Keyword arguments are named after imfit's command-line options (--psf, --noise,
--mask, --errors-are-variances, --gain, --sky, --overpsf_region, ...).
`makeimage_setup` is the counterpart for makeimage (core/makeimage_main.cpp):
no data, image size from NCOLS/NROWS or the arguments.
"""
import numpy as np

from . import (INTERPOLATION, MASK_ZERO_IS_BAD, MASK_ZERO_IS_GOOD, PRECISION_FP64, WEIGHTS_ARE_SIGMAS,
               WEIGHTS_ARE_VARIANCES, WEIGHTS_ARE_WEIGHTS, BatchGPU, ModelObjectGPU, init_genrand)
from .config import ConfigError, ReadConfigFile, SplitString
from .fits import DetermineImageOffset, FitsError, read_image, write_image


class SetupError(RuntimeError):
    pass


def _as_list(v):
    if v is None:
        return []
    return list(v) if isinstance(v, (list, tuple)) else [v]


def _region_coords(region):
    """(x1, x2, y1, y2) of an "x1:x2,y1:y2" region string (core/utilities.cpp:300-344)."""
    text = region.strip().lstrip("[").rstrip("]")
    parts = SplitString(text, ",")
    if len(parts) != 2:
        raise SetupError("incorrect region format %r" % region)
    out = []
    for p in parts:
        ab = SplitString(p, ":")
        if len(ab) != 2:
            raise SetupError("incorrect region format %r" % region)
        out += [int(ab[0]), int(ab[1])]
    return out[0], out[1], out[2], out[3]


def _read(name, what, shape=None):
    try:
        img = read_image(name)
    except FitsError as e:
        raise SetupError("unable to read %s file %r: %s" % (what, name, e))
    if shape is not None and img.shape != shape:
        raise SetupError("dimensions of %s image %r (%d columns, %d rows) do not match the data image "
                         "(%d columns, %d rows)" % (what, name, img.shape[1], img.shape[0], shape[1], shape[0]))
    return img


def _add_functions(model, cfg, subsampling):
    method = None
    for k, (name, new_set) in enumerate(cfg.functions()):
        model.AddFunction(name, new_set)
        # the first PointSource's "method" chooses the model's PSF interpolation
        # (core/model_object.cpp:294-305, func_pointsource.cpp:137-170)
        if name in ("PointSource", "PointSourceRot") and method is None:
            opt = cfg.optional_params[k] if k < len(cfg.optional_params) else {}
            m = str(opt.get("method", "bicubic")).lower()
            if m not in INTERPOLATION or (name == "PointSourceRot" and m == "lanczos3"):
                raise SetupError("unidentified interpolation type %r for %s" % (opt.get("method"), name))
            method = INTERPOLATION[m]
            model.SetupPsfInterpolation(method)
    model.SetSubsampling(subsampling)
    n = model.GetNParams()
    if n != len(cfg.parameters):
        raise SetupError("number of input parameters (%d) does not equal required number of parameters "
                         "for specified functions (%d)" % (len(cfg.parameters), n))


def _add_oversampled(model, overpsf, overpsf_scale, overpsf_region, xoff, yoff, normalize):
    files, scales, regions = _as_list(overpsf), _as_list(overpsf_scale), _as_list(overpsf_region)
    if not files:
        if regions:
            raise SetupError("oversampled regions given without an oversampled PSF")
        return
    if not regions:
        raise SetupError('an oversampled PSF needs at least one region ("overpsf_region")')
    if len(files) != len(scales):
        raise SetupError("number of oversampling scales (%d) is not the same as number of oversampled-PSF "
                         "images (%d)" % (len(scales), len(files)))
    if len(files) > 1 and len(files) != len(regions):
        raise SetupError("number of oversampled-PSF images (%d) must be 1 or the number of regions (%d)"
                         % (len(files), len(regions)))
    psf, scale = None, None
    for k, region in enumerate(regions):
        if k == 0 or len(files) > 1:
            psf, scale = _read(files[k], "oversampled PSF"), int(scales[k])
        x1, x2, y1, y2 = _region_coords(region)
        model.AddOversampledPsf(psf, scale, x1 - xoff, x2 - xoff, y1 - yoff, y2 - yoff, normalize)


class ImfitSetup:
    """A configured GPU model with the fit's starting point.

    params: initial values in the data image's frame (X0/Y0 minus the section offsets);
    limits: [(lower, upper) or None] in the same frame; fixed: bool per parameter;
    offsets: (X0 offset, Y0 offset); config: the parsed ConfigFile."""

    def __init__(self, model, config, params, limits, fixed, offsets, chars):
        self.model, self.config = model, config
        self.params, self.limits, self.fixed = params, limits, fixed
        self.offsets = offsets
        self.image_characteristics = chars
        self.n_free = int(len(params) - np.count_nonzero(fixed))

    def _xy(self):
        return [i for i, _ in self.config.xy_index], [j for _, j in self.config.xy_index if j >= 0]

    def to_file_frame(self, params):
        """Parameters with X0/Y0 in the full image's pixel frame, as imfit prints and
        saves them (the offsets added back)."""
        p = np.array(params, dtype=np.float64)
        xs, ys = self._xy()
        p[xs] += self.offsets[0]
        p[ys] += self.offsets[1]
        return p

    def fit_statistic(self, params=None):
        return self.model.GetFitStatistic(self.params if params is None else params)

    def to_data_frame(self, params):
        """Inverse of to_file_frame: X0/Y0 relative to the data image (section)."""
        p = np.array(params, dtype=np.float64)
        xs, ys = self._xy()
        p[xs] -= self.offsets[0]
        p[ys] -= self.offsets[1]
        return p

    def fit(self, ftol=1e-8, method="mpfit", jacobian=None, **kw):
        """Levenberg-Marquardt from the configuration's values, limits and fixed flags.
        Returns (best parameters in the file frame, LMResult)."""
        x, res = self.model.FitLM(self.params, limits=self.limits, fixed=self.fixed, ftol=ftol, method=method,
                                  jacobian=jacobian, **kw)
        return self.to_file_frame(x), res

    def clone(self):
        """A second, independent GPU model set up from the same files and options."""
        return imfit_setup(**self._kwargs)

    def bootstrap(self, best, n_iterations, seed, ftol=1e-8, batched=False):
        """Bootstrap resampling of the fit (core/bootstrap_errors.cpp:150-232): seed the
        generator, UseBootstrap (which draws a first sample), then per iteration a new
        sample and an L-M fit from `best` (file frame).  Returns the parameters of the
        successful fits in the file frame, one row per iteration.

        batched=False fits the samples one after another with mpfit and a finite-
        difference Jacobian, as imfit does.  batched=True draws the same samples into
        one model each (clones of this setup) and fits them all at once with the
        normal-equation LM of BatchGPU."""
        start = self.to_data_frame(best)
        rows = []
        if not batched:
            init_genrand(seed)
            self.model.UseBootstrap()
            try:
                for _ in range(n_iterations):
                    self.model.MakeBootstrapSample()
                    x, res = self.model.FitLM(start, limits=self.limits, fixed=self.fixed, ftol=ftol)
                    if res.status > 0:
                        rows.append(self.to_file_frame(x))
            finally:
                self.model.StopBootstrap()
            return np.array(rows)
        clones = [self.clone() for _ in range(n_iterations)]
        models = [c.model for c in clones]
        init_genrand(seed)
        models[0].UseBootstrap()
        for m in models:
            m.MakeBootstrapSample()
        batch = BatchGPU(models)
        xs, results = batch.FitLM([start] * n_iterations, limits=[self.limits] * n_iterations,
                                  fixed=[self.fixed] * n_iterations, ftol=ftol)
        for x, res in zip(xs, results):
            if res.status > 0:
                rows.append(self.to_file_frame(x))
        return np.array(rows)


def imfit_setup(config, image, psf=None, noise=None, mask=None, errors_are_variances=False,
                errors_are_weights=False, mask_zero_is_bad=False, gain=None, readnoise=None, exptime=None,
                ncombined=None, sky=None, overpsf=None, overpsf_scale=None, overpsf_region=None,
                no_subsampling=False, no_normalize=False, cashstat=False, poisson_mlr=False, model_errors=False,
                precision=PRECISION_FP64, device=0):
    """Configure a ModelObjectGPU from a configuration file and FITS file names as
    imfit would; returns an ImfitSetup.  Raises ConfigError, SetupError or the
    library's IfgError where imfit prints an error and exits."""
    kwargs = dict(locals())
    cfg = ReadConfigFile(config, mode2D=True, with_limits=True)
    chars = cfg.imfit_options(gain=gain, readnoise=readnoise, exptime=exptime, ncombined=ncombined,
                              original_sky=sky)
    data = _read(image, "data")
    xoff, yoff = DetermineImageOffset(image)
    mask_img = _read(mask, "mask", data.shape) if mask else None
    noise_img = _read(noise, "noise", data.shape) if noise else None
    psf_img = _read(psf, "PSF") if psf else None
    if cashstat and poisson_mlr:
        raise SetupError("choose one of the Cash statistic and the Poisson MLR statistic")

    m = ModelObjectGPU(device, precision)
    _add_functions(m, cfg, not no_subsampling)
    if psf_img is not None:
        m.AddPSFVector(psf_img, not no_normalize)
    m.AddImageDataVector(data)
    m.AddImageCharacteristics(chars["gain"], chars["readnoise"], chars["exptime"], chars["ncombined"],
                              chars["original_sky"])
    _add_oversampled(m, overpsf, overpsf_scale, overpsf_region, xoff, yoff, not no_normalize)
    if mask_img is not None:
        m.AddMaskVector(mask_img, MASK_ZERO_IS_BAD if mask_zero_is_bad else MASK_ZERO_IS_GOOD)
    if cashstat:
        m.UseCashStatistic()
    elif poisson_mlr:
        m.UsePoissonMLR()
    elif noise_img is not None:
        etype = WEIGHTS_ARE_WEIGHTS if errors_are_weights else (
            WEIGHTS_ARE_VARIANCES if errors_are_variances else WEIGHTS_ARE_SIGMAS)
        m.AddErrorVector(noise_img, etype)
    elif model_errors:
        m.UseModelErrors()
    m.FinalSetupForFitting()

    params = cfg.parameters.copy()
    limits = cfg.limit_pairs()
    for ix, iy in cfg.xy_index:
        for k, off in ((ix, xoff), (iy, yoff)):
            if k < 0:
                continue
            params[k] -= off
            cfg.limits[k].offset = off
            if limits[k] is not None:
                limits[k] = (limits[k][0] - off, limits[k][1] - off)
    s = ImfitSetup(m, cfg, params, limits, cfg.fixed(), (xoff, yoff), chars)
    s._kwargs = kwargs
    return s


class MakeimageSetup:
    def __init__(self, model, config, ncols, nrows):
        self.model, self.config = model, config
        self.params = config.parameters.copy()
        self.ncols, self.nrows = ncols, nrows

    def image(self, params=None):
        """Model image (rows, columns) for the configuration's or the given parameters."""
        self.model.CreateModelImage(self.params if params is None else params)
        return self.model.GetModelImageVector().reshape(self.nrows, self.ncols)

    def save(self, path, params=None, comments=()):
        """Write the model image as makeimage does: single-precision FITS."""
        write_image(self.image(params), path, comments)


def makeimage_setup(config, ncols=None, nrows=None, refimage=None, psf=None, overpsf=None, overpsf_scale=None,
                    overpsf_region=None, no_subsampling=False, no_normalize=False, precision=PRECISION_FP64,
                    device=0):
    """Configure a ModelObjectGPU for model images from a configuration file, as
    makeimage does (core/makeimage_main.cpp): size from the arguments, else from a
    reference image, else NCOLS/NROWS of the file."""
    cfg = ReadConfigFile(config, mode2D=True, with_limits=False)
    if refimage is not None and (ncols is None or nrows is None):
        ref = _read(refimage, "reference")
        ncols, nrows = ref.shape[1], ref.shape[0]
    ncols, nrows = cfg.makeimage_options(ncols, nrows)
    if not ncols or not nrows:
        raise SetupError("insufficient image dimensions (or no reference image) supplied")
    psf_img = _read(psf, "PSF") if psf else None
    m = ModelObjectGPU(device, precision)
    _add_functions(m, cfg, not no_subsampling)
    if psf_img is not None:
        m.AddPSFVector(psf_img, not no_normalize)
    m.SetupModelImage(ncols, nrows)
    _add_oversampled(m, overpsf, overpsf_scale, overpsf_region, 0, 0, not no_normalize)
    return MakeimageSetup(m, cfg, ncols, nrows)


__all__ = ["imfit_setup", "makeimage_setup", "ImfitSetup", "MakeimageSetup", "SetupError", "ConfigError"]
