"""Synthetic workloads of the benchmark configurations, built on the GPU library.

Each workload is a ModelObject call sequence (functions, PSF, data, oversampled
regions, image characteristics, statistic) plus truth and start parameters.
Data images are made the way the benchmark plan prescribes: the noise-free model
T (computed here by the GPU library in FP64 make-image mode), electrons
e ~ Poisson(gain*T), read noise r ~ N(0, rn) electrons, data = (e + r)/gain,
with numpy's PCG64 generator seeded per configuration.  PSFs are Moffat
functions evaluated by the same library (sub-pixel integration on) at the pixel
centres of an odd grid centred on pixel (P+1)/2, normalised to sum 1.
"""
import numpy as np

from . import ModelObjectGPU, PRECISION_FP64


class Workload:
    def __init__(self, name, funcs, truth, width, height, psf=None, osamp=(), gain=4.0, read_noise=0.0,
                 noise_seed=0, subsampling=True):
        self.name = name
        self.funcs = list(funcs)
        self.truth = np.asarray(truth, dtype=np.float64)
        self.width, self.height = width, height
        self.psf = psf
        self.osamp = list(osamp)   # (psf, scale, x1, x2, y1, y2)
        self.gain, self.read_noise = gain, read_noise
        self.noise_seed = noise_seed
        self.subsampling = subsampling
        self.data = None

    # -- the reference's setup order (core/setup_model_object.cpp:58-153)
    def _define(self, m):
        for name, new in self.funcs:
            m.AddFunction(name, new)
        m.SetSubsampling(self.subsampling)
        if self.psf is not None:
            m.AddPSFVector(self.psf)

    def _regions(self, m):
        for (psf, scale, x1, x2, y1, y2) in self.osamp:
            m.AddOversampledPsf(psf, scale, x1, x2, y1, y2)

    def model_image(self, params=None):
        m = ModelObjectGPU(0, PRECISION_FP64)
        self._define(m)
        m.SetupModelImage(self.width, self.height)
        self._regions(m)
        m.CreateModelImage(self.truth if params is None else params)
        return m.GetModelImageVector()

    def make_data(self):
        if self.data is None:
            T = self.model_image()
            rng = np.random.default_rng(self.noise_seed)
            e = rng.poisson(np.clip(self.gain * T, 0, None)).astype(np.float64)
            if self.read_noise > 0:
                e += rng.normal(0.0, self.read_noise, size=e.shape)
            self.data = e / self.gain
        return self.data

    def fit_model(self, precision=PRECISION_FP64):
        """A GPU ModelObject set up for fitting this workload's data."""
        return self.setup_fit(ModelObjectGPU(0, precision))

    def setup_fit(self, m):
        """Run the fitting setup sequence on any ModelObject-like `m`: data-based chi^2
        weights from GAIN and READNOISE (ORIGINAL_SKY 0, EXPTIME 1, NCOMBINED 1)."""
        data = self.make_data()
        self._define(m)
        m.AddImageDataVector(data)
        m.AddImageCharacteristics(self.gain, self.read_noise, 1.0, 1, 0.0)
        self._regions(m)
        m.FinalSetupForFitting()
        return m

    def start_params(self, seed=1, frac=0.2):
        rng = np.random.default_rng(seed)
        return self.truth * rng.uniform(1 - frac, 1 + frac, size=self.truth.size)

    def lm_start(self, seed=1):
        """LM starting point: every non-positional parameter scaled by U(0.9, 1.1), the
        centre shifted by U(-5, 5) pixels (seed 1), inside lm_limits()."""
        rng = np.random.default_rng(seed)
        x = self.truth * rng.uniform(0.9, 1.1, size=self.truth.size)
        off = 0
        for name, new in self.funcs:
            if new:
                x[off:off + 2] = self.truth[off:off + 2] + rng.uniform(-5, 5, size=2)
                off += 2
            off += {"Sersic": 5, "Exponential": 4, "FlatSky": 1, "Gaussian": 4, "Moffat": 5,
                    "ExponentialDisk3D": 6}[name]
        return x

    def lm_limits(self):
        """Parameter limits for the LM fits: centres within 20 px of the truth, PA in
        [-360, 360], ellipticities in [0, 0.95], Sersic n in [0.2, 10], the rest in
        [truth/10, truth*10]."""
        lim = []
        off = 0
        for name, new in self.funcs:
            if new:
                lim += [(self.truth[off] - 20, self.truth[off] + 20), (self.truth[off + 1] - 20,
                                                                        self.truth[off + 1] + 20)]
                off += 2
            n = {"Sersic": 5, "Exponential": 4, "FlatSky": 1, "Gaussian": 4, "Moffat": 5,
                 "ExponentialDisk3D": 6}[name]
            for k in range(n):
                t = self.truth[off + k]
                if name != "FlatSky" and k == 0:
                    lim.append((-360.0, 360.0))
                elif name == "ExponentialDisk3D" and k == 1:
                    lim.append((0.0, 90.0))
                elif name == "ExponentialDisk3D" and k == 4:
                    lim.append((1.0, 200.0))
                elif name != "FlatSky" and k == 1:
                    lim.append((0.0, 0.95))
                elif name == "Sersic" and k == 2:
                    lim.append((0.2, 10.0))
                else:
                    lim.append((t / 10.0, t * 10.0))
            off += n
        return lim


def moffat_psf(size, fwhm, beta):
    m = ModelObjectGPU(0, PRECISION_FP64)
    m.AddFunction("Moffat", True)
    m.SetupModelImage(size, size)
    c = (size + 1) / 2.0
    m.CreateModelImage(np.array([c, c, 0.0, 0.0, 1.0, fwhm, beta]))
    p = m.GetModelImageVector()
    return p / p.sum()


def cfg1():
    """256^2 Sersic + Exponential + FlatSky (sky 100), gain 4.5, read noise 0.5, no PSF."""
    truth = [129.5, 128.2, 20.0, 0.25, 2.5, 15.0, 20.0, 20.0, 0.4, 100.0, 40.0, 100.0]
    return Workload("cfg1_256_sersic_exp_sky", [("Sersic", True), ("Exponential", False), ("FlatSky", False)],
                    truth, 256, 256, gain=4.5, read_noise=0.5, noise_seed=0)


def cfg2():
    """1024^2 Sersic(n=4, r_e=60, ell=0.2) + Exponential(h=120, ell=0.5) + FlatSky,
    51^2 Moffat PSF (FWHM 4, beta 3), data pixels 493-532 oversampled 5x with a
    255^2 Moffat (FWHM 20 sub-pixels, beta 3), gain 4, noise seed 2."""
    truth = [512.5, 512.5, 20.0, 0.2, 4.0, 15.0, 60.0, 20.0, 0.5, 100.0, 120.0, 100.0]
    psf = moffat_psf(51, 4.0, 3.0)
    psf_os = moffat_psf(255, 20.0, 3.0)
    return Workload("cfg2_1024_sersic_exp_sky_psf51_osamp40x5", [("Sersic", True), ("Exponential", False),
                                                                 ("FlatSky", False)],
                    truth, 1024, 1024, psf=psf, osamp=[(psf_os, 5, 493, 532, 493, 532)], gain=4.0, noise_seed=2)


def cfg3_population(n=4096, seed=3):
    """512^2 Sersic + Exponential + FlatSky with a 35^2 Moffat PSF (FWHM 3, beta 2.5),
    and n parameter vectors uniform within +-30% of cfg1's truth."""
    base = cfg1()
    psf = moffat_psf(35, 3.0, 2.5)
    w = Workload("cfg3_512_population", base.funcs, base.truth, 512, 512, psf=psf, gain=4.5, read_noise=0.5,
                 noise_seed=3)
    rng = np.random.default_rng(seed)
    pop = base.truth * rng.uniform(0.7, 1.3, size=(n, base.truth.size))
    return w, pop


def gaussian_psf(size, sigma):
    """Gaussian PSF (func_gaussian.cpp) at the pixel centres of an odd size x size grid
    centred on pixel (size+1)/2, sub-pixel integration on, normalised to sum 1."""
    m = ModelObjectGPU(0, PRECISION_FP64)
    m.AddFunction("Gaussian", True)
    m.SetupModelImage(size, size)
    c = (size + 1) / 2.0
    m.CreateModelImage(np.array([c, c, 0.0, 0.0, 1.0, sigma]))
    p = m.GetModelImageVector()
    return p / p.sum()


def cfg4():
    """512^2 ExponentialDisk3D (PA 0, inclination 85, J_0 1, h 50, n 100, z_0 5) +
    Sersic bulge (PA 0, ell 0.1, n 1.5, I_e 20, r_e 15) sharing one centre, 25^2
    Gaussian PSF (sigma 1.5), gain 4, read noise 1 e-, noise seed 4."""
    truth = [256.5, 256.5, 0.0, 85.0, 1.0, 50.0, 100.0, 5.0, 0.0, 0.1, 1.5, 20.0, 15.0]
    return Workload("cfg4_512_expdisk3d_sersic_psf25", [("ExponentialDisk3D", True), ("Sersic", False)], truth,
                    512, 512, psf=gaussian_psf(25, 1.5), gain=4.0, read_noise=1.0, noise_seed=4)


def cutouts(n=2000, seed=4, size_range=(64, 512), psf=None, precision=None):
    """Many-cutout workload: n independent images, width and height each
    log-uniform in size_range (seed), with 1-3 components drawn from {Sersic
    (n in [0.5, 6]), Exponential, Gaussian} sharing a centre near the image centre,
    position angles, ellipticities and sizes scaled to the cutout, a 25^2 Moffat PSF
    (FWHM 3, beta 3), sky 20 as a FlatSky component, Poisson noise with gain 4 and
    read noise 2 e-.  Returns a list of Workload objects (truth = generating
    parameters)."""
    rng = np.random.default_rng(seed)
    if psf is None:
        psf = moffat_psf(25, 3.0, 3.0)
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    out = []
    for k in range(n):
        W, H = (int(round(np.exp(rng.uniform(lo, hi)))) for _ in range(2))
        s = min(W, H)
        ncomp = int(rng.integers(1, 4))
        names = list(rng.choice(["Sersic", "Exponential", "Gaussian"], size=ncomp, replace=True))
        x0 = W / 2.0 + 0.5 + rng.uniform(-0.05, 0.05) * W
        y0 = H / 2.0 + 0.5 + rng.uniform(-0.05, 0.05) * H
        funcs, truth = [], [x0, y0]
        for i, name in enumerate(names):
            funcs.append((str(name), i == 0))
            pa, ell = rng.uniform(0, 180), rng.uniform(0.0, 0.6)
            if name == "Sersic":
                truth += [pa, ell, rng.uniform(0.5, 6.0), rng.uniform(5.0, 50.0), rng.uniform(0.03, 0.15) * s]
            elif name == "Exponential":
                truth += [pa, ell, rng.uniform(20.0, 200.0), rng.uniform(0.04, 0.2) * s]
            else:
                truth += [pa, ell, rng.uniform(20.0, 200.0), rng.uniform(0.02, 0.08) * s]
        funcs.append(("FlatSky", False))
        truth.append(20.0)
        out.append(Workload("cutout%04d_%dx%d" % (k, W, H), funcs, truth, W, H, psf=psf, gain=4.0, read_noise=2.0,
                            noise_seed=int(rng.integers(0, 2**31))))
    return out
