#!/usr/bin/env python3
"""Extract the reference's own regression-test inputs and outputs into .npz fixtures.

Runs only in the build container, where the reference tree is mounted read-only at
/root/reference.  It reads FITS *data files* (images, masks, PSFs and archived
outputs) with a minimal FITS reader written here (astropy is not available), and
copies the archived numeric results (chi^2 / Cash values, iteration counts) from
the reference's printouts.  Nothing from the reference is imported or executed.

Each fixture records which reference test produced it:

  expdisk32      do_imfit_tests:63   -> tests/imfit_reference/linux_x86_64/imfit_textout1
  ic3478_64      do_imfit_tests:68   -> tests/imfit_reference/linux_x86_64/imfit_textout3
  n3073          do_imfit_tests:96-99 -> imfit_textout4, imfit_textout4c, imfit_textout4d
  sersic_exp512  do_makeimage_tests:51,155 -> tests/sersic+exp_orig.fits
  biggertest64   do_makeimage_tests:48 -> tests/biggertest_orig.fits
  gensersic512   do_makeimage_tests:49,55,129-178 -> tests/gensersictest_orig.fits,
                 tests/gensersictest612_conv_cutout512.fits (config
                 tests/makeimage_reference/config_makeimage_gensersic512.dat; PSF
                 tests/psf_moffat_35_n4699z.fits)
  oversampled    do_makeimage_tests:68,70,199,215 -> tests/oversampled_orig.fits,
                 tests/oversampled2both_orig.fits
  ic3478_64_errors  do_imfit_tests:56-58 -> imfit_textout3d2, imfit_textout3e, imfit_textout3e2
  n3073_cps      do_imfit_tests:61 -> imfit_textout4b
  poisson_lowsn20  do_imfit_tests:112 -> tests/imfit_reference/imfit_textout5c_tail
  twogauss_osamp do_imfit_tests:107 -> imfit_textout4e3
  conv_n5419     do_psfconvolve_tests:11,60 -> tests/convolve_out_n5419.fits
  conv_5x5       do_psfconvolve_tests:8-9 -> tests/psfconvolve_out{1,2}.txt (padded PSF)
  multimfit      do_multimfit_tests.sh:59-85 -> multimfit_textout1a, 1d, 1e, 3a, 3b
  ptsource       do_imfit_tests:178 -> imfit_textout8a (PointSource + fixed FlatSky on
                 tests/psf_moffat_fwhm2.fits, PSF tests/psf_moffat_fwhm2_35x35.fits)
  bootstrap_ic3478  do_imfit_tests:133-134 -> bootstrap_output_seed10_tail.dat,
                 bootstrap_output_seed10_2_tail.dat (seed 10, whole image and section)
"""
import os
import re
import sys

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def read_fits(path):
    """Primary-HDU image of a FITS file as a 2-D float64 array, shape (NAXIS2, NAXIS1).

    FITS stores the first axis (columns, x) fastest, so the C-order reshape gives
    row-major [row][col] exactly as imfit's pixel vectors (core/image_io.cpp:284).
    """
    with open(path, "rb") as f:
        raw = f.read()
    hdr = {}
    pos = 0
    done = False
    while not done:
        block = raw[pos:pos + 2880]
        pos += 2880
        for k in range(36):
            card = block[80 * k:80 * (k + 1)].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                val = card[10:].split("/")[0].strip()
                hdr[key] = val
    bitpix = int(hdr["BITPIX"])
    naxis = int(hdr["NAXIS"])
    dims = [int(hdr["NAXIS%d" % (i + 1)]) for i in range(naxis)]
    dtype = {-32: ">f4", -64: ">f8", 8: ">u1", 16: ">i2", 32: ">i4"}[bitpix]
    n = int(np.prod(dims))
    data = np.frombuffer(raw, dtype=dtype, count=n, offset=pos)
    native = data.astype(data.dtype.newbyteorder("="))
    arr = native.reshape(dims[::-1])
    bscale = float(hdr.get("BSCALE", "1"))
    bzero = float(hdr.get("BZERO", "0"))
    if bscale != 1.0 or bzero != 0.0:
        arr = arr.astype(np.float64) * bscale + bzero
    return arr


def grab(path, pattern, cast=float):
    with open(path) as f:
        txt = f.read()
    m = re.search(pattern, txt)
    if m is None:
        raise RuntimeError("pattern %r not found in %s" % (pattern, path))
    return cast(m.group(1))


def t(p):
    return os.path.join(REF, "tests", p)


def fit_record(path, stats=True):
    """The archived outcome of an imfit LM fit printout: initial and final statistic,
    NITER, NFEV (when printed) and the best-fit parameter values in parameter order
    (the "NAME  value # +/- err" lines)."""
    with open(path) as f:
        txt = f.read()
    best = [float(v) for v in re.findall(r"^\w+\s+(-?[0-9.]+(?:e[-+]?\d+)?) # \+/-", txt, re.M)]
    rec = {"best": np.array(best)}
    if stats:
        rec["initial"] = grab(path, r"INITIAL CHI\^2 = ([0-9.]+)")
        rec["final"] = grab(path, r"CHI-SQUARE = ([0-9.]+)")
        rec["niter"] = grab(path, r"NITER = (\d+)", int)
        rec["nfev"] = grab(path, r"NFEV = (\d+)", int)
    return rec


def flat(prefix, rec):
    return {prefix + "_" + k: v for k, v in rec.items()}


def main():
    lin = "imfit_reference/linux_x86_64/"
    # --- textout1: 32x32 Exponential, user noise image, no subsampling
    np.savez_compressed(
        os.path.join(OUT, "expdisk32.npz"),
        data=read_fits(t("testimage_expdisk32.fits")),
        noise=read_fits(t("uniform_image32.fits")),
        initial_chi2=grab(t(lin + "imfit_textout1"), r"INITIAL CHI\^2 = ([0-9.]+)"),
        final_chi2=grab(t(lin + "imfit_textout1"), r"CHI-SQUARE = ([0-9.]+)"),
        niter=grab(t(lin + "imfit_textout1"), r"NITER = (\d+)", int),
        nfev=grab(t(lin + "imfit_textout1"), r"NFEV = (\d+)", int),
    )
    # --- textout3: 64x64 Sersic, data-based errors
    np.savez_compressed(
        os.path.join(OUT, "ic3478_64.npz"),
        data=read_fits(t("ic3478rss_64x64.fits")),
        initial_chi2=grab(t(lin + "imfit_textout3"), r"INITIAL CHI\^2 = ([0-9.]+)"),
        final_chi2=grab(t(lin + "imfit_textout3"), r"CHI-SQUARE = ([0-9.]+)"),
        niter=grab(t(lin + "imfit_textout3"), r"NITER = (\d+)", int),
        nfev=grab(t(lin + "imfit_textout3"), r"NFEV = (\d+)", int),
        best=np.array([32.9439, 34.0933, 18.2612, 0.235989, 2.40028, 20.0094, 60.7613]),
    )
    # --- textout4/4c/4d: 150x200 Sersic+Exp, 35x35 PSF, mask
    np.savez_compressed(
        os.path.join(OUT, "n3073.npz"),
        data=read_fits(t("n3073rss_small.fits")),
        mask=read_fits(t("n3073rss_small_mask.fits")),
        psf=read_fits(t("psf_moffat_35_n4699z.fits")),
        chi2=grab(t("imfit_reference/imfit_textout4c"), r"CHI-SQUARE = ([0-9.]+)"),
        cash=grab(t("imfit_reference/imfit_textout4d"), r"CASH STATISTIC = (-?[0-9.]+)"),
        lm_final_chi2=grab(t(lin + "imfit_textout4"), r"CHI-SQUARE = ([0-9.]+)"),
        lm_niter=grab(t(lin + "imfit_textout4"), r"NITER = (\d+)", int),
        lm_nfev=grab(t(lin + "imfit_textout4"), r"NFEV = (\d+)", int),
        best=np.array([100.7073, 97.0877, 111.919, 0.184262, 3.2805, 78.2972, 26.7976,
                       51.724, 0.35226, 39.7753, 53.6058]),
    )
    # --- further LM fits of do_imfit_tests (archived printouts: outcome and counts)
    np.savez_compressed(   # do_imfit_tests:56-58 (config c: n fixed; data, sigma, variance errors)
        os.path.join(OUT, "ic3478_64_errors.npz"),
        sigma=read_fits(t("ic3478rss_64x64_sigma.fits")),
        variance=read_fits(t("ic3478rss_64x64_variance.fits")),
        **flat("t3d2", fit_record(t(lin + "imfit_textout3d2"))),
        **flat("t3e", fit_record(t(lin + "imfit_textout3e"))),
        **flat("t3e2", fit_record(t(lin + "imfit_textout3e2"))),
    )
    np.savez_compressed(   # do_imfit_tests:61 (counts per second, exposure time and sky)
        os.path.join(OUT, "n3073_cps.npz"),
        data=read_fits(t("n3073rss_small_cps.fits")),
        **flat("t4b", fit_record(t(lin + "imfit_textout4b"))),
    )
    np.savez_compressed(   # do_imfit_tests:112 (Poisson MLR with L-M; tail printout only)
        os.path.join(OUT, "poisson_lowsn20.npz"),
        data=read_fits(t("testimage_poisson_lowsn20.fits")),
        **flat("t5c", fit_record(t("imfit_reference/imfit_textout5c_tail"), stats=False)),
    )
    np.savez_compressed(   # do_imfit_tests:107 (two Gaussians, two oversampled regions, L-M)
        os.path.join(OUT, "twogauss_osamp.npz"),
        data=read_fits(t("twogaussian_psf+2osamp_noisy.fits")),
        psf=read_fits(t("psf_moffat_35.fits")),
        psf_osamp=read_fits(t("psf_moffat_35_oversamp3.fits")),
        **flat("t4e3", fit_record(t(lin + "imfit_textout4e3"))),
    )
    np.savez_compressed(   # do_imfit_tests:178 (PointSource with bicubic interpolation, L-M)
        os.path.join(OUT, "ptsource.npz"),
        data=read_fits(t("psf_moffat_fwhm2.fits")),
        psf=read_fits(t("psf_moffat_fwhm2_35x35.fits")),
        **flat("t8a", fit_record(t(lin + "imfit_textout8a"))),
    )
    # --- makeimage goldens (float32 images written by SaveVectorAsImage)
    np.savez_compressed(os.path.join(OUT, "sersic_exp512.npz"),
                        image=read_fits(t("sersic+exp_orig.fits")))
    np.savez_compressed(os.path.join(OUT, "biggertest64.npz"),
                        image=read_fits(t("biggertest_orig.fits")))
    np.savez_compressed(   # float32 as stored: the files are BITPIX -32
        os.path.join(OUT, "gensersic512.npz"),
        image=read_fits(t("gensersictest_orig.fits")).astype(np.float32),
        conv=read_fits(t("gensersictest612_conv_cutout512.fits")).astype(np.float32),
        psf=read_fits(t("psf_moffat_35_n4699z.fits")),
    )
    np.savez_compressed(
        os.path.join(OUT, "oversampled.npz"),
        image=read_fits(t("oversampled_orig.fits")),
        image2=read_fits(t("oversampled2both_orig.fits")),
        psf=read_fits(t("psf_standard.fits")),
        psf_osamp=read_fits(t("psf_oversamp.fits")),
    )
    # --- psfconvolve goldens
    np.savez_compressed(
        os.path.join(OUT, "conv_n5419.npz"),
        image=read_fits(t("n5419f555w_pc_derot_cutout.fits")),
        psf=read_fits(t("n5419_sin100_psf_pc2scale.fits")),
        out=read_fits(t("convolve_out_n5419.fits")),
    )

    def printed_block(path, title):
        # psfconvolve --printimages prints each image as rows of %f values after a
        # title line; return the numeric block that follows `title`.
        rows, grab_rows = [], False
        with open(path) as f:
            for line in f:
                if title in line:
                    grab_rows, rows = True, []
                    continue
                if grab_rows:
                    parts = line.split()
                    if not parts:
                        if rows:
                            break
                        continue
                    rows.append([float(v) for v in parts])
        return np.array(rows)

    # tests/psfconvolve_out2.txt: 5x5 Gaussian PSF, normalised then shifted and
    # wrapped into the 9x9 padded array (core/convolver.cpp:296-311,421-441).
    np.savez_compressed(
        os.path.join(OUT, "conv_5x5.npz"),
        gauss=read_fits(t("gaussian_5x5.fits")),
        delta=read_fits(t("deltafn_5x5.fits")),
        padded_psf_gauss=printed_block(t("psfconvolve_out2.txt"),
                                       "The whole padded, normalized PSF image"),
        padded_psf_delta=printed_block(t("psfconvolve_out1.txt"),
                                       "The whole padded, normalized PSF image"),
    )
    bootstrap_fixture()
    multimfit_fixture()
    copy_inputs()
    for fn in sorted(os.listdir(OUT)):
        if fn.endswith(".npz"):
            print("%-22s %8d bytes" % (fn, os.path.getsize(os.path.join(OUT, fn))))


# Input files of the reference's regression and unit tests, copied byte for byte so
# that the configuration-file and FITS readers run on the very files imfit read.
FITS_INPUTS = [
    # do_imfit_tests:63-99,107 and do_makeimage_tests:68-98 inputs
    "testimage_expdisk32.fits", "uniform_image32.fits",
    "ic3478rss_64x64.fits", "ic3478rss_64x64_sigma.fits", "ic3478rss_64x64_variance.fits",
    "n3073rss_small.fits", "n3073rss_small_mask.fits", "n3073rss_small_cps.fits", "psf_moffat_35_n4699z.fits",
    "twogaussian_psf+2osamp_noisy.fits", "psf_moffat_35.fits", "psf_moffat_35_oversamp3.fits",
    "oversamp_test4.fits", "psf_standard.fits", "psf_oversamp.fits",
    "testimage_3x3_nan.fits", "testimage_3x3_allzeros.fits",
    # do_imfit_tests:138-139,330 (data-based vs model-based weights)
    "flatsky_128x128.fits",
    # do_multimfit_tests.sh:59-85
    "multimfit_reference/smallDataImage1a.fits", "multimfit_reference/smallDataImage1b.fits",
    "multimfit_reference/smallgauss_5x5_0.fits", "multimfit_reference/smallgauss_5x5_1.fits",
    "multimfit_reference/smallgauss_5x5_sigma.fits",
    # do_multimfit_tests.sh:70 (two images with oversampled PSFs)
    "multimfit_reference/oversamp_test4b.fits",
    # do_imfit_tests:178 (PointSource), do_makeimage_tests:74,239 (oversampled PointSource)
    "psf_moffat_fwhm2.fits", "psf_moffat_fwhm2_35x35.fits", "oversampled_pointsource.fits",
    # unit_tests/unittest_psf_interpolators.t.h:20
    "../unit_tests/psf_gauss_sigma0.5_5.fits",
    # unit_tests/unittest_image_io.t.h:17-25
    "test_emptyhdu.fits", "test_table.fits", "test_multiextension_hdu0empty.fits",
    "test_multiextension_2images.fits", "multiextension_ic3478rss_64x64_twice.fits",
]
CONFIG_INPUTS = [
    # do_imfit_tests:63-107
    "imfit_reference/config_imfit_expdisk32.dat", "imfit_reference/imfit_config_ic3478_64x64.dat",
    "imfit_reference/imfit_config_ic3478_64x64b.dat", "imfit_reference/imfit_config_ic3478_64x64c.dat",
    "imfit_reference/imfit_config_n3073.dat", "imfit_reference/config_imfit_2gauss_small.dat",
    "imfit_reference/config_imfit_gauss-oversample-test2.dat", "imfit_reference/config_imfit_poisson.dat",
    "imfit_reference/config_imfit_badparamline.dat", "imfit_reference/config_imfit_flatsky.dat",
    # do_makeimage_tests:48-98
    "makeimage_reference/config_biggertest_4c.dat", "makeimage_reference/config_makeimage_gensersic512.dat",
    "makeimage_reference/config_makeimage_sersic+exp512.dat",
    "makeimage_reference/config_makeimage_sersic+exp512_nosize.dat",
    "makeimage_reference/config_makeimage_gauss-oversample.dat",
    "makeimage_reference/config_makeimage_gauss-oversample2.dat",
    "makeimage_reference/config_makeimage_sersictest512_bad1.dat",
    "makeimage_reference/config_makeimage_sersictest512_bad2.dat",
    "makeimage_reference/config_makeimage_sersictest512_bad3.dat",
    "makeimage_reference/config_makeimage_sersictest512_bad4.dat",
    "makeimage_reference/config_makeimage_sersictest512_bad5.dat",
    # do_multimfit_tests.sh:59-85
    "multimfit_reference/config_imfit_smalldataimages.dat",
    "multimfit_reference/config_imfit_smalldataimages_with-limits.dat",
    "multimfit_reference/config_imfit_smallgauss.dat", "multimfit_reference/imageinfo_multimfit0.txt",
    "multimfit_reference/imageinfo_multimfit0_with-limits.txt", "multimfit_reference/imageinfo_multimfit3.txt",
    "multimfit_reference/imageinfo_multimfit3b.txt", "multimfit_reference/imageinfo_multimfit3c.txt",
    "multimfit_reference/imageinfo_multimfit3_noerror.txt", "multimfit_reference/imageinfo_multimfit3d.txt",
    "multimfit_reference/imageinfo_multimfit+osample.txt",
    # do_imfit_tests:178, do_makeimage_tests:74 (PointSource)
    "imfit_reference/config_imfit_ptsource.dat", "makeimage_reference/config_makeimage_pointsource-oversample.dat",
    "makeimage_reference/config_imfit_pointsource-extra.dat",
    # unit_tests/unittest_config_parser.t.h:35-52
    "config_makeimage_sersictest512.dat", "config_makeimage_sersictest512_with_labels.dat",
    "config_imfit_sersictest512_badlimits4.dat",
    "config_makeimage-optional_params0.dat", "config_makeimage-optional_params1.dat",
    "config_makeimage-optional_params2.dat", "config_makeimage-optional_params3.dat",
    "config_makeimage-optional_params_bad1.dat",
]


def bootstrap_fixture():
    """do_imfit_tests:133-134,311-324: five bootstrap fits (seed 10) of the ic3478 fit with
    imfit_config_ic3478_64x64b.dat, on the whole image and on the section
    [10:64,10:64]; the archived tails list the fitted parameters, one fit per row."""
    def rows(path):
        with open(path) as f:
            return np.array([[float(v) for v in line.split()] for line in f
                             if line.strip() and not line.startswith("#")])
    lin = "imfit_reference/linux_x86_64/"
    np.savez_compressed(os.path.join(OUT, "bootstrap_ic3478.npz"),
                        seed10=rows(t(lin + "bootstrap_output_seed10_tail.dat")),
                        seed10_section=rows(t(lin + "bootstrap_output_seed10_2_tail.dat")))


def multimfit_fixture():
    """do_multimfit_tests.sh:59-95: multimfit printouts -- the statistic (1a: fit statistic
    only; 1d, 1e: L-M fits on two 2x2 images, 1e fitting FLUX_SCALE; 3a-3e: Nelder-Mead
    fits of a Gaussian on two 5x5 images: 3b, 3c on image sections, 3d without noise
    images, 3e with model errors; 2a: Nelder-Mead, Poisson MLR, two 200x200 images with
    oversampled PSFs; 3f: L-M, Poisson MLR, with 3g its five seed-10 bootstrap fits --
    3f and 3g are archived for macOS only, x86_64 taken, whose LM trajectory matches
    arm64's counts) and the best-fit main model and image-description parameters, in
    printed order; the L-M runs' NITER, NFEV and initial statistic."""
    def record(name):
        path = t("multimfit_reference/" + name)
        with open(path) as f:
            txt = f.read()
        if "POISSON-MLR STATISTIC" in txt:
            out = {"chi2": grab(path, r"POISSON-MLR STATISTIC = ([0-9.]+)")}
        else:
            out = {"chi2": grab(path, r"CHI-SQUARE = ([0-9.]+)")}
        if "NITER" in txt:
            out["niter"] = grab(path, r"NITER = (\d+)", int)
        if "NFEV" in txt:
            out["nfev"] = grab(path, r"NFEV = (\d+)", int)
        m = re.search(r"INITIAL (?:CHI\^2|POISSON-MLR STATISTIC) = ([0-9.]+)", txt)
        if m:
            out["initial"] = float(m.group(1))
        if "Statistics for parameter values from bootstrap" in txt:   # mean and std per printed parameter
            out["boot_mean_std"] = np.array([[float(a), float(b)] for a, b in
                                             re.findall(r"\(([-0-9.e+]+) \+/- ([-0-9.e+]+)\)", txt)])
        for key, head in (("model", "# Main model parameters"), ("image", "# Image-description parameters")):
            if head not in txt:
                continue
            vals = []
            for line in txt.split(head, 1)[1].split("\n")[1:]:
                line = line.strip()
                if not line or line.startswith("#") or line.startswith("Saving"):
                    break
                if line.startswith("FUNCTION"):
                    continue
                vals.append(float(line.split()[1]))
            out[key] = np.array(vals)
        return out
    recs = {}
    for k in ("1a", "1d", "1e", "3a", "3b", "3c", "3d", "3e", "2a"):
        for name, v in record("multimfit_textout" + k).items():
            recs["t%s_%s" % (k, name)] = v
    for k in ("3f", "3g"):
        for name, v in record("macos_x86_64/multimfit_textout" + k).items():
            recs["t%s_%s" % (k, name)] = v
    np.savez_compressed(os.path.join(OUT, "multimfit.npz"), **recs)


def copy_inputs():
    import shutil
    for sub, names in (("fits", FITS_INPUTS), ("configs", CONFIG_INPUTS)):
        dst = os.path.join(OUT, sub)
        os.makedirs(dst, exist_ok=True)
        for name in names:
            shutil.copyfile(t(name), os.path.join(dst, os.path.basename(name)))


if __name__ == "__main__":
    sys.exit(main())
