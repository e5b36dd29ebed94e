"""Configuration files and FITS images, on the CPU.

Known answers come from the reference's own unit tests
(unit_tests/unittest_config_parser.t.h, unit_tests/unittest_image_io.t.h) and from
its makeimage error printouts (tests/makeimage_reference/makeimage_textout6..15,
do_makeimage_tests:78-95); the files are the reference's, copied byte for byte into
tests/golden/fits and tests/golden/configs.
"""
import os

import numpy as np
import pytest

import cases
from imfit_amd import config as C
from imfit_amd import fits as F
from imfit_amd.imfit_setup import SetupError, makeimage_setup

FITS_DIR = os.path.join(cases.GOLDEN, "fits")
CFG_DIR = os.path.join(cases.GOLDEN, "configs")


def fits(name):
    return os.path.join(FITS_DIR, name)


def cfg(name):
    return os.path.join(CFG_DIR, name)


def vetted_lines(path):
    """Non-empty, non-comment trimmed lines with their line numbers, as the unit
    tests prepare them (unittest_config_parser.t.h:221-235)."""
    lines, numbers = [], []
    with open(path) as f:
        text = f.read().split("\n")
    if text and text[-1] == "":
        text = text[:-1]
    for k, line in enumerate(text, start=1):
        line = C.TrimWhitespace(line)
        if line and line[0] != "#":
            lines.append(line)
            numbers.append(k)
    return lines, numbers


# ---------------------------------------------------------------- parser pieces
def test_valid_parameter_lines():          # unittest_config_parser.t.h:61-126
    for line in ["X0   32.0    1,64\n", "X0   32.0\n", "X0   32.0   # this is a comment\n",
                 "X0   32.0   30,35  # this is a comment\n", "X0   32.0   fixed  # this is a comment\n"]:
        assert C.ValidParameterLine(line), line
    for line in ["X0\n", "X0\t\n", "X0   \n", "X0   fixed\n", "X0   33,36\n", "X0   32.0   30\n",
                 "X0   32.0   30,  \n", "X0   32.0   x\n", "X0   32.0   30,x\n"]:
        assert not C.ValidParameterLine(line), line
    assert C.ValidParameterLine("mode  alpha\n", True)
    assert C.ValidParameterLine("mode  alpha   # this is a comment\n", True)
    assert not C.ValidParameterLine("mode\n", True)
    assert not C.ValidParameterLine("mode  # this is a comment\n", True)


def test_function_labels():                # unittest_config_parser.t.h:131-204
    assert C.GetFunctionLabel("FUNCTION Sersic") == ""
    assert C.GetFunctionLabel("FUNCTION Sersic   # this is not a label") == ""
    assert C.GetFunctionLabel("FUNCTION Sersic # LABEL inner bulge") == "inner bulge"
    assert C.GetFunctionLabel("FUNCTION Sersic  LABEL inner bulge") == "inner bulge"
    for line, label in [("FUNCTION Sersic", ""), ("FUNCTION Sersic   # this is not a label", ""),
                        ("FUNCTION Sersic # LABEL inner bulge", "inner bulge"),
                        ("FUNCTION Sersic  LABEL inner bulge", "inner bulge")]:
        names, labels = [], []
        C.AddFunctionNameAndLabel(line, names, labels)
        assert names == ["Sersic"] and labels == [label]


def test_string_helpers_follow_c_library_rules():
    assert C.SplitString("") == [""] and C.SplitString("X0") == ["X0"]
    assert C.SplitString("  a \t b\n") == ["a", "b"]
    assert C.SplitString("30,", ",") == ["30"]
    assert C.TrimWhitespace("  x  ") == ""            # one non-blank character: emptied
    assert C.TrimWhitespace("  xy ") == "xy"
    assert C.IsNumeric("1e5") and C.IsNumeric("-.5") and C.IsNumeric("0x1p3") and C.IsNumeric("inf")
    assert not C.IsNumeric(" 1") and not C.IsNumeric("1e") and not C.IsNumeric("1_0") and not C.IsNumeric("")
    assert C.strtod("0x1.8p1")[0] == 3.0 and C.strtod("4.5abc") == (4.5, 3) and C.strtod("abc") == (0.0, 0)
    assert C.atoi(" 12x") == 12 and C.atoi("x") == 0


@pytest.mark.parametrize("name,mode2d", [
    ("config_makeimage_sersictest512.dat", False), ("config_makeimage_sersictest512.dat", True),
    ("config_makeimage_sersictest512_with_labels.dat", False),
    ("config_makeimage_sersictest512_with_labels.dat", True),
    ("config_makeimage-optional_params0.dat", True), ("config_makeimage-optional_params1.dat", True)])
def test_vet_good_files(name, mode2d):     # unittest_config_parser.t.h:211-345
    lines, numbers = vetted_lines(cfg(name))
    assert C.VetConfigFile(lines, numbers, mode2d)[0] == 2


@pytest.mark.parametrize("name,code,line", [
    ("config_makeimage_sersictest512_bad1.dat", C.CONFIG_FILE_ERROR_NOFUNCSECTION, None),
    ("config_makeimage_sersictest512_bad2.dat", C.CONFIG_FILE_ERROR_NOFUNCTIONS, None),
    ("config_makeimage_sersictest512_bad3.dat", C.CONFIG_FILE_ERROR_INCOMPLETEXY, 8),
    ("config_imfit_badparamline.dat", C.CONFIG_FILE_ERROR_BADPARAMLINE, 12),
    ("config_makeimage-optional_params_bad1.dat", C.CONFIG_FILE_ERROR_BADPARAMLINE, 13),
    ("config_imfit_sersictest512_badlimits4.dat", C.CONFIG_FILE_ERROR_BADPARAMLINE, 9)])
def test_vet_bad_files(name, code, line):  # unittest_config_parser.t.h:349-567
    lines, numbers = vetted_lines(cfg(name))
    got, bad = C.VetConfigFile(lines, numbers, True)
    assert got == code
    if line is not None:
        assert bad == line
    with pytest.raises(C.ConfigError) as e:   # and through ReadConfigFile
        C.ReadConfigFile(cfg(name))
    assert e.value.code == code


def test_parse_function_section():         # unittest_config_parser.t.h:570-666
    lines = ["X0    200.0\n", "Y0    200.0\n", "FUNCTION   Gaussian\n", "PA    10.0\n", "ell    0.1\n",
             "I_0    1000.0\n", "sigma      5.0\n", "X0    256.0\n", "Y0    256.0\n",
             "FUNCTION   Sersic_GenEllipse\n", "PA    30.0\n", "ell    0.5\n", "c0    -1.0\n", "n      2.5\n",
             "I_e  100.0\n", "r_e   50.0\n", "FUNCTION   FlatSky\n", "I_0  1.0\n"]
    fs = C.ParseFunctionSection(lines, True, list(range(1, 19)))
    assert fs.function_names == ["Gaussian", "Sersic_GenEllipse", "FlatSky"]
    assert fs.parameters == [200.0, 200.0, 10.0, 0.1, 1000.0, 5.0, 256.0, 256.0, 30.0, 0.5, -1.0, 2.5, 100.0,
                             50.0, 1.0]
    assert fs.fset_starts == [0, 1] and fs.limits_found is False
    assert fs.xy_index == [(0, 1), (6, 7)]
    with pytest.raises(C.ConfigError):       # missing Y0
        C.ParseFunctionSection(["X0    200.0\n", "FUNCTION   Gaussian\n", "PA    10.0\n", "ell    0.1\n",
                                "I_0    1000.0\n", "sigma      5.0\n"], True, list(range(1, 7)))
    with pytest.raises(C.ConfigError):       # lower limit above upper
        C.ParseFunctionSection(["X0    200.0\n", "Y0    200.0\n", "FUNCTION   Gaussian\n", "PA    10.0     20.0,1.0\n",
                                "ell    0.1\n", "I_0    1000.0\n", "sigma      5.0\n"], True, list(range(1, 8)), True)


GENSERSIC = [256.0, 256.0, 30.0, 0.5, -1.0, 2.5, 100.0, 50.0]


@pytest.mark.parametrize("with_limits", [True, False])
@pytest.mark.parametrize("name,label", [("config_makeimage_sersictest512.dat", ""),
                                        ("config_makeimage_sersictest512_with_labels.dat", "main component")])
def test_read_config_file(name, label, with_limits):   # unittest_config_parser.t.h:681-960
    c = C.ReadConfigFile(cfg(name), True, with_limits)
    assert c.function_names == ["Sersic_GenEllipse"] and c.function_labels == [label]
    assert list(c.parameters) == GENSERSIC
    assert c.fset_starts == [0] and c.limits_found is False
    assert c.nOptions == 2 and c.options == [("NCOLS", "512"), ("NROWS", "512")]
    assert c.optional_params == [{}]
    assert c.makeimage_options() == (512, 512)


def test_optional_parameters():            # unittest_config_parser.t.h:964-1166
    c = C.ReadConfigFile(cfg("config_makeimage-optional_params1.dat"), True, False)
    assert list(c.parameters) == GENSERSIC and c.optional_params == [{"mode": "alpha"}]
    c = C.ReadConfigFile(cfg("config_makeimage-optional_params2.dat"), True, False)
    assert list(c.parameters) == GENSERSIC and c.optional_params == [{"mode": "alpha", "degree": "12"}]
    c = C.ReadConfigFile(cfg("config_makeimage-optional_params3.dat"), True, False)
    assert c.function_names[0] == "Sersic_GenEllipse" and len(c.function_names) == 3
    assert list(c.parameters) == GENSERSIC + [20.0, 0.0, 200.0, 35.0, 30.0, 0.5, 100.0, 10.0]
    assert c.fset_starts == [0]
    assert c.optional_params == [{"mode": "alpha", "degree": "12"}, {}, {"floor": "1.5"}]


def test_imfit_config_limits_and_options():
    """tests/imfit_reference/imfit_config_n3073.dat: limits on every parameter, image
    characteristics in the file (imfit_textout4c: gain 4.725, read noise 4.3,
    original sky 154.33); values given as arguments win (imfit_main.cpp:997-1063)."""
    c = C.ReadConfigFile(cfg("imfit_config_n3073.dat"))
    assert c.function_names == ["Sersic", "Exponential"] and c.fset_starts == [0]
    assert list(c.parameters) == [101.0, 97.0, 110.0, 0.2, 1.5, 150, 15, 55.0, 0.2, 70.0, 45.0]
    assert c.limit_pairs() == cases.LIMITS["n3073"] and c.limits_found
    assert not c.fixed().any()
    o = c.imfit_options()
    assert (o["gain"], o["readnoise"], o["original_sky"], o["exptime"], o["ncombined"]) == (4.725, 4.3, 154.33, 1.0,
                                                                                          1)
    o = c.imfit_options(exptime=53.907456, original_sky=2.8628692847238053)   # do_imfit_tests:96
    assert o["exptime"] == 53.907456 and o["original_sky"] == 2.8628692847238053 and o["gain"] == 4.725
    c = C.ReadConfigFile(cfg("imfit_config_ic3478_64x64c.dat"))
    assert list(c.fixed()) == [bool(v) for v in cases.FIXED["ic3478_c"]]
    assert c.limit_pairs() == cases.LIMITS["ic3478_c"]
    c = C.ReadConfigFile(cfg("config_imfit_2gauss_small.dat"))
    assert list(c.fixed()) == [bool(v) for v in cases.FIXED["twogauss"]]
    assert c.limit_pairs() == cases.LIMITS["twogauss"]


def test_comment_mentioning_x0_does_not_start_a_function_set():
    """Vetting chops the comments off parameter lines before the X0 scan of the
    function section sees them (core/config_file_parser.cpp:73,306-318)."""
    text = "GAIN 2\nX0 10 # x\nY0 12\nFUNCTION Gaussian # LABEL core\nPA 0 # along X0\nell 0\nI_0 1\nsigma 2\n"
    c = C.ReadConfigFile(text)
    assert c.fset_starts == [0] and c.function_labels == ["core"] and len(c.parameters) == 6
    assert c.xy_index == [(0, 1)] and c.options == [("GAIN", "2")]


# ---------------------------------------------------------------- makeimage error printouts
@pytest.mark.parametrize("name,code", [
    ("config_makeimage_sersictest512_bad1.dat", C.CONFIG_FILE_ERROR_NOFUNCSECTION),   # makeimage_textout6
    ("config_makeimage_sersictest512_bad2.dat", C.CONFIG_FILE_ERROR_NOFUNCTIONS),     # makeimage_textout7
    ("config_makeimage_sersictest512_bad3.dat", C.CONFIG_FILE_ERROR_INCOMPLETEXY),    # makeimage_textout8
    ("config_makeimage_sersictest512_bad4.dat", C.CONFIG_FILE_ERROR_INCOMPLETEXY),    # makeimage_textout9
    ("config_imfit_badparamline.dat", C.CONFIG_FILE_ERROR_BADPARAMLINE)])              # makeimage_textout15
def test_makeimage_rejects_bad_config_files(name, code):
    with pytest.raises(C.ConfigError) as e:
        makeimage_setup(cfg(name))
    assert e.value.code == code
    if name.endswith("bad4.dat"):            # "X0 specification on input line 20"
        assert e.value.line == 20
    if name.endswith("badparamline.dat"):    # "Bad parameter specification at line 12"
        assert e.value.line == 12


def test_makeimage_rejects_missing_size_and_unreadable_images():
    with pytest.raises(SetupError, match="insufficient image dimensions"):       # makeimage_textout11
        makeimage_setup(cfg("config_makeimage_sersic+exp512_nosize.dat"))
    with pytest.raises(SetupError, match="unable to read reference"):            # makeimage_textout12
        makeimage_setup(cfg("config_makeimage_sersic+exp512_nosize.dat"),
                        refimage=cfg("config_makeimage_gensersic512.dat"))
    with pytest.raises(SetupError, match="unable to read PSF"):                  # makeimage_textout13
        makeimage_setup(cfg("config_makeimage_sersic+exp512.dat"), psf=cfg("config_makeimage_gensersic512.dat"))
    with pytest.raises(SetupError, match="unable to read PSF"):                  # makeimage_textout14
        makeimage_setup(cfg("config_makeimage_sersic+exp512.dat"), psf="_nofilewiththisname_wehope.fits")


# ---------------------------------------------------------------- FITS
def test_read_image_known_pixels():        # unittest_image_io.t.h:203-216
    v, nc, nr = F.ReadImageAsVector(fits("testimage_expdisk32.fits"))
    assert (nc, nr) == (32, 32)
    assert abs(v[0] - 20.903625) < 1e-6 and abs(v[32 * 32 - 1] - 14.947027) < 1e-6
    assert F.ReadImageAsVector("no_image_with_this_name.fits")[0] is None


def test_check_for_image_and_sizes():      # unittest_image_io.t.h:64-200,413-439
    assert F.CheckForImage(fits("testimage_expdisk32.fits")) == 1
    assert F.CheckForImage("no_image_with_this_name.fits") == -1
    assert F.CheckForImage(fits("test_emptyhdu.fits")) == -1
    assert F.CheckForImage(fits("test_table.fits")) == -1
    assert F.CheckForImage(fits("test_table.fits") + "[1]") == -1
    assert F.CheckForImage(fits("test_multiextension_hdu0empty.fits")) == -1
    assert F.CheckForImage(fits("test_multiextension_hdu0empty.fits") + "[1]") == 1
    assert F.CheckForImage(fits("test_multiextension_2images.fits")) == 1
    assert F.CheckForImage(fits("test_multiextension_2images.fits") + "[0]") == 1
    assert F.CheckForImage(fits("test_multiextension_2images.fits") + "[1]") == 1
    assert F.GetImageSize(fits("testimage_expdisk32.fits")) == (32, 32, 0)
    assert F.GetImageSize("no_image_with_this_name.fits")[2] == -1
    assert F.GetImageSize(fits("test_emptyhdu.fits"))[2] == -1
    assert F.GetImageSize(fits("test_multiextension_hdu0empty.fits"))[2] == -1
    assert F.GetImageSize(fits("test_table.fits"))[2] == -1
    assert F.GetImageSize(fits("test_multiextension_2images.fits") + "[0]")[:2] == (5, 5)
    assert F.GetImageSize(fits("test_multiextension_2images.fits") + "[1]")[:2] == (10, 10)
    twice = fits("multiextension_ic3478rss_64x64_twice.fits")
    assert F.GetImageSize(twice + "[0]")[:2] == (64, 64) and F.GetImageSize(twice + "[1]")[:2] == (64, 64)
    assert np.array_equal(F.read_image(twice + "[0]"), F.read_image(fits("ic3478rss_64x64.fits")))
    assert not np.array_equal(F.read_image(twice + "[1]"), F.read_image(twice + "[0]"))


def test_fits_readers_agree_with_extracted_fixtures():
    """The package reader and the fixture extractor's reader give the same pixels
    (two readers written separately; BITPIX -32 data, a 0/1 mask, and PSFs)."""
    d = cases.golden("n3073")
    assert np.array_equal(F.read_image(fits("n3073rss_small.fits")), d["data"])
    assert np.array_equal(F.read_image(fits("n3073rss_small_mask.fits")), d["mask"])
    assert np.array_equal(F.read_image(fits("psf_moffat_35_n4699z.fits")), d["psf"])
    e = cases.golden("ic3478_64_errors")
    assert np.array_equal(F.read_image(fits("ic3478rss_64x64_variance.fits")), e["variance"])


def test_sections_and_offsets():
    img = F.read_image(fits("oversamp_test4.fits"))
    sec = F.read_image(fits("oversamp_test4.fits") + "[51:200,41:190]")
    assert sec.shape == (150, 150) and np.array_equal(sec, img[40:190, 50:200])
    assert F.DetermineImageOffset("oversamp_test4.fits[51:200,41:190]") == (50, 40)
    assert F.DetermineImageOffset("oversamp_test4.fits") == (0, 0)
    assert F.DetermineImageOffset("x.fits[*,41:190]") == (0, 40)
    assert np.array_equal(F.read_image(fits("oversamp_test4.fits") + "[*,41:190]"), img[40:190])
    assert F.GetImageSize(fits("oversamp_test4.fits") + "[51:200,41:190]")[:2] == (150, 150)
    with pytest.raises(F.FitsError):
        F.read_image(fits("oversamp_test4.fits") + "[51:900,41:190]")


def test_write_and_read_back(tmp_path):    # unittest_image_io.t.h:232-305
    tiny = np.arange(1.0, 10.0)
    p = str(tmp_path / "tiny.fits")
    assert F.SaveVectorAsImage(tiny, p, 3, 3, ["Here is a comment", "this is a second comment"]) == 0
    v, nc, nr = F.ReadImageAsVector(p)
    assert (nc, nr) == (3, 3) and abs(v[0] - 1.0) < 1e-6 and abs(v[8] - 9.0) < 1e-6
    raw = open(p, "rb").read()
    assert len(raw) % 2880 == 0 and b"COMMENT Here is a comment" in raw and b"DATE    = '" in raw
    assert b"BITPIX  =                  -32" in raw
    assert F.SaveVectorAsImage(tiny, "/nopath/nohow/ZZZ/arglebargle.fits", 3, 3) == -1
    assert F.SaveVectorAsImage(None, str(tmp_path / "null.fits"), 3, 3) == -2
    rng = np.random.default_rng(0)
    a = rng.normal(size=(7, 5))
    F.write_image(a, p, bitpix=-64)
    assert np.array_equal(F.read_image(p), a)
    F.write_image(a, p)
    assert np.array_equal(F.read_image(p), a.astype(np.float32).astype(np.float64))


def test_image_info_files():
    """tests/multimfit_reference/imageinfo_multimfit3b.txt: two images (sections of the
    5x5 files), sky 100, the second image's scale, angle and flux scale fixed."""
    from imfit_amd.multimage import read_image_info
    infos = read_image_info(cfg("imageinfo_multimfit3b.txt"), lambda n: n)
    assert len(infos) == 2
    assert infos[0].data.endswith("smallgauss_5x5_0.fits[2:5,2:5]") and infos[0].original_sky == 100.0
    assert infos[1].params == [1.0, 0.0, 1.0, 2.0, 2.0]
    assert [l.fixed for l in infos[1].limits] == [1, 1, 1, 0, 0]
    info0 = read_image_info(cfg("imageinfo_multimfit0_with-limits.txt"), lambda n: n)
    assert info0[1].limits[2].limited == [1, 1] and info0[1].limits[2].limits == [0.0, 5.0]
