"""imfit / makeimage configuration files.

Follows core/config_file_parser.cpp (ReadConfigFile, VetConfigFile,
ParseFunctionSection, ValidParameterLine, AddParameterAndLimit,
AddFunctionNameAndLabel, GetFunctionLabel) and the string helpers it relies on
(core/utilities.cpp: SplitString 208-232, ChopComment 260-266, TrimWhitespace
273-284, IsNumeric 615-623), keeping their behaviour where it is unusual:

  - a line *containing* "X0" starts a function set, and a line containing
    "FUNCTION" declares a function (substring tests, core/config_file_parser.cpp:397,431);
  - a function label is read before the comment is chopped, so
    "FUNCTION Sersic # LABEL bulge" has the label "bulge" (:207-224, :232-246);
  - TrimWhitespace empties a line whose first and last non-blank characters
    coincide, so one-character lines are dropped (core/utilities.cpp:278-281);
  - numbers are read with C strtod rules (decimal, hex, inf, nan), and a token is
    numeric only if strtod consumes all of it;
  - in imfit mode (with limits) a limit pair must have lower < upper and bracket
    the initial value, else the file is rejected (:130-176).

The options before the first X0 line are kept as (name, value) pairs in file
order; `imfit_options` / `makeimage_options` apply them as
core/imfit_main.cpp:997-1063 and core/makeimage_main.cpp:679-722 do.
"""
import math
import re
import sys

import numpy as np

CONFIG_FILE_ERROR_NOFUNCSECTION = -1
CONFIG_FILE_ERROR_NOFUNCTIONS = -2
CONFIG_FILE_ERROR_INCOMPLETEXY = -3
CONFIG_FILE_ERROR_BADPARAMLINE = -4

OPTIONAL_PARAMS_START = "OPTIONAL_PARAMS_START"
OPTIONAL_PARAMS_END = "OPTIONAL_PARAMS_END"
FIXED = "fixed"
WHITESPACE = "\n\t "


class ConfigError(ValueError):
    def __init__(self, code, line, message):
        super().__init__(message)
        self.code, self.line = code, line


# ---------------------------------------------------------------- string helpers
def SplitString(s, delimiters=WHITESPACE):
    """Tokens of s separated by any of the delimiter characters.  A string without
    any delimiter comes back whole, even when empty."""
    if not any(c in delimiters for c in s):
        return [s]
    tokens, i, n = [], 0, len(s)
    while i < n:
        while i < n and s[i] in delimiters:
            i += 1
        if i >= n:
            break
        j = i
        while j < n and s[j] not in delimiters:
            j += 1
        tokens.append(s[i:j])
        i = j
    return tokens


def ChopComment(s, delimiter="#"):
    k = s.find(delimiter)
    return s if k < 0 else s[:k]


def TrimWhitespace(s):
    if not s:
        return s
    blank = " \t\n"
    start = next((k for k, c in enumerate(s) if c not in blank), -1)
    end = next((k for k in range(len(s) - 1, -1, -1) if s[k] not in blank), -1)
    if start == end:
        return ""
    return s[start:end + 1]


_DEC = r"(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?"
_HEX = r"0[xX](?:[0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)(?:[pP][+-]?\d+)?"
_SPECIAL = r"(?:[iI][nN][fF](?:[iI][nN][iI][tT][yY])?|[nN][aA][nN](?:\([0-9A-Za-z_]*\))?)"
_STRTOD = re.compile(r"[ \t\n\v\f\r]*([+-]?)(%s|%s|%s)" % (_HEX, _DEC, _SPECIAL))


def strtod(s):
    """(value, characters consumed) of C's strtod on s; (0.0, 0) when nothing parses."""
    m = _STRTOD.match(s)
    if m is None:
        return 0.0, 0
    sign, body = m.group(1), m.group(2)
    low = body.lower()
    if low.startswith("0x"):
        v = float.fromhex(body)
    elif low.startswith("inf"):
        v = math.inf
    elif low.startswith("nan"):
        v = math.nan
    else:
        v = float(body)
    return (-v if sign == "-" else v), m.end()


def IsNumeric(s):
    if not s or s[0].isspace():
        return False
    _, used = strtod(s)
    return used == len(s) and used > 0


def atoi(s):
    m = re.match(r"[ \t\n\v\f\r]*([+-]?\d+)", s)
    return int(m.group(1)) if m else 0


# ---------------------------------------------------------------- parameter limits
class ParamLimit:
    """mpfit's per-parameter settings as a config file gives them
    (core/param_struct.h:37-43): fixed, limited[2], limits[2], plus the image
    offset imfit stores for X0/Y0."""

    def __init__(self):
        self.fixed = 0
        self.limited = [0, 0]
        self.limits = [0.0, 0.0]
        self.offset = 0

    def __repr__(self):
        return "ParamLimit(fixed=%d, limited=%s, limits=%s, offset=%d)" % (
            self.fixed, self.limited, self.limits, self.offset)


# ---------------------------------------------------------------- line tests
def ValidParameterLine(line, optionalParams=False):
    """True if the line has a name and a numeric value, optionally followed by "fixed" or
    a "lower,upper" pair of numbers; in optional-parameter mode two tokens suffice
    (core/config_file_parser.cpp:66-97)."""
    pieces = SplitString(ChopComment(line))
    if len(pieces) < 2:
        return False
    if optionalParams:
        return True
    if not IsNumeric(pieces[1]):
        return False
    if len(pieces) > 2 and pieces[2] != FIXED:
        lim = SplitString(pieces[2], ",")
        if len(lim) != 2:
            return False
        if not IsNumeric(lim[0]) or not IsNumeric(lim[1]):
            return False
    return True


def GetFunctionLabel(line):
    """Text after "LABEL" on a FUNCTION line, trimmed; "" without one (:212-226)."""
    k = line.find("LABEL")
    if k < 0:
        return ""
    return TrimWhitespace(line[k + 5:])


def AddFunctionNameAndLabel(line, functionNameList, functionLabelList):
    functionLabelList.append(GetFunctionLabel(line))
    functionNameList.append(SplitString(ChopComment(line))[1])


def _add_parameter_and_limit(line, parameterList, parameterLimits, origLineNumber):
    """Value and limits of a parameter line: returns 1 if limits or "fixed" were given,
    0 if not, -1 for inconsistent limits (:130-176)."""
    pieces = SplitString(ChopComment(line))
    value = strtod(pieces[1])[0]
    parameterList.append(value)
    lim = ParamLimit()
    found = 0
    if len(pieces) > 2:
        found = 1
        extra = pieces[2]
        if extra == FIXED:
            lim.fixed = 1
        elif "," in extra:
            lo_hi = SplitString(extra, ",")
            lim.limited = [1, 1]
            lo, hi = strtod(lo_hi[0])[0], strtod(lo_hi[1])[0]
            if lo >= hi:
                print('*** WARNING: first parameter limit for "%s" (%g) must be < second limit (%g)!\n'
                      "    (Error on input line %d of configuration file)" % (pieces[0], lo, hi, origLineNumber),
                      file=sys.stderr)
                return -1
            if value < lo or value > hi:
                print('*** WARNING: initial parameter value for "%s" (%g) must lie between the limits (%g,%g)!\n'
                      "    (Error on input line %d of configuration file)" % (pieces[0], value, lo, hi,
                                                                             origLineNumber), file=sys.stderr)
                return -1
            lim.limits = [lo, hi]
    parameterLimits.append(lim)
    return found


def VetConfigFile(inputLines, origLineNumbers, mode2D):
    """(index of the first X0 line, bad line number or -1), or (negative error code,
    bad line number) (core/config_file_parser.cpp:264-340).  Like the reference,
    which checks each line through a reference to it, every parameter line it
    checks loses its comment in `inputLines` (FUNCTION lines keep theirs, and with
    them their labels)."""
    bad = -1
    start = -1
    y_ok = True
    for i, line in enumerate(inputLines):
        if "X0" in line:
            start = i
            if mode2D and (i + 1 == len(inputLines) or "Y0" not in inputLines[i + 1]):
                y_ok = False
                bad = origLineNumbers[i]
            break
    if start < 0:
        return CONFIG_FILE_ERROR_NOFUNCSECTION, bad
    if not y_ok:
        return CONFIG_FILE_ERROR_INCOMPLETEXY, bad
    if not any("FUNCTION" in line for line in inputLines[start:]):
        return CONFIG_FILE_ERROR_NOFUNCTIONS, bad
    in_optional = False
    for i in range(start, len(inputLines)):
        line = inputLines[i]
        if "FUNCTION" in line:
            continue
        if OPTIONAL_PARAMS_START in line:
            in_optional = True
            continue
        if OPTIONAL_PARAMS_END in line:
            in_optional = False
            continue
        inputLines[i] = ChopComment(line)
        if not ValidParameterLine(line, in_optional):
            return CONFIG_FILE_ERROR_BADPARAMLINE, origLineNumbers[i]
    return start, bad


class FunctionSection:
    """What ParseFunctionSection extracts: function names and labels, parameter values
    and limits, the index of the first function of each function set, the optional
    parameters of each function, and which parameters are X0 / Y0 of a set."""

    def __init__(self):
        self.function_names, self.function_labels = [], []
        self.parameters, self.limits = [], []
        self.fset_starts, self.optional_params = [], []
        self.limits_found = False
        self.xy_index = []            # (index of X0, index of Y0 or -1) per function set


def ParseFunctionSection(inputLines, mode2D, origLineNumbers, checkParameterLimits=False):
    """FunctionSection of the lines from the first X0 line on, or raises ConfigError
    where the reference returns -1 (core/config_file_parser.cpp:379-472)."""
    fs = FunctionSection()
    i, nfunc = 0, 0
    in_optional = False

    def add(line, lineno):
        found = _add_parameter_and_limit(line, fs.parameters, fs.limits, lineno)
        if checkParameterLimits and found < 0:
            raise ConfigError(CONFIG_FILE_ERROR_BADPARAMLINE, lineno, "bad parameter limits on line %d" % lineno)
        if found == 1:
            fs.limits_found = True

    while i < len(inputLines):
        line = inputLines[i]
        if "X0" in line:
            fs.fset_starts.append(nfunc)
            ix = len(fs.parameters)
            add(line, origLineNumbers[i])
            i += 1
            iy = -1
            if mode2D:
                if i >= len(inputLines) or "Y0" not in inputLines[i]:
                    lineno = origLineNumbers[i - 1]
                    raise ConfigError(CONFIG_FILE_ERROR_INCOMPLETEXY, lineno,
                                      "a 'Y0' line must follow each 'X0' line (X0 on input line %d)" % lineno)
                iy = len(fs.parameters)
                add(inputLines[i], origLineNumbers[i])
                i += 1
            fs.xy_index.append((ix, iy))
            continue
        if "FUNCTION" in line:
            AddFunctionNameAndLabel(line, fs.function_names, fs.function_labels)
            fs.optional_params.append({})
            nfunc += 1
            i += 1
            continue
        if OPTIONAL_PARAMS_START in line:
            in_optional = True
        elif OPTIONAL_PARAMS_END in line:
            in_optional = False
        elif in_optional:
            pieces = SplitString(ChopComment(line))
            if nfunc == 0:
                raise ConfigError(CONFIG_FILE_ERROR_BADPARAMLINE, origLineNumbers[i],
                                  "optional parameters before any FUNCTION line (line %d)" % origLineNumbers[i])
            fs.optional_params[nfunc - 1][pieces[0]] = pieces[1]
        else:
            add(line, origLineNumbers[i])
        i += 1
    return fs


_MESSAGES = {
    CONFIG_FILE_ERROR_NOFUNCSECTION: 'unable to find start of function section (no "X0" line found)',
    CONFIG_FILE_ERROR_NOFUNCTIONS: "no actual functions found in configuration file",
    CONFIG_FILE_ERROR_INCOMPLETEXY: "initial Y0 value (start of function section) missing (X0 on line %d)",
    CONFIG_FILE_ERROR_BADPARAMLINE: "bad parameter specification at line %d",
}


class ConfigFile:
    """A parsed configuration file.  Attributes follow ReadConfigFile's outputs:
    function_names, function_labels, parameters (float64 array), limits
    (ParamLimit per parameter; imfit mode only), fset_starts, options ([(name, value)]),
    optional_params (dict per function), limits_found, and xy_index (positions of
    X0 / Y0 in the parameter vector per function set)."""

    def __init__(self, fs, options, with_limits):
        self.function_names = fs.function_names
        self.function_labels = fs.function_labels
        self.parameters = np.array(fs.parameters, dtype=np.float64)
        self.limits = fs.limits if with_limits else None
        self.limits_found = fs.limits_found if with_limits else False
        self.fset_starts = fs.fset_starts
        self.optional_params = fs.optional_params
        self.options = options
        self.xy_index = fs.xy_index

    @property
    def nOptions(self):
        return len(self.options)

    def functions(self):
        """[(name, starts a new function set)] in file order."""
        starts = set(self.fset_starts)
        return [(name, k in starts) for k, name in enumerate(self.function_names)]

    def fixed(self):
        if self.limits is None:
            return np.zeros(len(self.parameters), dtype=bool)
        return np.array([lim.fixed == 1 for lim in self.limits], dtype=bool)

    def limit_pairs(self):
        """[(lower, upper) or None] per parameter, for LM fits."""
        if self.limits is None:
            return [None] * len(self.parameters)
        return [(lim.limits[0], lim.limits[1]) if lim.limited[0] and lim.limited[1] else None
                for lim in self.limits]

    def option(self, name):
        """Value text of the first option with this name, or None."""
        for k, v in self.options:
            if k == name:
                return v
        return None

    def imfit_options(self, gain=None, readnoise=None, exptime=None, ncombined=None, original_sky=None):
        """Image characteristics as imfit resolves them: a value given here (the command
        line) wins over the file's; GAIN, READNOISE, EXPTIME and ORIGINAL_SKY are read
        with strtod, NCOMBINED with atoi; other keywords are ignored
        (core/imfit_main.cpp:997-1063).  Defaults: gain 1, read noise 0, exposure
        time 1, 1 combined image, original sky 0."""
        out = {"gain": 1.0, "readnoise": 0.0, "exptime": 1.0, "ncombined": 1, "original_sky": 0.0}
        given = {"gain": gain, "readnoise": readnoise, "exptime": exptime, "ncombined": ncombined,
                 "original_sky": original_sky}
        keys = {"GAIN": "gain", "READNOISE": "readnoise", "EXPTIME": "exptime", "ORIGINAL_SKY": "original_sky",
                "NCOMBINED": "ncombined"}
        for name, value in self.options:
            key = keys.get(name)
            if key is None or given[key] is not None:
                continue
            out[key] = atoi(value) if key == "ncombined" else strtod(value)[0]
        for key, v in given.items():
            if v is not None:
                out[key] = v
        return out

    def makeimage_options(self, ncols=None, nrows=None):
        """(nColumns, nRows) as makeimage resolves them: NCOLS or NCOLUMNS and NROWS,
        which must be positive integers, unless given here (core/makeimage_main.cpp:679-722)."""
        out = {"ncols": ncols, "nrows": nrows}
        for name, value in self.options:
            key = "ncols" if name in ("NCOLS", "NCOLUMNS") else "nrows" if name == "NROWS" else None
            if key is None or (ncols if key == "ncols" else nrows) is not None:
                continue
            if not value or not value[0].isdigit() or value[0] == "0":
                raise ConfigError(0, -1, "%s should be a positive integer" % name)
            out[key] = atoi(value)
        return out["ncols"], out["nrows"]


def _lines(source):
    if "\n" not in source:
        try:
            with open(source) as f:
                return f.read().split("\n")
        except OSError as e:
            raise ConfigError(0, -1, "unable to open configuration file %r: %s" % (source, e))
    return source.split("\n")


def ReadConfigFile(source, mode2D=True, with_limits=True):
    """Parse a configuration file (a path, or the text itself when it has several
    lines).  with_limits=True is imfit's reading (limits parsed and checked),
    False makeimage's (core/config_file_parser.cpp:488-654).  Raises ConfigError
    with the reference's error code and line number."""
    raw = _lines(source)
    if raw and raw[-1] == "":
        raw = raw[:-1]           # getline sees no line after a final newline
    lines, numbers = [], []
    for k, line in enumerate(raw, start=1):
        line = TrimWhitespace(line)
        if line and line[0] != "#":
            lines.append(line)
            numbers.append(k)
    start, bad = VetConfigFile(lines, numbers, mode2D)
    if start < 0:
        msg = _MESSAGES[start]
        raise ConfigError(start, bad, msg % bad if "%d" in msg else msg)
    options = []
    for line in lines[:start]:
        pieces = SplitString(ChopComment(line))
        if len(pieces) == 2:
            options.append((pieces[0], pieces[1]))
    fs = ParseFunctionSection(lines[start:], mode2D, numbers[start:], checkParameterLimits=with_limits)
    return ConfigFile(fs, options, with_limits)
