"""The arcsine series behind the selection's chord-angle weights.

The fill kernel takes the central angle of an obs surely inside the support
as theta = 2 asin(|p - q| / 2), summed as asin(x) / x = sum_k c_k x^(2k) with
c_k = (2k)! / (4^k (k!)^2 (2k + 1)), 13 terms up to kChordMaxDeg and 8 terms
up to kChordShortMaxDeg.  This reads the coefficients and the two limits out
of the kernel source and checks, with exact rationals:

* every coefficient literal equals the rational its comment states, and that
  rational is the closed-form c_k of its position in the Horner chain;
* at each limit, the first term the chain leaves out is below 1e-18 of the
  sum, so the truncation is far below the 1e-12 weight tolerance the GPU
  tests hold the weights to (against the oracle's reference formula).
"""
import math
import os
import re
from fractions import Fraction

SRC = os.path.join(os.path.dirname(__file__), "..", "dida-thhpca-sc23_amd", "csrc",
                   "letkf_select.hip")


def _source():
    with open(SRC) as f:
        return f.read()


def _chord_body(src):
    start = src.index("__device__ __forceinline__ double chord_angle(")
    return src[start:src.index("\n}\n", start)]


def _limit(src, name):
    m = re.search(r"constexpr double %s = ([0-9.]+);" % name, src)
    assert m, name
    return float(m.group(1))


def _c(k):
    return Fraction(math.factorial(2 * k), 4 ** k * math.factorial(k) ** 2 * (2 * k + 1))


def _pairs(text):
    """(literal, stated rational) of every line whose comment states one."""
    out = []
    for line in text.splitlines():
        if "//" not in line:
            continue
        code, comment = line.split("//", 1)
        rat = re.findall(r"(\d+) / (\d+)", comment)
        if not rat:
            continue
        vals = re.findall(r"(\d+\.\d+(?:e-?\d+)?)", code)
        assert len(vals) == 1 and len(rat) == 1, line
        out.append((float(vals[0]), Fraction(int(rat[0][0]), int(rat[0][1]))))
    return out


def _chains(body):
    """The coefficients of the 13- and 8-term chains, highest power first."""
    i13 = body.index("if constexpr (kTerms == 13)")
    i8 = body.index("} else {", i13)
    iend = body.index("  }\n", i8)
    common = _pairs(body[iend:])
    return _pairs(body[i13:i8]) + common, _pairs(body[i8:iend]) + common


def test_coefficients_are_the_closed_form():
    c13, c8 = _chains(_chord_body(_source()))
    assert len(c13) == 12 and len(c8) == 7
    for chain in (c13, c8):
        n = len(chain)
        for i, (value, rat) in enumerate(chain):
            k = n - i  # highest power first, down to c_1
            assert rat == _c(k), (k, rat)
            assert value == float(rat), (k, value)


def test_truncation_below_1e_18_at_each_limit():
    src = _source()
    for name, terms in (("kChordMaxDeg", 13), ("kChordShortMaxDeg", 8)):
        theta = math.radians(_limit(src, name))
        x2 = Fraction(math.sin(theta / 2.0)) ** 2
        # first omitted term over the sum's lower bound 1
        assert float(_c(terms) * x2 ** terms) < 1e-18, name


def test_short_limit_within_long_limit():
    src = _source()
    assert _limit(src, "kChordShortMaxDeg") < _limit(src, "kChordMaxDeg")
