"""The mixed real space's FP32 erfc (kernels.h, pair32_fast): the fitted
coefficients in the source meet their stated bound against scipy's erfc,
and the form reproduces erfc's closed-form values.  CPU only."""
import os
import re

import numpy as np
from scipy.special import erfc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "molecular-dynamics-in-c_amd", "csrc", "kernels.h")


def _coefficients():
    """R(t)'s coefficients, highest degree first, as written in pair32_fast."""
    body = open(SRC).read()
    body = body[body.index("void pair32_fast"):]
    body = body[:body.index("const float g =")]
    first = re.search(r"float q = fmaf\(t, ([-0-9.e]+)f, ([-0-9.e]+)f\);", body)
    rest = re.findall(r"q = fmaf\(t, q, ([-0-9.e]+)f\);", body)
    return [float(first.group(1)), float(first.group(2))] + [float(x) for x in rest]


def _erfc_form(z, c, dtype):
    z = z.astype(dtype)
    t = (1 / (1 + dtype(0.5) * z)).astype(dtype)
    q = np.full_like(t, dtype(c[0]))
    for a in c[1:]:
        q = (q * t + dtype(a)).astype(dtype)
    return (t * np.exp(-(z * z).astype(dtype)).astype(dtype) * q).astype(dtype)


def test_mixed_erfc_fit_within_1e7():
    c = _coefficients()
    assert len(c) == 9   # degree 8
    z = np.linspace(0.0, 8.0, 400001)
    rel = np.abs(_erfc_form(z, c, np.float64) / erfc(z) - 1)
    assert rel.max() < 1e-7, rel.max()
    # a plausible slip (one coefficient's sign) fails the bound
    bad = list(c)
    bad[4] = -bad[4]
    assert np.abs(_erfc_form(z, bad, np.float64) / erfc(z) - 1).max() > 1e-3


def test_mixed_erfc_closed_forms_and_fp32():
    c = _coefficients()
    # erfc(0) = 1; erfc(1) = 0.157299207050285...
    assert abs(_erfc_form(np.array([0.0]), c, np.float64)[0] - 1.0) < 1e-7
    assert abs(_erfc_form(np.array([1.0]), c, np.float64)[0] / 0.15729920705028513 - 1) < 1e-7
    # in FP32 arithmetic, over the z of pairs at Ewald accuracies 1e-6 ... 1e-12
    z = np.linspace(0.2, 5.3, 100001)
    rel32 = np.abs(_erfc_form(z, c, np.float32).astype(np.float64) / erfc(z.astype(np.float32))
                   - 1)
    assert rel32.max() < 3e-6, rel32.max()
