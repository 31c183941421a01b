"""The quadrature tables (oracle and HIP copies) against their closed forms."""
import os
import re
from math import factorial

import numpy as np
from scipy.special import roots_jacobi, roots_legendre

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def table(path, name):
    txt = open(os.path.join(ROOT, path)).read()
    m = re.search(name + r"\s*\[[^=]*=\s*\{(.*?)\};", txt, re.S)
    nums = re.findall(r"[-+]?\d+\.\d+E[-+]\d+", m.group(1))
    return np.array([float(v) for v in nums])


def tables(path, prefix):
    return (table(path, prefix + "GL8").reshape(8, 2), table(path, prefix + "GJ8").reshape(8, 2),
            table(path, prefix + "T7").reshape(7, 3))


def test_oracle_and_hip_tables_identical():
    a = tables("oracle/orc_tables.h", "ORC_")
    b = tables("depsolver_amd/csrc/dep_common.h", "k")
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_gauss_legendre_8():
    gl, _, _ = tables("oracle/orc_tables.h", "ORC_")
    x, w = roots_legendre(8)
    assert np.abs(gl[:, 0] - x).max() < 1e-14
    assert np.abs(gl[:, 1] - w).max() < 1e-14


def test_gauss_jacobi_8_weight_one_plus_x():
    _, gj, _ = tables("oracle/orc_tables.h", "ORC_")
    x, w = roots_jacobi(8, 0.0, 1.0)        # weight (1-x)^0 (1+x)^1
    assert np.abs(gj[:, 0] - x).max() < 1e-14
    assert np.abs(gj[:, 1] - w).max() < 1e-14


def test_radon_7_point_rule():
    _, _, t7 = tables("oracle/orc_tables.h", "ORC_")
    s15 = np.sqrt(15.0)
    a1, a2 = (6 - s15) / 21, (6 + s15) / 21
    w1, w2 = (155 - s15) / 2400, (155 + s15) / 2400
    want = np.array([[1 / 3, 1 / 3, 9 / 80], [a2, a2, w2], [1 - 2 * a2, a2, w2], [a2, 1 - 2 * a2, w2],
                     [a1, a1, w1], [1 - 2 * a1, a1, w1], [a1, 1 - 2 * a1, w1]])
    assert np.abs(t7 - want).max() < 1e-14
    # exact for every monomial of degree <= 5 on the unit triangle
    for i in range(6):
        for j in range(6 - i):
            exact = factorial(i) * factorial(j) / factorial(i + j + 2)
            q = np.sum(t7[:, 2] * t7[:, 0] ** i * t7[:, 1] ** j)
            assert abs(q - exact) < 2e-15
