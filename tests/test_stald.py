"""The stability-limit detector of the GPU integrator (amici_amd/csrc/
amx_stald.h, compiled here for the host) against CVODES's own cvSLdet
(ThirdParty/sundials/src/cvodes/cvodes.c:8256, reached through
oracle/stald_probe.c): the same decision on every input, for

* the detector inputs of real integrations (every step at which CVODES
  called the detector while integrating 8 N40 and 8 Robertson conditions at
  rtol 1e-8 with stability-limit detection on),
* constructed inputs of the normal-matrix case (three orders sharing the
  root rr, magnitudes chosen so that the variance check holds, with rr on
  both sides of the 0.98 cut), and
* random inputs,

with every outcome of the detector (stable root 1-3, violation 4-6, and the
failure flags) occurring among them.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from amici_amd import networks

ROOT = Path(__file__).resolve().parent.parent
GOLD = Path(__file__).resolve().parent / "golden"

SHIM = r"""
#include "amx_stald.h"
extern "C" int amx_stald(const double* ss, int q) {
  double a[5][3];
  for (int i = 0; i < 5; ++i)
    for (int k = 0; k < 3; ++k) a[i][k] = ss[i * 3 + k];
  return amx::stald_flag(a, q);
}
"""


@pytest.fixture(scope="module")
def detector(tmp_path_factory):
    d = tmp_path_factory.mktemp("stald")
    (d / "shim.cpp").write_text(SHIM)
    lib = d / "libstald.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I",
                    str(ROOT / "amici_amd" / "csrc"), str(d / "shim.cpp"), "-o", str(lib)],
                   check=True)
    L = ctypes.CDLL(str(lib))
    L.amx_stald.restype = ctypes.c_int
    L.amx_stald.argtypes = [ctypes.c_void_p, ctypes.c_int]

    def flag(ss, q):
        a = np.ascontiguousarray(ss, np.float64).reshape(5, 3)
        return int(L.amx_stald(a.ctypes.data, int(q)))
    return flag


def _real_samples():
    from oracle import cvodes_oracle as O
    rows = []
    for net, fx in ((networks.network_n40(), "n40_conditions16"),
                    (networks.robertson(), "robertson_batch64")):
        g = np.load(GOLD / f"{fx}.npz")
        for b in range(8):
            kf = g["kfix"][b] if g["kfix"].size else np.zeros(0)
            rows.append(O.stald_samples(net, g["theta"][b], kf, g["ts"]))
    return np.concatenate(rows)


def _normal_case(rng, n):
    i = np.arange(5)[:, None]
    out = []
    for _ in range(n):
        q = int(rng.integers(3, 6))
        rr = rng.uniform(0.9, 1.06)
        qfac1, qfac2 = 0.25 * (q * q - 1.0), 2.0 / (q - 1.0)
        x = rng.uniform(0.2, 5.0)
        y = (1 + qfac1 * x + (1 - 1 / rr) / qfac2) / x
        ss = np.array([y, 1.0, x]) * rr ** (-i)
        ss = ss * (1 + rng.normal(0, 10 ** rng.uniform(-7, -4), (5, 3)))
        out.append((q, ss))
    return out


def test_detector_matches_cvodes(detector):
    from oracle import cvodes_oracle as O
    rng = np.random.default_rng(7)
    cases = [(int(r[0]), r[1:].reshape(5, 3)) for r in _real_samples()]
    n_real = len(cases)
    assert n_real > 1000
    cases += _normal_case(rng, 3000)
    cases += [(int(rng.integers(3, 6)), np.exp(rng.normal(0, 2, (5, 3)))) for _ in range(3000)]
    cases += [(int(rng.integers(3, 6)), rng.uniform(0.1, 10, (1, 3))
               * (1 + rng.uniform(-1, 1, (1, 3)) * rng.uniform(0.6, 1.4) ** (-np.arange(5)[:, None])))
              for _ in range(3000)]
    seen = set()
    for j, (q, ss) in enumerate(cases):
        want = O.sldet(ss, q)
        got = detector(ss, q)
        assert got == want, (j, j < n_real, q, ss.tolist(), got, want)
        seen.add(want)
    assert {1, 2, 3, 4, 5, 6} <= seen, seen
    assert {-1, -2, -3, -4, -5, -6, -7, -9} <= seen, seen
