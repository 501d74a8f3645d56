"""Writes tests/golden/oracle_n12.npz: small seeded inputs and what the FP64
oracle (oracle/gpe_oracle.py, the restatement of the reference's propagation
path) computes from them.  Run from the repository root:

    python tests/golden/make_fixtures.py

Contents (n = 12 on [-4, 4)^3, a side the engine runs with its generic
transform; no fast variant is involved):
  psi0                    seeded complex128 state, unit norm
  cases                   JSON list of (kind, static, dynamic, kcut, dt, beta)
  out<i>                  psi after 3 steps of case i from psi0 (t_i = i*dt)
  energy_keys, energies   norm, E_kin, E_pot, E_int, E, mu, Parseval of psi0
                          under case 1's potential at t = 0.37
  nk_counts, nk_Nk        shell histogram of fftn(psi0)

The oracle itself is pinned by the closed-form and brute-force checks of
tests/test_oracle.py; these fixtures freeze its outputs so that a change to
the oracle, or to numpy's FFT, shows up as a failed comparison.
"""

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import gpe_oracle as orc  # noqa: E402

N, LENGTH, STEPS = 12, 8.0, 3
CASES = [
    ("real", "x**2 + y**2 + z**2", None, None, 1e-3, 50.0),
    ("real", "x/2 + y**2/10 + abs(z)/5", "sin(t)*x + cos(pi)*z", None, 1e-3, 50.0),
    ("real", "0.002*(x*x + y*y + z*z)", "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)", 2.5, 1e-3, 50.0),
    ("real", "sqrt(x**2 + y**2 + 1) * z / 10", "cos(t)*x*z/20", None, 2e-3, 200.0),
    ("imaginary", "x**2 + y**2 + 16*z**2", None, None, 1e-3, 50.0),
]
ENERGY_KEYS = ("norm", "E_kin", "E_pot", "E_int", "E", "mu", "parseval")
DIAG_T = 0.37
PATH = Path(__file__).resolve().parent / "oracle_n12.npz"


def seeded_state():
    x, _ = orc.get_grids(N, LENGTH)
    rng = np.random.default_rng(42)
    psi = rng.normal(size=(N,) * 3) + 1j * rng.normal(size=(N,) * 3)
    return orc.normalise(psi, x)


def problem(case):
    kind, static, dynamic, kcut, dt, beta = case
    return orc.Problem(N, LENGTH, kind, dt, beta, static=static, dynamic=dynamic, kcut=kcut)


def compute():
    psi0 = seeded_state()
    out = {"psi0": psi0, "cases": np.array(json.dumps(CASES))}
    for i, case in enumerate(CASES):
        out[f"out{i}"] = problem(case).run(psi0, STEPS)
    energies = problem(CASES[1]).energies(psi0, DIAG_T)
    out["energy_keys"] = np.array(ENERGY_KEYS)
    out["energies"] = np.array([energies[k] for k in ENERGY_KEYS])
    x, k = orc.get_grids(N, LENGTH)
    spectrum = orc.nk_extraction(np.fft.fftn(psi0), x, k)
    out["nk_counts"] = spectrum["counts"]
    out["nk_Nk"] = spectrum["Nk"]
    return out


if __name__ == "__main__":
    np.savez_compressed(PATH, **compute())
    print(f"wrote {PATH.relative_to(ROOT)}")
