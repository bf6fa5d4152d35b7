"""Loader of tests/golden/rgf_golden_full.npz (full-transverse-size oracle
cases); eps is stored as the mask of its two values and rebuilt here."""
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rgf_golden_full.npz")


def load_case(name: str):
    """(syst, S_oracle) of a case: '2d_n1024' or '3d_n4096'."""
    z = np.load(PATH, allow_pickle=False)
    shape = tuple(int(v) for v in z[f"{name}/shape"])
    mask = np.unpackbits(z[f"{name}/mask"], count=int(np.prod(shape))).astype(bool).reshape(shape)
    k0dx, eL, eR = (complex(v) for v in z[f"{name}/params"])
    real = lambda v: v.real if v.imag == 0 else v  # noqa: E731
    if name == "2d_n1024":
        eps = 1 + 1.25 * mask + 0.01j
        syst = dict(epsilon=eps, epsilon_L=real(eL), epsilon_R=real(eR), k0dx=real(k0dx), yBC="Dirichlet")
    else:
        eps = np.where(mask, 2.1, 1.0)
        ka = [float(v) for v in z[f"{name}/xyBC"]]
        syst = dict(epsilon=eps, epsilon_L=real(eL), epsilon_R=real(eR), k0dx=real(k0dx), xyBC=ka)
    return syst, z[f"{name}/S"]
