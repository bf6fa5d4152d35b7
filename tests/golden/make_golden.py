"""Write tests/golden/rgf_golden.npz: small seeded S-matrix cases.

The MATLAB reference cannot run here and ships no vectors, so every case is
computed by the FP64 oracle (oracle/rgf_oracle.py) AND by an independent
brute-force sparse solve of the whole block-tridiagonal system; the script
refuses to write a case on which the two disagree by more than 1e-12.
Inputs (eps arrays, scalars, selections) are stored with the outputs so the
GPU tests consume identical data.

Run from the repository root:  python tests/golden/make_golden.py
(``--full`` writes rgf_golden_full.npz, the full-transverse-size cases).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle.rgf_oracle as o  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rgf_golden.npz")

CASES_2D = [
    # name, ny, nx, yBC, k0dx, eps_L, eps_R, complex eps, out, in
    ("p_full", 24, 12, "periodic", 2 * np.pi / 6, 1.0, 1.0, False, "LR", "LR"),
    ("d_full_abs", 20, 9, "Dirichlet", 2 * np.pi / 5, 1.0, 2.25, True, "LR", "LR"),
    ("n_t", 16, 7, "Neumann", 2 * np.pi / 5.5, 1.5, 1.0, True, "R", "L"),
    ("dn_rr", 13, 5, "DirichletNeumann", 1.0, 1.0, 1.3, False, "R", "R"),
    ("nd_ll", 13, 6, "NeumannDirichlet", 1.1, 1.2, 1.0, True, "L", "L"),
    ("bloch_inR", 15, 8, 1.7, 0.9, 1.0, 2.0, True, "LR", "R"),
    ("p_cplx_k", 18, 10, "periodic", 0.8 + 0.01j, 1.0, 1.0, False, "LR", "LR"),
    ("d_zero_len", 11, 0, "Dirichlet", 1.2, 1.0, 2.5, False, "LR", "LR"),
    ("hi_index", 10, 6, "periodic", 1.5, 2.2, 2.0, True, "LR", "LR"),
]

CASES_3D = [
    ("3d_bloch", 4, 5, 6, [0.4, 2.1], 1.1, 1.0, 1.0, False, "LR", "LR"),
    ("3d_dir", 5, 3, 4, "Dirichlet", 1.3, 1.0, 1.8, True, "R", "L"),
    ("3d_per", 4, 4, 3, "periodic", 1.2, 1.5, 1.0, False, "LR", "LR"),
]


def _cell(s):
    return [{"L": "L", "R": "R"}[c] for c in s]


def main():
    rng = np.random.default_rng(20240601)
    store = {}
    names = []
    for (name, ny, nx, bc, k0dx, eL, eR, cplx, out, inn) in CASES_2D:
        eps = 1 + 2 * rng.random((ny, nx))
        if cplx:
            eps = eps + 0.05j * rng.random((ny, nx))
        syst = dict(epsilon=eps, epsilon_L=eL, epsilon_R=eR, k0dx=k0dx, yBC=bc)
        S, ch = o.cal_smatrix_RGF(syst, _cell(out), _cell(inn))
        iL = np.arange(1, ch["L"]["N_prop"] + 1)
        iR = np.arange(1, ch["R"]["N_prop"] + 1)
        A0 = o.slice_operator_2d(ny, bc)
        sel = lambda s, n: n if s else np.zeros(0, np.int64)  # noqa: E731
        Sb = o.brute_force_smatrix([eps[:, n] for n in range(nx)], A0, complex(k0dx) ** 2, ch["phi"], ch["L"],
                                   ch["R"], eL, eR, sel("L" in inn, iL), sel("R" in inn, iR),
                                   sel("L" in out, iL), sel("R" in out, iR))
        err = np.abs(S - Sb).max()
        assert err < 1e-12, (name, err)
        store[f"{name}/eps"] = eps
        store[f"{name}/params"] = np.array([k0dx, eL, eR], dtype=complex)
        store[f"{name}/bc"] = np.array(str(bc))
        store[f"{name}/sel"] = np.array(out + "|" + inn)
        store[f"{name}/S"] = S
        names.append(name)
    for (name, nx, ny, nz, bc, k0dx, eL, eR, cplx, out, inn) in CASES_3D:
        eps = 1 + rng.random((nx, ny, nz))
        if cplx:
            eps = eps + 0.05j * rng.random((nx, ny, nz))
        syst = dict(epsilon=eps, epsilon_L=eL, epsilon_R=eR, k0dx=k0dx, xyBC=bc)
        S, ch = o.cal_smatrix_RGF_3D(syst, _cell(out), _cell(inn))
        iL = np.arange(1, ch["L"]["N_prop"] + 1)
        iR = np.arange(1, ch["R"]["N_prop"] + 1)
        A0 = o.slice_operator_3d(nx, ny, bc)
        sel = lambda s, n: n if s else np.zeros(0, np.int64)  # noqa: E731
        Sb = o.brute_force_smatrix([eps[:, :, n].ravel(order="F") for n in range(nz)], A0, complex(k0dx) ** 2,
                                   ch["phi"], ch["L"], ch["R"], eL, eR, sel("L" in inn, iL), sel("R" in inn, iR),
                                   sel("L" in out, iL), sel("R" in out, iR))
        err = np.abs(S - Sb).max()
        assert err < 1e-12, (name, err)
        store[f"{name}/eps"] = eps
        store[f"{name}/params"] = np.array([k0dx, eL, eR], dtype=complex)
        store[f"{name}/bc"] = np.array(repr(list(bc)) if not isinstance(bc, str) else bc)
        store[f"{name}/sel"] = np.array(out + "|" + inn)
        store[f"{name}/S"] = S
        names.append(name)
    store["names"] = np.array(names)
    np.savez_compressed(OUT, **store)
    print(f"wrote {len(names)} cases to {OUT}")


OUT_FULL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rgf_golden_full.npz")


def main_full():
    """Oracle S matrices at the benchmark workloads' full transverse sizes,
    on a few slices each (the oracle's general scheme costs 32 N^3 per slice):
      * 2d_n1024: the 2D large-slab workload (N = 1024, Dirichlet, eps_L = 1,
        eps_R = 2.25, lambda = 10 px, eps = 1 + 1.25 Bernoulli(0.3) + 0.01i),
        its first 32 slices, full S (528 x 528);
      * 3d_n4096: the 3D workload (64 x 64 transverse, 40 spheres, Bloch
        phases, lambda = 12 px), its first 6 slices, full S.
    These sizes are beyond the brute-force check of main(); the oracle's
    recursion is the same code that the small cases pin."""
    sys.path.insert(0, ROOT)
    from rgf_amd.workloads import slab_2d_large, spheres_3d
    store = {}
    syst, out, inn = slab_2d_large(nx=32)
    S, _ = o.cal_smatrix_RGF(syst, out, inn)
    eps = syst["epsilon"]
    # eps takes two values: store the Bernoulli mask (exact reconstruction)
    mask = np.real(eps) > 1.5
    assert np.array_equal(eps, 1 + 1.25 * mask + 0.01j)
    store["2d_n1024/mask"] = np.packbits(mask, axis=None)
    store["2d_n1024/shape"] = np.array(eps.shape)
    store["2d_n1024/params"] = np.array([syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"]], dtype=complex)
    store["2d_n1024/S"] = S
    print("2d_n1024", S.shape, flush=True)
    syst, out, inn = spheres_3d()
    eps = syst["epsilon"][:, :, :6].copy()
    S, _ = o.cal_smatrix_RGF_3D(dict(syst, epsilon=eps), out, inn)
    mask = eps > 1.5
    assert np.array_equal(eps, np.where(mask, 2.1, 1.0))
    store["3d_n4096/mask"] = np.packbits(mask, axis=None)
    store["3d_n4096/shape"] = np.array(eps.shape)
    store["3d_n4096/params"] = np.array([syst["k0dx"], syst["epsilon_L"], syst["epsilon_R"]], dtype=complex)
    store["3d_n4096/xyBC"] = np.array(syst["xyBC"], dtype=float)
    store["3d_n4096/S"] = S
    print("3d_n4096", S.shape, flush=True)
    np.savez_compressed(OUT_FULL, **store)
    print(f"wrote {OUT_FULL}")


if __name__ == "__main__":
    if "--full" in sys.argv:
        main_full()
    else:
        main()
