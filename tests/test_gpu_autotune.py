"""The slice-inverse block-width tuner (rgf_options.autotune): one timing
run per (device, N, precision, three-product bit, look-ahead) in a process,
a width from the candidates, and the same S as the fixed widths."""
import numpy as np
import pytest

import oracle.rgf_oracle as o
import rgf_amd

pytestmark = pytest.mark.gpu

CANDIDATES = {"fp64": (32, 64), "fp32_refined": (64, 128)}
FIXED = {"fp64": 64, "fp32_refined": 128}
TOL = {"fp64": 1e-10, "fp32_refined": 1e-9}


def _slab(ny, nx=12, seed=4):
    rng = np.random.default_rng(seed)
    eps = 1 + 2 * rng.random((ny, nx)) + 0.01j * rng.random((ny, nx))
    return dict(epsilon=eps, epsilon_L=1.0, epsilon_R=2.0, k0dx=2 * np.pi / 10, yBC="Dirichlet")


@pytest.mark.parametrize("prec", ["fp64", "fp32_refined"])
def test_tuned_once_per_size_and_same_S(prec):
    # widths no other test uses, so that this process has not tuned them yet
    for ny in (232, 296):
        syst = _slab(ny)
        So = o.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"])[0]
        S1, _, i1 = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=prec, return_info=True)
        S2, _, i2 = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=prec, return_info=True)
        S0, _, i0 = rgf_amd.cal_smatrix_RGF(syst, ["L", "R"], ["L", "R"], precision=prec, return_info=True,
                                            options={"autotune": 0})
        print(f"{prec} N={ny}: tuned b={i1['block_size']}", flush=True)
        assert i1["autotuned"] == 1 and i2["autotuned"] == 0 and i0["autotuned"] == 0
        assert i1["block_size"] in CANDIDATES[prec] and i2["block_size"] == i1["block_size"]
        assert i0["block_size"] == FIXED[prec]
        assert np.array_equal(S1, S2)      # same width, same kernels: the same numbers
        for S in (S1, S0):
            assert np.abs(S - So).max() < TOL[prec]


def test_explicit_block_size_is_not_tuned():
    syst = _slab(168)
    _, _, info = rgf_amd.cal_smatrix_RGF(syst, ["L"], ["L"], precision="fp64", return_info=True,
                                         options={"block_size": 48})
    assert info["block_size"] == 48 and info["autotuned"] == 0


def test_batched_workers_use_pretuned_widths():
    # problems outside the lockstep engine (FP64) on several worker contexts:
    # the widths are tuned before the workers start, and every S matches its
    # single solve
    items = [(_slab(ny, seed=ny), ["L", "R"], ["L", "R"]) for ny in (264, 264, 328, 200)]
    Sb, infos = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp64", return_info=True,
                                                options={"batch_concurrency": 3})
    for (syst, out, inn), S, info in zip(items, Sb, infos):
        S1, _, i1 = rgf_amd.cal_smatrix_RGF(syst, out, inn, return_info=True)
        assert info["autotuned"] == 0 and info["block_size"] == i1["block_size"]
        assert np.abs(S - S1).max() < 1e-12
