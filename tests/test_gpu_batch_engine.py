"""GPU parity of the lockstep batched engine (rgf_solve_batched with
mixed-precision problems of up to 1024 transverse sites).

Every S is checked against the FP64 oracle (<= 1e-9 max-entry error, the
mixed path's tolerance) and against rgf_solve of the same problem; both
launch modes of the engine (hipGraph replays of captured rounds, plain
launches) are covered.  The batched sweep workload (256 solves, widths
128..512, 500 slices, complex k0dx) runs at its full size.
"""
import time

import numpy as np
import pytest

import oracle.rgf_oracle as o
import rgf_amd

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-9


def _ragged_items(seed=21):
    rng = np.random.default_rng(seed)
    items = []
    # widths off the 64 grid (padded with the identity), one slice-less slab,
    # struct selections with repeated indices, both lead sides
    for k, (ny, nx) in enumerate([(40, 9), (100, 14), (64, 0), (130, 11), (192, 7), (24, 16), (70, 5)]):
        eps = 1 + 2 * rng.random((ny, nx)) + 0.02j * rng.random((ny, nx))
        lam = 7 + 10 * rng.random()
        syst = dict(epsilon=eps, epsilon_L=1.0, epsilon_R=[1.0, 1.7, 2.25][k % 3],
                    k0dx=(2 * np.pi / lam) * (1 + 1e-3j * rng.random()),
                    yBC=["periodic", "Dirichlet", 0.7, "Neumann"][k % 4])
        if k == 1:
            sel = ({"ind_out_L": [2, 1, 2], "ind_out_R": [3]}, {"ind_in_L": [1, 1], "ind_in_R": [2, 4]})
        elif k == 4:
            sel = (["R"], ["L"])
        else:
            sel = (["L", "R"], ["L", "R"])
        items.append((syst, *sel))
    eps3 = 1 + rng.random((8, 9, 6))
    items.append((dict(epsilon=eps3, epsilon_L=1.0, epsilon_R=1.3, k0dx=1.05, xyBC=[0.4, 2.0]), ["L", "R"], ["L", "R"]))
    eps3d = 1 + rng.random((5, 13, 4)) + 0.01j
    items.append((dict(epsilon=eps3d, epsilon_L=1.2, epsilon_R=1.0, k0dx=1.2, xyBC="Dirichlet"), ["R"], ["L", "R"]))
    return items


def _single(syst, out, inn, **kw):
    f = rgf_amd.cal_smatrix_RGF_3D if np.ndim(syst["epsilon"]) == 3 else rgf_amd.cal_smatrix_RGF
    return f(syst, out, inn, **kw)


def _oracle(syst, out, inn):
    f = o.cal_smatrix_RGF_3D if np.ndim(syst["epsilon"]) == 3 else o.cal_smatrix_RGF
    return f(syst, out, inn)[0]


@pytest.mark.parametrize("mode", [0, 1])
def test_engine_ragged_batch_vs_oracle_and_single(mode):
    items = _ragged_items()
    print(f"engine mode {mode}: {len(items)} problems", flush=True)
    Sb, infos = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined", return_info=True,
                                                options={"batch_engine": mode, "batch_graph_rounds": 5})
    for k, ((syst, out, inn), S, info) in enumerate(zip(items, Sb, infos)):
        So = _oracle(syst, out, inn)
        S1, _ = _single(syst, out, inn, precision="fp32_refined")
        assert S.shape == So.shape == S1.shape, k
        err = float(np.abs(S - So).max()) if S.size else 0.0
        print(f"  problem {k}: shape {S.shape}, iters {info['refine_iters']}, err {err:.2e}", flush=True)
        assert err < TOL_MIXED, k
        assert np.abs(S - S1).max() < TOL_MIXED, k
        assert info["refine_residual"] <= 1e-13


def test_engine_matches_graph_and_plain_launches_exactly():
    # the same kernels in the same order: hipGraph replays change nothing
    items = _ragged_items(5)[:4]
    a = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined", options={"batch_engine": 0})
    b = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined", options={"batch_engine": 1})
    c = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined",
                                        options={"batch_engine": 0, "batch_graph_rounds": 3})
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y) and np.array_equal(x, z)


def test_engine_validates_and_handles_empty_requests():
    items = _ragged_items()[:3]
    bad = (items[1][0], ["L"], {"ind_in_L": [999]})
    with pytest.raises(ValueError, match="N_prop_L"):
        rgf_amd.cal_smatrix_RGF_batched([items[0], bad, items[2]], precision="fp32_refined")
    # empty requests give real zeros, as the reference does
    Sb = rgf_amd.cal_smatrix_RGF_batched([items[0], (items[2][0], ["L"], {"ind_in_L": []})],
                                         precision="fp32_refined")
    assert Sb[1].shape[1] == 0 and not np.iscomplexobj(Sb[1])


def test_batched_sweep_subset_vs_oracle():
    """Eight problems of the batched sweep workload (widths 128..512, complex
    k0dx), cut to 30 slices, against the oracle."""
    from rgf_amd.workloads import batched_sweep
    items = [(dict(s, epsilon=s["epsilon"][:, :30]), out, inn) for s, out, inn in batched_sweep(count=8)]
    Sb = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined")
    for (syst, out, inn), S in zip(items, Sb):
        err = float(np.abs(S - _oracle(syst, out, inn)).max())
        print(f"  N={syst['epsilon'].shape[0]}: {err:.2e}", flush=True)
        assert err < TOL_MIXED


def test_batched_sweep_full_size():
    """The batched sweep workload at its stated size: 256 solves, widths in
    {128, 192, 256, 384, 512}, 500 slices, eps ~ U[1,3], lambda ~ U[8,20] px,
    Im(k0dx)/Re(k0dx) ~ U[0, 1e-3], full S, FP32 + refinement, in one call.
    Six problems (every width) against rgf_solve's mixed and FP64 paths;
    every S reciprocal (the slab is complex symmetric at complex frequency
    too: S[c, c'] = S[c', c] with c' = ind_prop_conj(c), example_script.m:32-35)."""
    from rgf_amd.workloads import batched_sweep
    items = batched_sweep()
    t0 = time.perf_counter()
    Sb, infos = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined", return_info=True)
    dt = time.perf_counter() - t0
    L = items[0][0]["epsilon"].shape[1]
    print(f"256 solves x {L} slices: {dt:.2f} s, {256 * L / dt:.0f} slice-updates/s "
          f"(slices {infos[0]['seconds_slices']:.2f} s, refinement {infos[0]['seconds_refine']:.2f} s, "
          f"iters {max(i['refine_iters'] for i in infos)})", flush=True)
    widths = [s["epsilon"].shape[0] for s, _, _ in items]
    picks = [widths.index(w) for w in (128, 192, 256, 384, 512)] + [255]
    for k in picks:
        syst, out, inn = items[k]
        S32, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn, precision="fp32_refined")
        S64, _ = rgf_amd.cal_smatrix_RGF(syst, out, inn)
        assert np.abs(Sb[k] - S32).max() < TOL_MIXED, k
        assert np.abs(Sb[k] - S64).max() < TOL_MIXED, k
    for (syst, _, _), S in zip(items, Sb):
        ch = rgf_amd.setup_channels(syst["epsilon"].shape[0], syst["yBC"], syst["k0dx"], syst["epsilon_L"],
                                    syst["epsilon_R"])
        perm = np.concatenate([ch["L"]["ind_prop_conj"], ch["L"]["N_prop"] + ch["R"]["ind_prop_conj"]]) - 1
        St = S[perm]
        assert np.abs(St.T - St).max() < 1e-9


def test_engine_s_error_stop():
    """refine_s_tol through the engine: the estimate bounds the error against
    the FP64 path and stays under the tolerance, with fewer sweeps than the
    residual criterion needs."""
    items = _ragged_items(9)[:6]
    S64 = [_single(s, o_, i_)[0] for s, o_, i_ in items]
    S32, infos = rgf_amd.cal_smatrix_RGF_batched(items, precision="fp32_refined", return_info=True,
                                                 options={"refine_s_tol": 1e-9, "refine_tol": 1e-16})
    for a, b, info in zip(S32, S64, infos):
        err = float(np.abs(a - b).max())
        assert err <= 1e-9
        assert info["refine_s_err_estimate"] <= 1e-9
        assert info["refine_iters"] <= 3
