"""The process-wide device handle under concurrent use and repeated large
solves (rgf_amd._lib.using, rgf_solve's factor-store accounting)."""
import threading

import numpy as np
import pytest
import torch

import rgf_amd

pytestmark = pytest.mark.gpu


def _slab(seed, ny=192, nx=60):
    rng = np.random.default_rng(seed)
    eps = 1 + 2 * rng.random((ny, nx))
    return dict(epsilon=eps, epsilon_L=1.0, epsilon_R=1.5, k0dx=2 * np.pi / 9, yBC="periodic")


def test_two_threads_share_the_handle():
    # ctypes releases the GIL inside the library: without the handle lock the
    # two solves would share the handle's cached device buffers
    systs = [_slab(1), _slab(2, ny=256)]
    precs = ["fp32_refined", "fp64"]
    ref = [rgf_amd.cal_smatrix_RGF(s, ["L", "R"], ["L", "R"], precision=p)[0] for s, p in zip(systs, precs)]
    out = [[None] * 3 for _ in range(2)]
    errs = []

    def work(k):
        try:
            for r in range(3):
                out[k][r] = rgf_amd.cal_smatrix_RGF(systs[k], ["L", "R"], ["L", "R"], precision=precs[k])[0]
        except Exception as e:  # noqa: BLE001
            errs.append(e)

    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    for k in range(2):
        for r in range(3):
            assert np.array_equal(out[k][r], ref[k]), (k, r)


def test_repeat_solve_counts_reused_factor_store():
    # A mixed solve whose factor store needs more than half of the free device
    # memory must succeed again on the same handle: the buffers it keeps from
    # the first solve are reused, not added to the need.
    rgf_amd.trim(0)
    torch.cuda.empty_cache()
    ny, nx = 1024, 200
    need = (nx + 2) * ny * ny * 8  # FP32 factor store, 1.7 GB
    free, _ = torch.cuda.mem_get_info(0)
    keep = int(1.45 * need) + (256 << 20)
    ballast = torch.empty(max(free - keep, 0), dtype=torch.uint8, device="cuda:0")
    try:
        rng = np.random.default_rng(3)
        syst = dict(epsilon=1 + rng.random((ny, nx)), epsilon_L=1.0, epsilon_R=1.0, k0dx=2 * np.pi / 12,
                    yBC="Dirichlet")
        sel = {"ind_in_L": [1, 2, 3, 4]}, {"ind_out_L": [1, 2, 3, 4]}  # 4 right-hand sides: the store dominates
        S1, _ = rgf_amd.cal_smatrix_RGF(syst, sel[1], sel[0], precision="fp32_refined")
        free2, _ = torch.cuda.mem_get_info(0)
        assert need > 0.5 * free2  # the old accounting would refuse the second solve
        S2, _ = rgf_amd.cal_smatrix_RGF(syst, sel[1], sel[0], precision="fp32_refined")
        assert np.array_equal(S1, S2)
    finally:
        del ballast
        rgf_amd.trim(0)
        torch.cuda.empty_cache()
