"""CPU model of the lockstep batched engine's schedule (rgf_amd/csrc/rgf_batch.cpp
and the batch_* kernels of rgf_kernels.hip).

The model repeats the kernels' index arithmetic element by element (round ->
slice and block, the row pass's write-back of block rows and columns and its
assembly of the next slice, the tiles skipped around the pivot block, the
backward sweep's per-problem slice) with explicit bounds checks, over a
ragged batch, and checks the result against the slice recursion done
directly: every stored slice inverse, and the solution of the whole
block-tridiagonal system after one forward/backward sweep through exact
inverses.  It pins the hand-over between launches (what the first launch of
round r may read of round r - 1's work), not the kernels' arithmetic, which
the GPU parity tests pin.
"""
import numpy as np
import pytest

BLK = 64
ROWS = 16  # rows per work group of the row-wise passes


def _rand_problem(rng, N, L):
    A0 = -np.diag(np.full(N, -4.0)) + 0j
    for i in range(N):
        A0[i, (i + 1) % N] = -1.0
        A0[(i + 1) % N, i] = -1.0
    eps = 1 + rng.random((N, L)) + 0.01j * rng.random((N, L))
    G0L = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * 0.05 / np.sqrt(N)
    G0R = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * 0.05 / np.sqrt(N)
    return dict(N=N, L=L, A0=A0, eps=eps, k2=0.7 + 0.01j, epsR=1.3, G0L=G0L, G0R=G0R)


def _slice_entry(p, n, i, j, g):
    """slice_entry(): the padded slice matrix entry given g_{n-1}(i, j)."""
    N, L = p["N"], p["L"]
    if i >= N or j >= N:
        return 1.0 if i == j else 0.0
    v = -g
    if n == L + 1:
        v -= p["G0R"][i, j]
    if i == j:
        e = p["eps"][i, n - 1] if n <= L else p["epsR"]
        v -= p["k2"] * e
    return v + p["A0"][i, j]


class Prob:
    def __init__(self, d):
        self.d = d
        self.N, self.L = d["N"], d["L"]
        self.NP = -(-self.N // BLK) * BLK
        self.nb = self.NP // BLK
        self.F = np.zeros((self.L + 2, self.NP, self.NP), complex)
        self.D = np.zeros((BLK, BLK), complex)
        self.R = np.full((BLK, self.NP), np.nan + 0j)  # columns P stay stale (NaN) as in the kernel
        self.C = np.full((self.NP, BLK), np.nan + 0j)
        F0 = np.eye(self.NP, dtype=complex)
        F0[:self.N, :self.N] = d["G0L"]
        self.F[0] = F0

    @property
    def T(self):
        return (self.L + 1) * self.nb


def _chk(*idx_lim):
    for v, lim in idx_lim:
        assert 0 <= v < lim, (v, lim)


def _prev_entry(pb, Prv, pending, p0w, i, j):
    NP = pb.NP
    _chk((i, NP), (j, NP))
    ri, cj = p0w <= i < p0w + BLK, p0w <= j < p0w + BLK
    if pending and ri != cj:
        if ri:
            _chk((i - p0w, BLK))
            return pb.R[i - p0w, j]
        _chk((j - p0w, BLK))
        return pb.C[i, j - p0w]
    return Prv[i, j]


def k1(probs, r):
    # diagonal groups first (they run beside the row pass, which never touches
    # what they read or write: the model runs them on a snapshot)
    snap = [pb.F.copy() for pb in probs]
    for w, pb in enumerate(probs):
        nb, NP = pb.nb, pb.NP
        if r >= pb.T:
            continue
        n, q = 1 + r // nb, r % nb
        Fn = snap[w][n]
        if q == 0:
            for i in range(BLK):
                for j in range(BLK):
                    g = _prev_entry(pb, snap[w][n - 1], False, 0, i, j)
                    Fn[i, j] = _slice_entry(pb.d, n, i, j, g)
            pb.F[n][:BLK, :BLK] = Fn[:BLK, :BLK]
        p0 = q * BLK
        blk = pb.F[n][p0:p0 + BLK, p0:p0 + BLK] if q else Fn[:BLK, :BLK]
        pb.D = np.linalg.inv(blk)
        pb.F[n][p0:p0 + BLK, p0:p0 + BLK] = pb.D
    for pb in probs:
        nb, NP, T = pb.nb, pb.NP, pb.T
        if r > T:
            continue
        asmb = r < T and r % nb == 0
        pending = r >= 1 and nb > 1
        if not asmb and not pending:
            continue
        ps = r - 1
        p0w = (ps % nb) * BLK if pending else 0
        Wsl = 1 + ps // nb if pending else 0
        n = 1 + r // nb
        for g in range(NP // ROWS):               # batch_rows_kernel: groups of rows
            for i in range(g * ROWS, (g + 1) * ROWS):
                _chk((i, NP))
                rowP = pending and p0w <= i < p0w + BLK
                if not asmb:
                    if rowP:
                        for j in range(NP):
                            if p0w <= j < p0w + BLK:
                                continue
                            v = pb.R[i - p0w, j]
                            assert not np.isnan(v)
                            pb.F[Wsl][i, j] = v
                    else:
                        for lane in range(BLK):
                            j = p0w + lane
                            v = pb.C[i, lane]
                            assert not np.isnan(v)
                            pb.F[Wsl][i, j] = v
                    continue
                for j in range(NP):
                    if i < BLK and j < BLK:
                        continue
                    g_ = _prev_entry(pb, pb.F[n - 1], pending, p0w, i, j)
                    colP = p0w <= j < p0w + BLK
                    if pending and rowP != colP:
                        assert Wsl == n - 1
                        pb.F[Wsl][i, j] = g_
                    pb.F[n][i, j] = _slice_entry(pb.d, n, i, j, g_)


def k2(probs, r):
    for pb in probs:
        nb = pb.nb
        if nb == 1 or r >= pb.T:
            continue
        n, p0 = 1 + r // nb, (r % nb) * BLK
        A = pb.F[n]
        for t in range(nb):                   # R tiles, columns P skipped
            if t * BLK == p0:
                continue
            cols = slice(t * BLK, (t + 1) * BLK)
            pb.R[:, cols] = pb.D @ A[p0:p0 + BLK, cols]
        for t in range(nb):                   # C tiles, rows P skipped
            if t * BLK == p0:
                continue
            rows = slice(t * BLK, (t + 1) * BLK)
            pb.C[rows, :] = -A[rows, p0:p0 + BLK] @ pb.D


def k3(probs, r):
    for pb in probs:
        nb = pb.nb
        if r >= pb.T:
            continue
        n, q = 1 + r // nb, r % nb
        p0 = q * BLK
        A = pb.F[n]
        old = A[:, p0:p0 + BLK].copy()
        for t in range((nb - 1) ** 2):
            ti0, tj0 = t // (nb - 1), t % (nb - 1)
            ti, tj = ti0 + (ti0 >= q), tj0 + (tj0 >= q)
            _chk((ti, nb), (tj, nb))
            assert ti != q and tj != q
            I, J = slice(ti * BLK, (ti + 1) * BLK), slice(tj * BLK, (tj + 1) * BLK)
            Rt = pb.R[:, J]
            assert not np.isnan(Rt).any()
            A[I, J] = A[I, J] - old[I, :] @ Rt


def _direct(d):
    """g_n by the plain recursion (FP64, unpadded)."""
    N, L = d["N"], d["L"]
    gs = [d["G0L"]]
    for n in range(1, L + 2):
        M = np.array([[_slice_entry(d, n, i, j, gs[-1][i, j]) for j in range(N)] for i in range(N)])
        gs.append(np.linalg.inv(M))
    return gs


@pytest.mark.parametrize("widths", [(64, 128, 192), (40, 130, 64)])
def test_dyson_rounds_match_the_recursion(widths):
    rng = np.random.default_rng(len(widths) + widths[1])
    Ls = [3, 1, 2]
    probs = [Prob(_rand_problem(rng, N, L)) for N, L in zip(widths, Ls)]
    probs.sort(key=lambda pb: -pb.T)  # the engine's order
    rounds = max(pb.T for pb in probs) + 1
    for r in range(rounds):
        k1(probs, r)
        k2(probs, r)
        k3(probs, r)
    for pb in probs:
        gs = _direct(pb.d)
        for n in range(pb.L + 2):
            assert np.abs(pb.F[n][:pb.N, :pb.N] - gs[n]).max() < 1e-10 * np.abs(gs[n]).max(), (pb.N, n)
            pad = pb.F[n][pb.N:, :]
            assert np.array_equal(pad[:, pb.N:], np.eye(pb.NP - pb.N)) and not pad[:, :pb.N].any()
            assert not pb.F[n][:pb.N, pb.N:].any()


def test_one_sweep_through_exact_inverses_solves_the_system():
    # forward z_n = g_n (r_n + z_{n-1}); backward step k: n = S - 2 - k,
    # z_n += g_n z_{n+1}; with exact g_n, z is the solution of H psi = b
    rng = np.random.default_rng(3)
    probs = [_rand_problem(rng, N, L) for N, L in ((7, 4), (5, 2))]
    for d in probs:
        N, L = d["N"], d["L"]
        S = L + 2
        gs = _direct(d)
        m = 3
        b = [np.zeros((N, m), complex) for _ in range(S)]
        b[0][:, :2] = rng.standard_normal((N, 2))
        b[S - 1][:, 2:] = rng.standard_normal((N, 1))
        z = [None] * S
        for n in range(S):
            t = b[n] + (z[n - 1] if n > 0 else 0)
            z[n] = gs[n] @ t
        smax = max(p["L"] + 2 for p in probs)
        for k in range(smax - 1):
            n = S - 2 - k
            if n < 0:
                continue
            z[n] = z[n] + gs[n] @ z[n + 1]
        # the block-tridiagonal H with the leads folded into the end blocks
        H = np.zeros((S * N, S * N), complex)
        for n in range(S):
            if n == 0:
                Dn = np.linalg.inv(gs[0])  # G0L = (A_lead - G0L)^-1 is what g_0 stands for
            else:
                Dn = np.array([[_slice_entry(d, n, i, j, 0.0) for j in range(N)] for i in range(N)])
            H[n * N:(n + 1) * N, n * N:(n + 1) * N] = Dn
            if n + 1 < S:
                H[n * N:(n + 1) * N, (n + 1) * N:(n + 2) * N] = -np.eye(N)
                H[(n + 1) * N:(n + 2) * N, n * N:(n + 1) * N] = -np.eye(N)
        psi = np.vstack(z)
        assert np.abs(H @ psi - np.vstack(b)).max() < 1e-10 * np.abs(psi).max()


def test_write_back_covers_the_block_rows_and_columns_once():
    # the row pass's write-back of block [p0w, p0w + 64): every entry outside
    # the block in its rows and columns exactly once, nothing else
    for NP in (128, 192, 320):
        for p0w in range(0, NP, 64):
            seen = np.zeros((NP, NP), int)
            for i in range(NP):
                if p0w <= i < p0w + BLK:
                    for j in range(NP):
                        if not p0w <= j < p0w + BLK:
                            seen[i, j] += 1
                else:
                    for lane in range(BLK):
                        seen[i, p0w + lane] += 1
            rows = np.zeros(NP, bool)
            rows[p0w:p0w + 64] = True
            want = (rows[:, None] ^ rows[None, :]).astype(int)
            assert np.array_equal(seen, want), (NP, p0w)
