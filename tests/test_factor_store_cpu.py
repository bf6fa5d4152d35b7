"""CPU model of the mixed path's factor store (FactorStore and
plan_factor_store in rgf_amd/csrc/rgf_api.cpp) when the slice inverses do
not all fit: the first R stay resident, the rest are rebuilt per segment of
C slices from checkpoints.  The model repeats the store's index arithmetic
with bounds-checked slots over the access pattern of the refinement (a
forward sweep 0..S-1, a backward sweep S-2..0, repeated) and checks every
slice it hands out against the recursion computed directly, the device
memory bound R + (segments - 1) + C <= K, and the rebuild count."""
import math

import pytest


def plan(S, K):
    if K >= S:
        return S, 1
    R, C = 0, 1
    for c in range(1, K):
        best = 0
        for r in range(1, S):
            if r + -(-(S - r) // c) - 1 + c <= K:
                best = r
        if best > R or (best == R and best > 0):
            R, C = best, c
    return R, C


class Slots:
    def __init__(self, n):
        self.v = [None] * n

    def __getitem__(self, k):
        assert 0 <= k < len(self.v), (k, len(self.v))
        assert self.v[k] is not None, k
        return self.v[k]

    def __setitem__(self, k, x):
        assert 0 <= k < len(self.v), (k, len(self.v))
        self.v[k] = x


def step(n, prev):
    # any recursion whose values identify (n, history) will do
    return (prev * 1000003 + 7 * n + 1) % 2305843009213693951


class Store:
    def __init__(self, S, K):
        self.S = S
        self.R, self.C = plan(S, K)
        assert self.R >= 1
        self.nseg = -(-(S - self.R) // self.C) if self.R < S else 0
        self.res = Slots(self.R)
        self.ck = Slots(max(self.nseg - 1, 0))
        self.seg = Slots(self.C if self.R < S else 0)
        self.cur = -1
        self.rebuilt = 0
        self.res[0] = 1  # g_0 (the left lead)
        for n in range(1, self.R):
            self.res[n] = step(n, self.res[n - 1])
        for k in range(self.nseg):
            self.build(k, True)

    def build(self, k, first):
        n0 = self.R + k * self.C
        n1 = min(n0 + self.C, self.S)
        for n in range(n0, n1):
            prev = self.seg[n - 1 - n0] if n > n0 else (self.res[self.R - 1] if k == 0 else self.ck[k - 1])
            self.seg[n - n0] = step(n, prev)
        if first and k + 1 < self.nseg:
            self.ck[k] = self.seg[n1 - 1 - n0]
        if not first:
            self.rebuilt += n1 - n0
        self.cur = k

    def at(self, n):
        assert 0 <= n < self.S
        if n < self.R:
            return self.res[n]
        k = (n - self.R) // self.C
        if k != self.cur:
            self.build(k, False)
        return self.seg[n - self.R - k * self.C]


@pytest.mark.parametrize("S", [2, 3, 7, 42, 130])
def test_store_hands_out_the_recursion(S):
    direct = [1]
    for n in range(1, S):
        direct.append(step(n, direct[-1]))
    kmin = next(K for K in range(1, S + 2) if plan(S, K)[0] >= 1)
    assert kmin <= 2 * math.isqrt(S) + 2      # about 2 sqrt(S) slices at least
    for K in sorted({kmin, kmin + 1, (kmin + S) // 2, S - 1, S, S + 5}):
        if K < kmin:
            continue
        st = Store(S, K)
        used = st.R + max(st.nseg - 1, 0) + (st.C if st.R < S else 0)
        assert used <= max(K, S) and (st.R == S) == (K >= S)
        sweeps = 3
        for _ in range(sweeps):
            for n in range(S):
                assert st.at(n) == direct[n], (S, K, n)
            for n in range(S - 2, -1, -1):
                assert st.at(n) == direct[n], (S, K, n)
        if st.R < S:
            # a forward sweep rebuilds every segment it does not find loaded
            # (all but segment 0 after a backward sweep), a backward sweep
            # every segment but the one it starts in
            tail = S - st.R
            assert sweeps * (tail - 2 * st.C) <= st.rebuilt <= sweeps * 2 * tail
        else:
            assert st.rebuilt == 0


def test_plan_prefers_resident_slices_then_long_segments():
    S = 1000
    for K in (70, 100, 400, 999):
        R, C = plan(S, K)
        nseg = -(-(S - R) // C)
        assert R + nseg - 1 + C <= K
        # no plan with more resident slices fits
        assert all(R + 1 + -(-(S - R - 1) // c) - 1 + c > K for c in range(1, K))
