"""CPU model of the look-ahead schedule of the slice inverse (invert_inplace in
rgf_amd/csrc/rgf_kernels.hip): blocked Gauss-Jordan on three HIP streams.

The model lists every launch of the schedule in the order the host issues
it, with its stream, the event records and waits between streams, and the
rectangles of each buffer (A, the column copies Aold[0/1], the row copy
Rold, the diagonal inverse D) that it reads and writes.  Two checks:

* races: any two launches on different streams that no chain of stream order
  and event waits orders must not write what the other reads or writes;
* results: executed in several orders that respect those dependencies, the
  launches (complex128 NumPy arithmetic) give A^-1.
The test also checks that taking out any one of the schedule's event waits
makes the checker find a race (the waits are all needed, and the checker
sees them).

The block widths follow the C++: width b, a 64-wide first block when
b = 128 and N > 2b, ragged last blocks.
"""
import itertools

import numpy as np
import pytest


class Model:
    def __init__(self):
        self.ops = []  # dicts: stream, name, reads, writes, fn, deps
        self.tail = {}  # stream -> index of its last op
        self.pending = {}  # stream -> ops it must wait for (events waited on)
        self.events = {}  # event -> op index captured at record time (None: nothing)

    def launch(self, stream, name, reads, writes, fn):
        deps = set(self.pending.pop(stream, set()))
        if stream in self.tail:
            deps.add(self.tail[stream])
        self.ops.append(dict(stream=stream, name=name, reads=reads, writes=writes, fn=fn, deps=deps))
        self.tail[stream] = len(self.ops) - 1

    def record(self, ev, stream):
        self.events[ev] = self.tail.get(stream)

    def wait(self, stream, ev):
        op = self.events.get(ev)
        if op is not None:
            self.pending.setdefault(stream, set()).add(op)

    def ancestors(self):
        anc = []
        for i, op in enumerate(self.ops):
            a = set()
            for d in op["deps"]:
                a.add(d)
                a |= anc[d]
            anc.append(a)
        return anc

    def races(self):
        anc = self.ancestors()
        found = []
        for i, j in itertools.combinations(range(len(self.ops)), 2):
            if i in anc[j] or j in anc[i]:
                continue
            a, b = self.ops[i], self.ops[j]
            for w in a["writes"]:
                for r in b["reads"] + b["writes"]:
                    if overlap(w, r):
                        found.append((a["name"], b["name"], w, r))
            for w in b["writes"]:
                for r in a["reads"]:
                    if overlap(w, r):
                        found.append((b["name"], a["name"], w, r))
        return found

    def run(self, bufs, order):
        done = set()
        for i in order:
            assert self.ops[i]["deps"] <= done
            self.ops[i]["fn"](bufs)
            done.add(i)

    def orders(self, rng, count):
        """Topological orders: the issue order, one that drains streams in
        turn, and random ones."""
        n = len(self.ops)
        yield list(range(n))
        anc_deps = [op["deps"] for op in self.ops]
        for k in range(count):
            done, order = set(), []
            ready = lambda: [i for i in range(n) if i not in done and anc_deps[i] <= done]
            while len(order) < n:
                r = ready()
                i = r[rng.integers(len(r))] if k else r[-1]
                order.append(i)
                done.add(i)
            yield order


def overlap(a, b):
    (ba, r0, r1, c0, c1), (bb, s0, s1, d0, d1) = a, b
    return ba == bb and max(r0, s0) < min(r1, s1) and max(c0, d0) < min(c1, d1)


def outside(lo, hi, n):
    """Index ranges of [0, n) outside [lo, hi)."""
    return [(a, b) for a, b in ((0, lo), (hi, n)) if b > a]


def blocks(N, b):
    b0 = 64 if (b == 128 and N > 2 * b) else b
    p, out = 0, []
    while p < N:
        w = min(b0 if not out else b, N - p)
        out.append((p, w))
        p += w
    return out


def lookahead_schedule(N, b):
    """The schedule invert_inplace issues with the look-ahead streams (s,
    s2, s3).  Block 0 and its panels, step 0 on block 1 and the inverse of
    block 1 are set up on s.  Per block q (P = q, Q = q + 1, Q2 = q + 2):
      s3: after P's panels and the bulk update of step q-1, one launch brings
          row and column Q up to date with step q into the copies Rold and
          Aold[(q+1)%2] and applies step q to the block (Q2, Q2);
      s2: the bulk update of step q outside P, Q and (Q2, Q2), after the step
          on (Q, Q) that precedes Q's inverse on s (so that the inverse, one
          work group that needs a whole CU, is queued first);
      s : Q's panels in place from the copies (D_Q from the inverse queued in
          the previous pass), step q+1 on (Q2, Q2), the inverse of Q2 into D.
    In FP32 the step on (Q2, Q2) is part of the panel launch (its tiles wait
    inside the launch for the row-panel tiles A[Q, Q2] they read); the model
    keeps it as the next launch on s, the order that hand-over enforces.
    """
    m = Model()
    bl = blocks(N, b)
    nbl = len(bl)
    p_of = lambda k: bl[k][0] if k < nbl else N
    nb_of = lambda k: bl[k][1] if k < nbl else 0

    def R(buf, rr, cc):
        return [(buf, r0, r1, c0, c1) for r0, r1 in rr for c0, c1 in cc]

    def diag(k):
        p, w = p_of(k), nb_of(k)

        def fn(B):
            Dk = np.linalg.inv(B["A"][p:p + w, p:p + w])
            B["A"][p:p + w, p:p + w] = Dk
            B["D"][:w, :w] = Dk

        m.launch("s", f"diag{k}", R("A", [(p, p + w)], [(p, p + w)]),
                 R("A", [(p, p + w)], [(p, p + w)]) + R("D", [(0, w)], [(0, w)]), fn)

    def panels(k):
        # A[Pk, J] = D Rold[:, J] (J outside Pk), A[I, Pk] = -Aold[k%2][I, :] D (I outside Pk);
        # A[Pk, P(k+1)] also to Aold[(k+1)%2][Pk, :], A[P(k+1), Pk] also to Rold[:, Pk]
        p, w = p_of(k), nb_of(k)
        p1, w1 = p_of(k + 1), nb_of(k + 1)
        Ao, An = f"Aold{k % 2}", f"Aold{(k + 1) % 2}"
        cols = outside(p, p + w, N)

        def fn(B):
            D = B["D"][:w, :w]
            for c0, c1 in cols:
                B["A"][p:p + w, c0:c1] = D @ B["Rold"][:w, c0:c1]
                B["A"][c0:c1, p:p + w] = -B[Ao][c0:c1, :w] @ D
            if w1:
                B[An][p:p + w, :w1] = B["A"][p:p + w, p1:p1 + w1]
                B["Rold"][:w1, p:p + w] = B["A"][p1:p1 + w1, p:p + w]

        reads = R("D", [(0, w)], [(0, w)]) + R("Rold", [(0, w)], cols) + R(Ao, cols, [(0, w)])
        writes = R("A", [(p, p + w)], cols) + R("A", cols, [(p, p + w)])
        if w1:
            writes += R(An, [(p, p + w)], [(0, w1)]) + R("Rold", [(0, w1)], [(p, p + w)])
        m.launch("s", f"panels{k}", reads, writes, fn)

    def schur(k, j):
        # step j on the diagonal block k: A[Pk, Pk] -= Aold[j%2][Pk, :] A[Pj, Pk]
        pk, nk, pj, nj = p_of(k), nb_of(k), p_of(j), nb_of(j)
        Ao = f"Aold{j % 2}"

        def fn(B):
            B["A"][pk:pk + nk, pk:pk + nk] -= B[Ao][pk:pk + nk, :nj] @ B["A"][pj:pj + nj, pk:pk + nk]

        m.launch("s", f"schur{k}", R(Ao, [(pk, pk + nk)], [(0, nj)]) + R("A", [(pj, pj + nj)], [(pk, pk + nk)]) +
                 R("A", [(pk, pk + nk)], [(pk, pk + nk)]), R("A", [(pk, pk + nk)], [(pk, pk + nk)]), fn)

    p, w = bl[0]

    def copy0(B, p=p, w=w):
        B["Aold0"][:, :w] = B["A"][:, p:p + w]
        B["Rold"][:w, :] = B["A"][p:p + w, :]

    diag(0)
    m.launch("s", "copy0", R("A", [(0, N)], [(p, p + w)]) + R("A", [(p, p + w)], [(0, N)]),
             R("Aold0", [(0, N)], [(0, w)]) + R("Rold", [(0, w)], [(0, N)]), copy0)
    panels(0)
    if nbl > 1:
        schur(1, 0)
        m.record("S0", "s")
        diag(1)
    for q in range(nbl):
        p0, nb = p_of(q), nb_of(q)
        p1, nb1 = p_of(q + 1), nb_of(q + 1)
        p2, nb2 = p_of(q + 2), nb_of(q + 2)
        p3 = p2 + nb2
        Ao, An = f"Aold{q % 2}", f"Aold{(q + 1) % 2}"
        rest = outside(p0, p2, N)
        m.wait("s2", f"S{q % 2}")
        if nb1:
            # s3: row and column Q brought up to date with step q
            if q > 0:
                m.wait("s3", f"B{(q - 1) % 2}")
            m.wait("s3", f"S{q % 2}")

            def s3fn(B, p0=p0, nb=nb, p1=p1, nb1=nb1, p2=p2, nb2=nb2, Ao=Ao, An=An, rest=rest):
                A = B["A"]
                for c0, c1 in rest:
                    B["Rold"][:nb1, c0:c1] = A[p1:p1 + nb1, c0:c1] - B[Ao][p1:p1 + nb1, :nb] @ A[p0:p0 + nb, c0:c1]
                    B[An][c0:c1, :nb1] = A[c0:c1, p1:p1 + nb1] - B[Ao][c0:c1, :nb] @ A[p0:p0 + nb, p1:p1 + nb1]
                if nb2:
                    A[p2:p2 + nb2, p2:p2 + nb2] -= B[Ao][p2:p2 + nb2, :nb] @ A[p0:p0 + nb, p2:p2 + nb2]

            reads = (R(Ao, [(p1, p1 + nb1)], [(0, nb)]) + R("A", [(p0, p0 + nb)], rest) +
                     R("A", [(p1, p1 + nb1)], rest) + R(Ao, rest, [(0, nb)]) +
                     R("A", [(p0, p0 + nb)], [(p1, p1 + nb1)]) + R("A", rest, [(p1, p1 + nb1)]))
            writes = R("Rold", [(0, nb1)], rest) + R(An, rest, [(0, nb1)])
            if nb2:
                reads += R(Ao, [(p2, p3)], [(0, nb)]) + R("A", [(p0, p0 + nb)], [(p2, p3)]) + R("A", [(p2, p3)], [(p2, p3)])
                writes += R("A", [(p2, p3)], [(p2, p3)])
            m.launch("s3", f"rowcol{q}", reads, writes, s3fn)
            m.record(f"T{q % 2}", "s3")

        # s2: bulk update of step q outside P, Q and (Q2, Q2)
        def bulk(B, p0=p0, nb=nb, p2=p2, p3=p3, Ao=Ao, rest=rest):
            A = B["A"]
            for r0, r1 in rest:
                for c0, c1 in rest:
                    upd = B[Ao][r0:r1, :nb] @ A[p0:p0 + nb, c0:c1]
                    # the block (Q2, Q2) is left to the third stream
                    if r0 <= p2 < r1 and c0 <= p2 < c1 and p3 > p2:
                        upd[p2 - r0:p3 - r0, p2 - c0:p3 - c0] = 0
                    A[r0:r1, c0:c1] -= upd

        reads = R(Ao, rest, [(0, nb)]) + R("A", [(p0, p0 + nb)], rest)
        writes = []
        for r0, r1 in rest:
            for c0, c1 in rest:
                writes += split_out(("A", r0, r1, c0, c1), (p2, p3))
        m.launch("s2", f"bulk{q}", reads + writes, writes, bulk)
        m.record(f"B{q % 2}", "s2")
        if nb1:
            # s: Q's panels, step q+1 on (Q2, Q2), the inverse of Q2
            m.wait("s", f"T{q % 2}")
            panels(q + 1)
            if nb2:
                schur(q + 2, q + 1)
                m.record(f"S{(q + 1) % 2}", "s")
                diag(q + 2)
            else:
                m.record(f"S{(q + 1) % 2}", "s")
    m.record("end", "s2")
    m.wait("s", "end")
    return m


def split_out(rect, blk):
    """A rectangle of A minus the square block blk x blk."""
    buf, r0, r1, c0, c1 = rect
    lo, hi = blk
    if hi <= lo or not (r0 < hi and lo < r1 and c0 < hi and lo < c1):
        return [rect]
    out = []
    for a, b2 in ((r0, max(r0, lo)), (min(r1, hi), r1)):
        if b2 > a:
            out.append((buf, a, b2, c0, c1))
    ra, rb = max(r0, lo), min(r1, hi)
    for a, b2 in ((c0, max(c0, lo)), (min(c1, hi), c1)):
        if b2 > a:
            out.append((buf, ra, rb, a, b2))
    return out


CASES = [(256, 64), (320, 64), (448, 128), (640, 128), (700, 128), (200, 64), (130, 64), (576, 128)]


@pytest.mark.parametrize("N,b", CASES)
def test_lookahead_schedule_has_no_races(N, b):
    m = lookahead_schedule(N, b)
    assert m.races() == []


@pytest.mark.parametrize("N,b", CASES)
def test_lookahead_schedule_inverts(N, b):
    rng = np.random.default_rng(N + b)
    A0 = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))) * 0.5 / np.sqrt(N) + 4 * np.eye(N)
    ref = np.linalg.inv(A0)
    m = lookahead_schedule(N, b)
    for order in m.orders(rng, 3):
        bufs = {"A": A0.copy(), "Aold0": np.zeros((N, b), complex), "Aold1": np.zeros((N, b), complex),
                "Rold": np.zeros((b, N), complex), "D": np.zeros((b, b), complex)}
        m.run(bufs, order)
        assert np.abs(bufs["A"] - ref).max() < 1e-12


def test_race_checker_catches_a_missing_wait():
    """The bulk update without its wait for the panels it reads is a race."""
    m = Model()
    m.launch("s", "panels", [], [("A", 0, 64, 0, 256)], lambda B: None)
    m.record("P", "s")
    m.launch("s2", "bulk", [("A", 0, 64, 0, 256)], [("A", 64, 256, 64, 256)], lambda B: None)
    assert m.races()
    m2 = Model()
    m2.launch("s", "panels", [], [("A", 0, 64, 0, 256)], lambda B: None)
    m2.record("P", "s")
    m2.wait("s2", "P")
    m2.launch("s2", "bulk", [("A", 0, 64, 0, 256)], [("A", 64, 256, 64, 256)], lambda B: None)
    assert m2.races() == []


@pytest.mark.parametrize("wait", ['m.wait("s2", f"S{q % 2}")', 'm.wait("s3", f"B{(q - 1) % 2}")',
                                  'm.wait("s3", f"S{q % 2}")', 'm.wait("s", f"T{q % 2}")'])
def test_every_wait_is_needed(wait):
    """Without any one of the event waits the schedule has races."""
    import inspect
    import textwrap
    src = textwrap.dedent(inspect.getsource(lookahead_schedule))
    assert src.count(wait) == 1
    # replace the wait by a statement that does nothing (keeps the block shape)
    src = src.replace(wait, "pass")
    ns = {}
    exec(compile(src, "mutated", "exec"), globals(), ns)
    assert ns["lookahead_schedule"](640, 128).races()
