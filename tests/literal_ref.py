"""Literal dense restatement of the reference N:M greedy and VENOM step.

Test-only.  Unlike the C++ oracle (sparse per-segment stores, incremental
bookkeeping) this follows the reference code shape directly: a dense
(row, segment) bitmap matrix rebuilt from the COO after every relabel, the
per-row gain loop of GetSegSwapScore (reorder/reorder.h:75-131), a Python
emulation of libstdc++'s std::priority_queue (push_heap / pop_heap, as in
bits/stl_heap.h) driven by the reference comparators
(reorder/swappair.h:112-120, reorder/reorder.h:421-434), and the dense
string-keyed sort of reordering.py:113-177.  It is only fast enough for tiny
matrices; tests compare the oracle against it.
"""


class StdPQ:
    """std::priority_queue over a vector with comparator less(a, b)."""

    def __init__(self, less):
        self.c = []
        self.less = less
        self.last_popped = None

    def __len__(self):
        return len(self.c)

    def copy(self):
        q = StdPQ(self.less)
        q.c = list(self.c)
        return q

    def top(self):
        return self.c[0]

    def _push_heap(self, hole, top, value):
        parent = int((hole - 1) / 2)
        while hole > top and self.less(self.c[parent], value):
            self.c[hole] = self.c[parent]
            hole = parent
            parent = int((hole - 1) / 2)
        self.c[hole] = value

    def push(self, x):
        self.c.append(x)
        self._push_heap(len(self.c) - 1, 0, x)

    def _adjust_heap(self, hole, length, value):
        top = hole
        second = hole
        while second < int((length - 1) / 2):
            second = 2 * (second + 1)
            if self.less(self.c[second], self.c[second - 1]):
                second -= 1
            self.c[hole] = self.c[second]
            hole = second
        if (length & 1) == 0 and second == int((length - 2) / 2):
            second = 2 * (second + 1)
            self.c[hole] = self.c[second - 1]
            hole = second - 1
        self._push_heap(hole, top, value)

    def pop(self):
        n = len(self.c)
        self.last_popped = self.c[0]
        if n > 1:
            last = n - 1
            value = self.c[last]
            self.c[last] = self.c[0]
            self._adjust_heap(0, last, value)
        self.c.pop()


def swap_less(p1, p2):
    # p = (v1, v2, lgain, rgain, idx)
    s1, s2 = p1[2] + p1[3], p2[2] + p2[3]
    if s1 == s2:
        return p1[3] <= p2[3]
    return s1 < s2


def cmp(a, b):
    if a[1] == b[1]:
        return a[0] > b[0]
    return a[1] < b[1]


def cmp_reverse(a, b):
    if a[1] == b[1]:
        return a[0] < b[0]
    return a[1] > b[1]


def popc(x):
    return bin(x).count("1")


def dense_bitmaps(R, M, rows, cols, vals, use_values=True):
    nseg = (R + M - 1) // M
    bm = [[0] * nseg for _ in range(R)]
    for r, c, v in zip(rows, cols, vals):
        if use_values and not (v != 0.0):
            continue
        if c // M < nseg:
            bm[r][c // M] |= 1 << (M - 1 - c % M)
    return bm


def seg_score(bm, N, M, seg1, seg2):
    score = [0] * (2 * M * M)
    for row in bm:
        lval, rval = row[seg1], row[seg2]
        for b1 in range(M):
            for b2 in range(M):
                bit1 = (lval >> (M - 1 - b1)) & 1
                bit2 = (rval >> (M - 1 - b2)) & 1
                lnew = (lval & ~(1 << (M - 1 - b1))) | (bit2 << (M - 1 - b1))
                rnew = (rval & ~(1 << (M - 1 - b2))) | (bit1 << (M - 1 - b2))
                score[(b1 * M + b2) * 2] += (popc(lval) > N) - (popc(lnew) > N)
                score[(b1 * M + b2) * 2 + 1] += (popc(rval) > N) - (popc(rnew) > N)
    return score


def pop_order(score, M):
    pq = StdPQ(swap_less)
    for i in range(M * M):
        pq.push((i // M, M + i % M, score[2 * i], score[2 * i + 1], i))
    out = []
    while len(pq):
        out.append(pq.top()[4])
        pq.pop()
    return out


def nm_reorder(R, rows, cols, vals, N, M, sched, maxiter=10):
    """run_onepair & siblings (sched 0..3) and run (sched >= 8).  Mutates
    rows/cols (lists) in place; returns the 5 stats."""
    nseg = (R + M - 1) // M
    newid = [-1] * R
    rev = sched in (2, 3)
    drain = sched in (1, 3)
    plain = sched >= 8
    state = {"bm": dense_bitmaps(R, M, rows, cols, vals)}

    def locate_list():
        bm = state["bm"]
        cnt = [0] * nseg
        for row in bm:
            for s in range(nseg):
                if popc(row[s]) > N:
                    cnt[s] += 1
        pq = StdPQ(cmp_reverse if rev else cmp)
        for s in range(nseg):
            if cnt[s] != 0:
                pq.push((s, cnt[s]))
        return pq

    def total(pq):
        q = pq.copy()
        t = 0
        while len(q):
            t += q.top()[1]
            q.pop()
        return t

    def min_outdeg_seg():
        bm = state["bm"]
        zero = [sum(1 for row in bm if row[s] == 0) for s in range(nseg)]
        pairs = sorted((zero[s], s) for s in range(nseg))
        return pairs[-1][1]

    def score_list(s1, s2):
        sc = seg_score(state["bm"], N, M, s1, s2)
        pq = StdPQ(swap_less)
        for i in range(M * M):
            pq.push((s1 * M + i // M, s2 * M + i % M, sc[2 * i], sc[2 * i + 1], i))
        return pq

    def visited(v):
        return newid[v] != -1

    def swap(a, b):
        newid[a] = b
        newid[b] = a

    lst = locate_list()
    init_cnt = total(lst)
    init_size = len(lst)
    rnd = 1
    MAXITER = max(maxiter, init_size)
    min_size = 2 if plain else 1
    while rnd <= MAXITER and len(lst) >= min_size:
        if not plain and len(lst) == 1:
            minseg = min_outdeg_seg()
            pq = score_list(lst.top()[0], minseg)
            p = pq.top()
            swap(p[0], p[1])
        while len(lst) >= 2:
            pseg, pcnt = lst.top()
            lst.pop()
            if drain:
                temp = StdPQ(cmp if rev else cmp_reverse)
                while len(lst):
                    temp.push(lst.top())
                    lst.pop()
            else:
                temp = lst.copy()
            upd = StdPQ(cmp_reverse if rev else cmp)
            while len(temp):
                if pcnt <= 0 or all(visited(pseg * M + i) for i in range(M)):
                    break
                tseg, tcnt = temp.top()
                pq = score_list(pseg, tseg)
                while len(pq) and (visited(pq.top()[0]) or visited(pq.top()[1])):
                    pq.pop()
                p = pq.top() if len(pq) else pq.last_popped
                swap(p[0], p[1])
                pcnt -= p[2]
                tcnt -= p[3]
                if tcnt > 0:
                    upd.push((tseg, tcnt))
                temp.pop()
            lst = upd
        for i in range(len(rows)):
            u, v = rows[i], cols[i]
            if newid[u] != -1:
                rows[i] = newid[u]
            if newid[v] != -1:
                cols[i] = newid[v]
        for i in range(R):
            newid[i] = -1
        state["bm"] = dense_bitmaps(R, M, rows, cols, vals)
        lst = locate_list()
        rnd += 1
    lst = locate_list()
    return [init_size, len(lst), init_cnt, total(lst), rnd - 1]


# ---------------------------------------------------------------- VENOM ----

def gray_code(n):
    if n == 0:
        return ["0"]
    if n == 1:
        return ["0", "1"]
    g = gray_code(n - 1)
    return ["0" + c for c in g] + ["1" + c for c in g[::-1]]


def hash_table(M):
    tab = {}
    for k, code in enumerate(gray_code(M)):
        tab[code] = k if code.count("1") <= 2 else -k
    return tab


def segmat(m, n, M, rows, cols):
    sm = [["0" * M for _ in range(n // M)] for _ in range(m)]
    for r, c in zip(rows, cols):
        t = list(sm[r][c // M])
        t[c % M] = "1"
        sm[r][c // M] = "".join(t)
    return sm


def venom_counts(sm, m, n, M, V):
    inv = inv_s = 0
    for seg in range(n // M):
        for i in range(0, (m + V - 1) // V * V, V):
            b = 0
            bad = False
            for j in range(i, i + V):
                if j < m:
                    if sm[j][seg].count("1") > 2:
                        bad = True
                    b |= int(sm[j][seg], 2)
            if bin(b).count("1") > 4:
                inv += 1
            if bin(b).count("1") > 4 or bad:
                inv_s += 1
    return inv, inv_s


def venom_new_id(sm, m, n, M):
    tab = hash_table(M)
    keyed = [(i, [tab[sm[i][j]] for j in range(n // M)]) for i in range(m)]
    return [x[0] for x in sorted(keyed, key=lambda x: x[1])]


def venom_reorder(nrows, ncols, rows, cols, M, V, maxiter=10):
    m = (nrows + M - 1) // M * M
    n = (ncols + M - 1) // M * M
    sm = segmat(m, n, M, rows, cols)
    init, init_s = venom_counts(sm, m, n, M, V)
    inv, prev, prev_s = init, init, init_s
    prow, pcol = list(rows), list(cols)
    rnd = 1
    MAXITER = max(maxiter, init)
    while inv > 0 and rnd <= MAXITER:
        new_id = venom_new_id(sm, m, n, M)
        nrow = [new_id[r] for r in prow]
        ncol = [new_id[c] for c in pcol]
        sm = segmat(m, n, M, nrow, ncol)
        inv, st = venom_counts(sm, m, n, M, V)
        if inv > prev:
            return [init, init_s, prev, prev_s, rnd], prow, pcol
        prev, prev_s = inv, st
        prow, pcol = nrow, ncol
        rnd += 1
    return [init, init_s, prev, prev_s, rnd - 1], prow, pcol
