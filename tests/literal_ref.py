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


# ------------------------------------------------------------ allpairs ----

# GetSegSwapScoreAll's bit compositions (reorder/reorder.h:185-227) and the
# SwapPairs pairtab (reorder/swappair.h:54-97), as data.
AP_POS = [
    [1, 2, 6, 7, 0, 3, 4, 5], [1, 2, 5, 7, 0, 3, 4, 6], [1, 2, 5, 6, 0, 3, 4, 7],
    [1, 2, 4, 7, 0, 3, 5, 6], [1, 2, 4, 6, 0, 3, 5, 7], [1, 2, 4, 5, 0, 3, 6, 7],
    [0, 1, 4, 5, 2, 3, 6, 7], [0, 1, 4, 6, 2, 3, 5, 7], [0, 1, 4, 7, 2, 3, 5, 6],
    [0, 1, 5, 6, 2, 3, 4, 7], [0, 1, 5, 7, 2, 3, 4, 6], [0, 1, 6, 7, 2, 3, 4, 5],
    [0, 2, 4, 5, 1, 3, 6, 7], [0, 2, 4, 6, 1, 3, 5, 7], [0, 2, 4, 7, 1, 3, 5, 6],
    [0, 2, 5, 6, 1, 3, 4, 7], [0, 2, 5, 7, 1, 3, 4, 6], [0, 2, 6, 7, 1, 3, 4, 5],
    [0, 3, 4, 5, 1, 2, 6, 7], [0, 3, 4, 6, 1, 2, 5, 7], [0, 3, 4, 7, 1, 2, 5, 6],
    [0, 3, 5, 6, 1, 2, 4, 7], [0, 3, 5, 7, 1, 2, 4, 6], [0, 3, 6, 7, 1, 2, 4, 5],
    [1, 3, 6, 7, 0, 2, 4, 5], [1, 3, 5, 7, 0, 2, 4, 6], [1, 3, 5, 6, 0, 2, 4, 7],
    [1, 3, 4, 7, 0, 2, 5, 6], [1, 3, 4, 6, 0, 2, 5, 7], [1, 3, 4, 5, 0, 2, 6, 7],
    [2, 3, 4, 5, 0, 1, 6, 7], [2, 3, 4, 6, 0, 1, 5, 7], [2, 3, 4, 7, 0, 1, 5, 6],
    [2, 3, 5, 6, 0, 1, 4, 7], [2, 3, 5, 7, 0, 1, 4, 6], [2, 3, 6, 7, 0, 1, 4, 5]]
AP_PAIRTAB = [
    [0, 6, 3, 7], [0, 5, 3, 7], [0, 5, 3, 6], [0, 4, 3, 7], [0, 4, 3, 6], [0, 4, 3, 5],
    [2, 4, 3, 5], [2, 4, 3, 6], [2, 4, 3, 7], [2, 5, 3, 6], [2, 5, 3, 7], [2, 6, 3, 7],
    [1, 4, 3, 5], [1, 4, 3, 6], [1, 4, 3, 7], [1, 5, 3, 6], [1, 5, 3, 7], [1, 6, 3, 7],
    [1, 4, 2, 5], [1, 4, 2, 6], [1, 4, 2, 7], [1, 5, 2, 6], [1, 5, 2, 7], [1, 6, 2, 7],
    [0, 6, 2, 7], [0, 5, 2, 7], [0, 5, 2, 6], [0, 4, 2, 7], [0, 4, 2, 6], [0, 4, 2, 5],
    [0, 4, 1, 5], [0, 4, 1, 6], [0, 4, 1, 7], [0, 5, 1, 7], [0, 5, 1, 6], [0, 6, 1, 5]]


def exchange_of_pos(i):
    """(l positions given away, r positions taken) of double swap i as scored."""
    kept = [p for p in AP_POS[i][:4] if p < 4]
    taken = sorted(p - 4 for p in AP_POS[i][:4] if p >= 4)
    return sorted(set(range(4)) - set(kept)), taken


def seg_score_all(bm, N, seg1, seg2):
    """GetSegSwapScoreAll per row (M = 4): 16 singles, 36 doubles."""
    M = 4
    sc = [0] * 104
    for row in bm:
        l, r = row[seg1], row[seg2]
        base = seg_score([[l, r]], N, M, 0, 1)
        for i in range(32):
            sc[i] += base[i]
        bits = [(l >> (M - 1 - i)) & 1 for i in range(M)] + [(r >> (M - 1 - i)) & 1 for i in range(M)]
        for i in range(36):
            ln = sum(bits[AP_POS[i][j]] << (M - 1 - j) for j in range(M))
            rn = sum(bits[AP_POS[i][M + j]] << (M - 1 - j) for j in range(M))
            sc[2 * (16 + i)] += (popc(l) > N) - (popc(ln) > N)
            sc[2 * (16 + i) + 1] += (popc(r) > N) - (popc(rn) > N)
    return sc


def allpairs_reorder(R, rows, cols, vals, N, sched, clean=False, maxiter=10):
    """run_allpairs & siblings (sched 4..7, M = 4), literal dense form.
    Returns (stats, number of swaps over an already relabelled node)."""
    M = 4
    nseg = (R + M - 1) // M
    newid = [-1] * R
    base = sched - 4
    rev = base in (2, 3)
    drain = base in (1, 3)
    state = {"bm": dense_bitmaps(R, M, rows, cols, vals), "over": 0}

    def locate_list():
        cnt = [0] * nseg
        for row in state["bm"]:
            for s_ in range(nseg):
                if popc(row[s_]) > N:
                    cnt[s_] += 1
        pq = StdPQ(cmp_reverse if rev else cmp)
        for s_ in range(nseg):
            if cnt[s_] != 0:
                pq.push((s_, cnt[s_]))
        return pq

    def total(pq):
        q = pq.copy()
        t = 0
        while len(q):
            t += q.top()[1]
            q.pop()
        return t

    def min_outdeg_seg():
        zero = [sum(1 for row in state["bm"] if row[s_] == 0) for s_ in range(nseg)]
        return sorted((zero[s_], s_) for s_ in range(nseg))[-1][1]

    def pair_list(s1, s2):
        sc = seg_score_all(state["bm"], N, s1, s2)
        pq = StdPQ(lambda a, b: swap_less((0, 0, a[4], a[5]), (0, 0, b[4], b[5])))
        for i in range(16):
            pq.push((s1 * M + i // M, s2 * M + i % M, -1, -1, sc[2 * i], sc[2 * i + 1]))
        for i in range(36):
            if clean:
                lo, tk = exchange_of_pos(i)
                a, b, c, d = lo[0], tk[0], lo[1], tk[1]
            else:
                a, b, c, d = AP_PAIRTAB[i][0], AP_PAIRTAB[i][1] - 4, AP_PAIRTAB[i][2], AP_PAIRTAB[i][3] - 4
            pq.push((s1 * M + a, s2 * M + b, s1 * M + c, s2 * M + d, sc[2 * (16 + i)], sc[2 * (16 + i) + 1]))
        return pq

    def visited(v):
        return v != -1 and newid[v] != -1

    def swap(a, b):
        if newid[a] != -1 or newid[b] != -1:
            state["over"] += 1
        newid[a] = b
        newid[b] = a

    def apply(p):
        swap(p[0], p[1])
        if p[2] != -1 and p[3] != -1:
            swap(p[2], p[3])

    lst = locate_list()
    init_cnt = total(lst)
    init_size = len(lst)
    rnd = 1
    MAXITER = max(maxiter, init_size)
    while rnd <= MAXITER and len(lst) >= 1:
        if len(lst) == 1:
            apply(pair_list(lst.top()[0], min_outdeg_seg()).top())
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
                pq = pair_list(pseg, tseg)
                if clean:
                    while len(pq) and any(visited(v) for v in pq.top()[:4]):
                        pq.pop()
                else:
                    while len(pq) and (visited(pq.top()[0]) or visited(pq.top()[1])) and \
                            (visited(pq.top()[2]) or visited(pq.top()[3])):
                        pq.pop()
                if len(pq):
                    p = pq.top()
                    apply(p)
                    pcnt -= p[4]
                    tcnt -= p[5]
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
    return [init_size, len(lst), init_cnt, total(lst), rnd - 1], state["over"]


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
