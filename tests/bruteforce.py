"""Brute-force enumeration of alignments (tiny inputs only).

Independent of every DP in this repository: it lists all alignments as op
strings, scores each maximal gap run on its own, and applies the tie-break
rules as an order on op strings read backward from the end:

  * local end cell: the first cell in row-major order among optimal endings;
  * after 'M' (or at the end): M before D before I   (diag > E > F);
  * after 'D': M before I before D when extending beats re-opening
    (gap_extend > gap_open), else M before D before I; after 'I': M, D, I;
  * local: an alignment that starts here beats one that continues.
"""


def run_score(k, go, ge):
    return max(go + (k - 1) * ge, k * go)


def score_ops(ops, a, b, i0, j0, S, go, ge):
    s, i, j, t = 0, i0, j0, 0
    while t < len(ops):
        op = ops[t]
        if op == "M":
            s += S[a[i]][b[j]]
            i += 1
            j += 1
            t += 1
            continue
        k = 1
        while t + k < len(ops) and ops[t + k] == op:
            k += 1
        s += run_score(k, go, ge)
        if op == "I":
            i += k
        else:
            j += k
        t += k
    return s


_PATHS = {}


def _paths(m, n):
    """All op strings from (0,0) consuming exactly m rows and n columns
    (cached: the 7 x 7 set holds D(7,7) = 48,639 strings)."""
    if (m, n) not in _PATHS:
        _PATHS[(m, n)] = _enumerate(m, n)
    return _PATHS[(m, n)]


def _enumerate(m, n):
    out = []

    def rec(i, j, acc):
        if i == m and j == n:
            out.append("".join(acc))
            return
        if i < m and j < n:
            acc.append("M"); rec(i + 1, j + 1, acc); acc.pop()
        if j < n:
            acc.append("D"); rec(i, j + 1, acc); acc.pop()
        if i < m:
            acc.append("I"); rec(i + 1, j, acc); acc.pop()

    rec(0, 0, [])
    return out


def _rank(ctx, op, go, ge):
    if ctx == "D" and ge > go:
        order = "MID"
    else:
        order = "MDI"
    return order.index(op)


def _better(x, y, go, ge):
    """True if op string x is preferred over y (both optimal, same end)."""
    ctx = None
    k = 1
    while True:
        xe = k > len(x)
        ye = k > len(y)
        if xe and ye:
            return False
        if xe != ye:
            return xe  # the one that has already started wins (local only)
        a, b = x[-k], y[-k]
        if a != b:
            return _rank(ctx, a, go, ge) < _rank(ctx, b, go, ge)
        ctx = a
        k += 1


def rle(ops):
    out, t = [], 0
    while t < len(ops):
        k = 1
        while t + k < len(ops) and ops[t + k] == ops[t]:
            k += 1
        out.append("%d%s" % (k, ops[t]))
        t += k
    return "".join(out)


def global_align(a, b, S, go, ge):
    m, n = len(a), len(b)
    best, choice = None, None
    for ops in _paths(m, n):
        s = score_ops(ops, a, b, 0, 0, S, go, ge)
        if best is None or s > best:
            best, choice = s, ops
        elif s == best and _better(ops, choice, go, ge):
            choice = ops
    return best, 0, m - 1, 0, n - 1, rle(choice)


def local_align(a, b, S, go, ge):
    m, n = len(a), len(b)
    cands = []  # (score, end_i, end_j, start_i, start_j, ops)
    for i0 in range(m):
        for j0 in range(n):
            for di in range(1, m - i0 + 1):
                for dj in range(1, n - j0 + 1):
                    for ops in _paths(di, dj):
                        if ops[0] != "M" or ops[-1] != "M":
                            continue
                        s = score_ops(ops, a, b, i0, j0, S, go, ge)
                        cands.append((s, i0 + di, j0 + dj, i0, j0, ops))
    best = max([c[0] for c in cands], default=0)
    if best <= 0:
        return 0, -1, -1, -1, -1, ""
    end = min((c[1], c[2]) for c in cands if c[0] == best)
    pool = [c for c in cands if c[0] == best and (c[1], c[2]) == end]
    choice = pool[0]
    for c in pool[1:]:
        if _better(c[5], choice[5], go, ge):
            choice = c
    s, ei, ej, si, sj, ops = choice
    return s, si, ei - 1, sj, ej - 1, rle(ops)


def linear_nw(a, b, S, g):
    """Textbook linear-gap Needleman-Wunsch score (gap cost g per position)."""
    m, n = len(a), len(b)
    prev = [j * g for j in range(n + 1)]
    for i in range(1, m + 1):
        cur = [i * g] + [0] * n
        for j in range(1, n + 1):
            cur[j] = max(prev[j - 1] + S[a[i - 1]][b[j - 1]], prev[j] + g, cur[j - 1] + g)
        prev = cur
    return prev[n]
