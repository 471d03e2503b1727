"""A CPU model of the fill's strip schedule (tweakseq_amd/csrc/fill.h,
fill_phase): why the first blocks of a strip keep the per-step column test.

A group of GG lanes sweeps a strip of GG * R query rows; lane l holds R rows
and works on column j = t - l at step t.  Its top input at step t is the
bottom row (Hbot, Fbot) the lane above left after step t - 1 (lane 0 reads
the previous strip's bottom row from the ring).  At a strip start a lane's
registers hold the left boundary, H(i, 0), and keep it until its first
column (j = 0).  Hbot and Fbot are NOT reset at a strip start: they still
hold the lanes' last values of the previous strip.

The masked form (the column test, `if (FULL || (j >= 0 && j < n))`) leaves a
lane untouched before its first column.  The straight-line form computes
every step; before its first column a lane then runs the recurrence on pad
scores with the lane above's stale Hbot / Fbot, which reach the lane's rows
through F (F = Hup + open) and overwrite the boundary it still needs.  Real
cells come out wrong from the second strip on (in the first strip the stale
values are the initial boundary, so nothing shows), the direction planes
with them, and a walk over wrong planes could loop -- the hang recorded when
local tasks' first and last blocks ran straight-line.  With Hbot / Fbot reset
to the boundary at each strip start the pre-column steps of a local sweep
leave every register at 0 and the straight-line form agrees: the missing
reset is the cause.

The model runs the same recurrence as the kernel (local, E and F floored at
0 by saturating gap terms, pad score -64) in both schedules and checks every
real cell against a row-major DP, and the best score against the oracle."""
import numpy as np
import pytest

from oracle import pyoracle as po

PAD = -64
GG, R = 4, 2  # a small group: 8-row strips


def row_major_local(a, b, smat, go, ge):
    m, n = len(a), len(b)
    H = np.zeros((m + 1, n + 1), dtype=np.int64)
    E = np.zeros((m + 1, n + 1), dtype=np.int64)
    F = np.zeros((m + 1, n + 1), dtype=np.int64)
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            E[i, j] = max(H[i, j - 1] + go, E[i, j - 1] + ge, 0)
            F[i, j] = max(H[i - 1, j] + go, F[i - 1, j] + ge, 0)
            H[i, j] = max(H[i - 1, j - 1] + int(smat[a[i - 1], b[j - 1]]), E[i, j], F[i, j])
    return H[1:, 1:]


def strip_sweep(a, b, smat, go, ge, straight, reset_bottom):
    """Real cells H(i, j) (0-based) as the lane schedule computes them."""
    m, n = len(a), len(b)
    gap = lambda x, g: max(x + g, 0)  # the local saturating gap term
    ST = GG * R
    S = -(-m // ST)
    steps = 8 * (-(-(n + GG) // 8))
    out = np.full((m, n), -10**9, dtype=np.int64)
    ring_h = np.zeros(steps + GG + 8, dtype=np.int64)
    ring_f = np.zeros(steps + GG + 8, dtype=np.int64)
    hbot = [0] * GG
    fbot = [0] * GG
    for s in range(S):
        Hr = [[0] * R for _ in range(GG)]
        Er = [[0] * R for _ in range(GG)]
        HoL = [[gap(0, go)] * R for _ in range(GG)]
        hd0 = [0] * GG
        if reset_bottom:
            hbot = [0] * GG
            fbot = [0] * GG
        new_ring_h = np.zeros_like(ring_h)
        new_ring_f = np.zeros_like(ring_f)
        for t in range(steps):
            nhbot, nfbot = list(hbot), list(fbot)
            for l in range(GG):
                j = t - l
                if l == 0:
                    hup = int(ring_h[j]) if s > 0 else 0
                    fup = int(ring_f[j]) if s > 0 else 0
                else:
                    hup, fup = hbot[l - 1], fbot[l - 1]
                if not (straight or 0 <= j < n):
                    continue
                hd, hou, F = hd0[l], gap(hup, go), fup
                for k in range(R):
                    i = s * ST + l * R + k
                    sc = int(smat[a[i], b[j]]) if (i < m and 0 <= j < n) else PAD
                    hds = hd + sc
                    En = max(HoL[l][k], gap(Er[l][k], ge))
                    Fn = max(hou, gap(F, ge))
                    Hn = max(hds, En, Fn)
                    hd = Hr[l][k]
                    Hr[l][k], Er[l][k], HoL[l][k] = Hn, En, gap(Hn, go)
                    F, hou = Fn, HoL[l][k]
                    if i < m and 0 <= j < n:
                        out[i, j] = Hn
                nhbot[l], nfbot[l], hd0[l] = Hr[l][R - 1], F, hup
                if l == GG - 1 and 0 <= j < len(new_ring_h):
                    new_ring_h[j], new_ring_f[j] = Hr[l][R - 1], F
            hbot, fbot = nhbot, nfbot
        ring_h, ring_f = new_ring_h, new_ring_f
    return out


def case(seed, m=37, n=29):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 20, m).astype(np.uint8)
    b = np.concatenate([a[5:5 + n // 2], rng.integers(0, 20, n - n // 2)]).astype(np.uint8)
    return a, b


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_masked_schedule_is_the_dp(seed):
    smat = po.protein_matrix()[0]
    a, b = case(seed)
    ref = row_major_local(a, b, smat, -11, -1)
    got = strip_sweep(a, b, smat, -11, -1, straight=False, reset_bottom=False)
    assert (got == ref).all()
    want = po.align(po.params(mode=1, gap_open=-11, gap_extend=-1, traceback=0), a, b)
    assert int(ref.max()) == want[0]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_straight_line_first_blocks_break_on_stale_bottom_rows(seed):
    smat = po.protein_matrix()[0]
    a, b = case(seed)
    ref = row_major_local(a, b, smat, -11, -1)
    bad = strip_sweep(a, b, smat, -11, -1, straight=True, reset_bottom=False)
    wrong = np.argwhere(bad != ref)
    assert len(wrong) > 0
    # only from the second strip on: the first strip's stale values are the boundary
    assert wrong[:, 0].min() >= GG * R
    # with the bottom rows reset at each strip start the straight-line form agrees
    fixed = strip_sweep(a, b, smat, -11, -1, straight=True, reset_bottom=True)
    assert (fixed == ref).all()
