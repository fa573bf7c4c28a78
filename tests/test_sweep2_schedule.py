"""The schedule of the two-sweep kernel (k_rb_stream3t in
basilisk_amd/csrc/bmg_kernels.h), run on the CPU: per step s of the stream
along i, A's red cells of plane s, A's black cells of plane s-1, B's red
cells of plane s-2 and B's black cells of plane s-3 for the column parity
red on plane s; a ring of width 3 around each tile; stages from three planes
below a chunk to three above; eight plane slots with a halo of 4; sigma
times each A value written into the ghosts of wall-adjacent cells.

The threads of each step run in a random order and the slots start as NaN,
so a read that the step's one barrier does not order, or a read of a value
never written, changes the result.  The model equals two plain red-black
sweeps (ghost = sigma times the cell before the sweep) bit for bit.  The
kernel's tile is 16 x 64; the schedule does not depend on the tile size, so
small tiles keep this fast."""
import numpy as np
import pytest

RING, HALO = 3, 4


def relax_nd(D2, res, lam, al, ah, ap, am):
    nn = -D2 * res
    dd = -lam * D2
    for d in range(3):
        nn = nn + ah[d] * ap[d] + al[d] * am[d]
        dd = dd + ah[d] + al[d]
    return nn / dd


def coefs(A, i, j, k):
    ax, ay, az, res, lam = A
    return (res[i, j, k], lam[i, j, k], (ax[i, j, k], ay[i, j, k], az[i, j, k]),
            (ax[i + 1, j, k], ay[i, j + 1, k], az[i, j, k + 1]))


def two_sweeps(d, A, sig, D2):
    n = d.shape[0]
    d = d.copy()

    def val(i, j, k, own):
        for v, lo, hi in ((i, 1, 0), (j, 3, 2), (k, 5, 4)):
            if v < 0:
                return sig[lo] * own
            if v >= n:
                return sig[hi] * own
        return d[i, j, k]

    for _ in range(2):
        for colour in (0, 1):
            new = {}
            for i, j, k in np.ndindex(n, n, n):
                if (i + j + k) & 1 != colour:
                    continue
                own = d[i, j, k]
                res, lam, al, ah = coefs(A, i, j, k)
                ap = (val(i + 1, j, k, own), val(i, j + 1, k, own), val(i, j, k + 1, own))
                am = (val(i - 1, j, k, own), val(i, j - 1, k, own), val(i, j, k - 1, own))
                new[(i, j, k)] = relax_nd(D2, res, lam, al, ah, ap, am)
            for key, v in new.items():
                d[key] = v
    return d


def block(dold, dnew, A, sig, D2, TJ, TK, nchunks, ct, jt, kt, rng):
    n = dold.shape[0]
    RJ, RK = TJ + 2 * HALO, TK + 2 * HALO
    i0, i1 = ct * n // nchunks, (ct + 1) * n // nchunks
    j0, k0 = jt * TJ, kt * TK
    buf = np.full((8, RJ, RK), np.nan)
    slot = lambda p: (p - (i0 - HALO) + 16) & 7
    clampn = lambda v: min(max(v, 0), n - 1)

    def commit(p):
        P = buf[slot(p)]
        for row in range(RJ):
            for c2 in range(RK // 2):
                gj, gk = j0 - HALO + row, k0 - HALO + 2 * c2
                mk = min(max(gk, 0), n - 2)
                wx, wy = dold[clampn(p), clampn(gj), mk], dold[clampn(p), clampn(gj), mk + 1]
                if p < 0 or p >= n:
                    s = sig[1 if p < 0 else 0]
                    wx, wy = s * wx, s * wy
                elif gj < 0 or gj >= n:
                    s = sig[3 if gj < 0 else 2]
                    wx, wy = s * wx, s * wy
                elif gk < 0:
                    wy = sig[5] * wx
                elif gk >= n:
                    wx = sig[4] * wy
                P[row, 2 * c2], P[row, 2 * c2 + 1] = wx, wy

    # the tile's pairs, then the ring's (rows outside the tile over the
    # region's width, then the pairs left and right of each tile row)
    pairs = [(j0 + r, k0 + 2 * c) for r in range(TJ) for c in range(TK // 2)]
    for r in range(2 * RING):
        for c in range(RK // 2):
            pairs.append((j0 - RING + r if r < RING else j0 + TJ + r - RING, k0 - HALO + 2 * c))
    for r in range(TJ):
        for side in range(4):
            pairs.append((j0 + r, k0 - HALO + 2 * side if side < 2 else k0 + TK + 2 * (side - 2)))
    dist = lambda x, lo, ln: lo - x if x < lo else (x - (lo + ln - 1) if x >= lo + ln else 0)
    info = []
    for pj, pk in pairs:
        inside = 0 <= pj < n and 0 <= pk < n
        wj = dist(pj, j0, TJ)
        info.append((pj, pk, max(wj, dist(pk, k0, TK)) if inside else 99,
                     max(wj, dist(pk + 1, k0, TK)) if inside else 99))

    def relax_at(p, pj, kc):
        qj, qk, sl = pj - j0 + HALO, kc - k0 + HALO, slot(p)
        assert 1 <= qj <= RJ - 2 and 1 <= qk <= RK - 2
        ap = (buf[slot(p + 1)][qj, qk], buf[sl][qj + 1, qk], buf[sl][qj, qk + 1])
        am = (buf[slot(p - 1)][qj, qk], buf[sl][qj - 1, qk], buf[sl][qj, qk - 1])
        res, lam, al, ah = coefs(A, p, pj, kc)
        return relax_nd(D2, res, lam, al, ah, ap, am)

    def put_a(p, pj, kc, v):
        qj, qk = pj - j0 + HALO, kc - k0 + HALO
        P = buf[slot(p)]
        P[qj, qk] = v
        if pj == 0:
            P[qj - 1, qk] = sig[3] * v
        if pj == n - 1:
            P[qj + 1, qk] = sig[2] * v
        if kc == 0:
            P[qj, qk - 1] = sig[5] * v
        if kc == n - 1:
            P[qj, qk + 1] = sig[4] * v
        if p == 0:
            buf[slot(-1)][qj, qk] = sig[1] * v
        if p == n - 1:
            buf[slot(n)][qj, qk] = sig[0] * v

    lo3, hi3 = max(0, i0 - 3), min(n - 1, i1 + 2)
    lo2, hi2 = max(0, i0 - 2), min(n - 1, i1 + 1)
    lo1, hi1 = max(0, i0 - 1), min(n - 1, i1)
    commit(i0 - 4)
    commit(i0 - 3)
    for s in range(i0 - 3, i1 + 3):
        commit(s + 1)  # then the step's barrier
        for t in rng.permutation(len(info)):
            pj, pk, w0, w1 = info[t]
            e = (s + pj + pk) & 1
            w, kc = (w1 if e else w0), pk + e
            if w <= 3 and lo3 <= s <= hi3:
                put_a(s, pj, kc, relax_at(s, pj, kc))
            if w <= 2 and lo2 <= s - 1 <= hi2:
                put_a(s - 1, pj, kc, relax_at(s - 1, pj, kc))
            if w <= 1 and lo1 <= s - 2 <= hi1:
                buf[slot(s - 2)][pj - j0 + HALO, kc - k0 + HALO] = relax_at(s - 2, pj, kc)
            if w == 0 and i0 <= s - 3 < i1:
                q = s - 3
                dnew[q, pj, kc] = relax_at(q, pj, kc)
                dnew[q, pj, pk + 1 - e] = buf[slot(q)][pj - j0 + HALO, pk + 1 - e - k0 + HALO]


@pytest.mark.parametrize("n,TJ,TK,nchunks", [(16, 4, 8, 2), (16, 8, 8, 4), (12, 4, 4, 3), (16, 4, 8, 1)])
@pytest.mark.parametrize("sig", [(1, -1, -1, 1, 1, -1), (-1,) * 6, (1,) * 6])
def test_schedule_equals_two_sweeps(n, TJ, TK, nchunks, sig):
    rng = np.random.default_rng(n + TJ + nchunks)
    d = rng.standard_normal((n, n, n))
    A = (np.exp(0.3 * rng.standard_normal((n + 1, n, n))), np.exp(0.3 * rng.standard_normal((n, n + 1, n))),
         np.exp(0.3 * rng.standard_normal((n, n, n + 1))), rng.standard_normal((n, n, n)),
         -rng.uniform(0.5, 2.0, (n, n, n)))
    dnew = np.full_like(d, np.nan)
    for ct in range(nchunks):
        for jt in range(n // TJ):
            for kt in range(n // TK):
                block(d, dnew, A, sig, 0.01, TJ, TK, nchunks, ct, jt, kt, rng)
    assert np.array_equal(dnew, two_sweeps(d, A, sig, 0.01))


def test_schedule_model_catches_a_missing_ghost_update():
    """Without the ghost writes of the A stages, B reads sigma times the
    value before A across the walls, and the model no longer equals two
    sweeps: the comparison above can fail."""
    import inspect
    n, TJ, TK = 12, 4, 4
    rng = np.random.default_rng(3)
    d = rng.standard_normal((n, n, n))
    A = (np.ones((n + 1, n, n)), np.ones((n, n + 1, n)), np.ones((n, n, n + 1)), rng.standard_normal((n, n, n)),
         -np.ones((n, n, n)))
    sig = (1, -1, -1, 1, 1, -1)
    # block() with put_a's in-plane (j) and plane (i) ghost writes dropped
    src = inspect.getsource(block)
    src = src.replace("        if pj == 0:\n", "        if False:\n", 1).replace("        if p == 0:\n", "        if False:\n", 1)
    ns = dict(globals())
    exec(compile(src, "block_without_ghosts", "exec"), ns)
    dnew = np.full_like(d, np.nan)
    for ct in range(2):
        for jt in range(n // TJ):
            for kt in range(n // TK):
                ns["block"](d, dnew, A, sig, 0.01, TJ, TK, 2, ct, jt, kt, rng)
    assert not np.array_equal(dnew, two_sweeps(d, A, sig, 0.01))
