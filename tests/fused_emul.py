"""CPU emulation of the temporally blocked kernel's schedule (lbm-cuda_amd/csrc/lbm_fused.hip, k_fused).

Test infrastructure: it replays, for every workgroup, the kernel's sweep -- ring slots, per-step row lags
and column/row margins, the load / step / write-back order inside an iteration -- vectorised over the strip
columns, with bounds assertions on every ring and lattice index, and an FP64 collision from
oracle/pyref.py.  Comparing its output with k FP64 oracle AA steps checks the schedule (which values feed
which collision) on the CPU before the kernel runs on a GPU.
"""
import numpy as np

from oracle import pyref

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])
F_SOLID, F_COMPUTE, F_INLET, F_OUTADJ = 1, 2, 4, 8


class Sched:
    def __init__(self, K, P0, LW=128):
        self.K, self.P0, self.LW = K, P0, LW
        self.D0 = 1 + P0
        self.LASTB = (P0 + K - 1) & 1
        self.DLAST = self.D0 + 2 * (K - 1)
        self.DOUT = self.DLAST + 1 + self.LASTB
        self.R = self.DOUT + 1
        self.HALO = 2 * self.nodd_before(K)
        self.CW = self.LW - 2 * self.HALO

    def d(self, s):
        return self.D0 + 2 * s

    def nodd_before(self, s):
        return ((self.P0 + s) >> 1) - (self.P0 >> 1)

    def margin(self, s):
        return 2 * self.nodd_before(s) + ((self.P0 + s) & 1)


def pick_chunk(W, H, k):
    CW = 128 - 2 * (k + 1)
    strips = (W + CW - 1) // CW
    ch = 256
    while ch > 16 and strips * ((H + ch - 1) // ch) < 512:
        ch //= 2
    return ch if ch < H else H


def _collide(g, inlet, uin, p):
    """FP64 collision of a [9, n] block (pyref's transcription of lbm_core4.cu / lbm_core_smemaa.cu)."""
    tau_inv, cs, model, clamp = p["tau_inv"], p["cs"], p["model"], p["clamp"]
    out = np.empty_like(g)
    fl = ~inlet
    if fl.any():
        gi = g[:, fl]
        rho = gi.sum(0)
        ux = (gi * CX[:, None]).sum(0) / rho
        uy = (gi * CY[:, None]).sum(0) / rho
        if model == 1:
            out[:, fl] = _bgk(gi, rho, ux, uy, tau_inv)
        else:
            if clamp:
                ux = np.clip(ux, -0.4, 0.4)
                uy = np.clip(uy, -0.4, 0.4)
                rho = np.clip(rho, 0.0, 1.5)
            om = np.full(rho.shape, tau_inv)
            if cs != 0.0:
                m7 = gi[1] - gi[2] + gi[3] - gi[4]
                m8 = gi[5] - gi[6] + gi[7] - gi[8]
                n7 = m7 - rho * (ux * ux - uy * uy)
                n8 = m8 - rho * ux * uy
                Q = np.sqrt(n7 * n7 + n8 * n8 + 1e-8)
                t0 = 1.0 / tau_inv
                om = 1.0 / (0.5 * (t0 + np.sqrt(t0 * t0 + 18.0 * cs * cs * Q / rho)))
            out[:, fl] = pyref._mrt(gi, rho, ux, uy, om, p["s_e"], p["s_eps"], p["s_q"])
    if inlet.any():
        gz, rz = pyref._zou_he(g[:, inlet], uin[inlet])
        u = uin[inlet]
        if model == 1:
            out[:, inlet] = _bgk(gz, rz, u, np.zeros_like(u), tau_inv)
        else:
            out[:, inlet] = pyref._mrt(gz, rz, u, np.zeros_like(u), tau_inv, p["s_e"], p["s_eps"], p["s_q"])
    return out


def _bgk(f, rho, ux, uy, tau_inv):
    usq = ux * ux + uy * uy
    out = np.empty_like(f)
    for q in range(9):
        cu = CX[q] * ux + CY[q] * uy
        feq = pyref.WQ[q] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
        out[q] = f[q] + tau_inv * (feq - f[q])
    return np.minimum(10.0, np.maximum(0.0, out))


def fused(f_in, mask, inlet_u, parity, K, p, periodic, strip=128, chunk=None):
    """k_fused on the whole grid; returns f_out (cells not written keep NaN)."""
    _, H, W = f_in.shape
    S = Sched(K, parity, strip)
    CH = pick_chunk(W, H, K) if chunk is None else min(chunk, H)
    out = np.full_like(f_in, np.nan)
    gx = (W + S.CW - 1) // S.CW
    gy_n = (H + CH - 1) // CH
    for bx in range(gx):
        for by in range(gy_n):
            _block(f_in, out, mask, inlet_u, S, CH, bx, by, p, periodic)
    return out


def _block(fin, fout, mask, inlet_u, S, CH, bx, by, p, periodic):
    _, H, W = fin.shape
    R, HALO, CW, K, P0, LW = S.R, S.HALO, S.CW, S.K, S.P0, S.LW
    sf = np.full((9, R, LW), np.nan)
    sflag = np.zeros((R, LW), np.uint8)
    xs = bx * CW - HALO
    y0 = by * CH
    ch = min(CH, H - y0)
    ybeg, yend = y0 - HALO, y0 + ch + HALO - 1
    niter = ch + HALO + S.DOUT
    cols = np.arange(LW)

    def load_row(j):
        y = ybeg + j
        slot = j % R
        assert 0 <= slot < R
        rowin = periodic or (0 <= y < H)
        gy = y % H if periodic else y
        x = xs + cols
        if periodic:
            x = x % W
            colin = np.ones(LW, bool)
        else:
            colin = (x >= 0) & (x < W)
        vals = np.zeros((9, LW))
        fl = np.zeros(LW, np.uint8)
        if rowin:
            xi = x[colin]
            assert xi.min(initial=0) >= 0 and xi.max(initial=0) < W and 0 <= gy < H
            vals[:, colin] = fin[:, gy, xi]
            m = np.zeros(LW, bool)
            m[colin] = mask[gy, xi] != 0
            if periodic:
                fl[:] = np.where(m, F_SOLID, F_COMPUTE)
            else:
                fl[colin & m] |= F_SOLID
                if 1 <= gy <= H - 2:
                    inner = colin & (x <= W - 2)
                    fl[inner & (x == 0)] |= F_COMPUTE | F_INLET
                    fl[inner & (x != 0) & ~m] |= F_COMPUTE
                    fl[inner & (x == W - 2)] |= F_OUTADJ
        sf[:, slot, :] = vals
        sflag[slot, :] = fl

    for i in range(niter):
        if i <= yend - ybeg:
            load_row(i)
        # all steps of this iteration read the ring state left by the previous iteration's barrier;
        # rows touched by different steps are disjoint, so applying them in order is equivalent
        touched = {}
        for s in range(K):
            jrow = i - S.d(s)
            m = S.margin(s)
            odd = (P0 + s) & 1
            if not (jrow >= m and ybeg + jrow <= yend - m):
                continue
            rows = (jrow - 1, jrow, jrow + 1) if odd else (jrow,)
            for r in rows:
                assert r not in touched, f"iteration {i}: row {r} touched by steps {touched[r]} and {s}"
                touched[r] = s
            c = np.arange(m, LW - m)
            sl = jrow % R
            fl = sflag[sl, c]
            comp = (fl & F_COMPUTE) != 0
            c = c[comp]
            fl = fl[comp]
            if c.size == 0:
                continue
            inlet = (fl & F_INLET) != 0
            y = ybeg + jrow
            gy = y % H if periodic else y
            uin = np.where(inlet, inlet_u[gy] if 0 <= gy < H else 0.0, 0.0)
            if not odd:
                g = sf[:, sl, c].copy()
                o = _collide(g, inlet, uin, p)
                sf[OPP[:, None], sl, c[None, :]] = o
                oa = (fl & F_OUTADJ) != 0
                if not periodic and oa.any():
                    cc = c[oa] + 1
                    keep = cc < LW
                    sf[OPP[:, None], sl, cc[keep][None, :]] = o[:, oa][:, keep]
            else:
                su, sd = (jrow + 1) % R, (jrow - 1) % R
                rs = [sl if CY[q] == 0 else (su if CY[q] > 0 else sd) for q in range(9)]
                assert c.min() >= 1 and c.max() <= LW - 2
                solid = np.zeros((9, c.size), bool)
                for q in range(1, 9):
                    sk = (not periodic) & inlet & (CX[q] < 0)
                    solid[q] = np.where(sk, False, (sflag[rs[q], c + CX[q]] & F_SOLID) != 0)
                g = np.empty((9, c.size))
                g[0] = sf[0, sl, c]
                for q in range(1, 9):
                    o_ = OPP[q]
                    v = np.where(solid[o_], sf[q, sl, c], sf[o_, rs[o_], c + CX[o_]])
                    if not periodic:
                        v = np.where(inlet & (CX[q] > 0), 0.0, v)
                    g[q] = v
                o = _collide(g, inlet, uin, p)
                sf[0, sl, c] = o[0]
                for q in range(1, 9):
                    drop = (inlet & (CX[q] < 0)) if not periodic else np.zeros(c.size, bool)
                    bb = solid[q] & ~drop
                    nb = ~solid[q] & ~drop
                    sf[OPP[q], sl, c[bb]] = o[q, bb]
                    sf[q, rs[q], c[nb] + CX[q]] = o[q, nb]
                oa = (fl & F_OUTADJ) != 0
                if not periodic and oa.any():
                    sf[3, sl, c[oa]] = o[3, oa]
                    nu = oa & ~solid[2]
                    sf[6, su, c[nu]] = o[6, nu]
                    nd = oa & ~solid[4]
                    sf[7, sd, c[nd]] = o[7, nd]
        assert i not in touched, f"iteration {i}: the row being loaded is also stepped"
        jout = i - S.DOUT
        if HALO <= jout < HALO + ch:
            assert jout not in touched
            sl = jout % R
            y = ybeg + jout
            cc = np.arange(CW)
            x = xs + HALO + cc
            ok = x < W
            if not periodic:
                ok &= x != W - 1  # the ghost is written by the owner of W-2
            assert 0 <= y < H
            fout[:, y, x[ok]] = sf[:, sl, HALO + cc[ok]]
            if not periodic:
                own = np.nonzero(ok & (x == W - 2))[0]
                if own.size:
                    c2 = HALO + own[0]
                    final_reversed = ((P0 + K) & 1) == 1
                    if final_reversed and 1 <= y <= H - 2:
                        fout[:, y, W - 1] = sf[:, sl, c2]
                    elif c2 + 1 < LW:
                        fout[:, y, W - 1] = sf[:, sl, c2 + 1]


def pair(f_in, mask, inlet_u, p, periodic, CH=8):
    """k_pair (lbm-cuda_amd/csrc/lbm_pair.hip) on the whole grid: per 128 x CH tile, phase A (even step of the
    130 x (CH + 2) region into an LDS image) and phase B (odd step of the core, scattered to the output).
    Returns f_out (slots not written keep NaN); every LDS and lattice index is bounds-checked."""
    _, H, W = f_in.shape
    CWp, RW, RH = 128, 130, CH + 2
    out = np.full_like(f_in, np.nan)
    written = np.zeros(f_in.shape, np.int32)
    for bx in range((W + CWp - 1) // CWp):
        for by in range((H + CH - 1) // CH):
            x0, y0 = bx * CWp, by * CH
            sf = np.full((9, RH, RW), np.nan)
            sm = np.ones((RH, RW), bool)
            for r in range(RH):
                for c in range(RW):
                    gx, gy = x0 - 1 + c, y0 - 1 + r
                    if periodic:
                        gx, gy, inside = gx % W, gy % H, True
                    else:
                        inside = 0 <= gx < W and 0 <= gy < H
                    solid = bool(mask[gy, gx]) if inside else True
                    sm[r, c] = solid
                    if periodic:
                        comp, inlet = not solid, False
                    else:
                        inlet = gx == 0
                        comp = inside and 1 <= gy <= H - 2 and gx <= W - 2 and (inlet or not solid)
                    if not comp:
                        continue
                    g = f_in[:, gy, gx].astype(np.float64)[:, None].copy()
                    o = _collide(g, np.array([inlet]), np.array([inlet_u[gy] if inlet else 0.0]), p)[:, 0]
                    sf[OPP, r, c] = o
                    if not periodic and gx == W - 2 and c + 1 < RW:
                        sf[OPP, r, c + 1] = o
            for r in range(1, CH + 1):
                for c in range(1, CWp + 1):
                    gx, gy = x0 + c - 1, y0 + r - 1
                    if gx >= W or gy >= H:
                        continue
                    inlet = False
                    if not periodic:
                        if gy == 0 or gy == H - 1 or gx == W - 1:
                            continue
                        inlet = gx == 0
                        if not inlet and sm[r, c]:
                            continue
                    elif sm[r, c]:
                        continue
                    solid = [False] + [bool(sm[r + CY[q], c + CX[q]]) and not (not periodic and inlet and CX[q] < 0)
                                       for q in range(1, 9)]
                    g = np.empty(9)
                    g[0] = sf[0, r, c]
                    for q in range(1, 9):
                        o = OPP[q]
                        rr, cc = (r, c) if solid[o] else (r - CY[q], c - CX[q])
                        assert 0 <= rr < RH and 0 <= cc < RW
                        g[q] = sf[q if solid[o] else o, rr, cc]
                    if inlet:
                        g[[1, 5, 8]] = 0.0  # arrive from x = -1: rebuilt by Zou-He, never read
                    o9 = _collide(g[:, None], np.array([inlet]), np.array([inlet_u[gy] if inlet else 0.0]), p)[:, 0]

                    def put(q, y, x, v):
                        assert 0 <= y < H and 0 <= x < W
                        out[q, y, x] = v
                        written[q, y, x] += 1

                    put(0, gy, gx, o9[0])
                    for q in range(1, 9):
                        back = solid[q] or (not periodic and inlet and CX[q] < 0)
                        if back:
                            put(OPP[q], gy, gx, o9[q])
                        else:
                            ty, tx = (gy + CY[q]) % H, (gx + CX[q]) % W
                            put(q, ty, tx, o9[q])
                    if not periodic and gx == W - 2:
                        put(3, gy, gx, o9[3])
                        if not solid[2]:
                            put(6, gy + 1, gx, o9[6])
                        if not solid[4]:
                            put(7, gy - 1, gx, o9[7])
    return out, written


def quad(f_in, mask, inlet_u, p, periodic, CH=8):
    """k_quad (lbm-cuda_amd/csrc/lbm_quad.hip) on the whole grid: per CW x CH tile with a 3-cell halo,
    A1 (even step of the (CW + 6) x (CH + 6) region into an LDS image), B1 (odd step of ring <= 2 in place in the
    image), A2 (even step of ring <= 1 in place), B2 (odd step of the core, scattered to the output).  Returns
    (f_out, writers): slots not written keep NaN.  Every LDS and lattice index is bounds-checked, and each
    in-place phase is checked race free: no LDS location is written by two cells, and no cell reads a
    location that another cell of the same phase writes."""
    _, H, W = f_in.shape
    CWq, R = (128 if CH in (8, 16) else 64), 3  # the kernel's shape table: 128 x {8, 16}, 64 x {12, 14, 20, 24, 32}
    RW, RH = CWq + 2 * R, CH + 2 * R
    out = np.full_like(f_in, np.nan)
    written = np.zeros(f_in.shape, np.int32)

    def coords(x0, y0, r, c):
        gx, gy = x0 - R + c, y0 - R + r
        if periodic:
            return gx % W, gy % H, True, gx, gy
        return gx, gy, 0 <= gx < W and 0 <= gy < H, gx, gy

    def updated(gx, gy, inside, solid):
        if periodic:
            return not solid, False
        inlet = gx == 0
        return inside and 1 <= gy <= H - 2 and 0 <= gx <= W - 2 and (inlet or not solid), inlet

    def col(g9, inlets, uins):
        return _collide(np.array(g9).T.copy(), np.array(inlets), np.array(uins), p)

    for bx in range((W + CWq - 1) // CWq):
        for by in range((H + CH - 1) // CH):
            x0, y0 = bx * CWq, by * CH
            sf = np.full((9, RH, RW), np.nan)
            sm = np.ones((RH, RW), bool)
            # ---- A1
            cells, g9, inl, uin = [], [], [], []
            for r in range(RH):
                for c in range(RW):
                    gx, gy, inside, _, _ = coords(x0, y0, r, c)
                    solid = bool(mask[gy, gx]) if inside else True
                    sm[r, c] = solid
                    comp, inlet = updated(gx, gy, inside, solid)
                    if comp:
                        cells.append((r, c, gx))
                        g9.append(f_in[:, gy, gx].astype(np.float64))
                        inl.append(inlet)
                        uin.append(inlet_u[gy] if inlet else 0.0)
            if cells:
                o = col(g9, inl, uin)
                for i, (r, c, gx) in enumerate(cells):
                    sf[OPP, r, c] = o[:, i]
                    if not periodic and gx == W - 2 and c + 1 < RW:
                        sf[OPP, r, c + 1] = o[:, i]

            def b_gather(r, c, gx, inlet):
                solid = [False] + [bool(sm[r + CY[q], c + CX[q]]) and not (not periodic and inlet and CX[q] < 0)
                                   for q in range(1, 9)]
                g, reads = np.empty(9), [(0, r, c)]
                g[0] = sf[0, r, c]
                for q in range(1, 9):
                    o = OPP[q]
                    loc = (q, r, c) if solid[o] else (o, r - CY[q], c - CX[q])
                    assert 0 <= loc[1] < RH and 0 <= loc[2] < RW
                    g[q] = sf[loc]
                    reads.append(loc)
                if inlet:
                    g[[1, 5, 8]] = 0.0  # arrive from x = -1: rebuilt by Zou-He, never read
                return g, solid, reads

            def phase_cells(ring):
                for r in range(R - ring, R + CH + ring):
                    for c in range(R - ring, R + CWq + ring):
                        gx, gy, inside, _, _ = coords(x0, y0, r, c)
                        comp, inlet = updated(gx, gy, inside, sm[r, c])
                        if comp:
                            yield r, c, gx, gy, inlet

            def race_free(reads, writes):
                owner = {}
                for i, ws in enumerate(writes):
                    for loc in ws:
                        assert loc not in owner, f"LDS {loc} written by two cells"
                        owner[loc] = i
                for i, rs in enumerate(reads):
                    for loc in rs:
                        assert owner.get(loc, i) == i, f"LDS {loc} read by one cell and written by another"

            # ---- B1: ring <= 2, in place
            cs_, g9, inl, uin, rds, sol = [], [], [], [], [], []
            for r, c, gx, gy, inlet in phase_cells(2):
                g, solid, reads = b_gather(r, c, gx, inlet)
                cs_.append((r, c, gx))
                g9.append(g)
                inl.append(inlet)
                uin.append(inlet_u[gy] if inlet else 0.0)
                rds.append(reads)
                sol.append(solid)
            wrs = []
            if cs_:
                o = col(g9, inl, uin)
                pend = []
                for i, (r, c, gx) in enumerate(cs_):
                    ws = [((0, r, c), o[0, i])]
                    for q in range(1, 9):
                        back = sol[i][q] or (not periodic and inl[i] and CX[q] < 0)
                        ws.append(((OPP[q], r, c) if back else (q, r + CY[q], c + CX[q]), o[q, i]))
                    if not periodic and gx == W - 2:
                        ws.append(((3, r, c), o[3, i]))
                        if not sol[i][2]:
                            ws.append(((6, r + 1, c), o[6, i]))
                        if not sol[i][4]:
                            ws.append(((7, r - 1, c), o[7, i]))
                    for loc, _ in ws:
                        assert 0 <= loc[1] < RH and 0 <= loc[2] < RW
                    wrs.append([loc for loc, _ in ws])
                    pend.append(ws)
                race_free(rds, wrs)
                for ws in pend:
                    for loc, v in ws:
                        sf[loc] = v
            # ---- A2: ring <= 1, in place (local)
            cs_, g9, inl, uin, rds, wrs = [], [], [], [], [], []
            for r, c, gx, gy, inlet in phase_cells(1):
                cs_.append((r, c, gx))
                g9.append(sf[:, r, c].copy())
                inl.append(inlet)
                uin.append(inlet_u[gy] if inlet else 0.0)
                rds.append([(q, r, c) for q in range(9)])
                w = [(q, r, c) for q in range(9)]
                if not periodic and gx == W - 2:
                    w += [(q, r, c + 1) for q in range(9)]
                wrs.append(w)
            if cs_:
                race_free(rds, wrs)
                o = col(g9, inl, uin)
                for i, (r, c, gx) in enumerate(cs_):
                    sf[OPP, r, c] = o[:, i]
                    if not periodic and gx == W - 2:
                        sf[OPP, r, c + 1] = o[:, i]
            # ---- B2: core -> output
            cs_, g9, inl, uin, sol = [], [], [], [], []
            for r in range(R, R + CH):
                for c in range(R, R + CWq):
                    gx, gy = x0 + c - R, y0 + r - R
                    if gx >= W or gy >= H:
                        continue
                    comp, inlet = updated(gx, gy, True, sm[r, c])
                    if not comp:
                        continue
                    g, solid, _ = b_gather(r, c, gx, inlet)
                    cs_.append((gx, gy))
                    g9.append(g)
                    inl.append(inlet)
                    uin.append(inlet_u[gy] if inlet else 0.0)
                    sol.append(solid)
            if not cs_:
                continue
            o = col(g9, inl, uin)

            def put(q, y, x, v):
                assert 0 <= y < H and 0 <= x < W
                out[q, y, x] = v
                written[q, y, x] += 1

            for i, (gx, gy) in enumerate(cs_):
                put(0, gy, gx, o[0, i])
                for q in range(1, 9):
                    back = sol[i][q] or (not periodic and inl[i] and CX[q] < 0)
                    if back:
                        put(OPP[q], gy, gx, o[q, i])
                    else:
                        put(q, (gy + CY[q]) % H, (gx + CX[q]) % W, o[q, i])
                if not periodic and gx == W - 2:
                    put(3, gy, gx, o[3, i])
                    if not sol[i][2]:
                        put(6, gy + 1, gx, o[6, i])
                    if not sol[i][4]:
                        put(7, gy - 1, gx, o[7, i])
    return out, written
