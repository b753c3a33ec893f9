"""Lane-level CPU model of k_acoustic's two-level cross-chunk scans
(moist2d_amd/csrc/moist_acoustic.h): thread t holds chunk c = t / TX of
column jj = t % TX; a wave holds CPW = 64 / TX consecutive chunks (1 when
TX >= 64), combined with lane shuffles (HIP __shfl_up / __shfl_down
semantics: a lane whose source is outside the wave keeps its own value);
group aggregates go through an LDS array of 6 NT doubles and each thread
folds the groups below (or above) its own.  The model uses the kernel's
index arithmetic and checks, for every workgroup size and tile width the
kernel can be launched with, that
  * the Moebius scan gives the c' entering each chunk (S_c-1 applied to (0, 1)),
  * the forward affine scan gives the carry into each chunk,
  * the backward affine scan gives x at each chunk's first face, the carry
    from above, and x at the last face of the chunk below,
against sequential composition, and that every LDS read hits a slot written
in the same phase and stays inside the array.
"""
import numpy as np
import pytest

def shfl_up(v, delta):  # v: per-thread array of one wave (64)
    out = v.copy()
    for l in range(64):
        if l - delta >= 0: out[l] = v[l - delta]
    return out
def shfl_down(v, delta):
    out = v.copy()
    for l in range(64):
        if l + delta < 64: out[l] = v[l + delta]
    return out

def run(NT, TX, rng):
    nch = NT // TX
    CPW = 1 if TX >= 64 else 64 // TX
    G = nch // CPW
    t = np.arange(NT); c = t // TX; jj = t % TX
    cw = c % CPW; grp = c // CPW; gi = grp * TX + jj
    assert (gi < NT).all() and G * TX <= NT
    lds = np.full(6 * NT, np.nan)
    # per-chunk maps (y, P) forward; (z, Q) back; Moebius 2x2
    y0 = rng.normal(size=NT); P0 = rng.normal(size=NT) * 0.5
    z0 = rng.normal(size=NT); Q0 = rng.normal(size=NT) * 0.5
    zl0 = rng.normal(size=NT); ql0 = rng.normal(size=NT)
    M0 = rng.normal(size=(NT, 4))
    waves = NT // 64
    def wavewise(f, nout, *arrs):
        outs = [np.zeros(NT) for _ in range(nout)]
        for w in range(waves):
            sl = slice(64 * w, 64 * w + 64)
            res = f(*[a[sl] for a in arrs], cw[sl])
            assert len(res) == nout
            for o, r in zip(outs, res): o[sl] = r
        return outs
    # ---- Moebius
    def mob_inwave(m00, m01, m10, m11, cwv):
        m = [m00.copy(), m01.copy(), m10.copy(), m11.copy()]
        d = 1
        while d < CPW:
            b = [shfl_up(x, d * TX) for x in m]
            n0 = m[0]*b[0] + m[1]*b[2]; n1 = m[0]*b[1] + m[1]*b[3]
            n2 = m[2]*b[0] + m[3]*b[2]; n3 = m[2]*b[1] + m[3]*b[3]
            sc = 1.0 / np.maximum(np.maximum(abs(n0), abs(n1)), np.maximum(abs(n2), abs(n3)))
            upd = cwv >= d
            for x, n in zip(m, (n0, n1, n2, n3)): x[upd] = (n * sc)[upd]
            d <<= 1
        e = [np.where(cwv > 0, shfl_up(x, TX), idv) if CPW > 1 else np.full(64, idv) for x, idv in zip(m, (1.0, 0.0, 0.0, 1.0))]
        return m + e
    m00, m01, m10, m11, e00, e01, e10, e11 = wavewise(mob_inwave, 8, M0[:, 0], M0[:, 1], M0[:, 2], M0[:, 3])
    for th in range(NT):
        if cw[th] == CPW - 1:
            lds[4 * gi[th]:4 * gi[th] + 4] = [m00[th], m01[th], m10[th], m11[th]]
    cin = np.zeros(NT)
    for th in range(NT):
        u0, u1 = 0.0, 1.0
        for gp in range(grp[th]):
            Mx = lds[4 * (gp * TX + jj[th]):4 * (gp * TX + jj[th]) + 4]
            assert not np.isnan(Mx).any()
            v0 = Mx[0]*u0 + Mx[1]*u1; v1 = Mx[2]*u0 + Mx[3]*u1
            sc = 1 / max(abs(v0), abs(v1)); u0, u1 = v0 * sc, v1 * sc
        s0 = e00[th]*u0 + e01[th]*u1; s1 = e10[th]*u0 + e11[th]*u1
        cin[th] = s0 / s1 if c[th] > 0 else 0.0
    # reference: sequential composite per column
    for j in range(TX):
        S = np.eye(2)
        for ch in range(nch):
            th = ch * TX + j
            if ch > 0:
                v = S @ np.array([0.0, 1.0]); ref = v[0] / v[1]
                assert abs(cin[th] - ref) <= 1e-9 * max(1, abs(ref)), (NT, TX, ch, j, cin[th], ref)
            Mc = M0[th].reshape(2, 2)
            S = Mc @ S
            S /= abs(S).max()
    # ---- forward affine
    def fwd_inwave(y, P, cwv):
        y = y.copy(); P = P.copy(); d = 1
        while d < CPW:
            yp = shfl_up(y, d * TX); Pp = shfl_up(P, d * TX)
            upd = cwv >= d
            y2 = y + P * yp; P2 = P * Pp
            y[upd] = y2[upd]; P[upd] = P2[upd]; d <<= 1
        if CPW > 1:
            yq = shfl_up(y, TX); Pq = shfl_up(P, TX)
            ye = np.where(cwv > 0, yq, 0.0); Pe = np.where(cwv > 0, Pq, 1.0)
        else:
            ye = np.zeros(64); Pe = np.ones(64)
        return y, P, ye, Pe
    y, P, ye, Pe = wavewise(fwd_inwave, 4, y0, P0)
    for th in range(NT):
        if cw[th] == CPW - 1:
            lds[4 * NT + 2 * gi[th]] = y[th]; lds[4 * NT + 2 * gi[th] + 1] = P[th]
    din = np.zeros(NT)
    for th in range(NT):
        E = 0.0
        for gp in range(grp[th]):
            a = lds[4 * NT + 2 * (gp * TX + jj[th])]; b = lds[4 * NT + 2 * (gp * TX + jj[th]) + 1]
            assert not np.isnan(a) and not np.isnan(b)
            E = a + b * E
        din[th] = ye[th] + Pe[th] * E
    for j in range(TX):
        x = 0.0
        for ch in range(nch):
            th = ch * TX + j
            assert abs(din[th] - x) <= 1e-9 * max(1, abs(x)), ('fwd', NT, TX, ch, j, din[th], x)
            x = y0[th] + P0[th] * x
    # ---- back affine + x_below
    lds[:4 * NT] = np.nan  # Moebius area reused
    def back_inwave(y, P, zl, ql, cwv):
        y = y.copy(); P = P.copy(); d = 1
        while d < CPW:
            yn = shfl_down(y, d * TX); Pn = shfl_down(P, d * TX)
            upd = cwv + d < CPW
            y2 = y + P * yn; P2 = P * Pn
            y[upd] = y2[upd]; P[upd] = P2[upd]; d <<= 1
        if CPW > 1:
            zq = shfl_up(zl, TX); qq = shfl_up(ql, TX)
            zlb = np.where(cwv > 0, zq, 0.0); qlb = np.where(cwv > 0, qq, 1.0)
        else:
            zlb = np.zeros(64); qlb = np.ones(64)
        return y, P, zlb, qlb
    yb, Pb, zlb, qlb = wavewise(back_inwave, 4, z0, Q0, zl0, ql0)
    for th in range(NT):
        if cw[th] == 0:
            lds[2 * gi[th]] = yb[th]; lds[2 * gi[th] + 1] = Pb[th]
        if cw[th] == CPW - 1:
            lds[2 * NT + 2 * gi[th]] = zl0[th]; lds[2 * NT + 2 * gi[th] + 1] = ql0[th]
    F = np.zeros(NT)
    for th in range(NT):
        f = 0.0
        for gp in range(G - 1, grp[th], -1):
            a = lds[2 * (gp * TX + jj[th])]; b = lds[2 * (gp * TX + jj[th]) + 1]
            assert not np.isnan(a)
            f = a + b * f
        F[th] = f
    xfirst = yb + Pb * F
    xin = F.copy()
    if CPW > 1:
        for w in range(waves):
            sl = slice(64 * w, 64 * w + 64)
            xn = shfl_down(xfirst[sl], TX)
            xin[sl] = np.where(cw[sl] < CPW - 1, xn, F[sl])
    xbel = np.zeros(NT)
    for th in range(NT):
        zb, qb = zlb[th], qlb[th]
        if cw[th] == 0 and grp[th] > 0:
            zb = lds[2 * NT + 2 * ((grp[th] - 1) * TX + jj[th])]; qb = lds[2 * NT + 2 * ((grp[th] - 1) * TX + jj[th]) + 1]
            assert not np.isnan(zb)
        if c[th] > 0: xbel[th] = zb + qb * xfirst[th]
    for j in range(TX):
        x = 0.0
        xf = {}
        for ch in range(nch - 1, -1, -1):
            th = ch * TX + j
            assert abs(xin[th] - x) <= 1e-9 * max(1, abs(x)), ('xin', NT, TX, ch, j)
            x = z0[th] + Q0[th] * x
            xf[ch] = x
            assert abs(xfirst[th] - x) <= 1e-9 * max(1, abs(x)), ('xfirst', NT, TX, ch, j)
        for ch in range(1, nch):
            th = ch * TX + j; tb = (ch - 1) * TX + j
            ref = zl0[tb] + ql0[tb] * xf[ch]
            assert abs(xbel[th] - ref) <= 1e-9 * max(1, abs(ref)), ('xbelow', NT, TX, ch, j)
    return G

@pytest.mark.parametrize("NT", [256, 512, 1024])
@pytest.mark.parametrize("TX", [1, 2, 4, 8, 16, 32, 64, 128, 256])
def test_two_level_scans_match_sequential_composition(NT, TX):
    if TX > NT:
        pytest.skip("tile wider than the workgroup")
    G = run(NT, TX, np.random.default_rng(NT * 1000 + TX))
    assert G <= NT // 64 or TX >= 64
