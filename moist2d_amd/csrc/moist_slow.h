// moist_slow.h — slow tendencies of one RK stage (moist.jl:1071-1332) and the
// acoustic start values (moist.jl:1339-1348).  Included by moist_kernels.h
// inside namespace mo.
//
// A block owns ST_X columns x ST_Z levels.  It stages u and w at faces (formed
// here from rho, rho_u and rho_w in the pair-layout variant, read from k_diag's
// output otherwise), theta, rho, rho_u and rho_w of the tile plus a 2-cell
// halo in LDS (theta, u and w take turns in one array, each written while the
// previous one's last phase runs, so a CU holds one more block), then evaluates
// every flux once at the face, centre or corner where the reference defines
// it and forms each tendency from flux differences:
//   R_rho_theta: x- and z-face fluxes (moist.jl:1287-1329)
//   R_rho_u:     centre momentum fluxes, corner fluxes, viscous fluxes (1102-1172)
//   R_rho_w:     corner fluxes with height weights (left of the face) and with
//                dz weights (right of the face) — the reference interpolates
//                rho_u differently on the two sides (moist.jl:1196, 1205), so
//                both are kept — centre fluxes and viscous fluxes (1193-1267)
// Each shared flux is the same expression the reference evaluates in both
// neighbouring cells, so evaluating it once gives the same numbers.

constexpr int SL_HX = ST_X + 4, SL_HZ = ST_Z + 4;

// 3rd-order upwind flux of the reference, w (-m + 5c + 2p) / 6 with
// (m, c, p) = (m0, c0, p0) when pos and (m1, c1, p1) when neg, else the
// centred value: the operands are selected and the expression evaluated
// once, so a wave whose points differ in direction does not run every
// branch.  The value is the branch's own, bit for bit.
template <class R>
__device__ __forceinline__ R upw_flux(R w, bool pos, bool neg, R m0, R c0, R p0, R m1, R c1, R p1, R centred) {
    const R m = pos ? m0 : m1, c = pos ? c0 : c1, p = pos ? p0 : p1;
    const R up = fdiv(w * upw3(m, c, p), R(6.0));
    return (pos || neg) ? up : centred;
}
constexpr int SL_FLUX = 5 * (ST_Z + 1) * (ST_X + 1);  // largest flux phase (w)

template <class S, bool P>
__global__ void __launch_bounds__(ST_THREADS) k_slow(const Dom *__restrict__ doms, const int4 *__restrict__ tiles,
                                                     ZVec zv, SSet<S> in, SSet<S> s0, Work<S> w, S *rtp0,
                                                     int u_phys_2d, int *step_bump) {
    using R = S;
    using KC = K<R>;
    // lX holds theta in the R_rho_theta phase, then u, then w
    __shared__ R lX[SL_HZ][SL_HX];
    // full rho of the tile with one more column and level than the others
    // (the pair-layout variant forms u and w at the faces from it)
    __shared__ R lR[SL_HZ + 1][SL_HX + 1];
    __shared__ R lRU[SL_HZ][SL_HX], lRW[SL_HZ][SL_HX];
    __shared__ R lF[SL_FLUX];
    __shared__ R lDZ[SL_HZ + 1], lZ[SL_HZ];  // dz of levels k0-2 .. k0+ST_Z+2, z of k0-2 .. k0+ST_Z+1

    // the first stage of a large step advances the step count of the NaN key
    // (no kernel of this launch reads it; the step's physics, which ran
    // before, ranks with the step's first stage)
    if (step_bump && blockIdx.x == 0 && threadIdx.x == 0 && *step_bump < STEP_CTR_MAX) *step_bump += 1;
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD ordering (uniform per block)
    const Dom &D = doms[t.x];
    const int i0 = t.y, k0 = t.z;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    const int tid = threadIdx.x;
    const double *zd = zv.z + D.offz, *zfd = zv.zf + D.offz, *dzd = zv.dz + D.offz;
    const R dx = R(D.dx), EPS = KC::EPS, nu = R(D.nu);
    auto gidx = [&](int k, int i) -> int64_t { return D.off2 + (int64_t)k * nx + i; };
    auto DZ = [&](int k) -> R { return lDZ[k - k0 + 2]; };
    auto Z = [&](int k) -> R { return lZ[k - k0 + 2]; };
    // the float form divides by dx as a times its reciprocal: taken once
    const R rdx = rcp_once(dx);
    // LDS accessors in absolute (k, i) of the tile neighbourhood
    auto U = [&](int k, int i) -> R { return lX[k - k0 + 2][i - i0 + 2]; };
    auto W = [&](int k, int i) -> R { return lX[k - k0 + 2][i - i0 + 2]; };
    auto TH = [&](int k, int i) -> R { return lX[k - k0 + 2][i - i0 + 2]; };
    auto RHL = [&](int k, int i) -> R { return lR[k - k0 + 2][i - i0 + 2]; };
    auto RU = [&](int k, int i) -> R { return lRU[k - k0 + 2][i - i0 + 2]; };
    auto RW = [&](int k, int i) -> R { return lRW[k - k0 + 2][i - i0 + 2]; };

    // Staged tile and every global value the cell phases read are loaded
    // up front (fixed trip counts), so the loads of a thread are in flight
    // together instead of one loop iteration at a time.  Fields are addressed
    // through the member's base pointer and 32-bit byte offsets.
    const int64_t o2 = D.off2;
    auto bo = [&](int k, int i) -> uint32_t { return (uint32_t)sizeof(S) * moff<false>(nx, k, i); };
    constexpr int NSL = (SL_HZ * SL_HX + ST_THREADS - 1) / ST_THREADS;
    constexpr int NSR = ((SL_HZ + 1) * (SL_HX + 1) + ST_THREADS - 1) / ST_THREADS;  // rho items
    static_assert(NSR == NSL, "the rho tile's extra column and level fit the staging items");
    constexpr int NCL = (ST_Z * ST_X + ST_THREADS - 1) / ST_THREADS;
    R xu[NSL], xw[NSL];  // staged u and w, held until lX takes them
    {
        R v[NSL][4];
#pragma unroll
        for (int u = 0; u < NSL; ++u) {
            // item u: cell (r, c) of the rho tile (SL_HX + 1 wide), and, when
            // inside the SL_HZ x SL_HX tile, the same cell of the other arrays
            const int e = tid + u * ST_THREADS;
            const int r = e / (SL_HX + 1), c = e - r * (SL_HX + 1);
            const int k = k0 - 2 + r, i = wrapx(i0 - 2 + c, nx);
            const bool okr = e < (SL_HZ + 1) * (SL_HX + 1) && k >= 0 && k < nz;
            const bool in_t = r < SL_HZ && c < SL_HX;
            const int kc = okr ? k : 0, ic = okr ? i : 0;
            const uint32_t q = bo(kc, ic);
            R rho = ldb(in.f[RHO] + o2, q);
            if constexpr (is_mixed<S>()) rho = rho + S(zv.rho_bar[D.offz + kc]);
            const R ru = ldb(in.f[RHOU] + o2, q), rw = ldb(in.f[RHOW] + o2, q);
            const R th = ldb(w.theta + o2, (uint32_t)sizeof(S) * moff<P>(nx, kc, ic));
            if constexpr (!P) {
                // k_diag's u and w at the faces, item u of the SL_HZ x SL_HX tile
                const int r2 = e / SL_HX, c2 = e - r2 * SL_HX;
                const int k2 = k0 - 2 + r2;
                const bool ok2 = e < SL_HZ * SL_HX && k2 >= 0 && k2 < nz;
                const uint32_t q2 = bo(ok2 ? k2 : 0, ok2 ? wrapx(i0 - 2 + c2, nx) : 0);
                xu[u] = ok2 ? ldb(w.uf + o2, q2) : R(0);
                xw[u] = ok2 ? ldb(w.wf + o2, q2) : R(0);
            }
            v[u][0] = okr && in_t ? th : R(0);
            v[u][1] = okr ? rho : R(0);
            v[u][2] = okr && in_t ? ru : R(0);
            v[u][3] = okr && in_t ? rw : R(0);
        }
        if (tid < SL_HZ + 1) {
            const int k = k0 - 2 + tid;
            const bool ok = k >= 0 && k < nz;
            lDZ[tid] = ok ? R(dzd[k]) : R(1);
            if (tid < SL_HZ) lZ[tid] = ok ? R(zd[k]) : R(0);
        }
#pragma unroll
        for (int u = 0; u < NSL; ++u) {
            const int e = tid + u * ST_THREADS;
            if (e >= (SL_HZ + 1) * (SL_HX + 1)) continue;
            const int r = e / (SL_HX + 1), c = e - r * (SL_HX + 1);
            (&lR[0][0])[e] = v[u][1];
            if (r < SL_HZ && c < SL_HX) {
                const int et = r * SL_HX + c;
                (&lX[0][0])[et] = v[u][0];
                (&lRU[0][0])[et] = v[u][2];
                (&lRW[0][0])[et] = v[u][3];
            }
        }
    }
    // interior cells of this thread: levels k0+cr0 and k0+cr0+1 (one level pair)
    // of column i0+cc0, so its outputs leave as level pairs
    static_assert(NCL == 2 && ST_Z * ST_X == 2 * ST_THREADS, "two interior cells per thread");
    const int cr0 = 2 * (tid / ST_X), cc0 = tid % ST_X;
    const int kc0 = k0 + cr0, ic0 = i0 + cc0;
    const bool pair_ok = kc0 < nz && ic0 < nx;
    // byte offset of the thread's level pair in the work fields' layout P
    const uint32_t bpw = (uint32_t)sizeof(S) * moff<P>(nx, min(kc0, ke), min(ic0, nx - 1));
    auto st_pair = [&](S *f, S a, S b) {
        if constexpr (P) stb2(f + o2, bpw, a, b);
        else {
            stb(f + o2, bpw, a);
            if (kc0 + 1 < nz) stb(f + o2, bpw + (uint32_t)sizeof(S) * (uint32_t)nx, b);
        }
    };
    // per interior cell of this thread: the cell phases' global inputs
    R c_drho[NCL], c_drt[NCL], c_rp[NCL], c_rup[NCL], c_rwp[NCL], c_rtp[NCL];
    R c_pres[NCL], c_presx[NCL], c_presz[NCL], c_dup[NCL], c_rho[NCL], c_rhoz[NCL], c_zf[NCL], c_bb[NCL];
#pragma unroll
    for (int u = 0; u < NCL; ++u) {
        const int r = cr0 + u, c = cc0;
        const int k = min(k0 + r, ke), i = min(i0 + c, nx - 1);
        const int kz = min(k + 1, ke);
        const uint32_t q = bo(k, i), qx = bo(k, wrapx(i + 1, nx)), qz = bo(kz, i);
        c_drho[u] = ldb(w.d_rho + o2, q);
        c_drt[u] = ldb(w.d_rt + o2, q);
        c_rp[u] = ldb(s0.f[RHO] + o2, q) - ldb(in.f[RHO] + o2, q);
        c_rup[u] = ldb(s0.f[RHOU] + o2, q) - ldb(in.f[RHOU] + o2, q);
        c_rwp[u] = ldb(s0.f[RHOW] + o2, q) - ldb(in.f[RHOW] + o2, q);
        c_rtp[u] = ldb(s0.f[RT] + o2, q) - ldb(in.f[RT] + o2, q);
        c_pres[u] = ldb(w.pres + o2, q);
        c_presx[u] = ldb(w.pres + o2, qx);
        c_presz[u] = ldb(w.pres + o2, qz);
        c_dup[u] = u_phys_2d ? ldb(w.d_u2d + o2, q) : (k == 0 ? w.d_u_sf[D.offx + i] : R(0.0));
        c_rho[u] = ldb(in.f[RHO] + o2, q);
        c_rhoz[u] = ldb(in.f[RHO] + o2, qz);
        c_zf[u] = R(zfd[k]);
        c_bb[u] = R(zv.B_bar[D.offz + k]);
    }
    __syncthreads();
    if constexpr (P) {
        // u and w at the faces from the staged rho, rho_u, rho_w and the rho of
        // the neighbours i+1 and k+1 (the stage diagnostics' expressions,
        // moist.jl:1091-1097, 1184-1189)
#pragma unroll
        for (int u = 0; u < NSL; ++u) {
            const int e = tid + u * ST_THREADS;
            const int r = e / SL_HX, c = e - r * SL_HX;
            const int k = k0 - 2 + r;
            const bool ok = e < SL_HZ * SL_HX && k >= 0 && k < nz;
            R uf = R(0), wf = R(0);
            if (ok) {
                const R rho = lR[r][c], rho_ip = lR[r][c + 1], rho_kp = lR[r + 1][c];
                uf = fdiv(lRU[r][c] * R(2.0), rho + rho_ip);
                wf = k < nz - 1 ? fdiv(lRW[r][c], f2h(rho, rho_kp, lDZ[r], lDZ[r + 1])) : R(0.0);
            }
            xu[u] = uf;
            xw[u] = wf;
        }
    }

    // ---------------------------------------------------------------- R_rho_theta
    R *FTx = lF;                          // [ST_Z][ST_X+1]: x-face between i0+f-1 and i0+f
    R *FTz = lF + ST_Z * (ST_X + 1);      // [ST_Z+1][ST_X]: z-face between k0+g-1 and k0+g
    for (int e = tid; e < ST_Z * (ST_X + 1); e += ST_THREADS) {
        const int r = e / (ST_X + 1), f = e - r * (ST_X + 1);
        const int k = k0 + r, L = i0 + f - 1;
        R fl = R(0);
        if (k < nz) {
            const R ru = RU(k, L);
            const R t_m = TH(k, L - 1), t_c = TH(k, L), t_p = TH(k, L + 1), t_pp = TH(k, L + 2);
            fl = upw_flux(ru, ru > EPS, ru < -EPS, t_m, t_c, t_p, t_pp, t_p, t_c, ru * (t_c + t_p) * R(0.5));
        }
        FTx[e] = fl;
    }
    for (int e = tid; e < (ST_Z + 1) * ST_X; e += ST_THREADS) {
        const int g = e / ST_X, c = e - g * ST_X;
        const int k = k0 + g - 1, i = i0 + c;
        R fl = R(0);
        if (k >= 0 && k <= ke - 1) {
            const R rw = RW(k, i);
            const R t_m = TH(k - 1, i), t_c = TH(k, i), t_p = TH(k + 1, i), t_pp = TH(k + 2, i);
            fl = upw_flux(rw, rw > EPS && k > 0, rw < -EPS && k < ke - 1, t_m, t_c, t_p, t_pp, t_p, t_c,
                          rw * f2h(t_c, t_p, DZ(k), DZ(k + 1)));
        }
        FTz[e] = fl;
    }
    __syncthreads();
    // theta's last reader is done: lX takes u for the R_rho_u phase
#pragma unroll
    for (int u = 0; u < NSL; ++u) {
        const int e = tid + u * ST_THREADS;
        if (e < SL_HZ * SL_HX) (&lX[0][0])[e] = xu[u];
    }
    {
        R o_rr[NCL], o_rt[NCL], o_mx[NCL], o_mz[NCL];
#pragma unroll
        for (int u = 0; u < NCL; ++u) {
            const int r = cr0 + u, c = cc0;
            const int k = k0 + r, i = i0 + c;
            o_rr[u] = o_rt[u] = o_mx[u] = o_mz[u] = R(0.0);
            if (k >= nz || i >= nx) continue;
            // R_rho (moist.jl:1079-1086)
            const R fkp = k == ke ? R(0.0) : RW(k, i);
            const R fkm = k == 0 ? R(0.0) : RW(k - 1, i);
            o_rr[u] = -(fdiv_r(RU(k, i) - RU(k, i - 1), dx, rdx) + fdiv(fkp - fkm, DZ(k))) + c_drho[u];
            // R_rho_theta
            const R adv = -(fdiv_r(FTx[r * (ST_X + 1) + c + 1] - FTx[r * (ST_X + 1) + c], dx, rdx) +
                            fdiv(FTz[(r + 1) * ST_X + c] - FTz[r * ST_X + c], DZ(k)));
            o_rt[u] = adv + c_drt[u];
            // mass-flux accumulators start from the stage input (moist.jl:1339-1348)
            o_mx[u] = RU(k, i);
            o_mz[u] = RW(k, i);
        }
        if (pair_ok) {
            st_pair(w.R_rho, o_rr[0], o_rr[1]);
            st_pair(w.R_rt, o_rt[0], o_rt[1]);
            // perturbations from the step start (moist.jl:1339-1348)
            st_pair(w.rp, c_rp[0], c_rp[1]);
            st_pair(w.rup, c_rup[0], c_rup[1]);
            st_pair(w.rwp, c_rwp[0], c_rwp[1]);
            st_pair(rtp0, c_rtp[0], c_rtp[1]);
            // the pair-layout acoustic kernel's first substep reads the stage
            // input itself for the accumulators' start
            if constexpr (!P) {
                st_pair(w.mom_x, o_mx[0], o_mx[1]);
                st_pair(w.mom_z, o_mz[0], o_mz[1]);
            }
        }
    }
    __syncthreads();

    // ---------------------------------------------------------------- R_rho_u
    R *CUx = lF;                                // [ST_Z][ST_X+1]: momentum flux at centre (k, i0+cc)
    R *GUx = lF + ST_Z * (ST_X + 1);            // [ST_Z][ST_X+1]: rho du/dx at centre (k, i0+cc)
    R *CUz = lF + 2 * ST_Z * (ST_X + 1);        // [ST_Z+1][ST_X]: corner flux at (k0+g-1/2, i+1/2)
    R *GUz = CUz + (ST_Z + 1) * ST_X;           // [ST_Z+1][ST_X]: rho du/dz at the same corners
    for (int e = tid; e < ST_Z * (ST_X + 1); e += ST_THREADS) {
        const int r = e / (ST_X + 1), cc = e - r * (ST_X + 1);
        const int k = k0 + r, i = i0 + cc;  // centre (k, i) lies between faces i-1/2 and i+1/2
        R fl = R(0), gx = R(0);
        if (k < nz) {
            const R rc = R(0.5) * (RU(k, i - 1) + RU(k, i));
            const R u_mm = U(k, i - 2), u_m = U(k, i - 1), u_c = U(k, i), u_p = U(k, i + 1);
            fl = upw_flux(rc, rc > EPS, rc < -EPS, u_mm, u_m, u_c, u_p, u_c, u_m, rc * (u_m + u_c) * R(0.5));
            gx = fdiv_r(U(k, i) - U(k, i - 1), dx, rdx) * RHL(k, i);
        }
        CUx[e] = fl;
        GUx[e] = gx;
    }
    for (int e = tid; e < (ST_Z + 1) * ST_X; e += ST_THREADS) {
        const int g = e / ST_X, c = e - g * ST_X;
        const int k = k0 + g - 1, i = i0 + c;  // corner (k+1/2, i+1/2)
        R fl = R(0), gz = R(0);
        if (k >= 0 && k <= ke - 1) {
            const R rw = R(0.5) * (RW(k, i) + RW(k, i + 1));
            const R u_m = U(k - 1, i), u_c = U(k, i), u_p = U(k + 1, i), u_pp = U(k + 2, i);
            fl = upw_flux(rw, rw > EPS && k > 0, rw < -EPS && k < ke - 1, u_m, u_c, u_p, u_pp, u_p, u_c,
                          rw * f2h(u_c, u_p, DZ(k), DZ(k + 1)));
            const R u_kp = fdiv(U(k + 1, i) - U(k, i), Z(k + 1) - Z(k));
            const R rho_kp = R(0.25) * (RHL(k, i) + RHL(k + 1, i) + RHL(k, i + 1) + RHL(k + 1, i + 1));
            gz = u_kp * rho_kp;
        }
        CUz[e] = fl;
        GUz[e] = gz;
    }
    __syncthreads();
    // u's last reader is done: lX takes w for the R_rho_w phase
#pragma unroll
    for (int u = 0; u < NSL; ++u) {
        const int e = tid + u * ST_THREADS;
        if (e < SL_HZ * SL_HX) (&lX[0][0])[e] = xw[u];
    }
    R o_ru[NCL];
#pragma unroll
    for (int u = 0; u < NCL; ++u) {
        const int r = cr0 + u, c = cc0;
        const int k = k0 + r, i = i0 + c;
        o_ru[u] = R(0.0);
        if (k >= nz || i >= nx) continue;
        const int x0 = r * (ST_X + 1) + c;
        const R druudx = fdiv_r(CUx[x0 + 1] - CUx[x0], dx, rdx);
        const R drwudz = fdiv(CUz[(r + 1) * ST_X + c] - CUz[r * ST_X + c], DZ(k));
        const R adv = -druudx - drwudz;
        const R u_xx = fdiv_r(GUx[x0 + 1] - GUx[x0], dx, rdx);
        const R u_zz = fdiv(GUz[(r + 1) * ST_X + c] - GUz[r * ST_X + c], DZ(k));
        const R diff = nu * (u_xx + u_zz);
        const R dpdx = fdiv_r(c_presx[u] - c_pres[u], dx, rdx);
        o_ru[u] = adv - dpdx + diff + c_dup[u];
    }
    if (pair_ok) st_pair(w.R_rhou, o_ru[0], o_ru[1]);
    __syncthreads();

    // ---------------------------------------------------------------- R_rho_w
    const int NX1 = ST_Z * (ST_X + 1);
    R *XZ = lF;                  // [ST_Z][ST_X+1] corner (k+1/2, i0+f-1/2), rho_u weighted by heights
    R *XD = lF + NX1;            // same corners, rho_u weighted by dz
    R *VX = lF + 2 * NX1;        // rho dw/dx at the same corners
    R *ZC = lF + 3 * NX1;        // [ST_Z+1][ST_X] centre flux at (k0+g, i): kp of face k0+g-1
    R *VZ = ZC + (ST_Z + 1) * ST_X;
    for (int e = tid; e < NX1; e += ST_THREADS) {
        const int r = e / (ST_X + 1), f = e - r * (ST_X + 1);
        const int k = k0 + r, L = i0 + f - 1;  // corner between columns L and L+1 on face k+1/2
        R xz = R(0), xd = R(0), vx = R(0);
        if (k <= ke - 1) {
            const R s_p = upw3(W(k, L - 1), W(k, L), W(k, L + 1));
            const R s_m = upw3(W(k, L + 2), W(k, L + 1), W(k, L));
            const R s_c = (W(k, L) + W(k, L + 1)) * R(0.5);
            const R uz = f2h(RU(k, L), RU(k + 1, L), Z(k), Z(k + 1));
            if (uz > EPS) xz = fdiv(uz * s_p, R(6.0));
            else if (uz < -EPS) xz = fdiv(uz * s_m, R(6.0));
            else xz = uz * s_c;
            const R ud = f2h(RU(k, L), RU(k + 1, L), DZ(k), DZ(k + 1));
            if (ud > EPS) xd = fdiv(ud * s_p, R(6.0));
            else if (ud < -EPS) xd = fdiv(ud * s_m, R(6.0));
            else xd = ud * s_c;
            const R w_ip = fdiv_r(W(k, L + 1) - W(k, L), dx, rdx);
            const R rho_ip = R(0.25) * (RHL(k, L) + RHL(k + 1, L) + RHL(k, L + 1) + RHL(k + 1, L + 1));
            vx = w_ip * rho_ip;
        }
        XZ[e] = xz;
        XD[e] = xd;
        VX[e] = vx;
    }
    for (int e = tid; e < (ST_Z + 1) * ST_X; e += ST_THREADS) {
        const int g = e / ST_X, c = e - g * ST_X;
        const int k = k0 + g - 1, i = i0 + c;  // face k, its kp centre is (k+1, i)
        R zc = R(0), vz = R(0);
        if (k >= 0 && k < ke - 1) {
            const R rw = f2h(RW(k, i), RW(k + 1, i), DZ(k), DZ(k + 1));
            const R w_m = W(k - 1, i), w_c = W(k, i), w_p = W(k + 1, i), w_pp = W(k + 2, i);
            zc = upw_flux(rw, rw > EPS && k > 0, rw < -EPS && k < ke - 2, w_m, w_c, w_p, w_pp, w_p, w_c,
                          rw * f2h(w_c, w_p, DZ(k), DZ(k + 1)));
            vz = fdiv(W(k + 1, i) - W(k, i), DZ(k + 1)) * RHL(k + 1, i);
        }
        ZC[e] = zc;
        VZ[e] = vz;
    }
    __syncthreads();
    const R zf_sp = R(D.zf_sponge), H = R(D.H);
    R o_rw[NCL];
#pragma unroll
    for (int u = 0; u < NCL; ++u) {
        const int r = cr0 + u, c = cc0;
        const int k = k0 + r, i = i0 + c;
        o_rw[u] = R(0.0);
        if (k >= nz || i >= nx || k == ke) continue;
        const int x0 = r * (ST_X + 1) + c;
        const R zz = Z(k + 1) - Z(k);
        const R drwudx = fdiv_r(XZ[x0 + 1] - XD[x0], dx, rdx);
        const R drwwdz = fdiv(ZC[(r + 1) * ST_X + c] - ZC[r * ST_X + c], zz);
        const R adv = -drwudx - drwwdz;
        const R w_xx = fdiv_r(VX[x0 + 1] - VX[x0], dx, rdx);
        const R w_zz = fdiv(VZ[(r + 1) * ST_X + c] - VZ[r * ST_X + c], zz);
        const R diff = nu * (w_xx + w_zz);
        R sigma = R(0.0);
        if (c_zf[u] > (R(1.0) - zf_sp) * H)
            sigma = (KC::GRAV / R(30.0)) * ((c_zf[u] - (R(1.0) - zf_sp) * H) / (zf_sp * H));
        const R sponge = -sigma * RW(k, i);
        const R dpdz = fdiv(c_presz[u] - c_pres[u], zz);
        R res;
        if constexpr (is_mixed<S>()) {
            // -dp/dz + buoyancy = B_bar - dp'/dz - g (rho'_k + rho'_k+1)/2
            const R buoy = -KC::GRAV * (c_rho[u] + c_rhoz[u]) * R(0.5);
            res = adv + ((c_bb[u] - dpdz) + buoy) + diff + sponge;
        } else {
            const R buoy = -KC::GRAV * (RHL(k, i) + RHL(k + 1, i)) * R(0.5);
            res = adv - dpdz + buoy + diff + sponge;
        }
        o_rw[u] = res;
    }
    if (pair_ok) st_pair(w.R_rhow, o_rw[0], o_rw[1]);
}
