// moist_scalar.h — scalar transport and final update of one RK stage
// (moist.jl:1469-1539).  Included by moist_kernels.h inside namespace mo.
//
// A block owns a tile of ST_X columns x ST_Z levels.  For each of qv, qc, qr
// it stages q = rho_q / rho and rho of the tile plus a 2-cell halo in LDS,
// evaluates every face flux once (Koren advective flux, moist.jl:850-945, and
// the diffusive flux, moist.jl:1494-1510), and updates each cell from the
// differences of its face fluxes.  The reference evaluates each face flux
// from both neighbouring cells with the same expression, so computing it once
// gives the same numbers.  All neighbour values come from the stage input
// (snapshot order), never from cells already updated.

constexpr int ST_X = 64, ST_Z = 8, ST_THREADS = 256;
constexpr int ST_HX = ST_X + 4, ST_HZ = ST_Z + 4;  // haloed tile

template <class S>
__global__ void __launch_bounds__(ST_THREADS) k_scalar(const Dom *__restrict__ doms, const int4 *__restrict__ tiles,
                                                       ZVec zv, SSet<S> in, SSet<S> s0, SSet<S> out, Work<S> w,
                                                       const S *__restrict__ rup_fin, const S *__restrict__ rtp_fin,
                                                       int div, double dt_ovr, int stage) {
    using R = S;
    __shared__ R l_rho[ST_HZ][ST_HX];
    __shared__ R l_q[ST_HZ][ST_HX];
    __shared__ R l_fx[ST_Z][ST_X + 1];   // advective flux at x-face i0+f-1/2
    __shared__ R l_dx[ST_Z][ST_X + 1];   // diffusive flux at the same faces
    __shared__ R l_fz[ST_Z + 1][ST_X];   // advective flux at z-face k0+g-1/2
    __shared__ R l_dz[ST_Z + 1][ST_X];   // diffusive flux at the same faces

    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD ordering (uniform per block)
    const int m = t.x;
    const Dom &D = doms[m];
    const int i0 = t.y, k0 = t.z;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    const int tid = threadIdx.x;
    const double *zd = zv.z + D.offz, *dzd = zv.dz + D.offz;
    const R dx = R(D.dx), kap = R(D.kappa), EPS = K<R>::EPS;
    const R dt = stage_dt<R>(D, div, dt_ovr);
    const R dtL = R(D.dt);  // the positivity limiters use the large step (moist.jl:917,930)
    auto gidx = [&](int k, int i) -> int64_t { return D.off2 + (int64_t)k * nx + i; };

    // rho of the haloed tile (full density), once for the three scalars
    for (int e = tid; e < ST_HZ * ST_HX; e += ST_THREADS) {
        const int r = e / ST_HX, cidx = e - r * ST_HX;
        const int k = k0 - 2 + r, i = wrapx(i0 - 2 + cidx, nx);
        R v = R(0.0);
        if (k >= 0 && k < nz) v = full_rho(in, gidx(k, i), zv, D, k);
        l_rho[r][cidx] = v;
    }

    const S *dph[3] = {w.d_qv, w.d_qc, w.d_qr};
    int nan_f = 0;
    for (int sc = 0; sc < 3; ++sc) {
        const S *qraw = in.f[QV + sc];
        __syncthreads();  // l_rho ready; previous scalar's fluxes consumed
        for (int e = tid; e < ST_HZ * ST_HX; e += ST_THREADS) {
            const int r = e / ST_HX, cidx = e - r * ST_HX;
            const int k = k0 - 2 + r, i = wrapx(i0 - 2 + cidx, nx);
            R v = R(0.0);
            if (k >= 0 && k < nz) v = fdiv(qraw[gidx(k, i)], l_rho[r][cidx]);
            l_q[r][cidx] = v;
        }
        __syncthreads();
        // x faces: face f lies between cells i0+f-1 and i0+f (compute_limited_flux_x at ii = i0+f-1)
        for (int e = tid; e < ST_Z * (ST_X + 1); e += ST_THREADS) {
            const int r = e / (ST_X + 1), f = e - r * (ST_X + 1);
            const int k = k0 + r;
            R fx = R(0.0), dfx = R(0.0);
            if (k < nz) {
                const int ii = wrapx(i0 + f - 1, nx);
                const R rhou_face = w.mom_x[gidx(k, ii)];
                const int lr = r + 2, lc = f + 1;  // LDS column of cell ii
                const R s_i = l_q[lr][lc], s_ip1 = l_q[lr][lc + 1];
                R s_face;
                if (rhou_face > EPS) {
                    const R s_im1 = l_q[lr][lc - 1];
                    const R r_num = s_ip1 - s_i, r_den = s_i - s_im1;
                    const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
                    s_face = s_i + R(0.5) * koren(rr) * r_den;
                } else if (rhou_face < -EPS) {
                    const R s_ip2 = l_q[lr][lc + 2];
                    const R r_num = s_i - s_ip1, r_den = s_ip1 - s_ip2;
                    const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
                    s_face = s_ip1 + R(0.5) * koren(rr) * r_den;
                } else {
                    s_face = R(0.5) * (s_i + s_ip1);
                }
                fx = s_face * rhou_face;
                // rho * dq/dx at the face (moist.jl:1494-1495)
                dfx = fdiv((l_rho[lr][lc + 1] + l_rho[lr][lc]) * R(0.5) * (s_ip1 - s_i), dx);
            }
            l_fx[r][f] = fx;
            l_dx[r][f] = dfx;
        }
        // z faces: face g lies between levels kk = k0+g-1 and kk+1 (compute_limited_flux_z at kk)
        for (int e = tid; e < (ST_Z + 1) * ST_X; e += ST_THREADS) {
            const int g = e / ST_X, c = e - g * ST_X;
            const int kk = k0 + g - 1, i = i0 + c;
            R fz = R(0.0), dfz = R(0.0);
            if (i < nx && kk >= 0 && kk <= ke - 1) {
                const R rhow_face = w.mom_z[gidx(kk, i)];
                const int lr = g + 1, lc = c + 2;  // LDS row of level kk
                const R s_k = l_q[lr][lc], s_kp1 = l_q[lr + 1][lc];
                const R dzk = R(dzd[kk]), dzk1 = R(dzd[kk + 1]);
                R s_face;
                if (rhow_face > EPS) {
                    if (kk == 0) {
                        s_face = f2h(s_k, s_kp1, dzk, dzk1);
                        // linear index into rho_q: row kk of the first halo column
                        s_face = jmin(s_face, fdiv(R(qraw[gidx(kk, D.lin_col)]), rhow_face * dtL));
                    } else {
                        const R s_km1 = l_q[lr - 1][lc];
                        const R r_num = s_kp1 - s_k, r_den = s_k - s_km1;
                        const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
                        s_face = s_k + R(0.5) * koren(rr) * r_den;
                    }
                } else if (rhow_face < -EPS) {
                    if (kk == ke - 1) {
                        s_face = f2h(s_k, s_kp1, dzk, dzk1);
                        s_face = jmin(s_face, fdiv(R(qraw[gidx(kk + 1, D.lin_col)]), -rhow_face * dtL));
                    } else {
                        const R s_kp2 = l_q[lr + 2][lc];
                        const R r_num = s_k - s_kp1, r_den = s_kp1 - s_kp2;
                        const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
                        s_face = s_kp1 - R(0.5) * koren(rr) * r_den;
                    }
                } else {
                    s_face = f2h(s_k, s_kp1, dzk, dzk1);
                }
                fz = s_face * rhow_face;
                // rho * dq/dz at the face (moist.jl:1507-1508)
                dfz = fdiv(f2h(l_rho[lr + 1][lc], l_rho[lr][lc], dzk1, dzk) * (s_kp1 - s_k),
                           R(zd[kk + 1]) - R(zd[kk]));
            }
            l_fz[g][c] = fz;
            l_dz[g][c] = dfz;
        }
        __syncthreads();
        // cells (moist.jl:1480-1516)
        for (int e = tid; e < ST_Z * ST_X; e += ST_THREADS) {
            const int r = e / ST_X, c = e - r * ST_X;
            const int k = k0 + r, i = i0 + c;
            if (k >= nz || i >= nx) continue;
            const int64_t q = gidx(k, i);
            const R dzk = R(dzd[k]);
            const R adv_x = fdiv(l_fx[r][c + 1] - l_fx[r][c], dx);
            const R adv_z = fdiv(l_fz[r + 1][c] - l_fz[r][c], dzk);
            const R adv = -(adv_x + adv_z);
            const R diff_x = fdiv(kap * (l_dx[r][c + 1] - l_dx[r][c]), dx);
            const R diff_z = fdiv(kap * (l_dz[r + 1][c] - l_dz[r][c]), dzk);
            const R v = s0.f[QV + sc][q] + dt * (adv + (diff_x + diff_z) + dph[sc][q]);
            out.f[QV + sc][q] = v;
            if (isnan(v) && !nan_f) nan_f = 5 + sc;
        }
    }

    // final update of rho, rho_u, rho_theta, rho_w (moist.jl:1522-1539) and NaN checks
    for (int e = tid; e < ST_Z * ST_X; e += ST_THREADS) {
        const int r = e / ST_X, c = e - r * ST_X;
        const int k = k0 + r, i = i0 + c;
        if (k >= nz || i >= nx) continue;
        const int64_t q = gidx(k, i);
        const R vr = in.f[RHO][q] + w.rp[q];
        const R vu = in.f[RHOU][q] + rup_fin[q];
        const R vt = in.f[RT][q] + rtp_fin[q];
        const R vw = k < ke ? in.f[RHOW][q] + w.rwp[q] : in.f[RHOW][q];
        out.f[RHO][q] = vr;
        out.f[RHOU][q] = vu;
        out.f[RT][q] = vt;
        out.f[RHOW][q] = vw;
        int f = 0;
        if (isnan(vr)) f = 1;
        else if (isnan(vu)) f = 2;
        else if (isnan(vw)) f = 3;
        else if (isnan(vt)) f = 4;
        if (f) report_nan(w.nan_key, ((*w.step_ctr * 3 + stage) << 4) | f);
        // stale halo values surface_fluxes! reads at the next step (moist.jl:723,736)
        if (k == 0 && i == nx - 1) w.sfh[2 * m + 0] = double(in.f[RHOU][q]);
        if (k == 0 && i == 0) w.sfh[2 * m + 1] = double(full_rho(in, q, zv, D, 0));
    }
    // scalar NaN codes (5..7) come after the state fields of the same stage
    if (nan_f) report_nan(w.nan_key, ((*w.step_ctr * 3 + stage) << 4) | nan_f);
}
