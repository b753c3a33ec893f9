// moist_scalar.h — scalar transport and final update of one RK stage
// (moist.jl:1469-1539).  Included by moist_kernels.h inside namespace mo.
//
// A block owns a tile of ST_X columns x ST_Z levels.  The work fields it reads
// (mom_x, mom_z and the final perturbations) are in layout P.  In the
// pair-layout variant the last acoustic substep has already done the final
// update of rho, rho_u, rho_theta and rho_w; each thread's two interior cells
// are the levels of one level pair of a column, and, given dg_th, it writes
// the stage diagnostics of its output (theta and rt2pres in level pairs,
// pressure x-fastest), which the next stage then reads instead of running
// k_diag on it.  For each of qv, qc, qr
// it stages q = rho_q / rho and rho of the tile plus a 2-cell halo in LDS,
// evaluates every face flux once (Koren advective flux, moist.jl:850-945, and
// the diffusive flux, moist.jl:1494-1510), and updates each cell from the
// differences of its face fluxes.  The reference evaluates each face flux
// from both neighbouring cells with the same expression, so computing it once
// gives the same numbers.  All neighbour values come from the stage input
// (snapshot order), never from cells already updated.

constexpr int ST_X = 64, ST_Z = 8, ST_THREADS = 256;
constexpr int ST_HX = ST_X + 4, ST_HZ = ST_Z + 4;  // haloed tile

// Koren face value at x-face i+1/2 (compute_limited_flux_x, moist.jl:850-892)
// from q = rho_q / rho of cells i-1 .. i+2, q[0] being cell i (a row of the
// staged tile).  Both upwind branches are b + 1/2 phi(r) (b - a) with
// r = (c - b) / (b - a): (a, b, c) = (q_i-1, q_i, q_i+1) for w > EPS and
// (q_i+2, q_i+1, q_i) for w < -EPS.  SEL selects the operands and evaluates
// that expression once (a wave whose faces differ in sign does not run both
// branches; the float kernels are VALU-bound); the FP64 form branches as the
// reference does, with fewer registers.  Both give the same value bit for bit.
template <bool SEL, class R> __device__ __forceinline__ R face_x(R w, const R *q, R EPS) {
    const R s_i = q[0], s_ip1 = q[1];
    if constexpr (SEL) {
        const R s_im1 = q[-1], s_ip2 = q[2];
        const bool pos = w > EPS, neg = w < -EPS;
        const R a = pos ? s_im1 : s_ip2, b = pos ? s_i : s_ip1, c = pos ? s_ip1 : s_i;
        const R r_num = c - b, r_den = b - a;
        const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
        const R lim = b + R(0.5) * koren(rr) * r_den;
        return (pos || neg) ? lim : R(0.5) * (s_i + s_ip1);
    } else {
        if (w > EPS) {
            const R r_num = s_ip1 - s_i, r_den = s_i - q[-1];
            const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
            return s_i + R(0.5) * koren(rr) * r_den;
        } else if (w < -EPS) {
            const R r_num = s_i - s_ip1, r_den = s_ip1 - q[2];
            const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
            return s_ip1 + R(0.5) * koren(rr) * r_den;
        }
        return R(0.5) * (s_i + s_ip1);
    }
}

// Koren face value at z-face kk+1/2 (compute_limited_flux_z, moist.jl:895-945),
// q[0] being level kk of a column of the staged tile (rows ST_HX apart).  The
// negative branch is q_kk+1 - 1/2 phi(r) (q_kk+1 - q_kk+2) (Q3: minus, where x
// has plus); SEL evaluates b + h phi(r) (b - a) with h = +-1/2, and
// -(1/2 phi) (b - a) is exactly -(1/2 phi (b - a)).  At the bottom face for
// w > EPS and the top face for w < -EPS the dz-weighted value is capped by the
// positivity limiter, which divides rho_q read through the reference's linear
// index (row kk of the first halo column: q_lo, q_hi) by the large-step flux.
// rs = rcp_once(dzk + dzk1).
template <bool SEL, class R>
__device__ __forceinline__ R face_z(R w, const R *q, int kk, int ke, R dzk, R dzk1, R rs, R q_lo, R q_hi, R dtL,
                                    R EPS) {
    const R s_k = q[0], s_kp1 = q[ST_HX];
    if constexpr (SEL) {
        const R s_km1 = q[-ST_HX], s_kp2 = q[2 * ST_HX];
        const bool pos = w > EPS, neg = w < -EPS;
        const R fh = f2h_r(s_k, s_kp1, dzk, dzk1, rs);
        const R a = pos ? s_km1 : s_kp2, b = pos ? s_k : s_kp1, c = pos ? s_kp1 : s_k;
        const R r_num = c - b, r_den = b - a;
        const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
        const R h = pos ? R(0.5) : R(-0.5);
        const R lim = b + h * koren(rr) * r_den;
        const bool bdy = (pos && kk == 0) || (neg && kk == ke - 1);
        R v = (pos || neg) ? lim : fh;
        // the positivity cap's division only where it applies (the bottom
        // and top faces: most waves skip it)
        if (bdy) v = jmin(fh, fdiv(pos ? q_lo : q_hi, (pos ? w : -w) * dtL));
        return v;
    } else {
        if (w > EPS) {
            if (kk == 0) return jmin(f2h_r(s_k, s_kp1, dzk, dzk1, rs), fdiv(q_lo, w * dtL));
            const R r_num = s_kp1 - s_k, r_den = s_k - q[-ST_HX];
            const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
            return s_k + R(0.5) * koren(rr) * r_den;
        } else if (w < -EPS) {
            if (kk == ke - 1) return jmin(f2h_r(s_k, s_kp1, dzk, dzk1, rs), fdiv(q_hi, -w * dtL));
            const R r_num = s_k - s_kp1, r_den = s_kp1 - q[2 * ST_HX];
            const R rr = m_fabs(r_den) < EPS ? R(1.0) : fdiv(r_num, r_den);
            return s_kp1 - R(0.5) * koren(rr) * r_den;
        }
        return f2h_r(s_k, s_kp1, dzk, dzk1, rs);
    }
}

// 6 waves per SIMD in float (80 VGPRs), 4 in double (128)
template <class S, bool P>
__global__ void __launch_bounds__(ST_THREADS, is_mixed<S>() ? 6 : 4) k_scalar(const Dom *__restrict__ doms, const int4 *__restrict__ tiles,
                                                       ZVec zv, SSet<S> in, SSet<S> s0, SSet<S> out, Work<S> w,
                                                       const S *__restrict__ rup_fin, const S *__restrict__ rtp_fin,
                                                       int div, double dt_ovr, int stage, S *dg_th, S *dg_pres,
                                                       S *dg_rt2, SSet<double> wide) {
    using R = S;
    __shared__ R l_rho[ST_HZ][ST_HX];
    __shared__ R l_q[ST_HZ][ST_HX];
    __shared__ R l_ri[ST_HZ][ST_HX];  // 1 / rho for q = rho_q / rho (float form)
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
    // member-relative byte offsets: x-fastest fields (bo) and work fields in layout P (bw)
    const int64_t o2 = D.off2;
    auto bo = [&](int k, int i) -> uint32_t { return (uint32_t)sizeof(S) * moff<false>(nx, k, i); };
    auto bw = [&](int k, int i) -> uint32_t { return (uint32_t)sizeof(S) * moff<P>(nx, k, i); };
    constexpr uint32_t NOCELL = 0xffffffffu;

    // Work items of the strided loops, fixed per thread: staged cells of the
    // haloed tile, x faces, z faces and interior cells.  Their offsets and the
    // face mass fluxes (the same for the three scalars) are set up once, so
    // every loop issues its loads together.
    constexpr int NSTG = (ST_HZ * ST_HX + ST_THREADS - 1) / ST_THREADS;
    constexpr int NFX = (ST_Z * (ST_X + 1) + ST_THREADS - 1) / ST_THREADS;
    constexpr int NFZ = ((ST_Z + 1) * ST_X + ST_THREADS - 1) / ST_THREADS;
    constexpr int NCL = (ST_Z * ST_X + ST_THREADS - 1) / ST_THREADS;
    static_assert(NCL == 2 && ST_Z * ST_X == 2 * ST_THREADS && ST_THREADS % ST_X == 0, "two cells a thread");
    // interior cell u of this thread as a flat tile index e = r ST_X + c:
    // rows tid/ST_X and tid/ST_X + ST_THREADS/ST_X, or in the pair-layout
    // variant the two levels 2 (tid/ST_X), +1 of one column, so their
    // work-field accesses move level pairs

    uint32_t gst[NSTG];  // staged cell (NOCELL: outside the domain's levels or the tile)
#pragma unroll
    for (int u = 0; u < NSTG; ++u) {
        const int e = tid + u * ST_THREADS;
        const int r = e / ST_HX, cidx = e - r * ST_HX;
        const int k = k0 - 2 + r, i = wrapx(i0 - 2 + cidx, nx);
        gst[u] = (e < ST_HZ * ST_HX && k >= 0 && k < nz) ? bo(k, i) : NOCELL;
    }
    // the thread's interior cells: offsets (clamped into the member) and whether they exist
    uint32_t ocl[NCL];
    bool okc[NCL];
#pragma unroll
    for (int u = 0; u < NCL; ++u) {
        const int e = P ? (2 * (tid / ST_X) + u) * ST_X + tid % ST_X : tid + u * ST_THREADS;
        const int r = e / ST_X, c = e - r * ST_X;
        const int k = k0 + r, i = i0 + c;
        okc[u] = e < ST_Z * ST_X && k < nz && i < nx;
        ocl[u] = bo(min(k, ke), min(i, nx - 1));
    }
    R mfx[NFX], mfz[NFZ];  // mass flux at the x and z faces of this thread
    R fdz0[NFZ], fdz1[NFZ], fzz[NFZ], cdz[NCL];  // dz below / above each z face, z spacing, dz of each cell
    // their reciprocals for the float form's divisions (rcp_once), the same
    // for the three scalars
    R rsz[NFZ], rzz[NFZ], rcz[NCL];
    const R rdx = rcp_once(dx);
#pragma unroll
    for (int u = 0; u < NFX; ++u) {
        const int e = tid + u * ST_THREADS;
        const int r = e / (ST_X + 1), f = e - r * (ST_X + 1);
        const int k = k0 + r;
        const R v = ldb(w.mom_x + o2, bw(min(k, ke), wrapx(i0 + f - 1, nx)));
        mfx[u] = (e < ST_Z * (ST_X + 1) && k < nz) ? v : R(0.0);
    }
#pragma unroll
    for (int u = 0; u < NFZ; ++u) {
        const int e = tid + u * ST_THREADS;
        const int g = e / ST_X, c = e - g * ST_X;
        const int kk = k0 + g - 1, i = i0 + c;
        const bool ok = e < (ST_Z + 1) * ST_X && i < nx && kk >= 0 && kk <= ke - 1;
        const R v = ldb(w.mom_z + o2, bw(min(max(kk, 0), ke), min(i, nx - 1)));
        mfz[u] = ok ? v : R(0.0);
        const int kc = ok ? kk : 0, kc1 = ok ? kk + 1 : 0;
        fdz0[u] = R(dzd[kc]);
        fdz1[u] = R(dzd[kc1]);
        fzz[u] = R(zd[kc1]) - R(zd[kc]);
        rsz[u] = rcp_once(fdz0[u] + fdz1[u]);
        rzz[u] = rcp_once(fzz[u]);
    }
#pragma unroll
    for (int u = 0; u < NCL; ++u) {
        const int e = P ? (2 * (tid / ST_X) + u) * ST_X + tid % ST_X : tid + u * ST_THREADS;
        const int k = k0 + e / ST_X;
        cdz[u] = R(dzd[min(k, ke)]);
        rcz[u] = rcp_once(cdz[u]);
    }

    // rho of the haloed tile (full density), once for the three scalars, and
    // its reciprocal for q = rho_q / rho (float form)
    {
        R v[NSTG];
#pragma unroll
        for (int u = 0; u < NSTG; ++u) {
            const int e = tid + u * ST_THREADS;
            const int k = k0 - 2 + e / ST_HX;
            R r = gst[u] != NOCELL ? ldb(in.f[RHO] + o2, gst[u]) : R(0.0);
            if constexpr (is_mixed<S>()) r = gst[u] != NOCELL ? r + S(zv.rho_bar[D.offz + min(max(k, 0), ke)]) : r;
            v[u] = r;
        }
#pragma unroll
        for (int u = 0; u < NSTG; ++u) {
            const int e = tid + u * ST_THREADS;
            if (e < ST_HZ * ST_HX) {
                (&l_rho[0][0])[e] = v[u];
                if constexpr (is_mixed<R>()) (&l_ri[0][0])[e] = rcp_once(v[u]);
            }
        }
    }

    // the final update's inputs (read-only here), loaded with the rest.  In
    // the pair-layout variant the last acoustic substep has done the final
    // update of rho, rho_u, rho_theta and rho_w already; it reads back only the
    // updated rho and rho_theta, for the next stage's diagnostics.
    R fin[NCL][P ? 2 : 4];
    if constexpr (P) {
        const int iq = min(i0 + tid % ST_X, nx - 1);
#pragma unroll
        for (int u = 0; u < NCL; ++u) {
            const int k = min(k0 + 2 * (tid / ST_X) + u, ke);
            const uint32_t q = bo(k, iq);
            fin[u][0] = dg_th ? ldb(out.f[RHO] + o2, q) : R(0.0);
            fin[u][1] = dg_th ? ldb(out.f[RT] + o2, q) : R(0.0);
        }
    } else {
#pragma unroll
        for (int u = 0; u < NCL; ++u) {
            const int e = tid + u * ST_THREADS;
            const int r = e / ST_X, c = e - r * ST_X;
            const int k = min(k0 + r, ke), i = min(i0 + c, nx - 1);
            const uint32_t q = bo(k, i);
            fin[u][0] = ldb(in.f[RHO] + o2, q) + ldb(w.rp + o2, q);
            fin[u][1] = ldb(in.f[RHOU] + o2, q) + ldb(rup_fin + o2, q);
            fin[u][2] = ldb(in.f[RT] + o2, q) + ldb(rtp_fin + o2, q);
            fin[u][3] = k < ke ? ldb(in.f[RHOW] + o2, q) + ldb(w.rwp + o2, q) : ldb(in.f[RHOW] + o2, q);
        }
    }

    // wide (x-fastest float variant, the adjoint's recording pass; null
    // pointers: none): the output as full FP64 values, rho and rho_theta with
    // their base state added, for the reverse sweep
    constexpr bool WIDE = !P && is_mixed<S>();
    const S *dph[3] = {w.d_qv, w.d_qc, w.d_qr};
    R qn[3][P ? NCL : 1];  // the updated rho_q of this thread's cells (pair variant's epilogue)
    int nan_f = 0;
    // a scalar's global values: staged cells, limiter inputs, and the s0 value
    // and tendency of the interior cells.  The next scalar's are loaded before
    // this scalar's stores (a load issued after a store to an array the
    // compiler cannot tell apart would wait for the store to complete).
    R qv_[NSTG], q_lo, q_hi, s0v[NCL], dpv[NCL];
    R nq_lo = R(0.0), nq_hi = R(0.0), ns0v[NCL] = {}, ndpv[NCL] = {};
    auto load_scalar = [&](int sc, R *qv, R &lo, R &hi, R *s0x, R *dpx) {
        const S *qraw = in.f[QV + sc] + o2;
#pragma unroll
        for (int u = 0; u < NSTG; ++u) qv[u] = gst[u] != NOCELL ? ldb(qraw, gst[u]) : R(0.0);
        lo = ldb(qraw, bo(0, D.lin_col));
        hi = ldb(qraw, bo(ke, D.lin_col));
#pragma unroll
        for (int u = 0; u < NCL; ++u) {
            const R a = ldb(s0.f[QV + sc] + o2, ocl[u]), b = ldb(dph[sc] + o2, ocl[u]);
            s0x[u] = okc[u] ? a : R(0.0);
            dpx[u] = okc[u] ? b : R(0.0);
        }
    };
    load_scalar(0, qv_, q_lo, q_hi, s0v, dpv);
#pragma unroll
    for (int sc = 0; sc < 3; ++sc) {
        __syncthreads();  // l_rho ready; previous scalar's fluxes consumed
#pragma unroll
        for (int u = 0; u < NSTG; ++u) {
            const int e = tid + u * ST_THREADS;
            if (e < ST_HZ * ST_HX)
                (&l_q[0][0])[e] = gst[u] != NOCELL ? fdiv_r(qv_[u], (&l_rho[0][0])[e], (&l_ri[0][0])[e]) : R(0.0);
        }
        __syncthreads();
        // x faces: face f lies between cells i0+f-1 and i0+f (compute_limited_flux_x at ii = i0+f-1)
#pragma unroll
        for (int u = 0; u < NFX; ++u) {
            const int e = tid + u * ST_THREADS;
            if (e >= ST_Z * (ST_X + 1)) continue;
            const int r = e / (ST_X + 1), f = e - r * (ST_X + 1);
            const int k = k0 + r;
            R fx = R(0.0), dfx = R(0.0);
            if (k < nz) {
                const R rhou_face = mfx[u];
                const int lr = r + 2, lc = f + 1;  // LDS column of cell ii
                const R s_i = l_q[lr][lc], s_ip1 = l_q[lr][lc + 1];
                fx = face_x<is_mixed<S>()>(rhou_face, &l_q[lr][lc], EPS) * rhou_face;
                // rho * dq/dx at the face (moist.jl:1494-1495)
                dfx = fdiv_r((l_rho[lr][lc + 1] + l_rho[lr][lc]) * R(0.5) * (s_ip1 - s_i), dx, rdx);
            }
            l_fx[r][f] = fx;
            l_dx[r][f] = dfx;
        }
        // z faces: face g lies between levels kk = k0+g-1 and kk+1 (compute_limited_flux_z at kk)
#pragma unroll
        for (int u = 0; u < NFZ; ++u) {
            const int e = tid + u * ST_THREADS;
            if (e >= (ST_Z + 1) * ST_X) continue;
            const int g = e / ST_X, c = e - g * ST_X;
            const int kk = k0 + g - 1, i = i0 + c;
            R fz = R(0.0), dfz = R(0.0);
            if (i < nx && kk >= 0 && kk <= ke - 1) {
                const R rhow_face = mfz[u];
                const int lr = g + 1, lc = c + 2;  // LDS row of level kk
                const R s_k = l_q[lr][lc], s_kp1 = l_q[lr + 1][lc];
                const R dzk = fdz0[u], dzk1 = fdz1[u];
                fz = face_z<is_mixed<S>()>(rhow_face, &l_q[lr][lc], kk, ke, dzk, dzk1, rsz[u], q_lo, q_hi, dtL,
                                           EPS) *
                     rhow_face;
                // rho * dq/dz at the face (moist.jl:1507-1508)
                dfz = fdiv_r(f2h_r(l_rho[lr + 1][lc], l_rho[lr][lc], dzk1, dzk, rsz[u]) * (s_kp1 - s_k), fzz[u],
                             rzz[u]);
            }
            l_fz[g][c] = fz;
            l_dz[g][c] = dfz;
        }
        __syncthreads();
        if (sc < 2) load_scalar(sc + 1, qv_, nq_lo, nq_hi, ns0v, ndpv);
        // cells (moist.jl:1480-1516)
#pragma unroll
        for (int u = 0; u < NCL; ++u) {
            const int e = P ? (2 * (tid / ST_X) + u) * ST_X + tid % ST_X : tid + u * ST_THREADS;
            if (e >= ST_Z * ST_X) continue;
            const int r = e / ST_X, c = e - r * ST_X;
            const int k = k0 + r, i = i0 + c;
            if (k >= nz || i >= nx) continue;
            const R dzk = cdz[u];
            const R adv_x = fdiv_r(l_fx[r][c + 1] - l_fx[r][c], dx, rdx);
            const R adv_z = fdiv_r(l_fz[r + 1][c] - l_fz[r][c], dzk, rcz[u]);
            const R adv = -(adv_x + adv_z);
            const R diff_x = fdiv_r(kap * (l_dx[r][c + 1] - l_dx[r][c]), dx, rdx);
            const R diff_z = fdiv_r(kap * (l_dz[r + 1][c] - l_dz[r][c]), dzk, rcz[u]);
            const R v = s0v[u] + dt * (adv + (diff_x + diff_z) + dpv[u]);
            stb(out.f[QV + sc] + o2, ocl[u], v);
            if constexpr (WIDE) {
                if (wide.f[0]) wide.f[QV + sc][gidx(k, i)] = double(v);
            }
            if constexpr (P) {
                if (sc == 0) qn[0][u] = v;
                else if (sc == 1) qn[1][u] = v;
                else qn[2][u] = v;
            }
            if (isnan(v) && !nan_f) nan_f = 5 + sc;
        }
        q_lo = nq_lo;
        q_hi = nq_hi;
#pragma unroll
        for (int u = 0; u < NCL; ++u) {
            s0v[u] = ns0v[u];
            dpv[u] = ndpv[u];
        }
    }

    // final update of rho, rho_u, rho_theta, rho_w (moist.jl:1522-1539) and NaN
    // checks (the pair variant: done by the last acoustic substep)
    R dth[P ? NCL : 1], drt2[P ? NCL : 1];  // the next stage's theta and rt2pres (pair variant)
    if constexpr (P)
#pragma unroll
        for (int u = 0; u < NCL; ++u) dth[u] = drt2[u] = R(0.0);
#pragma unroll
    for (int u = 0; u < NCL; ++u) {
        const int e = P ? (2 * (tid / ST_X) + u) * ST_X + tid % ST_X : tid + u * ST_THREADS;
        const int r = e / ST_X, c = e - r * ST_X;
        const int k = k0 + r, i = i0 + c;
        if (e >= ST_Z * ST_X || k >= nz || i >= nx) continue;
        const uint32_t q = ocl[u];
        if constexpr (P) {
            // the next stage's diagnostics of the updated cell (dg_th null: none)
            if (dg_th) {
                R pres;
                diag_point<S>(D, zv, k, fin[u][0], fin[u][1], qn[0][u], qn[1][u], qn[2][u], dth[u], pres, drt2[u]);
                stb(dg_pres + o2, q, pres);
            }
        } else {
            const R vr = fin[u][0], vu = fin[u][1], vt = fin[u][2], vw = fin[u][3];
            stb(out.f[RHO] + o2, q, vr);
            stb(out.f[RHOU] + o2, q, vu);
            stb(out.f[RT] + o2, q, vt);
            stb(out.f[RHOW] + o2, q, vw);
            if constexpr (WIDE) {
                if (wide.f[0]) {
                    const int64_t g = gidx(k, i);
                    wide.f[RHO][g] = double(vr) + zv.rho_bar[D.offz + k];
                    wide.f[RHOU][g] = double(vu);
                    wide.f[RT][g] = double(vt) + zv.rt_bar[D.offz + k];
                    wide.f[RHOW][g] = double(vw);
                }
            }
            int f = 0;
            if (isnan(vr)) f = 1;
            else if (isnan(vu)) f = 2;
            else if (isnan(vw)) f = 3;
            else if (isnan(vt)) f = 4;
            if (f) report_nan(w.nan_key, ((*w.step_ctr * 3 + stage) << 4) | f);
        }
        // stale halo values surface_fluxes! reads at the next step (moist.jl:723,736)
        if (k == 0 && i == nx - 1) w.sfh[2 * m + 0] = double(in.f[RHOU][gidx(0, i)]);
        if (k == 0 && i == 0) w.sfh[2 * m + 1] = double(full_rho(in, gidx(0, 0), zv, D, 0));
    }
    if constexpr (P) {
        const int kq = k0 + 2 * (tid / ST_X), iq = i0 + tid % ST_X;
        if (dg_th && kq < nz && iq < nx) {
            stb2(dg_th + o2, 4u * moff<true>(nx, kq, iq), dth[0], dth[1]);
            stb2(dg_rt2 + o2, 4u * moff<true>(nx, kq, iq), drt2[0], drt2[1]);
        }
    }
    // scalar NaN codes (5..7) come after the state fields of the same stage
    if (nan_f) report_nan(w.nan_key, ((*w.step_ctr * 3 + stage) << 4) | nan_f);
}
