// moist_acoustic_pairs.h — the acoustic substep of the mixed form over
// work fields in the level-pair layout (moist.jl:1351-1461).  Included by
// moist_kernels.h inside namespace mo, after moist_acoustic.h.
//
// Same algorithm, thread mapping, chunked Thomas solve and cross-chunk scans
// as k_acoustic (whose header describes them), for S = float: a workgroup of
// NT threads owns TX columns, thread t column t % TX and chunk t / TX of L
// levels.  With TX = 16 an x-fastest FP32 row segment is 64 bytes, half a
// line; in the level-pair layout (levels 2m, 2m+1 adjacent) each lane moves
// two of its levels with one 8-byte access and 16 lanes fill a 128-byte line,
// so one load instruction touches half as many lines.  Phase 1 runs per pair:
// the pair's inputs in flight together, then its two levels, then the
// pair's stores of rho_u' and mom_x.  The FP64 form keeps k_acoustic: its
// 16-column rows already fill a line.
//
// FIN (the last substep of a stage): the stage's final update of rho, rho_u,
// rho_theta and rho_w (moist.jl:1522-1539) is done here, on the values the
// substep holds, and written to the stage output (x-fastest, AcBuf::fin_*)
// with its NaN checks (moist.jl:1543-1556); the final perturbations, which
// nothing reads after this substep, are not stored.
//
// Addressing: every field is read and written through the member's base
// pointer and 32-bit byte offsets (ldb/ldb2/stb/stb2); the thread's level
// pairs at columns i, i-1 and i+1 have one offset register each, shared by
// all the fields.

// 1/x in FP64 from the hardware reciprocal and two Newton steps (within an
// ulp of the IEEE quotient; a third of the instructions of an IEEE division).
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

// At least 4 waves per SIMD (at most 128 VGPRs): two 512-thread workgroups
// share a CU.
template <int L, int NT, int WPE, bool FIN>
__global__ void __launch_bounds__(NT, (WPE > 4 ? WPE : 4))
    k_acoustic_pairs(const Dom *__restrict__ doms, const int4 *__restrict__ atiles, ZVec zv, SSet<float> in,
                     Work<float> w, AcBuf<float> ab, int div, double dt_ovr, int nsub, int first) {
    using S = float;
    using R = float;
    static_assert(L % 2 == 0, "chunks are whole level pairs");
    __shared__ double lds[6 * NT];  // group aggregates of the cross-chunk scans
    __shared__ S ldt[2 * NT];
    extern __shared__ double lgeo[];  // 5 doubles per level of the tallest member (dynamic)

    const int4 tl = atiles[blockIdx.x];
    if (tl.x < 0) return;  // padding tile of the XCD ordering (uniform per block)
    const Dom &D = doms[tl.x];
    const int TX = tl.z, nch = tl.w;
    const int t = threadIdx.x;
    const int c = t / TX, jj = t - c * TX;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    const int i = tl.y + jj;
    const bool col_ok = i < nx;
    const int k0 = c * L;
    const int nl = (col_ok && k0 < nz) ? min(L, nz - k0) : 0;  // levels owned
    const int nf = max(0, min(nl, nz - 1 - k0));                // faces owned (j <= nz-2)
    const int im = col_ok ? wrapx(i - 1, nx) : 0, ip = col_ok ? wrapx(i + 1, nx) : 0;
    const double *zd = zv.z + D.offz, *dzd = zv.dz + D.offz;
    const R dx = R(D.dx);
    const R EPS = K<R>::EPS;
    const R dt_s = stage_dt<R>(D, div, dt_ovr) / R(nsub);
    const R gam = R(D.gamma_d);
    // Repeated divisors: the mixed form divides by the hardware reciprocal
    // (fdiv's float form), so taking each reciprocal once gives the same values.
    auto rcp_of = [](R b) -> R { return __builtin_amdgcn_rcpf(b); };
    auto DIV = [](R a, R, R rb) -> R { return a * rb; };
    const R rdx = rcp_of(dx), rns = rcp_of(R(nsub)), r6 = rcp_of(R(6.0));
    const R beta = R(D.beta);
    // fact_tp of the explicit part is a Float64 expression (moist.jl:1375)
    const double fact_tp_d = double(R(0.5)) * (1.0 + double(beta));
    // implicit_correction!'s own factors, in T (moist.jl:987-989)
    const R ftp = R(0.5) * (R(1.0) + beta);
    const R ftm = R(0.5) * (R(1.0) - beta);
    const double dt_tp2 = double(dt_s * ftp) * double(dt_s * ftp);
    const double fact = double(ftm / ftp);
    const int64_t cidx = (int64_t)c * ab.cin_stride + D.offx + i;
    double cin_v = 0.0;
    if (!first && nl > 0) cin_v = ab.cin[cidx];
    const int ic = col_ok ? i : 0;
    const int kev = (nz - 1) & ~1;  // highest level-pair start of the member
    // the member's fields (uniform base pointers)
    const int64_t o2 = D.off2;
    float *const f_mx = w.mom_x + o2, *const f_mz = w.mom_z + o2, *const f_rp = w.rp + o2;
    const float *const f_cur = ab.rtp_cur + o2, *const f_old = ab.rtp_old + o2, *const f_rt2 = w.rt2 + o2,
                       *const f_rup = ab.rup_in + o2, *const f_Ru = w.R_rhou + o2, *const f_Rr = w.R_rho + o2,
                       *const f_th = w.theta + o2, *const f_Rt = w.R_rt + o2, *const f_Rw = w.R_rhow + o2,
                       *const f_rwi = ab.rwp_in + o2, *const f_srw = in.f[RHOW] + o2;
    // byte offsets of the thread's level pairs at columns i, i-1, i+1
    const uint32_t rowb = 8u * (uint32_t)nx;  // bytes per level-pair row
    uint32_t bc[L / 2], bm[L / 2], bp[L / 2];
#pragma unroll
    for (int m = 0; m < L / 2; ++m) {
        const uint32_t r = (uint32_t)(min(k0 + 2 * m, kev) >> 1) * rowb;
        bc[m] = r + 8u * (uint32_t)ic;
        bm[m] = r + 8u * (uint32_t)im;
        bp[m] = r + 8u * (uint32_t)ip;
    }
    // byte offset of level k (clamped to the member) of column i, x-fastest
    auto bx = [&](int k) -> uint32_t { return 4u * ((uint32_t)min(max(k, 0), ke) * (uint32_t)nx + (uint32_t)ic); };

    // Loads of the later phases (after the first barrier) are issued before
    // it, right after phase 1's: the per-level geometry goes to LDS, the few
    // per-cell values to registers, so no global load waits behind a barrier.
    R rp_e[L], rtp_e[L], drw[L], thx[L + 4], rt2c[L + 1];
    R rwv[L + 1];  // rho_w' at faces k0 .. k0+L
    R swv[L + 1];  // input rho_w at levels k0-1 .. k0+L-1 (upwind branch tests)
    R mzv[L];      // mom_z at faces k0 .. k0+L-1
    R inr[FIN ? L : 1], inrt[FIN ? L : 1];  // FIN: the stage input's rho and rho_theta
    int nanf = 5;  // FIN: lowest NaN code of this thread's final values (5: none)
    double dd[L], cpv[L], pv[L];
#pragma unroll
    for (int q = 0; q < L + 4; ++q) thx[q] = R(0.0);
#pragma unroll
    for (int q = 0; q < L + 1; ++q) rt2c[q] = R(0.0);
    // later-phase inputs: rho_w' at the faces, mom_z, the input rho_w, issued
    // after phase 1 (before it they would stay live across it and spill)
    auto load_late = [&]() {
#pragma unroll
        for (int m = 0; m < L / 2; ++m) {
            R r0, r1, m0 = R(0.0), m1 = R(0.0);
            ldb2(f_rwi, bc[m], r0, r1);
            if (!first) ldb2(f_mz, bc[m], m0, m1);
            rwv[2 * m] = r0;
            rwv[2 * m + 1] = r1;
            mzv[2 * m] = 2 * m < nf ? m0 : R(0.0);
            mzv[2 * m + 1] = 2 * m + 1 < nf ? m1 : R(0.0);
        }
        rwv[L] = ldb(f_rwi, 4u * moff<true>(nx, min(k0 + L, ke), ic));
#pragma unroll
        for (int q = 0; q < L + 1; ++q) {
            const int j = k0 + q, kq = k0 - 1 + q;
            const R sv = ldb(f_srw, bx(kq));
            if (!(nl > 0 && j <= nz - 2)) rwv[q] = R(0.0);
            swv[q] = (nl > 0 && kq >= 0 && kq < nz) ? sv : R(0.0);
        }
        if (first) {
            // mom_z starts from the stage input's rho_w (moist.jl:1347), which
            // swv holds at the chunk's faces
#pragma unroll
            for (int l = 0; l < L; ++l) mzv[l] = l < nf ? swv[l + 1] : R(0.0);
        }
        if constexpr (FIN) {
#pragma unroll
            for (int q = 0; q < L; ++q) {
                const uint32_t qq = bx(k0 + q);
                inr[q] = ldb(in.f[RHO] + o2, qq);
                inrt[q] = ldb(in.f[RT] + o2, qq);
            }
        }
    };
    // ---- phase 1: rho_u', explicit rho' and rho_theta', d rho_w'
    // Inputs arrive as level pairs, one pair's loads in flight together.  The
    // loads of the next pair (after the last pair: the later phases' inputs
    // and the level geometry) are issued before the current pair's stores of
    // rho_u' and mom_x: a load issued after a store to an array the compiler
    // cannot tell apart waits for that store to complete.
    constexpr int GL = 2, NP = L / GL;
    auto ldg = [&](const R *p, uint32_t b, R *v) { ldb2(p, b, v[0], v[1]); };
    R tp_r0 = R(0.0), tp_t0 = R(0.0), tp_rp = R(0.0), tp_tp = R(0.0);
    R rrw[L];  // R_rho_w per level
    // the level heights the buoyancy of phase 1 uses, loaded with the first
    // pair's inputs (a load issued inside the pair's arithmetic would cost
    // its own memory round trip)
    double zpre[L];
#pragma unroll
    for (int l = 0; l < L; ++l) zpre[l] = zd[min(k0 + l, ke)];
    R mx_[NP][GL], cc_[NP][GL], cm_[NP][GL], cx_[NP][GL], oc_[NP][GL], om_[NP][GL], ox_[NP][GL];
    R r2c_[NP][GL], r2m_[NP][GL], r2x_[NP][GL], ruc_[NP][GL], rum_[NP][GL], Fuc_[NP][GL], Fum_[NP][GL];
    R Fr_[NP][GL], rp_[NP][GL], thc_[NP][GL], thm_[NP][GL], thp_[NP][GL], Ft_[NP][GL];
    R iru_[NP][GL];  // FIN: the stage input's rho_u of the pair's levels
    auto load_group = [&](int m) {
        if constexpr (FIN) {
            const int kp = min(k0 + 2 * m, kev);
            iru_[m][0] = ldb(in.f[RHOU] + o2, bx(kp));
            iru_[m][1] = ldb(in.f[RHOU] + o2, bx(kp + 1));
        }
        const uint32_t oc = bc[m], om = bm[m], op = bp[m];
        if (first) {
            // mom_x starts from the stage input's rho_u (moist.jl:1346): read
            // it there rather than from a copy (0 on the padding level of odd nz)
            const int kp = min(k0 + 2 * m, kev);
            mx_[m][0] = ldb(in.f[RHOU] + o2, bx(kp));
            mx_[m][1] = kp + 1 < nz ? ldb(in.f[RHOU] + o2, bx(kp + 1)) : R(0.0);
        } else {
            ldg(f_mx, oc, mx_[m]);
        }
        ldg(f_cur, oc, cc_[m]);
        ldg(f_cur, om, cm_[m]);
        ldg(f_cur, op, cx_[m]);
        ldg(f_old, oc, oc_[m]);
        ldg(f_old, om, om_[m]);
        ldg(f_old, op, ox_[m]);
        ldg(f_rt2, oc, r2c_[m]);
        ldg(f_rt2, om, r2m_[m]);
        ldg(f_rt2, op, r2x_[m]);
        ldg(f_rup, oc, ruc_[m]);
        ldg(f_rup, om, rum_[m]);
        ldg(f_Ru, oc, Fuc_[m]);
        ldg(f_Ru, om, Fum_[m]);
        ldg(f_Rr, oc, Fr_[m]);
        ldg(f_rp, oc, rp_[m]);
        ldg(f_th, oc, thc_[m]);
        ldg(f_th, om, thm_[m]);
        ldg(f_th, op, thp_[m]);
        ldg(f_Rt, oc, Ft_[m]);
        ldg(f_Rw, oc, rrw + GL * m);
    };
    // theta two levels below and above the chunk (the level pairs k0-2 and
    // k0+L), rt2pres one level above, and the first level of the per-level
    // geometry this thread stages (z, dz, 1/dz, 1/(z_k+1 - z_k),
    // g/(dz_k + dz_k+1)); issued with the later phases' inputs
    R a0, a1, b0, b1, r2top;
    double4 g0;
    double z0, dz0;
    const int q0 = min(t, nz - 1);
    auto load_rest = [&]() {
        load_late();
        ldb2(f_th, 4u * moff<true>(nx, max(k0 - 2, 0), ic), a0, a1);
        ldb2(f_th, 4u * moff<true>(nx, min(k0 + L, kev), ic), b0, b1);
        r2top = ldb(f_rt2, 4u * moff<true>(nx, min(k0 + L, ke), ic));
        g0 = zv.geo[D.offz + q0];
        z0 = zd[q0];
        dz0 = dzd[q0];
    };
    load_group(0);
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        R rui[GL], mxo[GL];  // rho_u' at face i+1/2 and mom_x of the pair
#pragma unroll
        for (int h = 0; h < GL; ++h) {
            const int l = GL * m + h;
            rp_e[l] = R(0.0);
            rtp_e[l] = R(0.0);
            drw[l] = R(0.0);
            rui[h] = R(0.0);
            mxo[h] = R(0.0);
            if (l < nl) {
                const R cur_c = cc_[m][h], cur_m = cm_[m][h], cur_p = cx_[m][h];
                const R old_c = oc_[m][h], old_m = om_[m][h], old_p = ox_[m][h];
                const R rt2_c = r2c_[m][h], rt2_m = r2m_[m][h], rt2_p = r2x_[m][h];
                // rho_u' at i+1/2 and i-1/2 (moist.jl:1362-1367)
                R gx = DIV(cur_p - cur_c, dx, rdx);
                R gxo = DIV(old_p - old_c, dx, rdx);
                gx = gx + gam * (gx - gxo);
                const R ru_i = ruc_[m][h] + dt_s * (Fuc_[m][h] - (rt2_c + rt2_p) * R(0.5) * gx);
                R gm = DIV(cur_c - cur_m, dx, rdx);
                R gmo = DIV(old_c - old_m, dx, rdx);
                gm = gm + gam * (gm - gmo);
                const R ru_m = rum_[m][h] + dt_s * (Fum_[m][h] - (rt2_m + rt2_c) * R(0.5) * gm);
                // explicit rho' (moist.jl:1383-1386)
                R adv = DIV(-(ru_i - ru_m), dx, rdx);
                R inc = dt_s * (Fr_[m][h] + adv);
                const R rpc = rp_[m][h];
                const R rp_tp = rpc + R(fact_tp_d) * inc;
                rp_e[l] = rpc + inc;
                // explicit rho_theta' (moist.jl:1389-1394)
                const R th_c = thc_[m][h];
                const R fip = (thp_[m][h] + th_c) * R(0.5) * ru_i;
                const R fim = (th_c + thm_[m][h]) * R(0.5) * ru_m;
                adv = DIV(-(fip - fim), dx, rdx);
                inc = dt_s * (Ft_[m][h] + adv);
                const R rtp_tp = cur_c + R(fact_tp_d) * inc;
                rtp_e[l] = cur_c + inc;
                thx[l + 2] = th_c;
                rt2c[l] = rt2_c;
                if (l == 0) {
                    tp_r0 = rp_tp;
                    tp_t0 = rtp_tp;
                } else {
                    // d rho_w' at face k-1 (moist.jl:1399-1404)
                    const R buoy = fdiv(rtp_tp - tp_tp, R(zpre[l]) - R(zpre[l - 1]));
                    const R grav = -K<R>::GRAV * R(0.5) * (rp_tp + tp_rp);
                    const R rt2h = (rt2c[l - 1] + rt2_c) * R(0.5);
                    drw[l - 1] = dt_s * (rrw[l - 1] - rt2h * buoy + grav);
                }
                tp_rp = rp_tp;
                tp_tp = rtp_tp;
                // rho_u' and its running mean mom_x (moist.jl:1453-1455)
                rui[h] = ru_i;
                mxo[h] = mx_[m][h] + DIV(ru_i, R(nsub), rns);
            }
        }
        if (m + 1 < NP) load_group(m + 1 < NP ? m + 1 : 0);
        else load_rest();
        if (GL * m < nl) {
            if constexpr (FIN) {
#pragma unroll
                for (int h = 0; h < GL; ++h) {
                    const int l = GL * m + h;
                    if (l < nl) {
                        const R v = iru_[m][h] + rui[h];
                        stb(ab.fin_ru + o2, bx(k0 + l), v);
                        if (isnan(v)) nanf = min(nanf, 2);
                    }
                }
            } else {
                stb2(ab.rup_out + o2, bc[m], rui[0], rui[1]);
            }
            stb2(f_mx, bc[m], mxo[0], mxo[1]);
        }
    }
    // R_rho_w at the face below level k0+L, and the halo values now in place
    // (branch-free selects, so no use waits on a load inside a branch)
    const R rrw_top = rrw[L - 1];
    {
        const bool lo2 = nl > 0 && k0 >= 2, lo1 = nl > 0 && k0 >= 1;
        const bool hi = nl == L && k0 + L < nz, hi1 = nl == L && k0 + L + 1 < nz;
        thx[0] = lo2 ? a0 : thx[0];
        thx[1] = lo1 ? a1 : thx[1];
        thx[L + 2] = hi ? b0 : thx[L + 2];
        rt2c[L] = hi ? r2top : rt2c[L];
        thx[L + 3] = hi1 ? b1 : thx[L + 3];
    }
    if (t < nz) {
        double *lq = lgeo + 5u * (uint32_t)t;
        lq[0] = z0;
        lq[1] = dz0;
        lq[2] = g0.x;
        lq[3] = g0.y;
        lq[4] = g0.w;
    }
    for (int q = t + NT; q < nz; q += NT) {
        const double4 g = zv.geo[D.offz + q];
        double *lq = lgeo + 5u * (uint32_t)q;
        lq[0] = zd[q];
        lq[1] = dzd[q];
        lq[2] = g.x;
        lq[3] = g.y;
        lq[4] = g.w;
    }
    // ---- LDS round A: off-centred values of the first level of each chunk
    // (and the geometry staged above)
    ldt[t] = tp_r0;
    ldt[NT + t] = tp_t0;
    __syncthreads();
    // per-level geometry by 32-bit LDS index
    auto LZ = [&](int k) { return lgeo[5u * (uint32_t)k]; };
    auto LDZ = [&](int k) { return lgeo[5u * (uint32_t)k + 1u]; };
    if (nl == L && k0 + L < nz) {
        const int k = k0 + L;
        const R up_r = ldt[t + TX], up_t = ldt[NT + t + TX];
        const R buoy = fdiv(up_t - tp_tp, R(LZ(k)) - R(LZ(k - 1)));
        const R grav = -K<R>::GRAV * R(0.5) * (up_r + tp_rp);
        const R rt2h = (rt2c[L - 1] + rt2c[L]) * R(0.5);
        drw[L - 1] = dt_s * (rrw_top - rt2h * buoy + grav);
    }

    // theta at the faces k0-1 .. k0+L (f2h in T, moist.jl:982-985): thh[q] is face k0-1+q
    double thh[L + 2];
#pragma unroll
    for (int q = 0; q < L + 2; ++q) {
        const int j = k0 - 1 + q;
        thh[q] = 0.0;
        // faces up to nz-2 exist; their theta levels are all in thx (a partial
        // chunk is the top chunk, so it never needs levels above its own)
        if (nl > 0 && j >= 0 && j <= nz - 2) thh[q] = double(f2h(thx[q + 1], thx[q + 2], R(LDZ(j)), R(LDZ(j + 1))));
    }
    // tridiagonal coefficients of face j = k0 + l before the off-centering
    // update b += 1 (moist.jl:997-1006), with the grid factors 1/dz, 1/(z_j+1 - z_j)
    // and g/(dz_j + dz_j+1) from the staged per-level table
    auto coef = [&](int l, double &a, double &b, double &cc) {
        const int j = k0 + l;
        const double *lj = lgeo + 5u * (uint32_t)j;
        const double gx = lj[2], gy = lj[3], gw = lj[4];
        const double rdz1 = lj[5 + 2];
        const double r2j = double(rt2c[l]), r2k = double(rt2c[l + 1]);
        b = dt_tp2 * (r2j * gx + r2k * rdz1) * thh[l + 1] * gy;
        a = j > 0 ? -dt_tp2 * (r2j * thh[l] * (gy * gx) - gw) : 0.0;
        cc = j < nz - 2 ? -dt_tp2 * (r2k * thh[l + 2] * (gy * rdz1) + gw) : 0.0;
    };

    // coefficients of the chunk's faces, once for the right-hand sides, the
    // Moebius composites and the forward substitution
    double ca[L], cb[L], ccf[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        ca[l] = cb[l] = ccf[l] = 0.0;
        if (l < nf) coef(l, ca[l], cb[l], ccf[l]);
    }

    // ---- right-hand sides (moist.jl:1008-1021)
#pragma unroll
    for (int l = 0; l < L; ++l) {
        dd[l] = 0.0;
        if (l < nf) {
            const int j = k0 + l;
            const double a = ca[l], b = cb[l], cc = ccf[l];
            const double rwj = double(rwv[l]);
            double d = rwj + double(drw[l]);
            if (j > 0) d -= a * fact * rwj;
            d -= b * fact * rwj;
            if (j < nz - 2) d -= cc * fact * double(rwv[l + 1]);
            dd[l] = d;
        }
    }

    // ---- cross-chunk scans: two levels.  A wave holds CPW = 64 / TX
    // consecutive chunks of TX columns (CPW = 1 when TX >= 64); the chunks of
    // a wave combine with lane shuffles (no barrier), the groups of CPW chunks
    // through one LDS exchange per scan, each thread folding the aggregates of
    // the groups below (or above) its own in order.  G = nch / CPW <= NT / 64.
    const int CPW = TX >= 64 ? 1 : 64 / TX;
    const int cw = c % CPW;    // chunk within the wave
    const int grp = c / CPW;   // group of the chunk
    const int G = nch / CPW;
    const int gi = grp * TX + jj;  // this thread's (group, column) slot

    // ---- c' entering the chunk: Moebius composites scanned once per stage
    if (first) {
        double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            if (l < nf) {
                const double a = ca[l], b = cb[l] + 1.0, cc = ccf[l];
                // [num; den] <- [[0, c], [-a, b]] [num; den]
                const double n00 = cc * m10, n01 = cc * m11;
                const double n10 = -a * m00 + b * m10, n11 = -a * m01 + b * m11;
                const double sc = __builtin_amdgcn_rcp(fmax(fmax(fabs(n00), fabs(n01)), fmax(fabs(n10), fabs(n11))));
                m00 = n00 * sc;
                m01 = n01 * sc;
                m10 = n10 * sc;
                m11 = n11 * sc;
            }
        }
        // in-wave inclusive scan S_c = M_c S_c-1 ... M_first of the group
        for (int d = 1; d < CPW; d <<= 1) {
            const double b0 = __shfl_up(m00, d * TX), b1 = __shfl_up(m01, d * TX), b2 = __shfl_up(m10, d * TX),
                         b3 = __shfl_up(m11, d * TX);
            if (cw >= d) {
                const double n0 = m00 * b0 + m01 * b2, n1 = m00 * b1 + m01 * b3;
                const double n2 = m10 * b0 + m11 * b2, n3 = m10 * b1 + m11 * b3;
                const double sc = __builtin_amdgcn_rcp(fmax(fmax(fabs(n0), fabs(n1)), fmax(fabs(n2), fabs(n3))));
                m00 = n0 * sc;
                m01 = n1 * sc;
                m10 = n2 * sc;
                m11 = n3 * sc;
            }
        }
        // the group composite (last chunk of the group) to LDS
        double *agm = lds;  // 4 doubles per (group, column)
        if (cw == CPW - 1) {
            agm[4 * gi + 0] = m00;
            agm[4 * gi + 1] = m01;
            agm[4 * gi + 2] = m10;
            agm[4 * gi + 3] = m11;
        }
        // composite of the chunks of this group below this one (identity for the first)
        double e00 = 1.0, e01 = 0.0, e10 = 0.0, e11 = 1.0;
        if (CPW > 1) {
            const double p0 = __shfl_up(m00, TX), p1 = __shfl_up(m01, TX), p2 = __shfl_up(m10, TX),
                         p3 = __shfl_up(m11, TX);
            if (cw > 0) e00 = p0, e01 = p1, e10 = p2, e11 = p3;
        }
        __syncthreads();
        // the groups below applied to (0, 1), then the chunks of this group below this one
        double u0 = 0.0, u1 = 1.0;
        for (int gp = 0; gp < grp; ++gp) {
            const double *M = agm + 4 * (gp * TX + jj);
            const double v0 = M[0] * u0 + M[1] * u1, v1 = M[2] * u0 + M[3] * u1;
            const double sc = __builtin_amdgcn_rcp(fmax(fabs(v0), fabs(v1)));
            u0 = v0 * sc;
            u1 = v1 * sc;
        }
        const double s0 = e00 * u0 + e01 * u1, s1 = e10 * u0 + e11 * u1;
        cin_v = c > 0 ? s0 / s1 : 0.0;
        if (nl > 0) ab.cin[cidx] = cin_v;
    }

    // ---- forward substitution inside the chunk with a zero carry
    double ylast = 0.0, plast = 1.0;
    {
        double cprev = cin_v, yprev = 0.0, P = 1.0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            cpv[l] = 0.0;
            pv[l] = 0.0;
            if (l < nf) {
                const double a = ca[l], b = cb[l] + 1.0, cc = ccf[l];
                const double rm = rcp64(b - a * cprev);
                const double cp = cc * rm;
                const double y = (dd[l] - a * yprev) * rm;
                P = (-a * P) * rm;
                cpv[l] = cp;
                dd[l] = y;
                pv[l] = P;
                cprev = cp;
                yprev = y;
                ylast = y;
                plast = P;
            }
        }
    }
    // carries of the forward substitution: the chunk maps x -> y + P x
    // composed from the bottom; the carry into chunk c is S_c-1(0)
    {
        double y = ylast, P = plast;
        for (int d = 1; d < CPW; d <<= 1) {
            const double yp = __shfl_up(y, d * TX), Pp = __shfl_up(P, d * TX);
            if (cw >= d) {
                y = y + P * yp;
                P = P * Pp;
            }
        }
        double *agf = lds + 4 * NT;  // 2 doubles per (group, column)
        if (cw == CPW - 1) {
            agf[2 * gi + 0] = y;
            agf[2 * gi + 1] = P;
        }
        double ye = 0.0, Pe = 1.0;
        if (CPW > 1) {
            const double yq = __shfl_up(y, TX), Pq = __shfl_up(P, TX);
            if (cw > 0) ye = yq, Pe = Pq;
        }
        __syncthreads();
        double E = 0.0;
        for (int gp = 0; gp < grp; ++gp) E = agf[2 * (gp * TX + jj)] + agf[2 * (gp * TX + jj) + 1] * E;
        const double din = ye + Pe * E;
#pragma unroll
        for (int l = 0; l < L; ++l)
            if (l < nf) dd[l] = dd[l] + pv[l] * din;
    }

    // ---- back substitution inside the chunk with a zero carry from above
    double zfirst = 0.0, qfirst = 1.0, zlast = 0.0, qlast = 1.0;
    {
        double znext = 0.0, Q = 1.0;
#pragma unroll
        for (int l = L - 1; l >= 0; --l) {
            if (l < nf) {
                const double zl = dd[l] - cpv[l] * znext;
                Q = -cpv[l] * Q;
                if (l == nf - 1) {
                    zlast = zl;
                    qlast = Q;
                }
                dd[l] = zl;
                pv[l] = Q;
                znext = zl;
                zfirst = zl;
                qfirst = Q;
            }
        }
    }
    // carries of the back substitution: the maps x_in -> z_first + Q_first x_in
    // composed from the top; T_c(0) is x at chunk c's first face, the carry
    // into chunk c is T_c+1(0)
    double xin, x_below = 0.0;
    {
        double y = zfirst, P = qfirst;
        for (int d = 1; d < CPW; d <<= 1) {
            const double yn = __shfl_down(y, d * TX), Pn = __shfl_down(P, d * TX);
            if (cw + d < CPW) {
                y = y + P * yn;
                P = P * Pn;
            }
        }
        // group composites (first chunk of each group) and the (z_last, Q_last)
        // of each group's last chunk; the Moebius area is free again (the
        // forward scan's barrier lies between its reads and these writes)
        double *agb = lds, *lzq = lds + 2 * NT;
        if (cw == 0) {
            agb[2 * gi + 0] = y;
            agb[2 * gi + 1] = P;
        }
        if (cw == CPW - 1) {
            lzq[2 * gi + 0] = zlast;
            lzq[2 * gi + 1] = qlast;
        }
        double zl_b = 0.0, ql_b = 1.0;  // (z_last, Q_last) of chunk c-1 when it is in this wave
        if (CPW > 1) {
            const double zq = __shfl_up(zlast, TX), qq = __shfl_up(qlast, TX);
            if (cw > 0) zl_b = zq, ql_b = qq;
        }
        __syncthreads();
        double F = 0.0;
        for (int gp = G - 1; gp > grp; --gp) F = agb[2 * (gp * TX + jj)] + agb[2 * (gp * TX + jj) + 1] * F;
        const double xfirst = y + P * F;  // x at this chunk's first face
        xin = F;
        if (CPW > 1) {
            const double xn = __shfl_down(xfirst, TX);
            if (cw < CPW - 1) xin = xn;
        }
        if (cw == 0 && grp > 0) {
            zl_b = lzq[2 * ((grp - 1) * TX + jj)];
            ql_b = lzq[2 * ((grp - 1) * TX + jj) + 1];
        }
        // x at the last face of the chunk below = z_last + Q_last * (x at this chunk's first face)
        if (c > 0 && nl > 0) x_below = zl_b + ql_b * xfirst;
    }
    // rho_w' and mom_z at the chunk's faces; the pair of the lid level ke
    // (which has no face) writes 0 there: no kernel reads rho_w' or mom_z at ke
    R xr[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        xr[l] = R(0.0);
        if (l < nf) xr[l] = R(dd[l] + pv[l] * xin);
    }
#pragma unroll
    for (int m = 0; m < L / 2; ++m) {
        if (2 * m < nf) {
            if constexpr (!FIN) stb2(w.rwp + o2, bc[m], xr[2 * m], xr[2 * m + 1]);
            stb2(f_mz, bc[m], mzv[2 * m] + DIV(xr[2 * m], R(nsub), rns),
                      mzv[2 * m + 1] + DIV(xr[2 * m + 1], R(nsub), rns));
        }
    }
    if constexpr (FIN) {
        // rho_w of the stage output: the lid level keeps the input's value
#pragma unroll
        for (int l = 0; l < L; ++l) {
            if (l < nl) {
                const int k = k0 + l;
                const R v = k < ke ? swv[l + 1] + xr[l] : swv[l + 1];
                stb(ab.fin_rw + o2, bx(k), v);
                if (isnan(v)) nanf = min(nanf, 3);
            }
        }
    }
    const R xb = R(x_below);

    // ---- vertical flux updates of rho' and rho_theta' (moist.jl:1412-1446)
#pragma unroll
    for (int l = 0; l < L; ++l) {
        if (l < nl) {
            const int k = k0 + l;
            const R wkp = k == ke ? R(0.0) : xr[l];
            const R wkm = k == 0 ? R(0.0) : (l > 0 ? xr[l > 0 ? l - 1 : 0] : xb);
            const R dzk = R(LDZ(k)), rdzk = rcp_of(dzk);
            R adv = DIV(-(wkp - wkm), dzk, rdzk);
            rp_e[l] = rp_e[l] + dt_s * adv;
            const R t_m2 = thx[l], t_m1 = thx[l + 1], t_0 = thx[l + 2], t_p1 = thx[l + 3], t_p2 = thx[l + 4];
            R Fkp, Fkm;
            if (k == ke) Fkp = R(0.0);
            else {
                if (wkp > EPS && k > 0) Fkp = DIV(wkp * upw3(t_m1, t_0, t_p1), R(6.0), r6);
                else if (swv[l + 1] < -EPS && k < ke - 1) Fkp = DIV(wkp * upw3(t_p2, t_p1, t_0), R(6.0), r6);
                else Fkp = wkp * f2h(t_0, t_p1, dzk, R(LDZ(k + 1)));
            }
            if (k == 0) Fkm = R(0.0);
            else {
                const R sw = swv[l];
                if (sw > EPS && k > 1) Fkm = DIV(wkm * upw3(t_m2, t_m1, t_0), R(6.0), r6);
                else if (sw < -EPS && k < ke) Fkm = DIV(wkm * upw3(t_p1, t_0, t_m1), R(6.0), r6);
                else Fkm = wkm * f2h(t_m1, t_0, R(LDZ(k - 1)), dzk);
            }
            adv = DIV(-(Fkp - Fkm), dzk, rdzk);
            rtp_e[l] = rtp_e[l] + dt_s * adv;
            if constexpr (FIN) {
                const uint32_t q = bx(k);
                const R vr = inr[l] + rp_e[l], vt = inrt[l] + rtp_e[l];
                stb(ab.fin_r + o2, q, vr);
                stb(ab.fin_rt + o2, q, vt);
                if (isnan(vr)) nanf = 1;
                else if (isnan(vt)) nanf = min(nanf, 4);
            } else if ((l & 1) || l == nl - 1) {
                // stored as soon as a level pair is complete; an odd top level
                // is stored with the padding level above it
                const int l0 = l & ~1;
                stb2(f_rp, bc[l0 / 2], rp_e[l0], rp_e[l0 + 1]);
                stb2(ab.rtp_out + o2, bc[l0 / 2], rtp_e[l0], rtp_e[l0 + 1]);
            }
        }
    }
    if constexpr (FIN) {
        // the NaN codes of the state fields, as the final update reports them
        if (nanf < 5) report_nan(w.nan_key, ((*w.step_ctr * 3 + ab.stage) << 4) | nanf);
    }
}
