// moist_acoustic.h — one acoustic substep as one kernel (moist.jl:1351-1461).
//
// Included by moist_kernels.h inside namespace mo.
//
// A workgroup of NT threads owns a tile of TX columns and every level
// of those columns.  Thread t handles column t % TX and chunk t / TX, a run
// of L consecutive levels (and the faces above them) kept in registers, so
// one wave reads 64/TX level rows of TX consecutive columns (full 128-byte
// lines for TX >= 32 in FP32).  nch = NT / TX chunks cover nz.  L and NT are
// template parameters (the runtime picks them per handle, or times them).
//
// Per substep, for every cell of the chunk:
//   rho_u' at faces i-1/2 and i+1/2 with divergence damping (moist.jl:1360-1369;
//   the i-1/2 value is recomputed, not read back, so the kernel needs no
//   grid-wide barrier), its running mean mom_x (moist.jl:1453-1455),
//   the explicit rho'/rho_theta' increments and their off-centred values
//   (moist.jl:1381-1396), d rho_w' (moist.jl:1398-1405), the w-p system
//   (implicit_correction!, moist.jl:974-1030) and the vertical flux updates
//   (moist.jl:1411-1446), the running mean mom_z (moist.jl:1456-1458).
//
// The tridiagonal system is the reference's Thomas algorithm (moist.jl:952-971)
// in FP64, split over the chunks of a column:
//   * the LU factor recurrence c'_j = c_j / (b_j - a_j c'_{j-1}) is a Moebius
//     map of c'_{j-1}; the chunk composites (2x2 matrices) are scanned once per
//     stage (the matrix does not change between substeps) to give the c'
//     entering each chunk, stored in `cin`; each substep re-runs the recurrence
//     inside its chunk from that value;
//   * forward and back substitution are affine recurrences: each chunk solves
//     with a zero carry and keeps the carry coefficients, the carries are
//     combined across chunks, and every level is corrected.
// The cross-chunk combinations are two-level scans: lane shuffles over the
// chunks one wave holds, then one LDS exchange of the waves' aggregates.

#ifndef MOIST_AC_THREADS
#define MOIST_AC_THREADS 1024
#endif
constexpr int AC_THREADS = MOIST_AC_THREADS;
// tallest column the kernel takes: its per-level geometry (40 bytes a level)
// is staged in dynamic LDS within the 64 KiB a launch gets without opt-in
constexpr int AC_MAX_NZ = 1600;

template <class S> struct AcBuf {
    const S *rup_in;   // rho_u' before this substep
    S *rup_out;        // rho_u' after it
    const S *rtp_cur;  // rho_theta' at the start of the substep
    const S *rtp_old;  // rho_theta' at the start of the previous substep
    S *rtp_out;        // rho_theta' after the substep
    double *cin;       // c' entering each chunk, [chunk][column]
    int64_t cin_stride;
    const S *rwp_in;   // rho_w' before this substep (the result goes to Work::rwp; may be the same array)
    // the last substep of a stage in the pair-layout kernel: the stage's
    // final update (moist.jl:1522-1539), state = stage input + perturbation
    // of rho, rho_u, rho_theta and rho_w, goes straight to these fields
    // instead of the perturbations going back to memory
    S *fin_r, *fin_ru, *fin_rt, *fin_rw;
    int stage;  // RK stage of its NaN report
};

// WPE: minimum waves per SIMD the register allocation must allow (1: no
// constraint; 8 lets two 1024-thread workgroups share a CU).  The float
// kernel (the mixed form's recording pass) keeps at least 4 (at most 128
// VGPRs): with FP contraction it asked for 132, which left one 512-thread
// workgroup per CU instead of two.
template <class S, int L, int NT, int WPE>
__global__ void __launch_bounds__(NT, (is_mixed<S>() && WPE < 4) ? 4 : WPE) k_acoustic(const Dom *__restrict__ doms, const int4 *__restrict__ atiles,
                                                 ZVec zv, SSet<S> in, Work<S> w, AcBuf<S> ab, int div,
                                                 double dt_ovr, int nsub, int first) {
    using R = S;
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
    const int64_t base = D.off2;
    const double *zd = zv.z + D.offz, *dzd = zv.dz + D.offz;
    const R dx = R(D.dx);
    const R EPS = K<R>::EPS;
    const R dt_s = stage_dt<R>(D, div, dt_ovr) / R(nsub);
    const R gam = R(D.gamma_d);
    // Repeated divisors: the mixed form divides by the hardware reciprocal
    // (fdiv's float form), so taking each reciprocal once gives the same
    // values; the FP64 form keeps IEEE division.
    auto rcp_of = [](R b) -> R {
        if constexpr (is_mixed<S>()) return __builtin_amdgcn_rcpf(b);
        else return R(1.0) / b;
    };
    auto DIV = [](R a, R b, R rb) -> R {
        if constexpr (is_mixed<S>()) return a * rb;
        else return a / b;
    };
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

    // Loads of the later phases (after the first barrier) are issued early:
    // the per-level geometry goes to LDS, the few per-cell values to
    // registers, so no global load waits behind a barrier or a store.
    R rp_e[L], rtp_e[L], drw[L], thx[L + 4], rt2c[L + 1];
    R rwv[L + 1];  // rho_w' at faces k0 .. k0+L
    R swv[L + 1];  // input rho_w at levels k0-1 .. k0+L-1 (upwind branch tests)
    R mzv[L];      // mom_z at faces k0 .. k0+L-1
    double dd[L], cpv[L], pv[L];
#pragma unroll
    for (int q = 0; q < L + 4; ++q) thx[q] = R(0.0);
#pragma unroll
    for (int q = 0; q < L + 1; ++q) rt2c[q] = R(0.0);
    {
        const int ic = col_ok ? i : 0;
        auto at = [&](int k) -> int64_t { return base + (int64_t)min(max(k, 0), ke) * nx + ic; };
#pragma unroll
        for (int q = 0; q < L + 1; ++q) {
            const int j = k0 + q, kq = k0 - 1 + q;
            const R rv = ab.rwp_in[at(j)], sv = in.f[RHOW][at(kq)];
            rwv[q] = (nl > 0 && j <= nz - 2) ? rv : R(0.0);
            swv[q] = (nl > 0 && kq >= 0 && kq < nz) ? sv : R(0.0);
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const R mv = w.mom_z[at(k0 + l)];
            mzv[l] = l < nf ? mv : R(0.0);
        }
    }

    // ---- phase 1: rho_u', explicit rho' and rho_theta', d rho_w'
    R tp_r0 = R(0.0), tp_t0 = R(0.0), tp_rp = R(0.0), tp_tp = R(0.0);
#pragma unroll
    for (int l = 0; l < L; ++l) {
        rp_e[l] = R(0.0);
        rtp_e[l] = R(0.0);
        drw[l] = R(0.0);
        if (l < nl) {
            const int k = k0 + l;
            const int64_t row = base + (int64_t)k * nx;
            const int64_t cc = row + i, cm = row + im, cx = row + ip;
            const R mx_in = w.mom_x[cc];
            const R cur_c = ab.rtp_cur[cc], cur_m = ab.rtp_cur[cm], cur_p = ab.rtp_cur[cx];
            const R old_c = ab.rtp_old[cc], old_m = ab.rtp_old[cm], old_p = ab.rtp_old[cx];
            const R rt2_c = w.rt2[cc], rt2_m = w.rt2[cm], rt2_p = w.rt2[cx];
            // rho_u' at i+1/2 and i-1/2 (moist.jl:1362-1367)
            R gx = DIV(cur_p - cur_c, dx, rdx);
            R gxo = DIV(old_p - old_c, dx, rdx);
            gx = gx + gam * (gx - gxo);
            const R ru_i = ab.rup_in[cc] + dt_s * (w.R_rhou[cc] - (rt2_c + rt2_p) * R(0.5) * gx);
            R gm = DIV(cur_c - cur_m, dx, rdx);
            R gmo = DIV(old_c - old_m, dx, rdx);
            gm = gm + gam * (gm - gmo);
            const R ru_m = ab.rup_in[cm] + dt_s * (w.R_rhou[cm] - (rt2_m + rt2_c) * R(0.5) * gm);
            // explicit rho' (moist.jl:1383-1386)
            R adv = DIV(-(ru_i - ru_m), dx, rdx);
            R inc = dt_s * (w.R_rho[cc] + adv);
            const R rpc = w.rp[cc];
            R rp_tp;
            if constexpr (is_mixed<S>()) rp_tp = rpc + R(fact_tp_d) * inc;
            else rp_tp = R(double(rpc) + fact_tp_d * double(inc));
            rp_e[l] = rpc + inc;
            // explicit rho_theta' (moist.jl:1389-1394)
            const R th_c = w.theta[cc];
            const R fip = (w.theta[cx] + th_c) * R(0.5) * ru_i;
            const R fim = (th_c + w.theta[cm]) * R(0.5) * ru_m;
            adv = DIV(-(fip - fim), dx, rdx);
            inc = dt_s * (w.R_rt[cc] + adv);
            R rtp_tp;
            if constexpr (is_mixed<S>()) rtp_tp = cur_c + R(fact_tp_d) * inc;
            else rtp_tp = R(double(cur_c) + fact_tp_d * double(inc));
            rtp_e[l] = cur_c + inc;
            thx[l + 2] = th_c;
            rt2c[l] = rt2_c;
            if (l == 0) {
                tp_r0 = rp_tp;
                tp_t0 = rtp_tp;
            } else {
                // d rho_w' at face k-1 (moist.jl:1399-1404)
                const R buoy = fdiv(rtp_tp - tp_tp, R(zd[k]) - R(zd[k - 1]));
                const R grav = -K<R>::GRAV * R(0.5) * (rp_tp + tp_rp);
                const R rt2h = (rt2c[l - 1] + rt2_c) * R(0.5);
                drw[l - 1] = dt_s * (w.R_rhow[cc - nx] - rt2h * buoy + grav);
            }
            tp_rp = rp_tp;
            tp_tp = rtp_tp;
            // rho_u' and its running mean mom_x (moist.jl:1453-1455)
            ab.rup_out[cc] = ru_i;
            w.mom_x[cc] = mx_in + DIV(ru_i, R(nsub), rns);
        }
    }
    // theta two levels below and above the chunk, rt2pres one level above
    R rrw_top = R(0.0);  // R_rho_w at the face below level k0+L
    if (nl > 0) {
        const int64_t ci = base + i;
        if (k0 >= 2) thx[0] = w.theta[ci + (int64_t)(k0 - 2) * nx];
        if (k0 >= 1) thx[1] = w.theta[ci + (int64_t)(k0 - 1) * nx];
        if (nl == L) {
            if (k0 + L < nz) {
                thx[L + 2] = w.theta[ci + (int64_t)(k0 + L) * nx];
                rt2c[L] = w.rt2[ci + (int64_t)(k0 + L) * nx];
                rrw_top = w.R_rhow[ci + (int64_t)(k0 + L - 1) * nx];
            }
            if (k0 + L + 1 < nz) thx[L + 3] = w.theta[ci + (int64_t)(k0 + L + 1) * nx];
        }
    }

    // geometry for the phases after the barrier, per level: z, dz, 1/dz, 1/(z_k+1 - z_k), g/(dz_k + dz_k+1)
    for (int q = t; q < nz; q += NT) {
        const double4 g = zv.geo[D.offz + q];
        lgeo[5 * q + 0] = zd[q];
        lgeo[5 * q + 1] = dzd[q];
        lgeo[5 * q + 2] = g.x;
        lgeo[5 * q + 3] = g.y;
        lgeo[5 * q + 4] = g.w;
    }
    // ---- LDS round A: off-centred values of the first level of each chunk
    // (and the geometry staged above)
    ldt[t] = tp_r0;
    ldt[NT + t] = tp_t0;
    __syncthreads();
    auto LZ = [&](int k) { return lgeo[5 * k]; };
    auto LDZ = [&](int k) { return lgeo[5 * k + 1]; };
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
        const double gx = lgeo[5 * j + 2], gy = lgeo[5 * j + 3], gw = lgeo[5 * j + 4];
        const double rdz1 = lgeo[5 * (j + 1) + 2];
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
                const double sc = 1.0 / fmax(fmax(fabs(n00), fabs(n01)), fmax(fabs(n10), fabs(n11)));
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
                const double sc = 1.0 / fmax(fmax(fabs(n0), fabs(n1)), fmax(fabs(n2), fabs(n3)));
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
            const double sc = 1.0 / fmax(fabs(v0), fabs(v1));
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
                const double rm = 1.0 / (b - a * cprev);
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
    R xr[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        xr[l] = R(0.0);
        if (l < nf) {
            const double x = dd[l] + pv[l] * xin;
            const int64_t cj = base + (int64_t)(k0 + l) * nx + i;
            xr[l] = R(x);
            w.rwp[cj] = xr[l];
            w.mom_z[cj] = mzv[l] + DIV(xr[l], R(nsub), rns);
        }
    }
    const R xb = R(x_below);

    // ---- vertical flux updates of rho' and rho_theta' (moist.jl:1412-1446)
#pragma unroll
    for (int l = 0; l < L; ++l) {
        if (l < nl) {
            const int k = k0 + l;
            const int64_t cc = base + (int64_t)k * nx + i;
            const R wkp = k == ke ? R(0.0) : xr[l];
            const R wkm = k == 0 ? R(0.0) : (l > 0 ? xr[l > 0 ? l - 1 : 0] : xb);
            const R dzk = R(LDZ(k)), rdzk = rcp_of(dzk);
            R adv = DIV(-(wkp - wkm), dzk, rdzk);
            w.rp[cc] = rp_e[l] + dt_s * adv;
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
            ab.rtp_out[cc] = rtp_e[l] + dt_s * adv;
        }
    }
}
