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
// The cross-chunk combinations are Hillis-Steele scans in LDS (log2(nch)
// steps, every thread active).

#ifndef MOIST_AC_THREADS
#define MOIST_AC_THREADS 1024
#endif
constexpr int AC_THREADS = MOIST_AC_THREADS;

template <class S> struct AcBuf {
    const S *rup_in;   // rho_u' before this substep
    S *rup_out;        // rho_u' after it
    const S *rtp_cur;  // rho_theta' at the start of the substep
    const S *rtp_old;  // rho_theta' at the start of the previous substep
    S *rtp_out;        // rho_theta' after the substep
    double *cin;       // c' entering each chunk, [chunk][column]
    int64_t cin_stride;
};

// WPE: minimum waves per SIMD the register allocation must allow (1: no
// constraint; 8 lets two 1024-thread workgroups share a CU).
template <class S, int L, int NT, int WPE>
__global__ void __launch_bounds__(NT, WPE) k_acoustic(const Dom *__restrict__ doms, const int4 *__restrict__ atiles,
                                                 ZVec zv, SSet<S> in, Work<S> w, AcBuf<S> ab, int div,
                                                 double dt_ovr, int nsub, int first) {
    using R = S;
    __shared__ double lds[8 * NT];  // two buffers of 4 doubles per thread for the scans
    __shared__ S ldt[2 * NT];

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
    const R beta = R(D.beta);
    // fact_tp of the explicit part is a Float64 expression (moist.jl:1375)
    const double fact_tp_d = double(R(0.5)) * (1.0 + double(beta));
    // implicit_correction!'s own factors, in T (moist.jl:987-989)
    const R ftp = R(0.5) * (R(1.0) + beta);
    const R ftm = R(0.5) * (R(1.0) - beta);
    const double dt_tp2 = double(dt_s * ftp) * double(dt_s * ftp);
    const double fact = double(ftm / ftp);
    const double G = K<double>::GRAV;
    const int64_t cidx = (int64_t)c * ab.cin_stride + D.offx + i;
    double cin_v = 0.0;
    if (!first && nl > 0) cin_v = ab.cin[cidx];

    R rp_e[L], rtp_e[L], drw[L], thx[L + 4], rt2c[L + 1];
    double dd[L], cpv[L], pv[L];
#pragma unroll
    for (int q = 0; q < L + 4; ++q) thx[q] = R(0.0);
#pragma unroll
    for (int q = 0; q < L + 1; ++q) rt2c[q] = R(0.0);

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
            const R cur_c = ab.rtp_cur[cc], cur_m = ab.rtp_cur[cm], cur_p = ab.rtp_cur[cx];
            const R old_c = ab.rtp_old[cc], old_m = ab.rtp_old[cm], old_p = ab.rtp_old[cx];
            const R rt2_c = w.rt2[cc], rt2_m = w.rt2[cm], rt2_p = w.rt2[cx];
            // rho_u' at i+1/2 and i-1/2 (moist.jl:1362-1367)
            R gx = fdiv(cur_p - cur_c, dx);
            R gxo = fdiv(old_p - old_c, dx);
            gx = gx + gam * (gx - gxo);
            const R ru_i = ab.rup_in[cc] + dt_s * (w.R_rhou[cc] - (rt2_c + rt2_p) * R(0.5) * gx);
            R gm = fdiv(cur_c - cur_m, dx);
            R gmo = fdiv(old_c - old_m, dx);
            gm = gm + gam * (gm - gmo);
            const R ru_m = ab.rup_in[cm] + dt_s * (w.R_rhou[cm] - (rt2_m + rt2_c) * R(0.5) * gm);
            ab.rup_out[cc] = ru_i;
            w.mom_x[cc] = w.mom_x[cc] + fdiv(ru_i, R(nsub));
            // explicit rho' (moist.jl:1383-1386)
            R adv = fdiv(-(ru_i - ru_m), dx);
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
            adv = fdiv(-(fip - fim), dx);
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
        }
    }
    // theta two levels below and above the chunk, rt2pres one level above
    if (nl > 0) {
        const int64_t ci = base + i;
        if (k0 >= 2) thx[0] = w.theta[ci + (int64_t)(k0 - 2) * nx];
        if (k0 >= 1) thx[1] = w.theta[ci + (int64_t)(k0 - 1) * nx];
        if (nl == L) {
            if (k0 + L < nz) {
                thx[L + 2] = w.theta[ci + (int64_t)(k0 + L) * nx];
                rt2c[L] = w.rt2[ci + (int64_t)(k0 + L) * nx];
            }
            if (k0 + L + 1 < nz) thx[L + 3] = w.theta[ci + (int64_t)(k0 + L + 1) * nx];
        }
    }

    // ---- LDS round A: off-centred values of the first level of each chunk
    ldt[t] = tp_r0;
    ldt[NT + t] = tp_t0;
    __syncthreads();
    if (nl == L && k0 + L < nz) {
        const int k = k0 + L;
        const R up_r = ldt[t + TX], up_t = ldt[NT + t + TX];
        const R buoy = fdiv(up_t - tp_tp, R(zd[k]) - R(zd[k - 1]));
        const R grav = -K<R>::GRAV * R(0.5) * (up_r + tp_rp);
        const R rt2h = (rt2c[L - 1] + rt2c[L]) * R(0.5);
        drw[L - 1] = dt_s * (w.R_rhow[base + (int64_t)(k - 1) * nx + i] - rt2h * buoy + grav);
    }

    // theta at the faces k0-1 .. k0+L (f2h in T, moist.jl:982-985): thh[q] is face k0-1+q
    double thh[L + 2];
#pragma unroll
    for (int q = 0; q < L + 2; ++q) {
        const int j = k0 - 1 + q;
        thh[q] = 0.0;
        // faces up to nz-2 exist; their theta levels are all in thx (a partial
        // chunk is the top chunk, so it never needs levels above its own)
        if (nl > 0 && j >= 0 && j <= nz - 2) thh[q] = double(f2h(thx[q + 1], thx[q + 2], R(dzd[j]), R(dzd[j + 1])));
    }
    const double4 *geo = zv.geo + D.offz;
    // tridiagonal coefficients of face j = k0 + l before the off-centering
    // update b += 1 (moist.jl:997-1006), with the grid factors 1/dz, 1/(z_j+1 - z_j)
    // and g/(dz_j + dz_j+1) taken from the per-level table
    auto coef = [&](int l, double &a, double &b, double &cc) {
        const int j = k0 + l;
        const double4 gj = geo[j];
        const double rdz1 = geo[j + 1].x;
        const double r2j = double(rt2c[l]), r2k = double(rt2c[l + 1]);
        b = dt_tp2 * (r2j * gj.x + r2k * rdz1) * thh[l + 1] * gj.y;
        a = j > 0 ? -dt_tp2 * (r2j * thh[l] * (gj.y * gj.x) - gj.w) : 0.0;
        cc = j < nz - 2 ? -dt_tp2 * (r2k * thh[l + 2] * (gj.y * rdz1) + gj.w) : 0.0;
    };

    // ---- right-hand sides (moist.jl:1008-1021)
#pragma unroll
    for (int l = 0; l < L; ++l) {
        dd[l] = 0.0;
        if (l < nf) {
            const int j = k0 + l;
            double a, b, cc;
            coef(l, a, b, cc);
            const double rwj = double(w.rwp[base + (int64_t)j * nx + i]);
            double d = rwj + double(drw[l]);
            if (j > 0) d -= a * fact * rwj;
            d -= b * fact * rwj;
            if (j < nz - 2) d -= cc * fact * double(w.rwp[base + (int64_t)(j + 1) * nx + i]);
            dd[l] = d;
        }
    }

    // ---- c' entering the chunk: Moebius composites scanned once per stage
    if (first) {
        double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            if (l < nf) {
                double a, b, cc;
                coef(l, a, b, cc);
                b += 1.0;
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
        // inclusive Hillis-Steele scan over the chunks of each column:
        // S_c = M_c S_{c-1} ... M_0; the c' entering chunk c is S_{c-1} applied to (0, 1)
        double *A = lds, *B = lds + 4 * NT;
        A[4 * t + 0] = m00;
        A[4 * t + 1] = m01;
        A[4 * t + 2] = m10;
        A[4 * t + 3] = m11;
        __syncthreads();
        for (int dst = 1; dst < nch; dst <<= 1) {
            double a0 = A[4 * t], a1 = A[4 * t + 1], a2 = A[4 * t + 2], a3 = A[4 * t + 3];
            if (c >= dst) {
                const int o = 4 * (t - dst * TX);
                const double b0 = A[o], b1 = A[o + 1], b2 = A[o + 2], b3 = A[o + 3];
                const double n0 = a0 * b0 + a1 * b2, n1 = a0 * b1 + a1 * b3;
                const double n2 = a2 * b0 + a3 * b2, n3 = a2 * b1 + a3 * b3;
                const double sc = 1.0 / fmax(fmax(fabs(n0), fabs(n1)), fmax(fabs(n2), fabs(n3)));
                a0 = n0 * sc;
                a1 = n1 * sc;
                a2 = n2 * sc;
                a3 = n3 * sc;
            }
            B[4 * t + 0] = a0;
            B[4 * t + 1] = a1;
            B[4 * t + 2] = a2;
            B[4 * t + 3] = a3;
            __syncthreads();
            double *tmp = A;
            A = B;
            B = tmp;
        }
        cin_v = 0.0;
        if (c > 0) {
            const int o = 4 * (t - TX);
            cin_v = A[o + 1] / A[o + 3];
        }
        if (nl > 0) ab.cin[cidx] = cin_v;
        __syncthreads();
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
                double a, b, cc;
                coef(l, a, b, cc);
                b += 1.0;
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
    // carries of the forward substitution: inclusive scan of the chunk maps
    // x -> y + P x from the bottom; the carry into chunk c is S_{c-1}(0)
    {
        double *A = lds, *B = lds + 4 * NT;
        A[2 * t + 0] = ylast;
        A[2 * t + 1] = plast;
        __syncthreads();
        for (int dst = 1; dst < nch; dst <<= 1) {
            double y = A[2 * t], P = A[2 * t + 1];
            if (c >= dst) {
                const int o = 2 * (t - dst * TX);
                y = y + P * A[o];
                P = P * A[o + 1];
            }
            B[2 * t + 0] = y;
            B[2 * t + 1] = P;
            __syncthreads();
            double *tmp = A;
            A = B;
            B = tmp;
        }
        const double din = c > 0 ? A[2 * (t - TX)] : 0.0;
#pragma unroll
        for (int l = 0; l < L; ++l)
            if (l < nf) dd[l] = dd[l] + pv[l] * din;
        __syncthreads();
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
    // carries of the back substitution: inclusive scan of the maps
    // x_in -> z_first + Q_first x_in from the top; T_c(0) is x at chunk c's
    // first face, the carry into chunk c is T_{c+1}(0)
    double xin, x_below = 0.0;
    {
        double *A = lds, *B = lds + 4 * NT;
        A[2 * t + 0] = zfirst;
        A[2 * t + 1] = qfirst;
        __syncthreads();
        for (int dst = 1; dst < nch; dst <<= 1) {
            double y = A[2 * t], P = A[2 * t + 1];
            if (c + dst < nch) {
                const int o = 2 * (t + dst * TX);
                y = y + P * A[o];
                P = P * A[o + 1];
            }
            B[2 * t + 0] = y;
            B[2 * t + 1] = P;
            __syncthreads();
            double *tmp = A;
            A = B;
            B = tmp;
        }
        xin = c + 1 < nch ? A[2 * (t + TX)] : 0.0;
        const double xfirst = A[2 * t];  // x at this chunk's first face
        // x at the last face of the chunk below = z_last + Q_last * (x at this chunk's first face)
        B[2 * t + 0] = zlast;
        B[2 * t + 1] = qlast;
        __syncthreads();
        if (c > 0 && nl > 0) {
            const int o = 2 * (t - TX);
            x_below = B[o] + B[o + 1] * xfirst;
        }
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
            w.mom_z[cj] = w.mom_z[cj] + fdiv(xr[l], R(nsub));
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
            const R dzk = R(dzd[k]);
            R adv = fdiv(-(wkp - wkm), dzk);
            w.rp[cc] = rp_e[l] + dt_s * adv;
            const R t_m2 = thx[l], t_m1 = thx[l + 1], t_0 = thx[l + 2], t_p1 = thx[l + 3], t_p2 = thx[l + 4];
            R Fkp, Fkm;
            if (k == ke) Fkp = R(0.0);
            else {
                if (wkp > EPS && k > 0) Fkp = fdiv(wkp * upw3(t_m1, t_0, t_p1), R(6.0));
                else if (in.f[RHOW][cc] < -EPS && k < ke - 1) Fkp = fdiv(wkp * upw3(t_p2, t_p1, t_0), R(6.0));
                else Fkp = wkp * f2h(t_0, t_p1, dzk, R(dzd[k + 1]));
            }
            if (k == 0) Fkm = R(0.0);
            else {
                const R sw = in.f[RHOW][cc - nx];
                if (sw > EPS && k > 1) Fkm = fdiv(wkm * upw3(t_m2, t_m1, t_0), R(6.0));
                else if (sw < -EPS && k < ke) Fkm = fdiv(wkm * upw3(t_p1, t_0, t_m1), R(6.0));
                else Fkm = wkm * f2h(t_m1, t_0, R(dzd[k - 1]), dzk);
            }
            adv = fdiv(-(Fkp - Fkm), dzk);
            ab.rtp_out[cc] = rtp_e[l] + dt_s * adv;
        }
    }
}
