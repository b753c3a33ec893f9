// moist_history.h — the derived fields and budget of output_history
// (moist.jl:1816-1858) as device kernels.  Included by moist_api.hip inside
// namespace mo.
//
// Per cell, in FP64 from the state in either form: u and w at the cell
// centre, theta, qv, qc, qr, relative humidity, temperature, pressure (hPa),
// equivalent potential temperature, and the kinetic, potential and moist
// internal energy.  Per member: the total energy and the extrema the
// reference logs (w min/max, qc and qr max in g/kg, RH min/max).  Blocks
// write partial sums and extrema; one block per member combines them in a
// fixed order, so results are deterministic.

enum { HF_RHO = 0, HF_RHOU, HF_RHOW, HF_U, HF_W, HF_THETA, HF_QV, HF_QC, HF_QR, HF_RH, HF_TEMP, HF_PRES, HF_THETAE,
       HF_COUNT };
enum { HS_ETOT = 0, HS_WMIN, HS_WMAX, HS_QCMAX, HS_QRMAX, HS_RHMIN, HS_RHMAX, HS_COUNT };

template <class S>
__global__ void __launch_bounds__(256)
    k_history(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<S> s, double *fields,
              int64_t n2, double *partial) {
    const int4 t = tiles[blockIdx.x];
    __shared__ double red[HS_COUNT][256];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double e = 0.0, wmin = 1e300, wmax = -1e300, qcm = -1e300, qrm = -1e300, rhmin = 1e300, rhmax = -1e300;
    if (t.x >= 0) {
        const Dom &D = doms[t.x];
        const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
        if (i < D.nx && k < D.nz) {
            const K<double> C;
            const int64_t c = D.off2 + (int64_t)k * D.nx + i;
            const int64_t cm = D.off2 + (int64_t)k * D.nx + wrapx(i - 1, D.nx);
            const bool mixed = is_mixed<S>();
            const double rho = double(s.f[RHO][c]) + (mixed ? zv.rho_bar[D.offz + k] : 0.0);
            const double rt = double(s.f[RT][c]) + (mixed ? zv.rt_bar[D.offz + k] : 0.0);
            const double rinv = 1.0 / rho;
            const double th = rt * rinv;
            const double qv = double(s.f[QV][c]) * rinv, qc = double(s.f[QC][c]) * rinv, qr = double(s.f[QR][c]) * rinv;
            const double u = (double(s.f[RHOU][c]) + double(s.f[RHOU][cm])) * 0.5 * rinv;
            const double w = k > 0 ? (double(s.f[RHOW][c]) + double(s.f[RHOW][c - D.nx])) * 0.5 * rinv
                                   : double(s.f[RHOW][c]) * 0.5 * rinv;
            const double qt = qv + qc + qr, qd = 1.0 - qt;
            const double r = C.Rd * qd + C.Rv * qv;
            double cp = C.Cpd * qd + C.Cpv * qv + C.Cpl * (qc + qr);
            const double cv = cp - r;
            const double pi = pow(r * rt / C.P0, r / cv);
            const double pres = C.P0 * pow(pi, cp / r);
            const double temp = th * pi;
            double es, des, qs, dqs;
            sat_vp(temp, es, des);
            sat_q(pres, es, qs, dqs);
            const double rh = qs > 1e-10 ? qv / qs * 100.0 : 0.0;
            const double presv = rho * qv * C.Rv * temp;
            const double presd = pres - presv;
            cp = C.Cpd * qd + C.Cpl * qt;
            const double the = temp * pow(presd / C.P0, -C.Rd / cp) * exp(C.Lv * qv / (cp * temp));
            const double ek = 0.5 * rho * (u * u + w * w);
            const double ep = rho * C.GRAV * zv.z[D.offz + k];
            const double ei = rho * (cv * temp + qv * C.Lv);
            e = ek + ep + ei;
            wmin = wmax = w;
            qcm = qc * 1000.0;
            qrm = qr * 1000.0;
            rhmin = rhmax = rh;
            if (fields) {
                const double v[HF_COUNT] = {rho, double(s.f[RHOU][c]), double(s.f[RHOW][c]), u, w, th, qv, qc, qr,
                                            rh, temp, pres / 100.0, the};
                for (int f = 0; f < HF_COUNT; ++f) fields[(int64_t)f * n2 + c] = v[f];
            }
        }
    }
    red[HS_ETOT][tid] = e;
    red[HS_WMIN][tid] = wmin;
    red[HS_WMAX][tid] = wmax;
    red[HS_QCMAX][tid] = qcm;
    red[HS_QRMAX][tid] = qrm;
    red[HS_RHMIN][tid] = rhmin;
    red[HS_RHMAX][tid] = rhmax;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) {
            red[HS_ETOT][tid] += red[HS_ETOT][tid + st];
            red[HS_WMIN][tid] = fmin(red[HS_WMIN][tid], red[HS_WMIN][tid + st]);
            red[HS_RHMIN][tid] = fmin(red[HS_RHMIN][tid], red[HS_RHMIN][tid + st]);
            for (int a : {HS_WMAX, HS_QCMAX, HS_QRMAX, HS_RHMAX}) red[a][tid] = fmax(red[a][tid], red[a][tid + st]);
        }
        __syncthreads();
    }
    if (tid < HS_COUNT) partial[(int64_t)blockIdx.x * HS_COUNT + tid] = red[tid][0];
}

// One block per member: combine the partials of the member's tiles.
__global__ void __launch_bounds__(256)
    k_history_reduce(const int4 *__restrict__ tiles, int ntiles, const double *partial, double *stats) {
    const int m = blockIdx.x, tid = threadIdx.x;
    __shared__ double red[HS_COUNT][256];
    double a[HS_COUNT] = {0.0, 1e300, -1e300, -1e300, -1e300, 1e300, -1e300};
    for (int b = tid; b < ntiles; b += blockDim.x) {
        if (tiles[b].x != m) continue;
        const double *p = partial + (int64_t)b * HS_COUNT;
        a[HS_ETOT] += p[HS_ETOT];
        a[HS_WMIN] = fmin(a[HS_WMIN], p[HS_WMIN]);
        a[HS_RHMIN] = fmin(a[HS_RHMIN], p[HS_RHMIN]);
        for (int f : {HS_WMAX, HS_QCMAX, HS_QRMAX, HS_RHMAX}) a[f] = fmax(a[f], p[f]);
    }
    for (int f = 0; f < HS_COUNT; ++f) red[f][tid] = a[f];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) {
            red[HS_ETOT][tid] += red[HS_ETOT][tid + st];
            red[HS_WMIN][tid] = fmin(red[HS_WMIN][tid], red[HS_WMIN][tid + st]);
            red[HS_RHMIN][tid] = fmin(red[HS_RHMIN][tid], red[HS_RHMIN][tid + st]);
            for (int f : {HS_WMAX, HS_QCMAX, HS_QRMAX, HS_RHMAX}) red[f][tid] = fmax(red[f][tid], red[f][tid + st]);
        }
        __syncthreads();
    }
    if (tid < HS_COUNT) stats[m * HS_COUNT + tid] = red[tid][0];
}

// ---- water and mass budget (moist_diagnostics): FP64 block partials -----------------
// Per tile: sum (rho - water) dz, sum water dz, sum rho dz.
template <class S>
__global__ void __launch_bounds__(256)
    k_budget(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<S> s, double *partial) {
    const int4 t = tiles[blockIdx.x];
    __shared__ double red[3][256];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (t.x >= 0) {
        const Dom &D = doms[t.x];
        const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
        if (i < D.nx && k < D.nz) {
            const int64_t c = D.off2 + (int64_t)k * D.nx + i;
            double rho = double(s.f[RHO][c]);
            if (is_mixed<S>()) rho += zv.rho_bar[D.offz + k];
            const double wat = double(s.f[QV][c]) + double(s.f[QC][c]) + double(s.f[QR][c]);
            const double dz = zv.dz[D.offz + k];
            a0 = (rho - wat) * dz;
            a1 = wat * dz;
            a2 = rho * dz;
        }
    }
    red[0][tid] = a0;
    red[1][tid] = a1;
    red[2][tid] = a2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st)
            for (int j = 0; j < 3; ++j) red[j][tid] += red[j][tid + st];
        __syncthreads();
    }
    if (tid < 3) partial[(int64_t)blockIdx.x * 3 + tid] = red[tid][0];
}

// One block per member: dry mass, total water, accumulated rain, total mass (times dx).
__global__ void __launch_bounds__(256) k_budget_reduce(const Dom *__restrict__ doms, const int4 *__restrict__ tiles,
                                                       int ntiles, const double *partial, const double *prec,
                                                       double *out) {
    const int m = blockIdx.x, tid = threadIdx.x;
    const Dom &D = doms[m];
    __shared__ double red[4][256];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = tid; b < ntiles; b += blockDim.x) {
        if (tiles[b].x != m) continue;
        a[0] += partial[(int64_t)b * 3 + 0];
        a[1] += partial[(int64_t)b * 3 + 1];
        a[3] += partial[(int64_t)b * 3 + 2];
    }
    for (int i = tid; i < D.nx; i += blockDim.x) a[2] += prec[D.offx + i];
    for (int j = 0; j < 4; ++j) red[j][tid] = a[j];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st)
            for (int j = 0; j < 4; ++j) red[j][tid] += red[j][tid + st];
        __syncthreads();
    }
    if (tid < 4) out[4 * m + tid] = red[tid][0] * D.dx;
}
