// moist_adjoint.h — reverse-mode (adjoint) kernels of the FP64 time step.
// Included by moist_api.hip after moist_kernels.h, inside namespace mo.
//
// The adjoint of rk3_step! (moist.jl:1589-1683) runs stage by stage in
// reverse: scalar transport + final update (moist.jl:1469-1539), the
// acoustic substeps (moist.jl:1351-1461, with the transposed tridiagonal
// solve of implicit_correction!, moist.jl:974-1030), the perturbation
// initialisation, the slow tendencies (moist.jl:1071-1332) and the stage
// diagnostics (moist.jl:1044-1069); then Kessler microphysics and rain
// sedimentation (moist.jl:554-707).  Every branch the forward pass takes
// (upwind direction, Koren limiter piece, positivity limiter, saturation
// clipping, caps) is frozen: the kernels differentiate the piece taken.
//
// All kernels are gathers: the thread that owns an input cell sums the
// partial derivatives of every flux or update that reads it, so results are
// deterministic and no atomics are needed.  Partial derivatives of the
// nonlinear point functions (Koren faces, diagnostics, Kessler) come from
// forward-mode dual numbers evaluated on the same formula.
//
// Adjoint arrays live in the same x-fastest pools as the state (Dom::off2,
// periodic x through wrapn).  FP64 only.

constexpr int AD_X = 64, AD_Z = 4;  // cell tile of the gather kernels

// ---- dual numbers -------------------------------------------------------------
// Value and N derivative parts.  The value parts take exactly the operations
// the forward step takes (so frozen branch tests see its numbers); the
// derivative parts of a product or quotient are contracted to fused
// multiply-adds (`#pragma clang fp contract(fast)` in their loops only), and a
// dual combined with a plain number touches only its own derivative parts.
template <int N> struct Dl {
    double v;
    double d[N];
    __device__ Dl() : v(0.0) {
#pragma unroll
        for (int j = 0; j < N; ++j) d[j] = 0.0;
    }
    __device__ Dl(double a) : v(a) {
#pragma unroll
        for (int j = 0; j < N; ++j) d[j] = 0.0;
    }
    __device__ static Dl var(double a, int j) {
        Dl r(a);
        r.d[j] = 1.0;
        return r;
    }
};
template <int N> __device__ __forceinline__ Dl<N> operator+(const Dl<N> &a, const Dl<N> &b) {
    Dl<N> r(a.v + b.v);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = a.d[j] + b.d[j];
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator-(const Dl<N> &a, const Dl<N> &b) {
    Dl<N> r(a.v - b.v);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = a.d[j] - b.d[j];
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator-(const Dl<N> &a) {
    Dl<N> r(-a.v);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = -a.d[j];
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator*(const Dl<N> &a, const Dl<N> &b) {
    Dl<N> r(a.v * b.v);
#pragma unroll
    for (int j = 0; j < N; ++j) {
#pragma clang fp contract(fast)
        r.d[j] = a.d[j] * b.v + a.v * b.d[j];
    }
    return r;
}
// Quotients: the value is the correctly rounded quotient (the forward's
// value, so frozen branch tests see the same numbers); the derivative parts
// share one reciprocal.
template <int N> __device__ __forceinline__ Dl<N> operator/(const Dl<N> &a, const Dl<N> &b) {
    Dl<N> r(a.v / b.v);
    const double rb = 1.0 / b.v;
#pragma unroll
    for (int j = 0; j < N; ++j) {
#pragma clang fp contract(fast)
        r.d[j] = (a.d[j] - r.v * b.d[j]) * rb;
    }
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator+(const Dl<N> &a, double b) {
    Dl<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator+(double a, const Dl<N> &b) {
    Dl<N> r = b;
    r.v = a + b.v;
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator-(const Dl<N> &a, double b) {
    Dl<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator-(double a, const Dl<N> &b) {
    Dl<N> r(a - b.v);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = -b.d[j];
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator*(const Dl<N> &a, double b) {
    Dl<N> r(a.v * b);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = a.d[j] * b;
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator*(double a, const Dl<N> &b) {
    Dl<N> r(a * b.v);
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = a * b.d[j];
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator/(const Dl<N> &a, double b) {
    Dl<N> r(a.v / b);
    const double rb = 1.0 / b;
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = a.d[j] * rb;
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> operator/(double a, const Dl<N> &b) {
    Dl<N> r(a / b.v);
    const double s = -r.v / b.v;
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = s * b.d[j];
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> d_exp(const Dl<N> &a) {
    Dl<N> r(exp(a.v));
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = r.v * a.d[j];
    return r;
}
template <int N> __device__ __forceinline__ Dl<N> d_sqrt(const Dl<N> &a) {
    Dl<N> r(sqrt(a.v));
    const double h = 0.5 / r.v;
#pragma unroll
    for (int j = 0; j < N; ++j) r.d[j] = a.d[j] * h;
    return r;
}
// a^b for a > 0
template <int N> __device__ __forceinline__ Dl<N> d_pow(const Dl<N> &a, const Dl<N> &b) {
    Dl<N> r(pow(a.v, b.v));
    const double la = log(a.v), ba = b.v / a.v;
#pragma unroll
    for (int j = 0; j < N; ++j) {
#pragma clang fp contract(fast)
        r.d[j] = r.v * (b.d[j] * la + ba * a.d[j]);
    }
    return r;
}
template <int N> __device__ __forceinline__ double val(const Dl<N> &a) { return a.v; }
// Julia Base.min/max on values; the derivative follows the operand chosen
template <class V> __device__ __forceinline__ V vmin(const V &x, const V &y) {
    return signbit(val(x) - val(y)) ? x : y;
}
template <class V> __device__ __forceinline__ V vmax(const V &x, const V &y) {
    return signbit(val(x) - val(y)) ? y : x;
}

__device__ __forceinline__ int wrapn(int j, int nx) {
    j %= nx;
    return j < 0 ? j + nx : j;
}

// ---- Koren faces: value and partial derivatives (moist.jl:850-945) -------------
// On each branch of the limiter the face value is linear in the four cell
// values, so its partials are constants.  The branch itself is decided on the
// forward's values: r = rn / rd and phi(r) = max(0, min(2r, min((1+2r)/3, 2)))
// are evaluated exactly as the forward does, with Julia's min/max operand
// choice (min(x, y) = signbit(x - y) ? x : y, max(x, y) = signbit(x - y) ? y : x).
//
// One-sided reconstruction s = u0 + sg * 0.5 * phi(r) * (u0 - um), r = (u1 - u0) / (u0 - um),
// sg = +1 (moist.jl koren_face) or -1 (the z face's downward branch, which
// subtracts).  Returns s; (am, a0, a1) = ds / d(um, u0, u1).
__device__ __forceinline__ double koren_side(double um, double u0, double u1, double sg, double &am, double &a0,
                                             double &a1) {
    const double EPS = 1e-12;
    const double rn = u1 - u0, rd = u0 - um;
    double pm, p0, p1, phi;  // 0.5 phi rd = pm um + p0 u0 + p1 u1 on the branch taken
    if (fabs(rd) < EPS) {    // r = 1, phi = 1
        phi = 1.0;
        pm = -0.5, p0 = 0.5, p1 = 0.0;
    } else {
        const double r = rn / rd;
        const double t = (1.0 + 2.0 * r) / 3.0, r2 = 2.0 * r;
        int b1 = signbit(t - 2.0) ? 1 : 2;  // (1+2r)/3 or 2
        const double t1 = b1 == 1 ? t : 2.0;
        const int b2 = signbit(r2 - t1) ? 0 : b1;  // 2r or the above
        const double t2 = b2 == 0 ? r2 : t1;
        const int br = signbit(0.0 - t2) ? b2 : 3;  // phi > 0 or phi = 0
        phi = br == 3 ? 0.0 : t2;
        if (br == 0) pm = 0.0, p0 = -1.0, p1 = 1.0;                          // 0.5 (2r) rd = rn
        else if (br == 1) pm = -1.0 / 6.0, p0 = -1.0 / 6.0, p1 = 2.0 / 6.0;  // (rd + 2 rn) / 6
        else if (br == 2) pm = -1.0, p0 = 1.0, p1 = 0.0;                     // rd
        else pm = 0.0, p0 = 0.0, p1 = 0.0;
    }
    am = sg * pm;
    a0 = 1.0 + sg * p0;
    a1 = sg * p1;
    return sg > 0.0 ? u0 + 0.5 * phi * rd : u0 - 0.5 * phi * rd;
}

// x face between q1 and q2 (q0 upstream of q1, q3 downstream of q2), mass
// flux m: the face value s (flux s m) and ds/dq0..dq3.
__device__ __forceinline__ double kcoef_x(double q0, double q1, double q2, double q3, double m, double c[4]) {
    const double EPS = 1e-12;
    if (m > EPS) {
        c[3] = 0.0;
        return koren_side(q0, q1, q2, 1.0, c[0], c[1], c[2]);
    }
    if (m < -EPS) {
        c[0] = 0.0;
        return koren_side(q3, q2, q1, 1.0, c[3], c[2], c[1]);
    }
    c[0] = c[3] = 0.0;
    c[1] = c[2] = 0.5;
    return 0.5 * (q1 + q2);
}

// z face kk+1/2 (moist.jl:903-945): first-order f2h faces with the
// positivity limiters at the bottom (upward flow) and top (downward flow)
// faces, which cap the face value by lo / (m dtL) or hi / (-m dtL).  Returns
// s; c = ds/dq0..dq3; on a capped face `capped` is set and the flux s m has
// d/dlo = clo, d/dhi = chi and no dependence on m or q.
__device__ __forceinline__ double kcoef_z(double q0, double q1, double q2, double q3, double m, double lo, double hi,
                                          bool bottom, bool top, double dzk, double dzk1, double rden, double dtL,
                                          double c[4], bool &capped, double &clo, double &chi) {
    const double EPS = 1e-12;
    capped = false;
    clo = chi = 0.0;
    c[0] = c[1] = c[2] = c[3] = 0.0;
    auto f2h = [&](double &s) {
        s = (q1 * dzk1 + q2 * dzk) / (dzk + dzk1);
        c[1] = dzk1 * rden;
        c[2] = dzk * rden;
    };
    double s;
    if (m > EPS) {
        if (bottom) {
            f2h(s);
            const double lim = lo / (m * dtL);
            if (!signbit(s - lim)) {  // min(s, lim) takes lim
                capped = true;
                c[1] = c[2] = 0.0;
                clo = 1.0 / dtL;
                s = lim;
            }
        } else {
            s = koren_side(q0, q1, q2, 1.0, c[0], c[1], c[2]);
        }
    } else if (m < -EPS) {
        if (top) {
            f2h(s);
            const double lim = hi / (-m * dtL);
            if (!signbit(s - lim)) {
                capped = true;
                c[1] = c[2] = 0.0;
                chi = -1.0 / dtL;
                s = lim;
            }
        } else {
            s = koren_side(q3, q2, q1, -1.0, c[3], c[2], c[1]);
        }
    } else {
        f2h(s);
    }
    return s;
}

// ---- third-order upwind fluxes: coefficient of one input ------------------------
// A flux m * g at the interface between cell slots 0 and 1 with g from the
// values at slots -1, 0, 1, 2: positive m uses (-v[-1] + 5 v[0] + 2 v[1]) / 6,
// negative m (-v[2] + 5 v[1] + 2 v[0]) / 6, otherwise w0 v[0] + w1 v[1]
// (1/2, 1/2 or the f2h weights).  `pos` / `neg` carry the forward pass's
// branch decision (including its boundary conditions).
__device__ __forceinline__ double upw_coef(bool pos, bool neg, int off, double w0, double w1) {
    if (pos) return off == -1 ? -1.0 / 6.0 : off == 0 ? 5.0 / 6.0 : off == 1 ? 2.0 / 6.0 : 0.0;
    if (neg) return off == 2 ? -1.0 / 6.0 : off == 1 ? 5.0 / 6.0 : off == 0 ? 2.0 / 6.0 : 0.0;
    return off == 0 ? w0 : off == 1 ? w1 : 0.0;
}
__device__ __forceinline__ double upw_val(bool pos, bool neg, double vm1, double v0, double v1, double v2, double w0,
                                          double w1) {
    if (pos) return upw3(vm1, v0, v1) * (1.0 / 6.0);
    if (neg) return upw3(v2, v1, v0) * (1.0 / 6.0);
    return v0 * w0 + v1 * w1;
}

// ---- adjoint work arrays of one stage --------------------------------------------
// RC: type of the forward record (the recording pass's form: double, or
// float in the mixed form); every adjoint is FP64.
template <class RC> struct AdjW {
    // forward record of the stage
    const RC *theta, *pres, *rt2, *R_rho, *R_rhou, *R_rhow, *R_rt, *mom_x, *mom_z;
    // adjoints
    double *Lth, *Lrt, *Lpres, *LU, *LW, *LRr, *LRu, *LRw, *LRt, *Lmx, *Lmz;
    double *Lrp, *Lrwp, *Lrtp, *pend, *Linc_r, *Linc_t;
    double *Y, *cpr, *rmi, *aco;  // column scratch: transposed-solve vector; c', 1/m, a of the forward Thomas
    double *q[3], *uf, *wf;       // stage input's q = rho_q / rho and face velocities (moist.jl:1091-1097, 1184-1189)
    double *linb, *lint;       // per column: positivity-limiter adjoints (3 scalars x columns)
};

// Fill of `n` doubles (graph-safe zeroing).
__global__ void ka_fill(double *p, int64_t n, double v) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        p[q] = v;
}

// Arrays a per-cell kernel clears at its cells (up to 24).
struct FillList {
    double *p[24];
    int n;
};

// Full FP64 state from a mixed-form state (FP32 deviations; rho and
// rho_theta about the FP64 horizontal-mean base state), one cell per thread.
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_widen(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<float> in,
             SSet<double> out) {
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
    if (i >= D.nx || k >= D.nz) return;
    const int64_t c = D.off2 + (int64_t)k * D.nx + i;
#pragma unroll
    for (int f = 0; f < 7; ++f) {
        double v = double(in.f[f][c]);
        if (f == RHO) v += zv.rho_bar[D.offz + k];
        if (f == RT) v += zv.rt_bar[D.offz + k];
        out.f[f][c] = v;
    }
}

// ---- loss: J = dx sum prec + 1/2 sum (theta - theta_target)^2 -------------------
// One block per member: writes J[m] and the adjoint of the final state.
__global__ void __launch_bounds__(256) ka_loss(const Dom *__restrict__ doms, SSet<double> s, const double *prec,
                                               const double *__restrict__ tgt, double *lam_rho, double *lam_rt,
                                               double *J) {
    const int m = blockIdx.x;
    const Dom &D = doms[m];
    const int64_t n = (int64_t)D.nx * D.nz;
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        const int64_t q = D.off2 + e;
        const double th = s.f[RT][q] / s.f[RHO][q];
        const double d = th - tgt[q];
        acc += 0.5 * d * d;
        lam_rt[q] = d / s.f[RHO][q];
        lam_rho[q] = -d * th / s.f[RHO][q];
    }
    for (int i = threadIdx.x; i < D.nx; i += blockDim.x) acc += D.dx * prec[D.offx + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) J[m] = red[0];
}

// Sum of the positivity-limiter adjoints of every column into rho_q at
// (0, lin_col) and (ke, lin_col) of one member, by one 256-thread block.
__device__ __forceinline__ void lin_sum(const Dom &D, const double *linb, const double *lint, SSet<double> ls,
                                        int tid, double *rb, double *rt) {
    for (int sc = 0; sc < 3; ++sc) {
        double b = 0.0, tp = 0.0;
        for (int i = tid; i < D.nx; i += 256) {
            b += linb[D.offx * 3 + (int64_t)sc * D.nx + i];
            tp += lint[D.offx * 3 + (int64_t)sc * D.nx + i];
        }
        rb[tid] = b;
        rt[tid] = tp;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) {
                rb[tid] += rb[tid + w];
                rt[tid] += rt[tid + w];
            }
            __syncthreads();
        }
        if (tid == 0 && D.nz >= 2) {
            ls.f[QV + sc][D.off2 + D.lin_col] += rb[0];
            ls.f[QV + sc][D.off2 + (int64_t)(D.nz - 1) * D.nx + D.lin_col] += rt[0];
        }
        __syncthreads();
    }
}
// The limiter sums of the stage reversed before, run by ka_pre's blocks past
// its ntiles tiles (one per member, when the launch has them): they add into
// that stage's input adjoint, which the next stage's ka_scalar reads after
// ka_pre.
struct LinSum {
    const double *linb, *lint;
    SSet<double> ls;
    int ntiles;
};

// ---- per-stage point values the gathers read many times -----------------------------
// q = rho_q / rho of the three scalars, u = 2 rho_u / (rho_i + rho_i+1) and
// w = rho_w / f2h(rho_k, rho_k+1) of the stage input, with the forward's
// expressions.  The stage's accumulators (`zero`) are cleared at
// the cell: ka_pre is the stage's first kernel and none of them is read or
// accumulated before it has run everywhere.
template <class RC>
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_pre(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<double> s, AdjW<RC> a,
           FillList zero, LinSum lin) {
    if ((int)blockIdx.x >= lin.ntiles) {  // the blocks past the tiles: limiter sums
        __shared__ double rb[256], rt[256];
        lin_sum(doms[blockIdx.x - lin.ntiles], lin.linb, lin.lint, lin.ls, threadIdx.y * AD_X + threadIdx.x, rb,
                rt);
        return;
    }
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
    const int nx = D.nx, nz = D.nz;
    if (i >= nx || k >= nz) return;
    const double *dzd = zv.dz + D.offz;
    const int64_t c = D.off2 + (int64_t)k * nx + i, cp = D.off2 + (int64_t)k * nx + wrapn(i + 1, nx);
    const double *rho = s.f[RHO];
    for (int sc = 0; sc < 3; ++sc) a.q[sc][c] = s.f[QV + sc][c] / rho[c];
    a.uf[c] = s.f[RHOU][c] * 2.0 / (rho[c] + rho[cp]);
    a.wf[c] = k < nz - 1 ? s.f[RHOW][c] / f2h(rho[c], rho[c + nx], dzd[k], dzd[k + 1]) : 0.0;
    for (int f = 0; f < zero.n; ++f) zero.p[f][c] = 0.0;
}

// ---- scalar transport + final update (moist.jl:1469-1539) ------------------------
// lo: adjoint of the stage output; writes ls (stage input, first writer of the
// stage), adds into ls0 (rho_q) and lph (q tendencies) (writes them in the
// step's first reverse stage, `first`: nothing cleared them); writes Lmx, Lmz and
// the substep adjoints Lrp, Lrtp, Lrwp, Lrup (= lo of rho, rho_theta, rho_w, rho_u).
//
// A block owns the AD_X x AD_Z cells of its tile.  Per scalar it stages in
// LDS q = rho_q / rho with a 3-cell halo and the scalar's output adjoint L
// with a 2-cell halo (rho with a 1-cell halo once for the three), evaluates
// the Koren partials of every x and z face that touches the tile once, and
// keeps per face the four products (L-difference * mass flux) * partial that
// its four cells gather; each cell then sums them in the order of the
// per-cell expression sum_faces LF * m * c[slot], so the sums are the same
// numbers as if every cell evaluated its four faces itself.
template <class RC>
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_scalar(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<double> s, SSet<double> lo,
              SSet<double> ls, SSet<double> ls0, SSet<double> lph, AdjW<RC> a, double *Lrup, int div, int first) {
    constexpr int QX = AD_X + 6, QZ = AD_Z + 6;  // q: columns i0-3 .. i0+AD_X+2, levels k0-3 .. k0+AD_Z+2
    constexpr int LX = AD_X + 4, LZ = AD_Z + 4;  // L: i0-2 .. i0+AD_X+1, k0-2 .. k0+AD_Z+1
    constexpr int RX = AD_X + 2, RZ = AD_Z + 2;  // rho: i0-1 .. i0+AD_X, k0-1 .. k0+AD_Z
    constexpr int FX = AD_X + 3;                 // x faces j+1/2, j = i0-2 .. i0+AD_X, of levels k0 .. k0+AD_Z-1
    constexpr int FZ = AD_Z + 3;                 // z faces kk+1/2, kk = k0-2 .. k0+AD_Z, of columns i0 ..
    constexpr int NT = AD_X * AD_Z;
    __shared__ double sQ[QZ][QX], sL[LZ][LX], sR[RZ][RX];
    __shared__ double sPx[AD_Z][FX][4], sSx[AD_Z][FX];  // x faces: (LF m) c[slot], the face value
    __shared__ double sPz[FZ][AD_X][4];                  // z faces: (LF m) c[slot]
    __shared__ double sOz[AD_Z][AD_X][3];                // each cell's own z face: value (0 if capped), clo, chi
    __shared__ double sMx[AD_Z][FX], sMz[FZ][AD_X];       // face mass fluxes, the same for the three scalars
    __shared__ double sG[QZ][4];                         // levels k0-3 ..: 1/dz, 1/dzh, 1/(dz+dz'), dz

    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD ordering (uniform per block)
    const Dom &D = doms[t.x];
    const int i0 = t.y, k0 = t.z;
    const int tx = threadIdx.x, tz = threadIdx.y, tid = tz * AD_X + tx;
    const int i = i0 + tx, k = k0 + tz;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    const bool cell = i < nx && k < nz;
    const double *dzd = zv.dz + D.offz;
    const double dx = D.dx, kap = D.kappa, dt = D.dt / double(div), rdx = 1.0 / dx;
    const double4 *gz = zv.geo + D.offz;  // per level: 1/dz, 1/(z_k+1 - z_k), 1/(dz_k + dz_k+1)
    auto G = [&](int kk, int ii) -> int64_t { return D.off2 + (int64_t)kk * nx + wrapn(ii, nx); };
    const int64_t c = G(min(k, ke), i);
    const double *rho = s.f[RHO];
    // once for the three scalars: rho of the tile with a 1-cell halo, the
    // face mass fluxes and the per-level geometry
    for (int e = tid; e < RZ * RX; e += NT) {
        const int r = e / RX, kk = k0 - 1 + r;
        (&sR[0][0])[e] = kk >= 0 && kk <= ke ? rho[G(kk, i0 - 1 + e - r * RX)] : 0.0;
    }
    for (int e = tid; e < AD_Z * FX; e += NT) {
        const int r = e / FX, kk = k0 + r;
        (&sMx[0][0])[e] = kk <= ke ? double(a.mom_x[G(kk, i0 - 2 + e - r * FX)]) : 0.0;
    }
    for (int e = tid; e < FZ * AD_X; e += NT) {
        const int g = e / AD_X, kk = k0 - 2 + g, ii = i0 + e - g * AD_X;
        (&sMz[0][0])[e] = kk >= 0 && kk <= ke - 1 && ii < nx ? double(a.mom_z[G(kk, ii)]) : 0.0;
    }
    if (tid < QZ) {
        const int kk = k0 - 3 + tid;
        if (kk >= 0 && kk <= ke) {
            const double4 gv = gz[kk];
            sG[tid][0] = gv.x;
            sG[tid][1] = gv.y;
            sG[tid][2] = gv.z;
            sG[tid][3] = dzd[kk];
        }
    }
    // the final update's adjoints, read before any store
    double lo_r = 0.0, lo_u = 0.0, lo_t = 0.0, lo_w = 0.0, rrho = 0.0;
    if (cell) {
        lo_r = lo.f[RHO][c];
        lo_u = lo.f[RHOU][c];
        lo_t = lo.f[RT][c];
        lo_w = lo.f[RHOW][c];
        rrho = 1.0 / rho[c];
    }
    double lrho = lo_r;  // final update: rho_out = rho + rho'
    double lmx = 0.0, lmz = 0.0;
    auto SQ = [&](int kk, int ii) { return sQ[kk - k0 + 3][ii - i0 + 3]; };
    auto SL = [&](int kk, int ii) { return sL[kk - k0 + 2][ii - i0 + 2]; };
    auto SR = [&](int kk, int ii) { return sR[kk - k0 + 1][ii - i0 + 1]; };
    // staged level geometry: 1/dz, 1/(z_k+1 - z_k), 1/(dz_k + dz_k+1), dz of level kk
    auto GX = [&](int kk) { return sG[kk - k0 + 3][0]; };
    auto GY = [&](int kk) { return sG[kk - k0 + 3][1]; };
    auto GZ = [&](int kk) { return sG[kk - k0 + 3][2]; };
    auto DZ = [&](int kk) { return sG[kk - k0 + 3][3]; };
    // a scalar's q and L tiles are loaded into registers one scalar ahead
    // (while the previous scalar's faces and cells are computed) and written
    // to LDS at the top of its iteration
    constexpr int NQ = (QZ * QX + NT - 1) / NT, NL = (LZ * LX + NT - 1) / NT;
    double qn[NQ], ln[NL], s0n = 0.0, phn = 0.0;
    auto prefetch = [&](int sc) {
        const double *qa = a.q[sc], *L = lo.f[QV + sc];
        // with the scalar's read-modify-write values, ahead of the previous
        // scalar's stores (a load issued after a store waits for it); in the
        // step's first reverse stage this kernel is their first writer
        if (cell && !first) {
            s0n = ls0.f[QV + sc][c];
            phn = lph.f[3 + sc][c];
        }
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            const int e = tid + u * NT, r = e / QX, kk = k0 - 3 + r;
            qn[u] = e < QZ * QX && kk >= 0 && kk <= ke ? qa[G(kk, i0 - 3 + e - r * QX)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int e = tid + u * NT, r = e / LX, kk = k0 - 2 + r;
            ln[u] = e < LZ * LX && kk >= 0 && kk <= ke ? L[G(kk, i0 - 2 + e - r * LX)] : 0.0;
        }
    };
    prefetch(0);
    // ---- the three scalars
#pragma unroll 1
    for (int sc = 0; sc < 3; ++sc) {
        const double *qr = s.f[QV + sc];
        // this scalar's read-modify-write values (prefetched)
        const double s0v = s0n, phv = phn;
#pragma unroll
        for (int u = 0; u < NQ; ++u)
            if (tid + u * NT < QZ * QX) (&sQ[0][0])[tid + u * NT] = qn[u];
#pragma unroll
        for (int u = 0; u < NL; ++u)
            if (tid + u * NT < LZ * LX) (&sL[0][0])[tid + u * NT] = ln[u];
        __syncthreads();
        if (sc < 2) prefetch(sc + 1);
        // x faces j+1/2 of the tile's levels (compute_limited_flux_x at j)
        for (int e = tid; e < AD_Z * FX; e += NT) {
            const int r = e / FX, f = e - r * FX;
            const int kk = k0 + r, j = i0 - 2 + f;
            double p[4] = {0.0, 0.0, 0.0, 0.0}, sf = 0.0;
            if (kk <= ke) {
                double cf[4];
                const double m = sMx[r][f];
                sf = kcoef_x(SQ(kk, j - 1), SQ(kk, j), SQ(kk, j + 1), SQ(kk, j + 2), m, cf);
                const double A = dt * rdx * (SL(kk, j + 1) - SL(kk, j)) * m;
#pragma unroll
                for (int q = 0; q < 4; ++q) p[q] = A * cf[q];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) sPx[r][f][q] = p[q];
            sSx[r][f] = sf;
        }
        // z faces kk+1/2 of the tile's columns (compute_limited_flux_z at kk)
        for (int e = tid; e < FZ * AD_X; e += NT) {
            const int g = e / AD_X, cc = e - g * AD_X;
            const int kk = k0 - 2 + g, ii = i0 + cc;
            double p[4] = {0.0, 0.0, 0.0, 0.0}, own[3] = {0.0, 0.0, 0.0};
            if (kk >= 0 && kk <= ke - 1 && ii < nx) {
                double cf[4], clo, chi;
                bool capped;
                const double m = sMz[g][cc];
                // the positivity limiters' rho_q (read through the linear index) only
                // matter on the bottom and top faces
                const bool bot = kk == 0, top = kk == ke - 1;
                const double qlo = bot ? qr[G(kk, D.lin_col)] : 0.0, qhi = top ? qr[G(kk + 1, D.lin_col)] : 0.0;
                const double sf = kcoef_z(kk > 0 ? SQ(kk - 1, ii) : 0.0, SQ(kk, ii), SQ(kk + 1, ii),
                                          kk + 2 <= ke ? SQ(kk + 2, ii) : 0.0, m, qlo, qhi, bot, top, DZ(kk),
                                          DZ(kk + 1), GZ(kk), D.dt, cf, capped, clo, chi);
                const double B = dt * (SL(kk + 1, ii) * GX(kk + 1) - SL(kk, ii) * GX(kk)) * m;
#pragma unroll
                for (int q = 0; q < 4; ++q) p[q] = B * cf[q];
                own[0] = capped ? 0.0 : sf;
                own[1] = clo;
                own[2] = chi;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) sPz[g][cc][q] = p[q];
            if (g >= 2 && g < 2 + AD_Z)
#pragma unroll
                for (int q = 0; q < 3; ++q) sOz[g - 2][cc][q] = own[q];
        }
        __syncthreads();
        if (cell) {
            double Lq = 0.0;
            // x faces j+1/2, j = i-2..i+1; cell i is slot i-j+1
#pragma unroll
            for (int o = 0; o < 4; ++o) Lq += sPx[tz][tx + o][3 - o];
            const double s_own_x = sSx[tz][tx + 2];
            // z faces kk+1/2, kk = k-2..k+1 within 0..ke-1; cell k is slot k-kk+1
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const int kk = k - 2 + o;
                if (kk < 0 || kk > ke - 1) continue;
                Lq += sPz[tz + o][tx][3 - o];
            }
            const double s_own_z = sOz[tz][tx][0], clo_own = sOz[tz][tx][1], chi_own = sOz[tz][tx][2];
            // diffusive x faces j+1/2, j = i-1, i: D = (rho_j + rho_j+1) / 2 (q_j+1 - q_j) / dx
            for (int j = i - 1; j <= i; ++j) {
                const double lD = kap * dt * rdx * (SL(k, j) - SL(k, j + 1));
                const double r0 = SR(k, j), r1 = SR(k, j + 1);
                Lq += lD * (r1 + r0) * 0.5 * rdx * (j == i - 1 ? 1.0 : -1.0);
                lrho += lD * 0.5 * (SQ(k, j + 1) - SQ(k, j)) * rdx;
            }
            // diffusive z faces kk+1/2, kk = k-1, k: D = f2h(rho) (q_kk+1 - q_kk) / (z_kk+1 - z_kk)
            for (int kk = k - 1; kk <= k; ++kk) {
                if (kk < 0 || kk > ke - 1) continue;
                const double lD = kap * dt * (SL(kk, i) * GX(kk) - SL(kk + 1, i) * GX(kk + 1));
                const double r0 = SR(kk, i), r1 = SR(kk + 1, i);
                const double rden = GZ(kk), rzz = GY(kk);
                const double rf = (r1 * DZ(kk) + r0 * DZ(kk + 1)) * rden;
                const double gq = (SQ(kk + 1, i) - SQ(kk, i)) * rzz;
                Lq += lD * rf * rzz * (kk == k - 1 ? 1.0 : -1.0);
                lrho += lD * gq * (kk == k - 1 ? DZ(kk) : DZ(kk + 1)) * rden;
            }
            const double Lqr = Lq * rrho;
            lrho -= Lqr * SQ(k, i);
            // this cell owns face i+1/2 (mom_x) and face k+1/2 (mom_z)
            lmx += dt * rdx * (SL(k, i + 1) - SL(k, i)) * s_own_x;
            if (k <= ke - 1) {
                const double lf = dt * (SL(k + 1, i) * GX(k + 1) - SL(k, i) * GX(k));
                lmz += lf * s_own_z;
                // linear-index positivity limiters: bottom face reads (0, lin_col), top face (ke, lin_col)
                if (k == 0) a.linb[D.offx * 3 + (int64_t)sc * nx + i] = lf * clo_own;
                if (k == ke - 1) a.lint[D.offx * 3 + (int64_t)sc * nx + i] = lf * chi_own;
            }
            const double Lc = SL(k, i);
            ls.f[QV + sc][c] = Lqr;
            ls0.f[QV + sc][c] = s0v + Lc;
            lph.f[3 + sc][c] = phv + dt * Lc;
        }
        __syncthreads();  // the next scalar restages sQ and sL
    }
    if (!cell) return;
    a.Lmx[c] = lmx;
    a.Lmz[c] = lmz;
    if (ke == 0) {
        for (int sc = 0; sc < 3; ++sc) {
            a.linb[D.offx * 3 + (int64_t)sc * nx + i] = 0.0;
            a.lint[D.offx * 3 + (int64_t)sc * nx + i] = 0.0;
        }
    }
    // final update (moist.jl:1522-1539): identity parts and the perturbation adjoints
    ls.f[RHO][c] = lrho;
    ls.f[RHOU][c] = lo_u;
    ls.f[RT][c] = lo_t;
    ls.f[RHOW][c] = lo_w;
    a.Lrp[c] = lo_r;
    Lrup[c] = lo_u;
    a.Lrtp[c] = lo_t;
    a.Lrwp[c] = k <= ke - 1 ? lo_w : 0.0;
}

// Sum of the positivity-limiter adjoints into rho_q (lin_sum), one block per member.
template <class RC>
__global__ void __launch_bounds__(256) ka_lin(const Dom *__restrict__ doms, AdjW<RC> a, SSet<double> ls) {
    __shared__ double rb[256], rt[256];
    lin_sum(doms[blockIdx.x], a.linb, a.lint, ls, threadIdx.x, rb, rt);
}

// ---- acoustic substep, vertical part (moist.jl:1377-1446) ---------------------------
// Coefficients of the w-p system of column i at row kk (0 <= kk <= ke-1):
// A = tridiag(a, b + 1, c) (implicit_correction!, moist.jl:974-1030).
// Divisions by grid quantities use the per-level reciprocals of the geometry
// table gz: {1/dz_k, 1/(z_k+1 - z_k), 1/(dz_k + dz_k+1), g/(dz_k + dz_k+1)}.
template <class RC> struct WpCoef {
    const Dom &D;
    const double *zd, *dzd;
    const double4 *gz;
    const RC *TH, *RT;
    int nx, i;
    double dt_tp2;
    __device__ int64_t G(int kk) const { return D.off2 + (int64_t)kk * nx + i; }
    __device__ double th(int kk) const { return (TH[G(kk)] * dzd[kk + 1] + TH[G(kk + 1)] * dzd[kk]) * gz[kk].z; }
    __device__ double z1(int kk) const { return zd[kk + 1] - zd[kk]; }
    __device__ double rz1(int kk) const { return gz[kk].y; }
    __device__ double a(int kk) const {  // kk > 0
        return -dt_tp2 * (RT[G(kk)] * th(kk - 1) * (gz[kk].y * gz[kk].x) - gz[kk].w);
    }
    __device__ double b(int kk) const {
        return dt_tp2 * (RT[G(kk)] * gz[kk].x + RT[G(kk + 1)] * gz[kk + 1].x) * th(kk) * gz[kk].y;
    }
    __device__ double c(int kk) const {  // kk < ke - 1
        return -dt_tp2 * (RT[G(kk + 1)] * th(kk + 1) * (gz[kk].y * gz[kk + 1].x) + gz[kk].w);
    }
};

// (1) Adjoint of rho_w' after the solve: the momentum average, the rho'
// and rho_theta' vertical updates (with the Q5 branch tests of the
// forward); theta adjoint of the rho_theta' vertical flux.  Per cell.  A
// cell reads levels k-2 .. k+2 of its column: the block stages the fields
// it reads for its tile and two levels either side in LDS, in their own
// types, and the cells evaluate the per-cell expressions on them.
template <class RC>
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_ac_v1(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<double> s, AdjW<RC> a,
             const RC *__restrict__ rwp_n, int div, int nsub) {
    constexpr int HZ = AD_Z + 4, NH = HZ * AD_X, NT = AD_X * AD_Z, NS = (NH + NT - 1) / NT;
    __shared__ RC sR[2][HZ][AD_X];      // rwp_n, theta
    __shared__ double sD[4][HZ][AD_X];  // rho_w, Lrtp, Lrp, Lmz
    __shared__ double4 sG[HZ];          // per-level geometry of the tile's levels
    __shared__ double sDZ[HZ];
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i0 = t.y, k0 = t.z;
    const int i = i0 + threadIdx.x, k = k0 + threadIdx.y, tid = threadIdx.y * AD_X + threadIdx.x;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    // this cell's read-modify-write values and the tile's level geometry,
    // loaded with the staged fields (before the barrier)
    const bool cell_ok = i < nx && k < nz;
    const int64_t qc = D.off2 + (int64_t)min(k, ke) * nx + min(i, nx - 1);
    const double lth0 = cell_ok ? a.Lth[qc] : 0.0, lrwp0 = cell_ok && k <= ke - 1 ? a.Lrwp[qc] : 0.0;
    if (tid < HZ) {
        const int kk = min(max(k0 - 2 + tid, 0), ke);
        sG[tid] = zv.geo[D.offz + kk];
        sDZ[tid] = zv.dz[D.offz + kk];
    }
    {
        RC xr[NS][2];
        double xd[NS][4];
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT, r = e / AD_X, kk = k0 - 2 + r, ii = i0 + e - r * AD_X;
            const bool ok = e < NH && kk >= 0 && kk <= ke && ii < nx;
            const int64_t q = ok ? D.off2 + (int64_t)kk * nx + ii : 0;
            xr[u][0] = ok ? rwp_n[q] : RC(0);
            xr[u][1] = ok ? a.theta[q] : RC(0);
            xd[u][0] = ok ? s.f[RHOW][q] : 0.0;
            xd[u][1] = ok ? a.Lrtp[q] : 0.0;
            xd[u][2] = ok ? a.Lrp[q] : 0.0;
            xd[u][3] = ok ? a.Lmz[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT;
            if (e < NH) {
#pragma unroll
                for (int f = 0; f < 2; ++f) (&sR[f][0][0])[e] = xr[u][f];
#pragma unroll
                for (int f = 0; f < 4; ++f) (&sD[f][0][0])[e] = xd[u][f];
            }
        }
    }
    __syncthreads();
    if (i >= nx || k >= nz) return;
    // per level (staged): 1/dz, 1/(z_k+1 - z_k), 1/(dz_k + dz_k+1), and dz
    auto gzl = [&](int kk) { return sG[kk - k0 + 2]; };
    auto dzl = [&](int kk) { return sDZ[kk - k0 + 2]; };
    const double EPS = 1e-12, rns = 1.0 / double(nsub);
    const double dt_s = D.dt / double(div) / double(nsub);
    auto G = [&](int kk) -> int64_t { return D.off2 + (int64_t)kk * nx + i; };
    // staged values of level kk of this column
    auto RWN = [&](int kk) { return sR[0][kk - k0 + 2][threadIdx.x]; };
    auto TH = [&](int kk) { return sR[1][kk - k0 + 2][threadIdx.x]; };
    auto RW = [&](int kk) { return sD[0][kk - k0 + 2][threadIdx.x]; };
    auto LRTP = [&](int kk) { return sD[1][kk - k0 + 2][threadIdx.x]; };
    auto LRP = [&](int kk) { return sD[2][kk - k0 + 2][threadIdx.x]; };
    auto LMZ = [&](int kk) { return sD[3][kk - k0 + 2][threadIdx.x]; };
    auto THv = [&](int kk) { return kk >= 0 && kk <= ke ? TH(kk) : 0.0; };
    // face f+1/2 as fkp of cell f: positive branch tests rwp_n[f], negative s.rho_w[f]
    auto brp = [&](int f, bool &pos, bool &neg) {
        pos = RWN(f) > EPS && f > 0;
        neg = !pos && RW(f) < -EPS && f < ke - 1;
    };
    // face f+1/2 as fkm of cell f+1: both branches test s.rho_w[f]
    auto brm = [&](int f, bool &pos, bool &neg) {
        const double sw = RW(f);
        pos = sw > EPS && f + 1 > 1;
        neg = !pos && sw < -EPS && f + 1 < ke;
    };
    double lth = 0.0;
    for (int f = k - 2; f <= k + 1; ++f) {
        if (f < 0 || f > ke - 1) continue;
        const double rden = gzl(f).z;
        const double w0 = dzl(f + 1) * rden, w1 = dzl(f) * rden;
        const double w = RWN(f);
        bool pos, neg;
        brp(f, pos, neg);
        lth += -dt_s * gzl(f).x * LRTP(f) * w * upw_coef(pos, neg, k - f, w0, w1);
        brm(f, pos, neg);
        lth += dt_s * gzl(f + 1).x * LRTP(f + 1) * w * upw_coef(pos, neg, k - f, w0, w1);
    }
    double lrwp = 0.0;
    if (k <= ke - 1) {
        const double rden = gzl(k).z;
        const double w0 = dzl(k + 1) * rden, w1 = dzl(k) * rden;
        bool pos, neg;
        brp(k, pos, neg);
        const double gp = upw_val(pos, neg, THv(k - 1), THv(k), THv(k + 1), THv(k + 2), w0, w1);
        brm(k, pos, neg);
        const double gm = upw_val(pos, neg, THv(k - 1), THv(k), THv(k + 1), THv(k + 2), w0, w1);
        lrwp = LMZ(k) * rns - dt_s * gzl(k).x * LRTP(k) * gp + dt_s * gzl(k + 1).x * LRTP(k + 1) * gm -
               dt_s * gzl(k).x * LRP(k) + dt_s * gzl(k + 1).x * LRP(k + 1);
    }
    // read-modify-writes last, after every load of the kernel
    a.Lth[G(k)] = lth0 + lth;
    if (k <= ke - 1) a.Lrwp[G(k)] = lrwp0 + lrwp;
}
// Inclusive scans over the CH chunks of one column inside a wave (lane =
// column + CW * chunk, CW = 64 / CH): affine maps x -> A x + B from the
// bottom chunk up or from the top chunk down, and 2x2 matrices (Moebius
// maps, scaled to unit max norm).
template <int CH> __device__ __forceinline__ void scan_affine_up(double &A, double &B, int ch) {
    constexpr int CW = 64 / CH;
#pragma unroll
    for (int d = 1; d < CH; d <<= 1) {
        const double Ap = __shfl_up(A, CW * d), Bp = __shfl_up(B, CW * d);
        if (ch >= d) {
            B = A * Bp + B;
            A = A * Ap;
        }
    }
}
template <int CH> __device__ __forceinline__ void scan_affine_down(double &A, double &B, int ch) {
    constexpr int CW = 64 / CH;
#pragma unroll
    for (int d = 1; d < CH; d <<= 1) {
        const double Ap = __shfl_down(A, CW * d), Bp = __shfl_down(B, CW * d);
        if (ch + d < CH) {
            B = A * Bp + B;
            A = A * Ap;
        }
    }
}
__device__ __forceinline__ void mob_norm(double &m00, double &m01, double &m10, double &m11) {
    const double s = fmax(fmax(fabs(m00), fabs(m01)), fmax(fabs(m10), fabs(m11)));
    if (s > 0.0) {
        const double r = 1.0 / s;
        m00 *= r;
        m01 *= r;
        m10 *= r;
        m11 *= r;
    }
}

// (2) Transposed solve A^T y = Lx of every column.  A wave owns CW = 64/CH
// adjacent columns, each split into CH chunks of rows (lane = column +
// CW * chunk), so a load instruction touches CH runs of CW adjacent columns.
// With the forward factorization A = L U (L: diagonal m_j, sub-diagonal a_j;
// U: unit, super-diagonal c'_j): U^T w = Lx is w_j = Lx_j - c'_j-1 w_j-1,
// L^T y = w is y_j = (w_j - a_j+1 y_j+1) / m_j; both are affine scans over
// the chunks.  `factor` (first reverse substep of a stage) first computes
// c'_j = c_j / (b_j + 1 - a_j c'_j-1) with a Moebius scan and stores c', 1/m
// and a.  groups: {member, first column} of every wave.
template <int CH, class RC>
__global__ void __launch_bounds__(256)
    ka_ac_solve(const Dom *__restrict__ doms, const int2 *__restrict__ groups, int ngroups, ZVec zv, AdjW<RC> a, int div,
                int nsub, int factor) {
    constexpr int CW = 64 / CH;
    const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (wid >= ngroups) return;
    const int2 t = groups[wid];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int nx = D.nx, nz = D.nz, n = nz - 1;
    if (n <= 0) return;
    const int col = lane % CW, ch = lane / CW;
    const int i = t.y + col;
    const bool live = i < nx;
    const double dt_s = D.dt / double(div) / double(nsub);
    const double fact_tp = 0.5 * (1.0 + D.beta);
    const WpCoef<RC> W{D, zv.z + D.offz, zv.dz + D.offz, zv.geo + D.offz, a.theta, a.rt2, nx, live ? i : 0,
                   (dt_s * fact_tp) * (dt_s * fact_tp)};
    const int L = (n + CH - 1) / CH, j0 = ch * L, j1 = live ? min(j0 + L, n) : j0;
    // Chunks of at most LR rows (one column per wave up to nz = 257) keep
    // their c', 1/m, a, right-hand side and w in registers: one load phase
    // and one store phase instead of a memory round trip between the scans.
    // The arithmetic is the same, in the same order.  (L is the same for the
    // whole wave: its columns belong to one member.)
    constexpr int LR = 4;
    const bool regs = L <= LR;
    double cpv[LR], rmv[LR], acv[LR];
#pragma unroll
    for (int q = 0; q < LR; ++q) cpv[q] = rmv[q] = acv[q] = 0.0;
    if (factor) {
        double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
#pragma unroll 4
        for (int j = j0; j < j1; ++j) {
            const double aj = j > 0 ? W.a(j) : 0.0, bj = W.b(j) + 1.0, cj = j < n - 1 ? W.c(j) : 0.0;
            // M_j = [[0, c], [-a, b+1]] applied after the chunk composite
            const double n00 = cj * m10, n01 = cj * m11;
            const double n10 = -aj * m00 + bj * m10, n11 = -aj * m01 + bj * m11;
            m00 = n00;
            m01 = n01;
            m10 = n10;
            m11 = n11;
            mob_norm(m00, m01, m10, m11);
        }
#pragma unroll
        for (int d = 1; d < CH; d <<= 1) {
            const double p00 = __shfl_up(m00, CW * d), p01 = __shfl_up(m01, CW * d), p10 = __shfl_up(m10, CW * d),
                         p11 = __shfl_up(m11, CW * d);
            if (ch >= d) {
                const double n00 = m00 * p00 + m01 * p10, n01 = m00 * p01 + m01 * p11;
                const double n10 = m10 * p00 + m11 * p10, n11 = m10 * p01 + m11 * p11;
                m00 = n00;
                m01 = n01;
                m10 = n10;
                m11 = n11;
                mob_norm(m00, m01, m10, m11);
            }
        }
        // c' entering this chunk: the chunks below applied to c'_-1 = 0, i.e. (p, q) = (0, 1)
        const double e01 = __shfl_up(m01, CW), e11 = __shfl_up(m11, CW);
        double cprev = ch == 0 ? 0.0 : e01 / e11;
        if (regs) {
#pragma unroll
            for (int q = 0; q < LR; ++q) {
                const int j = j0 + q;
                if (j < j1) {
                    const double aj = j > 0 ? W.a(j) : 0.0;
                    const double m = W.b(j) + 1.0 - aj * cprev;
                    const double rm = 1.0 / m;
                    const double cp = j < n - 1 ? W.c(j) * rm : 0.0;
                    a.cpr[W.G(j)] = cp;
                    a.rmi[W.G(j)] = rm;
                    a.aco[W.G(j)] = aj;
                    cpv[q] = cp;
                    rmv[q] = rm;
                    acv[q] = aj;
                    cprev = cp;
                }
            }
        } else {
#pragma unroll 4
            for (int j = j0; j < j1; ++j) {
                const double aj = j > 0 ? W.a(j) : 0.0;
                const double m = W.b(j) + 1.0 - aj * cprev;
                const double rm = 1.0 / m;
                const double cp = j < n - 1 ? W.c(j) * rm : 0.0;
                a.cpr[W.G(j)] = cp;
                a.rmi[W.G(j)] = rm;
                a.aco[W.G(j)] = aj;
                cprev = cp;
            }
        }
    }
    if (regs) {
        double lx[LR];  // right-hand side
#pragma unroll
        for (int q = 0; q < LR; ++q) {
            const int j = j0 + q;
            lx[q] = j < j1 ? a.Lrwp[W.G(j)] : 0.0;
            if (!factor && j < j1) {
                cpv[q] = a.cpr[W.G(j)];
                rmv[q] = a.rmi[W.G(j)];
                acv[q] = a.aco[W.G(j)];
            }
        }
        // c' of row j0-1 (last row of the chunk below, which is full) and a of
        // row j1 (first row of the chunk above), from the neighbouring lanes
        double cp_last = 0.0;
#pragma unroll
        for (int q = 0; q < LR; ++q)
            if (q == L - 1) cp_last = cpv[q];
        const double cp_below = __shfl_up(cp_last, CW), a_above = __shfl_down(acv[0], CW);
        // forward sweep w_j = Lx_j - c'_j-1 w_j-1
        double A = 1.0, B = 0.0;
#pragma unroll
        for (int q = 0; q < LR; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                const double al = j > 0 ? -(q == 0 ? cp_below : cpv[q > 0 ? q - 1 : 0]) : 0.0;
                B = al * B + lx[q];
                A = al * A;
            }
        }
        scan_affine_up<CH>(A, B, ch);
        double wv = __shfl_up(B, CW);
        if (ch == 0) wv = 0.0;
#pragma unroll
        for (int q = 0; q < LR; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                const double al = j > 0 ? -(q == 0 ? cp_below : cpv[q > 0 ? q - 1 : 0]) : 0.0;
                wv = al * wv + lx[q];
                lx[q] = wv;  // w
            }
        }
        // back substitution y_j = rm_j w_j - rm_j a_j+1 y_j+1, chunks from the top
        A = 1.0;
        B = 0.0;
#pragma unroll
        for (int q = LR - 1; q >= 0; --q) {
            const int j = j0 + q;
            if (j < j1) {
                const double rm = rmv[q];
                const double al = j < n - 1 ? -rm * (q + 1 < L ? acv[q < LR - 1 ? q + 1 : 0] : a_above) : 0.0;
                B = al * B + rm * lx[q];
                A = al * A;
            }
        }
        scan_affine_down<CH>(A, B, ch);
        double yv = __shfl_down(B, CW);
        if (ch == CH - 1) yv = 0.0;
#pragma unroll
        for (int q = LR - 1; q >= 0; --q) {
            const int j = j0 + q;
            if (j < j1) {
                const double rm = rmv[q];
                const double al = j < n - 1 ? -rm * (q + 1 < L ? acv[q < LR - 1 ? q + 1 : 0] : a_above) : 0.0;
                yv = al * yv + rm * lx[q];
                a.Y[W.G(j)] = yv;
            }
        }
        return;
    }
    // forward sweep w_j = Lx_j - c'_j-1 w_j-1 (each sweep unrolled by four:
    // the rows' loads do not depend on the recurrence, so four rows' loads are
    // in flight together instead of one row's at a time)
    double A = 1.0, B = 0.0;
#pragma unroll 4
    for (int j = j0; j < j1; ++j) {
        const double al = j > 0 ? -a.cpr[W.G(j - 1)] : 0.0;
        B = al * B + a.Lrwp[W.G(j)];
        A = al * A;
    }
    scan_affine_up<CH>(A, B, ch);
    double wv = __shfl_up(B, CW);
    if (ch == 0) wv = 0.0;
#pragma unroll 4
    for (int j = j0; j < j1; ++j) {
        const double al = j > 0 ? -a.cpr[W.G(j - 1)] : 0.0;
        wv = al * wv + a.Lrwp[W.G(j)];
        a.Y[W.G(j)] = wv;  // w, replaced by y below
    }
    // back substitution y_j = rm_j w_j - rm_j a_j+1 y_j+1, chunks from the top
    A = 1.0;
    B = 0.0;
#pragma unroll 4
    for (int j = j1 - 1; j >= j0; --j) {
        const double rm = a.rmi[W.G(j)];
        const double al = j < n - 1 ? -rm * a.aco[W.G(j + 1)] : 0.0;
        B = al * B + rm * a.Y[W.G(j)];
        A = al * A;
    }
    scan_affine_down<CH>(A, B, ch);
    double yv = __shfl_down(B, CW);
    if (ch == CH - 1) yv = 0.0;
#pragma unroll 4
    for (int j = j1 - 1; j >= j0; --j) {
        const double rm = a.rmi[W.G(j)];
        const double al = j < n - 1 ? -rm * a.aco[W.G(j + 1)] : 0.0;
        yv = al * yv + rm * a.Y[W.G(j)];
        a.Y[W.G(j)] = yv;
    }
}

// (3) Matrix and right-hand-side adjoints of the solve, the adjoint of
// rho_w' before it, drwp = dt_s (R_rho_w - rt2h buoy + grav) and the explicit
// part (rp_tp = rp + f inc, rp += inc; the same for rho_theta').  Per cell.
// A cell reads levels k-2 .. k+1 of its column and, for the off-centred
// rho_theta', the columns i-1 and i+1: the block stages the eight fields it
// reads for its tile (levels k0-2 .. k0+AD_Z, columns i0-1 .. i0+AD_X) in
// LDS, in their own types, and the cells evaluate the per-cell expressions
// on them.
template <class RC>
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_ac_v2(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, AdjW<RC> a,
             const RC *__restrict__ rtp_s, const RC *__restrict__ rup_n, const RC *__restrict__ rwp_s,
             const RC *__restrict__ rwp_n, int div, int nsub) {
    constexpr int HX = AD_X + 2, HZ = AD_Z + 3, NH = HX * HZ, NT = AD_X * AD_Z, NS = (NH + NT - 1) / NT;
    __shared__ double sY[HZ][HX];
    __shared__ RC sF[7][HZ][HX];  // rwp_n, rwp_s, theta, rt2pres, rup_n, R_rho_theta, rtp_s
    constexpr int GZ = AD_Z + 6;  // levels k0-3 .. k0+AD_Z+2 of the staged geometry
    __shared__ double4 sG[GZ];
    __shared__ double sDZ[GZ];
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i0 = t.y, k0 = t.z;
    const int i = i0 + threadIdx.x, k = k0 + threadIdx.y, tid = threadIdx.y * AD_X + threadIdx.x;
    const int nx = D.nx, nz = D.nz, n = nz - 1;
    if (n <= 0) return;  // uniform per block
    auto Gx = [&](int kk, int ii) -> int64_t { return D.off2 + (int64_t)kk * nx + wrapn(ii, nx); };
    // this cell's read-modify-write values and the tile's level geometry,
    // loaded with the staged fields (before the barrier)
    const bool cell_ok = i < nx && k < nz;
    const int64_t q = D.off2 + (int64_t)min(k, n) * nx + min(i, nx - 1);
    const double Lrt0 = cell_ok ? a.Lrt[q] : 0.0, Lth0 = cell_ok ? a.Lth[q] : 0.0, lr = cell_ok ? a.Lrp[q] : 0.0,
                 lt = cell_ok ? a.Lrtp[q] : 0.0, LRr0 = cell_ok ? a.LRr[q] : 0.0, LRt0 = cell_ok ? a.LRt[q] : 0.0,
                 LRw0 = cell_ok && k <= n - 1 ? a.LRw[q] : 0.0;
    if (tid < GZ) {
        const int kk = min(max(k0 - 3 + tid, 0), n);
        sG[tid] = zv.geo[D.offz + kk];
        sDZ[tid] = zv.dz[D.offz + kk];
    }
    {
        const RC *src[7] = {rwp_n, rwp_s, a.theta, a.rt2, rup_n, a.R_rt, rtp_s};
        double y[NS];
        RC x[NS][7];
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT, r = e / HX, kk = k0 - 2 + r;
            const bool ok = e < NH && kk >= 0 && kk <= n;
            const int64_t q = ok ? Gx(kk, i0 - 1 + e - r * HX) : 0;
            y[u] = ok ? a.Y[q] : 0.0;
#pragma unroll
            for (int f = 0; f < 7; ++f) x[u][f] = ok ? src[f][q] : RC(0);
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT;
            if (e < NH) {
                (&sY[0][0])[e] = y[u];
#pragma unroll
                for (int f = 0; f < 7; ++f) (&sF[f][0][0])[e] = x[u][f];
            }
        }
    }
    __syncthreads();
    if (i >= nx || k >= nz) return;
    // the staged per-level geometry, indexed by level: gz[kk] and dzd[kk]
    struct LG {
        const double4 *g;
        int o;
        __device__ const double4 &operator[](int kk) const { return g[kk - o]; }
    };
    struct LD {
        const double *g;
        int o;
        __device__ double operator[](int kk) const { return g[kk - o]; }
    };
    const LG gz{sG, k0 - 3};
    const LD dzd{sDZ, k0 - 3};
    const double GRAV = K<double>::GRAV, dx = D.dx, rdx = 1.0 / dx;
    const double dt_s = D.dt / double(div) / double(nsub);
    const double fact_tp = 0.5 * (1.0 + D.beta), fact_tm = 0.5 * (1.0 - D.beta);
    const double fact = fact_tm / fact_tp;
    const double dt_tp2 = (dt_s * fact_tp) * (dt_s * fact_tp);
    // staged values of level kk (column i, or ii for theta and rup_n)
    auto SF = [&](int f, int kk, int ii) -> RC { return sF[f][kk - k0 + 2][ii - i0 + 1]; };
    auto RWN = [&](int kk) { return SF(0, kk, i); };
    auto RWS = [&](int kk) { return SF(1, kk, i); };
    auto TH = [&](int kk, int ii) { return SF(2, kk, ii); };
    auto RT = [&](int kk) { return SF(3, kk, i); };
    auto RUN = [&](int kk, int ii) { return SF(4, kk, ii); };
    auto RRT = [&](int kk) { return SF(5, kk, i); };
    auto RTS = [&](int kk) { return SF(6, kk, i); };
    // the w-p system's coefficients of row kk (implicit_correction!)
    auto th = [&](int kk) { return (TH(kk, i) * dzd[kk + 1] + TH(kk + 1, i) * dzd[kk]) * gz[kk].z; };
    auto Wa = [&](int kk) { return -dt_tp2 * (RT(kk) * th(kk - 1) * (gz[kk].y * gz[kk].x) - gz[kk].w); };
    auto Wb = [&](int kk) { return dt_tp2 * (RT(kk) * gz[kk].x + RT(kk + 1) * gz[kk + 1].x) * th(kk) * gz[kk].y; };
    auto Wc = [&](int kk) { return -dt_tp2 * (RT(kk + 1) * th(kk + 1) * (gz[kk].y * gz[kk + 1].x) + gz[kk].w); };
    auto yv = [&](int j) { return j >= 0 && j <= n - 1 ? sY[j - k0 + 2][i - i0 + 1] : 0.0; };
    auto xv = [&](int j) { return j >= 0 && j <= n - 1 ? double(RWN(j)) : 0.0; };
    auto La = [&](int j) { return j >= 1 && j <= n - 1 ? -yv(j) * xv(j - 1) - fact * yv(j) * RWS(j) : 0.0; };
    auto Lb = [&](int j) { return j >= 0 && j <= n - 1 ? -yv(j) * xv(j) - fact * yv(j) * RWS(j) : 0.0; };
    auto Lc = [&](int j) { return j >= 0 && j < n - 1 ? -yv(j) * xv(j + 1) - fact * yv(j) * RWS(j + 1) : 0.0; };
    // adjoint of the face theta th_j (j = 0..n-1)
    auto Lthc = [&](int j) {
        if (j < 0 || j > n - 1) return 0.0;
        double v = Lb(j) * dt_tp2 * (RT(j) * gz[j].x + RT(j + 1) * gz[j + 1].x) * gz[j].y;
        if (j + 1 <= n - 1) v += La(j + 1) * (-dt_tp2 * RT(j + 1) * (gz[j + 1].y * gz[j + 1].x));
        if (j >= 1) v += Lc(j - 1) * (-dt_tp2 * RT(j) * (gz[j - 1].y * gz[j].x));
        return v;
    };
    // off-centred rho_theta' of the forward explicit part
    auto rtp_tp = [&](int kk) {
        const double fip = (TH(kk, i + 1) + TH(kk, i)) * 0.5 * RUN(kk, i);
        const double fim = (TH(kk, i) + TH(kk, i - 1)) * 0.5 * RUN(kk, i - 1);
        const double inc = dt_s * (RRT(kk) + (-(fip - fim) * rdx));
        return RTS(kk) + fact_tp * inc;
    };
    double lrt = 0.0, lth = 0.0, Lrp_tp = 0.0, Lrtp_tp = 0.0, lws = 0.0, lRw = 0.0;
    if (k <= n - 1) {
        const double y = yv(k);
        const double thk = th(k), rz1 = gz[k].y, rzd = rz1 * gz[k].x;
        lrt += Lb(k) * dt_tp2 * thk * rzd;
        if (k >= 1) lrt += La(k) * (-dt_tp2 * th(k - 1) * rzd);
        const double buoy = (rtp_tp(k + 1) - rtp_tp(k)) * rz1;
        lrt += -0.5 * dt_s * y * buoy;
        const double rt2h = (RT(k) + RT(k + 1)) * 0.5;
        Lrtp_tp += dt_s * y * rt2h * rz1;
        Lrp_tp += -dt_s * y * GRAV * 0.5;
        lRw = dt_s * y;
        lth += Lthc(k) * dzd[k + 1] * gz[k].z;
        // rho_w' before the solve
        lws = y * (1.0 - fact * Wb(k));
        if (k >= 1) lws += -fact * Wa(k) * y - fact * Wc(k - 1) * yv(k - 1);
    }
    if (k >= 1) {
        const int kb = k - 1;
        const double yb = yv(kb);
        const double rz1 = gz[kb].y, rzd = rz1 * gz[k].x;
        lrt += Lb(kb) * dt_tp2 * th(kb) * rzd;
        if (kb < n - 1) lrt += Lc(kb) * (-dt_tp2 * th(k) * rzd);
        const double buoy = (rtp_tp(k) - rtp_tp(kb)) * rz1;
        lrt += -0.5 * dt_s * yb * buoy;
        const double rt2h = (RT(kb) + RT(k)) * 0.5;
        Lrtp_tp += -dt_s * yb * rt2h * rz1;
        Lrp_tp += -dt_s * yb * GRAV * 0.5;
        lth += Lthc(kb) * dzd[kb] * gz[kb].z;
    }
    const double lir = lr + fact_tp * Lrp_tp, lit = lt + fact_tp * Lrtp_tp;
    if (k <= n - 1) {
        a.LRw[q] = LRw0 + lRw;
        a.Lrwp[q] = lws;
    }
    a.Lrt[q] = Lrt0 + lrt;
    a.Lth[q] = Lth0 + lth;
    a.Lrp[q] = lr + Lrp_tp;
    a.Lrtp[q] = lt + Lrtp_tp;
    a.Linc_r[q] = lir;
    a.Linc_t[q] = lit;
    a.LRr[q] = LRr0 + dt_s * lir;
    a.LRt[q] = LRt0 + dt_s * lit;
}

// ---- acoustic substep, vertical part, one kernel per column ------------------------
// (1) + (2) + (3) of one reverse substep in one launch: NW waves (one or two)
// per column, thread t owning the consecutive levels [t Lc, t Lc + Lc) and the
// solve's rows [t Ls, t Ls + Ls) (Lc = ceil(nz / 64 NW), Ls = ceil((nz - 1) /
// 64 NW) <= 4).  The column's inputs of the three parts are staged in LDS
// once; the solve's right-hand side and its solution stay in LDS instead of a
// global round trip, and two launches per substep are saved.  The per-level
// geometry (dz and the four grid factors) is staged in LDS with the column, so
// the loops over neighbouring faces wait on no global load.  With one wave the
// arithmetic of every value is that of ka_ac_v1, ka_ac_solve (one column per
// wave, CH = 64) and ka_ac_v2, in the same order.  Two waves per column (taller
// columns: half the levels per thread, and twice the waves on few columns)
// split the solve into 128 chunks: the scans run within each wave by shuffles,
// and the lower wave's aggregate (bottom-up scans) or the upper wave's
// (top-down) is composed in through LDS, which regroups the scans' products
// (rounding differences only).  cols: {member, column} per block.
// Dynamic LDS: AC_COL_ARRAYS doubles per level of the tallest member.
constexpr int AC_COL_ARRAYS = 16;
constexpr int AC_COL_MAX_NZ = 256;  // Lc, Ls <= 4: the solve's rows stay in registers
// LCM: levels a thread may own; NW: waves per column (members up to 64 levels:
// <1, 1>; up to 128: <1, 2>; up to 256: <2, 2>)
template <class RC, int LCM, int NW>
__global__ void __launch_bounds__(64 * NW)
    ka_ac_col(const Dom *__restrict__ doms, const int2 *__restrict__ cols, ZVec zv, SSet<double> s, AdjW<RC> a,
              const RC *__restrict__ rtp_s, const RC *__restrict__ rup_n, const RC *__restrict__ rwp_s,
              const RC *__restrict__ rwp_n, int div, int nsub, int factor, int max_ka) {
    extern __shared__ double lcol[];
    const int2 t = cols[blockIdx.x];
    if (t.x < 0) return;  // padding entry of the XCD ordering (uniform per block)
    const Dom &D = doms[t.x];
    const int nx = D.nx, nz = D.nz, ke = nz - 1, n = nz - 1;
    if (n <= 0) return;
    static_assert(NW == 1 || NW == 2, "one or two waves per column");
    const int i = t.y, lane = threadIdx.x, wl = lane & 63, wave = lane >> 6;
    // the scans' values that cross the wave boundary (NW = 2): the Moebius
    // aggregate of the lower wave (0-3), the bottom-up affine aggregate of the
    // lower wave (4-5), the top-down one of the upper wave (6-7), c' of the
    // lower wave's last chunk (8) and a of the upper wave's first row (9)
    __shared__ double xw[10];
    double *sRWN = lcol, *sRWS = lcol + max_ka, *sRW = lcol + 2 * max_ka, *sTH = lcol + 3 * max_ka,
           *sRT = lcol + 4 * max_ka, *sLRTP = lcol + 5 * max_ka, *sLRP = lcol + 6 * max_ka,
           *sLMZ = lcol + 7 * max_ka, *sTP = lcol + 8 * max_ka, *sRHS = lcol + 9 * max_ka, *sY = lcol + 10 * max_ka;
    double *sGX = lcol + 11 * max_ka, *sGY = lcol + 12 * max_ka, *sGZ = lcol + 13 * max_ka, *sGW = lcol + 14 * max_ka,
           *sDZ = lcol + 15 * max_ka;
    const double *dzd = zv.dz + D.offz;
    const double4 *gz = zv.geo + D.offz;
    // the staged geometry of level kk
    auto gzl = [&](int kk) { return make_double4(sGX[kk], sGY[kk], sGZ[kk], sGW[kk]); };
    const double EPS = 1e-12, rns = 1.0 / double(nsub);
    const double GRAV = K<double>::GRAV, rdx = 1.0 / D.dx;
    const double dt_s = D.dt / double(div) / double(nsub);
    const double fact_tp = 0.5 * (1.0 + D.beta), fact_tm = 0.5 * (1.0 - D.beta);
    const double fact = fact_tm / fact_tp;
    const double dt_tp2 = (dt_s * fact_tp) * (dt_s * fact_tp);
    const RC *TH = a.theta;
    auto G = [&](int kk) -> int64_t { return D.off2 + (int64_t)kk * nx + i; };
    auto Gx = [&](int kk, int ii) -> int64_t { return D.off2 + (int64_t)kk * nx + wrapn(ii, nx); };
    const int Lc = (nz + 64 * NW - 1) / (64 * NW), kc0 = lane * Lc;

    // ---- stage the column (and read this lane's read-modify-write values)
    double lth0[LCM], lrwp0[LCM], lrt0[LCM], lr0[LCM], lt0[LCM], lRr0[LCM], lRt0[LCM], lRw0[LCM];
#pragma unroll
    for (int u = 0; u < LCM; ++u) {
        const int k = kc0 + u;
        lth0[u] = lrwp0[u] = lrt0[u] = lr0[u] = lt0[u] = lRr0[u] = lRt0[u] = lRw0[u] = 0.0;
        if (u < Lc && k < nz) {
            const int64_t q = G(k);
            sRWN[k] = rwp_n[q];
            sRWS[k] = rwp_s[q];
            sRW[k] = s.f[RHOW][q];
            sTH[k] = TH[q];
            sRT[k] = a.rt2[q];
            sLRTP[k] = a.Lrtp[q];
            sLRP[k] = a.Lrp[q];
            sLMZ[k] = a.Lmz[q];
            const double4 g = gz[k];
            sGX[k] = g.x;
            sGY[k] = g.y;
            sGZ[k] = g.z;
            sGW[k] = g.w;
            sDZ[k] = dzd[k];
            // off-centred rho_theta' of the forward explicit part (ka_ac_v2's rtp_tp)
            const double fip = (TH[Gx(k, i + 1)] + TH[q]) * 0.5 * rup_n[q];
            const double fim = (TH[q] + TH[Gx(k, i - 1)]) * 0.5 * rup_n[Gx(k, i - 1)];
            const double inc = dt_s * (a.R_rt[q] + (-(fip - fim) * rdx));
            sTP[k] = rtp_s[q] + fact_tp * inc;
            lth0[u] = a.Lth[q];
            lrwp0[u] = k <= ke - 1 ? a.Lrwp[q] : 0.0;
            lrt0[u] = a.Lrt[q];
            lr0[u] = a.Lrp[q];
            lt0[u] = a.Lrtp[q];
            lRr0[u] = a.LRr[q];
            lRt0[u] = a.LRt[q];
            lRw0[u] = k <= n - 1 ? a.LRw[q] : 0.0;
        }
    }
    __syncthreads();

    // ---- (1) adjoint of rho_w' after the solve (ka_ac_v1)
    auto THv = [&](int kk) { return kk >= 0 && kk <= ke ? sTH[kk] : 0.0; };
    auto brp = [&](int f, bool &pos, bool &neg) {
        pos = sRWN[f] > EPS && f > 0;
        neg = !pos && sRW[f] < -EPS && f < ke - 1;
    };
    auto brm = [&](int f, bool &pos, bool &neg) {
        const double sw = sRW[f];
        pos = sw > EPS && f + 1 > 1;
        neg = !pos && sw < -EPS && f + 1 < ke;
    };
    double lth1[LCM];
#pragma unroll
    for (int u = 0; u < LCM; ++u) {
        const int k = kc0 + u;
        lth1[u] = 0.0;
        if (!(u < Lc && k < nz)) continue;
        double lth = 0.0;
        for (int f = k - 2; f <= k + 1; ++f) {
            if (f < 0 || f > ke - 1) continue;
            const double rden = gzl(f).z;
            const double w0 = sDZ[f + 1] * rden, w1 = sDZ[f] * rden;
            const double w = sRWN[f];
            bool pos, neg;
            brp(f, pos, neg);
            lth += -dt_s * gzl(f).x * sLRTP[f] * w * upw_coef(pos, neg, k - f, w0, w1);
            brm(f, pos, neg);
            lth += dt_s * gzl(f + 1).x * sLRTP[f + 1] * w * upw_coef(pos, neg, k - f, w0, w1);
        }
        lth1[u] = lth0[u] + lth;
        if (k <= ke - 1) {
            const double rden = gzl(k).z;
            const double w0 = sDZ[k + 1] * rden, w1 = sDZ[k] * rden;
            bool pos, neg;
            brp(k, pos, neg);
            const double gp = upw_val(pos, neg, THv(k - 1), THv(k), THv(k + 1), THv(k + 2), w0, w1);
            brm(k, pos, neg);
            const double gm = upw_val(pos, neg, THv(k - 1), THv(k), THv(k + 1), THv(k + 2), w0, w1);
            const double lrwp = sLMZ[k] * rns - dt_s * gzl(k).x * sLRTP[k] * gp + dt_s * gzl(k + 1).x * sLRTP[k + 1] * gm -
                                dt_s * gzl(k).x * sLRP[k] + dt_s * gzl(k + 1).x * sLRP[k + 1];
            sRHS[k] = lrwp0[u] + lrwp;
        }
    }
    __syncthreads();

    // ---- (2) transposed solve A^T y = rhs (ka_ac_solve, one column per wave)
    auto th = [&](int kk) { return (sTH[kk] * sDZ[kk + 1] + sTH[kk + 1] * sDZ[kk]) * gzl(kk).z; };
    auto Wa = [&](int kk) { return -dt_tp2 * (sRT[kk] * th(kk - 1) * (gzl(kk).y * gzl(kk).x) - gzl(kk).w); };
    auto Wb = [&](int kk) { return dt_tp2 * (sRT[kk] * gzl(kk).x + sRT[kk + 1] * gzl(kk + 1).x) * th(kk) * gzl(kk).y; };
    auto Wc = [&](int kk) { return -dt_tp2 * (sRT[kk + 1] * th(kk + 1) * (gzl(kk).y * gzl(kk + 1).x) + gzl(kk).w); };
    {
        constexpr int CH = 64 * NW, LR = 4;
        const int ch = lane;
        const int L = (n + CH - 1) / CH, j0 = ch * L, j1 = min(j0 + L, n);
        double cpv[LR], rmv[LR], acv[LR];
#pragma unroll
        for (int q = 0; q < LR; ++q) cpv[q] = rmv[q] = acv[q] = 0.0;
        if (factor) {
            double m00 = 1.0, m01 = 0.0, m10 = 0.0, m11 = 1.0;
            for (int j = j0; j < j1; ++j) {
                const double aj = j > 0 ? Wa(j) : 0.0, bj = Wb(j) + 1.0, cj = j < n - 1 ? Wc(j) : 0.0;
                const double n00 = cj * m10, n01 = cj * m11;
                const double n10 = -aj * m00 + bj * m10, n11 = -aj * m01 + bj * m11;
                m00 = n00;
                m01 = n01;
                m10 = n10;
                m11 = n11;
                mob_norm(m00, m01, m10, m11);
            }
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const double p00 = __shfl_up(m00, d), p01 = __shfl_up(m01, d), p10 = __shfl_up(m10, d),
                             p11 = __shfl_up(m11, d);
                if (wl >= d) {
                    const double n00 = m00 * p00 + m01 * p10, n01 = m00 * p01 + m01 * p11;
                    const double n10 = m10 * p00 + m11 * p10, n11 = m10 * p01 + m11 * p11;
                    m00 = n00;
                    m01 = n01;
                    m10 = n10;
                    m11 = n11;
                    mob_norm(m00, m01, m10, m11);
                }
            }
            if constexpr (NW == 2) {
                // the upper wave's composites applied after the lower wave's aggregate
                if (lane == 63) {
                    xw[0] = m00;
                    xw[1] = m01;
                    xw[2] = m10;
                    xw[3] = m11;
                }
                __syncthreads();
                if (wave == 1) {
                    const double p00 = xw[0], p01 = xw[1], p10 = xw[2], p11 = xw[3];
                    const double n00 = m00 * p00 + m01 * p10, n01 = m00 * p01 + m01 * p11;
                    const double n10 = m10 * p00 + m11 * p10, n11 = m10 * p01 + m11 * p11;
                    m00 = n00;
                    m01 = n01;
                    m10 = n10;
                    m11 = n11;
                    mob_norm(m00, m01, m10, m11);
                }
            }
            double e01 = __shfl_up(m01, 1), e11 = __shfl_up(m11, 1);
            if (NW == 2 && lane == 64) {
                e01 = xw[1];
                e11 = xw[3];
            }
            double cprev = ch == 0 ? 0.0 : e01 / e11;
#pragma unroll
            for (int q = 0; q < LR; ++q) {
                const int j = j0 + q;
                if (j < j1) {
                    const double aj = j > 0 ? Wa(j) : 0.0;
                    const double m = Wb(j) + 1.0 - aj * cprev;
                    const double rm = 1.0 / m;
                    const double cp = j < n - 1 ? Wc(j) * rm : 0.0;
                    a.cpr[G(j)] = cp;
                    a.rmi[G(j)] = rm;
                    a.aco[G(j)] = aj;
                    cpv[q] = cp;
                    rmv[q] = rm;
                    acv[q] = aj;
                    cprev = cp;
                }
            }
        }
        double lx[LR];
#pragma unroll
        for (int q = 0; q < LR; ++q) {
            const int j = j0 + q;
            lx[q] = j < j1 ? sRHS[j] : 0.0;
            if (!factor && j < j1) {
                cpv[q] = a.cpr[G(j)];
                rmv[q] = a.rmi[G(j)];
                acv[q] = a.aco[G(j)];
            }
        }
        double cp_last = 0.0;
#pragma unroll
        for (int q = 0; q < LR; ++q)
            if (q == L - 1) cp_last = cpv[q];
        double cp_below = __shfl_up(cp_last, 1), a_above = __shfl_down(acv[0], 1);
        if constexpr (NW == 2) {
            if (lane == 63) xw[8] = cp_last;
            if (lane == 64) xw[9] = acv[0];
            __syncthreads();
            if (lane == 64) cp_below = xw[8];
            if (lane == 63) a_above = xw[9];
        }
        double A = 1.0, B = 0.0;
#pragma unroll
        for (int q = 0; q < LR; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                const double al = j > 0 ? -(q == 0 ? cp_below : cpv[q > 0 ? q - 1 : 0]) : 0.0;
                B = al * B + lx[q];
                A = al * A;
            }
        }
        scan_affine_up<64>(A, B, wl);
        if constexpr (NW == 2) {
            if (lane == 63) {
                xw[4] = A;
                xw[5] = B;
            }
            __syncthreads();
            if (wave == 1) {
                B = A * xw[5] + B;
                A = A * xw[4];
            }
        }
        double wv = __shfl_up(B, 1);
        if (NW == 2 && lane == 64) wv = xw[5];
        if (ch == 0) wv = 0.0;
#pragma unroll
        for (int q = 0; q < LR; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                const double al = j > 0 ? -(q == 0 ? cp_below : cpv[q > 0 ? q - 1 : 0]) : 0.0;
                wv = al * wv + lx[q];
                lx[q] = wv;
            }
        }
        A = 1.0;
        B = 0.0;
#pragma unroll
        for (int q = LR - 1; q >= 0; --q) {
            const int j = j0 + q;
            if (j < j1) {
                const double rm = rmv[q];
                const double al = j < n - 1 ? -rm * (q + 1 < L ? acv[q < LR - 1 ? q + 1 : 0] : a_above) : 0.0;
                B = al * B + rm * lx[q];
                A = al * A;
            }
        }
        scan_affine_down<64>(A, B, wl);
        if constexpr (NW == 2) {
            if (lane == 64) {
                xw[6] = A;
                xw[7] = B;
            }
            __syncthreads();
            if (wave == 0) {
                B = A * xw[7] + B;
                A = A * xw[6];
            }
        }
        double yv = __shfl_down(B, 1);
        if (NW == 2 && lane == 63) yv = xw[7];
        if (ch == CH - 1) yv = 0.0;
#pragma unroll
        for (int q = LR - 1; q >= 0; --q) {
            const int j = j0 + q;
            if (j < j1) {
                const double rm = rmv[q];
                const double al = j < n - 1 ? -rm * (q + 1 < L ? acv[q < LR - 1 ? q + 1 : 0] : a_above) : 0.0;
                yv = al * yv + rm * lx[q];
                sY[j] = yv;
            }
        }
    }
    __syncthreads();

    // ---- (3) matrix, right-hand-side and explicit-part adjoints (ka_ac_v2)
    auto yv = [&](int j) { return j >= 0 && j <= n - 1 ? sY[j] : 0.0; };
    auto xv = [&](int j) { return j >= 0 && j <= n - 1 ? sRWN[j] : 0.0; };
    auto La = [&](int j) { return j >= 1 && j <= n - 1 ? -yv(j) * xv(j - 1) - fact * yv(j) * sRWS[j] : 0.0; };
    auto Lb = [&](int j) { return j >= 0 && j <= n - 1 ? -yv(j) * xv(j) - fact * yv(j) * sRWS[j] : 0.0; };
    auto Lc_ = [&](int j) { return j >= 0 && j < n - 1 ? -yv(j) * xv(j + 1) - fact * yv(j) * sRWS[j + 1] : 0.0; };
    auto Lthc = [&](int j) {
        if (j < 0 || j > n - 1) return 0.0;
        double v = Lb(j) * dt_tp2 * (sRT[j] * gzl(j).x + sRT[j + 1] * gzl(j + 1).x) * gzl(j).y;
        if (j + 1 <= n - 1) v += La(j + 1) * (-dt_tp2 * sRT[j + 1] * (gzl(j + 1).y * gzl(j + 1).x));
        if (j >= 1) v += Lc_(j - 1) * (-dt_tp2 * sRT[j] * (gzl(j - 1).y * gzl(j).x));
        return v;
    };
#pragma unroll
    for (int u = 0; u < LCM; ++u) {
        const int k = kc0 + u;
        if (!(u < Lc && k < nz)) continue;
        const int64_t q = G(k);
        const double lr = lr0[u], lt = lt0[u];
        double lrt = 0.0, lth = 0.0, Lrp_tp = 0.0, Lrtp_tp = 0.0, lws = 0.0, lRw = 0.0;
        if (k <= n - 1) {
            const double y = yv(k);
            const double thk = th(k), rz1 = gzl(k).y, rzd = rz1 * gzl(k).x;
            lrt += Lb(k) * dt_tp2 * thk * rzd;
            if (k >= 1) lrt += La(k) * (-dt_tp2 * th(k - 1) * rzd);
            const double buoy = (sTP[k + 1] - sTP[k]) * rz1;
            lrt += -0.5 * dt_s * y * buoy;
            // the records' sum in their own type, as ka_ac_v2 forms it
            const double rt2h = (RC(sRT[k]) + RC(sRT[k + 1])) * 0.5;
            Lrtp_tp += dt_s * y * rt2h * rz1;
            Lrp_tp += -dt_s * y * GRAV * 0.5;
            lRw = dt_s * y;
            lth += Lthc(k) * sDZ[k + 1] * gzl(k).z;
            lws = y * (1.0 - fact * Wb(k));
            if (k >= 1) lws += -fact * Wa(k) * y - fact * Wc(k - 1) * yv(k - 1);
        }
        if (k >= 1) {
            const int kb = k - 1;
            const double yb = yv(kb);
            const double rz1 = gzl(kb).y, rzd = rz1 * gzl(k).x;
            lrt += Lb(kb) * dt_tp2 * th(kb) * rzd;
            if (kb < n - 1) lrt += Lc_(kb) * (-dt_tp2 * th(k) * rzd);
            const double buoy = (sTP[k] - sTP[kb]) * rz1;
            lrt += -0.5 * dt_s * yb * buoy;
            const double rt2h = (RC(sRT[kb]) + RC(sRT[k])) * 0.5;
            Lrtp_tp += -dt_s * yb * rt2h * rz1;
            Lrp_tp += -dt_s * yb * GRAV * 0.5;
            lth += Lthc(kb) * sDZ[kb] * gzl(kb).z;
        }
        const double lir = lr + fact_tp * Lrp_tp, lit = lt + fact_tp * Lrtp_tp;
        if (k <= n - 1) {
            a.LRw[q] = lRw0[u] + lRw;
            a.Lrwp[q] = lws;
        }
        a.Lrt[q] = lrt0[u] + lrt;
        a.Lth[q] = lth1[u] + lth;
        a.Lrp[q] = lr + Lrp_tp;
        a.Lrtp[q] = lt + Lrtp_tp;
        a.Linc_r[q] = lir;
        a.Linc_t[q] = lit;
        a.LRr[q] = lRr0[u] + dt_s * lir;
        a.LRt[q] = lRt0[u] + dt_s * lit;
    }
}

// ---- acoustic substep, horizontal part (moist.jl:1360-1408) -----------------------
// Per cell.  Lrup_in: adjoint of rho_u' after this substep (from later
// substeps and the final update); Lrup_out: adjoint before it.  pend: the
// damping term's adjoint for rho_theta' at the start of the previous substep.
// A cell reads columns i-1 .. i+1 of its level: the block stages the fields
// it reads for its tile and a column either side in LDS, in their own types,
// and the cells evaluate the per-cell expressions on them.
template <class RC>
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_ac_hor(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, AdjW<RC> a,
              const double *__restrict__ Lrup_in, double *__restrict__ Lrup_out, const RC *__restrict__ rtp_s,
              const RC *__restrict__ rtp_prev, const RC *__restrict__ rup_n, int div, int nsub, int first_substep) {
    constexpr int HX = AD_X + 2, NH = HX * AD_Z, NT = AD_X * AD_Z, NS = (NH + NT - 1) / NT;
    __shared__ RC sR[5][AD_Z][HX];      // theta, rt2pres, rtp_s, rtp_prev, rup_n
    __shared__ double sD[4][AD_Z][HX];  // Linc_t, Linc_r, Lrup_in, Lmx
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i0 = t.y, k0 = t.z;
    const int i = i0 + threadIdx.x, k = k0 + threadIdx.y, tid = threadIdx.y * AD_X + threadIdx.x;
    const int nx = D.nx, nz = D.nz;
    auto Gk = [&](int kk, int ii) -> int64_t { return D.off2 + (int64_t)kk * nx + wrapn(ii, nx); };
    {
        const RC *sr[5] = {a.theta, a.rt2, rtp_s, rtp_prev, rup_n};
        const double *sd[4] = {a.Linc_t, a.Linc_r, Lrup_in, a.Lmx};
        RC xr[NS][5];
        double xd[NS][4];
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT, r = e / HX, kk = k0 + r;
            const bool ok = e < NH && kk < nz;
            const int64_t q = ok ? Gk(kk, i0 - 1 + e - r * HX) : 0;
#pragma unroll
            for (int f = 0; f < 5; ++f) xr[u][f] = ok ? sr[f][q] : RC(0);
#pragma unroll
            for (int f = 0; f < 4; ++f) xd[u][f] = ok ? sd[f][q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT;
            if (e < NH) {
#pragma unroll
                for (int f = 0; f < 5; ++f) (&sR[f][0][0])[e] = xr[u][f];
#pragma unroll
                for (int f = 0; f < 4; ++f) (&sD[f][0][0])[e] = xd[u][f];
            }
        }
    }
    __syncthreads();
    if (i >= nx || k >= nz) return;
    const double dx = D.dx, gam = D.gamma_d, rdx = 1.0 / dx;
    const double dt_s = D.dt / double(div) / double(nsub), rns = 1.0 / double(nsub);
    // staged values of column ii (i-1 .. i+1) of this level
    const int r = threadIdx.y;
    auto TH = [&](int ii) { return sR[0][r][ii - i0 + 1]; };
    auto RT = [&](int ii) { return sR[1][r][ii - i0 + 1]; };
    auto RTS = [&](int ii) { return sR[2][r][ii - i0 + 1]; };
    auto RTP = [&](int ii) { return sR[3][r][ii - i0 + 1]; };
    auto RUN = [&](int ii) { return sR[4][r][ii - i0 + 1]; };
    auto LIT = [&](int ii) { return sD[0][r][ii - i0 + 1]; };
    auto LIR = [&](int ii) { return sD[1][r][ii - i0 + 1]; };
    auto LRUP = [&](int ii) { return sD[2][r][ii - i0 + 1]; };
    auto LMX = [&](int ii) { return sD[3][r][ii - i0 + 1]; };
    // adjoint of rho_u'(j) after the substep, all contributions of this substep included
    auto lrup = [&](int j) {
        const double lF = dt_s * rdx * (LIT(j + 1) - LIT(j));
        return LRUP(j) + LMX(j) * rns - dt_s * rdx * LIR(j) + dt_s * rdx * LIR(j + 1) + lF * 0.5 * (TH(j) + TH(j + 1));
    };
    const int64_t c = Gk(k, i);
    // this cell's read-modify-write values, read before any store
    const double Lth0 = a.Lth[c], LRu0 = a.LRu[c], Lrt0 = a.Lrt[c], Lrtp0 = a.Lrtp[c], pend0 = a.pend[c];
    // theta adjoint of the rho_theta' flux (faces i+1/2 and i-1/2)
    double lth;
    {
        const double lFi = dt_s * rdx * (LIT(i + 1) - LIT(i));
        const double lFm = dt_s * rdx * (LIT(i) - LIT(i - 1));
        lth = 0.5 * (lFi * RUN(i) + lFm * RUN(i - 1));
    }
    const double Li = lrup(i), Lm = lrup(i - 1);
    // rho_theta' at the start of this substep: add the damping adjoint from the next substep
    double lt = Lrtp0 + pend0;
    // rho_u' update: rup_n = rup_s + dt_s (R_rho_u - rt2h ((1 + g) gx - g gxo))
    auto Gx = [&](int j) {
        const double gx = (RTS(j + 1) - RTS(j)) * rdx;
        const double gxo = (RTP(j + 1) - RTP(j)) * rdx;
        return gx + gam * (gx - gxo);
    };
    const double lrt = -0.5 * dt_s * (Li * Gx(i) + Lm * Gx(i - 1));
    const double lGi = -dt_s * (RT(i) + RT(i + 1)) * 0.5 * Li;
    const double lGm = -dt_s * (RT(i - 1) + RT(i)) * 0.5 * Lm;
    lt += (1.0 + gam) * rdx * (lGm - lGi);
    const double lo = -gam * rdx * (lGm - lGi);
    if (first_substep) lt += lo;
    a.Lth[c] = Lth0 + lth;
    a.LRu[c] = LRu0 + dt_s * Li;
    a.Lrt[c] = Lrt0 + lrt;
    a.pend[c] = first_substep ? 0.0 : lo;
    a.Lrtp[c] = lt;
    Lrup_out[c] = Li;
}

// ---- perturbation initialisation + slow tendencies (moist.jl:1071-1348) ----------
// Gathers every contribution of R_rho, R_rho_u, R_rho_w, R_rho_theta to the
// stage input (rho, rho_u, rho_w), to the face velocities (LU, LW), pres and
// theta; adds the perturbation adjoints (rho' = s0 - s etc.) and the momentum
// copies.  Its s0 and physics-tendency adjoints (ls0, lph) accumulate over the
// three stages; in the step's first reverse stage (`first`) it is their first
// writer and does not read them.
//
// Every neighbour value a cell reads lies within two cells of it in x and z:
// the block stages the ten fields it gathers from (the four tendency
// adjoints, rho, rho_u, rho_w, u and w at the faces, theta) for its tile with
// a 2-cell halo in LDS, loads issued together, and the cells then read them
// there; the arithmetic is the per-cell gather's, value for value.
template <class RC>
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_slow(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<double> s, SSet<double> ls,
            SSet<double> ls0, SSet<double> lph, AdjW<RC> a, const double *__restrict__ Lrup, int first) {
    constexpr int HX = AD_X + 4, HZ = AD_Z + 4, NH = HX * HZ, NT = AD_X * AD_Z, NS = (NH + NT - 1) / NT;
    constexpr int NF = 10;
    __shared__ double sh[NF][HZ][HX];  // LRu, LRw, LRt, LRr, rho, rho_u, rho_w, u, w, theta
    // per-level geometry of levels k0-2 .. k0+AD_Z+1: {1/dz, 1/(z_k+1 - z_k), 1/(dz_k + dz_k+1), g/..}, z, zf, dz
    __shared__ double4 sG[HZ];
    __shared__ double sZ[HZ], sZF[HZ], sDZ[HZ];
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i0 = t.y, k0 = t.z;
    const int i = i0 + threadIdx.x, k = k0 + threadIdx.y, tid = threadIdx.y * AD_X + threadIdx.x;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    auto G = [&](int kk, int ii) -> int64_t { return D.off2 + (int64_t)kk * nx + wrapn(ii, nx); };
    // this cell's read-modify-write values (stored at the end) and the other
    // per-cell inputs, loaded with the staged tile (before the barrier)
    const bool cell_ok = i < nx && k < nz;
    const int64_t c = D.off2 + (int64_t)min(k, ke) * nx + min(i, nx - 1);
    double s0r = 0.0, s0u = 0.0, s0t = 0.0, s0w = 0.0, sr = 0.0, su = 0.0, sw = 0.0, st = 0.0, p0 = 0.0, p1 = 0.0,
           p2 = 0.0, Lth0 = 0.0, Lrp = 0.0, Lrt = 0.0, Lrw = 0.0, Lru = 0.0, Lmx = 0.0, Lmz = 0.0;
    if (cell_ok) {
        if (!first) {  // in the step's first reverse stage this kernel is their first writer
            s0r = ls0.f[RHO][c];
            s0u = ls0.f[RHOU][c];
            s0t = ls0.f[RT][c];
            s0w = ls0.f[RHOW][c];
            p0 = lph.f[0][c];
            p1 = lph.f[1][c];
            p2 = lph.f[2][c];
        }
        sr = ls.f[RHO][c];
        su = ls.f[RHOU][c];
        sw = ls.f[RHOW][c];
        st = ls.f[RT][c];
        Lth0 = a.Lth[c];
        Lrp = a.Lrp[c];
        Lrt = a.Lrtp[c];
        Lrw = k <= ke - 1 ? a.Lrwp[c] : 0.0;
        Lru = Lrup[c];
        Lmx = a.Lmx[c];
        Lmz = a.Lmz[c];
    }
    if (tid < HZ) {
        const int kk = min(max(k0 - 2 + tid, 0), ke);
        sG[tid] = zv.geo[D.offz + kk];
        sZ[tid] = zv.z[D.offz + kk];
        sZF[tid] = zv.zf[D.offz + kk];
        sDZ[tid] = zv.dz[D.offz + kk];
    }
    {
        const double *src[NF] = {a.LRu, a.LRw, a.LRt, a.LRr, s.f[RHO], s.f[RHOU], s.f[RHOW], a.uf, a.wf, nullptr};
        double x[NS][NF];
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT, r = e / HX, kk = k0 - 2 + r;
            const bool ok = e < NH && kk >= 0 && kk <= ke;
            const int64_t q = ok ? G(kk, i0 - 2 + e - r * HX) : 0;
#pragma unroll
            for (int f = 0; f < NF - 1; ++f) x[u][f] = ok ? src[f][q] : 0.0;
            x[u][NF - 1] = ok ? double(a.theta[q]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int e = tid + u * NT;
            if (e < NH)
#pragma unroll
                for (int f = 0; f < NF; ++f) (&sh[f][0][0])[e] = x[u][f];
        }
    }
    __syncthreads();
    if (!cell_ok) return;
    // the staged per-level geometry, indexed by level (kk - 2 <= ... <= k + 2)
    struct LG {
        const double4 *g;
        int o;
        __device__ const double4 &operator[](int kk) const { return g[kk - o]; }
    };
    struct LD {
        const double *g;
        int o;
        __device__ double operator[](int kk) const { return g[kk - o]; }
    };
    const LD zd{sZ, k0 - 2}, zfd{sZF, k0 - 2}, dzd{sDZ, k0 - 2};
    const double dx = D.dx, EPS = 1e-12, GRAV = K<double>::GRAV, nu = D.nu, rdx = 1.0 / dx;
    const LG gz{sG, k0 - 2};  // per level: 1/dz, 1/(z_k+1 - z_k), 1/(dz_k + dz_k+1)
    // the staged value of field f at (kk, ii), |kk - k| <= 2, |ii - i| <= 2
    auto SH = [&](int f, int kk, int ii) { return sh[f][kk - k0 + 2][ii - i0 + 2]; };
    auto LRU = [&](int kk, int ii) { return SH(0, kk, ii); };
    auto LRT = [&](int kk, int ii) { return SH(2, kk, ii); };
    auto LRR = [&](int kk, int ii) { return SH(3, kk, ii); };
    auto RHOv = [&](int kk, int ii) { return SH(4, kk, ii); };
    auto RU = [&](int kk, int ii) { return SH(5, kk, ii); };
    auto RW = [&](int kk, int ii) { return SH(6, kk, ii); };
    auto U = [&](int kk, int ii) { return SH(7, kk, ii); };
    auto W = [&](int kk, int ii) { return SH(8, kk, ii); };  // kk <= ke - 1
    auto THv = [&](int kk, int ii) { return SH(9, kk, ii); };
    double l_r = 0.0, l_u = 0.0, l_w = 0.0, LUc = 0.0, LWc = 0.0, lp = 0.0, lth = 0.0;

    // this cell's tendency adjoints (staged)
    const double LRr_c = LRR(k, i), LRu_c = LRU(k, i), LRt_c = LRT(k, i);
    // perturbation init and momentum copies
    s0r += Lrp;
    s0u += Lru;
    s0t += Lrt;
    s0w += Lrw;
    l_r -= Lrp;
    l_u += -Lru + Lmx;
    st -= Lrt;
    l_w += -Lrw + Lmz;
    // physics tendencies enter R_rho, R_rho_u, R_rho_theta additively
    p0 += LRr_c;
    p1 += LRu_c;
    p2 += LRt_c;

    // ---- R_rho (moist.jl:1079-1086)
    l_u += -LRr_c * rdx + LRR(k, i + 1) * rdx;
    if (k < ke) l_w += -LRr_c * gz[k].x + LRR(k + 1, i) * gz[k + 1].x;

    // ---- R_rho_u (moist.jl:1098-1179)
    // x flux F(j) between u-faces j and j+1, m = (rho_u_j + rho_u_j+1)/2
    for (int j = i - 2; j <= i + 1; ++j) {
        const double lF = (LRU(k, j + 1) - LRU(k, j)) * rdx;
        const double m = 0.5 * (RU(k, j) + RU(k, j + 1));
        const bool pos = m > EPS, neg = m < -EPS;
        LUc += lF * m * upw_coef(pos, neg, i - j, 0.5, 0.5);
        if (j == i - 1 || j == i) {
            const double g = upw_val(pos, neg, U(k, j - 1), U(k, j), U(k, j + 1), U(k, j + 2), 0.5, 0.5);
            l_u += lF * g * 0.5;
        }
    }
    // z flux at corner (kk+1/2, face i), kk = k-2..k+1 within 0..ke-1
    for (int kk = k - 2; kk <= k + 1; ++kk) {
        if (kk < 0 || kk > ke - 1) continue;
        const double lF = LRU(kk + 1, i) * gz[kk + 1].x - LRU(kk, i) * gz[kk].x;
        const double m = 0.5 * (RW(kk, i) + RW(kk, i + 1));
        const bool pos = m > EPS && kk > 0, neg = !pos && m < -EPS && kk < ke - 1;
        const double rden = gz[kk].z;
        LUc += lF * m * upw_coef(pos, neg, k - kk, dzd[kk + 1] * rden, dzd[kk] * rden);
    }
    // rho_w at (k, i) enters the corner fluxes (k+1/2, i) and (k+1/2, i-1)
    if (k <= ke - 1) {
        for (int ii = i - 1; ii <= i; ++ii) {
            const double lF = LRU(k + 1, ii) * gz[k + 1].x - LRU(k, ii) * gz[k].x;
            const double m = 0.5 * (RW(k, ii) + RW(k, ii + 1));
            const bool pos = m > EPS && k > 0, neg = !pos && m < -EPS && k < ke - 1;
            const double rden = gz[k].z;
            const double g = upw_val(pos, neg, k > 0 ? U(k - 1, ii) : 0.0, U(k, ii), U(k + 1, ii),
                                     k + 2 <= ke ? U(k + 2, ii) : 0.0, dzd[k + 1] * rden, dzd[k] * rden);
            l_w += lF * g * 0.5;
        }
    }
    // viscous x flux Dx(j) = (U_j+1 - U_j) / dx * rho_j+1, j = i-1, i
    for (int j = i - 1; j <= i; ++j) {
        const double lD = nu * (LRU(k, j) - LRU(k, j + 1)) * rdx;
        LUc += lD * RHOv(k, j + 1) * rdx * (j == i - 1 ? 1.0 : -1.0);
        if (j == i - 1) l_r += lD * (U(k, i) - U(k, i - 1)) * rdx;
    }
    // viscous z flux Dz(kk) = (U_kk+1 - U_kk) / zz * rho at the corner, kk = k-1, k
    for (int kk = k - 1; kk <= k; ++kk) {
        if (kk < 0 || kk > ke - 1) continue;
        const double lD = nu * (LRU(kk, i) * gz[kk].x - LRU(kk + 1, i) * gz[kk + 1].x);
        const double rc = 0.25 * (RHOv(kk, i) + RHOv(kk + 1, i) + RHOv(kk, i + 1) + RHOv(kk + 1, i + 1));
        LUc += lD * rc * gz[kk].y * (kk == k - 1 ? 1.0 : -1.0);
    }
    // rho(k, i) in the corner densities of the z viscous fluxes (kk, ii), kk = k-1, k; ii = i-1, i
    for (int kk = k - 1; kk <= k; ++kk) {
        if (kk < 0 || kk > ke - 1) continue;
        for (int ii = i - 1; ii <= i; ++ii) {
            const double lD = nu * (LRU(kk, ii) * gz[kk].x - LRU(kk + 1, ii) * gz[kk + 1].x);
            l_r += lD * (U(kk + 1, ii) - U(kk, ii)) * gz[kk].y * 0.25;
        }
    }
    // pressure gradient (p_i+1 - p_i) / dx
    lp += LRu_c * rdx - LRU(k, i - 1) * rdx;

    // ---- R_rho_w (moist.jl:1190-1279), cells k' <= ke - 1
    const double sponge_base = (1.0 - D.zf_sponge) * D.H;
    auto LRw = [&](int kk, int ii) { return SH(1, kk, ii); };
    if (k <= ke - 1) {
        // fip(k, j): m = f2h(rho_u(k,j), rho_u(k+1,j), z_k, z_k+1), reads W(j-1..j+2)
        const double rzs = 1.0 / (zd[k] + zd[k + 1]);
        for (int j = i - 2; j <= i + 1; ++j) {
            const double lF = -LRw(k, j) * rdx;
            const double m = (RU(k, j) * zd[k + 1] + RU(k + 1, j) * zd[k]) * rzs;
            LWc += lF * m * upw_coef(m > EPS, m < -EPS, i - j, 0.5, 0.5);
        }
        // fim(k, j): m = f2h(rho_u(k,j-1), rho_u(k+1,j-1), dz), reads W(j-2..j+1): interface between j-1 and j
        const double rds = gz[k].z;
        for (int j = i - 1; j <= i + 2; ++j) {
            const double lF = LRw(k, j) * rdx;
            const double m = (RU(k, j - 1) * dzd[k + 1] + RU(k + 1, j - 1) * dzd[k]) * rds;
            LWc += lF * m * upw_coef(m > EPS, m < -EPS, i - (j - 1), 0.5, 0.5);
        }
        // viscous x flux Ex(j) = (W_j+1 - W_j) / dx * rho at corner (k+1/2, j+1/2), j = i-1, i
        for (int j = i - 1; j <= i; ++j) {
            const double lE = nu * (LRw(k, j) - LRw(k, j + 1)) * rdx;
            const double rc = 0.25 * (RHOv(k, j) + RHOv(k + 1, j) + RHOv(k, j + 1) + RHOv(k + 1, j + 1));
            LWc += lE * rc * rdx * (j == i - 1 ? 1.0 : -1.0);
        }
        // sponge
        double sigma = 0.0;
        if (zfd[k] > sponge_base) sigma = (GRAV / 30.0) * ((zfd[k] - sponge_base) / (D.zf_sponge * D.H));
        l_w += -sigma * LRw(k, i);
    }
    // z flux Fz(kk) between w levels kk and kk+1 (kk < ke - 1), shared by cells kk and kk+1
    for (int kk = k - 2; kk <= k + 1; ++kk) {
        if (kk < 0 || kk >= ke - 1) continue;
        const double lF = LRw(kk + 1, i) * gz[kk + 1].y - LRw(kk, i) * gz[kk].y;
        const double rds = gz[kk].z;
        const double m = (RW(kk, i) * dzd[kk + 1] + RW(kk + 1, i) * dzd[kk]) * rds;
        const bool pos = m > EPS && kk > 0, neg = !pos && m < -EPS && kk < ke - 2;
        if (k <= ke - 1) LWc += lF * m * upw_coef(pos, neg, k - kk, dzd[kk + 1] * rds, dzd[kk] * rds);
        if (kk == k || kk == k - 1) {
            const double g = upw_val(pos, neg, kk > 0 ? W(kk - 1, i) : 0.0, W(kk, i), W(kk + 1, i),
                                     kk + 2 <= ke - 1 ? W(kk + 2, i) : 0.0, dzd[kk + 1] * rds, dzd[kk] * rds);
            l_w += lF * g * (kk == k ? dzd[kk + 1] : dzd[kk]) * rds;
        }
    }
    // viscous z flux Ez(kk) = (W_kk+1 - W_kk) / dz_kk+1 * rho_kk+1, kk = k-1, k (kk < ke - 1)
    for (int kk = k - 1; kk <= k; ++kk) {
        if (kk < 0 || kk >= ke - 1) continue;
        const double lE = nu * (LRw(kk, i) * gz[kk].y - LRw(kk + 1, i) * gz[kk + 1].y);
        if (k <= ke - 1) LWc += lE * RHOv(kk + 1, i) * gz[kk + 1].x * (kk == k - 1 ? 1.0 : -1.0);
        if (kk == k - 1) l_r += lE * (W(k, i) - W(k - 1, i)) * gz[k].x;
    }
    // rho(k, i) in the corner densities of the x viscous fluxes Ex(k', j), k' = k-1, k; j = i-1, i
    for (int kp = k - 1; kp <= k; ++kp) {
        if (kp < 0 || kp > ke - 1) continue;
        for (int j = i - 1; j <= i; ++j) {
            const double lE = nu * (LRw(kp, j) - LRw(kp, j + 1)) * rdx;
            l_r += lE * (W(kp, j + 1) - W(kp, j)) * rdx * 0.25;
        }
    }
    // rho_u(k, i) in the R_rho_w x fluxes: fip(k', i) and fim(k', i+1), k' = k-1, k
    for (int kp = k - 1; kp <= k; ++kp) {
        if (kp < 0 || kp > ke - 1) continue;
        {
            const double lF = -LRw(kp, i) * rdx;
            const double rzs = 1.0 / (zd[kp] + zd[kp + 1]);
            const double m = (RU(kp, i) * zd[kp + 1] + RU(kp + 1, i) * zd[kp]) * rzs;
            const double g = upw_val(m > EPS, m < -EPS, W(kp, i - 1), W(kp, i), W(kp, i + 1), W(kp, i + 2), 0.5, 0.5);
            l_u += lF * g * (kp == k ? zd[kp + 1] : zd[kp]) * rzs;
        }
        {
            const double lF = LRw(kp, i + 1) * rdx;
            const double rds = gz[kp].z;
            const double m = (RU(kp, i) * dzd[kp + 1] + RU(kp + 1, i) * dzd[kp]) * rds;
            const double g = upw_val(m > EPS, m < -EPS, W(kp, i - 1), W(kp, i), W(kp, i + 1), W(kp, i + 2), 0.5, 0.5);
            l_u += lF * g * (kp == k ? dzd[kp + 1] : dzd[kp]) * rds;
        }
    }
    // pressure gradient (p_k+1 - p_k) / zz and buoyancy -g (rho_k + rho_k+1) / 2
    if (k <= ke - 1) {
        lp += LRw(k, i) * gz[k].y;
        l_r += -GRAV * 0.5 * LRw(k, i);
    }
    if (k >= 1 && k - 1 <= ke - 1) {
        lp += -LRw(k - 1, i) * gz[k - 1].y;
        l_r += -GRAV * 0.5 * LRw(k - 1, i);
    }

    // ---- R_rho_theta (moist.jl:1284-1332)
    for (int j = i - 2; j <= i + 1; ++j) {
        const double lF = (LRT(k, j + 1) - LRT(k, j)) * rdx;
        const double m = RU(k, j);
        const bool pos = m > EPS, neg = m < -EPS;
        lth += lF * m * upw_coef(pos, neg, i - j, 0.5, 0.5);
        if (j == i) l_u += lF * upw_val(pos, neg, THv(k, i - 1), THv(k, i), THv(k, i + 1), THv(k, i + 2), 0.5, 0.5);
    }
    for (int kk = k - 2; kk <= k + 1; ++kk) {
        if (kk < 0 || kk > ke - 1) continue;
        const double lF = LRT(kk + 1, i) * gz[kk + 1].x - LRT(kk, i) * gz[kk].x;
        const double m = RW(kk, i);
        const bool pos = m > EPS && kk > 0, neg = !pos && m < -EPS && kk < ke - 1;
        const double rden = gz[kk].z;
        lth += lF * m * upw_coef(pos, neg, k - kk, dzd[kk + 1] * rden, dzd[kk] * rden);
        if (kk == k)
            l_w += lF * upw_val(pos, neg, kk > 0 ? THv(kk - 1, i) : 0.0, THv(kk, i), THv(kk + 1, i),
                                kk + 2 <= ke ? THv(kk + 2, i) : 0.0, dzd[kk + 1] * rden, dzd[kk] * rden);
    }

    ls0.f[RHO][c] = s0r;
    ls0.f[RHOU][c] = s0u;
    ls0.f[RT][c] = s0t;
    ls0.f[RHOW][c] = s0w;
    lph.f[0][c] = p0;
    lph.f[1][c] = p1;
    lph.f[2][c] = p2;
    ls.f[RT][c] = st;
    ls.f[RHO][c] = sr + l_r;
    ls.f[RHOU][c] = su + l_u;
    ls.f[RHOW][c] = sw + l_w;
    a.LU[c] = LUc;
    a.LW[c] = k <= ke - 1 ? LWc : 0.0;
    a.Lpres[c] = lp;
    a.Lth[c] = Lth0 + lth;
}

// ---- face velocities and stage diagnostics (moist.jl:1044-1069, 1091-1097, 1184-1189)
template <class RC>
__global__ void __launch_bounds__(AD_X *AD_Z)
    ka_diag(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<double> s, SSet<double> ls,
            AdjW<RC> a) {
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    if (i >= nx || k >= nz) return;
    const double *dzd = zv.dz + D.offz;
    auto G = [&](int kk, int ii) -> int64_t { return D.off2 + (int64_t)kk * nx + wrapn(ii, nx); };
    const int64_t c = G(k, i), cm = G(k, i - 1), cp = G(k, i + 1);
    const double *rho = s.f[RHO], *ru = s.f[RHOU], *rw = s.f[RHOW];
    double l_r = 0.0, l_u = 0.0, l_w = 0.0;
    // U(j) = 2 rho_u(j) / (rho_j + rho_j+1)
    {
        const double rden = 1.0 / (rho[c] + rho[cp]);
        const double Uc = ru[c] * 2.0 * rden;
        l_u += a.LU[c] * 2.0 * rden;
        l_r += -a.LU[c] * Uc * rden;
        const double rdm = 1.0 / (rho[cm] + rho[c]);
        l_r += -a.LU[cm] * (ru[cm] * 2.0 * rdm) * rdm;
    }
    // W(kk) = rho_w(kk) / f2h(rho_kk, rho_kk+1)
    const double4 *gz = zv.geo + D.offz;  // per level: 1/dz, 1/(z_k+1 - z_k), 1/(dz_k + dz_k+1)
    if (k <= ke - 1) {
        const double rds = gz[k].z;
        const double rrf = 1.0 / ((rho[c] * dzd[k + 1] + rho[G(k + 1, i)] * dzd[k]) * rds);
        const double Wc = rw[c] * rrf;
        l_w += a.LW[c] * rrf;
        l_r += -a.LW[c] * Wc * rrf * dzd[k + 1] * rds;
    }
    if (k >= 1) {
        const int kb = k - 1;
        const double rds = gz[kb].z;
        const double rrf = 1.0 / ((rho[G(kb, i)] * dzd[k] + rho[c] * dzd[kb]) * rds);
        const double Wb = rw[G(kb, i)] * rrf;
        l_r += -a.LW[G(kb, i)] * Wb * rrf * dzd[kb] * rds;
    }
    // diagnostics pres, theta, rt2pres
    typedef Dl<5> V;
    const K<double> C;
    V rr = V::var(rho[c], 0), rt = V::var(s.f[RT][c], 1), rqv = V::var(s.f[QV][c], 2), rqc = V::var(s.f[QC][c], 3),
      rqr = V::var(s.f[QR][c], 4);
    const V rinv = 1.0 / rr;
    V qv = rqv * rinv, qc = rqc * rinv, qr = rqr * rinv;
    V qd = V(1.0) - qv - qc - qr;
    V r = C.Rd * qd + C.Rv * qv;
    V cp_ = C.Cpd * qd + C.Cpv * qv + C.Cpl * (qc + qr);
    V cv = cp_ - r;
    V pres = C.P0 * d_pow(r * rt / C.P0, cp_ / cv);
    V theta = rt * rinv;
    V rt2 = cp_ / cv * pres / rt;
    const double lp = a.Lpres[c], lt = a.Lth[c], lr2 = a.Lrt[c];
    double g[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) g[j] = lp * pres.d[j] + lt * theta.d[j] + lr2 * rt2.d[j];
    double o[7];
#pragma unroll
    for (int f = 0; f < 7; ++f) o[f] = ls.f[f][c];  // all read before the first store
    ls.f[RHO][c] = o[RHO] + (l_r + g[0]);
    ls.f[RHOU][c] = o[RHOU] + l_u;
    ls.f[RHOW][c] = o[RHOW] + l_w;
    ls.f[RT][c] = o[RT] + g[1];
    ls.f[QV][c] = o[QV] + g[2];
    ls.f[QC][c] = o[QC] + g[3];
    ls.f[QR][c] = o[QR] + g[4];
}

// ---- Kessler microphysics of one cell with dual numbers (moist.jl:554-671) -------
// In two parts: the saturation adjustment (three iterations), which the
// three Kessler parameters do not enter, and autoconversion, accretion and
// rain evaporation, which they do.  KessSat carries what the second part
// reads of the first.
template <class V> struct KessSat {
    V rho_k, theta_k, qv, qc, qr, qd, qsat, dqv, dqc, dth;
};
template <class V>
__device__ __forceinline__ KessSat<V> kessler_sat(const Dom &D, V rho_k, V rt, V rqv, V rqc, V rqr) {
    const K<double> c;
    const double dt = D.dt;
    V dqv(0.0), dqc(0.0), dth(0.0);
    V rinv = V(1.0) / rho_k;
    V theta_k = rt * rinv;
    V qv = rqv * rinv, qc = rqc * rinv, qr = rqr * rinv;
    V qd = V(1.0) - qv - qc - qr;
    V r = c.Rd * qd + c.Rv * qv;
    V cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
    V cv = cp - r;
    V Pi = d_pow(r * rt / c.P0, r / cv);
    V Tk = theta_k * Pi;
    V pk = c.P0 * d_pow(Pi, cp / r);
    V lhe = c.Lv / (cp * Pi);
    V qv_init = qv;
    V ce(0.0), qsat(0.0);
    for (int it = 0; it < 3; ++it) {
        V T_c = Tk - c.T0;
        V tmp = c.B_SAT + T_c;
        V es = c.ES0 * d_exp(c.A_SAT * T_c / tmp);
        V des = es * c.A_SAT * (tmp - T_c) / (tmp * tmp);
        V tq = pk - (1.0 - c.EPSvap) * es;
        V dqs_des;
        if (val(tq) > c.EPS) {
            qsat = c.EPSvap * es / tq;
            dqs_des = c.EPSvap * pk / (tq * tq);
        } else {
            qsat = V(0.0);
            dqs_des = V(1.0);
        }
        V dqs_dT = dqs_des * des;
        ce = (qv - qsat) / (V(1.0) + (c.Lv / cp) * dqs_dT);
        ce = vmin(ce, qv);
        ce = vmax(ce, -qc);
        qv = qv - ce;
        qc = qc + ce;
        theta_k = theta_k + ce * lhe;
        r = c.Rd * qd + c.Rv * qv;
        cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
        cv = cp - r;
        Pi = d_pow(r * rho_k * theta_k / c.P0, r / cv);
        Tk = theta_k * Pi;
        pk = c.P0 * d_pow(Pi, cp / r);
        lhe = c.Lv / (cp * Pi);
    }
    ce = qv_init - qv;
    V rate = ce * rho_k / dt;
    dqv = dqv - rate;
    dqc = dqc + rate;
    dth = dth + rate * lhe;
    return KessSat<V>{rho_k, theta_k, qv, qc, qr, qd, qsat, dqv, dqc, dth};
}
template <class V>
__device__ __forceinline__ void kessler_rest(const Dom &D, const KessSat<V> &k0, V Ka, V Qc0, V Kacc, V out[4]) {
    const K<double> c;
    const double dt = D.dt;
    const bool do_auto = D.flags & F_AUTO, do_accr = D.flags & F_ACCR, do_revap = D.flags & F_REVAP;
    const V rho_k = k0.rho_k, theta_k = k0.theta_k, qv = k0.qv, qd = k0.qd, qsat = k0.qsat;
    V qc = k0.qc, qr = k0.qr, dqv = k0.dqv, dqc = k0.dqc, dqr(0.0), dth = k0.dth;
    if (do_auto) {
        V ra = Ka * vmax(V(0.0), qc - Qc0);
        ra = vmin(ra, qc / dt);
        qc = qc - ra * dt;
        qr = qr + ra * dt;
        ra = ra * rho_k;
        dqc = dqc - ra;
        dqr = dqr + ra;
    }
    if (do_accr) {
        V rc(0.0);
        if (val(qr) > 1e-12 && val(qc) > 1e-12) {
            rc = Kacc * qc * qr;
            rc = vmin(rc, qc / dt);
        }
        qc = qc - rc * dt;
        qr = qr + rc * dt;
        rc = rc * rho_k;
        dqc = dqc - rc;
        dqr = dqr + rc;
    }
    if (do_revap) {
        V re(0.0);
        if (val(qr) > 1e-12 && val(qv) < val(qsat)) {
            re = c.K_evap * (qsat - qv) * d_sqrt(rho_k * qr);
            re = vmin(re, qr) / dt;
        }
        re = re * rho_k;
        dqr = dqr - re;
        dqv = dqv + re;
        const V r = c.Rd * qd + c.Rv * qv;
        const V cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
        const V cv = cp - r;
        const V Pi = d_pow(r * rho_k * theta_k / c.P0, r / cv);
        const V lhe = c.Lv / (cp * Pi);
        dth = dth - re * lhe;
    }
    out[0] = dqv;
    out[1] = dqc;
    out[2] = dqr;
    out[3] = dth;
}
// The saturation part's values as constants of another dual type.
template <class W, class V> __device__ __forceinline__ KessSat<W> kessler_const(const KessSat<V> &k) {
    return KessSat<W>{W(val(k.rho_k)), W(val(k.theta_k)), W(val(k.qv)), W(val(k.qc)), W(val(k.qr)), W(val(k.qd)),
                      W(val(k.qsat)), W(val(k.dqv)), W(val(k.dqc)), W(val(k.dth))};
}

// ---- surface fluxes (moist.jl:710-752): adjoint, one thread per column ---------
// Layout of the per-column buffer sfb (SF_SLOTS x nxs doubles):
//   0..5  adjoints of the surface cell's rho, rho_u, rho_theta, rho_qv, rho_qc, rho_qr
//   6     adjoint of rho_u at the face left of the column (column i-1's)
//   7     adjoint of rho of the cell right of the column (column i+1's)
//   8     the forward d_rho_qv_sf (enters d_rho, and so d_rho_theta's d_rho rho_theta / rho)
// At the domain's ends the left / right values are the stale halo copies
// `stale` (2 per member) the forward read; their adjoints go to carry
// (2 per member), to be added to the previous step's stage-3 input.
constexpr int SF_SLOTS = 9;
template <class V> __device__ __forceinline__ V vabs(const V &x) { return val(x) < 0.0 ? -x : x; }
template <class V>
__device__ __forceinline__ void surface_v(const Dom &D, double dz0, V rho, V ru, V rt, V rqv, V rqc, V rqr, V ru_l,
                                          V rho_r, V out[3]) {
    const K<double> c;
    V qv = rqv / rho, qc = rqc / rho, qr = rqr / rho;
    V qd = V(1.0) - qv - qc - qr;
    V r = c.Rd * qd + c.Rv * qv;
    V cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
    V cv = cp - r;
    V uabs = vabs(ru_l + ru) * 0.5 / rho;
    uabs = vmax(uabs, V(D.uabs_min));
    V theta_ks = rt / rho;
    V pi_ks = d_pow(r * rt / c.P0, r / cv);
    V pi_surf = pi_ks + dz0 * 0.5 * c.GRAV / (cp * theta_ks);
    V p_surf = c.P0 * d_pow(pi_surf, cp / r);
    V rho_surf = p_surf / (r * D.temp_surf);
    V theta_surf = D.temp_surf / pi_surf;
    V u = ru / (rho + rho_r) * 2.0;
    V mflux = -rho_surf * D.Cd * u * vabs(u);
    out[0] = mflux / dz0;
    V sh = rho_surf * D.Ch * uabs * (theta_surf - theta_ks) * pi_surf;
    out[1] = sh / (pi_ks * dz0);
    double es, des;
    sat_vp<double>(D.temp_surf, es, des);
    V tq = p_surf - (1.0 - c.EPSvap) * es;
    V qsat = val(tq) > c.EPS ? c.EPSvap * es / tq : V(0.0);
    V lh = rho_surf * D.Ce * uabs * (qsat - qv);
    out[2] = lh / dz0;
}

__global__ void __launch_bounds__(64)
    ka_surface(const Dom *__restrict__ doms, const int2 *__restrict__ ctiles, ZVec zv, SSet<double> s,
               const double *__restrict__ stale, SSet<double> lph, double *sfb, int64_t nxs, double *carry) {
    const int2 t = ctiles[blockIdx.x];
    if (t.x < 0) return;
    const int m = t.x;
    const Dom &D = doms[m];
    const int i = t.y + threadIdx.x, nx = D.nx;
    if (i >= nx || !(D.flags & F_SF)) return;
    typedef Dl<8> V;
    const int64_t c = D.off2 + i;
    const double ru_l = i == 0 ? stale[2 * m + 0] : s.f[RHOU][c - 1];
    const double rho_r = i == nx - 1 ? stale[2 * m + 1] : s.f[RHO][c + 1];
    V out[3];
    surface_v<V>(D, zv.dz[D.offz], V::var(s.f[RHO][c], 0), V::var(s.f[RHOU][c], 1), V::var(s.f[RT][c], 2),
                 V::var(s.f[QV][c], 3), V::var(s.f[QC][c], 4), V::var(s.f[QR][c], 5), V::var(ru_l, 6),
                 V::var(rho_r, 7), out);
    // d_rho_u_phys, d_rho_theta_phys and d_qv_phys at the ground take the three
    // fluxes; d_rho takes d_qv_sf, which d_rho_theta sees through d_rho rho_theta / rho
    const double lqv = lph.f[0][c] + lph.f[2][c] * s.f[RT][c] / s.f[RHO][c] + lph.f[3][c];
    const double lo[3] = {lph.f[1][c], lph.f[2][c], lqv};
    const int64_t x = D.offx + i;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        double g = 0.0;
#pragma unroll
        for (int o = 0; o < 3; ++o) g += lo[o] * out[o].d[j];
        sfb[j * nxs + x] = g;
    }
    sfb[8 * nxs + x] = out[2].v;
    if (i == 0) carry[2 * m + 0] = sfb[6 * nxs + x];
    if (i == nx - 1) carry[2 * m + 1] = sfb[7 * nxs + x];
}

// The adjoints of the two stale halo reads of the step reversed just before
// (carry) enter this step's stage-3 input adjoint l at (0, nx-1) for rho_u
// and (0, 0) for rho.  One thread per member.
__global__ void ka_stale_add(const Dom *__restrict__ doms, int nm, SSet<double> l, const double *carry) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const Dom &D = doms[m];
    if (!(D.flags & F_SF)) return;
    l.f[RHOU][D.off2 + D.nx - 1] += carry[2 * m + 0];
    l.f[RHO][D.off2] += carry[2 * m + 1];
}

// ---- gray radiation (moist.jl:776-840): adjoint, one thread per column ----------
// g = adjoint of d_rho_theta_rd (= that of d_rho_theta_phys); adds into l
// (rho, rho_theta, rho_qv, rho_qc, rho_qr).  The upward long-wave recursion is
// reversed from the top, the downward one from the ground; scratch (5 x n2):
// the forward fluxes up at face k+1 and dn at face k, and the partial
// adjoints of tau, eps and B = sigma T^4 of the first reversed recursion.
__device__ __forceinline__ void rad_cell(const Dom &D, const ZVec &zv, const SSet<double> &s, int k, int64_t q,
                                         double &eps, double &ex, double &B, Dl<5> &Td, Dl<5> &CP) {
    const K<double> c;
    typedef Dl<5> V;
    V rho = V::var(s.f[RHO][q], 0), rt = V::var(s.f[RT][q], 1), rqv = V::var(s.f[QV][q], 2),
      rqc = V::var(s.f[QC][q], 3), rqr = V::var(s.f[QR][q], 4);
    V qv = rqv / rho, qc = rqc / rho, qr = rqr / rho;
    V qd = V(1.0) - qv - qc - qr;
    V r = c.Rd * qd + c.Rv * qv;
    V cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
    V cv = cp - r;
    V pex = d_pow(r * rt / c.P0, r / cv);
    Td = rt / rho * pex;
    CP = cp * pex;
    const double dzk = zv.dz[D.offz + k];
    const double dopt = D.k_abs_qv * s.f[QV][q] * dzk + D.k_abs_qc * s.f[QC][q] * dzk + D.k_abs_qr * s.f[QR][q] * dzk;
    ex = exp(-dopt);
    eps = 1.0 - ex;
    const double tv = Td.v;
    B = c.SIGMA * (tv * tv * tv * tv);
}

__global__ void __launch_bounds__(64)
    ka_radiation(const Dom *__restrict__ doms, const int2 *__restrict__ ctiles, ZVec zv, SSet<double> s,
                 const double *__restrict__ theta_ref, const double *__restrict__ g, SSet<double> l, double *sc0,
                 double *sc1, double *sc2, double *sc3, double *sc4) {
    const int2 t = ctiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x, nx = D.nx, nz = D.nz;
    if (i >= nx || !(D.flags & F_RD)) return;
    const K<double> c;
    const double *dzd = zv.dz + D.offz, *zd = zv.z + D.offz;
    auto G = [&](int kk) -> int64_t { return D.off2 + (int64_t)kk * nx + i; };
    double *UP = sc0, *DN = sc1, *LTAU = sc2, *LEPS = sc3, *LB = sc4;
    const double ts = D.temp_surf;
    const double up0 = D.surf_emissivity * c.SIGMA * (ts * ts * ts * ts);
    // forward: up at faces 1..nz (stored at levels 0..nz-1), dn at faces 0..nz-1
    double up = up0;
    for (int k = 0; k < nz; ++k) {
        double eps, ex, B;
        Dl<5> Td, CP;
        rad_cell(D, zv, s, k, G(k), eps, ex, B, Td, CP);
        up = up * (1.0 - eps) + eps * B;
        UP[G(k)] = up;
    }
    double dn = 0.0;
    for (int k = nz - 1; k >= 0; --k) {
        double eps, ex, B;
        Dl<5> Td, CP;
        rad_cell(D, zv, s, k, G(k), eps, ex, B, Td, CP);
        dn = dn * (1.0 - eps) + eps * B;
        DN[G(k)] = dn;
    }
    auto upf = [&](int j) { return j == 0 ? up0 : UP[G(j - 1)]; };   // up at face j
    auto dnf = [&](int j) { return j == nz ? 0.0 : DN[G(j)]; };     // dn at face j
    // direct adjoint of the net flux F = up - dn at face j: from the heating of
    // the levels below (j-1) and above (j); heating hr_k = -divF_k / (cp Pi)_k
    auto ldiv = [&](int k) {
        double eps, ex, B;
        Dl<5> Td, CP;
        rad_cell(D, zv, s, k, G(k), eps, ex, B, Td, CP);
        return -g[G(k)] / CP.v / dzd[k];
    };
    auto lF = [&](int j) { return (j >= 1 ? ldiv(j - 1) : 0.0) - (j <= nz - 1 ? ldiv(j) : 0.0); };
    // up[k+1] = up[k] tau_k + eps_k B_k, reversed from the top
    double Lup = lF(nz);  // total adjoint of up at face nz
    for (int k = nz - 1; k >= 0; --k) {
        double eps, ex, B;
        Dl<5> Td, CP;
        rad_cell(D, zv, s, k, G(k), eps, ex, B, Td, CP);
        LTAU[G(k)] = Lup * upf(k);
        LEPS[G(k)] = Lup * B;
        LB[G(k)] = Lup * eps;
        Lup = lF(k) + Lup * (1.0 - eps);
    }
    // dn[k] = dn[k+1] tau_k + eps_k B_k, reversed from the ground; then the
    // cell's adjoints: tau = 1 - eps, eps = 1 - exp(-dopt), B = sigma T^4, cp Pi,
    // and the Newtonian relaxation -(rho_theta - rho theta_ref) alpha
    double Ldn = -lF(0);
    const double dz_trans = 2.0e3;
    for (int k = 0; k < nz; ++k) {
        const int64_t q = G(k);
        double eps, ex, B;
        Dl<5> Td, CP;
        rad_cell(D, zv, s, k, q, eps, ex, B, Td, CP);
        const double ltau = LTAU[q] + Ldn * dnf(k + 1);
        const double leps = LEPS[q] + Ldn * B;
        const double lB = LB[q] + Ldn * eps;
        Ldn = -lF(k + 1) + Ldn * (1.0 - eps);
        const double divF = ((upf(k + 1) - dnf(k + 1)) - (upf(k) - dnf(k))) / dzd[k];
        const double gk = g[q];
        const double lcpp = gk * divF / (CP.v * CP.v);
        const double ldopt = (leps - ltau) * ex;
        const double tv = Td.v;
        const double lT = lB * c.SIGMA * 4.0 * tv * tv * tv;
        const double alpha =
            0.5 * (1.0 + tanh((zd[k] - (D.z_tropopose + dz_trans * 0.5)) / dz_trans)) / D.tau_cooling;
        const double dzk = dzd[k];
        const double a0 = l.f[RHO][q], a1 = l.f[RT][q], a2 = l.f[QV][q], a3 = l.f[QC][q], a4 = l.f[QR][q];
        l.f[RHO][q] = a0 + (lT * Td.d[0] + lcpp * CP.d[0]) + alpha * theta_ref[D.offz + k] * gk;
        l.f[RT][q] = a1 + (lT * Td.d[1] + lcpp * CP.d[1]) - alpha * gk;
        l.f[QV][q] = a2 + (lT * Td.d[2] + lcpp * CP.d[2]) + ldopt * D.k_abs_qv * dzk;
        l.f[QC][q] = a3 + (lT * Td.d[3] + lcpp * CP.d[3]) + ldopt * D.k_abs_qc * dzk;
        l.f[QR][q] = a4 + (lT * Td.d[4] + lcpp * CP.d[4]) + ldopt * D.k_abs_qr * dzk;
    }
}

// ---- physics adjoint (moist.jl:554-707, 1632-1644) ----------------------------------
// lph: adjoint of the six tendencies; ldprec: adjoint of d_prec (dx * dt).
// lout = ls + ls0 + the physics adjoints (lout may be another buffer than ls:
// each cell reads and writes only itself); accumulates dJ/dK per cell into gk
// (3 fields).  sfb (surface fluxes on): ka_surface's per-column adjoints,
// gathered at the ground (own cell, the right column's left face, the left
// column's right cell) and its forward d_qv_sf; null otherwise.  The gray
// radiation's adjoint is added afterwards by ka_radiation.
// At least 2 waves per SIMD (at most 256 registers; the dual-number Kessler
// pass wants 262, and at one wave per SIMD nothing hides its latencies): a few
// values spill, and the 64-member 1024x128 reverse step ran 4% faster.
__global__ void __launch_bounds__(AD_X *AD_Z, 2)
    ka_physics(const Dom *__restrict__ doms, const int4 *__restrict__ tiles, ZVec zv, SSet<double> s,
               SSet<double> lph, SSet<double> ls, SSet<double> ls0, SSet<double> lout, double *gk0, double *gk1,
               double *gk2, const double *__restrict__ sfb, int64_t nxs) {
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    if (i >= nx || k >= nz) return;
    const double *dzd = zv.dz + D.offz;
    const K<double> C;
    auto G = [&](int kk) -> int64_t { return D.off2 + (int64_t)kk * nx + i; };
    const int64_t c = G(k);
    const double *rho = s.f[RHO], *rt = s.f[RT], *rqr = s.f[QR];
    const bool mp = D.flags & F_MP, sed = mp && (D.flags & F_SED);
    auto vt = [&](int kk) {
        const int64_t q = G(kk);
        const double eff = jmax(0.0, rqr[q]);
        const double v = C.A_term * pow(eff / rho[q], C.B_term);
        return eff > 1e-12 ? v : 0.0;
    };
    auto flux = [&](int kk) { return rqr[G(kk)] * vt(kk); };
    double sedc = 0.0;
    if (sed) sedc = ((k < ke ? flux(k + 1) : 0.0) - flux(k)) / dzd[k];
    const bool sf = sfb && (D.flags & F_SF) && k == 0;
    const int64_t x = D.offx + i;
    // d_rho = d_qr_sed (+ d_qv_sf at the ground)
    const double drho = sf ? sedc + sfb[8 * nxs + x] : sedc;
    // ls += ls0 (the step input is also every stage's s0) plus the physics
    // adjoints below; all seven read first, written at the end
    double o[7];
#pragma unroll
    for (int f = 0; f < 7; ++f) o[f] = ls.f[f][c] + ls0.f[f][c];
    double l_r = 0.0, l_rt = 0.0, l_qr = 0.0;
    // d_rho_theta = d_th_mp (+ d_th_rd) + d_rho rho_theta / rho (+ d_th_sf)
    l_rt += lph.f[2][c] * drho / rho[c];
    l_r += -lph.f[2][c] * drho * rt[c] / (rho[c] * rho[c]);
    if (sf) {
        // the surface fluxes' adjoints at this column, and those of the
        // neighbours that read this cell (periodic: the ends read stale copies)
        o[RHO] += sfb[x];
        o[RHOU] += sfb[nxs + x];
        o[RT] += sfb[2 * nxs + x];
        o[QV] += sfb[3 * nxs + x];
        o[QC] += sfb[4 * nxs + x];
        o[QR] += sfb[5 * nxs + x];
        if (i < nx - 1) o[RHOU] += sfb[6 * nxs + x + 1];
        if (i > 0) o[RHO] += sfb[7 * nxs + x - 1];
    }
    if (mp) {
        // partials w.r.t. the five state fields, then w.r.t. the three parameters
        // (two narrower dual evaluations keep the registers in bounds)
        const double lo[4] = {lph.f[3][c], lph.f[4][c], lph.f[5][c], lph.f[2][c]};
        // (the saturation adjustment, which the parameters do not enter, runs
        // once: the parameter pass starts from its values)
        KessSat<Dl<3>> satK;
        {
            typedef Dl<5> V;
            V out[4];
            const KessSat<V> sat = kessler_sat<V>(D, V::var(rho[c], 0), V::var(rt[c], 1), V::var(s.f[QV][c], 2),
                                                  V::var(s.f[QC][c], 3), V::var(rqr[c], 4));
            satK = kessler_const<Dl<3>>(sat);
            kessler_rest<V>(D, sat, V(D.K_auto), V(D.Qc_threshold), V(D.K_accr), out);
            double g[5];
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                g[j] = 0.0;
#pragma unroll
                for (int o = 0; o < 4; ++o) g[j] += lo[o] * out[o].d[j];
            }
            l_r += g[0];
            l_rt += g[1];
            o[QV] += g[2];
            o[QC] += g[3];
            l_qr += g[4];
        }
        {
            typedef Dl<3> V;
            V out[4];
            kessler_rest<V>(D, satK, V::var(D.K_auto, 0), V::var(D.Qc_threshold, 1), V::var(D.K_accr, 2), out);
            double g[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                g[j] = 0.0;
#pragma unroll
                for (int o = 0; o < 4; ++o) g[j] += lo[o] * out[o].d[j];
            }
            gk0[c] += g[0];
            gk1[c] += g[1];
            gk2[c] += g[2];
        }
    }
    if (sed) {
        // F(k) = rho_qr Vt; d_sed(k) = (F(k+1) [k<ke] - F(k)) / dz_k; d_prec = F(0)
        auto lsed = [&](int kk) {
            const int64_t q = G(kk);
            return lph.f[0][q] + lph.f[5][q] + lph.f[2][q] * rt[q] / rho[q];
        };
        double lF = -lsed(k) / dzd[k];
        if (k > 0) lF += lsed(k - 1) / dzd[k - 1];
        if (k == 0) lF += D.dx * D.dt;  // the rain term of the loss: dJ/dd_prec = dx dt
        const double eff = jmax(0.0, rqr[c]);
        if (eff > 1e-12) {
            const double v = vt(k);
            l_qr += lF * v * (1.0 + C.B_term);
            l_r += -lF * C.B_term * rqr[c] * v / rho[c];
        }
    }
    o[RHO] += l_r;
    o[RT] += l_rt;
    o[QR] += l_qr;
#pragma unroll
    for (int f = 0; f < 7; ++f) lout.f[f][c] = o[f];
}

// Per-member sums of the parameter gradients (one block per member).
__global__ void __launch_bounds__(256) ka_reduce_k(const Dom *__restrict__ doms, const double *gk0, const double *gk1,
                                                   const double *gk2, double *out) {
    const Dom &D = doms[blockIdx.x];
    const int64_t n = (int64_t)D.nx * D.nz;
    __shared__ double r[3][256];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        a0 += gk0[D.off2 + e];
        a1 += gk1[D.off2 + e];
        a2 += gk2[D.off2 + e];
    }
    r[0][threadIdx.x] = a0;
    r[1][threadIdx.x] = a1;
    r[2][threadIdx.x] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < 3; ++j) r[j][threadIdx.x] += r[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 3; ++j) out[3 * blockIdx.x + j] = r[j][0];
}
