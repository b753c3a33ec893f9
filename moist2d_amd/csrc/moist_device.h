// moist_device.h — device-side types and point functions of the gfx950
// time-stepping core.  Everything here is __device__ inline and shared by
// the kernels in moist_kernels.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mo {

// ---- physical constants in the compute type (moist.jl:30-59) --------------
template <class R> struct K {
    static constexpr R GRAV = R(9.81);
    static constexpr R Rd = R(287.04);
    static constexpr R Rv = R(461.5);
    static constexpr R Cpd = R(1004.64);
    static constexpr R Cpv = R(1846.0);
    static constexpr R Cpl = R(4218.0);
    static constexpr R Cvd = Cpd - Rd;
    static constexpr R P0 = R(1e5);
    static constexpr R Lv = R(2.501e6);
    static constexpr R SIGMA = R(5.67e-8);
    static constexpr R EPSvap = Rd / Rv;
    static constexpr R K_evap = R(1e-2);
    static constexpr R A_term = R(36.3);
    static constexpr R B_term = R(0.1346);
    static constexpr R T0 = R(273.15);
    static constexpr R ES0 = R(610.78);
    static constexpr R A_SAT = R(17.67);
    static constexpr R B_SAT = R(243.5);
    static constexpr R EPS = R(1e-12);
};

// ---- per-domain descriptor (one per ensemble member) -----------------------
enum : int {
    F_MP = 1, F_AUTO = 2, F_ACCR = 4, F_REVAP = 8, F_SED = 16, F_RD = 32, F_SF = 64
};

struct Dom {
    int nx, nz, halo, ns;
    int flags;
    int lin_col;        // interior column the positivity limiters read (halo column 1)
    int64_t off2;       // first cell of this domain in every 2-D pool
    int64_t offx;       // first column in the 1-D x pools
    int64_t offz;       // first level in the 1-D z pools
    double dx, dt, H, nu, kappa, gamma_d, beta, zf_sponge;
    double Cd, Ch, Ce, uabs_min, temp_surf;
    double k_abs_qv, k_abs_qc, k_abs_qr, surf_emissivity, tau_cooling, z_tropopose;
    double K_auto, Qc_threshold, K_accr;
};

// 1-D z vectors of all domains, concatenated (offset Dom::offz).
struct ZVec {
    const double *z, *zf, *dz;
    // per-level geometry of the w-p system: {1/dz_k, 1/(z_k+1 - z_k),
    // 1/(dz_k + dz_k+1), g/(dz_k + dz_k+1)} (zeros on the top level)
    const double4 *geo;
    // base state of the mixed form (zero/unused in the FP64 form)
    const double *rho_bar, *rt_bar, *p_bar, *L_bar, *B_bar;
};

// ---- Julia Base.min/max for IEEE floats -------------------------------------
template <class R> __device__ __forceinline__ R jmax(R x, R y) {
    R diff = x - y;
    R am = signbit(diff) ? y : x;
    return (isnan(x) || isnan(y)) ? diff : am;
}
template <class R> __device__ __forceinline__ R jmin(R x, R y) {
    R diff = x - y;
    R am = signbit(diff) ? x : y;
    return (isnan(x) || isnan(y)) ? diff : am;
}
// In float, the IEEE 754-2019 minimum and maximum (gfx950 v_minimum3_f32,
// v_maximum3_f32: one instruction) have exactly these semantics: a NaN
// operand gives NaN, -0 orders below +0, and otherwise the smaller (larger)
// value; only the NaN's payload can differ.
__device__ __forceinline__ float jmax(float x, float y) { return __builtin_elementwise_maximum(x, y); }
__device__ __forceinline__ float jmin(float x, float y) { return __builtin_elementwise_minimum(x, y); }

// ---- math in the compute type -----------------------------------------------
// Division: IEEE division in double (the FP64 form performs the reference's
// operations); in float (mixed form) a times the hardware reciprocal
// v_rcp_f32 (about 1 ulp).
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// dz-weighted face-to-centre interpolation (f2h, moist.jl:243-251)
template <class R> __device__ __forceinline__ R f2h(R v1, R v2, R d1, R d2) {
    return fdiv(v1 * d2 + v2 * d1, d1 + d2);
}

// Division by a divisor used several times: rb = rcp_once(b) once, then
// fdiv_r(a, b, rb).  The float form's division is a times the hardware
// reciprocal, so this gives the same values as fdiv with one reciprocal; in
// double it is the IEEE division (rb is not used).
__device__ __forceinline__ float rcp_once(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rcp_once(double) { return 0.0; }
__device__ __forceinline__ float fdiv_r(float a, float, float rb) { return a * rb; }
__device__ __forceinline__ double fdiv_r(double a, double b, double) { return a / b; }
// f2h with rs = rcp_once(d1 + d2) (the same for d2 + d1)
template <class R> __device__ __forceinline__ R f2h_r(R v1, R v2, R d1, R d2, R rs) {
    return fdiv_r(v1 * d2 + v2 * d1, d1 + d2, rs);
}

// Koren (1993) limiter (moist.jl:844-847)
template <class R> __device__ __forceinline__ R koren(R r) {
    return jmax(R(0.0), jmin(R(2.0) * r, jmin(fdiv(R(1.0) + R(2.0) * r, R(3.0)), R(2.0))));
}

// (-m + 5c + 2p), the numerator of the 3rd-order upwind face value
template <class R> __device__ __forceinline__ R upw3(R m, R c, R p) {
    return -m + R(5.0) * c + R(2.0) * p;
}

// Periodic wrap of i + o for |o| <= 2 (exchange_halo!, moist.jl:233-241).
__device__ __forceinline__ int wrapx(int j, int nx) {
    if (j < 0) j += nx;
    if (j < 0) j += nx;
    if (j >= nx) j -= nx;
    if (j >= nx) j -= nx;
    return j;
}

__device__ __forceinline__ double m_pow(double a, double b) { return pow(a, b); }
// float pow/exp of the mixed form from the hardware base-2 log/exp
// (v_log_f32, v_exp_f32); arguments of pow here are positive or zero.
__device__ __forceinline__ float m_pow(float a, float b) {
    return __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(a));
}
__device__ __forceinline__ double m_exp(double a) { return exp(a); }
__device__ __forceinline__ float m_exp(float a) { return __builtin_amdgcn_exp2f(a * 1.4426950408889634f); }
__device__ __forceinline__ double m_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float m_sqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double m_tanh(double a) { return tanh(a); }
__device__ __forceinline__ float m_tanh(float a) { return tanhf(a); }
__device__ __forceinline__ double m_fabs(double a) { return fabs(a); }
__device__ __forceinline__ float m_fabs(float a) { return fabsf(a); }
__device__ __forceinline__ double m_log1p(double a) { return log1p(a); }
__device__ __forceinline__ float m_log1p(float a) { return log1pf(a); }
// log1p(a) + log1p(b) = log1p(a + b + ab).  In float from the hardware log2
// (v_log_f32) by Goldberg's form log(u) x / (u - 1), u = 1 + x, whose rounding
// of 1 + x cancels (Goldberg 1991, "What every computer scientist should know
// about floating-point arithmetic", Theorem 4): a few ulp, where the library
// log1pf takes some 120 instructions.
__device__ __forceinline__ double m_log1p_sum(double a, double b) { return log1p(a) + log1p(b); }
__device__ __forceinline__ float m_log1p_sum(float a, float b) {
    const float x = a + b + a * b;
    const float u = 1.0f + x, d = u - 1.0f;
    const float l = __builtin_amdgcn_logf(u) * 0.69314718055994531f;
    return d == 0.0f ? x : l * (x * __builtin_amdgcn_rcpf(d));
}
__device__ __forceinline__ double m_expm1(double a) { return expm1(a); }
__device__ __forceinline__ float m_expm1(float a) { return expm1f(a); }

// Tetens saturation vapour pressure and its T derivative (moist.jl:254-260)
template <class R> __device__ __forceinline__ void sat_vp(R Tk, R &es, R &des) {
    R Tc = Tk - K<R>::T0;
    R tmp = K<R>::B_SAT + Tc;
    es = K<R>::ES0 * m_exp(fdiv(K<R>::A_SAT * Tc, tmp));
    des = fdiv(es * K<R>::A_SAT * (tmp - Tc), tmp * tmp);
}

// Saturation specific humidity and d qsat / d es (moist.jl:263-273)
template <class R> __device__ __forceinline__ void sat_q(R p, R es, R &qs, R &dqs) {
    R tmp = p - (R(1.0) - K<R>::EPSvap) * es;
    if (tmp > K<R>::EPS) {
        qs = fdiv(K<R>::EPSvap * es, tmp);
        dqs = fdiv(K<R>::EPSvap * p, tmp * tmp);
    } else {
        qs = R(0.0);
        dqs = R(1.0);
    }
}

// NaN report: the smallest (step, stage, field) key wins, like the
// reference's first error() (moist.jl:1543-1578).
__device__ __forceinline__ void report_nan(int *nan_key, int key) { atomicMin(nan_key, key); }

}  // namespace mo
