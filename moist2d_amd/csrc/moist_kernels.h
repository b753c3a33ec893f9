// moist_kernels.h — gfx950 kernels of one large step (rk3_step!, moist.jl:1589-1683).
//
// Device layout: every 2-D field is a pool holding all ensemble members;
// member m's cell (k, i) is at Dom::off2 + k*nx + i (x fastest, no halo
// columns; periodic x is an index wrap).  A block takes one entry of a tile
// work-list: 2-D tiles are 64 columns x 4 levels (blockDim 64x4, one wave
// per level row, coalesced along x); column tiles are 64 columns (one thread
// per column, sweeping k).
//
// S is the storage type.  S = double is the FP64 form, which restates the
// reference's arithmetic operation by operation.  S = float is the mixed
// form: rho and rho*theta are stored as deviations from an FP64 1-D base
// state (rho_bar, rt_bar), pressure is carried as its deviation from
// p_bar = P0 (Rd rt_bar / P0)^(Cpd/Cvd) computed through log1p/expm1 so that
// horizontal and vertical pressure differences keep full FP32 precision,
// and the hydrostatic residual of the base state (B_bar) is added in FP64.
// The w-p tridiagonal system is assembled and eliminated in FP64 in both forms.
#pragma once

#include "moist_device.h"

namespace mo {

template <class S> struct SSet {  // the 7 prognostic fields (moist.jl:295-301)
    S *f[7];
};
enum { RHO = 0, RHOU = 1, RHOW = 2, RT = 3, QV = 4, QC = 5, QR = 6 };

template <class S> struct Work {
    S *theta, *pres, *rt2, *uf, *wf;      // stage diagnostics (moist.jl:1044-1097,1184-1189)
    S *R_rho, *R_rhou, *R_rhow, *R_rt;    // slow tendencies (moist.jl:1071-1332)
    S *rp, *rup, *rwp;                    // acoustic perturbations (moist.jl:1339-1342)
    S *mom_x, *mom_z;                     // time-averaged mass fluxes (moist.jl:1347-1348)
    S *d_rho, *d_rt, *d_qv, *d_qc, *d_qr; // physics tendencies, frozen over a step (moist.jl:1632-1644)
    S *d_u2d;                             // 2-D d_rho_u_phys (stage-level entry only)
    S *d_u_sf, *d_th_sf, *d_qv_sf, *mflux, *shflux, *lhflux;  // 1-D x (moist.jl:710-752)
    S *dg_th_mp, *dg_qv_mp, *dg_qr_mp, *dg_qr_sed, *d_th_rd;  // State diagnostics
    double *prec;                         // accumulated surface rain (1-D x)
    double *sfh;                          // stale halo reads of surface_fluxes!, 2 per member
    const double *theta_ref;              // 1-D z (gray radiation)
    int *nan_key;
    const int *step_ctr;
};

// NaN reports rank by (step of the call, stage, field): key ((step * 3 + stage)
// << 4) | field, atomically minimized.  The step count is advanced by the
// first stage's k_slow, so a step's physics (which runs before it, and ranks
// with that step's first stage) reads step - 1 and reports at 3 (step - 1) + 3.
// The count saturates below the largest step whose key fits an int: beyond
// 44.7 M steps in one call, later steps share its rank.
constexpr int STEP_CTR_MAX = (INT_MAX >> 4) / 3 - 1;

template <class S> constexpr bool is_mixed() { return sizeof(S) == 4; }

// Layout of the fields the acoustic substeps stream (theta, rt2pres, the four
// slow tendencies, rho', rho_u', rho_w', rho_theta', mom_x, mom_z).  P = true
// is the level-pair layout: levels 2m and 2m+1 of a column are adjacent, cell
// (k, i) of a member at off2 + ((k >> 1) nx + i) 2 + (k & 1).  A lane that
// holds two consecutive levels moves them with one 8-byte (FP32) or 16-byte
// (FP64) access, and 16 consecutive columns of a level pair fill a 128-byte
// line, where an x-fastest row of 16 FP32 columns fills half of one.  The
// mixed form's forward path uses it; P = false is the x-fastest layout of
// every other pool, of the FP64 form (whose 16-column rows already fill a
// line) and of the adjoint's recording pass.  Member offsets are even and
// an odd nz is padded by one level, so the pair of the top level stays
// inside its member.
template <bool P> __device__ __forceinline__ int64_t widx(const Dom &D, int k, int i) {
    if constexpr (P) return D.off2 + ((int64_t)(k >> 1) * D.nx + i) * 2 + (k & 1);
    else return D.off2 + (int64_t)k * D.nx + i;
}
template <class S> struct Vec2;
template <> struct Vec2<float> { using T = float2; };
template <> struct Vec2<double> { using T = double2; };
// levels k and k+1 (k even) of column i; x-fastest reads level k+1 only when
// it exists.  In FP32 the pair is read as one relaxed wavefront-scope atomic
// 8-byte load: the compiler splits a plain float2 load whose halves are used
// apart into two 4-byte loads, and an atomic load is never split.
template <bool P, class S> __device__ __forceinline__ void ld2(const S *p, const Dom &D, int k, int i, S &a, S &b) {
    if constexpr (P && sizeof(S) == 4) {
        const uint64_t u = __hip_atomic_load((const uint64_t *)(p + widx<true>(D, k, i)), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WAVEFRONT);
        a = __builtin_bit_cast(float, (uint32_t)u);
        b = __builtin_bit_cast(float, (uint32_t)(u >> 32));
    } else if constexpr (P) {
        const typename Vec2<S>::T v = *(const typename Vec2<S>::T *)(p + widx<true>(D, k, i));
        a = v.x;
        b = v.y;
    } else {
        a = p[widx<false>(D, k, i)];
        b = p[widx<false>(D, min(k + 1, D.nz - 1), i)];
    }
}
// stores levels k and k+1 (k even); x-fastest stores level k+1 only when it exists
template <bool P, class S> __device__ __forceinline__ void st2(S *p, const Dom &D, int k, int i, S a, S b) {
    if constexpr (P) {
        typename Vec2<S>::T v;
        v.x = a;
        v.y = b;
        *(typename Vec2<S>::T *)(p + widx<true>(D, k, i)) = v;
    } else {
        p[widx<false>(D, k, i)] = a;
        if (k + 1 < D.nz) p[widx<false>(D, k + 1, i)] = b;
    }
}

// ---- member-relative addressing ----------------------------------------------
// A block works on one member, so it can address a field through the member's
// base pointer (uniform: one scalar register pair) and a 32-bit byte offset
// per lane, the form gfx950 loads and stores take directly (global_load v,
// v_off, s[base]): no 64-bit address arithmetic per access, and one offset
// register serves every field at the same cell.  moist_create keeps each
// member below 2^32 bytes a field.
// element offset of cell (k, i) within its member, layout P
template <bool P> __device__ __forceinline__ uint32_t moff(int nx, int k, int i) {
    if constexpr (P) return ((uint32_t)(k >> 1) * (uint32_t)nx + (uint32_t)i) * 2u + (uint32_t)(k & 1);
    else return (uint32_t)k * (uint32_t)nx + (uint32_t)i;
}
template <class S> __device__ __forceinline__ S ldb(const S *b, uint32_t boff) {
    return *(const S *)((const char *)b + boff);
}
template <class S> __device__ __forceinline__ void stb(S *b, uint32_t boff, S v) { *(S *)((char *)b + boff) = v; }
// a level pair (8 bytes; the same relaxed atomic load as ld2, never split)
__device__ __forceinline__ void ldb2(const float *b, uint32_t boff, float &x, float &y) {
    const uint64_t u = __hip_atomic_load((const uint64_t *)((const char *)b + boff), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WAVEFRONT);
    x = __builtin_bit_cast(float, (uint32_t)u);
    y = __builtin_bit_cast(float, (uint32_t)(u >> 32));
}
__device__ __forceinline__ void stb2(float *b, uint32_t boff, float x, float y) {
    float2 v;
    v.x = x;
    v.y = y;
    *(float2 *)((char *)b + boff) = v;
}

// Stage time step: dt/3, dt/2, dt in the compute type (moist.jl:1649-1657),
// or an explicit dt for the stage-level entry.
template <class R> __device__ __forceinline__ R stage_dt(const Dom &D, int div, double dt_ovr) {
    if (dt_ovr > 0.0) return R(dt_ovr);
    return R(D.dt) / R(div);
}

template <class S>
__device__ __forceinline__ S full_rho(const SSet<S> &s, int64_t c, const ZVec &zv, const Dom &D, int k) {
    if constexpr (is_mixed<S>()) return s.f[RHO][c] + S(zv.rho_bar[D.offz + k]);
    else return s.f[RHO][c];
}
template <class S>
__device__ __forceinline__ S full_rt(const SSet<S> &s, int64_t c, const ZVec &zv, const Dom &D, int k) {
    if constexpr (is_mixed<S>()) return s.f[RT][c] + S(zv.rt_bar[D.offz + k]);
    else return s.f[RT][c];
}

// =============================================================================
// Stage diagnostics: theta, pressure (or its deviation), rt2pres
// (moist.jl:1048-1062, 1091-1097), and in the x-fastest variant also u and w
// at faces (moist.jl:1091-1097, 1184-1189); the pair-layout variant (the
// mixed form) leaves those to k_slow, which forms them from the tile it stages.
// =============================================================================
// One cell from its stored state values (rho and rho_theta are deviations
// from the base state in the mixed form).  k_diag, k_physics (the mixed
// form's first stage) and the epilogue of k_scalar (the next stage's
// diagnostics) share it.
template <class S>
__device__ __forceinline__ void diag_point(const Dom &D, const ZVec &zv, int k, S rho_s, S rt_s, S rqv, S rqc,
                                           S rqr, S &theta, S &pres, S &rt2) {
    using R = S;
    using KC = K<R>;
    R rho = rho_s, rt = rt_s;
    if constexpr (is_mixed<S>()) {
        rho = rho_s + S(zv.rho_bar[D.offz + k]);
        rt = rt_s + S(zv.rt_bar[D.offz + k]);
    }
    const R qv = fdiv(rqv, rho), qc = fdiv(rqc, rho), qr = fdiv(rqr, rho);
    const R qd = R(1.0) - qv - qc - qr;
    const R r = KC::Rd * qd + KC::Rv * qv;
    const R cp = KC::Cpd * qd + KC::Cpv * qv + KC::Cpl * (qc + qr);
    const R cv = cp - r;
    R pfull;
    if constexpr (is_mixed<S>()) {
        // p = P0 exp(g ln(r rt/P0)), p_bar = P0 exp(g_bar ln(Rd rt_bar/P0));
        // ln(r rt / (Rd rt_bar)) = log1p(dr) + log1p(rt'/rt_bar),
        // g - g_bar = (Cpd dR - Rd dCp) / (cv Cvd).
        const R dR = (KC::Rv - KC::Rd) * qv - KC::Rd * (qc + qr);
        const R dCp = (KC::Cpv - KC::Cpd) * qv + (KC::Cpl - KC::Cpd) * (qc + qr);
        const R dg = fdiv(KC::Cpd * dR - KC::Rd * dCp, cv * KC::Cvd);
        const R gam = fdiv(cp, cv);
        const R rtb = R(zv.rt_bar[D.offz + k]);
        const R E = dg * R(zv.L_bar[D.offz + k]) + gam * m_log1p_sum(fdiv(dR, KC::Rd), fdiv(rt_s, rtb));
        const double pb = zv.p_bar[D.offz + k];
        pres = R(pb) * m_expm1(E);
        pfull = R(pb) + pres;
    } else {
        pres = KC::P0 * m_pow(r * rt / KC::P0, cp / cv);
        pfull = pres;
    }
    theta = fdiv(rt, rho);
    rt2 = fdiv(fdiv(cp, cv) * pfull, rt);
}

// theta and rt2pres are stored in layout P (level pairs when true), the
// pressure and the face velocities x-fastest.
template <class S, bool P>
__global__ void __launch_bounds__(256) k_diag(const Dom *__restrict__ doms, const int4 *__restrict__ tiles,
                                              ZVec zv, SSet<S> in, Work<S> w) {
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD ordering
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
    const int nx = D.nx, nz = D.nz;
    if (i >= nx || k >= nz) return;
    const int64_t c = D.off2 + (int64_t)k * nx + i;
    S th, pres, rt2;
    diag_point<S>(D, zv, k, in.f[RHO][c], in.f[RT][c], in.f[QV][c], in.f[QC][c], in.f[QR][c], th, pres, rt2);
    S uf = S(0.0), wf = S(0.0);
    if constexpr (!P) {
        // neighbour reads (i+1 and k+1) happen before the first store
        const S rho = full_rho(in, c, zv, D, k);
        const int64_t cip = D.off2 + (int64_t)k * nx + wrapx(i + 1, nx);
        uf = fdiv(in.f[RHOU][c] * S(2.0), rho + full_rho(in, cip, zv, D, k));
        if (k < nz - 1) {
            const double *dz = zv.dz + D.offz;
            wf = fdiv(in.f[RHOW][c], f2h(rho, full_rho(in, c + nx, zv, D, k + 1), S(dz[k]), S(dz[k + 1])));
        }
    }
    w.pres[c] = pres;
    w.theta[widx<P>(D, k, i)] = th;
    w.rt2[widx<P>(D, k, i)] = rt2;
    if constexpr (!P) {
        w.uf[c] = uf;
        w.wf[c] = wf;
    }
}

#include "moist_acoustic.h"
#include "moist_acoustic_pairs.h"

#include "moist_scalar.h"
#include "moist_slow.h"

// =============================================================================
// Surface fluxes (surface_fluxes!, moist.jl:710-752), one thread per column.
// =============================================================================
template <class S>
__global__ void __launch_bounds__(64) k_surface(const Dom *__restrict__ doms, const int2 *__restrict__ ctiles,
                                                ZVec zv, SSet<S> s, Work<S> w) {
    using R = S;
    using KC = K<R>;
    const int2 t = ctiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD ordering
    const int m = t.x;
    const Dom &D = doms[m];
    const int i = t.y + threadIdx.x;
    const int nx = D.nx;
    if (i >= nx || !(D.flags & F_SF)) return;
    const int64_t c = D.off2 + i;
    const R rho = full_rho(s, c, zv, D, 0), rt = full_rt(s, c, zv, D, 0);
    const R qv = s.f[QV][c] / rho, qc = s.f[QC][c] / rho, qr = s.f[QR][c] / rho;
    const R qd = R(1.0) - qv - qc - qr;
    const R r = KC::Rd * qd + KC::Rv * qv;
    const R cp = KC::Cpd * qd + KC::Cpv * qv + KC::Cpl * (qc + qr);
    const R cv = cp - r;
    const R ru_left = i == 0 ? R(w.sfh[2 * m + 0]) : s.f[RHOU][c - 1];
    const R rho_right = i == nx - 1 ? R(w.sfh[2 * m + 1]) : full_rho(s, c + 1, zv, D, 0);
    R uabs = m_fabs(ru_left + s.f[RHOU][c]) * R(0.5) / rho;
    uabs = jmax(uabs, R(D.uabs_min));
    const R theta_ks = rt / rho;
    const R pi_ks = m_pow(r * rt / KC::P0, r / cv);
    const R dz0 = R(zv.dz[D.offz]);
    const R pi_surf = pi_ks + dz0 * R(0.5) * KC::GRAV / (cp * theta_ks);
    const R p_surf = KC::P0 * m_pow(pi_surf, cp / r);
    const R Ts = R(D.temp_surf);
    const R rho_surf = p_surf / (r * Ts);
    const R theta_surf = Ts / pi_surf;
    const R u = s.f[RHOU][c] / (rho + rho_right) * R(2.0);
    const R mflux = -rho_surf * R(D.Cd) * u * m_fabs(u);
    const R d_u = mflux / dz0;
    const R sh = rho_surf * R(D.Ch) * uabs * (theta_surf - theta_ks) * pi_surf;
    R es, des, qs, dq;
    sat_vp(Ts, es, des);
    sat_q(p_surf, es, qs, dq);
    const R lh = rho_surf * R(D.Ce) * uabs * (qs - qv);
    const int64_t x = D.offx + i;
    w.mflux[x] = mflux;
    w.d_u_sf[x] = d_u;
    w.d_th_sf[x] = sh / (pi_ks * dz0);
    w.shflux[x] = sh * cp;
    w.d_qv_sf[x] = lh / dz0;
    w.lhflux[x] = lh * KC::Lv;
    if (isnan(d_u)) report_nan(w.nan_key, ((*w.step_ctr * 3 + 3) << 4) | 8);
}

// =============================================================================
// Gray long-wave radiation (gray_radiation!, moist.jl:776-840), one thread
// per column.  The upward flux is parked in d_th_rd during the sweep up and
// replaced by the tendency on the sweep down.
// =============================================================================
template <class S>
__global__ void __launch_bounds__(64) k_radiation(const Dom *__restrict__ doms, const int2 *__restrict__ ctiles,
                                                  ZVec zv, SSet<S> s, Work<S> w) {
    using R = S;
    using KC = K<R>;
    const int2 t = ctiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD ordering
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x;
    const int nx = D.nx, nz = D.nz;
    if (i >= nx || !(D.flags & F_RD)) return;
    const double *zd = zv.z + D.offz, *dzd = zv.dz + D.offz;
    const double *th_ref = w.theta_ref + D.offz;
    auto cellv = [&](int k, R &eps, R &tau, R &T4, R &cpk, R &pex, R &rho, R &rt) {
        const int64_t c = D.off2 + (int64_t)k * nx + i;
        rho = full_rho(s, c, zv, D, k);
        rt = full_rt(s, c, zv, D, k);
        const R qv = s.f[QV][c] / rho, qc = s.f[QC][c] / rho, qr = s.f[QR][c] / rho;
        const R qd = R(1.0) - qv - qc - qr;
        const R r = KC::Rd * qd + KC::Rv * qv;
        cpk = KC::Cpd * qd + KC::Cpv * qv + KC::Cpl * (qc + qr);
        const R cv = cpk - r;
        pex = m_pow(r * rt / KC::P0, r / cv);
        const R temp = rt / rho * pex;
        const R dzk = R(dzd[k]);
        const R dopt = R(D.k_abs_qv) * s.f[QV][c] * dzk + R(D.k_abs_qc) * s.f[QC][c] * dzk +
                       R(D.k_abs_qr) * s.f[QR][c] * dzk;
        eps = R(1.0) - m_exp(-dopt);
        tau = R(1.0) - eps;
        T4 = temp * temp * temp * temp;
    };
    const R Ts = R(D.temp_surf);
    const R up0 = R(D.surf_emissivity) * KC::SIGMA * (Ts * Ts * Ts * Ts);
    R up = up0;
    for (int k = 0; k < nz; ++k) {
        R eps, tau, T4, cpk, pex, rho, rt;
        cellv(k, eps, tau, T4, cpk, pex, rho, rt);
        up = up * tau + eps * KC::SIGMA * T4;
        w.d_th_rd[D.off2 + (int64_t)k * nx + i] = up;  // LW_up at face k+1
    }
    R dn_above = R(0.0);  // LW_dn[k+1]
    const R dz_trans = R(2.0e3);
    for (int k = nz - 1; k >= 0; --k) {
        R eps, tau, T4, cpk, pex, rho, rt;
        cellv(k, eps, tau, T4, cpk, pex, rho, rt);
        const R dn = dn_above * tau + eps * KC::SIGMA * T4;
        const int64_t c = D.off2 + (int64_t)k * nx + i;
        const R up_kp = w.d_th_rd[c];
        const R up_k = k == 0 ? up0 : w.d_th_rd[c - nx];
        const R divF = ((up_kp - dn_above) - (up_k - dn)) / R(dzd[k]);
        const R hr = -divF / (cpk * pex);
        const R alpha = R(0.5) * (R(1.0) + m_tanh((R(zd[k]) - (R(D.z_tropopose) + dz_trans * R(0.5))) / dz_trans)) /
                        R(D.tau_cooling);
        const R cr = -(rt - rho * R(th_ref[k])) * alpha;
        w.d_th_rd[c] = hr + cr;
        dn_above = dn;
    }
}

// =============================================================================
// Kessler microphysics (cloud_microphysics!, moist.jl:554-671), rain
// sedimentation (rain_sedimentation!, moist.jl:674-707) and the physics
// source terms (moist.jl:1632-1644); accumulates surface rain
// (moist.jl:1680-1682).  One thread per cell of the step-start state.
// =============================================================================
template <class S>
__device__ __forceinline__ S term_velocity(S rho_qr, S rho) {
    using R = S;
    const R eff = jmax(R(0.0), rho_qr);
    const R v = K<R>::A_term * m_pow(fdiv(eff, rho), K<R>::B_term);
    return eff > R(1e-12) ? v : R(0.0);
}

template <class S>
__global__ void __launch_bounds__(256) k_physics(const Dom *__restrict__ doms, const int4 *__restrict__ tiles,
                                                 ZVec zv, SSet<S> s, Work<S> w, int store_diag,
                                                 int diag_pairs, SSet<double> wide) {
    // No FP contraction here, in either form: the condensate of the
    // saturation adjustment is qv_init - qv, and with qv_init = rho_qv / rho
    // fused into that subtraction it would be the product's rounding error
    // (about 1e-9 kg/m^3 of cloud water in every cell of a subsaturated
    // column) instead of zero.
#pragma clang fp contract(off)
    using R = S;
    using KC = K<R>;
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD ordering
    const Dom &D = doms[t.x];
    const int i = t.y + threadIdx.x, k = t.z + threadIdx.y;
    const int nx = D.nx, nz = D.nz, ke = nz - 1;
    if (i >= nx || k >= nz) return;
    const int64_t c = D.off2 + (int64_t)k * nx + i;
    // fields by the member's base pointer and 32-bit byte offsets
    const int64_t o2 = D.off2;
    const uint32_t q = (uint32_t)sizeof(S) * moff<false>(nx, k, i);
    auto L = [&](const S *f, uint32_t b) -> R { return ldb(f + o2, b); };
    const R dt = R(D.dt);
    const int fl = D.flags;
    const R s_rho = L(s.f[RHO], q), s_rt = L(s.f[RT], q);
    R rho_k = s_rho, rt = s_rt;
    if constexpr (is_mixed<S>()) {
        rho_k = rho_k + S(zv.rho_bar[D.offz + k]);
        rt = rt + S(zv.rt_bar[D.offz + k]);
    }
    // every other global read of the kernel, before its first store: the
    // rain and density one level up (sedimentation), radiation and surface terms
    const int kup = k < ke ? k + 1 : k;
    const uint32_t qu = (uint32_t)sizeof(S) * moff<false>(nx, kup, i);
    const R qr_c = L(s.f[QR], q), qr_u = L(s.f[QR], qu);
    R rho_u = L(s.f[RHO], qu);
    if constexpr (is_mixed<S>()) rho_u = rho_u + S(zv.rho_bar[D.offz + kup]);
    const R s_qv = L(s.f[QV], q), s_qc = L(s.f[QC], q);
    const R th_rd = (fl & F_RD) ? L(w.d_th_rd, q) : R(0.0);
    const bool sf = (k == 0) && (fl & F_SF);
    const int64_t x = D.offx + i;
    const R qv_sf = sf ? w.d_qv_sf[x] : R(0.0), th_sf = sf ? w.d_th_sf[x] : R(0.0);
    // wide (the mixed form's recording pass of the reverse sweep; null
    // pointers: none): this state as full FP64 values, rho and rho_theta with
    // their base state added
    double wv[is_mixed<S>() ? 7 : 1];
    if constexpr (is_mixed<S>()) {
        if (wide.f[0]) {
#pragma unroll
            for (int f = 0; f < 7; ++f) wv[f] = double(s.f[f][c]);
            wv[RHO] += zv.rho_bar[D.offz + k];
            wv[RT] += zv.rt_bar[D.offz + k];
        }
    }
    // mixed form: the first stage's diagnostics of this same state (what
    // k_diag would compute), stored in the level-pair layout with the rest
    R dg_th = R(0.0), dg_pres = R(0.0), dg_rt2 = R(0.0);
    if constexpr (is_mixed<S>()) {
        if (diag_pairs)
            diag_point<S>(D, zv, k, s_rho, s_rt, s_qv, s_qc, qr_c, dg_th, dg_pres, dg_rt2);
    }
    R dqv = R(0.0), dqc = R(0.0), dqr = R(0.0), dth = R(0.0);
    if (fl & F_MP) {
        const R rinv = fdiv(R(1.0), rho_k);
        R theta_k = rt * rinv;
        R qv = s_qv * rinv, qc = s_qc * rinv, qr = qr_c * rinv;
        const R qd = R(1.0) - qv - qc - qr;
        R r = KC::Rd * qd + KC::Rv * qv;
        R cp = KC::Cpd * qd + KC::Cpv * qv + KC::Cpl * (qc + qr);
        R cv = cp - r;
        R Pi = m_pow(fdiv(r * rt, KC::P0), fdiv(r, cv));
        R Tk = theta_k * Pi;
        R pk = KC::P0 * m_pow(Pi, fdiv(cp, r));
        R lhe = fdiv(KC::Lv, cp * Pi);
        const R qv_init = qv;
        R qsat = R(0.0);
        for (int it = 0; it < 3; ++it) {
            R es, des, dqs_des;
            sat_vp(Tk, es, des);
            sat_q(pk, es, qsat, dqs_des);
            const R dqs_dT = dqs_des * des;
            R ce = fdiv(qv - qsat, R(1.0) + fdiv(KC::Lv, cp) * dqs_dT);
            ce = jmin(ce, qv);
            ce = jmax(ce, -qc);
            qv -= ce;
            qc += ce;
            theta_k += ce * lhe;
            r = KC::Rd * qd + KC::Rv * qv;
            cp = KC::Cpd * qd + KC::Cpv * qv + KC::Cpl * (qc + qr);
            cv = cp - r;
            Pi = m_pow(fdiv(r * rho_k * theta_k, KC::P0), fdiv(r, cv));
            Tk = theta_k * Pi;
            pk = KC::P0 * m_pow(Pi, fdiv(cp, r));
            lhe = fdiv(KC::Lv, cp * Pi);
        }
        const R ce = qv_init - qv;
        const R rate = fdiv(ce * rho_k, dt);
        dqv -= rate;
        dqc += rate;
        dth += rate * lhe;
        if (fl & F_AUTO) {
            R ra = R(D.K_auto) * jmax(R(0.0), qc - R(D.Qc_threshold));
            ra = jmin(ra, fdiv(qc, dt));
            qc -= ra * dt;
            qr += ra * dt;
            ra *= rho_k;
            dqc -= ra;
            dqr += ra;
        }
        if (fl & F_ACCR) {
            R ra = R(0.0);
            if (qr > R(1e-12) && qc > R(1e-12)) {
                ra = R(D.K_accr) * qc * qr;
                ra = jmin(ra, fdiv(qc, dt));
            }
            qc -= ra * dt;
            qr += ra * dt;
            ra *= rho_k;
            dqc -= ra;
            dqr += ra;
        }
        if (fl & F_REVAP) {
            R re = R(0.0);
            if (qr > R(1e-12) && qv < qsat) {
                re = KC::K_evap * (qsat - qv) * m_sqrt(rho_k * qr);
                re = fdiv(jmin(re, qr), dt);
            }
            re *= rho_k;
            dqr -= re;
            dqv += re;
            r = KC::Rd * qd + KC::Rv * qv;
            cp = KC::Cpd * qd + KC::Cpv * qv + KC::Cpl * (qc + qr);
            cv = cp - r;
            Pi = m_pow(fdiv(r * rho_k * theta_k, KC::P0), fdiv(r, cv));
            lhe = fdiv(KC::Lv, cp * Pi);
            dth -= re * lhe;
        }
    }
    R dsed = R(0.0), fkm = R(0.0);
    const bool sed = (fl & F_MP) && (fl & F_SED);
    if (sed) {
        const R vt_k = term_velocity(qr_c, rho_k);
        fkm = qr_c * vt_k;
        R fkp = R(0.0);
        if (k < ke) fkp = qr_u * term_velocity(qr_u, rho_u);
        dsed = fdiv(fkp - fkm, R(zv.dz[D.offz + k]));
    }
    R d_rho = dsed;
    if (sf) d_rho += qv_sf;
    R d_rt = dth + th_rd + fdiv(d_rho * rt, rho_k);
    if (sf) d_rt += th_sf;
    R d_qv = dqv;
    if (sf) d_qv += qv_sf;
    const R d_qr = dqr + dsed;
    if (sed && k == 0) {  // accumulated rain (FP64)
        if constexpr (is_mixed<S>()) w.prec[x] += double(fkm) * D.dt;
        else w.prec[x] += fkm * dt;
    }
    stb(w.d_rho + o2, q, d_rho);
    stb(w.d_rt + o2, q, d_rt);
    stb(w.d_qv + o2, q, d_qv);
    stb(w.d_qc + o2, q, dqc);
    stb(w.d_qr + o2, q, d_qr);
    if constexpr (is_mixed<S>()) {
        if (diag_pairs) {
            const uint32_t qp = 4u * moff<true>(nx, k, i);
            stb(w.pres + o2, q, dg_pres);
            stb(w.theta + o2, qp, dg_th);
            stb(w.rt2 + o2, qp, dg_rt2);
        }
    }
    if constexpr (is_mixed<S>()) {
        if (wide.f[0])
#pragma unroll
            for (int f = 0; f < 7; ++f) wide.f[f][c] = wv[f];
    }
    if (store_diag) {
        stb(w.dg_th_mp + o2, q, dth);
        stb(w.dg_qv_mp + o2, q, dqv);
        stb(w.dg_qr_mp + o2, q, dqr);
        stb(w.dg_qr_sed + o2, q, dsed);
    }
    int f = 0;
    if (isnan(d_rt)) f = 9;
    else if (isnan(d_qv)) f = 10;
    else if (isnan(dqc)) f = 11;
    else if (isnan(d_qr)) f = 12;
    if (f) report_nan(w.nan_key, ((*w.step_ctr * 3 + 3) << 4) | f);
}


// =============================================================================
// Layout conversion between the host State arrays ((ka, ia) column-major,
// x halos) and the device pools.
// =============================================================================
template <class S, class H>
__global__ void k_upload(Dom D, const H *__restrict__ host, S *__restrict__ dst, const double *__restrict__ base,
                         int halo, int ka) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= D.nx || k >= D.nz) return;
    double v = double(host[(int64_t)(i + halo) * ka + k]);
    if (base) v -= base[k];
    dst[D.off2 + (int64_t)k * D.nx + i] = S(v);
}

// Writes host-layout columns: interior from `src`; halo columns from the
// periodic partner column of `halo_src` (when given).
template <class S, class H>
__global__ void k_download(Dom D, const S *__restrict__ src, const S *__restrict__ halo_src, H *__restrict__ host,
                           const double *__restrict__ base, int halo, int ka) {
    const int ih = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int ia = D.nx + 2 * halo;
    if (ih >= ia || k >= D.nz) return;
    const int i = ih - halo;
    const S *from;
    int j;
    if (i >= 0 && i < D.nx) {
        from = src;
        j = i;
    } else {
        if (!halo_src) return;
        from = halo_src;
        j = ((i % D.nx) + D.nx) % D.nx;
    }
    double v = double(from[D.off2 + (int64_t)k * D.nx + j]);
    if (base) v += base[k];
    host[(int64_t)ih * ka + k] = H(v);
}

// =============================================================================
// The float instantiations of the step kernels (the mixed form's step, and the
// FP32 recording pass of its adjoint) live in a translation unit of their own,
// moist_mixed.hip, compiled with FP contraction: a * b + c becomes one fused
// multiply-add, rounded once instead of twice (fewer instructions, and no less
// accurate).  The FP64 form keeps the reference's separately rounded
// operations: moist_api.hip, compiled without contraction, declares these
// instantiations extern (MO_MIXED_KERNELS(extern)) and so instantiates none of
// them itself.  Every float step kernel the host launches must be listed here.
// =============================================================================
#define MO_AC_ARGS const Dom *, const int4 *, ZVec, SSet<float>, Work<float>, AcBuf<float>, int, double, int, int
#define MO_AC_SHAPE(EXT, L_, NT_, W_)                                                       \
    EXT template __global__ void k_acoustic<float, L_, NT_, W_>(MO_AC_ARGS);                \
    EXT template __global__ void k_acoustic_pairs<L_, NT_, W_, false>(MO_AC_ARGS);          \
    EXT template __global__ void k_acoustic_pairs<L_, NT_, W_, true>(MO_AC_ARGS);
#define MO_SCALAR_ARGS                                                                                       \
    const Dom *, const int4 *, ZVec, SSet<float>, SSet<float>, SSet<float>, Work<float>, const float *,   \
        const float *, int, double, int, float *, float *, float *, SSet<double>
#define MO_MIXED_KERNELS(EXT)                                                                                   \
    EXT template __global__ void k_diag<float, false>(const Dom *, const int4 *, ZVec, SSet<float>, Work<float>); \
    EXT template __global__ void k_diag<float, true>(const Dom *, const int4 *, ZVec, SSet<float>, Work<float>);  \
    EXT template __global__ void k_surface<float>(const Dom *, const int2 *, ZVec, SSet<float>, Work<float>);     \
    EXT template __global__ void k_radiation<float>(const Dom *, const int2 *, ZVec, SSet<float>, Work<float>);   \
    EXT template __global__ void k_physics<float>(const Dom *, const int4 *, ZVec, SSet<float>, Work<float>, int, \
                                                  int, SSet<double>);                                            \
    EXT template __global__ void k_slow<float, false>(const Dom *, const int4 *, ZVec, SSet<float>, SSet<float>,  \
                                                      Work<float>, float *, int, int *);                         \
    EXT template __global__ void k_slow<float, true>(const Dom *, const int4 *, ZVec, SSet<float>, SSet<float>,   \
                                                     Work<float>, float *, int, int *);                          \
    EXT template __global__ void k_scalar<float, false>(MO_SCALAR_ARGS);                                        \
    EXT template __global__ void k_scalar<float, true>(MO_SCALAR_ARGS);                                         \
    MO_AC_SHAPE(EXT, 2, 256, 1)                                                                                   \
    MO_AC_SHAPE(EXT, 2, 512, 1)                                                                                   \
    MO_AC_SHAPE(EXT, 2, 1024, 1)                                                                                  \
    MO_AC_SHAPE(EXT, 4, 256, 1)                                                                                   \
    MO_AC_SHAPE(EXT, 4, 512, 1)                                                                                   \
    MO_AC_SHAPE(EXT, 4, 1024, 1)                                                                                  \
    MO_AC_SHAPE(EXT, 2, 512, 8)                                                                                   \
    MO_AC_SHAPE(EXT, 2, 1024, 8)                                                                                  \
    MO_AC_SHAPE(EXT, 4, 512, 8)                                                                                   \
    MO_AC_SHAPE(EXT, 4, 1024, 8)

}  // namespace mo
