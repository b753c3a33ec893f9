// moist_oracle.cpp — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU restatement of the reference's time-stepping core (seiya/moist2d,
// moist.jl), used only as the checker for the HIP path: by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing in
// moist2d_amd/ links, loads or calls it.
//
// It follows, as text (no code was run or copied):
//   constants            moist.jl:26-60
//   exchange_halo!, f2h  moist.jl:233-251
//   Tetens, qsat         moist.jl:254-273
//   cloud_microphysics!  moist.jl:554-671
//   rain_sedimentation!  moist.jl:674-707
//   surface_fluxes!      moist.jl:710-752
//   gray_radiation!      moist.jl:776-840
//   koren_limiter        moist.jl:844-847
//   compute_limited_flux_x / _z  moist.jl:850-945
//   solve_tridiagonal!   moist.jl:952-971
//   implicit_correction! moist.jl:974-1030
//   compute_step!        moist.jl:1039-1580
//   rk3_step!            moist.jl:1589-1683
// Julia semantics reproduced: literal integer powers as multiplications,
// Base.min/max for IEEE floats (NaN-propagating, signed-zero rule), left
// associativity of + - * /, Float64 promotion of fact_tp in the Float32
// path (moist.jl:1375), and every quirk of the reference's loop nests
// (linear-index positivity limiter, minus sign of the downward z Koren face,
// height-weighted f2h in R_rho_w, mixed branch tests in the acoustic
// rho_theta flux, off-centering term a*rho_w[k]).
//
// Array layout is the reference's: column-major (ka, ia), z fastest, x halos
// included; 0-based here (i = 0..ia-1, k = 0..ka-1; is = halo,
// ie = halo + nx - 1, ks = 0, ke = nz - 1).
//
// Scalar transport order: the reference writes rho_q* in place while reading
// neighbours (moist.jl:1469-1518).  ORDER_SERIAL reproduces one thread's
// sweep (i outer, k inner); ORDER_SNAPSHOT reads every neighbour from the
// stage input.  Build with -ffp-contract=off so that no FMA contraction
// happens (Julia does not contract outside @turbo/@fastmath).
//
// Parity status of each function is recorded in the oracle's test module
// tests/test_oracle.py (which pins what).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/moist2d.h"
#include "moist_dual.h"

namespace orc {

enum { ORDER_SNAPSHOT = 0, ORDER_SERIAL = 1 };

// ---- Julia Base.min / Base.max for IEEE floats (base/math.jl) -------------
template <class T> static inline T jmax(T x, T y) {
    T diff = x - y;
    T am = std::signbit(diff) ? y : x;
    return (std::isnan(x) || std::isnan(y)) ? diff : am;
}
template <class T> static inline T jmin(T x, T y) {
    T diff = x - y;
    T am = std::signbit(diff) ? x : y;
    return (std::isnan(x) || std::isnan(y)) ? diff : am;
}

// ---- constants (moist.jl:30-59), built in T like set_constants!(T) -------
template <class T> struct C {
    T GRAV, Rd, Rv, Cpd, Cpv, Cpl, Cvd, P0, Lv, SIGMA, EPSvap;
    T K_auto, Qc_threshold, K_accr, K_evap, A_term, B_term;
    T T0, ES0, A_SAT, B_SAT, EPS;
    explicit C(const moist_params *p) {
        GRAV = T(9.81); Rd = T(287.04); Rv = T(461.5); Cpd = T(1004.64);
        Cpv = T(1846.0); Cpl = T(4218.0); Cvd = Cpd - Rd; P0 = T(1e5);
        Lv = T(2.501e6); SIGMA = T(5.67e-8); EPSvap = Rd / Rv;
        K_auto = T(p ? p->K_auto : 1e-3);
        Qc_threshold = T(p ? p->Qc_threshold : 1e-3);
        K_accr = T(p ? p->K_accr : 2.2);
        K_evap = T(1e-2); A_term = T(36.3); B_term = T(0.1346);
        T0 = T(273.15); ES0 = T(610.78); A_SAT = T(17.67); B_SAT = T(243.5);
        EPS = T(1e-12);
    }
};

// ---- grid view of Params{T} -----------------------------------------------
template <class T> struct G {
    int nx, nz, halo, ia, ka, is, ie, ks, ke, ns;
    bool do_mp, do_auto, do_accr, do_revap, do_sed, do_rd, do_sf;
    T dx, dt, H, nu, kappa, gamma_d, beta, zf_sponge;
    T Cd, Ch, Ce, uabs_min, temp_surf, k_abs_qv, k_abs_qc, k_abs_qr,
        surf_emissivity, tau_cooling, z_tropopose;
    std::vector<T> z, zf, dz;
    C<T> c;
    explicit G(const moist_params *p) : c(p) {
        nx = p->nx; nz = p->nz; halo = p->halo; ns = p->ns;
        ia = nx + 2 * halo; ka = nz; is = halo; ie = halo + nx - 1; ks = 0; ke = nz - 1;
        do_mp = p->do_mp; do_auto = p->do_auto; do_accr = p->do_accr;
        do_revap = p->do_revap; do_sed = p->do_sed; do_rd = p->do_rd; do_sf = p->do_sf;
        dx = T(p->dx); dt = T(p->dt); H = T(p->H); nu = T(p->nu); kappa = T(p->kappa);
        gamma_d = T(p->gamma_d); beta = T(p->beta_offcenter); zf_sponge = T(p->zf_sponge);
        Cd = T(p->Cd); Ch = T(p->Ch); Ce = T(p->Ce); uabs_min = T(p->uabs_min);
        temp_surf = T(p->temp_surf); k_abs_qv = T(p->k_abs_qv); k_abs_qc = T(p->k_abs_qc);
        k_abs_qr = T(p->k_abs_qr); surf_emissivity = T(p->surf_emissivity);
        tau_cooling = T(p->tau_cooling); z_tropopose = T(p->z_tropopose);
        z.resize(nz); zf.resize(nz); dz.resize(nz);
        for (int k = 0; k < nz; ++k) { z[k] = T(p->z[k]); zf[k] = T(p->zf[k]); dz[k] = T(p->dz[k]); }
    }
    size_t n2() const { return (size_t)ka * ia; }
    size_t ix(int k, int i) const { return (size_t)i * ka + k; }
};

// ---- helpers (moist.jl:233-273) -------------------------------------------
template <class T> static void exchange_halo(T *A, const G<T> &p) {
    for (int h = 0; h < p.halo; ++h) {
        std::memcpy(A + p.ix(0, h), A + p.ix(0, p.ie - p.halo + 1 + h), sizeof(T) * p.ka);
        std::memcpy(A + p.ix(0, p.ie + 1 + h), A + p.ix(0, p.is + h), sizeof(T) * p.ka);
    }
}

template <class T> static inline T f2h(T vf1, T vf2, T dz1, T dz2) {
    return (vf1 * dz2 + vf2 * dz1) / (dz1 + dz2);
}

template <class T> static inline void sat_vp(const C<T> &c, T Tk, T &es, T &des_dT) {
    T T_c = Tk - c.T0;
    T tmp = c.B_SAT + T_c;
    es = c.ES0 * std::exp(c.A_SAT * T_c / tmp);
    des_dT = es * c.A_SAT * (tmp - T_c) / (tmp * tmp);
}

template <class T> static inline void sat_q(const C<T> &c, T p, T es, T &qsat, T &dqsat_des) {
    T tmp = p - (T(1.0) - c.EPSvap) * es;
    if (tmp > c.EPS) {
        qsat = c.EPSvap * es / tmp;
        dqsat_des = c.EPSvap * p / (tmp * tmp);
    } else {
        qsat = T(0.0);
        dqsat_des = T(1.0);
    }
}

// ---- Koren advection numerics (moist.jl:844-945) --------------------------
template <class T> static inline T koren(T r) {
    return jmax(T(0.0), jmin(T(2.0) * r, jmin((T(1.0) + T(2.0) * r) / T(3.0), T(2.0))));
}

// Flux at face i+1/2 of row k (moist.jl:850-892).
template <class T>
static inline T flux_x(const G<T> &p, const T *sr, const T *rho, const T *rho_u, int k, int i) {
    const C<T> &c = p.c;
    T rhou_face = rho_u[p.ix(k, i)];
    T s_i = sr[p.ix(k, i)] / rho[p.ix(k, i)];
    T s_ip1 = sr[p.ix(k, i + 1)] / rho[p.ix(k, i + 1)];
    T s_face;
    if (rhou_face > c.EPS) {
        T s_im1 = sr[p.ix(k, i - 1)] / rho[p.ix(k, i - 1)];
        T r_num = s_ip1 - s_i, r_den = s_i - s_im1;
        T r = std::fabs(r_den) < c.EPS ? T(1.0) : r_num / r_den;
        s_face = s_i + T(0.5) * koren(r) * r_den;
    } else if (rhou_face < -c.EPS) {
        T s_ip2 = sr[p.ix(k, i + 2)] / rho[p.ix(k, i + 2)];
        T r_num = s_i - s_ip1, r_den = s_ip1 - s_ip2;
        T r = std::fabs(r_den) < c.EPS ? T(1.0) : r_num / r_den;
        s_face = s_ip1 + T(0.5) * koren(r) * r_den;
    } else {
        s_face = T(0.5) * (s_i + s_ip1);
    }
    return s_face * rhou_face;
}

// Flux at face k+1/2 of column i (moist.jl:895-945).  `lin` is the array the
// positivity limiters index linearly (element (k, 0) = first halo column).
template <class T>
static inline T flux_z(const G<T> &p, const T *sr, const T *lin, const T *rho, const T *rho_w,
                       int k, int i) {
    const C<T> &c = p.c;
    T rhow_face = rho_w[p.ix(k, i)];
    T s_k = sr[p.ix(k, i)] / rho[p.ix(k, i)];
    T s_kp1 = sr[p.ix(k + 1, i)] / rho[p.ix(k + 1, i)];
    T s_face;
    if (rhow_face > c.EPS) {
        if (k == p.ks) {
            s_face = f2h(s_k, s_kp1, p.dz[k], p.dz[k + 1]);
            s_face = jmin(s_face, lin[k] / (rhow_face * p.dt));
        } else {
            T s_km1 = sr[p.ix(k - 1, i)] / rho[p.ix(k - 1, i)];
            T r_num = s_kp1 - s_k, r_den = s_k - s_km1;
            T r = std::fabs(r_den) < c.EPS ? T(1.0) : r_num / r_den;
            s_face = s_k + T(0.5) * koren(r) * r_den;
        }
    } else if (rhow_face < -c.EPS) {
        if (k == p.ke - 1) {
            s_face = f2h(s_k, s_kp1, p.dz[k], p.dz[k + 1]);
            s_face = jmin(s_face, lin[k + 1] / (-rhow_face * p.dt));
        } else {
            T s_kp2 = sr[p.ix(k + 2, i)] / rho[p.ix(k + 2, i)];
            T r_num = s_k - s_kp1, r_den = s_kp1 - s_kp2;
            T r = std::fabs(r_den) < c.EPS ? T(1.0) : r_num / r_den;
            s_face = s_kp1 - T(0.5) * koren(r) * r_den;
        }
    } else {
        s_face = f2h(s_k, s_kp1, p.dz[k], p.dz[k + 1]);
    }
    return s_face * rhow_face;
}

// ---- Thomas algorithm (moist.jl:952-971) ----------------------------------
template <class T> static void thomas(int n, const T *a, const T *b, T *c, T *d) {
    T c1 = c[0] / b[0];
    T d1 = d[0] / b[0];
    c[0] = c1;
    d[0] = d1;
    for (int i = 1; i < n; ++i) {
        T temp = b[i] - a[i] * c[i - 1];
        T ci = c[i] / temp;
        T di = (d[i] - a[i] * d[i - 1]) / temp;
        c[i] = ci;
        d[i] = di;
    }
    for (int i = n - 2; i >= 0; --i) d[i] = d[i] - c[i] * d[i + 1];
}

// ---- w-p system of one column (moist.jl:974-1030) -------------------------
// rho_w, theta, rt2pres are the column (length ka); drho_w length ka.
template <class T>
static void implicit_correction(const G<T> &p, T *rho_w, const T *drho_w, const T *theta,
                                const T *rt2pres, T dt) {
    const int ks = p.ks, ke = p.ke, n = p.nz - 1;
    const C<T> &c0 = p.c;
    if (n <= 0) return;
    std::vector<T> th(p.ka), a(n), b(n), c(n), d(n);
    for (int k = ks; k <= ke - 1; ++k) th[k] = f2h(theta[k], theta[k + 1], p.dz[k], p.dz[k + 1]);
    T fact_tp = T(0.5) * (T(1.0) + p.beta);
    T fact_tm = T(0.5) * (T(1.0) - p.beta);
    T dt_tp2 = (dt * fact_tp) * (dt * fact_tp);
    for (int k = ks; k <= ke - 1; ++k) {
        int j = k - ks;
        T zz = p.z[k + 1] - p.z[k];
        b[j] = dt_tp2 * (rt2pres[k] / p.dz[k] + rt2pres[k + 1] / p.dz[k + 1]) * th[k] / zz;
        if (k > ks)
            a[j] = -dt_tp2 * (rt2pres[k] * th[k - 1] / (zz * p.dz[k]) - c0.GRAV / (p.dz[k] + p.dz[k + 1]));
        else
            a[j] = T(0.0);  // a[1] unused by the solver
        if (k < ke - 1)
            c[j] = -dt_tp2 * (rt2pres[k + 1] * th[k + 1] / (zz * p.dz[k + 1]) + c0.GRAV / (p.dz[k] + p.dz[k + 1]));
        else
            c[j] = T(0.0);  // c[n] is read by the solver and discarded
        d[j] = rho_w[k] + drho_w[k];
    }
    T fact = fact_tm / fact_tp;
    for (int k = ks; k <= ke - 1; ++k) {
        int j = k - ks;
        if (k > ks) d[j] -= a[j] * fact * rho_w[k];  // a multiplies rho_w[k] (sic)
        d[j] -= b[j] * fact * rho_w[k];
        b[j] += T(1.0);
        if (k < ke - 1) d[j] -= c[j] * fact * rho_w[k + 1];
    }
    thomas(n, a.data(), b.data(), c.data(), d.data());
    for (int k = ks; k <= ke - 1; ++k) rho_w[k] = d[k - ks];
}

// ---- state views ------------------------------------------------------------
template <class T> struct St {  // the 7 prognostic fields of State
    T *f[7];
    T *rho() const { return f[0]; }
    T *rho_u() const { return f[1]; }
    T *rho_w() const { return f[2]; }
    T *rho_theta() const { return f[3]; }
    T *rho_qv() const { return f[4]; }
    T *rho_qc() const { return f[5]; }
    T *rho_qr() const { return f[6]; }
};

template <class T> static inline T upw3(T m, T c0, T p1) {  // (-m + 5c + 2p)/6 numerator
    return -m + T(5.0) * c0 + T(2.0) * p1;
}

// What the adjoint of one compute_step! needs from its forward pass: the
// stage diagnostics and slow tendencies, the averaged momenta, and per
// acoustic substep rho_theta' and rho_w' at its start, rho_u' after its
// update and rho_w' after the implicit solve.
template <class T> struct Rec {
    std::vector<T> pres, theta, rt2pres, uf, wf, R_rho, R_rho_u, R_rho_w, R_rho_theta, mom_x, mom_z;
    std::vector<std::vector<T>> rtp_start, rup_new, rwp_start, rwp_new;
};

// ---- compute_step! (moist.jl:1039-1580) -----------------------------------
template <class T>
static int compute_step(const G<T> &p, const T *d_rho_phys, const T *d_rho_u_phys,
                        const T *d_rho_theta_phys, const T *d_qv_phys, const T *d_qc_phys,
                        const T *d_qr_phys, const St<T> &s, const St<T> &s0, T dt, int ns,
                        int order, Rec<T> *rec = nullptr) {
    const C<T> &c = p.c;
    const int is = p.is, ie = p.ie, ks = p.ks, ke = p.ke;
    const size_t N = p.n2();
    const T dx = p.dx;
    T *rho = s.rho(), *rho_u = s.rho_u(), *rho_w = s.rho_w(), *rho_theta = s.rho_theta();

    std::vector<T> pres(N), theta(N), rt2pres(N);
#pragma omp parallel for schedule(static)
    for (int i = is; i <= ie + 1; ++i) {
        for (int k = ks; k <= ke; ++k) {
            size_t q = p.ix(k, i);
            T qv = s.rho_qv()[q] / rho[q];
            T qc = s.rho_qc()[q] / rho[q];
            T qr = s.rho_qr()[q] / rho[q];
            T qd = T(1.0) - qv - qc - qr;
            T r = c.Rd * qd + c.Rv * qv;
            T cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
            T cv = cp - r;
            pres[q] = c.P0 * std::pow(r * rho_theta[q] / c.P0, cp / cv);
            theta[q] = rho_theta[q] / rho[q];
            rt2pres[q] = cp / cv * pres[q] / rho_theta[q];
        }
    }
    exchange_halo(pres.data(), p);
    exchange_halo(theta.data(), p);
    exchange_halo(rt2pres.data(), p);
    exchange_halo(rho, p);
    exchange_halo(rho_u, p);
    exchange_halo(rho_w, p);
    if (rec) {
        rec->pres = pres;
        rec->theta = theta;
        rec->rt2pres = rt2pres;
        rec->rtp_start.clear();
        rec->rup_new.clear();
        rec->rwp_start.clear();
        rec->rwp_new.clear();
    }

    std::vector<T> R_rho(N, T(0)), R_rho_u(N, T(0)), R_rho_w(N, T(0)), R_rho_theta(N, T(0));

    // R_rho (moist.jl:1079-1086)
#pragma omp parallel for schedule(static)
    for (int i = is; i <= ie; ++i) {
        for (int k = ks; k <= ke; ++k) {
            T flux_kp = k == ke ? T(0.0) : rho_w[p.ix(k, i)];
            T flux_km = k == ks ? T(0.0) : rho_w[p.ix(k - 1, i)];
            T adv = -((rho_u[p.ix(k, i)] - rho_u[p.ix(k, i - 1)]) / dx + (flux_kp - flux_km) / p.dz[k]);
            R_rho[p.ix(k, i)] = adv + d_rho_phys[p.ix(k, i)];
        }
    }

    // u at x-faces (moist.jl:1091-1097)
    std::vector<T> uf(N, T(0));
    for (int i = is; i <= ie; ++i)
        for (int k = ks; k <= ke; ++k)
            uf[p.ix(k, i)] = rho_u[p.ix(k, i)] * T(2.0) / (rho[p.ix(k, i)] + rho[p.ix(k, i + 1)]);
    exchange_halo(uf.data(), p);
    auto U = [&](int k, int i) { return uf[p.ix(k, i)]; };
    auto RHO = [&](int k, int i) { return rho[p.ix(k, i)]; };

    // R_rho_u (moist.jl:1098-1179)
#pragma omp parallel for schedule(static)
    for (int i = is; i <= ie; ++i) {
        for (int k = ks; k <= ke; ++k) {
            T rho_u_ip = T(0.5) * (rho_u[p.ix(k, i)] + rho_u[p.ix(k, i + 1)]);
            T fip, fim;
            if (rho_u_ip > c.EPS) fip = rho_u_ip * upw3(U(k, i - 1), U(k, i), U(k, i + 1)) / T(6.0);
            else if (rho_u_ip < -c.EPS) fip = rho_u_ip * upw3(U(k, i + 2), U(k, i + 1), U(k, i)) / T(6.0);
            else fip = rho_u_ip * (U(k, i) + U(k, i + 1)) * T(0.5);
            T rho_u_im = T(0.5) * (rho_u[p.ix(k, i - 1)] + rho_u[p.ix(k, i)]);
            if (rho_u_im > c.EPS) fim = rho_u_im * upw3(U(k, i - 2), U(k, i - 1), U(k, i)) / T(6.0);
            else if (rho_u_im < -c.EPS) fim = rho_u_im * upw3(U(k, i + 1), U(k, i), U(k, i - 1)) / T(6.0);
            else fim = rho_u_im * (U(k, i - 1) + U(k, i)) * T(0.5);
            T druudx = (fip - fim) / dx;

            T fkp, fkm;
            if (k == ke) fkp = T(0.0);
            else {
                T rwkp = T(0.5) * (rho_w[p.ix(k, i)] + rho_w[p.ix(k, i + 1)]);
                if (rwkp > c.EPS && k > ks) fkp = rwkp * upw3(U(k - 1, i), U(k, i), U(k + 1, i)) / T(6.0);
                else if (rwkp < -c.EPS && k < ke - 1) fkp = rwkp * upw3(U(k + 2, i), U(k + 1, i), U(k, i)) / T(6.0);
                else fkp = rwkp * f2h(U(k, i), U(k + 1, i), p.dz[k], p.dz[k + 1]);
            }
            if (k == ks) fkm = T(0.0);
            else {
                T rwkm = T(0.5) * (rho_w[p.ix(k - 1, i)] + rho_w[p.ix(k - 1, i + 1)]);
                if (rwkm > c.EPS && k > ks + 1) fkm = rwkm * upw3(U(k - 2, i), U(k - 1, i), U(k, i)) / T(6.0);
                else if (rwkm < -c.EPS && k < ke) fkm = rwkm * upw3(U(k + 1, i), U(k, i), U(k - 1, i)) / T(6.0);
                else fkm = rwkm * f2h(U(k - 1, i), U(k, i), p.dz[k - 1], p.dz[k]);
            }
            T drwudz = (fkp - fkm) / p.dz[k];
            T adv = -druudx - drwudz;

            T u_ip = (U(k, i + 1) - U(k, i)) / dx;
            T u_im = (U(k, i) - U(k, i - 1)) / dx;
            T u_xx = (u_ip * RHO(k, i + 1) - u_im * RHO(k, i)) / dx;
            T u_kp = k == ke ? T(0.0) : (U(k + 1, i) - U(k, i)) / (p.z[k + 1] - p.z[k]);
            T u_km = k == ks ? T(0.0) : (U(k, i) - U(k - 1, i)) / (p.z[k] - p.z[k - 1]);
            T rho_kp = k == ke ? T(0.0) : T(0.25) * (RHO(k, i) + RHO(k + 1, i) + RHO(k, i + 1) + RHO(k + 1, i + 1));
            T rho_km = k == ks ? T(0.0) : T(0.25) * (RHO(k - 1, i) + RHO(k, i) + RHO(k - 1, i + 1) + RHO(k, i + 1));
            T u_zz = (u_kp * rho_kp - u_km * rho_km) / p.dz[k];
            T diff = p.nu * (u_xx + u_zz);
            T dpdx = (pres[p.ix(k, i + 1)] - pres[p.ix(k, i)]) / dx;
            R_rho_u[p.ix(k, i)] = adv - dpdx + diff + d_rho_u_phys[p.ix(k, i)];
        }
    }

    // w at z-faces (moist.jl:1184-1189); row ke is never read
    std::vector<T> wf(N, T(0));
    for (int i = is - 2; i <= ie + 2; ++i)
        for (int k = ks; k <= ke - 1; ++k)
            wf[p.ix(k, i)] = rho_w[p.ix(k, i)] / f2h(RHO(k, i), RHO(k + 1, i), p.dz[k], p.dz[k + 1]);
    auto W = [&](int k, int i) { return wf[p.ix(k, i)]; };

    // R_rho_w (moist.jl:1190-1279)
    const T sponge_base = (T(1.0) - p.zf_sponge) * p.H;
#pragma omp parallel for schedule(static)
    for (int i = is; i <= ie; ++i) {
        for (int k = ks; k <= ke - 1; ++k) {
            T rho_u_ip = f2h(rho_u[p.ix(k, i)], rho_u[p.ix(k + 1, i)], p.z[k], p.z[k + 1]);
            T fip, fim;
            if (rho_u_ip > c.EPS) fip = rho_u_ip * upw3(W(k, i - 1), W(k, i), W(k, i + 1)) / T(6.0);
            else if (rho_u_ip < -c.EPS) fip = rho_u_ip * upw3(W(k, i + 2), W(k, i + 1), W(k, i)) / T(6.0);
            else fip = rho_u_ip * (W(k, i) + W(k, i + 1)) * T(0.5);
            T rho_u_im = f2h(rho_u[p.ix(k, i - 1)], rho_u[p.ix(k + 1, i - 1)], p.dz[k], p.dz[k + 1]);
            if (rho_u_im > c.EPS) fim = rho_u_im * upw3(W(k, i - 2), W(k, i - 1), W(k, i)) / T(6.0);
            else if (rho_u_im < -c.EPS) fim = rho_u_im * upw3(W(k, i + 1), W(k, i), W(k, i - 1)) / T(6.0);
            else fim = rho_u_im * (W(k, i - 1) + W(k, i)) * T(0.5);
            T drwudx = (fip - fim) / dx;

            T fkp, fkm;
            if (k == ke - 1) fkp = T(0.0);
            else {
                T rwkp = f2h(rho_w[p.ix(k, i)], rho_w[p.ix(k + 1, i)], p.dz[k], p.dz[k + 1]);
                if (rwkp > c.EPS && k > ks) fkp = rwkp * upw3(W(k - 1, i), W(k, i), W(k + 1, i)) / T(6.0);
                else if (rwkp < -c.EPS && k < ke - 2) fkp = rwkp * upw3(W(k + 2, i), W(k + 1, i), W(k, i)) / T(6.0);
                else fkp = rwkp * f2h(W(k, i), W(k + 1, i), p.dz[k], p.dz[k + 1]);
            }
            if (k == ks) fkm = T(0.0);
            else {
                T rwkm = f2h(rho_w[p.ix(k - 1, i)], rho_w[p.ix(k, i)], p.dz[k - 1], p.dz[k]);
                if (rwkm > c.EPS && k > ks + 1) fkm = rwkm * upw3(W(k - 2, i), W(k - 1, i), W(k, i)) / T(6.0);
                else if (rwkm < -c.EPS && k < ke - 1) fkm = rwkm * upw3(W(k + 1, i), W(k, i), W(k - 1, i)) / T(6.0);
                else fkm = rwkm * f2h(W(k - 1, i), W(k, i), p.dz[k - 1], p.dz[k]);
            }
            T zz = p.z[k + 1] - p.z[k];
            T drwwdz = (fkp - fkm) / zz;
            T adv = -drwudx - drwwdz;

            T w_ip = (W(k, i + 1) - W(k, i)) / dx;
            T w_im = (W(k, i) - W(k, i - 1)) / dx;
            T rho_ip = T(0.25) * (RHO(k, i) + RHO(k + 1, i) + RHO(k, i + 1) + RHO(k + 1, i + 1));
            T rho_im = T(0.25) * (RHO(k, i - 1) + RHO(k + 1, i - 1) + RHO(k, i) + RHO(k + 1, i));
            T w_xx = (w_ip * rho_ip - w_im * rho_im) / dx;
            T w_kp = k == ke - 1 ? T(0.0) : (W(k + 1, i) - W(k, i)) / p.dz[k + 1];
            T w_km = k == ks ? T(0.0) : (W(k, i) - W(k - 1, i)) / p.dz[k];
            T w_zz = (w_kp * RHO(k + 1, i) - w_km * RHO(k, i)) / zz;
            T diff = p.nu * (w_xx + w_zz);

            T sigma = T(0.0);
            if (p.zf[k] > sponge_base)
                sigma = (c.GRAV / T(30.0)) * ((p.zf[k] - sponge_base) / (p.zf_sponge * p.H));
            T sponge = -sigma * rho_w[p.ix(k, i)];
            T dpdz = (pres[p.ix(k + 1, i)] - pres[p.ix(k, i)]) / zz;
            T buoy = -c.GRAV * (RHO(k, i) + RHO(k + 1, i)) * T(0.5);
            R_rho_w[p.ix(k, i)] = adv - dpdz + buoy + diff + sponge;
        }
    }

    // R_rho_theta (moist.jl:1284-1332)
    auto TH = [&](int k, int i) { return theta[p.ix(k, i)]; };
#pragma omp parallel for schedule(static)
    for (int i = is; i <= ie; ++i) {
        for (int k = ks; k <= ke; ++k) {
            T ru = rho_u[p.ix(k, i)], rum = rho_u[p.ix(k, i - 1)];
            T fip, fim, fkp, fkm;
            if (ru > c.EPS) fip = ru * upw3(TH(k, i - 1), TH(k, i), TH(k, i + 1)) / T(6.0);
            else if (ru < -c.EPS) fip = ru * upw3(TH(k, i + 2), TH(k, i + 1), TH(k, i)) / T(6.0);
            else fip = ru * (TH(k, i) + TH(k, i + 1)) * T(0.5);
            if (rum > c.EPS) fim = rum * upw3(TH(k, i - 2), TH(k, i - 1), TH(k, i)) / T(6.0);
            else if (rum < -c.EPS) fim = rum * upw3(TH(k, i + 1), TH(k, i), TH(k, i - 1)) / T(6.0);
            else fim = rum * (TH(k, i - 1) + TH(k, i)) * T(0.5);
            if (k == ke) fkp = T(0.0);
            else {
                T rw = rho_w[p.ix(k, i)];
                if (rw > c.EPS && k > ks) fkp = rw * upw3(TH(k - 1, i), TH(k, i), TH(k + 1, i)) / T(6.0);
                else if (rw < -c.EPS && k < ke - 1) fkp = rw * upw3(TH(k + 2, i), TH(k + 1, i), TH(k, i)) / T(6.0);
                else fkp = rw * f2h(TH(k, i), TH(k + 1, i), p.dz[k], p.dz[k + 1]);
            }
            if (k == ks) fkm = T(0.0);
            else {
                T rw = rho_w[p.ix(k - 1, i)];
                if (rw > c.EPS && k > ks + 1) fkm = rw * upw3(TH(k - 2, i), TH(k - 1, i), TH(k, i)) / T(6.0);
                else if (rw < -c.EPS && k < ke) fkm = rw * upw3(TH(k + 1, i), TH(k, i), TH(k - 1, i)) / T(6.0);
                else fkm = rw * f2h(TH(k - 1, i), TH(k, i), p.dz[k - 1], p.dz[k]);
            }
            T adv = -((fip - fim) / dx + (fkp - fkm) / p.dz[k]);
            R_rho_theta[p.ix(k, i)] = adv + d_rho_theta_phys[p.ix(k, i)];
        }
    }

    // perturbations (moist.jl:1339-1348)
    std::vector<T> rp(N), rup(N), rwp(N), rtp(N);
    for (size_t q = 0; q < N; ++q) {
        rp[q] = s0.rho()[q] - rho[q];
        rup[q] = s0.rho_u()[q] - rho_u[q];
        rwp[q] = s0.rho_w()[q] - rho_w[q];
        rtp[q] = s0.rho_theta()[q] - rho_theta[q];
    }
    exchange_halo(rtp.data(), p);
    std::vector<T> rtp_old(rtp);
    std::vector<T> mom_x(rho_u, rho_u + N), mom_z(rho_w, rho_w + N);
    if (rec) {
        rec->uf = uf;
        rec->wf = wf;
        rec->R_rho = R_rho;
        rec->R_rho_u = R_rho_u;
        rec->R_rho_w = R_rho_w;
        rec->R_rho_theta = R_rho_theta;
    }

    // acoustic substeps (moist.jl:1351-1461)
    const T dt_s = dt / T(ns);
    const double fact_tp = double(T(0.5)) * (1.0 + double(p.beta));  // Float64 (moist.jl:1375)
    for (int n = 1; n <= ns; ++n) {
        if (n > 1) exchange_halo(rtp.data(), p);
        if (rec) {
            rec->rtp_start.push_back(rtp);
            rec->rwp_start.push_back(rwp);
        }
#pragma omp parallel for schedule(static)
        for (int i = is; i <= ie; ++i) {
            for (int k = ks; k <= ke; ++k) {
                T gx = (rtp[p.ix(k, i + 1)] - rtp[p.ix(k, i)]) / dx;
                T gxo = (rtp_old[p.ix(k, i + 1)] - rtp_old[p.ix(k, i)]) / dx;
                gx = gx + p.gamma_d * (gx - gxo);
                T rt2h = (rt2pres[p.ix(k, i)] + rt2pres[p.ix(k, i + 1)]) * T(0.5);
                rup[p.ix(k, i)] += dt_s * (R_rho_u[p.ix(k, i)] - rt2h * gx);
            }
        }
        exchange_halo(rup.data(), p);
        if (rec) rec->rup_new.push_back(rup);
        rtp_old = rtp;

#pragma omp parallel for schedule(static)
        for (int i = is; i <= ie; ++i) {
            std::vector<T> rp_tp(p.ka), rtp_tp(p.ka), drwp(p.ka, T(0)), colw(p.ka), colth(p.ka), colrt(p.ka);
            for (int k = ks; k <= ke; ++k) {
                size_t q = p.ix(k, i);
                T adv = -(rup[q] - rup[p.ix(k, i - 1)]) / dx;
                T inc = dt_s * (R_rho[q] + adv);
                rp_tp[k] = off_centre(rp[q], fact_tp, inc);
                rp[q] += inc;
                T fip = (TH(k, i + 1) + TH(k, i)) * T(0.5) * rup[q];
                T fim = (TH(k, i) + TH(k, i - 1)) * T(0.5) * rup[p.ix(k, i - 1)];
                adv = -(fip - fim) / dx;
                inc = dt_s * (R_rho_theta[q] + adv);
                rtp_tp[k] = off_centre(rtp[q], fact_tp, inc);
                rtp[q] += inc;
            }
            for (int k = ks; k <= ke - 1; ++k) {
                T buoy = (rtp_tp[k + 1] - rtp_tp[k]) / (p.z[k + 1] - p.z[k]);
                T grav = -c.GRAV * T(0.5) * (rp_tp[k + 1] + rp_tp[k]);
                T rt2h = (rt2pres[p.ix(k, i)] + rt2pres[p.ix(k + 1, i)]) * T(0.5);
                drwp[k] = dt_s * (R_rho_w[p.ix(k, i)] - rt2h * buoy + grav);
            }
            for (int k = 0; k < p.ka; ++k) {
                colw[k] = rwp[p.ix(k, i)];
                colth[k] = TH(k, i);
                colrt[k] = rt2pres[p.ix(k, i)];
            }
            implicit_correction(p, colw.data(), drwp.data(), colth.data(), colrt.data(), dt_s);
            for (int k = 0; k < p.ka; ++k) rwp[p.ix(k, i)] = colw[k];
            for (int k = ks; k <= ke; ++k) {
                T fkp = k == ke ? T(0.0) : rwp[p.ix(k, i)];
                T fkm = k == ks ? T(0.0) : rwp[p.ix(k - 1, i)];
                T adv = -(fkp - fkm) / p.dz[k];
                rp[p.ix(k, i)] += dt_s * adv;
            }
            for (int k = ks; k <= ke; ++k) {
                T fkp, fkm;
                if (k == ke) fkp = T(0.0);
                else {
                    T w = rwp[p.ix(k, i)];
                    if (w > c.EPS && k > ks) fkp = w * upw3(TH(k - 1, i), TH(k, i), TH(k + 1, i)) / T(6.0);
                    else if (rho_w[p.ix(k, i)] < -c.EPS && k < ke - 1) fkp = w * upw3(TH(k + 2, i), TH(k + 1, i), TH(k, i)) / T(6.0);
                    else fkp = w * f2h(TH(k, i), TH(k + 1, i), p.dz[k], p.dz[k + 1]);
                }
                if (k == ks) fkm = T(0.0);
                else {
                    T w = rwp[p.ix(k - 1, i)];
                    T sw = rho_w[p.ix(k - 1, i)];
                    if (sw > c.EPS && k > ks + 1) fkm = w * upw3(TH(k - 2, i), TH(k - 1, i), TH(k, i)) / T(6.0);
                    else if (sw < -c.EPS && k < ke) fkm = w * upw3(TH(k + 1, i), TH(k, i), TH(k - 1, i)) / T(6.0);
                    else fkm = w * f2h(TH(k - 1, i), TH(k, i), p.dz[k - 1], p.dz[k]);
                }
                T adv = -(fkp - fkm) / p.dz[k];
                rtp[p.ix(k, i)] += dt_s * adv;
            }
        }
#pragma omp parallel for schedule(static)
        for (int i = is; i <= ie; ++i) {
            for (int k = ks; k <= ke; ++k) mom_x[p.ix(k, i)] += rup[p.ix(k, i)] / T(ns);
            for (int k = ks; k <= ke - 1; ++k) mom_z[p.ix(k, i)] += rwp[p.ix(k, i)] / T(ns);
        }
        if (rec) rec->rwp_new.push_back(rwp);
    }

    exchange_halo(mom_x.data(), p);
    if (rec) {
        rec->mom_x = mom_x;
        rec->mom_z = mom_z;
    }
    exchange_halo(s.rho_qv(), p);
    exchange_halo(s.rho_qc(), p);
    exchange_halo(s.rho_qr(), p);

    // scalar transport (moist.jl:1469-1518)
    T *scal[3] = {s.rho_qv(), s.rho_qc(), s.rho_qr()};
    const T *scal0[3] = {s0.rho_qv(), s0.rho_qc(), s0.rho_qr()};
    const T *dph[3] = {d_qv_phys, d_qc_phys, d_qr_phys};
    std::vector<T> snap[3];
    for (int sc = 0; sc < 3; ++sc) snap[sc].assign(scal[sc], scal[sc] + N);
    auto cell = [&](int sc, int k, int i) {
        // read source: in place (serial) or the stage-input snapshot
        const T *src = order == ORDER_SERIAL ? scal[sc] : snap[sc].data();
        const T *lin = snap[sc].data();  // column 0 is a halo column: never written here
        T fxp = flux_x(p, src, rho, mom_x.data(), k, i);
        T fxm = flux_x(p, src, rho, mom_x.data(), k, i - 1);
        T adv_x = (fxp - fxm) / dx;
        T fzp = k == ke ? T(0.0) : flux_z(p, src, lin, rho, mom_z.data(), k, i);
        T fzm = k == ks ? T(0.0) : flux_z(p, src, lin, rho, mom_z.data(), k - 1, i);
        T adv_z = (fzp - fzm) / p.dz[k];
        T adv = -(adv_x + adv_z);
        T sc_c = src[p.ix(k, i)] / RHO(k, i);
        T sc_ip = src[p.ix(k, i + 1)] / RHO(k, i + 1);
        T sc_im = src[p.ix(k, i - 1)] / RHO(k, i - 1);
        T sc_kp = k == ke ? T(0.0) : src[p.ix(k + 1, i)] / RHO(k + 1, i);
        T sc_km = k == ks ? T(0.0) : src[p.ix(k - 1, i)] / RHO(k - 1, i);
        T dip = (RHO(k, i + 1) + RHO(k, i)) * T(0.5) * (sc_ip - sc_c) / dx;
        T dim = (RHO(k, i) + RHO(k, i - 1)) * T(0.5) * (sc_c - sc_im) / dx;
        T diff_x = p.kappa * (dip - dim) / dx;
        T dkp = k == ke ? T(0.0)
                        : f2h(RHO(k + 1, i), RHO(k, i), p.dz[k + 1], p.dz[k]) * (sc_kp - sc_c) / (p.z[k + 1] - p.z[k]);
        T dkm = k == ks ? T(0.0)
                        : f2h(RHO(k, i), RHO(k - 1, i), p.dz[k], p.dz[k - 1]) * (sc_c - sc_km) / (p.z[k] - p.z[k - 1]);
        T diff_z = p.kappa * (dkp - dkm) / p.dz[k];
        T diffu = diff_x + diff_z;
        scal[sc][p.ix(k, i)] = scal0[sc][p.ix(k, i)] + dt * (adv + diffu + dph[sc][p.ix(k, i)]);
    };
    if (order == ORDER_SERIAL) {
        for (int i = is; i <= ie; ++i)
            for (int k = ks; k <= ke; ++k)
                for (int sc = 0; sc < 3; ++sc) cell(sc, k, i);
    } else {
#pragma omp parallel for schedule(static)
        for (int i = is; i <= ie; ++i)
            for (int k = ks; k <= ke; ++k)
                for (int sc = 0; sc < 3; ++sc) cell(sc, k, i);
    }

    // final update (moist.jl:1522-1539)
#pragma omp parallel for schedule(static)
    for (int i = is; i <= ie; ++i) {
        for (int k = ks; k <= ke; ++k) {
            size_t q = p.ix(k, i);
            rho[q] = rho[q] + rp[q];
            rho_u[q] = rho_u[q] + rup[q];
            rho_theta[q] = rho_theta[q] + rtp[q];
        }
        for (int k = ks; k <= ke - 1; ++k) rho_w[p.ix(k, i)] = rho_w[p.ix(k, i)] + rwp[p.ix(k, i)];
    }

    // NaN checks (moist.jl:1543-1578)
    const T *chk[12] = {rho, rho_u, rho_w, rho_theta, s.rho_qv(), s.rho_qc(), s.rho_qr(),
                        d_rho_u_phys, d_rho_theta_phys, d_qv_phys, d_qc_phys, d_qr_phys};
    for (int f = 0; f < 12; ++f)
        for (int i = is; i <= ie; ++i)
            for (int k = ks; k <= ke; ++k)
                if (std::isnan(chk[f][p.ix(k, i)])) return f + 1;
    return 0;
}

// ---- cloud_microphysics! (moist.jl:554-671) -------------------------------
template <class T>
static void cloud_microphysics(const G<T> &p, T *d_qv, T *d_qc, T *d_qr, T *d_th, const T *rho,
                               const T *rho_theta, const T *rho_qv, const T *rho_qc, const T *rho_qr) {
    const C<T> &c = p.c;
#pragma omp parallel for schedule(static)
    for (int i = p.is; i <= p.ie; ++i) {
        for (int k = p.ks; k <= p.ke; ++k) {
            size_t q = p.ix(k, i);
            d_qv[q] = T(0.0); d_qc[q] = T(0.0); d_qr[q] = T(0.0); d_th[q] = T(0.0);
            T rho_k = rho[q];
            T rinv = T(1.0) / rho_k;
            T theta_k = rho_theta[q] * rinv;
            T qv = rho_qv[q] * rinv, qc = rho_qc[q] * rinv, qr = rho_qr[q] * rinv;
            T qd = T(1.0) - qv - qc - qr;
            T r = c.Rd * qd + c.Rv * qv;
            T cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
            T cv = cp - r;
            T Pi = std::pow(r * rho_theta[q] / c.P0, r / cv);
            T Tk = theta_k * Pi;
            T pk = c.P0 * std::pow(Pi, cp / r);
            T lhe = c.Lv / (cp * Pi);
            T qv_init = qv;
            T ce = T(0), qsat = T(0);
            for (int it = 0; it < 3; ++it) {
                T es, des, dqs_des;
                sat_vp(c, Tk, es, des);
                sat_q(c, pk, es, qsat, dqs_des);
                T dqs_dT = dqs_des * des;
                ce = (qv - qsat) / (T(1.0) + (c.Lv / cp) * dqs_dT);
                ce = jmin(ce, qv);
                ce = jmax(ce, -qc);
                qv -= ce;
                qc += ce;
                theta_k += ce * lhe;
                r = c.Rd * qd + c.Rv * qv;
                cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
                cv = cp - r;
                Pi = std::pow(r * rho_k * theta_k / c.P0, r / cv);
                Tk = theta_k * Pi;
                pk = c.P0 * std::pow(Pi, cp / r);
                lhe = c.Lv / (cp * Pi);
            }
            ce = qv_init - qv;
            T rate = ce * rho_k / p.dt;
            d_qv[q] -= rate;
            d_qc[q] += rate;
            d_th[q] += rate * lhe;
            if (p.do_auto) {
                T ra = c.K_auto * jmax(T(0.0), qc - c.Qc_threshold);
                ra = jmin(ra, qc / p.dt);
                qc -= ra * p.dt;
                qr += ra * p.dt;
                ra *= rho_k;
                d_qc[q] -= ra;
                d_qr[q] += ra;
            }
            if (p.do_accr) {
                T rc = T(0.0);
                if (qr > T(1e-12) && qc > T(1e-12)) {
                    rc = c.K_accr * qc * qr;
                    rc = jmin(rc, qc / p.dt);
                }
                qc -= rc * p.dt;
                qr += rc * p.dt;
                rc *= rho_k;
                d_qc[q] -= rc;
                d_qr[q] += rc;
            }
            if (p.do_revap) {
                T re = T(0.0);
                if (qr > T(1e-12) && qv < qsat) {
                    re = c.K_evap * (qsat - qv) * std::sqrt(rho_k * qr);
                    re = jmin(re, qr) / p.dt;
                }
                re *= rho_k;
                d_qr[q] -= re;
                d_qv[q] += re;
                r = c.Rd * qd + c.Rv * qv;
                cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
                cv = cp - r;
                Pi = std::pow(r * rho[q] * theta_k / c.P0, r / cv);
                lhe = c.Lv / (cp * Pi);
                d_th[q] -= re * lhe;
            }
        }
    }
}

// ---- rain_sedimentation! (moist.jl:674-707) -------------------------------
template <class T>
static void rain_sedimentation(const G<T> &p, T *d_sed, T *d_prec, const T *rho_qr, const T *rho) {
    const C<T> &c = p.c;
    std::vector<T> Vt(p.n2(), T(0));
#pragma omp parallel for schedule(static)
    for (int i = p.is; i <= p.ie; ++i) {
        for (int k = p.ks; k <= p.ke; ++k) {
            size_t q = p.ix(k, i);
            T eff = jmax(T(0.0), rho_qr[q]);
            T v = c.A_term * std::pow(eff / rho[q], c.B_term);
            Vt[q] = eff > T(1e-12) ? v : T(0.0);
        }
        for (int k = p.ks; k <= p.ke; ++k) {
            T fkp = k == p.ke ? T(0.0) : rho_qr[p.ix(k + 1, i)] * Vt[p.ix(k + 1, i)];
            T fkm = rho_qr[p.ix(k, i)] * Vt[p.ix(k, i)];
            d_sed[p.ix(k, i)] = (fkp - fkm) / p.dz[k];
        }
        d_prec[i] = rho_qr[p.ix(p.ks, i)] * Vt[p.ix(p.ks, i)];
    }
}

// ---- surface_fluxes! (moist.jl:710-752) -----------------------------------
template <class T>
static void surface_fluxes(const G<T> &p, T *d_u_sf, T *d_th_sf, T *d_qv_sf, T *mflux, T *shflux,
                           T *lhflux, const T *rho, const T *rho_u, const T *rho_theta,
                           const T *rho_qv, const T *rho_qc, const T *rho_qr) {
    const C<T> &c = p.c;
    const int ks = p.ks;
    for (int i = p.is; i <= p.ie; ++i) {
        size_t q = p.ix(ks, i);
        T qv = rho_qv[q] / rho[q], qc = rho_qc[q] / rho[q], qr = rho_qr[q] / rho[q];
        T qd = T(1.0) - qv - qc - qr;
        T r = c.Rd * qd + c.Rv * qv;
        T cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
        T cv = cp - r;
        T uabs = std::fabs(rho_u[p.ix(ks, i - 1)] + rho_u[q]) * T(0.5) / rho[q];
        uabs = jmax(uabs, p.uabs_min);
        T theta_ks = rho_theta[q] / rho[q];
        T pi_ks = std::pow(r * rho_theta[q] / c.P0, r / cv);
        T pi_surf = pi_ks + p.dz[ks] * T(0.5) * c.GRAV / (cp * theta_ks);
        T p_surf = c.P0 * std::pow(pi_surf, cp / r);
        T rho_surf = p_surf / (r * p.temp_surf);
        T theta_surf = p.temp_surf / pi_surf;
        T u = rho_u[q] / (rho[q] + rho[p.ix(ks, i + 1)]) * T(2.0);
        mflux[i] = -rho_surf * p.Cd * u * std::fabs(u);
        d_u_sf[i] = mflux[i] / p.dz[ks];
        T sh = rho_surf * p.Ch * uabs * (theta_surf - theta_ks) * pi_surf;
        d_th_sf[i] = sh / (pi_ks * p.dz[ks]);
        shflux[i] = sh * cp;
        T es, des, qsat_s, dq;
        sat_vp(c, p.temp_surf, es, des);
        sat_q(c, p_surf, es, qsat_s, dq);
        T lh = rho_surf * p.Ce * uabs * (qsat_s - qv);
        d_qv_sf[i] = lh / p.dz[ks];
        lhflux[i] = lh * c.Lv;
    }
}

// ---- gray_radiation! (moist.jl:776-840) -----------------------------------
template <class T>
static void gray_radiation(const G<T> &p, T *d_rd, const T *rho, const T *rho_theta, const T *rho_qv,
                           const T *rho_qc, const T *rho_qr, const T *theta_ref) {
    const C<T> &c = p.c;
    const int ks = p.ks, ke = p.ke;
    std::fill(d_rd, d_rd + p.n2(), T(0.0));
#pragma omp parallel for schedule(static)
    for (int i = p.is; i <= p.ie; ++i) {
        std::vector<T> up(p.ka + 1), dn(p.ka + 1), eps(p.ka), tau(p.ka), temp(p.ka), pex(p.ka), cpv(p.ka);
        for (int k = ks; k <= ke; ++k) {
            size_t q = p.ix(k, i);
            T qv = rho_qv[q] / rho[q], qc = rho_qc[q] / rho[q], qr = rho_qr[q] / rho[q];
            T qd = T(1.0) - qv - qc - qr;
            T r = c.Rd * qd + c.Rv * qv;
            cpv[k] = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
            T cv = cpv[k] - r;
            pex[k] = std::pow(r * rho_theta[q] / c.P0, r / cv);
            temp[k] = rho_theta[q] / rho[q] * pex[k];
            T dq = p.k_abs_qv * rho_qv[q] * p.dz[k];
            T dc = p.k_abs_qc * rho_qc[q] * p.dz[k];
            T dr = p.k_abs_qr * rho_qr[q] * p.dz[k];
            T dopt = dq + dc + dr;
            eps[k] = T(1.0) - std::exp(-dopt);
            tau[k] = T(1.0) - eps[k];
        }
        T ts = p.temp_surf;
        up[ks] = p.surf_emissivity * c.SIGMA * (ts * ts * ts * ts);
        for (int k = ks; k <= ke; ++k) {
            T t = temp[k];
            up[k + 1] = up[k] * tau[k] + eps[k] * c.SIGMA * (t * t * t * t);
        }
        dn[ke + 1] = T(0.0);
        for (int k = ke; k >= ks; --k) {
            T t = temp[k];
            dn[k] = dn[k + 1] * tau[k] + eps[k] * c.SIGMA * (t * t * t * t);
        }
        const T dz_trans = T(2.0e3);
        for (int k = ks; k <= ke; ++k) {
            size_t q = p.ix(k, i);
            T divF = ((up[k + 1] - dn[k + 1]) - (up[k] - dn[k])) / p.dz[k];
            T hr = -divF / (cpv[k] * pex[k]);
            T alpha = T(0.5) * (T(1.0) + std::tanh((p.z[k] - (p.z_tropopose + dz_trans * T(0.5))) / dz_trans)) / p.tau_cooling;
            T cr = -(rho_theta[q] - rho[q] * theta_ref[k]) * alpha;
            d_rd[q] = hr + cr;
        }
    }
}

// State in the reference's field order (moist.jl:295-316).
enum {
    F_RHO, F_RHO_U, F_RHO_W, F_RHO_THETA, F_QV, F_QC, F_QR, F_PREC, F_MFLUX, F_SHFLUX, F_LHFLUX,
    F_THETA_REF, F_D_TH_MP, F_D_QV_MP, F_D_QC_MP, F_D_QR_MP, F_D_QR_SED, F_D_U_SF, F_D_TH_SF,
    F_D_QV_SF, F_D_TH_RD, F_COUNT
};

// ---- rk3_step! (moist.jl:1589-1683) ----------------------------------------
template <class T> static int rk3_step(const G<T> &p, T *const *F, int order) {
    const size_t N = p.n2();
    St<T> s;
    for (int f = 0; f < 7; ++f) s.f[f] = F[f];
    if (p.do_mp)
        cloud_microphysics(p, F[F_D_QV_MP], F[F_D_QC_MP], F[F_D_QR_MP], F[F_D_TH_MP], s.rho(),
                           s.rho_theta(), s.rho_qv(), s.rho_qc(), s.rho_qr());
    else
        for (int f : {F_D_QV_MP, F_D_QC_MP, F_D_QR_MP, F_D_TH_MP}) std::fill(F[f], F[f] + N, T(0));
    std::vector<T> d_prec(p.ia, T(0));
    if (p.do_mp && p.do_sed)
        rain_sedimentation(p, F[F_D_QR_SED], d_prec.data(), s.rho_qr(), s.rho());
    else
        std::fill(F[F_D_QR_SED], F[F_D_QR_SED] + N, T(0));
    if (p.do_sf)
        surface_fluxes(p, F[F_D_U_SF], F[F_D_TH_SF], F[F_D_QV_SF], F[F_MFLUX], F[F_SHFLUX], F[F_LHFLUX],
                       s.rho(), s.rho_u(), s.rho_theta(), s.rho_qv(), s.rho_qc(), s.rho_qr());
    else
        for (int f : {F_D_U_SF, F_D_TH_SF, F_D_QV_SF, F_MFLUX, F_SHFLUX, F_LHFLUX})
            std::fill(F[f], F[f] + p.ia, T(0));
    if (p.do_rd)
        gray_radiation(p, F[F_D_TH_RD], s.rho(), s.rho_theta(), s.rho_qv(), s.rho_qc(), s.rho_qr(),
                       F[F_THETA_REF]);
    else
        std::fill(F[F_D_TH_RD], F[F_D_TH_RD] + N, T(0));

    // physical source terms (moist.jl:1632-1644)
    std::vector<T> d_rho(F[F_D_QR_SED], F[F_D_QR_SED] + N), d_u(N, T(0)), d_th(N), d_qv(F[F_D_QV_MP], F[F_D_QV_MP] + N),
        d_qr(N);
    for (int i = 0; i < p.ia; ++i) d_rho[p.ix(p.ks, i)] += F[F_D_QV_SF][i];
    for (int i = 0; i < p.ia; ++i) d_u[p.ix(p.ks, i)] += F[F_D_U_SF][i];
    for (size_t q = 0; q < N; ++q)
        d_th[q] = F[F_D_TH_MP][q] + F[F_D_TH_RD][q] + d_rho[q] * s.rho_theta()[q] / s.rho()[q];
    for (int i = 0; i < p.ia; ++i) d_th[p.ix(p.ks, i)] += F[F_D_TH_SF][i];
    for (int i = 0; i < p.ia; ++i) d_qv[p.ix(p.ks, i)] += F[F_D_QV_SF][i];
    const T *d_qc = F[F_D_QC_MP];
    for (size_t q = 0; q < N; ++q) d_qr[q] = F[F_D_QR_MP][q] + F[F_D_QR_SED][q];

    std::vector<T> s0buf(7 * N);
    St<T> s0;
    for (int f = 0; f < 7; ++f) {
        s0.f[f] = s0buf.data() + f * N;
        std::memcpy(s0.f[f], s.f[f], sizeof(T) * N);
    }
    const T dt = p.dt;
    int rc;
    rc = compute_step(p, d_rho.data(), d_u.data(), d_th.data(), d_qv.data(), d_qc, d_qr.data(), s, s0,
                      dt / T(3), std::max(1, (p.ns + 2) / 3), order);
    if (rc) return rc;
    rc = compute_step(p, d_rho.data(), d_u.data(), d_th.data(), d_qv.data(), d_qc, d_qr.data(), s, s0,
                      dt / T(2), std::max(1, (p.ns + 1) / 2), order);
    if (rc) return rc;
    rc = compute_step(p, d_rho.data(), d_u.data(), d_th.data(), d_qv.data(), d_qc, d_qr.data(), s, s0, dt,
                      p.ns, order);
    if (rc) return rc;
    for (int i = p.is; i <= p.ie; ++i) F[F_PREC][i] += d_prec[i] * dt;
    return 0;
}

#include "moist_adjoint.inc"

// ---- output_history's derived fields and budget (moist.jl:1816-1858) -------
// out[13]: nx*nz arrays, x fastest; stats[7]: Etot sum, w min/max, qc/qr max
// (g/kg), RH min/max.  rho_u's halo is exchanged first, as the reference does.
template <class T>
static void output_history(const G<T> &p, T *const *F, T *const *out, double *stats) {
    const C<T> &c = p.c;
    T *rho = F[0], *rho_u = F[1], *rho_w = F[2], *rho_theta = F[3];
    exchange_halo(rho_u, p);
    const int nx = p.nx, nz = p.nz;
    double ek = 0.0, epot = 0.0, ei = 0.0;
    double wmin = 1e300, wmax = -1e300, qcm = -1e300, qrm = -1e300, rhmin = 1e300, rhmax = -1e300;
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
            const size_t q = p.ix(k, p.is + i);
            const T rinv = T(1.0) / rho[q];
            const T th = rho_theta[q] * rinv;
            const T qv = F[4][q] * rinv, qc = F[5][q] * rinv, qr = F[6][q] * rinv;
            const T u = (rho_u[q] + rho_u[p.ix(k, p.is + i - 1)]) * T(0.5) * rinv;
            const T w = k > 0 ? (rho_w[q] + rho_w[p.ix(k - 1, p.is + i)]) * T(0.5) * rinv : rho_w[q] * T(0.5) * rinv;
            const T qt = qv + qc + qr, qd = T(1.0) - qt;
            const T r = c.Rd * qd + c.Rv * qv;
            T cp = c.Cpd * qd + c.Cpv * qv + c.Cpl * (qc + qr);
            const T cv = cp - r;
            const T pi = std::pow(r * rho_theta[q] / c.P0, r / cv);
            const T pres = c.P0 * std::pow(pi, cp / r);
            const T temp = th * pi;
            T es, des, qs, dqs;
            sat_vp(c, temp, es, des);
            sat_q(c, pres, es, qs, dqs);
            const T rh = qs > T(1e-10) ? qv / qs * T(100.0) : T(0.0);
            const T presv = rho[q] * qv * c.Rv * temp;
            const T presd = pres - presv;
            cp = c.Cpd * qd + c.Cpl * qt;
            const T the = temp * std::pow(presd / c.P0, -c.Rd / cp) * std::exp(c.Lv * qv / (cp * temp));
            ek += double(T(0.5) * rho[q] * (u * u + w * w));
            epot += double(rho[q] * c.GRAV * p.z[k]);
            ei += double(rho[q] * (cv * temp + qv * c.Lv));
            wmin = std::min(wmin, double(w));
            wmax = std::max(wmax, double(w));
            qcm = std::max(qcm, double(qc) * 1000.0);
            qrm = std::max(qrm, double(qr) * 1000.0);
            rhmin = std::min(rhmin, double(rh));
            rhmax = std::max(rhmax, double(rh));
            const T v[13] = {rho[q], rho_u[q], rho_w[q], u, w, th, qv, qc, qr, rh, temp, pres / T(100.0), the};
            for (int f = 0; f < 13; ++f) out[f][(size_t)k * nx + i] = v[f];
        }
    stats[0] = ek + epot + ei;
    stats[1] = wmin;
    stats[2] = wmax;
    stats[3] = qcm;
    stats[4] = qrm;
    stats[5] = rhmin;
    stats[6] = rhmax;
}

}  // namespace orc

// ============================================================================
// C entry points (ctypes).  Suffix _f64 / _f32 selects T.  Arrays are host
// arrays in the reference layout.
// ============================================================================
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

#define ORC_BOTH(X) X(double, f64) X(float, f32)

#define DEF_SCALARS(T, S)                                                                       \
    ORC_EXPORT T orc_f2h_##S(T a, T b, T c, T d) { return orc::f2h<T>(a, b, c, d); }           \
    ORC_EXPORT T orc_koren_##S(T r) { return orc::koren<T>(r); }                               \
    ORC_EXPORT void orc_sat_vp_##S(T Tk, T *es, T *des) {                                       \
        orc::C<T> c(nullptr);                                                                   \
        orc::sat_vp(c, Tk, *es, *des);                                                          \
    }                                                                                           \
    ORC_EXPORT void orc_sat_q_##S(T p, T es, T *qs, T *dq) {                                   \
        orc::C<T> c(nullptr);                                                                   \
        orc::sat_q(c, p, es, *qs, *dq);                                                         \
    }                                                                                           \
    ORC_EXPORT void orc_thomas_##S(int n, const T *a, const T *b, T *c, T *d) {                \
        orc::thomas<T>(n, a, b, c, d);                                                          \
    }                                                                                           \
    ORC_EXPORT void orc_implicit_correction_##S(const moist_params *mp, T *rho_w,               \
                                                const T *drho_w, const T *theta,                \
                                                const T *rt2pres, T dt) {                       \
        orc::G<T> g(mp);                                                                        \
        orc::implicit_correction(g, rho_w, drho_w, theta, rt2pres, dt);                         \
    }                                                                                           \
    ORC_EXPORT T orc_flux_x_##S(const moist_params *mp, const T *sr, const T *rho,             \
                                const T *rho_u, int k, int i) {                                 \
        orc::G<T> g(mp);                                                                        \
        return orc::flux_x(g, sr, rho, rho_u, k, i);                                            \
    }                                                                                           \
    ORC_EXPORT T orc_flux_z_##S(const moist_params *mp, const T *sr, const T *rho,             \
                                const T *rho_w, int k, int i) {                                 \
        orc::G<T> g(mp);                                                                        \
        return orc::flux_z(g, sr, sr, rho, rho_w, k, i);                                        \
    }                                                                                           \
    ORC_EXPORT void orc_exchange_halo_##S(const moist_params *mp, T *A) {                      \
        orc::G<T> g(mp);                                                                        \
        orc::exchange_halo(A, g);                                                               \
    }                                                                                           \
    ORC_EXPORT void orc_cloud_microphysics_##S(const moist_params *mp, T *d_qv, T *d_qc,        \
                                               T *d_qr, T *d_th, const T *rho, const T *rt,     \
                                               const T *qv, const T *qc, const T *qr) {         \
        orc::G<T> g(mp);                                                                        \
        orc::cloud_microphysics(g, d_qv, d_qc, d_qr, d_th, rho, rt, qv, qc, qr);                \
    }                                                                                           \
    ORC_EXPORT void orc_rain_sedimentation_##S(const moist_params *mp, T *d_sed, T *d_prec,     \
                                               const T *rho_qr, const T *rho) {                 \
        orc::G<T> g(mp);                                                                        \
        orc::rain_sedimentation(g, d_sed, d_prec, rho_qr, rho);                                 \
    }                                                                                           \
    ORC_EXPORT void orc_surface_fluxes_##S(const moist_params *mp, T *const *o, const T *const *s) { \
        orc::G<T> g(mp);                                                                        \
        orc::surface_fluxes(g, o[0], o[1], o[2], o[3], o[4], o[5], s[0], s[1], s[3], s[4], s[5], \
                            s[6]);                                                              \
    }                                                                                           \
    ORC_EXPORT void orc_gray_radiation_##S(const moist_params *mp, T *d_rd, const T *const *s,  \
                                           const T *theta_ref) {                                \
        orc::G<T> g(mp);                                                                        \
        orc::gray_radiation(g, d_rd, s[0], s[3], s[4], s[5], s[6], theta_ref);                   \
    }                                                                                           \
    ORC_EXPORT int orc_compute_step_##S(const moist_params *mp, const T *const *phys,           \
                                        T *const *s, const T *const *s0, T dt, int ns,          \
                                        int order) {                                            \
        orc::G<T> g(mp);                                                                        \
        orc::St<T> a, b;                                                                        \
        for (int f = 0; f < 7; ++f) {                                                           \
            a.f[f] = s[f];                                                                      \
            b.f[f] = const_cast<T *>(s0[f]);                                                    \
        }                                                                                       \
        return orc::compute_step(g, phys[0], phys[1], phys[2], phys[3], phys[4], phys[5], a, b, \
                                 dt, ns, order);                                                \
    }                                                                                           \
    ORC_EXPORT int orc_rk3_step_##S(const moist_params *mp, T *const *fields, int nsteps,       \
                                    int order) {                                                \
        orc::G<T> g(mp);                                                                        \
        for (int n = 0; n < nsteps; ++n) {                                                      \
            int rc = orc::rk3_step(g, fields, order);                                           \
            if (rc) return rc;                                                                  \
        }                                                                                       \
        return 0;                                                                               \
    }

ORC_BOTH(DEF_SCALARS)

// ---- adjoint (FP64 only) ------------------------------------------------------
// One compute_step!: forward with record on copies of s, then the adjoint.
ORC_EXPORT int orc_compute_step_adj_f64(const moist_params *mp, const double *const *phys, const double *const *s,
                                        const double *const *s0, double dt, int ns, const double *const *lo,
                                        double *const *ls, double *const *ls0, double *const *lph) {
    orc::G<double> g(mp);
    const size_t N = g.n2();
    orc::StBuf a, b;
    for (int f = 0; f < 7; ++f) {
        a.f[f].assign(s[f], s[f] + N);
        b.f[f].assign(s0[f], s0[f] + N);
    }
    orc::StBuf in = a;
    orc::St<double> av = a.view(), bv = b.view();
    orc::Rec<double> rec;
    int rc = orc::compute_step(g, phys[0], phys[1], phys[2], phys[3], phys[4], phys[5], av, bv, dt, ns,
                               orc::ORDER_SNAPSHOT, &rec);
    if (rc) return rc;
    orc::St<double> iv = in.view();
    orc::compute_step_adj(g, iv, bv, dt, ns, rec, lo, ls, ls0, lph);
    return 0;
}

// The slow tendencies of compute_step! (moist.jl:1071-1332) at (phys, s):
// out[0..3] = R_rho, R_rho_u, R_rho_w, R_rho_theta, out[4..5] = u and w at
// the faces.  One compute_step! with ns = 1 runs on copies of s and its
// record is returned, so the tendencies are exactly the step's own.
ORC_EXPORT int orc_slow_tendencies_f64(const moist_params *mp, const double *const *phys, const double *const *s,
                                       double *const *out) {
    orc::G<double> g(mp);
    const size_t N = g.n2();
    orc::StBuf a, b;
    for (int f = 0; f < 7; ++f) {
        a.f[f].assign(s[f], s[f] + N);
        b.f[f].assign(s[f], s[f] + N);
    }
    orc::St<double> av = a.view(), bv = b.view();
    orc::Rec<double> rec;
    int rc = orc::compute_step(g, phys[0], phys[1], phys[2], phys[3], phys[4], phys[5], av, bv, g.dt, 1,
                               orc::ORDER_SNAPSHOT, &rec);
    const std::vector<double> *src[6] = {&rec.R_rho, &rec.R_rho_u, &rec.R_rho_w, &rec.R_rho_theta, &rec.uf, &rec.wf};
    for (int f = 0; f < 6; ++f) std::copy(src[f]->begin(), src[f]->end(), out[f]);
    return rc;
}

ORC_EXPORT void orc_physics_fwd_f64(const moist_params *mp, const double *const *s, double *const *ph,
                                    double *dprec, const double *theta_ref) {
    orc::G<double> g(mp);
    orc::St<double> sv;
    for (int f = 0; f < 7; ++f) sv.f[f] = const_cast<double *>(s[f]);
    std::vector<double> v[6], dp;
    orc::physics_fwd(g, sv, theta_ref, v, dp);
    for (int f = 0; f < 6; ++f) std::copy(v[f].begin(), v[f].end(), ph[f]);
    std::copy(dp.begin(), dp.end(), dprec);
}

ORC_EXPORT void orc_physics_adj_f64(const moist_params *mp, const double *const *s, const double *const *lph,
                                    const double *ldprec, double *const *ls, double *lK, const double *theta_ref,
                                    double *stale) {
    orc::G<double> g(mp);
    orc::St<double> sv;
    for (int f = 0; f < 7; ++f) sv.f[f] = const_cast<double *>(s[f]);
    orc::physics_adj(g, sv, theta_ref, lph, ldprec, ls, lK, stale);
}

// s: the step input with its halo columns as the forward step sees them
// (with surface fluxes on, two of them are read: they hold the previous
// step's stage-3 input).  stale[2]: in/out, see rk3_step_adj.
ORC_EXPORT int orc_rk3_step_adj_f64(const moist_params *mp, const double *const *s, double *const *lam,
                                    double ldprec, double *lK, const double *theta_ref, double *stale) {
    orc::G<double> g(mp);
    orc::StBuf a;
    for (int f = 0; f < 7; ++f) a.f[f].assign(s[f], s[f] + g.n2());
    orc::St<double> av = a.view();
    return orc::rk3_step_adj(g, av, theta_ref, lam, ldprec, lK, stale);
}

ORC_EXPORT void orc_output_history_f64(const moist_params *mp, double *const *fields, double *const *out,
                                       double *stats) {
    orc::G<double> g(mp);
    orc::output_history(g, fields, out, stats);
}

ORC_EXPORT int orc_adjoint_run_f64(const moist_params *mp, double *const *fields, int nsteps, int ckpt,
                                   const double *theta_tgt, double *J, double *const *grad, double *gK) {
    orc::G<double> g(mp);
    return orc::adjoint_run(g, fields, nsteps, ckpt, theta_tgt, J, grad, gK);
}

// ---- forward-mode derivative of the 600-step loss (test infrastructure) --------
// The loss of adjoint_run, J = dx sum(prec) + 1/2 sum (theta - theta_tgt)^2
// after nsteps rk3_step!s, and its directional derivative along the tangent
// `dir` of the 7 prognostic fields (with their halo columns) and dK of
// (K_auto, Qc_threshold, K_accr): rk3_step! instantiated with dual numbers (FwdAD)
// (moist_dual.h), branches following the primal trajectory.  fields: the
// State in the order of the F_ enum (read only).
ORC_EXPORT int orc_jvp_run_f64(const moist_params *mp, const double *const *fields, const double *const *dir,
                               const double *dK, int nsteps, const double *theta_tgt, double *J, double *dJ) {
    orc::G<orc::FwdAD> g(mp);
    g.c.K_auto.d = dK[0];
    g.c.Qc_threshold.d = dK[1];
    g.c.K_accr.d = dK[2];
    const size_t N = g.n2();
    const size_t len[orc::F_COUNT] = {N, N, N, N, N, N, N, (size_t)g.ia, (size_t)g.ia, (size_t)g.ia, (size_t)g.ia,
                                      (size_t)g.ka, N, N, N, N, N, (size_t)g.ia, (size_t)g.ia, (size_t)g.ia, N};
    std::vector<std::vector<orc::FwdAD>> buf(orc::F_COUNT);
    orc::FwdAD *F[orc::F_COUNT];
    for (int f = 0; f < orc::F_COUNT; ++f) {
        buf[f].resize(len[f]);
        for (size_t q = 0; q < len[f]; ++q) buf[f][q] = orc::FwdAD(fields[f][q], f < 7 ? dir[f][q] : 0.0);
        F[f] = buf[f].data();
    }
    for (int n = 0; n < nsteps; ++n) {
        int rc = orc::rk3_step(g, F, orc::ORDER_SNAPSHOT);
        if (rc) return rc;
    }
    orc::FwdAD j(0.0);
    for (int i = g.is; i <= g.ie; ++i) j += g.dx * F[orc::F_PREC][i];
    for (int i = g.is; i <= g.ie; ++i)
        for (int k = g.ks; k <= g.ke; ++k) {
            const size_t q = g.ix(k, i);
            const orc::FwdAD d = F[3][q] / F[0][q] - orc::FwdAD(theta_tgt[q]);
            j += orc::FwdAD(0.5) * d * d;
        }
    *J = j.v;
    *dJ = j.d;
    return 0;
}

ORC_EXPORT int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
