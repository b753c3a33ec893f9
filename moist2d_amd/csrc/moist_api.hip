// moist_api.hip — host runtime and C ABI of libmoist2d_hip.so (include/moist2d.h).
//
// A handle owns one HIP stream, the device pools of every ensemble member,
// the tile work-lists and (optionally) a hipGraph of one large step.  The
// large step is the device restatement of rk3_step! (moist.jl:1589-1683):
//   surface fluxes, radiation (flag-gated), Kessler + sedimentation + source
//   terms; then three stages (dt/3, dt/2, dt with max(1,ceil(ns/3)),
//   max(1,ceil(ns/2)), ns substeps), each: diagnostics, slow tendencies,
//   the acoustic substeps (horizontal kernel + column kernel), scalar
//   transport + final update.
// State buffer rotation inside a step (stage input -> output):
//   S0 -> SA, SA -> SB, SB -> S0; S0 always holds the step-start state s0,
//   and stage 3 writes S0 in place because that kernel reads S0 only at the
//   cell it writes.  After a step SB still holds the stage-3 input, whose
//   periodic partner columns are what the reference leaves in its halos.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/moist2d.h"
#include "moist_kernels.h"
namespace mo {
MO_MIXED_KERNELS(extern)  // compiled in moist_mixed.hip
#include "moist_adjoint.h"
#include "moist_history.h"
}

using namespace mo;

namespace {

constexpr int TILE_X = 64, TILE_Z = 4, COL_TILE = 64;
// levels per acoustic chunk (FP64 keeps fewer levels in registers)
#ifndef MOIST_AC_L_MIXED
#define MOIST_AC_L_MIXED 4
#endif
#ifndef MOIST_AC_L_FP64
#define MOIST_AC_L_FP64 4
#endif
template <class S> constexpr int ac_levels() { return sizeof(S) == 4 ? MOIST_AC_L_MIXED : MOIST_AC_L_FP64; }

// Workgroups are handed to the 8 XCDs of the chip round-robin by block index.
// Reorder a tile list (natural order: neighbours adjacent) so that block b
// runs tile (b % 8) * per + b / 8: each XCD then works on one contiguous
// band of tiles and the halo lines its tiles share stay in that XCD's L2.
// The list is padded to a multiple of 8 with tiles whose member is -1.
constexpr int NXCD = 8;
template <class T> std::vector<T> xcd_order(const std::vector<T> &v, T pad) {
    const size_t n = v.size(), per = (n + NXCD - 1) / NXCD, np = per * NXCD;
    std::vector<T> out(np, pad);
    for (size_t b = 0; b < np; ++b) {
        const size_t t = (b % NXCD) * per + b / NXCD;
        if (t < n) out[b] = v[t];
    }
    return out;
}

}  // namespace

// ---- static scratch of moist_diagnostics_noalloc ----------------------------
// The block partials and per-member sums of the budget reduction live in this
// file-scope device array instead of memory the library allocates.  Each GPU
// holds its own copy until the process ends; host code uses the per-device
// address from hipGetSymbolAddress (taken once per device in moist_create),
// never the array's name, which on the host is only a stand-in symbol.
static __device__ __align__(256) double scratch[MOIST_NOALLOC_SCRATCH_DOUBLES];
namespace {
constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;
}  // namespace

struct AdjState;

struct moist_handle {
    int precision = MOIST_FP64;
    int device = 0;  // HIP device current at moist_create
    int nm = 0;
    std::string err;
    hipStream_t stream = nullptr;
    std::vector<Dom> dom_h;
    Dom *dom_d = nullptr;
    std::vector<int> ns_stage;  // substeps per stage
    int ns = 0;
    int any_sf = 0, any_rd = 0;
    int store_diag = 1;
    int64_t n2 = 0, nxs = 0, nzs = 0;  // pool sizes
    int max_ka = 0, max_ia = 0;
    int4 *tiles2d = nullptr;
    int ntiles2d = 0;
    int4 *stiles = nullptr;  // scalar-transport tiles (ST_X x ST_Z)
    int nstiles = 0;
    int2 *ctiles = nullptr;
    int nctiles = 0;
    int4 *atiles = nullptr;  // acoustic tiles {member, i0, TX, nch}
    int ac_L = 4, ac_NT = 1024, ac_W = 1;  // acoustic kernel: levels per chunk, threads per workgroup,
                                           // minimum waves per SIMD of its register budget
    int tuned = 0;
    int2 *wgroups = nullptr;  // {member, i0} of the adjoint column solve's waves
    int2 *acols = nullptr;    // {member, i} of every column (the fused acoustic adjoint, one wave each)
    int nacols = 0;
    int adj_col = 1;          // reverse substeps with the one-kernel-per-column vertical part
    int nwgroups = 0, wch = 8;  // chunks per column in those waves
    int natiles = 0;
    int nch_max = 0;
    double *cin = nullptr;   // c' entering each acoustic chunk, [chunk][column]
    // 1-D z data (double)
    double *z = nullptr, *zf = nullptr, *dz = nullptr, *theta_ref = nullptr;
    double4 *geo = nullptr;
    double *rho_bar = nullptr, *rt_bar = nullptr, *p_bar = nullptr, *L_bar = nullptr, *B_bar = nullptr;
    std::vector<double> rho_bar_h, rt_bar_h;
    // pools (void: typed views below)
    void *pool = nullptr;        // all 2-D fields of type S
    size_t elem = 8;
    double *prec = nullptr, *sfh = nullptr;
    void *pool1 = nullptr;       // 1-D x fields of type S
    int *nan_key = nullptr, *step_ctr = nullptr;
    void *staging = nullptr;
    double *diag_out = nullptr;
    int halo_src = -1;  // index of the buffer set whose partner columns fill halos (-1: none)
    // graph
    int use_graph = 0;
    hipGraphExec_t gexec = nullptr;
    // timing of the acoustic column kernel
    int ev_on = 0;
    std::vector<hipEvent_t> ev_pool;
    int ev_used = 0;
    double ac_ms = 0.0;
    int64_t ac_launches = 0;
    // reverse mode (allocated by the first moist_adjoint_run)
    AdjState *adj = nullptr;
    // forward records of the reverse sweep: 0 automatic, 1 per step, 2 per segment,
    // 3 per segment pipelined, 4 per step pipelined
    int adj_mode = 0;
    int64_t dev_bytes = 0;  // device memory of the state pools and tables
    // output_history: derived fields (allocated on first request), block partials, per-member stats
    double *hist_fields = nullptr, *hist_part = nullptr, *hist_stats = nullptr, *diag_all = nullptr;
};

namespace {

#define HIPCHK(h, x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            (h)->err = std::string(#x) + ": " + hipGetErrorString(e_);                     \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

// Field slots of the 2-D pool.
enum Slot {
    S0_ = 0, SA_ = 7, SB_ = 14,
    THETA = 21, PRES, RT2, UF, WF, R_RHO, R_RHOU, R_RHOW, R_RT, RP, RUP, RUP2, RWP, RTP0, RTP1, RTP2, MOMX, MOMZ,
    D_RHO, D_RT, D_QV, D_QC, D_QR, D_U2D, DG_TH_MP, DG_QV_MP, DG_QR_MP, DG_QR_SED, D_TH_RD,
    NSLOT
};
enum Slot1 { X_DU_SF = 0, X_DTH_SF, X_DQV_SF, X_MFLUX, X_SHFLUX, X_LHFLUX, NSLOT1 };

template <class S> S *slot(moist_handle *h, int s) { return (S *)h->pool + (size_t)s * h->n2; }
template <class S> S *slot1(moist_handle *h, int s) { return (S *)h->pool1 + (size_t)s * h->nxs; }

template <class S> SSet<S> sset(moist_handle *h, int first) {
    SSet<S> s;
    for (int f = 0; f < 7; ++f) s.f[f] = slot<S>(h, first + f);
    return s;
}

// Views of the work slots in the handle's pools, or in pools of the same
// layout given by `pool` / `pool1`.
template <class S> Work<S> work(moist_handle *h, void *pool = nullptr, void *pool1 = nullptr) {
    S *const p2 = pool ? (S *)pool : (S *)h->pool, *const p1 = pool1 ? (S *)pool1 : (S *)h->pool1;
    auto P = [&](int s) { return p2 + (size_t)s * h->n2; };
    auto P1 = [&](int s) { return p1 + (size_t)s * h->nxs; };
    Work<S> w;
    w.theta = P(THETA);
    w.pres = P(PRES);
    w.rt2 = P(RT2);
    w.uf = P(UF);
    w.wf = P(WF);
    w.R_rho = P(R_RHO);
    w.R_rhou = P(R_RHOU);
    w.R_rhow = P(R_RHOW);
    w.R_rt = P(R_RT);
    w.rp = P(RP);
    w.rup = P(RUP);
    w.rwp = P(RWP);
    w.mom_x = P(MOMX);
    w.mom_z = P(MOMZ);
    w.d_rho = P(D_RHO);
    w.d_rt = P(D_RT);
    w.d_qv = P(D_QV);
    w.d_qc = P(D_QC);
    w.d_qr = P(D_QR);
    w.d_u2d = P(D_U2D);
    w.d_u_sf = P1(X_DU_SF);
    w.d_th_sf = P1(X_DTH_SF);
    w.d_qv_sf = P1(X_DQV_SF);
    w.mflux = P1(X_MFLUX);
    w.shflux = P1(X_SHFLUX);
    w.lhflux = P1(X_LHFLUX);
    w.dg_th_mp = P(DG_TH_MP);
    w.dg_qv_mp = P(DG_QV_MP);
    w.dg_qr_mp = P(DG_QR_MP);
    w.dg_qr_sed = P(DG_QR_SED);
    w.d_th_rd = P(D_TH_RD);
    w.prec = h->prec;
    w.sfh = h->sfh;
    w.theta_ref = h->theta_ref;
    w.nan_key = h->nan_key;
    w.step_ctr = h->step_ctr;
    return w;
}

ZVec zvec(moist_handle *h) {
    ZVec z;
    z.z = h->z;
    z.zf = h->zf;
    z.dz = h->dz;
    z.geo = h->geo;
    z.rho_bar = h->rho_bar;
    z.rt_bar = h->rt_bar;
    z.p_bar = h->p_bar;
    z.L_bar = h->L_bar;
    z.B_bar = h->B_bar;
    return z;
}

hipEvent_t next_event(moist_handle *h) {
    if (h->ev_used >= (int)h->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        h->ev_pool.push_back(e);
    }
    return h->ev_pool[h->ev_used++];
}

// ---- acoustic kernel shapes ------------------------------------------------------
// Tiles {member, i0, TX, nch} for L levels per chunk and NT threads per
// workgroup: nch = smallest power of two with nch * L >= nz, TX = NT / nch.
int make_atiles(const std::vector<Dom> &doms, int L, int NT, std::vector<int4> &out, int &nch_max) {
    out.clear();
    nch_max = 0;
    for (size_t m = 0; m < doms.size(); ++m) {
        const Dom &D = doms[m];
        int nch = 1;
        while (nch * L < D.nz) nch *= 2;
        if (nch > NT || D.nz > AC_MAX_NZ) return -1;
        const int TX = NT / nch;
        nch_max = std::max(nch_max, nch);
        for (int i0 = 0; i0 < D.nx; i0 += TX) out.push_back(make_int4((int)m, i0, TX, nch));
    }
    out = xcd_order(out, make_int4(-1, 0, 0, 0));
    return 0;
}

int set_acoustic_shape(moist_handle *h, int L, int NT, int W = 1) {
    std::vector<int4> ta;
    int nch_max = 0;
    if (make_atiles(h->dom_h, L, NT, ta, nch_max)) {
        h->err = "nz too large for the acoustic kernel shape";
        return -2;
    }
    if (nch_max > h->nch_max) {
        h->err = "acoustic shape needs more chunks than allocated";
        return -2;
    }
    if (h->atiles) HIPCHK(h, hipFree(h->atiles));
    h->atiles = nullptr;
    HIPCHK(h, hipMalloc((void **)&h->atiles, sizeof(int4) * ta.size()));
    HIPCHK(h, hipMemcpy(h->atiles, ta.data(), sizeof(int4) * ta.size(), hipMemcpyHostToDevice));
    h->natiles = (int)ta.size();
    h->ac_L = L;
    h->ac_NT = NT;
    h->ac_W = W;
    if (h->gexec) {
        hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    return 0;
}

template <class S, bool P, int L, int NT, int W>
void launch_ac(moist_handle *h, const ZVec &zv, const SSet<S> &in, const Work<S> &w, const AcBuf<S> &ab, int div,
               double dt_ovr, int nsub, int first, bool fin) {
    // dynamic LDS: the per-level geometry (5 doubles per level of the tallest member)
    if constexpr (P) {
        static_assert(is_mixed<S>(), "the level-pair layout is the mixed form's");
        if (fin)
            hipLaunchKernelGGL((k_acoustic_pairs<L, NT, W, true>), dim3(h->natiles), dim3(NT),
                               (unsigned)(40 * h->max_ka), h->stream, h->dom_d, h->atiles, zv, in, w, ab, div, dt_ovr,
                               nsub, first);
        else
            hipLaunchKernelGGL((k_acoustic_pairs<L, NT, W, false>), dim3(h->natiles), dim3(NT),
                               (unsigned)(40 * h->max_ka), h->stream, h->dom_d, h->atiles, zv, in, w, ab, div, dt_ovr,
                               nsub, first);
    } else {
        hipLaunchKernelGGL((k_acoustic<S, L, NT, W>), dim3(h->natiles), dim3(NT), (unsigned)(40 * h->max_ka),
                           h->stream, h->dom_d, h->atiles, zv, in, w, ab, div, dt_ovr, nsub, first);
    }
}

// P: layout of the work fields (level pairs on the mixed form's forward
// path, x-fastest in the FP64 form and the adjoint's recording pass).
// fin (P only): the last substep of a stage, which also does the stage's
// final update of rho, rho_u, rho_theta, rho_w into ab.fin_*.
template <class S, bool P>
void launch_acoustic(moist_handle *h, const ZVec &zv, const SSet<S> &in, const Work<S> &w, const AcBuf<S> &ab,
                     int div, double dt_ovr, int nsub, int first, bool fin = false) {
    const int key = (h->ac_L * 10000 + h->ac_NT) * 10 + (h->ac_W == 8 ? 8 : 1);
#define AC_CASE(L_, NT_, W_)                                                                 \
    case (L_ * 10000 + NT_) * 10 + W_:                                                       \
        launch_ac<S, P, L_, NT_, W_>(h, zv, in, w, ab, div, dt_ovr, nsub, first, fin);       \
        break;
    switch (key) {
        AC_CASE(2, 256, 1)
        AC_CASE(2, 512, 1)
        AC_CASE(2, 1024, 1)
        AC_CASE(4, 256, 1)
        AC_CASE(4, 512, 1)
        AC_CASE(2, 512, 8)
        AC_CASE(2, 1024, 8)
        AC_CASE(4, 512, 8)
        AC_CASE(4, 1024, 8)
    default: launch_ac<S, P, 4, 1024, 1>(h, zv, in, w, ab, div, dt_ovr, nsub, first, fin); break;
    }
#undef AC_CASE
}

// ---- one RK stage ------------------------------------------------------------
// diag_in: the stage diagnostics of the input are still to be computed
// (k_diag); otherwise they are already there (in the mixed form's step,
// written by k_physics for stage 1 and by the previous stage's k_scalar for
// stages 2 and 3).  diag_out: this stage's k_scalar writes those of its
// output for the next stage.
template <class S>
void enqueue_stage(moist_handle *h, int in_set, int out_set, int s0_set, int div, double dt_ovr, int nsub,
                   int stage, int u_phys_2d, bool diag_in, bool diag_out, bool bump = false) {
    const dim3 b2(TILE_X, TILE_Z);
    const ZVec zv = zvec(h);
    const Work<S> w = work<S>(h);
    const SSet<S> in = sset<S>(h, in_set), out = sset<S>(h, out_set), s0 = sset<S>(h, s0_set);
    S *rtp[3] = {slot<S>(h, RTP0), slot<S>(h, RTP1), slot<S>(h, RTP2)};
    S *rup[2] = {slot<S>(h, RUP), slot<S>(h, RUP2)};
    // the work fields of the mixed form are in the level-pair layout
    constexpr bool PL = is_mixed<S>();
    // only the mixed form's k_scalar writes the next stage's diagnostics
    if (!PL) diag_in = true, diag_out = false;
    if (diag_in)
        hipLaunchKernelGGL((k_diag<S, PL>), dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, in, w);
    hipLaunchKernelGGL((k_slow<S, PL>), dim3(h->nstiles), dim3(ST_THREADS), 0, h->stream, h->dom_d, h->stiles, zv,
                       in, s0, w, rtp[0], u_phys_2d, bump ? h->step_ctr : (int *)nullptr);
    // rho_theta' rotates over three buffers (current, previous, next) and
    // rho_u' over two, because every column reads its neighbours' values
    int cur = 0, old = 0, ru = 0;
    for (int n = 0; n < nsub; ++n) {
        const int nxt = cur == old ? (cur + 1) % 3 : 3 - cur - old;
        // in the mixed form the last substep also does the stage's final
        // update of rho, rho_u, rho_theta and rho_w (k_scalar then does the
        // scalars' only)
        const bool fin = PL && n == nsub - 1;
        AcBuf<S> ab{};
        ab.rup_in = rup[ru];
        ab.rup_out = rup[1 - ru];
        ab.rtp_cur = rtp[cur];
        ab.rtp_old = rtp[old];
        ab.rtp_out = rtp[nxt];
        ab.cin = h->cin;
        ab.cin_stride = h->nxs;
        ab.rwp_in = w.rwp;  // updated in place
        if (fin) {
            ab.fin_r = out.f[RHO];
            ab.fin_ru = out.f[RHOU];
            ab.fin_rt = out.f[RT];
            ab.fin_rw = out.f[RHOW];
            ab.stage = stage;
        }
        // the per-launch events time the plain substeps (one kernel variant)
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (h->ev_on && !fin) {
            e0 = next_event(h);
            e1 = next_event(h);
            if (e0) hipEventRecord(e0, h->stream);
        }
        launch_acoustic<S, PL>(h, zv, in, w, ab, div, dt_ovr, nsub, n == 0 ? 1 : 0, fin);
        if (e1) hipEventRecord(e1, h->stream);
        old = cur;
        cur = nxt;
        ru = 1 - ru;
    }
    hipLaunchKernelGGL((k_scalar<S, PL>), dim3(h->nstiles), dim3(ST_THREADS), 0, h->stream, h->dom_d, h->stiles,
                       zv, in, s0, out, w, (const S *)rup[ru], (const S *)rtp[cur], div, dt_ovr, stage,
                       diag_out ? w.theta : nullptr, w.pres, w.rt2, SSet<double>{});
}

// Surface fluxes and gray radiation of the step-start state s (flag-gated).
template <class S> void enqueue_sf_rd(moist_handle *h, const ZVec &zv, const SSet<S> &s, const Work<S> &w) {
    if (h->any_sf)
        hipLaunchKernelGGL(k_surface<S>, dim3(h->nctiles), dim3(COL_TILE), 0, h->stream, h->dom_d, h->ctiles, zv, s, w);
    if (h->any_rd)
        hipLaunchKernelGGL(k_radiation<S>, dim3(h->nctiles), dim3(COL_TILE), 0, h->stream, h->dom_d, h->ctiles, zv,
                           s, w);
}

template <class S> void enqueue_step(moist_handle *h) {
    const ZVec zv = zvec(h);
    const Work<S> w = work<S>(h);
    const SSet<S> s = sset<S>(h, S0_);
    enqueue_sf_rd<S>(h, zv, s, w);
    // in the mixed form k_physics also writes the first stage's diagnostics,
    // and stages 2 and 3 read those the previous stage's k_scalar wrote
    constexpr bool PL = is_mixed<S>();
    hipLaunchKernelGGL(k_physics<S>, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d, h->tiles2d,
                       zv, s, w, h->store_diag, PL ? 1 : 0, SSet<double>{});
    enqueue_stage<S>(h, S0_, SA_, S0_, 3, 0.0, h->ns_stage[0], 0, 0, !PL, true, true);
    enqueue_stage<S>(h, SA_, SB_, S0_, 2, 0.0, h->ns_stage[1], 1, 0, false, true);
    enqueue_stage<S>(h, SB_, S0_, S0_, 1, 0.0, h->ns_stage[2], 2, 0, false, false);
}

void enqueue_step_any(moist_handle *h) {
    if (h->precision == MOIST_MIXED) enqueue_step<float>(h);
    else enqueue_step<double>(h);
}

int build_graph(moist_handle *h) {
    if (h->gexec) return 0;
    hipGraph_t g;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step_any(h);
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        h->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
        h->gexec = nullptr;
        return -1;
    }
    return 0;
}

int run_steps(moist_handle *h, int nsteps) {
    h->ev_used = 0;  // the per-launch events of this call reuse the pool
    if (h->use_graph && !h->ev_on) {
        if (build_graph(h)) return -1;
        for (int n = 0; n < nsteps; ++n) HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    } else {
        for (int n = 0; n < nsteps; ++n) enqueue_step_any(h);
    }
    HIPCHK(h, hipGetLastError());
    h->halo_src = SB_;
    return 0;
}

// HIP events that are destroyed with the scope, whatever path leaves it.
struct ScopedEvents {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    hipError_t create(int n) {
        for (int j = 0; j < n; ++j)
            if (hipError_t r = hipEventCreate(&e[j])) return r;
        return hipSuccess;
    }
    ~ScopedEvents() {
        for (hipEvent_t x : e)
            if (x) hipEventDestroy(x);
    }
};

int decode_nan(moist_handle *h) {
    int key = INT_MAX;
    HIPCHK(h, hipMemcpyAsync(&key, h->nan_key, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (key == INT_MAX) return 0;
    return key & 15;
}

// The NaN key orders reports by (step of this call, stage, field): the step
// counter restarts with the key at every call.
int reset_nan(moist_handle *h) {
    static const int key = INT_MAX, zero = 0;
    HIPCHK(h, hipMemcpyAsync(h->nan_key, &key, sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->step_ctr, &zero, sizeof(int), hipMemcpyHostToDevice, h->stream));
    return 0;
}

int check_member(moist_handle *h, int m) {
    if (m < 0 || m >= h->nm) {
        h->err = "member index out of range";
        return -2;
    }
    return 0;
}

// ---- base state of the mixed form -----------------------------------------
void base_from_state(const Dom &D, const double *z, const std::vector<double> &rb, const std::vector<double> &rtb,
                     double *pb, double *Lb, double *Bb) {
    const double Rd = 287.04, Cpd = 1004.64, Cvd = Cpd - Rd, P0 = 1e5, G = 9.81;
    for (int k = 0; k < D.nz; ++k) {
        Lb[k] = std::log(Rd * rtb[k] / P0);
        pb[k] = P0 * std::exp(Cpd / Cvd * Lb[k]);
    }
    for (int k = 0; k < D.nz; ++k) {
        if (k < D.nz - 1)
            Bb[k] = -(pb[k + 1] - pb[k]) / (z[k + 1] - z[k]) - G * (rb[k] + rb[k + 1]) * 0.5;
        else
            Bb[k] = 0.0;
    }
}

template <class S, class H>
int upload_t(moist_handle *h, int m, const void *const fields[7], const void *prec, const void *theta_ref) {
    const Dom &D = h->dom_h[m];
    const int ka = D.nz, ia = D.nx + 2 * D.halo;
    const size_t bytes = (size_t)ka * ia * sizeof(H);
    const bool mixed = h->precision == MOIST_MIXED;
    if (mixed) {
        // base state: horizontal means of rho and rho*theta (FP64)
        std::vector<double> rb(ka, 0.0), rtb(ka, 0.0), pb(ka), Lb(ka), Bb(ka), zz(ka);
        const H *r = (const H *)fields[0], *rt = (const H *)fields[3];
        for (int i = 0; i < D.nx; ++i)
            for (int k = 0; k < ka; ++k) {
                rb[k] += double(r[(size_t)(i + D.halo) * ka + k]);
                rtb[k] += double(rt[(size_t)(i + D.halo) * ka + k]);
            }
        for (int k = 0; k < ka; ++k) {
            rb[k] /= D.nx;
            rtb[k] /= D.nx;
        }
        HIPCHK(h, hipMemcpy(zz.data(), h->z + D.offz, ka * sizeof(double), hipMemcpyDeviceToHost));
        base_from_state(D, zz.data(), rb, rtb, pb.data(), Lb.data(), Bb.data());
        HIPCHK(h, hipMemcpy(h->rho_bar + D.offz, rb.data(), ka * 8, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->rt_bar + D.offz, rtb.data(), ka * 8, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->p_bar + D.offz, pb.data(), ka * 8, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->L_bar + D.offz, Lb.data(), ka * 8, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->B_bar + D.offz, Bb.data(), ka * 8, hipMemcpyHostToDevice));
    }
    for (int f = 0; f < 7; ++f) {
        HIPCHK(h, hipMemcpyAsync(h->staging, fields[f], bytes, hipMemcpyHostToDevice, h->stream));
        const double *base = nullptr;
        if (mixed && f == RHO) base = h->rho_bar + D.offz;
        if (mixed && f == RT) base = h->rt_bar + D.offz;
        hipLaunchKernelGGL((k_upload<S, H>), dim3((D.nx + 255) / 256, ka), dim3(256), 0, h->stream, D,
                           (const H *)h->staging, slot<S>(h, S0_ + f), base, D.halo, ka);
        HIPCHK(h, hipGetLastError());
    }
    // stale halo reads of surface_fluxes!: rho_u[ks, is-1], rho[ks, ie+1]
    double sf[2] = {double(((const H *)fields[1])[(size_t)(D.halo - 1) * ka]),
                    double(((const H *)fields[0])[(size_t)(D.halo + D.nx) * ka])};
    HIPCHK(h, hipMemcpyAsync(h->sfh + 2 * m, sf, sizeof(sf), hipMemcpyHostToDevice, h->stream));
    std::vector<double> pr(D.nx), tr(ka);
    if (prec)
        for (int i = 0; i < D.nx; ++i) pr[i] = double(((const H *)prec)[i + D.halo]);
    else
        std::fill(pr.begin(), pr.end(), 0.0);
    HIPCHK(h, hipMemcpyAsync(h->prec + D.offx, pr.data(), D.nx * 8, hipMemcpyHostToDevice, h->stream));
    if (theta_ref) {
        for (int k = 0; k < ka; ++k) tr[k] = double(((const H *)theta_ref)[k]);
        HIPCHK(h, hipMemcpyAsync(h->theta_ref + D.offz, tr.data(), ka * 8, hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->halo_src = -1;
    return 0;
}

template <class S, class H>
int download_field(moist_handle *h, const Dom &D, const S *src, const S *hsrc, void *dst, const double *base,
                   bool halos) {
    const int ka = D.nz, ia = D.nx + 2 * D.halo;
    hipLaunchKernelGGL((k_download<S, H>), dim3((ia + 255) / 256, ka), dim3(256), 0, h->stream, D, src,
                       halos ? hsrc : (const S *)nullptr, (H *)h->staging, base, D.halo, ka);
    HIPCHK(h, hipGetLastError());
    char *d = (char *)dst;
    const size_t col = (size_t)ka * sizeof(H);
    if (halos && hsrc) {
        HIPCHK(h, hipMemcpyAsync(d, h->staging, col * D.halo, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(d + col * (D.halo + D.nx), (char *)h->staging + col * (D.halo + D.nx),
                                 col * D.halo, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(h, hipMemcpyAsync(d + col * D.halo, (char *)h->staging + col * D.halo, col * D.nx,
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

template <class S, class H> int download_t(moist_handle *h, int m, void *const fields[7], void *prec) {
    const Dom &D = h->dom_h[m];
    const bool mixed = h->precision == MOIST_MIXED;
    for (int f = 0; f < 7; ++f) {
        if (!fields[f]) continue;
        const double *base = nullptr;
        if (mixed && f == RHO) base = h->rho_bar + D.offz;
        if (mixed && f == RT) base = h->rt_bar + D.offz;
        const S *hs = h->halo_src >= 0 ? slot<S>(h, h->halo_src + f) : nullptr;
        int rc = download_field<S, H>(h, D, slot<S>(h, S0_ + f), hs, fields[f], base, f != RT);
        if (rc) return rc;
    }
    if (prec) {
        std::vector<double> pr(D.nx);
        HIPCHK(h, hipMemcpy(pr.data(), h->prec + D.offx, D.nx * 8, hipMemcpyDeviceToHost));
        for (int i = 0; i < D.nx; ++i) ((H *)prec)[i + D.halo] = H(pr[i]);
    }
    return 0;
}

template <class S, class H> int download_phys_t(moist_handle *h, int m, void *const o2[6], void *const o1[6]) {
    const Dom &D = h->dom_h[m];
    const int s2[6] = {DG_TH_MP, DG_QV_MP, D_QC, DG_QR_MP, DG_QR_SED, D_TH_RD};
    for (int f = 0; f < 6; ++f) {
        if (!o2 || !o2[f]) continue;
        int rc = download_field<S, H>(h, D, slot<S>(h, s2[f]), nullptr, o2[f], nullptr, false);
        if (rc) return rc;
    }
    const int s1[6] = {X_DU_SF, X_DTH_SF, X_DQV_SF, X_MFLUX, X_SHFLUX, X_LHFLUX};
    for (int f = 0; f < 6; ++f) {
        if (!o1 || !o1[f]) continue;
        std::vector<S> v(D.nx);
        HIPCHK(h, hipMemcpy(v.data(), slot1<S>(h, s1[f]) + D.offx, D.nx * sizeof(S), hipMemcpyDeviceToHost));
        for (int i = 0; i < D.nx; ++i) ((H *)o1[f])[i + D.halo] = H(v[i]);
    }
    return 0;
}

template <class S, class H>
int compute_step_t(moist_handle *h, int m, const void *const phys[6], const void *const s0[7], double dt, int ns) {
    const Dom &D = h->dom_h[m];
    const int ka = D.nz, ia = D.nx + 2 * D.halo;
    const size_t bytes = (size_t)ka * ia * sizeof(H);
    const bool mixed = h->precision == MOIST_MIXED;
    const int pslot[6] = {D_RHO, D_U2D, D_RT, D_QV, D_QC, D_QR};
    for (int f = 0; f < 6; ++f) {
        HIPCHK(h, hipMemcpyAsync(h->staging, phys[f], bytes, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL((k_upload<S, H>), dim3((D.nx + 255) / 256, ka), dim3(256), 0, h->stream, D,
                           (const H *)h->staging, slot<S>(h, pslot[f]), (const double *)nullptr, D.halo, ka);
    }
    for (int f = 0; f < 7; ++f) {
        HIPCHK(h, hipMemcpyAsync(h->staging, s0[f], bytes, hipMemcpyHostToDevice, h->stream));
        const double *base = nullptr;
        if (mixed && f == RHO) base = h->rho_bar + D.offz;
        if (mixed && f == RT) base = h->rt_bar + D.offz;
        hipLaunchKernelGGL((k_upload<S, H>), dim3((D.nx + 255) / 256, ka), dim3(256), 0, h->stream, D,
                           (const H *)h->staging, slot<S>(h, SB_ + f), base, D.halo, ka);
    }
    HIPCHK(h, hipGetLastError());
    if (reset_nan(h)) return -1;
    enqueue_stage<S>(h, S0_, SA_, SB_, 1, dt, ns, 0, 1, true, false);
    // the output becomes the current state; the input's partner columns are the halos
    for (int f = 0; f < 7; ++f)
        HIPCHK(h, hipMemcpyAsync(slot<S>(h, SB_ + f), slot<S>(h, S0_ + f), h->n2 * sizeof(S),
                                 hipMemcpyDeviceToDevice, h->stream));
    for (int f = 0; f < 7; ++f)
        HIPCHK(h, hipMemcpyAsync(slot<S>(h, S0_ + f), slot<S>(h, SA_ + f), h->n2 * sizeof(S),
                                 hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipGetLastError());
    h->halo_src = SB_;
    int code = decode_nan(h);
    if (code) return code;
    // NaN in the given tendencies (moist.jl:1564-1578), checked after the state
    const int order[5] = {1, 2, 3, 4, 5};  // d_rho_u, d_rho_theta, d_qv, d_qc, d_qr
    for (int f = 0; f < 5; ++f) {
        const H *a = (const H *)phys[order[f]];
        for (int i = D.halo; i < D.halo + D.nx; ++i)
            for (int k = 0; k < ka; ++k)
                if (std::isnan(double(a[(size_t)i * ka + k]))) return 8 + f;
    }
    return 0;
}

// ---- output_history's derived fields and budget -------------------------------------
template <class S, class H> int history_t(moist_handle *h, int m, double *stats, void *const fields[HF_COUNT]) {
    bool want = false;
    for (int f = 0; fields && f < HF_COUNT; ++f) want |= fields[f] != nullptr;
    if (want && !h->hist_fields) HIPCHK(h, hipMalloc((void **)&h->hist_fields, sizeof(double) * HF_COUNT * h->n2));
    if (!h->hist_part) HIPCHK(h, hipMalloc((void **)&h->hist_part, sizeof(double) * HS_COUNT * h->ntiles2d));
    if (!h->hist_stats) HIPCHK(h, hipMalloc((void **)&h->hist_stats, sizeof(double) * HS_COUNT * h->nm));
    hipLaunchKernelGGL(k_history<S>, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d, h->tiles2d,
                       zvec(h), sset<S>(h, S0_), want ? h->hist_fields : (double *)nullptr, (int64_t)h->n2,
                       h->hist_part);
    hipLaunchKernelGGL(k_history_reduce, dim3(h->nm), dim3(256), 0, h->stream, (const int4 *)h->tiles2d,
                       h->ntiles2d, (const double *)h->hist_part, h->hist_stats);
    HIPCHK(h, hipGetLastError());
    if (stats)
        HIPCHK(h, hipMemcpyAsync(stats, h->hist_stats + (size_t)HS_COUNT * m, sizeof(double) * HS_COUNT,
                                 hipMemcpyDeviceToHost, h->stream));
    const Dom &D = h->dom_h[m];
    const size_t n = (size_t)D.nx * D.nz;
    std::vector<double> tmp(n);
    for (int f = 0; want && f < HF_COUNT; ++f) {
        if (!fields[f]) continue;
        HIPCHK(h, hipMemcpyAsync(tmp.data(), h->hist_fields + (size_t)f * h->n2 + D.off2, n * sizeof(double),
                                 hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (size_t q = 0; q < n; ++q) ((H *)fields[f])[q] = H(tmp[q]);
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// ---- reverse mode ------------------------------------------------------------------
// Buffers of one adjoint run: checkpoints every `ckpt` steps, the states of
// one segment (recomputed from its checkpoint), the forward record of one
// step (stage diagnostics and tendencies, per-substep rho_theta', rho_u',
// rho_w'), and the adjoint fields.
//
// The trajectory (forward pass, checkpoints, segment recomputation) advances
// in the handle's form.  Each step is then recorded by an FP64 forward step
// from that trajectory's state and reversed in FP64, so in the mixed form the
// gradient is the FP64 step's adjoint along the mixed-form trajectory.
enum AdjWork {
    W_TH = 0, W_RT, W_PRES, W_LU, W_LW, W_LRR, W_LRU, W_LRW, W_LRT, W_LMX, W_LMZ, W_LRP, W_LRWP, W_LRTP, W_PEND,
    W_LINCR, W_LINCT, W_Y, W_CPR, W_RMI, W_ACO, W_LRUP0, W_LRUP1, W_Q0, W_Q1, W_Q2, W_UF, W_WF, NWORK
};
enum { REC_TH = 0, REC_PRES, REC_RT2, REC_RR, REC_RRU, REC_RRW, REC_RRT, REC_MX, REC_MZ, NREC };

}  // namespace

struct AdjState {
    int nsteps = 0, ckpt = 0, ncp = 0, mode = 0;
    // checkpoint: the 7 state slots in the handle's form (cp_words doubles), then prec
    size_t cp_stride = 0, cp_words = 0;
    double *cp = nullptr, *SC = nullptr;
    // step inputs of one segment in the handle's form (7 fields per step,
    // seg_bytes apart)
    char *seg = nullptr;
    size_t seg_bytes = 0;
    // forward records: blocks of {stage-2 input, stage-3 input (FP64), stage
    // records, substep records (in the handle's form, rec_elem bytes)}; one
    // block per step of a segment when they fit in memory (`segrec`), else one
    // block recomputed per step
    char *recs = nullptr, *cur = nullptr;
    size_t rb_bytes = 0;
    int rec_elem = 8;
    int segrec = 0, nblocks = 1;
    // pipelined segments: two halves of the segment buffers (records, step
    // inputs, stale halo pairs); segment sg uses half sg & 1.  The recording
    // pass of segment sg - 1 runs on stream s2 while stream `stream` reverses
    // segment sg; events order each half's writer after its last reader.
    // With per-step records, pipelining keeps two record blocks: the recording
    // pass of step j - 1 (stream s2, block (j - 1) mod 2) runs beside the
    // reverse of step j (block j mod 2).
    int overlap = 0, half = 0;
    hipStream_t s2 = nullptr;
    hipEvent_t ev_go = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr};
    double *lam = nullptr, *lam2 = nullptr, *ls0 = nullptr, *lph = nullptr, *wk = nullptr, *lim = nullptr;
    double *gk = nullptr, *tgt = nullptr, *fin = nullptr, *J = nullptr, *gK = nullptr;
    // surface fluxes: the stale halo pair of every step of a segment (2 per
    // member), the adjoints of a step's stale reads carried to the step
    // before, and ka_surface's per-column buffer
    double *sfh_seg = nullptr, *carry = nullptr, *sfb = nullptr;
    std::vector<double *> allocs;
    size_t bytes = 0;
    std::vector<double> J_h, gK_h;
    double ms_fwd = 0.0, ms_bwd = 0.0;
    int sub_off[3] = {0, 0, 0};  // first substep record of each stage
};

namespace {

void adj_free(moist_handle *h) {
    if (!h->adj) return;
    AdjState &A = *h->adj;
    if (A.s2) {
        hipStreamSynchronize(A.s2);
        hipStreamDestroy(A.s2);
    }
    for (hipEvent_t e : {A.ev_go, A.ev_rec[0], A.ev_rec[1], A.ev_rev[0], A.ev_rev[1]})
        if (e) hipEventDestroy(e);
    for (double *p : A.allocs)
        if (p) hipFree(p);
    delete h->adj;
    h->adj = nullptr;
}

// FP64 inputs kept in a record block: those of stages 2 and 3, and in the
// mixed form also the step input (widened)
int rec_nin(const moist_handle *h) { return h->precision == MOIST_MIXED ? 3 : 2; }
size_t rec_in_bytes(const moist_handle *h) { return 7 * (size_t)rec_nin(h) * h->n2 * sizeof(double); }

int adj_alloc_buffers(moist_handle *h, int nsteps, int ckpt);
// The adjoint's buffers for (nsteps, ckpt) and the record mode; kept between
// runs with the same three.  A failed allocation frees what it got, so a
// later call starts over instead of finding a half-built state.
int adj_alloc(moist_handle *h, int nsteps, int ckpt) {
    if (h->adj && h->adj->nsteps == nsteps && h->adj->ckpt == ckpt && h->adj->mode == h->adj_mode) return 0;
    adj_free(h);
    if (adj_alloc_buffers(h, nsteps, ckpt)) {
        adj_free(h);
        (void)hipGetLastError();  // the failed call's error is in h->err; later checks must not see it
        return -1;
    }
    return 0;
}
int adj_alloc_buffers(moist_handle *h, int nsteps, int ckpt) {
    h->adj = new AdjState();
    AdjState &A = *h->adj;
    A.nsteps = nsteps;
    A.ckpt = ckpt;
    A.mode = h->adj_mode;
    A.ncp = (nsteps + ckpt - 1) / ckpt;
    const size_t n2 = h->n2;
    A.cp_words = (7 * n2 * h->elem + 7) / 8;
    A.cp_stride = A.cp_words + h->nxs + 2 * (size_t)h->nm;  // state, prec, the stale halo pair
    A.rec_elem = (int)h->elem;
    A.seg_bytes = (7 * n2 * h->elem + 255) / 256 * 256;
    int nrec_sub = 0;
    for (int st = 0; st < 3; ++st) {
        A.sub_off[st] = nrec_sub;
        nrec_sub += h->ns_stage[st] + 1;
    }
    A.rb_bytes = (rec_in_bytes(h) + (3 * (size_t)NREC + 3 * (size_t)nrec_sub) * n2 * A.rec_elem + 255) / 256 * 256;
    {
        // records of a whole segment when they fit in the free memory with a
        // margin (a tenth of it, at least 8 GiB): one recording pass per step
        // instead of a state pass and a recording pass; those of two segments
        // when they fit, so that one segment's recording pass runs beside the
        // reverse sweep of the next (pipelined segments)
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
            const size_t seg1 = (A.rb_bytes + A.seg_bytes) * (size_t)ckpt;
            const size_t other = (A.cp_stride * A.ncp + 64 * n2) * sizeof(double);
            const size_t margin = std::max(fr / 10, (size_t)8 << 30);
            A.segrec = fr > margin && seg1 + other < fr - margin ? 1 : 0;
            // pipelined: two segments' buffers, or (per-step records) two record blocks
            const size_t step2 = 2 * A.rb_bytes + A.seg_bytes * (size_t)ckpt;
            A.overlap = (A.segrec ? 2 * seg1 : step2) + other < fr - margin ? 1 : 0;
        }
        if (A.mode) {
            A.segrec = A.mode == 2 || A.mode == 3 ? 1 : 0;
            A.overlap = A.mode == 3 || A.mode == 4 ? 1 : 0;
        }
        A.nblocks = A.segrec ? ckpt : A.overlap ? 2 : 1;
    }
    const size_t halves = A.overlap && A.segrec ? 2 : 1;
    if (A.overlap) {
        HIPCHK(h, hipStreamCreate(&A.s2));
        for (hipEvent_t *e : {&A.ev_go, &A.ev_rec[0], &A.ev_rec[1], &A.ev_rev[0], &A.ev_rev[1]})
            HIPCHK(h, hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    auto al = [&](double *&p, size_t n) -> int {
        HIPCHK(h, hipMalloc((void **)&p, n * sizeof(double)));
        A.allocs.push_back(p);
        A.bytes += n * sizeof(double);
        return 0;
    };
    double *seg = nullptr, *recs = nullptr;
    if (al(A.cp, A.cp_stride * A.ncp) || al(seg, halves * A.seg_bytes * (size_t)ckpt / 8) || al(A.SC, 7 * n2) ||
        al(recs, halves * A.rb_bytes * (size_t)A.nblocks / 8) || al(A.lam, 7 * n2) ||
        al(A.lam2, 7 * n2) || al(A.ls0, 7 * n2) || al(A.lph, 7 * n2) || al(A.wk, NWORK * n2) ||
        al(A.lim, 6 * h->nxs) || al(A.gk, 3 * n2) || al(A.tgt, n2) || al(A.fin, A.cp_stride) || al(A.J, h->nm) ||
        al(A.gK, 3 * h->nm))
        return -1;
    if (al(A.sfh_seg, halves * 2 * (size_t)h->nm * ckpt) || al(A.carry, 2 * (size_t)h->nm) ||
        al(A.sfb, SF_SLOTS * (size_t)h->nxs))
        return -1;
    A.seg = (char *)seg;
    A.recs = (char *)recs;
    A.cur = A.recs;
    A.J_h.assign(h->nm, 0.0);
    A.gK_h.assign(3 * h->nm, 0.0);
    return 0;
}

SSet<double> dset(double *base, size_t n2) {
    SSet<double> s;
    for (int f = 0; f < 7; ++f) s.f[f] = base + f * n2;
    return s;
}
SSet<float> fset(float *base, size_t n2) {
    SSet<float> s;
    for (int f = 0; f < 7; ++f) s.f[f] = base + f * n2;
    return s;
}

// Layout of a record block (current block h->adj->cur): the FP64 inputs
// (7 fields each: in the mixed form the step input, then the stage-2 and
// stage-3 inputs; in the FP64 form the last two), per stage the NREC stage
// records, then per stage and substep n (n = 0: initial) rho_theta', rho_u',
// rho_w' after it; the records are of the recording pass's type RC (the
// handle's).
double *rec_in(moist_handle *h, int st) {  // st = 1, 2; 0 in the mixed form
    return (double *)(h->adj->cur + (size_t)(st - 3 + rec_nin(h)) * 7 * h->n2 * sizeof(double));
}
template <class RC> RC *rec_st(moist_handle *h, int st, int r) {
    return (RC *)(h->adj->cur + rec_in_bytes(h) + ((size_t)st * NREC + r) * h->n2 * sizeof(RC));
}
template <class RC> RC *rec_sub(moist_handle *h, int st, int n, int f) {
    return (RC *)(h->adj->cur + rec_in_bytes(h) +
                  (3 * (size_t)NREC + (size_t)3 * (h->adj->sub_off[st] + n) + f) * h->n2 * sizeof(RC));
}
template <class RC> RC *rec_rtp(moist_handle *h, int st, int n) { return rec_sub<RC>(h, st, n, 0); }
template <class RC> RC *rec_rup(moist_handle *h, int st, int n) { return rec_sub<RC>(h, st, n, 1); }
template <class RC> RC *rec_rwp(moist_handle *h, int st, int n) { return rec_sub<RC>(h, st, n, 2); }
double *wk(moist_handle *h, int w) { return h->adj->wk + (size_t)w * h->n2; }

template <class RC> AdjW<RC> adj_w(moist_handle *h, int st) {
    AdjW<RC> a;
    a.theta = rec_st<RC>(h, st, REC_TH);
    a.pres = rec_st<RC>(h, st, REC_PRES);
    a.rt2 = rec_st<RC>(h, st, REC_RT2);
    a.R_rho = rec_st<RC>(h, st, REC_RR);
    a.R_rhou = rec_st<RC>(h, st, REC_RRU);
    a.R_rhow = rec_st<RC>(h, st, REC_RRW);
    a.R_rt = rec_st<RC>(h, st, REC_RRT);
    a.mom_x = rec_st<RC>(h, st, REC_MX);
    a.mom_z = rec_st<RC>(h, st, REC_MZ);
    a.Lth = wk(h, W_TH);
    a.Lrt = wk(h, W_RT);
    a.Lpres = wk(h, W_PRES);
    a.LU = wk(h, W_LU);
    a.LW = wk(h, W_LW);
    a.LRr = wk(h, W_LRR);
    a.LRu = wk(h, W_LRU);
    a.LRw = wk(h, W_LRW);
    a.LRt = wk(h, W_LRT);
    a.Lmx = wk(h, W_LMX);
    a.Lmz = wk(h, W_LMZ);
    a.Lrp = wk(h, W_LRP);
    a.Lrwp = wk(h, W_LRWP);
    a.Lrtp = wk(h, W_LRTP);
    a.pend = wk(h, W_PEND);
    a.Linc_r = wk(h, W_LINCR);
    a.Linc_t = wk(h, W_LINCT);
    a.Y = wk(h, W_Y);
    a.cpr = wk(h, W_CPR);
    a.rmi = wk(h, W_RMI);
    a.aco = wk(h, W_ACO);
    for (int sc = 0; sc < 3; ++sc) a.q[sc] = wk(h, W_Q0 + sc);
    a.uf = wk(h, W_UF);
    a.wf = wk(h, W_WF);
    a.linb = h->adj->lim;
    a.lint = h->adj->lim + 3 * h->nxs;
    return a;
}

void fill(moist_handle *h, double *p, int64_t n, double v = 0.0) {
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(ka_fill, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, h->stream, p, n, v);
}

// Zeroes the arrays of `ps` (each `len` doubles) in one launch.
void copy_bytes(moist_handle *h, void *dst, const void *src, size_t n) {
    hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, h->stream);
}
void copy_set(moist_handle *h, double *dst, const double *src, size_t n) { copy_bytes(h, dst, src, n * sizeof(double)); }

// One stage of the forward pass in the handle's form S with every
// intermediate the adjoint reads kept in the stage record (the kernels of
// enqueue_stage, their work arrays pointed into the record).
// wide: the stage output as full FP64 values (the mixed form's stage-2 and -3
// inputs for the reverse sweep; null pointers: none).
template <class S>
void enqueue_stage_rec(moist_handle *h, int st, SSet<S> in, SSet<S> out, SSet<S> s0, int div,
                       SSet<double> wide = SSet<double>{}) {
    const int nsub = h->ns_stage[st];
    const dim3 b2(TILE_X, TILE_Z);
    const ZVec zv = zvec(h);
    Work<S> w = work<S>(h);
    w.theta = rec_st<S>(h, st, REC_TH);
    w.pres = rec_st<S>(h, st, REC_PRES);
    w.rt2 = rec_st<S>(h, st, REC_RT2);
    w.R_rho = rec_st<S>(h, st, REC_RR);
    w.R_rhou = rec_st<S>(h, st, REC_RRU);
    w.R_rhow = rec_st<S>(h, st, REC_RRW);
    w.R_rt = rec_st<S>(h, st, REC_RRT);
    w.mom_x = rec_st<S>(h, st, REC_MX);
    w.mom_z = rec_st<S>(h, st, REC_MZ);
    w.rup = rec_rup<S>(h, st, 0);
    w.rwp = rec_rwp<S>(h, st, 0);  // rho_w' of every substep goes straight to its record
    hipLaunchKernelGGL((k_diag<S, false>), dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, in, w);
    hipLaunchKernelGGL((k_slow<S, false>), dim3(h->nstiles), dim3(ST_THREADS), 0, h->stream, h->dom_d, h->stiles, zv,
                       in, s0, w, rec_rtp<S>(h, st, 0), 0, (int *)nullptr);
    for (int n = 0; n < nsub; ++n) {
        AcBuf<S> ab{};
        ab.rup_in = rec_rup<S>(h, st, n);
        ab.rup_out = rec_rup<S>(h, st, n + 1);
        ab.rtp_cur = rec_rtp<S>(h, st, n);
        ab.rtp_old = rec_rtp<S>(h, st, n > 0 ? n - 1 : 0);
        ab.rtp_out = rec_rtp<S>(h, st, n + 1);
        ab.cin = h->cin;
        ab.cin_stride = h->nxs;
        ab.rwp_in = rec_rwp<S>(h, st, n);
        w.rwp = rec_rwp<S>(h, st, n + 1);
        launch_acoustic<S, false>(h, zv, in, w, ab, div, 0.0, nsub, n == 0 ? 1 : 0);
    }
    hipLaunchKernelGGL((k_scalar<S, false>), dim3(h->nstiles), dim3(ST_THREADS), 0, h->stream, h->dom_d, h->stiles,
                       zv, in, s0, out, w, (const S *)rec_rup<S>(h, st, nsub), (const S *)rec_rtp<S>(h, st, nsub), div,
                       0.0, st, (S *)nullptr, (S *)nullptr, (S *)nullptr, wide);
}

// FP64 form: forward pass of one step from the state `sn` into the current
// record block; the stage-2 and -3 inputs are kept in the block; stage 3
// writes the separate set SC (the step's output).
void enqueue_step_rec64(moist_handle *h, const double *sn) {
    const size_t n2 = h->n2;
    copy_set(h, slot<double>(h, S0_), sn, 7 * n2);  // slots S0_..S0_+6 are contiguous
    const ZVec zv = zvec(h);
    const Work<double> w = work<double>(h);
    const SSet<double> s0 = sset<double>(h, S0_), ia = dset(rec_in(h, 1), n2), ib = dset(rec_in(h, 2), n2);
    enqueue_sf_rd<double>(h, zv, s0, w);
    hipLaunchKernelGGL(k_physics<double>, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d,
                       h->tiles2d, zv, s0, w, 0, 0, SSet<double>{});
    enqueue_stage_rec<double>(h, 0, s0, ia, s0, 3);
    enqueue_stage_rec<double>(h, 1, ia, ib, s0, 2);
    enqueue_stage_rec<double>(h, 2, ib, dset(h->adj->SC, n2), s0, 1);
}

// Mixed form: one step of the mixed form from the handle's state, recording
// into the current block: FP32 stage and substep records, and the step input
// and the stage-2 and -3 inputs as FP64 (written by k_physics and the first
// two stages' k_scalar, which hold those values); the step's output replaces
// the state, as a plain step does.
void enqueue_step_rec_mixed(moist_handle *h) {
    const size_t n2 = h->n2;
    const ZVec zv = zvec(h);
    const dim3 b2(TILE_X, TILE_Z);
    const Work<float> w = work<float>(h);
    const SSet<float> s0 = sset<float>(h, S0_), sa = sset<float>(h, SA_), sb = sset<float>(h, SB_);
    enqueue_sf_rd<float>(h, zv, s0, w);
    hipLaunchKernelGGL(k_physics<float>, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, s0, w, 0, 0,
                       dset(rec_in(h, 0), n2));
    enqueue_stage_rec<float>(h, 0, s0, sa, s0, 3, dset(rec_in(h, 1), n2));
    enqueue_stage_rec<float>(h, 1, sa, sb, s0, 2, dset(rec_in(h, 2), n2));
    enqueue_stage_rec<float>(h, 2, sb, s0, s0, 1);
}

// Adjoint of one step from its record (the current block, records of type
// RC): A.lam holds the adjoint of the step output on entry and of the step
// input `sn` (FP64) on exit.
// stale: the step's stale halo pair (surface fluxes).
template <class RC> void enqueue_step_adj(moist_handle *h, const double *sn, const double *stale) {
    AdjState &A = *h->adj;
    const size_t n2 = h->n2;
    const ZVec zv = zvec(h);
    const dim3 b2(TILE_X, TILE_Z);
    const SSet<double> ls0 = dset(A.ls0, n2), lph = dset(A.lph, n2);
    const int divs[3] = {3, 2, 1};
    double *lo = A.lam, *li = A.lam2;
    for (int st = 2; st >= 0; --st) {
        const SSet<double> in = dset(st == 0 ? const_cast<double *>(sn) : rec_in(h, st), n2);
        const int nsub = h->ns_stage[st], div = divs[st];
        const AdjW<RC> a = adj_w<RC>(h, st);
        // the stage's accumulators are cleared by ka_pre; the step-level ones
        // (ls0, lph) are not cleared: in the step's first reverse stage (st = 2)
        // ka_scalar and ka_slow write them instead of adding to them (`first`)
        FillList zero{};
        for (double *p : {wk(h, W_TH), wk(h, W_RT), wk(h, W_LRR), wk(h, W_LRU), wk(h, W_LRW), wk(h, W_LRT),
                          wk(h, W_PEND)})
            zero.p[zero.n++] = p;
        double *Lrup[2] = {wk(h, W_LRUP0), wk(h, W_LRUP1)};
        // the limiter sums of the stage reversed before (into its input adjoint, lo
        // here) run in ka_pre's extra blocks; those of stage 1 in their own launch,
        // since the physics adjoint reads lo next
        const LinSum lin{a.linb, a.lint, dset(lo, n2), h->ntiles2d};
        hipLaunchKernelGGL(ka_pre<RC>, dim3(h->ntiles2d + (st < 2 ? h->nm : 0)), b2, 0, h->stream, h->dom_d,
                           h->tiles2d, zv, in, a, zero, lin);
        hipLaunchKernelGGL(ka_scalar<RC>, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, in,
                           dset(lo, n2), dset(li, n2), ls0, lph, a, Lrup[0], div, st == 2 ? 1 : 0);
        if (st == 0)
            hipLaunchKernelGGL(ka_lin<RC>, dim3(h->nm), dim3(256), 0, h->stream, h->dom_d, a, dset(li, n2));
        int cur = 0;
        for (int n = nsub; n >= 1; --n) {
            const int m = n - 1;
            // the one-kernel-per-column form when one wave per column is also the
            // three-kernel form's layout (few columns: the kernels are bound by
            // latency, and launches and round trips are what the fusion saves);
            // many columns keep the three kernels, whose waves hold 8 or 16
            // adjacent columns and so coalesce their loads
            if (h->adj_col && h->max_ka <= AC_COL_MAX_NZ && h->wch == 64) {
                // the vertical part in one kernel per column
                const unsigned lds = (unsigned)(AC_COL_ARRAYS * sizeof(double) * h->max_ka);
                const RC *rtp_s = rec_rtp<RC>(h, st, m), *rup_n = rec_rup<RC>(h, st, m + 1),
                         *rwp_s = rec_rwp<RC>(h, st, m), *rwp_n = rec_rwp<RC>(h, st, m + 1);
                const int fac = n == nsub ? 1 : 0;
                // one wave per column up to 64 levels, two above (half the levels per
                // thread, twice the waves)
                if (h->max_ka <= 64)
                    hipLaunchKernelGGL((ka_ac_col<RC, 1, 1>), dim3(h->nacols), dim3(64), lds, h->stream, h->dom_d,
                                       h->acols, zv, in, a, rtp_s, rup_n, rwp_s, rwp_n, div, nsub, fac, h->max_ka);
                else if (h->max_ka <= 128)
                    hipLaunchKernelGGL((ka_ac_col<RC, 1, 2>), dim3(h->nacols), dim3(128), lds, h->stream, h->dom_d,
                                       h->acols, zv, in, a, rtp_s, rup_n, rwp_s, rwp_n, div, nsub, fac, h->max_ka);
                else
                    hipLaunchKernelGGL((ka_ac_col<RC, 2, 2>), dim3(h->nacols), dim3(128), lds, h->stream, h->dom_d,
                                       h->acols, zv, in, a, rtp_s, rup_n, rwp_s, rwp_n, div, nsub, fac, h->max_ka);
            } else {
                hipLaunchKernelGGL(ka_ac_v1<RC>, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, in,
                                   a, (const RC *)rec_rwp<RC>(h, st, m + 1), div, nsub);
                // one wave (one column group) per block: the waves are independent,
                // and small grids then reach every CU
                const dim3 gs(h->nwgroups), bs(64);
                if (h->wch == 4)
                    hipLaunchKernelGGL((ka_ac_solve<4, RC>), gs, bs, 0, h->stream, h->dom_d, h->wgroups, h->nwgroups,
                                       zv, a, div, nsub, n == nsub ? 1 : 0);
                else if (h->wch == 64)
                    hipLaunchKernelGGL((ka_ac_solve<64, RC>), gs, bs, 0, h->stream, h->dom_d, h->wgroups,
                                       h->nwgroups, zv, a, div, nsub, n == nsub ? 1 : 0);
                else
                    hipLaunchKernelGGL((ka_ac_solve<8, RC>), gs, bs, 0, h->stream, h->dom_d, h->wgroups, h->nwgroups,
                                       zv, a, div, nsub, n == nsub ? 1 : 0);
                hipLaunchKernelGGL(ka_ac_v2<RC>, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, a,
                                   (const RC *)rec_rtp<RC>(h, st, m), (const RC *)rec_rup<RC>(h, st, m + 1),
                                   (const RC *)rec_rwp<RC>(h, st, m), (const RC *)rec_rwp<RC>(h, st, m + 1), div,
                                   nsub);
            }
            hipLaunchKernelGGL(ka_ac_hor<RC>, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, a,
                               (const double *)Lrup[cur], Lrup[1 - cur], (const RC *)rec_rtp<RC>(h, st, m),
                               (const RC *)rec_rtp<RC>(h, st, m > 0 ? m - 1 : 0),
                               (const RC *)rec_rup<RC>(h, st, m + 1), div, nsub, n == 1 ? 1 : 0);
            cur = 1 - cur;
        }
        hipLaunchKernelGGL(ka_slow<RC>, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, in,
                           dset(li, n2), ls0, lph, a, (const double *)Lrup[cur], st == 2 ? 1 : 0);
        hipLaunchKernelGGL(ka_diag<RC>, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, in,
                           dset(li, n2), a);
        // the stage-3 input is what the next step's surface fluxes read in the
        // halos: the adjoints of those reads (reversed just before) enter here
        if (st == 2 && h->any_sf)
            hipLaunchKernelGGL(ka_stale_add, dim3((h->nm + 63) / 64), dim3(64), 0, h->stream, h->dom_d, h->nm,
                               dset(li, n2), (const double *)A.carry);
        std::swap(lo, li);
    }
    // lo: adjoint of the stage-1 input (A.lam2: the three stages alternate
    // the two buffers, starting from A.lam); the step input also enters as s0
    // of every stage and through physics.  The step input's adjoint goes to
    // A.lam, so every reverse step reads and writes the same buffers.
    const SSet<double> snv = dset(const_cast<double *>(sn), n2);
    if (h->any_sf)  // before ka_physics, which gathers its per-column adjoints
        hipLaunchKernelGGL(ka_surface, dim3(h->nctiles), dim3(COL_TILE), 0, h->stream, h->dom_d, h->ctiles, zv, snv,
                           stale, lph, A.sfb, (int64_t)h->nxs, A.carry);
    hipLaunchKernelGGL(ka_physics, dim3(h->ntiles2d), b2, 0, h->stream, h->dom_d, h->tiles2d, zv, snv, lph,
                       dset(lo, n2), ls0, dset(A.lam, n2), A.gk, A.gk + n2, A.gk + 2 * n2,
                       h->any_sf ? (const double *)A.sfb : (const double *)nullptr, (int64_t)h->nxs);
    if (h->any_rd)  // after it: the column kernel adds into the step input's adjoint
        hipLaunchKernelGGL(ka_radiation, dim3(h->nctiles), dim3(COL_TILE), 0, h->stream, h->dom_d, h->ctiles, zv,
                           snv, (const double *)h->theta_ref, (const double *)A.lph + 2 * n2, dset(A.lam, n2),
                           wk(h, W_Y), wk(h, W_CPR), wk(h, W_RMI), wk(h, W_ACO), wk(h, W_PEND));
}

// Checkpoints hold the state slots (S0_.. S0_+6, contiguous at the front of
// the pool) in the handle's form, and the accumulated rain.
void save_state(moist_handle *h, double *dst) {
    copy_bytes(h, dst, h->pool, 7 * h->n2 * h->elem);
    copy_set(h, dst + h->adj->cp_words, h->prec, h->nxs);
    copy_set(h, dst + h->adj->cp_words + h->nxs, h->sfh, 2 * h->nm);
}
void load_state(moist_handle *h, const double *src) {
    copy_bytes(h, h->pool, src, 7 * h->n2 * h->elem);
    copy_set(h, h->prec, src + h->adj->cp_words, h->nxs);
    copy_set(h, h->sfh, src + h->adj->cp_words + h->nxs, 2 * h->nm);
}
// The stale halo pair surface_fluxes! reads at step j of the segment.
double *seg_sfh(moist_handle *h, int j) {
    return h->adj->sfh_seg + 2 * (size_t)h->nm * ((size_t)h->adj->half * h->adj->ckpt + j);
}
// The current state as a full FP64 state (7 fields) at `dst`.
void state64(moist_handle *h, double *dst) {
    if (h->precision == MOIST_MIXED)
        hipLaunchKernelGGL(ka_widen, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d, h->tiles2d,
                           zvec(h), sset<float>(h, S0_), dset(dst, h->n2));
    else
        copy_set(h, dst, slot<double>(h, S0_), 7 * h->n2);
}
// Step input j of the segment: the state slots in the handle's form.
void *seg_step(moist_handle *h, int j) {
    return h->adj->seg + h->adj->seg_bytes * ((size_t)h->adj->half * h->adj->ckpt + j);
}
// Record block j of the current half.
char *rec_block(moist_handle *h, int j) {
    return h->adj->recs + h->adj->rb_bytes * ((size_t)h->adj->half * h->adj->nblocks + j);
}
// Reverse step j of the segment (its record block is current): FP64 input
// as stored, or in the mixed form as its recording pass widened it.
void reverse_step(moist_handle *h, int j) {
    if (h->precision == MOIST_MIXED) {
        enqueue_step_adj<float>(h, rec_in(h, 0), seg_sfh(h, j));
    } else {
        enqueue_step_adj<double>(h, (const double *)seg_step(h, j), seg_sfh(h, j));
    }
}
// Recording pass of step j of the segment (the mixed form from the handle's
// state, which the pass advances; the FP64 form from the stored input, and
// with `advance` its output becomes the state).
void record_step(moist_handle *h, int j, bool advance) {
    if (h->precision == MOIST_MIXED) {
        enqueue_step_rec_mixed(h);
    } else {
        enqueue_step_rec64(h, (const double *)seg_step(h, j));
        if (advance) copy_set(h, slot<double>(h, S0_), h->adj->SC, 7 * h->n2);
    }
}

// Step j of a segment's recording pass with segment records: the step input
// saved, then the step recorded (which advances the state).
void seg_record(moist_handle *h, int j) {
    AdjState &A = *h->adj;
    // the mixed form's reverse step reads its input from the record block
    if (h->precision != MOIST_MIXED) copy_bytes(h, seg_step(h, j), h->pool, 7 * h->n2 * h->elem);
    copy_set(h, seg_sfh(h, j), h->sfh, 2 * h->nm);
    A.cur = rec_block(h, j);
    record_step(h, j, true);
}
// Reverse step j of a segment from its records, or, with per-step records,
// its recording pass first (from the stored step input).
void seg_reverse(moist_handle *h, int j) {
    AdjState &A = *h->adj;
    if (A.segrec) {
        A.cur = rec_block(h, j);
    } else {
        if (h->precision == MOIST_MIXED) copy_bytes(h, h->pool, seg_step(h, j), 7 * h->n2 * h->elem);
        copy_set(h, h->sfh, seg_sfh(h, j), 2 * h->nm);
        A.cur = A.recs;
        record_step(h, j, false);
    }
    reverse_step(h, j);
}

template <class H> int adjoint_run_t(moist_handle *h, int nsteps, int ckpt, const void *const *tgt, double *J) {
    if (adj_alloc(h, nsteps, ckpt)) return -1;
    AdjState &A = *h->adj;
    const size_t n2 = h->n2;
    // theta targets
    for (int m = 0; m < h->nm; ++m) {
        const Dom &D = h->dom_h[m];
        const int ka = D.nz, ia = D.nx + 2 * D.halo;
        HIPCHK(h, hipMemcpyAsync(h->staging, tgt[m], (size_t)ka * ia * sizeof(H), hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL((k_upload<double, H>), dim3((D.nx + 255) / 256, ka), dim3(256), 0, h->stream, D,
                           (const H *)h->staging, A.tgt, (const double *)nullptr, D.halo, ka);
        HIPCHK(h, hipGetLastError());
    }
    ScopedEvents ev;
    HIPCHK(h, ev.create(3));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    if (reset_nan(h)) return -1;
    HIPCHK(h, hipEventRecord(e0, h->stream));
    // forward with checkpoints
    for (int n = 0; n < nsteps; ++n) {
        if (n % ckpt == 0) save_state(h, A.cp + A.cp_stride * (n / ckpt));
        if (run_steps(h, 1)) return -1;
    }
    HIPCHK(h, hipEventRecord(e1, h->stream));
    int code = decode_nan(h);
    if (code) return code;
    save_state(h, A.fin);
    // loss and the adjoint of the final state
    fill(h, A.lam, 7 * n2);
    fill(h, A.gk, 3 * n2);
    state64(h, A.SC);
    hipLaunchKernelGGL(ka_loss, dim3(h->nm), dim3(256), 0, h->stream, h->dom_d, dset(A.SC, n2),
                       (const double *)h->prec, (const double *)A.tgt, A.lam + RHO * n2, A.lam + RT * n2, A.J);
    // reverse sweep over the segments; the last step's stale halo reads feed
    // no later step
    fill(h, A.carry, 2 * h->nm);
    const size_t state_bytes = 7 * n2 * h->elem;
    if (A.overlap && A.segrec) {
        // pipelined segments: segment sg - 1's recording pass (stream s2, the
        // half of segment sg + 1, once that segment's reverse sweep is done)
        // runs beside segment sg's reverse sweep (the handle's stream, once
        // sg's recording pass is done).  The two share no buffer: the
        // recording pass writes the state, the work arrays and its half's
        // segment buffers; the reverse sweep reads its half and writes the
        // adjoint arrays.
        hipStream_t sA = h->stream, sB = A.s2;
        auto record_segment = [&](int sg) -> int {
            const int n0 = sg * ckpt, n1 = std::min(nsteps, n0 + ckpt);
            A.half = sg & 1;
            h->stream = sB;
            load_state(h, A.cp + A.cp_stride * sg);
            for (int n = n0; n < n1; ++n) seg_record(h, n - n0);
            h->stream = sA;
            HIPCHK(h, hipEventRecord(A.ev_rec[sg & 1], sB));
            return 0;
        };
        HIPCHK(h, hipEventRecord(A.ev_go, sA));
        HIPCHK(h, hipStreamWaitEvent(sB, A.ev_go, 0));
        if (record_segment(A.ncp - 1)) return -1;
        for (int sg = A.ncp - 1; sg >= 0; --sg) {
            const int n0 = sg * ckpt, n1 = std::min(nsteps, n0 + ckpt);
            if (sg > 0) {
                if (sg + 1 < A.ncp) HIPCHK(h, hipStreamWaitEvent(sB, A.ev_rev[(sg + 1) & 1], 0));
                if (record_segment(sg - 1)) return -1;
            }
            HIPCHK(h, hipStreamWaitEvent(sA, A.ev_rec[sg & 1], 0));
            A.half = sg & 1;
            for (int n = n1 - 1; n >= n0; --n) seg_reverse(h, n - n0);
            HIPCHK(h, hipEventRecord(A.ev_rev[sg & 1], sA));
            HIPCHK(h, hipGetLastError());
        }
        A.half = 0;
    }
    if (A.overlap && !A.segrec) {
        // pipelined per-step records: the states of a segment (the handle's
        // stream), then per step in reverse its recording pass on stream s2
        // (block j mod 2, once the reverse of step j + 2 has read that block)
        // beside the reverse of step j + 1.  Both passes of the pool (the state
        // pass and the recording passes) belong to one segment at a time: the
        // next segment's state pass follows the reverse of this one's step 0,
        // which waited for its recording pass.
        hipStream_t sA = h->stream, sB = A.s2;
        auto record_async = [&](int j) -> int {
            h->stream = sB;
            if (h->precision == MOIST_MIXED) copy_bytes(h, h->pool, seg_step(h, j), state_bytes);
            copy_set(h, h->sfh, seg_sfh(h, j), 2 * h->nm);
            A.cur = rec_block(h, j & 1);
            record_step(h, j, false);
            h->stream = sA;
            HIPCHK(h, hipEventRecord(A.ev_rec[j & 1], sB));
            return 0;
        };
        for (int sg = A.ncp - 1; sg >= 0; --sg) {
            const int n0 = sg * ckpt, n1 = std::min(nsteps, n0 + ckpt), J = n1 - n0;
            load_state(h, A.cp + A.cp_stride * sg);
            for (int n = n0; n < n1; ++n) {
                copy_bytes(h, seg_step(h, n - n0), h->pool, state_bytes);
                copy_set(h, seg_sfh(h, n - n0), h->sfh, 2 * h->nm);
                if (n + 1 >= n1) continue;
                if (h->precision == MOIST_MIXED) {
                    A.cur = rec_block(h, 0);
                    record_step(h, n - n0, true);
                } else if (run_steps(h, 1)) {
                    return -1;
                }
            }
            HIPCHK(h, hipEventRecord(A.ev_go, sA));
            HIPCHK(h, hipStreamWaitEvent(sB, A.ev_go, 0));
            if (record_async(J - 1)) return -1;
            for (int j = J - 1; j >= 0; --j) {
                if (j > 0) {
                    if (j + 1 <= J - 1) HIPCHK(h, hipStreamWaitEvent(sB, A.ev_rev[(j - 1) & 1], 0));
                    if (record_async(j - 1)) return -1;
                }
                HIPCHK(h, hipStreamWaitEvent(sA, A.ev_rec[j & 1], 0));
                A.cur = rec_block(h, j & 1);
                reverse_step(h, j);
                HIPCHK(h, hipEventRecord(A.ev_rev[j & 1], sA));
            }
            HIPCHK(h, hipGetLastError());
        }
    }
    for (int sg = A.overlap ? -1 : A.ncp - 1; sg >= 0; --sg) {
        const int n0 = sg * ckpt, n1 = std::min(nsteps, n0 + ckpt);
        load_state(h, A.cp + A.cp_stride * sg);
        if (A.segrec) {
            // one recording pass over the segment, then the reverse sweep from the records
            for (int n = n0; n < n1; ++n) seg_record(h, n - n0);
        } else {
            // states of the segment, then per step (in reverse) its record and adjoint.
            // The mixed form advances by its recording pass (into the one record
            // block, rewritten later): those kernels, not the step's level-pair
            // ones, make the states the reverse steps start from, as with segment
            // records, so both modes replay the same values (the two layouts'
            // kernels round differently under FP contraction).
            for (int n = n0; n < n1; ++n) {
                copy_bytes(h, seg_step(h, n - n0), h->pool, state_bytes);
                copy_set(h, seg_sfh(h, n - n0), h->sfh, 2 * h->nm);
                if (n + 1 >= n1) continue;
                if (h->precision == MOIST_MIXED) {
                    A.cur = A.recs;
                    record_step(h, n - n0, true);
                } else if (run_steps(h, 1)) {
                    return -1;
                }
            }
        }
        for (int n = n1 - 1; n >= n0; --n) seg_reverse(h, n - n0);
        HIPCHK(h, hipGetLastError());
    }
    // the first step's stale halo reads are the initial state's halo columns,
    // taken as periodic copies of its interior
    if (h->any_sf)
        hipLaunchKernelGGL(ka_stale_add, dim3((h->nm + 63) / 64), dim3(64), 0, h->stream, h->dom_d, h->nm,
                           dset(A.lam, n2), (const double *)A.carry);
    hipLaunchKernelGGL(ka_reduce_k, dim3(h->nm), dim3(256), 0, h->stream, h->dom_d, (const double *)A.gk,
                       (const double *)(A.gk + n2), (const double *)(A.gk + 2 * n2), A.gK);
    HIPCHK(h, hipEventRecord(e2, h->stream));
    load_state(h, A.fin);
    HIPCHK(h, hipMemcpyAsync(A.J_h.data(), A.J, h->nm * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(A.gK_h.data(), A.gK, 3 * h->nm * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipGetLastError());
    float f1 = 0.f, f2 = 0.f;
    hipEventElapsedTime(&f1, e0, e1);
    hipEventElapsedTime(&f2, e1, e2);
    A.ms_fwd = f1;
    A.ms_bwd = f2;
    if (J)
        for (int m = 0; m < h->nm; ++m) J[m] = A.J_h[m];
    h->halo_src = -1;
    return decode_nan(h);
}

}  // namespace

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

int moist_create(const moist_params *const *members, int n_members, int precision, moist_handle **out) {
    if (!out) return -2;
    *out = nullptr;
    if (!members || n_members < 1 || (precision != MOIST_FP64 && precision != MOIST_MIXED)) return -2;
    auto h = std::make_unique<moist_handle>();
    h->precision = precision;
    h->nm = n_members;
    h->elem = precision == MOIST_MIXED ? 4 : 8;
    std::vector<double> zh, zfh, dzh;
    int64_t off2 = 0, offx = 0, offz = 0;
    for (int m = 0; m < n_members; ++m) {
        const moist_params *p = members[m];
        if (!p || p->nx < 2 || p->nz < 1 || p->halo < 1 || p->halo > p->nx || p->ns < 1 || !p->z || !p->zf ||
            !p->dz) {
            h->err = "invalid member params";
            *out = h.release();
            return -2;
        }
        // kernels address a member's cells (nz padded to even) by 32-bit byte
        // offsets from its base: a field of one member stays below 2^32 bytes
        if ((int64_t)p->nx * (p->nz + 1) * h->elem > (int64_t)UINT32_MAX) {
            h->err = "member too large: nx * (nz + 1) * (4 or 8 bytes) must stay below 2^32 bytes";
            *out = h.release();
            return -2;
        }
        if (m == 0) h->ns = p->ns;
        if (p->ns != h->ns) {
            h->err = "all members must share ns";
            *out = h.release();
            return -2;
        }
        Dom D{};
        D.nx = p->nx;
        D.nz = p->nz;
        D.halo = p->halo;
        D.ns = p->ns;
        D.flags = (p->do_mp ? F_MP : 0) | (p->do_auto ? F_AUTO : 0) | (p->do_accr ? F_ACCR : 0) |
                  (p->do_revap ? F_REVAP : 0) | (p->do_sed ? F_SED : 0) | (p->do_rd ? F_RD : 0) |
                  (p->do_sf ? F_SF : 0);
        D.lin_col = ((p->nx - p->halo) % p->nx + p->nx) % p->nx;
        D.off2 = off2;
        D.offx = offx;
        D.offz = offz;
        D.dx = p->dx;
        D.dt = p->dt;
        D.H = p->H;
        D.nu = p->nu;
        D.kappa = p->kappa;
        D.gamma_d = p->gamma_d;
        D.beta = p->beta_offcenter;
        D.zf_sponge = p->zf_sponge;
        D.Cd = p->Cd;
        D.Ch = p->Ch;
        D.Ce = p->Ce;
        D.uabs_min = p->uabs_min;
        D.temp_surf = p->temp_surf;
        D.k_abs_qv = p->k_abs_qv;
        D.k_abs_qc = p->k_abs_qc;
        D.k_abs_qr = p->k_abs_qr;
        D.surf_emissivity = p->surf_emissivity;
        D.tau_cooling = p->tau_cooling;
        D.z_tropopose = p->z_tropopose;
        D.K_auto = p->K_auto;
        D.Qc_threshold = p->Qc_threshold;
        D.K_accr = p->K_accr;
        h->any_sf |= p->do_sf;
        h->any_rd |= p->do_rd;
        for (int k = 0; k < p->nz; ++k) {
            zh.push_back(p->z[k]);
            zfh.push_back(p->zf[k]);
            dzh.push_back(p->dz[k]);
        }
        // an odd nz is padded by one level: the level-pair layout of the
        // acoustic's work fields stores the top level with a partner
        off2 += (int64_t)p->nx * (p->nz + (p->nz & 1));
        offx += p->nx;
        offz += p->nz;
        h->max_ka = std::max(h->max_ka, p->nz);
        h->max_ia = std::max(h->max_ia, p->nx + 2 * p->halo);
        h->dom_h.push_back(D);
    }
    h->n2 = off2;
    h->nxs = offx;
    h->nzs = offz;
    h->ns_stage = {std::max(1, (h->ns + 2) / 3), std::max(1, (h->ns + 1) / 2), h->ns};

    std::vector<int4> t2, ta, ts;
    std::vector<int2> tc, wc;
    const int acL = precision == MOIST_MIXED ? ac_levels<float>() : ac_levels<double>();
    for (int m = 0; m < n_members; ++m) {
        const Dom &D = h->dom_h[m];
        for (int k0 = 0; k0 < D.nz; k0 += TILE_Z)
            for (int i0 = 0; i0 < D.nx; i0 += TILE_X) t2.push_back(make_int4(m, i0, k0, 0));
        for (int k0 = 0; k0 < D.nz; k0 += ST_Z)
            for (int i0 = 0; i0 < D.nx; i0 += ST_X) ts.push_back(make_int4(m, i0, k0, 0));
        for (int i0 = 0; i0 < D.nx; i0 += COL_TILE) tc.push_back(make_int2(m, i0));
        // c' storage for the largest chunk count any acoustic shape can use (L = 2)
        int nch2 = 1;
        while (nch2 * 2 < D.nz) nch2 *= 2;
        h->nch_max = std::max(h->nch_max, std::min(nch2, AC_THREADS));
    }
    {
        int nch_def = 0;
        if (make_atiles(h->dom_h, acL, AC_THREADS, ta, nch_def)) {
            h->err = "nz too large for the acoustic kernel";
            *out = h.release();
            return -2;
        }
        h->nch_max = std::max(h->nch_max, nch_def);
        h->ac_L = acL;
        h->ac_NT = AC_THREADS;
    }
    t2 = xcd_order(t2, make_int4(-1, 0, 0, 0));
    ts = xcd_order(ts, make_int4(-1, 0, 0, 0));
    h->nstiles = (int)ts.size();
    tc = xcd_order(tc, make_int2(-1, 0));
    {
        // adjoint column solve: one wave per column for few columns (parallelism
        // over rows), 8 columns x 8 chunks or 16 x 4 per wave for many (coalescing)
        int64_t ncols = 0;
        for (const Dom &D : h->dom_h) ncols += D.nx;
        h->wch = ncols < 4096 ? 64 : ncols < 16384 ? 8 : 4;
        const int cw = 64 / h->wch;
        for (int m = 0; m < n_members; ++m)
            for (int i0 = 0; i0 < h->dom_h[m].nx; i0 += cw) wc.push_back(make_int2(m, i0));
        wc = xcd_order(wc, make_int2(-1, 0));
        h->nwgroups = (int)wc.size();
    }
    std::vector<int2> acol;  // one entry per column of every member
    for (int m = 0; m < n_members; ++m)
        for (int i = 0; i < h->dom_h[m].nx; ++i) acol.push_back(make_int2(m, i));
    acol = xcd_order(acol, make_int2(-1, 0));
    h->nacols = (int)acol.size();
    h->natiles = (int)ta.size();
    h->ntiles2d = (int)t2.size();
    h->nctiles = (int)tc.size();

    moist_handle *hp = h.get();
    // Device set-up.  On a failure the handle is still returned (with the
    // error text), so that the caller can read it and moist_destroy frees
    // the stream and whatever was allocated before the failure.
    auto setup = [&]() -> int {
#define ALLOC(ptr, bytes)                                          \
    do {                                                           \
        HIPCHK(hp, hipMalloc((void **)&(ptr), (bytes)));           \
        hp->dev_bytes += (int64_t)(bytes);                         \
    } while (0)
    HIPCHK(hp, hipGetDevice(&hp->device));
    if (hp->device < 0 || hp->device >= MAX_DEVICES) {
        hp->err = "device index beyond the scratch table";
        return -2;
    }
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch[hp->device]) {
            void *ptr = nullptr;
            HIPCHK(hp, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            g_scratch[hp->device] = (double *)ptr;
        }
    }
    HIPCHK(hp, hipStreamCreate(&hp->stream));
    ALLOC(hp->dom_d, sizeof(Dom) * n_members);
    ALLOC(hp->tiles2d, sizeof(int4) * t2.size());
    ALLOC(hp->ctiles, sizeof(int2) * tc.size());
    ALLOC(hp->atiles, sizeof(int4) * ta.size());
    ALLOC(hp->stiles, sizeof(int4) * ts.size());
    ALLOC(hp->wgroups, sizeof(int2) * wc.size());
    ALLOC(hp->acols, sizeof(int2) * acol.size());
    HIPCHK(hp, hipMemcpy(hp->acols, acol.data(), sizeof(int2) * acol.size(), hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->wgroups, wc.data(), sizeof(int2) * wc.size(), hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->stiles, ts.data(), sizeof(int4) * ts.size(), hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->atiles, ta.data(), sizeof(int4) * ta.size(), hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->dom_d, hp->dom_h.data(), sizeof(Dom) * n_members, hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->tiles2d, t2.data(), sizeof(int4) * t2.size(), hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->ctiles, tc.data(), sizeof(int2) * tc.size(), hipMemcpyHostToDevice));
    double **zarr[] = {&hp->z, &hp->zf, &hp->dz, &hp->theta_ref, &hp->rho_bar, &hp->rt_bar, &hp->p_bar, &hp->L_bar,
                       &hp->B_bar};
    for (double **a : zarr) {
        ALLOC(*a, sizeof(double) * offz);
        HIPCHK(hp, hipMemset(*a, 0, sizeof(double) * offz));
    }
    HIPCHK(hp, hipMemcpy(hp->z, zh.data(), 8 * offz, hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->zf, zfh.data(), 8 * offz, hipMemcpyHostToDevice));
    HIPCHK(hp, hipMemcpy(hp->dz, dzh.data(), 8 * offz, hipMemcpyHostToDevice));
    {
        std::vector<double4> g(offz);
        for (int m = 0; m < n_members; ++m) {
            const Dom &D = hp->dom_h[m];
            for (int k = 0; k < D.nz; ++k) {
                const size_t q = D.offz + k;
                if (k < D.nz - 1)
                    g[q] = make_double4(1.0 / dzh[q], 1.0 / (zh[q + 1] - zh[q]), 1.0 / (dzh[q] + dzh[q + 1]),
                                        9.81 / (dzh[q] + dzh[q + 1]));
                else
                    g[q] = make_double4(1.0 / dzh[q], 0.0, 0.0, 0.0);
            }
        }
        ALLOC(hp->geo, sizeof(double4) * offz);
        HIPCHK(hp, hipMemcpy(hp->geo, g.data(), sizeof(double4) * offz, hipMemcpyHostToDevice));
    }
    const size_t pool_bytes = (size_t)NSLOT * hp->n2 * hp->elem;
    ALLOC(hp->pool, pool_bytes);
    HIPCHK(hp, hipMemset(hp->pool, 0, pool_bytes));
    ALLOC(hp->pool1, (size_t)NSLOT1 * offx * hp->elem);
    HIPCHK(hp, hipMemset(hp->pool1, 0, (size_t)NSLOT1 * offx * hp->elem));
    ALLOC(hp->cin, sizeof(double) * (size_t)hp->nch_max * offx);
    HIPCHK(hp, hipMemset(hp->cin, 0, sizeof(double) * (size_t)hp->nch_max * offx));
    ALLOC(hp->prec, sizeof(double) * offx);
    HIPCHK(hp, hipMemset(hp->prec, 0, sizeof(double) * offx));
    ALLOC(hp->sfh, sizeof(double) * 2 * n_members);
    HIPCHK(hp, hipMemset(hp->sfh, 0, sizeof(double) * 2 * n_members));
    ALLOC(hp->nan_key, sizeof(int));
    ALLOC(hp->step_ctr, sizeof(int));
    HIPCHK(hp, hipMemset(hp->step_ctr, 0, sizeof(int)));
    ALLOC(hp->staging, (size_t)hp->max_ka * hp->max_ia * 8);
    ALLOC(hp->diag_out, sizeof(double) * 4);
#undef ALLOC
    if (reset_nan(hp)) return -1;
    HIPCHK(hp, hipStreamSynchronize(hp->stream));
    return 0;
    };
    const int rc = setup();
    *out = h.release();
    return rc;
}

int moist_upload(moist_handle *h, int member, const void *const fields[7], const void *prec, const void *theta_ref,
                 int elem_bytes) {
    if (!h || !fields) return -2;
    if (int rc = check_member(h, member)) return rc;
    for (int f = 0; f < 7; ++f)
        if (!fields[f]) return -2;
    const bool mixed = h->precision == MOIST_MIXED;
    if (elem_bytes == 8)
        return mixed ? upload_t<float, double>(h, member, fields, prec, theta_ref)
                     : upload_t<double, double>(h, member, fields, prec, theta_ref);
    if (elem_bytes == 4)
        return mixed ? upload_t<float, float>(h, member, fields, prec, theta_ref)
                     : upload_t<double, float>(h, member, fields, prec, theta_ref);
    h->err = "elem_bytes must be 4 or 8";
    return -2;
}

int moist_download(moist_handle *h, int member, void *const fields[7], void *prec, int elem_bytes) {
    if (!h || !fields) return -2;
    if (int rc = check_member(h, member)) return rc;
    const bool mixed = h->precision == MOIST_MIXED;
    if (elem_bytes == 8)
        return mixed ? download_t<float, double>(h, member, fields, prec) : download_t<double, double>(h, member, fields, prec);
    if (elem_bytes == 4)
        return mixed ? download_t<float, float>(h, member, fields, prec) : download_t<double, float>(h, member, fields, prec);
    h->err = "elem_bytes must be 4 or 8";
    return -2;
}

int moist_download_phys(moist_handle *h, int member, void *const out2d[6], void *const out1d[6], int elem_bytes) {
    if (!h) return -2;
    if (int rc = check_member(h, member)) return rc;
    const bool mixed = h->precision == MOIST_MIXED;
    if (elem_bytes == 8)
        return mixed ? download_phys_t<float, double>(h, member, out2d, out1d)
                     : download_phys_t<double, double>(h, member, out2d, out1d);
    if (elem_bytes == 4)
        return mixed ? download_phys_t<float, float>(h, member, out2d, out1d)
                     : download_phys_t<double, float>(h, member, out2d, out1d);
    h->err = "elem_bytes must be 4 or 8";
    return -2;
}

int moist_rk3_step(moist_handle *h, int nsteps) {
    if (!h || nsteps < 0) return -2;
    if (reset_nan(h)) return -1;
    if (run_steps(h, nsteps)) return -1;
    return decode_nan(h);
}

int moist_compute_step(moist_handle *h, int member, const void *const phys[6], const void *const s0[7], double dt,
                       int ns, int elem_bytes) {
    if (!h || !phys || !s0 || ns < 1 || !(dt > 0.0)) return -2;
    if (int rc = check_member(h, member)) return rc;
    if (h->nm != 1) {
        h->err = "moist_compute_step takes a single-domain handle";
        return -2;
    }
    const bool mixed = h->precision == MOIST_MIXED;
    if (elem_bytes == 8)
        return mixed ? compute_step_t<float, double>(h, member, phys, s0, dt, ns)
                     : compute_step_t<double, double>(h, member, phys, s0, dt, ns);
    if (elem_bytes == 4)
        return mixed ? compute_step_t<float, float>(h, member, phys, s0, dt, ns)
                     : compute_step_t<double, float>(h, member, phys, s0, dt, ns);
    h->err = "elem_bytes must be 4 or 8";
    return -2;
}

int moist_diagnostics(moist_handle *h, int member, double *out) {
    if (!h || !out) return -2;
    if (int rc = check_member(h, member)) return rc;
    const ZVec zv = zvec(h);
    if (!h->hist_part) HIPCHK(h, hipMalloc((void **)&h->hist_part, sizeof(double) * HS_COUNT * h->ntiles2d));
    if (!h->diag_all) HIPCHK(h, hipMalloc((void **)&h->diag_all, sizeof(double) * 4 * h->nm));
    if (h->precision == MOIST_MIXED)
        hipLaunchKernelGGL(k_budget<float>, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d,
                           h->tiles2d, zv, sset<float>(h, S0_), h->hist_part);
    else
        hipLaunchKernelGGL(k_budget<double>, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d,
                           h->tiles2d, zv, sset<double>(h, S0_), h->hist_part);
    hipLaunchKernelGGL(k_budget_reduce, dim3(h->nm), dim3(256), 0, h->stream, h->dom_d, (const int4 *)h->tiles2d,
                       h->ntiles2d, (const double *)h->hist_part, (const double *)h->prec, h->diag_all);
    HIPCHK(h, hipGetLastError());
    double v[4];
    HIPCHK(h, hipMemcpyAsync(v, h->diag_all + 4 * member, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    int key = INT_MAX;
    HIPCHK(h, hipMemcpyAsync(&key, h->nan_key, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    out[MOIST_DIAG_DRY_MASS] = v[0];
    out[MOIST_DIAG_TOTAL_WATER] = v[1];
    out[MOIST_DIAG_RAIN] = v[2];
    out[MOIST_DIAG_MASS] = v[3];
    out[MOIST_DIAG_NAN] = key == INT_MAX ? 0.0 : double(key & 15);
    return 0;
}

int moist_diagnostics_noalloc(moist_handle *h, int member, double *out) {
    if (!h || !out) return -2;
    if (int rc = check_member(h, member)) return rc;
    const size_t need = 3 * (size_t)h->ntiles2d + 4 * (size_t)h->nm;
    if (need > (size_t)MOIST_NOALLOC_SCRATCH_DOUBLES) {
        h->err = "moist_diagnostics_noalloc: the handle needs " + std::to_string(need) +
                 " doubles of scratch, above the static maximum of " +
                 std::to_string((size_t)MOIST_NOALLOC_SCRATCH_DOUBLES);
        return -2;
    }
    double *part = g_scratch[h->device];
    if (!part) {
        h->err = "moist_diagnostics_noalloc: no scratch address for this device";
        return -1;
    }
    double *sums = part + 3 * (size_t)h->ntiles2d;
    // every partial and sum is written before it is read within the call,
    // so nothing needs clearing
    const ZVec zv = zvec(h);
    if (h->precision == MOIST_MIXED)
        hipLaunchKernelGGL(k_budget<float>, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d,
                           h->tiles2d, zv, sset<float>(h, S0_), part);
    else
        hipLaunchKernelGGL(k_budget<double>, dim3(h->ntiles2d), dim3(TILE_X, TILE_Z), 0, h->stream, h->dom_d,
                           h->tiles2d, zv, sset<double>(h, S0_), part);
    hipLaunchKernelGGL(k_budget_reduce, dim3(h->nm), dim3(256), 0, h->stream, h->dom_d, (const int4 *)h->tiles2d,
                       h->ntiles2d, (const double *)part, (const double *)h->prec, sums);
    HIPCHK(h, hipGetLastError());
    double v[4];
    HIPCHK(h, hipMemcpyAsync(v, sums + 4 * member, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    int key = INT_MAX;
    HIPCHK(h, hipMemcpyAsync(&key, h->nan_key, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    out[MOIST_DIAG_DRY_MASS] = v[0];
    out[MOIST_DIAG_TOTAL_WATER] = v[1];
    out[MOIST_DIAG_RAIN] = v[2];
    out[MOIST_DIAG_MASS] = v[3];
    out[MOIST_DIAG_NAN] = key == INT_MAX ? 0.0 : double(key & 15);
    return 0;
}

int moist_use_graph(moist_handle *h, int on) {
    if (!h) return -2;
    h->use_graph = on ? 1 : 0;
    return 0;
}

int moist_set_option(moist_handle *h, int option, int value) {
    if (!h) return -2;
    if (option == MOIST_OPT_PHYS_DIAG) {
        h->store_diag = value ? 1 : 0;
        if (h->gexec) {
            hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
        return 0;
    }
    if (option == MOIST_OPT_ADJ_COLUMN_KERNEL) {
        h->adj_col = value ? 1 : 0;
        return 0;
    }
    if (option == MOIST_OPT_ADJ_RECORDS) {
        if (value < 0 || value > 4) {
            h->err = "adjoint record mode must be 0 to 4";
            return -2;
        }
        h->adj_mode = value;
        return 0;
    }
    h->err = "unknown option";
    return -2;
}

int moist_time_steps(moist_handle *h, int nsteps, double *ms) {
    if (!h || !ms || nsteps < 1) return -2;
    ScopedEvents ev;
    HIPCHK(h, ev.create(2));
    hipEvent_t a = ev.e[0], b = ev.e[1];
    if (reset_nan(h)) return -1;
    if (h->use_graph && build_graph(h)) return -1;
    HIPCHK(h, hipEventRecord(a, h->stream));
    if (run_steps(h, nsteps)) return -1;
    HIPCHK(h, hipEventRecord(b, h->stream));
    HIPCHK(h, hipEventSynchronize(b));
    float f = 0.0f;
    HIPCHK(h, hipEventElapsedTime(&f, a, b));
    *ms = f;
    if (h->ev_on) {
        double tot = 0.0;
        int64_t n = 0;
        for (int e = 0; e + 1 < h->ev_used; e += 2) {
            float d = 0.0f;
            if (hipEventElapsedTime(&d, h->ev_pool[e], h->ev_pool[e + 1]) == hipSuccess) {
                tot += d;
                ++n;
            }
        }
        h->ac_ms = n ? tot / n : 0.0;
        h->ac_launches = n;
    }
    return decode_nan(h);
}

int moist_kernel_events(moist_handle *h, int on) {
    if (!h) return -2;
    h->ev_on = on ? 1 : 0;
    return 0;
}

int moist_kernel_stats(moist_handle *h, double *acoustic_ms_per_launch, int64_t *acoustic_launches) {
    if (!h) return -2;
    if (acoustic_ms_per_launch) *acoustic_ms_per_launch = h->ac_ms;
    if (acoustic_launches) *acoustic_launches = h->ac_launches;
    return 0;
}

int moist_autotune(moist_handle *h, int reps, int *levels, int *threads, double *us_per_launch) {
    if (!h || reps < 1) return -2;
    // shapes: {L, NT, W}; W = 8 asks for registers that let 8 waves share a SIMD
    const int shapes[10][3] = {{2, 256, 1}, {2, 512, 1}, {2, 1024, 1}, {4, 256, 1}, {4, 512, 1},
                               {4, 1024, 1}, {2, 512, 8}, {2, 1024, 8}, {4, 512, 8}, {4, 1024, 8}};
    const int L0 = h->ac_L, NT0 = h->ac_NT, W0 = h->ac_W;
    int bestL = L0, bestNT = NT0, bestW = W0;
    double best = 1e300, t0 = 1e300;
    ScopedEvents ev;
    HIPCHK(h, ev.create(2));
    hipEvent_t a = ev.e[0], b = ev.e[1];
    const ZVec zv = zvec(h);
    for (const auto &sh : shapes) {
        const int L = sh[0], NT = sh[1], W = sh[2];
        {
            if (set_acoustic_shape(h, L, NT, W)) continue;  // shape does not fit every member
            // one stage's substeps on the current state and work arrays (outputs are work arrays only)
            auto run = [&]() {
                for (int n = 0; n < reps; ++n) {
                    if (h->precision == MOIST_MIXED) {
                        const Work<float> w = work<float>(h);
                        AcBuf<float> ab{slot<float>(h, RUP), slot<float>(h, RUP2), slot<float>(h, RTP0),
                                        slot<float>(h, RTP0), slot<float>(h, RTP1), h->cin, h->nxs, w.rwp};
                        launch_acoustic<float, true>(h, zv, sset<float>(h, S0_), w, ab, 1, 0.0, 1, n == 0 ? 1 : 0);
                    } else {
                        const Work<double> w = work<double>(h);
                        AcBuf<double> ab{slot<double>(h, RUP), slot<double>(h, RUP2), slot<double>(h, RTP0),
                                         slot<double>(h, RTP0), slot<double>(h, RTP1), h->cin, h->nxs, w.rwp};
                        launch_acoustic<double, false>(h, zv, sset<double>(h, S0_), w, ab, 1, 0.0, 1,
                                                       n == 0 ? 1 : 0);
                    }
                }
            };
            run();  // warm-up
            HIPCHK(h, hipEventRecord(a, h->stream));
            run();
            HIPCHK(h, hipEventRecord(b, h->stream));
            HIPCHK(h, hipEventSynchronize(b));
            HIPCHK(h, hipGetLastError());
            float ms = 0.f;
            HIPCHK(h, hipEventElapsedTime(&ms, a, b));
            if (L == L0 && NT == NT0 && W == W0) t0 = ms;
            if (ms < best) {
                best = ms;
                bestL = L;
                bestNT = NT;
                bestW = W;
            }
        }
    }
    // keep the current shape unless another is faster by more than 3 %
    if (!(best < 0.97 * t0)) {
        bestL = L0;
        bestNT = NT0;
        bestW = W0;
        best = t0;
    }
    if (int rc = set_acoustic_shape(h, bestL, bestNT, bestW)) return rc;
    h->tuned = 1;
    if (levels) *levels = bestL;
    if (threads) *threads = bestNT * (bestW == 8 ? -1 : 1);
    if (us_per_launch) *us_per_launch = best * 1e3 / reps;
    return 0;
}

int moist_set_acoustic_shape(moist_handle *h, int levels, int threads) {
    if (!h) return -2;
    // threads < 0: the same workgroup size with registers for 8 waves per SIMD
    const int W = threads < 0 ? 8 : 1, NT = threads < 0 ? -threads : threads;
    const bool ok = (levels == 2 || levels == 4) &&
                    (W == 1 ? (NT == 256 || NT == 512 || NT == 1024) : (NT == 512 || NT == 1024));
    if (!ok) {
        h->err = "acoustic shape: levels 2 or 4; threads 256, 512 or 1024, or -512 / -1024 for 8 waves per SIMD";
        return -2;
    }
    return set_acoustic_shape(h, levels, NT, W);
}

int moist_adjoint_run(moist_handle *h, int nsteps, int ckpt, const void *const *theta_target, int elem_bytes,
                      double *J) {
    if (!h || !theta_target || nsteps < 1 || ckpt < 1) return -2;
    for (int m = 0; m < h->nm; ++m)
        if (!theta_target[m]) return -2;
    if (elem_bytes == 8) return adjoint_run_t<double>(h, nsteps, ckpt, theta_target, J);
    if (elem_bytes == 4) return adjoint_run_t<float>(h, nsteps, ckpt, theta_target, J);
    h->err = "elem_bytes must be 4 or 8";
    return -2;
}

int moist_adjoint_grad(moist_handle *h, int member, void *const grad[7], double *grad_params, int elem_bytes) {
    if (!h) return -2;
    if (int rc = check_member(h, member)) return rc;
    if (!h->adj) {
        h->err = "no adjoint run on this handle";
        return -2;
    }
    const Dom &D = h->dom_h[member];
    for (int f = 0; f < 7; ++f) {
        if (!grad || !grad[f]) continue;
        const double *src = h->adj->lam + (size_t)f * h->n2;
        int rc = elem_bytes == 8 ? download_field<double, double>(h, D, src, nullptr, grad[f], nullptr, false)
                 : elem_bytes == 4 ? download_field<double, float>(h, D, src, nullptr, grad[f], nullptr, false)
                                   : -2;
        if (rc) return rc;
    }
    if (grad_params)
        for (int j = 0; j < 3; ++j) grad_params[j] = h->adj->gK_h[3 * member + j];
    return 0;
}

int moist_history(moist_handle *h, int member, double *stats, void *const fields[13], int elem_bytes) {
    if (!h) return -2;
    if (int rc = check_member(h, member)) return rc;
    const bool mixed = h->precision == MOIST_MIXED;
    if (elem_bytes == 8)
        return mixed ? history_t<float, double>(h, member, stats, fields) : history_t<double, double>(h, member, stats, fields);
    if (elem_bytes == 4)
        return mixed ? history_t<float, float>(h, member, stats, fields) : history_t<double, float>(h, member, stats, fields);
    h->err = "elem_bytes must be 4 or 8";
    return -2;
}

int moist_device_bytes(moist_handle *h, int64_t *state_bytes, int64_t *adjoint_bytes) {
    if (!h) return -2;
    if (state_bytes) *state_bytes = h->dev_bytes;
    if (adjoint_bytes) *adjoint_bytes = h->adj ? (int64_t)h->adj->bytes : 0;
    return 0;
}

int moist_adjoint_times(moist_handle *h, double *ms_forward, double *ms_reverse) {
    if (!h || !h->adj) return -2;
    if (ms_forward) *ms_forward = h->adj->ms_fwd;
    if (ms_reverse) *ms_reverse = h->adj->ms_bwd;
    return 0;
}

const char *moist_last_error(const moist_handle *h) { return h ? h->err.c_str() : "null handle"; }

void moist_destroy(moist_handle *h) {
    if (!h) return;
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    for (auto e : h->ev_pool) hipEventDestroy(e);
    adj_free(h);
    void *ptrs[] = {h->dom_d, h->tiles2d, h->stiles, h->ctiles, h->wgroups, h->acols, h->hist_fields, h->hist_part, h->hist_stats, h->diag_all, h->z, h->zf, h->dz, h->geo, h->theta_ref, h->rho_bar, h->rt_bar,
                    h->p_bar, h->L_bar, h->B_bar, h->pool, h->pool1, h->atiles, h->cin, h->prec, h->sfh, h->nan_key,
                    h->step_ctr, h->staging, h->diag_out};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
