/*
 * moist2d.h — C ABI of libmoist2d_hip.so, the gfx950 time-stepping core of the
 * 2-D moist compressible C-grid model (split-explicit acoustic substeps with a
 * vertically implicit w–p tridiagonal solve, Koren scalar transport, Kessler
 * warm rain).
 *
 * The reference (seiya/moist2d, Julia) has no FFI: its step routine is the
 * plain method `rk3_step!(s::State, p::Params{T})` (moist.jl:1589) and, one
 * RK stage below it, `compute_step!(...)` (moist.jl:1039-1040).  The entry
 * points below are what a `ccall` layer in place of those methods binds.
 *
 * Conventions
 *  - Host arrays are in the reference layout: Julia column-major (ka, ia),
 *    z fastest, x halo columns included (moist.jl:319-345).  1-D x arrays have
 *    length ia = nx + 2*halo, 1-D z arrays length ka = nz.  Element type is
 *    double (elem_bytes == 8) or float (elem_bytes == 4); host buffers are
 *    only borrowed for the duration of the call.
 *  - Device layout is x fastest without halos: field[k*nx + i].  Periodic x
 *    is an index wrap inside the kernels (replaces exchange_halo!,
 *    moist.jl:233-241).
 *  - Return codes: 0 OK; 1..12 a NaN was found in a field, numbered in the
 *    order of the reference's checks (moist.jl:1543-1578): rho, rho_u, rho_w,
 *    rho_theta, rho_qv, rho_qc, rho_qr, d_rho_u_phys, d_rho_theta_phys,
 *    d_rho_qv_phys, d_rho_qc_phys, d_rho_qr_phys; < 0 HIP/runtime/argument
 *    error, with text from moist_last_error().
 *  - One handle owns one HIP stream and its device buffers; a handle is not
 *    reentrant, different handles are independent.
 */
#ifndef MOIST2D_H
#define MOIST2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision forms. */
#define MOIST_FP64  0  /* every field and every operation in double          */
#define MOIST_MIXED 1  /* FP32 prognostic deviations from an FP64 1-D base
                          state, FP32 stencils/microphysics, FP64 tridiagonal
                          elimination, Exner/base state and accumulators      */

/* Scalar-transport ordering.  The reference updates the scalars in place
 * while reading neighbours (moist.jl:1469-1518), so its result depends on
 * the loop order; the device always reads a snapshot of the stage input.   */
#define MOIST_ORDER_SNAPSHOT 0

/* Flattened Params{T} (moist.jl:70-222).  Values are passed as double; for
 * the reference's Float32 form they are the exact Float32 values widened.
 * z, zf, dz are the grid vectors of length nz built by the Params
 * constructor (moist.jl:203-216).  K_auto, Qc_threshold, K_accr are globals
 * in the reference (moist.jl:46-48) and parameters here.                   */
typedef struct moist_params {
    int32_t nx, nz, halo, ns;
    int32_t do_mp, do_auto, do_accr, do_revap, do_sed, do_rd, do_sf;
    int32_t reserved0;
    double dx, dt, H;
    double nu, kappa, gamma_d, beta_offcenter, zf_sponge;
    double Cd, Ch, Ce, uabs_min, temp_surf;
    double k_abs_qv, k_abs_qc, k_abs_qr, surf_emissivity, tau_cooling, z_tropopose;
    double K_auto, Qc_threshold, K_accr;
    const double *z, *zf, *dz;
} moist_params;

typedef struct moist_handle moist_handle;

/* Diagnostics returned by moist_diagnostics (FP64 device reductions over the
 * interior, weighted by cell area dx*dz[k]). */
#define MOIST_DIAG_DRY_MASS    0  /* sum (rho - rho_qv - rho_qc - rho_qr) dx dz */
#define MOIST_DIAG_TOTAL_WATER 1  /* sum (rho_qv + rho_qc + rho_qr) dx dz       */
#define MOIST_DIAG_RAIN        2  /* sum prec dx  (accumulated surface rain)     */
#define MOIST_DIAG_MASS        3  /* sum rho dx dz                               */
#define MOIST_DIAG_NAN         4  /* NaN field code of the last step (0 = none)  */
#define MOIST_DIAG_COUNT       5

/* Create a handle for an ensemble of n_members independent domains (n=1 for
 * a single domain).  Each member has its own params (nx, nz and grid may
 * differ; dt, ns and the switches may differ too).  Replaces allocate_state
 * (moist.jl:319) for the device.  On failure *out may still hold a handle
 * that carries the error text (moist_last_error); pass it to moist_destroy,
 * which frees whatever was allocated before the failure.                    */
int moist_create(const moist_params *const *members, int n_members,
                 int precision, moist_handle **out);

/* Upload one member's State: fields[0..6] = rho, rho_u, rho_w, rho_theta,
 * rho_qv, rho_qc, rho_qr (ka x ia), prec (ia), theta_ref (ka).  The halo
 * columns of rho_u and rho at the surface row seed the stale-halo reads of
 * surface_fluxes! (moist.jl:723,736).                                        */
int moist_upload(moist_handle *h, int member, const void *const fields[7],
                 const void *prec, const void *theta_ref, int elem_bytes);

/* Download one member's State.  Interior columns hold the state; halo
 * columns hold what the reference leaves there after a step (periodic
 * copies taken at the last stage's exchanges, moist.jl:1063-1069 and
 * 1463-1466; rho_theta's halo columns are not written).                     */
int moist_download(moist_handle *h, int member, void *const fields[7],
                   void *prec, int elem_bytes);

/* Download the physics diagnostics the reference keeps in State
 * (moist.jl:306-316), as computed at the start of the last step:
 * out2d[0..5] = d_rho_theta_mp, d_rho_qv_mp, d_rho_qc_mp, d_rho_qr_mp,
 * d_rho_qr_sed, d_rho_theta_rd (ka x ia, interior written);
 * out1d[0..5] = d_rho_u_sf, d_rho_theta_sf, d_rho_qv_sf, mflux, shflux,
 * lhflux (ia, interior written).  Any pointer may be NULL.                  */
int moist_download_phys(moist_handle *h, int member, void *const out2d[6],
                        void *const out1d[6], int elem_bytes);

/* Advance every member by nsteps large steps: the device equivalent of
 * nsteps calls of rk3_step!(s, p) (moist.jl:1589-1683).  Enqueues on the
 * handle's stream and synchronizes; returns the NaN code of the first field
 * found NaN (checked after each step, in the reference's order).           */
int moist_rk3_step(moist_handle *h, int nsteps);

/* One RK stage: compute_step!(d_rho_phys, d_rho_u_phys, d_rho_theta_phys,
 * d_rho_qv_phys, d_rho_qc_phys, d_rho_qr_phys, s, s0, p, dt, ns)
 * (moist.jl:1039-1580) on member `member`, with the handle's current state
 * as s.  phys[0..5] are the six tendency arrays (ka x ia) and s0[0..6] the
 * seven s0 fields (ka x ia), both in host layout.                           */
int moist_compute_step(moist_handle *h, int member, const void *const phys[6],
                       const void *const s0[7], double dt, int ns,
                       int elem_bytes);

/* FP64 reductions of the current state of one member into out[MOIST_DIAG_COUNT]. */
int moist_diagnostics(moist_handle *h, int member, double *out);

/* moist_diagnostics without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under stream or graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The block
 * partials go to a static device array of MOIST_NOALLOC_SCRATCH_DOUBLES
 * doubles (8 MiB, one per GPU, held until the process ends); a handle that
 * needs more (3 per 64x4 cell tile plus 4 per member: about 38 M cells)
 * is rejected with -2.  Calls on one GPU share that array: calls that overlap
 * in time on one GPU race, and must be serialized by the caller.  Same
 * results as moist_diagnostics, bit for bit.                               */
#define MOIST_NOALLOC_SCRATCH_DOUBLES (1 << 20)
int moist_diagnostics_noalloc(moist_handle *h, int member, double *out);

/* Capture one large step of all members as a hipGraph and replay it for
 * later moist_rk3_step calls (on = 1) or go back to direct launches (0).   */
int moist_use_graph(moist_handle *h, int on);

/* Options: MOIST_OPT_PHYS_DIAG (default 1) stores the State physics
 * diagnostics of moist_download_phys at every step; 0 skips those writes.
 * MOIST_OPT_ADJ_RECORDS selects how moist_adjoint_run keeps the forward
 * records of its reverse sweep: 1 per step (each step's record recomputed
 * from the segment's states); 2 whole segment; 3 whole segment, pipelined
 * (the records of two segments: one segment's recording pass runs on a
 * second stream beside the reverse sweep of the next); 4 per step, pipelined
 * (two record blocks: one step's recording pass on a second stream beside
 * the reverse of the step after it); 0 (default) the first of 3, 2, 4, 1
 * whose buffers fit in the free device memory less a margin (a tenth of it,
 * at least 8 GiB).  All give the same gradient.
 * MOIST_OPT_ADJ_COLUMN_KERNEL
 * (default 1) runs the vertical part of each reverse acoustic substep as one
 * kernel per column (members up to 257 levels); 0 runs it as three kernels.
 * The arithmetic is the same.                                              */
#define MOIST_OPT_PHYS_DIAG 0
#define MOIST_OPT_ADJ_RECORDS 1
#define MOIST_OPT_ADJ_COLUMN_KERNEL 2
int moist_set_option(moist_handle *h, int option, int value);

/* Time nsteps large steps with HIP events on the handle's stream (state
 * resident in HBM); *ms receives the elapsed milliseconds.                  */
int moist_time_steps(moist_handle *h, int nsteps, double *ms);

/* Record a HIP event pair around every acoustic column kernel launch during
 * moist_time_steps (on = 1); direct launches are used while it is on.     */
int moist_kernel_events(moist_handle *h, int on);

/* Average device duration (ms) of the acoustic column kernel launches of
 * the last moist_time_steps call, from HIP events around those launches.   */
int moist_kernel_stats(moist_handle *h, double *acoustic_ms_per_launch,
                       int64_t *acoustic_launches);

/* The derived fields and budget of output_history (moist.jl:1816-1858) for
 * one member, computed on the device in FP64 from the current state.
 * stats[7] (may be NULL): total energy (kinetic + potential + moist
 * internal, J/m per unit y, before the reference subtracts its first value),
 * w min, w max (m/s), qc max, qr max (g/kg), RH min, RH max (%).
 * fields[13] (NULL entries skipped): host arrays of nx*nz elements, x
 * fastest (the history file's layout): rho, rho_u, rho_w, u, w, theta, qv,
 * qc, qr, RH (%), temperature (K), pressure (hPa), theta_e (K). */
#define MOIST_HIST_FIELDS 13
#define MOIST_HIST_STATS 7
int moist_history(moist_handle *h, int member, double *stats, void *const fields[13], int elem_bytes);

/* Tile-shape autotuner (no counterpart in the reference, whose loops have
 * no tiles): times the acoustic substep kernel on this handle's domains for
 * every shape of L levels per chunk (2, 4) and workgroup size (256, 512,
 * 1024 threads, and 512 / 1024 with registers for 8 waves per SIMD) that fits
 * every member, `reps` launches each, and keeps the
 * fastest.  Outputs (may be NULL): the chosen L and threads and the
 * microseconds per launch.  Opt-in: different shapes combine the chunked
 * tridiagonal solve in a different order (results differ at rounding
 * level), so a handle keeps its default shape unless this is called. */
int moist_autotune(moist_handle *h, int reps, int *levels, int *threads, double *us_per_launch);
/* Sets the acoustic kernel's shape explicitly: levels 2 or 4; threads 256,
 * 512 or 1024, or -512 / -1024 for the same workgroup size compiled with
 * registers for 8 waves per SIMD (two or four workgroups per CU).  The
 * autotuner reports threads the same way. */
int moist_set_acoustic_shape(moist_handle *h, int levels, int threads);

/* Reverse mode (every physics switch, surface fluxes and gray radiation
 * included; the surface fluxes' halo reads at the domain's ends, copies of
 * the previous step's stage-3 input, are differentiated as such, and the
 * initial State's halo columns are taken as periodic copies).  Replaces
 * ForwardDiff / a ChainRulesCore rrule around the reference's driver loop
 * (run_sim, moist.jl:1930-1963, calling rk3_step!, moist.jl:1589): runs
 * nsteps forward from the uploaded state with checkpoints every `ckpt`
 * steps, evaluates per member
 *   J = dx * sum_i prec_i + 1/2 * sum_cells (rho_theta/rho - theta_target)^2
 * and sweeps back, recomputing each segment from its checkpoint and each
 * step's acoustic substeps.  theta_target[m]: host array (ka, ia) in the
 * reference layout (halo columns ignored).  J[m] receives the loss.  The
 * handle's state afterwards is the final forward state.  The trajectory
 * (forward pass, checkpoints, segment recomputation) and the forward records
 * are in the handle's form (FP32 records in the mixed form); the adjoint
 * arithmetic is FP64 in both, so a mixed-form handle returns the gradient of
 * its own loss, linearised at the mixed step's stage values. */
int moist_adjoint_run(moist_handle *h, int nsteps, int ckpt, const void *const *theta_target, int elem_bytes,
                      double *J);
/* dJ/d(initial rho, rho_u, rho_w, rho_theta, rho_qv, rho_qc, rho_qr) of one
 * member into host arrays (ka, ia) (interior columns written), and
 * dJ/d(K_auto, Qc_threshold, K_accr) into grad_params[3]; either may be NULL. */
int moist_adjoint_grad(moist_handle *h, int member, void *const grad[7], double *grad_params, int elem_bytes);
/* Device memory held by the handle: state pools and tables, and the
 * buffers of the last adjoint run (checkpoints, segment states, forward
 * record, adjoint fields). */
int moist_device_bytes(moist_handle *h, int64_t *state_bytes, int64_t *adjoint_bytes);
/* Device time of the last adjoint run: forward pass (with checkpoint copies)
 * and reverse sweep (segment recomputation included), in ms. */
int moist_adjoint_times(moist_handle *h, double *ms_forward, double *ms_reverse);

const char *moist_last_error(const moist_handle *h);
void moist_destroy(moist_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* MOIST2D_H */
