/*
 * ns_hip.h -- C ABI of the MI355X (gfx950) FP64 staggered-grid Navier-Stokes
 * time step.  Drop-in boundary for the Fortran host of NS_3D_Staggered_dp.
 *
 * Conventions
 *  - Every array is FP64, Fortran column-major, with the reference's ghosted
 *    shapes: u(nx+1,ny+2,nz+2), v(nx+2,ny+1,nz+2), w(nx+2,ny+2,nz+1),
 *    p,dp(nx+2,ny+2,nz+2) (big_periodic_3D.f90:14-18,27-28); interior-only
 *    arrays are (nx,ny,nz).
 *  - Pointers may be host or device (HIP unified addressing); host arrays are
 *    staged through the context.  Device-resident callers pay no copies.
 *  - A boundary-plane set is one contiguous FP64 vector: the 18 velocity
 *    planes in vel_bc_staggered's argument order (NS_functions.f90:213-217:
 *    bx_u_1, bx_u_nx, by_u_1, by_u_ny, bz_u_1, bz_u_nz, then v, then w), the
 *    6 pressure planes (bx_p_1 ... bz_p_nz), then the scalar mean_dp.  Its
 *    length is ns_plane_set_length().
 *  - Every entry point returns an int status (NS_OK = 0, negative on error);
 *    nothing aborts across the ABI (the reference stops with stop(1),
 *    re_simulation_sub_tstep_interBC_combine.f90:393-394).
 *  - One context per host thread; calls on one context are serialised on its
 *    HIP stream.  Contexts share no mutable state.
 *  - The context stream is non-blocking: device buffers passed in are read
 *    (and device outputs written) on it, so the caller completes the work
 *    that produces an input before the call (synchronise the producing
 *    stream, or make the context stream wait on an event), and reads a
 *    device output after ns_synchronize.  Host buffers need nothing.
 */
#ifndef NS_HIP_H
#define NS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define NS_OK 0
#define NS_ERR_INVALID (-1)     /* bad dimensions / codes / null pointer */
#define NS_ERR_UNSUPPORTED (-2) /* combination the reference does not implement either */
#define NS_ERR_HIP (-3)         /* a HIP runtime call failed */
#define NS_ERR_NONFINITE (-4)   /* NaN/Inf detected by ns_check_finite */
#define NS_ERR_NOMEM (-5)

#define NS_EULER 0
#define NS_AB2 1
#define NS_AB2_CN 2

/* dp = dp - dp(1,1,1) after the pressure solve, as the periodic driver does
 * (big_periodic_3D.f90:562). */
#define NS_FLAG_GAUGE_CORNER 1
/* Periodic CN rows use the wrap stencil.  Without it the periodic CN matrix
 * is mat_CN's (NS_functions.f90:1273-1278), whose row 1 couples to column
 * nx+1 (== u(1)) instead of nx. */
#define NS_FLAG_CN_CYCLIC 2
/* ns_step runs every step as individual kernel launches.  Without it, steady
 * steps (not the first, no plane capture active) are replayed two at a time
 * from a HIP graph captured once per context; the results are identical. */
#define NS_FLAG_NO_GRAPH 4
/* Project every step (u, v, w, p formed each step) instead of folding the
 * projection of u into the next predictor.  Without it the projection is
 * folded with dp_flag = 0, and with dp_flag = 1 where the pressure codes are
 * all non-periodic (DCT solve) and NS_FLAG_GAUGE_CORNER is off; p is then
 * formed where dp is. */
#define NS_FLAG_EAGER_PROJECTION 8

/* Run configuration: the driver parameter block (big_periodic_3D.f90:9-21,
 * 104-108; re_simulation_sub_tstep_interBC_combine.f90:100-112).
 * bc/pbc codes (NS_functions.f90:219-220): 1 periodic, 2 Dirichlet on the
 * cell wall, 3 Neumann on the cell wall, 4 Dirichlet on the ghost cell. */
typedef struct ns_config {
    int nx, ny, nz;
    double dx, dy, dz;
    double nu, dt;
    int scheme;  /* NS_EULER, NS_AB2, NS_AB2_CN */
    int bc[3];   /* velocity: 1, 2 or 4 */
    int pbc[3];  /* pressure: all 1 (FFT) or all in {2,3} (DST-II/DCT-II) */
    int dp_flag; /* 1: p accumulates dp and -dt*grad p enters the predictor */
    int flags;   /* NS_FLAG_* */
} ns_config;

typedef struct ns_ctx ns_ctx;

/* ---- lifecycle and state (replaces the drivers' array setup,
 *      big_periodic_3D.f90:26-33,157-169, and the HDF5 restart set
 *      u,v,w,p,rhs_*_previous, big_periodic_3D.f90:309-325) ---- */
int ns_create(const ns_config *cfg, ns_ctx **out);
int ns_destroy(ns_ctx *ctx);
long ns_plane_set_length(int nx, int ny, int nz);
/* first != 0: the next AB2/AB2-CN step is the Euler-type first step
 * (big_periodic_3D.f90:377, ...interBC_combine.f90:569); 0 = restart. */
int ns_upload_state(ns_ctx *ctx, const double *u, const double *v, const double *w, const double *p,
                    const double *hx, const double *hy, const double *hz, int first);
int ns_download_state(ns_ctx *ctx, double *u, double *v, double *w, double *p, double *hx, double *hy, double *hz);
/* u*, v*, w* and dp of the last step (dp gauged like the driver's). Any may be NULL.
 * All-periodic lazy AB2/Euler steps leave the ghost entries of u*, v*, w*
 * unfilled (their kernels read the periodic images instead); this call, like
 * ns_state_pointers and the downloads, fills them first. */
int ns_download_work(ns_ctx *ctx, double *u_star, double *v_star, double *w_star, double *dp);
int ns_synchronize(ns_ctx *ctx);
/* device pointers of the resident state, in the order u,v,w,p,hx,hy,hz,u*,v*,w*,dp; valid
 * until the next step (u* alternates between two buffers when dp_flag == 0, and u,v,w,p
 * are then formed from u*, dp only by ns_download_state / ns_max_div / ns_check_finite).
 * The resident arrays keep the reference's index order and extents, but their x rows
 * are ns_state_pitch() elements apart (whole 128-byte lines): entry (i,j,k) of an array
 * with y extent n2 is at ptr[(i-1) + pitch*((j-1) + n2*(k-1))].  The copy calls
 * (ns_upload_state, ns_download_state, ns_download_work) take and return dense arrays. */
int ns_state_pointers(ns_ctx *ctx, double **ptrs11);
/* row pitch (elements) of the resident arrays; no counterpart in the reference (its
 * arrays are dense), for hosts that read ns_state_pointers' arrays themselves */
int ns_state_pitch(ns_ctx *ctx, int *pitch);

/* ---- fused time step (the loop body of big_periodic_3D.f90:326-571 and
 *      re_simulation_sub_tstep_interBC_combine.f90:508-802) ---- */
/* current boundary-plane set used by the following steps (host or device) */
int ns_set_bc_planes(ns_ctx *ctx, const double *planes);
int ns_get_bc_planes(ns_ctx *ctx, double *planes);
/* current set = (1-theta)*a + theta*b on the device: inter-boundary time
 * interpolation (MATLAB interp1 of sub_tstep_BC_combined_interpolation.m:61-64,
 * linear form) */
int ns_bc_interp(ns_ctx *ctx, const double *set_a, const double *set_b, double theta);
int ns_step(ns_ctx *ctx, int nsteps);
/* ns_step with HIP events around each phase on the context stream; adds the
 * per-phase GPU milliseconds, summed over the steps, into phase_ms[NS_NPHASES]:
 * 0 predictor, 1 velocity BC fill or CN2 rows + ADI sweeps, 2 divergence +
 * x forward transform, 3 y forward, 4 z forward + modal solve + z inverse,
 * 5 y inverse, 6 x inverse (dp), 7 pressure ghost fill, 8 projection.
 * (The reference times the same phases with SYSTEM_CLOCK,
 * big_periodic_3D.f90:328-347,491-492,515-519,545-546,568-569.) */
#define NS_NPHASES 9
int ns_step_profiled(ns_ctx *ctx, int nsteps, double *phase_ms);
/* steps with a plane set per step, linearly interpolated between parent
 * snapshots: snapshot s is at time s*snap_dt, step m uses
 * t = t0 + (m+1)*dt (the boundary data of the step's end time). */
int ns_resim_advance(ns_ctx *ctx, const double *snapshots, int nsnap, double snap_dt, double t0, int nsteps);
/* The same with MATLAB interp1(...,'spline') in time, i.e. the not-a-knot
 * cubic spline through a window of six snapshots, three at or before t,
 * kept inside the history (sub_tstep_BC_combined_interpolation.m:47-64;
 * the reference interpolates u*, dp boxes offline and extracts planes, which
 * is the same linear map applied in the other order). */
int ns_resim_advance_spline(ns_ctx *ctx, const double *snapshots, int nsnap, double snap_dt, double t0,
                            int nsteps);
/* the current plane set := that spline interpolation at time t */
int ns_bc_interp_spline(ns_ctx *ctx, const double *snapshots, int nsnap, double snap_dt, double t);
/* Host only (no GPU): weights w[0..m-1] with spline(x) = sum_i w[i]*y[i] for
 * the not-a-knot cubic spline through (t[i], y[i]), 2 <= m <= 8, t increasing
 * (m = 3: the parabola, m = 2: the line, as MATLAB's spline). */
int ns_spline_weights(const double *t, int m, double x, double *w);
/* Time step of the following steps (the re-simulation's sub-steps use
 * dt0/sub_tstep, re_simulation_sub_tstep_interBC_combine.f90:446-448).
 * Rebuilds the CN line systems; a pending lazy projection is applied first
 * with the old dt. */
int ns_set_dt(ns_ctx *ctx, double dt);
/* Save / restore the AB2 history rhs_*_previous: the reference saves it
 * after the first sub-step and restores it after the last, so the first full
 * step's AB2 extrapolation spans one dt0 (…interBC_combine.f90:602-612). */
int ns_history_save(ns_ctx *ctx);
int ns_history_restore(ns_ctx *ctx);
/* Parent-run capture of a sub-box's boundary data, the per-step sub-box
 * dump of big_periodic_3D.f90:433-489,664-712 reduced to what a re-simulation
 * reads: after ns_capture_begin, every `every`-th step writes one plane set
 * of the sub-box (its length is ns_plane_set_length(sb->nx,sb->ny,sb->nz))
 * into dst + k*length for the k-th capture (k < max_sets; dst is device
 * memory).  Velocity planes follow get_velpr_bc with wall code 2 on u*, v*, w*
 * (NS_functions.f90:446-503); for AB2-CN they are taken after each ADI
 * sweep of their axis (the post-sweep planes bx_u_1s ... of
 * big_periodic_3D.f90:433-489).  Pressure planes are the code-3 gradients of
 * the gauged dp (get_pr_bc, NS_functions.f90:534-536) and the trailing
 * scalar is mean(dp) over the sub-box interior (...interBC_combine.f90:771).
 * The sub-box origin (x0,y0,z0) is 1-based, as the reference's idx_* vectors. */
typedef struct ns_subbox {
    int x0, y0, z0;
    int nx, ny, nz;
} ns_subbox;
int ns_capture_begin(ns_ctx *ctx, const ns_subbox *sb, double *dst, int max_sets, int every);
/* Sub-box of one resident field, ghosts included, as the parent's idx_*
 * vectors select it (big_periodic_3D.f90:21-25,652-657): face axes x0..x0+n,
 * centred axes x0..x0+n+1.  dst (host or device) gets the Fortran-ordered
 * box, e.g. u_sub(n+1,ny+2,nz+2).  dp is gauged when NS_FLAG_GAUGE_CORNER. */
#define NS_FIELD_U 0
#define NS_FIELD_V 1
#define NS_FIELD_W 2
#define NS_FIELD_P 3
#define NS_FIELD_USTAR 4
#define NS_FIELD_VSTAR 5
#define NS_FIELD_WSTAR 6
#define NS_FIELD_DP 7
#define NS_FIELD_HX 8
#define NS_FIELD_HY 9
#define NS_FIELD_HZ 10
int ns_extract_subbox(ns_ctx *ctx, const ns_subbox *sb, int field, double *dst);

/* ---- discrete adjoint of ns_step (no counterpart in the reference: the
 * gradient side of the re-simulation) ----
 * Euler, AB2 and AB2-CN on every code the forward step takes (the ADI
 * sweeps' adjoints are solves with the transposed tridiagonal or reduced
 * cyclic systems), lazily projected contexts included: the adjoint works on
 * the state ns_download_state returns and on the forward velocity passed to
 * ns_adjoint_step, not on what the context stores between steps.  The
 * adjoint state is (u_b, v_b, w_b, p_b, hx_b, hy_b, hz_b), the gradients of a
 * scalar objective with respect to the state (u, v, w, p, rhs_*_previous)
 * at the current point of the reverse sweep, all in the ghosted shapes. */
int ns_adjoint_set_state(ns_ctx *ctx, const double *u_b, const double *v_b, const double *w_b, const double *p_b,
                         const double *hx_b, const double *hy_b, const double *hz_b); /* null = zero */
int ns_adjoint_get_state(ns_ctx *ctx, double *u_b, double *v_b, double *w_b, double *p_b, double *hx_b,
                         double *hy_b, double *hz_b);                                  /* null = skip */
/* One reverse step: the adjoint state of the end of a forward step becomes
 * that of its start.  u, v, w: the forward velocity at the start of that step
 * (host or device); first: the step was the Euler first step; planes_b
 * (host or device, plane-set length, may be null) receives the gradient with
 * respect to that step's plane set (mean_dp included). */
int ns_adjoint_step(ns_ctx *ctx, const double *u, const double *v, const double *w, int first, double *planes_b);

/* ---- batched re-simulation (no counterpart in the reference, which runs one
 * sub-box per process): many contexts stepped together, each kernel of the
 * step launched once for all boxes (a block finds its box from a prefix of
 * per-box block counts), steady steps replayed from one HIP graph ----
 * The contexts must share scheme (Euler, AB2 or AB2-CN), bc/pbc codes,
 * dp_flag and flags and use the DCT Poisson solver (walls); sizes may
 * differ.  Lazily projected boxes step lazily in the batch too; where the
 * boxes differ in projection mode, the lazy ones form u, v, w, p when they
 * join and project every step from then on.  A batch of 16 or more boxes
 * deals them into groups of about equal cell counts (n/8 groups, at most 4),
 * each with its own stream, tables and graphs, and queues every group's steps
 * before waiting for all, so the device runs the groups' kernels side by
 * side.  Each context keeps its own state and current plane set;
 * ns_batch_step(b, n) leaves every context as ns_step(ctx, n) would.  A batch
 * holds its contexts' device pointers and time steps: destroy it before any
 * of them, and re-create it after ns_set_dt on any of them. */
typedef struct ns_batch ns_batch;
int ns_batch_create(ns_ctx *const *ctxs, int n, ns_batch **out);
int ns_batch_destroy(ns_batch *b);
int ns_batch_step(ns_batch *b, int nsteps);
/* Autotuner of the batch's launch shapes (no counterpart in the reference):
 * per size class of the boxes (cells < 2^17, < 2^20, larger) it picks the
 * predictor's z-chunk length (in rounds of resident blocks: 4, 8, 16, 32),
 * the DCT passes' lines per block (default, /2, /4), the predictor's tile
 * (each box's shape with fewer idle lanes, 64 x 4, or 32 x 8) and, when the
 * x sweep is not partitioned (lines over 640 rows), its lines per block
 * (64, 32, 16, 8), by coordinate descent over steady batch steps
 * timed on the device (trial_steps steps per trial, best of two); the boxes'
 * states are restored after every trial.  No choice changes any computed
 * value.  knobs_out (12 ints or null): rounds, DCT drop, x-sweep log2 lines,
 * tile (0 per box, 64, 32), per class. */
int ns_batch_tune(ns_batch *b, int trial_steps, int *knobs_out);
int ns_capture_end(ns_ctx *ctx, int *nsets_written);
/* max |div u| over interior cells (big_periodic_3D.f90:576-579) */
int ns_max_div(ns_ctx *ctx, double *out);
int ns_check_finite(ns_ctx *ctx);
/* Re-simulation error metrics of the context's current u, v, w, p against the
 * parent's sub-box fields (host or device arrays of the ghosted shapes), over
 * the full ghosted arrays (re_simulation_sub_tstep_interBC_combine.f90:826-833,
 * other_utility.f90:110-129):
 *   rms[0..3] = rms(u_sub), rms(v_sub), rms(w_sub), rms(p_sub)
 *   err[0..3] = max|x - x_sub| / rms(x_sub)   (u, v, w, p)
 *   err[4..6] = mean|x - x_sub| / rms(x_sub)  (u, v, w)
 *   err[7]    = max|oz - oz_sub| / rms(oz_sub), oz = diff(v,1,1)/dx - diff(u,1,2)/dy */
int ns_error_metrics(ns_ctx *ctx, const double *u_sub, const double *v_sub, const double *w_sub,
                     const double *p_sub, double err[8], double rms[4]);
/* ns_error_metrics without any allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory, e.g. a re-simulation loop
 * computing its errors every step under a no-allocation rule).  u_sub..p_sub
 * must be device arrays (NS_ERR_INVALID otherwise) and the context's device
 * must be current.  The block partials live in a static 229 KB device array,
 * one per GPU, shared by every call on that GPU: calls that overlap on one
 * GPU (from several host threads or contexts) race and must be serialised
 * by the caller.  Every box ns_create accepts fits (at most 2048 partial
 * rows; the kernel walks larger boxes grid-stride).  Sums are combined on the
 * device in double precision, so results agree with ns_error_metrics to
 * rounding (its host combination uses extended precision). */
int ns_error_metrics_noalloc(ns_ctx *ctx, const double *u_sub, const double *v_sub, const double *w_sub,
                             const double *p_sub, double err[8], double rms[4]);

/* ---- measurement support (no counterpart in the reference) ----
 * dst[i] = src[i] for i < n (device arrays), one launch on the context's
 * stream: a hand-written streaming kernel whose traffic is known (8n bytes
 * read, 8n written), used to scale the hardware traffic counters. */
int ns_stream_copy(ns_ctx *ctx, const double *src, double *dst, long n);

/* ---- per-subroutine mirrors with the reference's arguments ---- */
/* cal_conv NS_functions.f90:53-59; conv_* are (nx or nx-1,ny,nz) etc. */
int ns_cal_conv(ns_ctx *ctx, const double *u, const double *v, const double *w, int bc_x, int bc_y, int bc_z,
                double dx, double dy, double dz, double *conv_x, double *conv_y, double *conv_z);
/* cal_diff NS_functions.f90:161-168 */
int ns_cal_diff(ns_ctx *ctx, const double *u, const double *v, const double *w, int bc_x, int bc_y, int bc_z,
                double dx2, double dy2, double dz2, double *diff_x, double *diff_y, double *diff_z);
/* vel_bc_staggered NS_functions.f90:213-217 (planes: packed set) */
int ns_vel_bc_staggered(ns_ctx *ctx, double *u_star, double *v_star, double *w_star, const double *planes,
                        int bc_x, int bc_y, int bc_z);
/* vel_bc_staggered_CN2 NS_functions.f90:313 (the dir-th plane pairs of the packed set) */
int ns_vel_bc_staggered_cn2(ns_ctx *ctx, double *u_star, double *v_star, double *w_star, const double *planes,
                            int bc, int dir);
/* ADI sweep along dir with the context's CN matrices (getrs/dttrsb loops,
 * big_periodic_3D.f90:425-483, ...interBC_combine.f90:656-701) */
int ns_adi_sweep(ns_ctx *ctx, double *u_star, double *v_star, double *w_star, int dir);
/* get_velpr_bc NS_functions.f90:414-420 (writes the packed set; p_previous
 * used only when !(dp_flag || using_ustar)) */
int ns_get_velpr_bc(ns_ctx *ctx, const double *u_bc, const double *v_bc, const double *w_bc, const double *p_bc,
                    double dt, int dp_flag, int using_ustar, const double *p_previous, double *planes);
/* get_pr_bc NS_functions.f90:518-521 */
int ns_get_pr_bc(ns_ctx *ctx, const double *p_bc, double *planes);
/* divergence RHS (big_periodic_3D.f90:498-501): rhs(nx,ny,nz) = div(u*)/dt */
int ns_divergence(ns_ctx *ctx, const double *u_star, const double *v_star, const double *w_star, double dt,
                  double *rhs);
/* pr_bc_staggered_modify_rhs NS_functions.f90:1345 */
int ns_pr_bc_staggered_modify_rhs(ns_ctx *ctx, double *rhs, const double *planes);
/* periodic FFT Poisson (big_periodic_3D.f90:526-548): rhs(nx,ny,nz) -> sol(nx,ny,nz) */
int ns_poisson_fft(ns_ctx *ctx, const double *rhs, double *sol);
/* DCT_poisson_solver NS_functions.f90:1505 on rhs(nx,ny,nz) in place */
int ns_dct_poisson_solver(ns_ctx *ctx, double *rhs, double mean_dp);
/* pr_bc_staggered NS_functions.f90:1403 */
int ns_pr_bc_staggered(ns_ctx *ctx, double *dp, const double *planes);
/* projection (big_periodic_3D.f90:564-567) over the full ghosted arrays */
int ns_project(ns_ctx *ctx, const double *u_star, const double *v_star, const double *w_star, const double *dp,
               double dt, int dp_flag, double *u, double *v, double *w, double *p);

/* ---- adjoint mirrors (no counterpart in the reference: the transposed
 * kernels of ns_adjoint_step, one per linear mirror above) ----
 * Each applies the transpose of its forward mirror's linear map, with the
 * same context and codes.  Arrays the forward transforms in place are
 * transformed in place; the adjoints of the forward's plane-set input are
 * ADDED to planes_b (packed set, plane-set length); other adjoint outputs
 * are noted per entry.  Host or device arrays, as for the forward mirrors. */
/* (u_b, v_b, w_b) := F_u^T (u_b, v_b, w_b), planes_b += F_p^T (...) */
int ns_vel_bc_staggered_adj(ns_ctx *ctx, double *u_b, double *v_b, double *w_b, double *planes_b, int bc_x,
                            int bc_y, int bc_z);
/* bc 2, 4: the ends' adjoints move to planes_b; bc 1: ghost rows' adjoints zeroed; bc 3: identity */
int ns_vel_bc_staggered_cn2_adj(ns_ctx *ctx, double *u_b, double *v_b, double *w_b, double *planes_b, int bc,
                                int dir);
/* solve with the transposed CN systems (AB2-CN contexts) */
int ns_adi_sweep_adj(ns_ctx *ctx, double *u_b, double *v_b, double *w_b, int dir);
/* (u_b, v_b, w_b) += div^T rhs_b */
int ns_divergence_adj(ns_ctx *ctx, const double *rhs_b, double dt, double *u_b, double *v_b, double *w_b);
/* planes_b += the boundary terms' transpose applied to rhs_b (rhs itself passes through unchanged) */
int ns_pr_bc_staggered_modify_rhs_adj(ns_ctx *ctx, const double *rhs_b, double *planes_b);
/* rhs_b := P^T sol_b (the periodic solve is symmetric) */
int ns_poisson_fft_adj(ns_ctx *ctx, const double *sol_b, double *rhs_b);
/* in place: the solution's adjoint becomes the RHS's adjoint (solve with the
 * zero mode dropped); *mean_dp_b += its sum when every pressure code is 3 */
int ns_dct_poisson_solver_adj(ns_ctx *ctx, double *rhs_b, double *mean_dp_b);
/* dp_b := F^T dp_b, planes_b += ... */
int ns_pr_bc_staggered_adj(ns_ctx *ctx, double *dp_b, double *planes_b);
/* from the adjoints of (u, v, w, p): us_b.. := u_b.., dp_b := p_b + grad^T(dt u_b),
 * p_b := dp_flag * p_b (the adjoint of the input p) */
int ns_project_adj(ns_ctx *ctx, const double *u_b, const double *v_b, const double *w_b, double *p_b, double dt,
                   int dp_flag, double *us_b, double *vs_b, double *ws_b, double *dp_b);

#ifdef __cplusplus
}
#endif
#endif
