/*
 * aeromhd.h -- C ABI of libaeromhd_hip.so, the gfx950 implementation of the
 * 2-D ideal-MHD update of Semedo2002/AeroMHD-CUDA (MUSCL / van Leer
 * reconstruction, HLLD flux, GLM psi transport with Powell source, SSP-RK3
 * stage combine, CFL max-wavespeed reduction).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on failure;
 *     amhd_last_error() then describes the failure (thread-local text).
 *       -1 invalid argument, -2 HIP runtime error, -3 out of memory,
 *       -4 unsupported configuration.
 *   - State arrays crossing the ABI have the reference's layout: the ghosted
 *     conserved array U of shape (10, nx+4, ny+4), C order, float64, with
 *     variable order [rho, mx, my, mz, Bx, By, Bz, E, psi, rhoC]
 *     (config.py:13-16); primitive arrays use [rho, vx, vy, vz, Bx, By, Bz,
 *     p, psi, C] (config.py:19-22).
 *   - `stream` arguments are hipStream_t handles passed as void* (NULL = the
 *     null stream).  Functions that return a host value synchronise.
 *   - A context (amhd_ctx) owns its device buffers and its own stream.  It is
 *     not thread-safe; use one context per thread.
 */
#ifndef AEROMHD_H
#define AEROMHD_H

#ifdef __cplusplus
extern "C" {
#endif

#define AMHD_ABI_VERSION 1

enum { AMHD_F64 = 0, AMHD_F32 = 1 };

/* Solver parameters: the fields of config.Config (config.py:47-89) that the
 * update reads.  bc_*_periodic: 1 for "periodic", 0 for "extrapolation" and
 * "characteristic" (identical in effect, solver.py:167-176).  powell must
 * already be (powell_source && B_transverse > 0), the gate of
 * solver.py:304. */
typedef struct {
    double gamma;
    double dx, dy;
    double x_min, y_min;
    double glm_alpha;
    int bc_x_periodic;
    int bc_y_periodic;
    int powell;
    int reserved;
} amhd_params;

const char* amhd_last_error(void);
int amhd_abi_version(void);

/* ---- solver context: replaces MHDSolver's device state (solver.py:72-87) */

/* Allocate a context for an nx x ny interior grid (ng = 2, solver.py:74).
 * dtype: AMHD_F64 (reference arithmetic) or AMHD_F32. */
int amhd_create(void** ctx, int nx, int ny, int dtype, const amhd_params* p);
int amhd_destroy(void* ctx);

/* Rows per wavefront segment of the fused stage kernel (0 = default). */
int amhd_set_rows(void* ctx, int rows);
int amhd_get_rows(void* ctx);

/* Packing of a narrow last y strip: when the last strip of 60 columns
 * holds at most 28 cells, up to max_groups of its x segments (2, 4 or 8;
 * 0 = no packing) share one wavefront.  Default: 0.  A tuning switch like
 * amhd_set_rows (no counterpart in the reference); results do not depend
 * on it. */
int amhd_set_packing(void* ctx, int max_groups);

/* Cost-sized segments (no counterpart in the reference): each step the
 * last stage records, per strip, which rows hold the same bits as the next
 * row (such a row takes its y-face fluxes from the row before, and its x
 * faces are cheap: the physical flux in FP32, the previous face's flux in
 * FP64 once five rows repeat), and the segment boundaries move so that
 * every wavefront's segment costs about the same.  Default 1.  Results do
 * not depend on it. */
int amhd_set_balance(void* ctx, int on);

/* MHDSolver.U assignment / read (solver.py:83, 139).  U: ghosted float64
 * (10, nx+4, ny+4), host or device memory.  On read, ghost cells are filled
 * as apply_bc (solver.py:158-197) leaves them. */
int amhd_upload_state(void* ctx, const double* U, int on_device);
int amhd_download_state(void* ctx, double* U, int on_device);

/* MHDSolver.compute_dt on the context's own state (solver.py:206-223):
 * returns dt and glm_ch (the value the reference stores in cfg.glm_ch). */
int amhd_compute_dt(void* ctx, double cfl, double* dt, double* ch);

/* MHDSolver.step_ssprk3 (solver.py:310-333) with self.dt = dt and the
 * frozen cfg.glm_ch = ch.  psi_decay = exp(-glm_alpha*ch*dt/min(dx,dy)),
 * evaluated by the caller as the reference does on the host
 * (solver.py:331-333).  Asynchronous on the context stream. */
int amhd_step_ssprk3(void* ctx, double dt, double ch, double psi_decay);

/* ---- x-slab decomposition (one slab per context; the per-GPU part of the
 * multi-GPU layout, no counterpart in the reference).  A slab is a context
 * over rows [i0, i1) of the global grid (x_min and dx of the global grid).
 * amhd_set_halo: caller-owned halo buffers for the rows just outside the
 * slab, each 10 planes of 2 rows of ny cells in the context's precision
 * (plane stride 2*ny): lo = rows i0-2, i0-1, hi = rows i1, i1+1 of the stage
 * input; NULL = that side is the domain boundary (the params' BC).
 * amhd_stage: one SSP-RK3 stage (1, 2, 3) of amhd_step_ssprk3 on its own, so
 * that halos can be refreshed between stages; stage 3 completes the step.
 * amhd_edge_rows: copy the slab's first (side 0) or last (side 1) two rows
 * of the given stage's input into a neighbour's halo buffer, on the
 * context's stream.
 * amhd_cfl_maxima: the four CFL maxima of the current state, {max cf^2,
 * max(|vx|+|vy|), max|vx|, max|vy|} (synchronises); the maxima of the whole
 * grid are their maxima over the slabs.
 * amhd_dt_from_maxima: dt and ch of solver.py:206-223 from such maxima (the
 * device step control's expressions, on the host).
 * amhd_synchronize: wait for the context's stream. */
int amhd_set_halo(void* ctx, const void* lo, const void* hi);
int amhd_stage(void* ctx, int stage, double dt, double ch, double psi_decay);
int amhd_edge_rows(void* ctx, int stage, int side, void* dst);
int amhd_cfl_maxima(void* ctx, double m[4]);
int amhd_dt_from_maxima(const double m[4], double cfl, double dx, double dy, double* dt, double* ch);
int amhd_synchronize(void* ctx);

/* Device-resident time loop of MHDSolver.run (solver.py:475-482) without
 * diagnostics: up to n_steps steps, stopping at t_end or max_steps (counted
 * in *step_inout) like the reference loop.  dt, ch and the psi decay are
 * computed on the device; no host round trip per step.  A step is three
 * launches: the stage kernels carry the step control (the first stage
 * reduces the CFL maxima and begins the step, the third closes it).
 * use_graph: replay a captured hipGraph of the step sequence.  Synchronises
 * at the end. */
int amhd_run(void* ctx, int n_steps, double cfl, double t_end, int max_steps, int use_graph, double* t_inout,
             int* step_inout, double* last_dt);

/* amhd_run without graph replay, every launch on the context stream
 * carrying a HIP event pair that its dispatch fills with the kernel's start
 * and end: stage_ms[s] receives the summed device time of the stage-(s+1)
 * kernel over the steps run; *ctl_ms (may be NULL) receives 0, the device
 * loop having no separate step-control launch.  For measurement. */
int amhd_run_profiled(void* ctx, int n_steps, double cfl, double t_end, int max_steps, double* t_inout,
                      int* step_inout, double* last_dt, double stage_ms[3], double* ctl_ms);

/* Accumulated boundary energy flux (solver.py:316-317) and the last step's
 * flux. */
int amhd_boundary_energy(void* ctx, double* cumulative, double* last);
int amhd_reset_boundary_energy(void* ctx, double value);

/* MHDSolver.compute_rhs(U) (solver.py:268-308) for a ghosted float64 device
 * array U, which receives apply_bc in place as in the reference; R (same
 * shape, device) is zero in the ghosts.  ch: the frozen glm_ch. */
int amhd_compute_rhs(void* ctx, double* U_dev, double* R_dev, double ch);

/* FP64 global sums of the interior state: out[0] = sum(rho)*dx*dy,
 * out[1] = sum(E)*dx*dy (solver.py:430), out[2] = max|divB|,
 * out[3] = sqrt(mean(divB^2)) with divB as solver.py:419-428. */
int amhd_global_sums(void* ctx, double out[4]);

/* The concentration profile behind the mixing diagnostics
 * (solver.py:340, 352-364): for each interior x row i, with
 * C = clip(rhoC * (1/max(rho, 1e-12)), 0, 1) as cons_to_prim gives it,
 * mean[i] = mean over y of C and var[i] = mean over y of (C - mean[i])^2,
 * FP64 (host arrays of nx; synchronises). */
int amhd_concentration_profile(void* ctx, double* mean, double* var);

/* The rest of compute_diagnostics (solver.py:335-444) reduced on the device
 * in FP64, from cons_to_prim with its 1e-12 floors:
 * row_ens[i] = sum over y of rho * omega^2 for interior row i, omega =
 * dvy/dx - dvx/dy with np.gradient's differences (solver.py:396-399);
 * row_stag[i] = max over y of p + rho |v|^2 / 2 + |B|^2 / 2 (solver.py:
 * 412-414; NaN if any cell is NaN);
 * col_pos[j] = the interface x position of column j as solver.py:372-384
 * finds it (the |dC/dx|-weighted mean of x over the 2 hw + 1 rows around
 * the first maximum), NaN where it finds none.
 * Host arrays of nx, nx and ny doubles; synchronises. */
int amhd_flow_diagnostics(void* ctx, double* row_ens, double* row_stag, double* col_pos);

/* Learned-closure hooks (no counterpart in the reference).
 * amhd_state_check: out = {min rho, min p, max|divB|, non-finite cells}.
 * amhd_state_ptr: device address of the current interior-layout state
 * (10 planes of nx*ny cells, the context's precision; valid until the next
 * step).  The caller may edit the state through it before the next step
 * (the cached CFL partials are dropped).
 * amhd_get_stream: the context's stream (for work that must be ordered
 * with its steps without a host synchronisation). */
int amhd_state_check(void* ctx, double out[4]);
int amhd_state_ptr(void* ctx, void** state);
int amhd_get_stream(void* ctx, void** stream);

/* ---- standalone operators of physics.py / solver.py (float64, device
 * pointers, asynchronous on `stream` unless a host result is returned) */

/* physics.py:284-301 / 304-320 on (10, n) planes. */
int amhd_cons_to_prim(const double* U, double* W, long n, double gamma, void* stream);
int amhd_prim_to_cons(const double* W, double* U, long n, double gamma, void* stream);

/* physics.py:323-350: W (10, nxt, nyt) -> WL, WR of shape (10, nxt-3, nyt)
 * (dir 0, muscl_x) or (10, nxt, nyt-3) (dir 1, muscl_y). */
int amhd_muscl(const double* W, int nxt, int nyt, int dir, double* WL, double* WR, void* stream);

/* The HLLD kernel hlld_flux_x_kernel(WL, WR, F, smax_arr, n_faces, gamma,
 * ch, floor_rho, floor_pr) of physics.py:44-55 with the wrappers of
 * physics.py:353-379: (10, n_faces) planes, dir 0 = x faces, dir 1 = y faces
 * (vx<->vy, Bx<->By rotation).  smax_arr (device, n_faces values, may be
 * NULL) receives each face's max(|SL|,|SR|) as the reference kernel writes
 * it (physics.py:108-109); smax (host, may be NULL) their maximum, as
 * float(xp.max(smax_arr)) does (synchronises).  dtype AMHD_F64 or AMHD_F32
 * for WL/WR/F/smax_arr. */
int amhd_hlld_flux(const void* WL, const void* WR, void* F, void* smax_arr, double* smax, long n_faces, int dir,
                   double gamma, double ch, double floor_rho, double floor_pr, int dtype, void* stream);

/* solver.py:158-197 in place on a ghosted (10, nx+4, ny+4) device array. */
int amhd_apply_bc(double* U, int nx, int ny, int bc_x_periodic, int bc_y_periodic, void* stream);

/* solver.py:199-204 in place on (10, n) planes. */
int amhd_enforce_scalar_bounds(double* U, long n, void* stream);

/* solver.py:206-223 over a primitive (10, n) device array W: dt and ch. */
int amhd_compute_dt_w(const double* W, long n, double gamma, double cfl, double dx, double dy, double* dt,
                      double* ch, void* stream);

/* amhd_compute_dt_w without allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same results
 * bit for bit; its CFL partials (at most 1024) and step-control block live
 * in a file-scope device array of the library, 32 KiB + 128 B per device,
 * held until the process ends.  The partial grid is capped at 1024
 * workgroups (grid-stride), so every n >= 1 fits.  Calls on one device share
 * that array: overlapping calls on one device race.  Synchronises. */
int amhd_compute_dt_w_noalloc(const double* W, long n, double gamma, double cfl, double dx, double dy, double* dt,
                              double* ch, void* stream);

/* ---- batched form (no counterpart in the reference): many independent
 * instances of uneven size advanced together, one persistent launch per RK
 * stage over all instances, each with its own device-side dt, glm_ch, psi
 * decay, t and step count; the step is captured once in a hipGraph. */

/* n instances; instance k has an nx[k] x ny[k] interior and params[k]. */
int amhd_batch_create(void** batch, int n, const int* nx, const int* ny, int dtype, const amhd_params* params);
int amhd_batch_destroy(void* batch);
/* Rows per task (0 = sized so that the task count matches the resident
 * wavefront slots). */
int amhd_batch_set_rows(void* batch, int rows);
int amhd_batch_info(void* batch, int* n_tasks, int* resident_waves, int* rows);
/* Ghosted float64 (10, nx[k]+4, ny[k]+4) state of instance k, as
 * amhd_upload_state / amhd_download_state. */
int amhd_batch_upload(void* batch, int k, const double* U, int on_device);
int amhd_batch_download(void* batch, int k, double* U, int on_device);
/* Up to n_steps steps of every instance (each stops at its own t_end[k] /
 * max_steps[k], counted in step_inout[k]); t_inout, step_inout: [n]. */
int amhd_batch_run(void* batch, int n_steps, const double* cfl, const double* t_end, const int* max_steps,
                   int use_graph, double* t_inout, int* step_inout);
/* Device addresses of every instance's current state (cur[k]) and of the
 * state before its last step (prev[k], may be NULL), interior layout, valid
 * until the next amhd_batch_run (synchronises).  The caller may edit the
 * current states before the next run (the CFL partials are recomputed). */
int amhd_batch_state_ptrs(void* batch, void** cur, void** prev);
/* dt of every instance's last step of the last amhd_batch_run (host array
 * of n; meaningful for the instances that took a step; synchronises). */
int amhd_batch_last_dt(void* batch, double* dt);

/* ---- learned flux closure (no counterpart in the reference): 3x3
 * convolutions on BF16 matrix cores (v_mfma_f32_16x16x32_bf16) with FP32
 * accumulation.  NHWC tensors over the coarse grid, (n, nx, ny, C) with
 * C in {16, 32}; zero padding in x, periodic in y.  Weights W[co][k],
 * k = tap*Cin + ci, tap = (di+1)*3 + (dj+1), padded to kpad = roundup(9*Cin, 32);
 * transposed weights WT[ci][tap*Cout + co] padded to roundup(9*Cout, 32). */
int amhd_conv3x3_kpad(int cin);
/* y = act(conv(x, W) + bias): BF16 x, W; y BF16 (relu = 1) or FP32 (y_f32 = 1). */
int amhd_conv3x3_fwd(const void* x, int n, int nx, int ny, int cin, const void* w, const float* bias, int cout,
                     int relu, void* y, int y_f32, void* stream);
/* dx = mask * convT(dy, W): dy FP32 (n,nx,ny,cout); mask BF16 (n,nx,ny,cin)
 * or NULL (dx where mask > 0); dx FP32. */
int amhd_conv3x3_bwd_data(const float* dy, int n, int nx, int ny, int cout, const void* wt, int cin, const void* mask,
                          float* dx, void* stream);
/* dW[co][k] = sum_p dy[p][co] x[p + off(tap)][ci], dbias = sum_p dy; the
 * pixel sum is split over n_split workgroup columns with FP32 partials in
 * workspace (n_split * cout * kpad floats), summed in a fixed order. */
int amhd_conv3x3_bwd_weight(const void* x, const float* dy, int n, int nx, int ny, int cin, int cout, float* dw,
                            float* dbias, float* workspace, int n_split, void* stream);
/* D (16x16) = A (16x32) * B (32x16), row-major FP32, operands rounded to BF16:
 * one v_mfma_f32_16x16x32_bf16 (checks the lane layout). */
int amhd_mfma_probe(const float* A, const float* B, float* D, void* stream);
/* Closure inference glue over n instances of nx*ny cells.  states[m]: device
 * address of instance m's 10 planes of nx*ny cells (dtype AMHD_F64/AMHD_F32;
 * the interior layout of amhd_state_ptr); fields[9]: the plane of each of
 * the 9 network fields; mean, std, fscale: host float[9].
 * amhd_closure_features: x (n, nx, ny, 16) BF16, channel f < 9 =
 * (U[fields[f]] - mean[f]) / std[f] in FP32, channels 9..15 zero.
 * amhd_closure_source: from the network output dF (n, nx, ny, 32) FP32,
 * 16-byte aligned --
 * channel f the x-face flux correction at i-1/2, 9 + f the y-face one at
 * j-1/2, both in units of fscale[f]; fields[0] must be rho (0) and
 * fields[8] rhoC (9), whose channels carry the mixing flux G -- the
 * face fluxes F_f = w * dF_f * fscale[f] for f < 8 and
 * F_rhoC = C_donor * F_rho + wm * G (C = rhoC/rho of states, taken from
 * the cell upwind of F_rho), and the source
 * S[fields[f]] = -((F_f(i+1/2) - F_f(i-1/2)) / dx + (F_f(j+1/2) - F_f(j-1/2)) / dy)
 * (x faces on both x boundaries carry nothing, y is periodic).
 * accumulate = 0: srcs[m] (10 planes, dtype) = S, other planes 0;
 * accumulate = 1: srcs[m] += dt * S on the named planes (srcs[m] may be
 * states[m]: the post-step correction of a state in place).
 * dt: host array of n step sizes (needed by the limiter and by accumulate;
 * may be NULL otherwise).
 * theta: NULL (w = wm = 1) or 2*n*nx*ny floats of workspace for the face
 * weights w = min over the face's two cells of theta_c, wm likewise of
 * theta_mix_c.  theta_c = min over the cell's faces of the largest t in
 * {1, 1/2, ..., 1/32, 0} for which states + 4 t dt S_face alone keeps rho
 * and p at or above keep times their values; theta_mix_c likewise keeps C
 * within the min/max of C over the cell and its face neighbours, taken in
 * states and, if bounds (n device addresses, or NULL) is given, in those
 * states as well (the states before the step, for a post-step
 * correction: Zalesak's bounds).  By convexity every cell of states + dt*S
 * keeps both bounds, and S stays a flux difference, so it conserves
 * exactly. */
int amhd_closure_features(const void* const* states, int dtype, const int* fields, int n, int nx, int ny,
                          const float* mean, const float* std, void* x, void* stream);
int amhd_closure_source(const float* dF, const void* const* states, void* const* srcs, int dtype, const int* fields,
                        int n, int nx, int ny, const float* fscale, double dx, double dy, const double* dt,
                        double gamma, double keep, float* theta, const void* const* bounds, int accumulate,
                        void* stream);
/* The closure's guard without writing the state: for instance m, out[4m..4m+3]
 * = {min rho, min p, max|divB|, non-finite cells} of states[m] + dt[m] * srcs[m]
 * (10 planes each; p = (gamma-1)(E - |m|^2/(2 rho) - |B|^2/2), divB as
 * amhd_state_check), FP64; work: device workspace of 32*64*4 doubles;
 * host out; synchronises.  dt[m] = 0 checks the state itself.
 * amhd_closure_apply: states[m] += dt[m] * srcs[m] on all 10 planes. */
int amhd_closure_check(const void* const* states, const void* const* srcs, const double* dt, int dtype, int n,
                       int nx, int ny, double gamma, double dx, double dy, double* work, double* out, void* stream);
int amhd_closure_apply(void* const* states, const void* const* srcs, const double* dt, int dtype, int n, int nx,
                       int ny, void* stream);
/* Limited anti-diffusion of the concentration across x faces, one pass of
 * flux-corrected transport (Zalesak 1979) on states[m] in place: face
 * i+1/2 carries A = kappa * (rho_i + rho_{i+1}) / 2 * (C_{i+1} - C_i) of
 * rhoC (x-boundary faces nothing), limited so that every cell keeps C within
 * the min/max of C over the cell and its face neighbours, in states and,
 * if bounds (n device addresses, or NULL) is given, in those states as
 * well; rho and every other plane are untouched and rhoC is conserved row
 * by row.  kappa >= 0 (0: nothing runs); work: device workspace of
 * 3 * min(n, 32) * nx * ny doubles. */
int amhd_closure_sharpen(void* const* states, const void* const* bounds, int dtype, int n, int nx, int ny,
                         double kappa, double* work, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* AEROMHD_H */
