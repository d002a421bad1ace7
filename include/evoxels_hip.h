/*
 * evoxels_hip.h - C ABI of libevoxels_hip.so, the MI355X (gfx950) kernels
 * behind evoxels' time-stepping hot path.
 *
 * The reference (evoxels 1.0.0) is pure Python on torch; it has no FFI.  Each
 * entry point below replaces one reference operator or one fused composition
 * of them, cited as file:line of the reference package.
 *
 * Conventions
 *   - All pointers are device pointers into caller-owned memory, except the
 *     evx_grid / evx_bc / evx_epilogue / evx_imex_params descriptors and the
 *     out-parameters of plan functions, which are host pointers.
 *   - Fields are C-contiguous (channels, nx, ny, nz), z fastest; padded
 *     fields are (channels, nx+2, ny+2, nz+2).
 *   - dtype is EVX_F32 or EVX_F64 and applies to every real buffer of a call.
 *   - stream is a hipStream_t (NULL = legacy default stream).  Every call is
 *     asynchronous on that stream, performs no device allocation and no host
 *     synchronisation, and is therefore legal under hipGraph capture.
 *   - Return value: EVX_OK or a negative status; evx_last_error() gives the
 *     message of the last failing call on the calling thread.
 */
#ifndef EVOXELS_HIP_H
#define EVOXELS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EVX_ABI_VERSION 2

enum evx_status {
  EVX_OK = 0,
  EVX_ERR_INVALID = -1,     /* bad argument, shape or descriptor            */
  EVX_ERR_HIP = -2,         /* a HIP runtime call or kernel launch failed   */
  EVX_ERR_UNSUPPORTED = -3, /* valid request this build does not implement  */
  EVX_ERR_NOMEM = -4        /* plan-time allocation failed                  */
};

enum evx_dtype { EVX_F32 = 0, EVX_F64 = 1 };

/* Boundary kinds of one axis (problem_definition.py:60-103).  The first
 * three are the cell_center rules; the last two are the staggered_x rules of
 * the x axis, where the ghost plane is the boundary node itself:
 *   EVX_BC_NODE_VALUE   ghost := v          (StaggeredXBCs.pad_dirichlet_periodic,
 *                                            boundary_conditions.py:96-106)
 *   EVX_BC_NODE_EXTRAP  ghost := 4/3 u1 - 1/3 u2 (zero flux,
 *                                            StaggeredXBCs.pad_zero_flux_periodic,
 *                                            boundary_conditions.py:108-121)
 * The staggered kinds are valid on x only, with y and z periodic (the only
 * staggered_x paddings the reference implements), and are accepted by
 * evx_pad_bc, evx_rd_rhs, evx_ac_rhs and evx_coupled_rd_rhs. */
enum evx_bc_kind {
  EVX_BC_PERIODIC = 0,
  EVX_BC_NEUMANN = 1,
  EVX_BC_DIRICHLET = 2,
  EVX_BC_NODE_VALUE = 3,
  EVX_BC_NODE_EXTRAP = 4
};

/* Per-axis boundary description.  lo/hi are the Dirichlet (or boundary node)
 * values, ignored otherwise.  cell_center ghosts follow
 * CellCenteredBCs.pad_bc (boundary_conditions.py:33-59): periodic wrap, then
 * x, y, z rules in that order (Neumann u0 := u1, Dirichlet u0 := 2v - u1). */
typedef struct evx_bc {
  int32_t kind[3];
  double lo[3];
  double hi[3];
} evx_bc;

/* Grid geometry.  div_dx / div_dx2 must already be rounded to the working
 * precision exactly as VoxelGrid.__init__ does (voxelgrid.py:29-30):
 * 1/float(spacing) and 1/float(spacing)**2 in that precision. */
typedef struct evx_grid {
  int64_t channels, nx, ny, nz;
  double div_dx[3];
  double div_dx2[3];
} evx_grid;

/* Fused output of a right-hand-side kernel (r = rhs(u) at each voxel):
 *   EVX_OUT_RHS        out = r
 *   EVX_OUT_EULER      out = u + coef*r                 (timesteppers.py:42-43)
 *   EVX_OUT_RK4_FIRST  out = base + coef*r, acc_out = r  (timesteppers.py:57-58)
 *   EVX_OUT_RK4_MID    out = base + coef*r, acc_out = acc_in + 2*r
 *   EVX_OUT_RK4_LAST   out = base + coef*(acc_in + r)    (timesteppers.py:61)
 * coef is the Python-double scalar of the reference expression (dt, 0.5*dt,
 * dt/6); it is rounded to the working precision in-kernel.  base == NULL
 * means u.  out must not alias u (stencil reads neighbours). */
enum evx_out_mode {
  EVX_OUT_RHS = 0, EVX_OUT_EULER = 1, EVX_OUT_RK4_FIRST = 2,
  EVX_OUT_RK4_MID = 3, EVX_OUT_RK4_LAST = 4
};
typedef struct evx_epilogue {
  int32_t mode;
  double coef;
  const void* base;
  const void* acc_in;
  void* acc_out;
} evx_epilogue;

/* ---- library ------------------------------------------------------------ */
int evx_version(void);
const char* evx_last_error(void);

/* ---- boundary padding and stencils (operator level) --------------------- */

/* out(C,nx+2,ny+2,nz+2) = pad_bc(u).  Replaces CellCenteredBCs.pad_bc /
 * pad_periodic / pad_dirichlet_periodic / pad_zero_flux_periodic
 * (boundary_conditions.py:9-59; voxelgrid.py:194-195). */
int evx_pad_bc(int32_t dtype, const void* u, void* out, const evx_grid* g,
               const evx_bc* bc, void* stream);

/* out(C,nx,ny,nz) = laplace(padded) for an already padded input.
 * Replaces FDStencils.laplace (fd_stencils.py:62-75). */
int evx_laplace_padded(int32_t dtype, const void* padded, void* out,
                       const evx_grid* g, void* stream);

/* out(C,nx,ny,nz) = normal_laplace(padded).  Replaces
 * FDStencils.normal_laplace (fd_stencils.py:77-103). */
int evx_normal_laplace_padded(int32_t dtype, const void* padded, void* out,
                              const evx_grid* g, void* stream);

/* ---- fused right-hand sides --------------------------------------------- */

/* ReactionDiffusion with f == 0: r = D*laplace(pad_bc(u)).
 * Replaces ReactionDiffusion.rhs (problem_definition.py:226-229) composed
 * with ForwardEuler / RungeKutta4 through the epilogue. */
int evx_rd_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g,
               const evx_bc* bc, double D, const evx_epilogue* ep, void* stream);

/* CahnHilliard with the default mu_hom:
 * r = D*div(cf(1-cf) grad mu), mu = 18/eps c(1-c)(1-2c) - 2 eps lap c,
 * c = clip(u,0,1).  Replaces CahnHilliard.rhs (problem_definition.py:328-371). */
int evx_ch_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g,
               const evx_bc* bc, double eps, double D, const evx_epilogue* ep,
               void* stream);

/* TwoPhaseAllenCahn with the default potential.  Replaces
 * TwoPhaseAllenCahn.rhs (problem_definition.py:440-447). */
int evx_ac_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g,
               const evx_bc* bc, double eps, double gab, double M, double force,
               double curvature, const evx_epilogue* ep, void* stream);

/* CoupledReactionDiffusion with the default interaction u[0]*u[1]^2 on a
 * two-channel state (A = channel 0, B = channel 1):
 * r_A = D_A*lap(A) - A*B^2 + feed*(1-A),  r_B = D_B*lap(B) + A*B^2 - kill*B.
 * Replaces CoupledReactionDiffusion.rhs (problem_definition.py:612-633);
 * g->c must be 2. */
int evx_coupled_rd_rhs(int32_t dtype, const void* u, void* out, const evx_grid* g,
                       const evx_bc* bc, double D_A, double D_B, double feed,
                       double kill, const evx_epilogue* ep, void* stream);

/* Simplex projection of the phase fractions of MultiPhaseAllenCahn, per
 * voxel over n_phases channels of `voxels` values each: euclidean = 0 is the
 * clip-and-normalise projection (problem_definition.py:508-512), 1 the
 * Euclidean projection by sort and cumulative sum (problem_definition.py:
 * 514-530).  1 <= n_phases <= 8. */
int evx_simplex_project(int32_t dtype, const void* phis, void* out, int32_t n_phases,
                        int64_t voxels, int32_t euclidean, void* stream);

/* MultiPhaseAllenCahn.rhs with the 'well' potential (problem_definition.py:
 * 532-573) from the projected phases `phis` (g->c = n_phases channels) and
 * their padded copy (C, Nx+2, Ny+2, Nz+2):
 * df = -curv*lap - (1-curv)*normal_lap + 9/(2eps^2)*(3phi(sum phi^2 - phi^2)
 * + phi^3 - phi), r = -M*gab*(df - mean_phases(df)).  The normal Laplacian is
 * skipped when curvature == 1. */
int evx_multiphase_ac_rhs_padded(int32_t dtype, const void* phis, const void* padded,
                                 void* out, const evx_grid* g, double eps, double gab,
                                 double M, double curvature, void* stream);

/* ReactionDiffusionSBM.rhs without the source term (problem_definition.py:
 * 273-287) from the padded field and the padded (clipped) mask:
 * r = D * sum_axes grad_face(grad_face(z) - to_face(z/m)*grad_face(m)) + norm*bc_flux,
 * with `norm` an interior field (Nx, Ny, Nz) or NULL for the scalar
 * norm_scalar.  The mask has one channel, shared by the g->c channels of z. */
int evx_sbm_rhs_padded(int32_t dtype, const void* z_padded, const void* mask_padded,
                       const void* norm, double norm_scalar, void* out, const evx_grid* g,
                       double D, double bc_flux, void* stream);

/* Elementwise out = a + coef*b (coef rounded to the working precision).
 * Used for generic-path Euler / RK4 stage updates (timesteppers.py:42-61). */
int evx_axpy(int32_t dtype, const void* a, const void* b, double coef,
             void* out, int64_t n, void* stream);

/* ---- spectral (real 3-D FFT) ---------------------------------------------- */
typedef struct evx_fft_plan evx_fft_plan;

/* Plan for per-channel real 3-D transforms of a (channels,nx,ny,nz) field,
 * norm "backward" (unscaled forward, 1/N inverse).  Replaces
 * VoxelGridTorch.rfftn / irfftn (voxelgrid.py:203-207) for C == 1 and their
 * per-channel equivalent for C > 1.  Owns only twiddle tables. */
int evx_fft_plan_create(evx_fft_plan** plan, int32_t dtype, int64_t channels,
                        int64_t nx, int64_t ny, int64_t nz);
int evx_fft_plan_destroy(evx_fft_plan* plan);
/* Bytes of caller-provided device workspace needed by evx_ch_imex_step and
 * evx_rd_imex_step (spectrum plus real staging). */
int evx_fft_plan_workspace_bytes(const evx_fft_plan* plan, size_t* bytes);
/* Number of complex values of one spectrum (channels*nx*ny*(nz/2+1)). */
int evx_fft_plan_spectrum_elems(const evx_fft_plan* plan, int64_t* elems);

/* spec(C,nx,ny,nz/2+1) complex = rfftn(x) (voxelgrid.py:203-204). */
int evx_rfftn(evx_fft_plan* plan, const void* x, void* spec, void* stream);
/* y(C,nx,ny,nz) = irfftn(spec) (voxelgrid.py:206-207); spec is overwritten.
 * Imaginary parts of the kz=0 (and kz=nz/2) bins are ignored as C2R does. */
int evx_irfftn(evx_fft_plan* plan, void* spec, void* y, void* stream);

/* Semi-implicit step parameters.  The k-space symbol follows the reference
 * quirk: k^2 is float32 (voxelgrid.py:84-114), sigma and
 * prefac = dt/(1-dt*sigma) are float32 (timesteppers.py:77). */
typedef struct evx_imex_params {
  double dt;
  double spacing[3];  /* grid spacing (fftfreq's d) */
  double eps;         /* CahnHilliard eps  (problem_definition.py:293)    */
  double D;           /* diffusivity / mobility prefactor                 */
  double A;           /* stabilisation constant (problem_definition.py:296) */
  /* Optional device array of 2*channels doubles {eps_c, D_c}: per-channel
   * parameters for batched independent trajectories (each channel then
   * follows the reference's scalar-parameter step with its own eps, D).
   * NULL: every channel uses eps and D above. */
  const double* chan_params;
} evx_imex_params;

/* u_out = u + irfftn(prefac * rfftn(ch_rhs(u)))  for a fully periodic
 * CahnHilliard + PseudoSpectralIMEX step (timesteppers.py:85-89,
 * problem_definition.py:303,328-371).  u_out may alias u. */
int evx_ch_imex_step(evx_fft_plan* plan, const void* u, void* u_out,
                     void* workspace, const evx_imex_params* p, void* stream);

/* Same for ReactionDiffusion (f == 0; symbol -D*A*k^2, problem_definition.py:206). */
int evx_rd_imex_step(evx_fft_plan* plan, const void* u, void* u_out,
                     void* workspace, const evx_imex_params* p, void* stream);

/* Exponential Euler (ETD1) steps, u_out = u + irfftn(dt*phi_1(dt*sigma) *
 * rfftn(rhs(u))), phi_1(z) = (exp(z)-1)/z with the (6,6) Pade form for
 * |z| < 0.5, on the same float32 symbol as the semi-implicit steps.
 * Replace ExponentialEuler.step (timesteppers.py:136-199) for CahnHilliard
 * and ReactionDiffusion (f == 0) on periodic grids.  Same workspace and
 * parameters as evx_ch_imex_step / evx_rd_imex_step. */
int evx_ch_etd1_step(evx_fft_plan* plan, const void* u, void* u_out,
                     void* workspace, const evx_imex_params* p, void* stream);
int evx_rd_etd1_step(evx_fft_plan* plan, const void* u, void* u_out,
                     void* workspace, const evx_imex_params* p, void* stream);

/* PseudoSpectralIMEX.step / ExponentialEuler.step (etd = 1) of CahnHilliard
 * (ch entry) or ReactionDiffusion (rd entry) with default closures, one
 * channel, x non-periodic and y, z periodic (timesteppers.py:85-89,195-199,
 * problem_definition.py:138-177): the right-hand side of u (1, n, Ny, Nz) is
 * evaluated with the ghost rules of `bc`, mirror-extended along x as
 * pad_fft_bc does (boundary_conditions.py:61-71,133-144: [f, flip f] zero
 * flux, [f, -flip f] Dirichlet, [f, 0, -flip f, 0] staggered_x Dirichlet),
 * multiplied in k-space on the extended grid and the first n planes are
 * added to u.  The plan is made for (1, 2n, Ny, Nz) (cell_center) or
 * (1, 2n+2, Ny, Nz) (staggered = 1, Dirichlet only; the CH entry needs
 * cell_center); its workspace as for the periodic steps. */
int evx_ch_imex_step_mirror(evx_fft_plan* plan, const void* u, void* u_out, void* workspace,
                            const evx_imex_params* p, const evx_bc* bc, int32_t staggered,
                            int32_t etd, void* stream);
int evx_rd_imex_step_mirror(evx_fft_plan* plan, const void* u, void* u_out, void* workspace,
                            const evx_imex_params* p, const evx_bc* bc, int32_t staggered,
                            int32_t etd, void* stream);

/* ---- adjoints (reverse mode) ----------------------------------------------
 * The reference differentiates its solvers with JAX (inversion.py:20-233);
 * these entry points give the vector-Jacobian products of the torch path's
 * CahnHilliard step for gradient-based parameter estimation.  Periodic
 * boundaries only.  Parameter gradients are accumulated (+=) into a device
 * double array dparams[2*channels] = {dJ/deps_c, dJ/dD_c}, in a fixed
 * reduction order (deterministic).  The clip of CahnHilliard.rhs passes
 * gradient where 0 <= u <= 1, as torch.clip's backward does. */

/* Device workspace bytes of evx_ch_rhs_vjp for a (channels,nx,ny,nz) field. */
int evx_ch_rhs_vjp_workspace_bytes(int32_t dtype, const evx_grid* g,
                                   size_t* bytes);

/* grad_u = base + (d rhs/d u)^T w for r = ch_rhs(u) (problem_definition.py:
 * 328-371, periodic); base may be NULL (zero) and may alias grad_u.
 * dparams += (w . dr/deps, w . dr/dD) per channel.  chan_params as in
 * evx_imex_params (NULL: eps, D for every channel). */
int evx_ch_rhs_vjp(int32_t dtype, const void* u, const void* w,
                   const void* base, void* grad_u, double* dparams,
                   const evx_grid* g, double eps, double D,
                   const double* chan_params, void* workspace, void* stream);

/* Device workspace bytes of evx_ch_imex_step_vjp for this plan. */
int evx_ch_imex_vjp_workspace_bytes(const evx_fft_plan* plan, size_t* bytes);

/* Reverse step of evx_ch_imex_step, u1 = u0 + L(rhs(u0)) with
 * L = irfftn(prefac * rfftn(.)) (timesteppers.py:85-89):
 *   lam0 = lam1 + (d rhs/d u0)^T L(lam1)
 *   dparams += the eps / D gradients through rhs and through prefac.
 * L is self-adjoint (real, even multiplier).  The prefac term uses
 * d prefac/d theta = prefac*(prefac-dt)/(dt*theta) for theta in {eps, D}
 * (sigma is proportional to eps*D), so it needs u1 - u0 and no extra
 * transform.  u0 and u1 are the forward step's input and output; lam0 may
 * alias lam1. */
int evx_ch_imex_step_vjp(evx_fft_plan* plan, const void* u0, const void* u1,
                         const void* lam1, void* lam0, double* dparams,
                         void* workspace, const evx_imex_params* p,
                         void* stream);

/* ---- batched grids --------------------------------------------------------
 * Many independent periodic Cahn-Hilliard grids of different shapes, each
 * advanced by the semi-implicit step of evx_ch_imex_step (bit-identical per
 * grid), float32.  No reference counterpart: the reference advances one
 * grid per solver (solvers.py:80-114).  Edges must be in
 * {16, 32, 48, 64, 96, 128}.  The state is one packed float array: grid i
 * starts at element offsets[i] of evx_batch_state_layout. */
typedef struct evx_batch_plan evx_batch_plan;
int evx_batch_plan_create(evx_batch_plan** plan, int32_t dtype, int32_t n_grids,
                          const int64_t* shapes /* n_grids x 3 */,
                          const double* spacing /* n_grids x 3 */);
int evx_batch_plan_destroy(evx_batch_plan* plan);
int evx_batch_state_layout(const evx_batch_plan* plan, int64_t* offsets,
                           int64_t* total_elems);
int evx_batch_workspace_bytes(const evx_batch_plan* plan, size_t* bytes);
/* Per-grid (eps, D) as a host array eps_D[2*i], eps_D[2*i+1]; dt and A as in
 * evx_imex_params.  Set-up call (synchronous upload). */
int evx_batch_set_params(evx_batch_plan* plan, const double* eps_D, double dt,
                         double A);
/* x planes per Cahn-Hilliard tile (default 8); the other sweeps' tiles
 * follow from the FFT sizes.  Set-up call: no step of this plan may be in
 * flight. */
int evx_batch_set_tiles(evx_batch_plan* plan, int xchunk);
int evx_batch_get_tiles(const evx_batch_plan* plan, int* xchunk, int* n_items);
/* Consecutive z tiles or y/x bricks of one grid per workgroup of the five
 * FFT sweeps (1..64, default 1; a workgroup runs them back to back).  The
 * results do not depend on it.  Set-up call, as evx_batch_set_tiles. */
int evx_batch_set_fft_group(evx_batch_plan* plan, int tiles_per_item);
int evx_batch_get_fft_group(const evx_batch_plan* plan, int* tiles_per_item);
/* The step as one cooperative launch (1) instead of one launch per sweep (0,
 * default): every block resident, the blocks walking each sweep's items and
 * meeting at the runtime's grid barrier between sweeps.  Same results as the
 * six launches.  EVX_ERR_UNSUPPORTED where the device has no cooperative
 * launches.  Set-up call, as evx_batch_set_tiles. */
int evx_batch_set_one_launch(evx_batch_plan* plan, int one_launch);
int evx_batch_get_one_launch(const evx_batch_plan* plan, int* one_launch);
/* Times the candidate schedules on scratch_state (overwritten) and keeps
 * the fastest: the Cahn-Hilliard x-chunk (4, 6, 8, 12, 16, 24, 32, 48, 64), then the FFT
 * group (1, 2, 4, 8) at that x-chunk, both with one launch per sweep, then
 * one cooperative launch per step against one launch per sweep; best_ms
 * (may be NULL) receives the chosen schedule's time per step.  On an error
 * the launch schedule is left as it was.  Synchronises the stream. */
int evx_batch_autotune(evx_batch_plan* plan, void* scratch_state,
                       void* workspace, int steps, void* stream,
                       double* best_ms);
/* One semi-implicit step of every grid, in place on the packed state: six
 * launches (one per sweep), each covering every grid.  Asynchronous,
 * allocation-free, capturable in a graph. */
int evx_batch_ch_imex_step(evx_batch_plan* plan, void* state, void* workspace,
                           void* stream);
/* Launch timing for benchmarks (no reference counterpart; the reference
 * times whole runs with a wall clock, solvers.py:108-114).  Arms the next
 * evx_ch_imex_step / evx_rd_imex_step on this plan to record the caller's
 * hipEvent_t handles: events[0] before its first kernel and events[i] after
 * its i-th kernel, for as many kernels as there are events (at most 64).
 * The arming is consumed by that one step.  evx_fft_plan_trace_kinds then
 * gives the kernel of each recorded interval: 0 right-hand side, 1 z R2C,
 * 2 forward y, 3 x forward*prefac*inverse, 4 inverse y, 5 z C2R + add,
 * 6 right-hand side fused with the z R2C, 7 other. */
int evx_fft_plan_trace_next_step(evx_fft_plan* plan, void* const* events,
                                 int n_events);
int evx_fft_plan_trace_kinds(const evx_fft_plan* plan, int32_t* kinds,
                             int32_t* n_intervals);

/* Options of a plan (no reference counterpart).
 *   EVX_OPT_SPECIALIZED   1 (default): size-specialised FFT kernels where one
 *                         exists; 0: runtime-radix kernels for every size
 *                         (tests and benchmarks compare the two).
 *   EVX_OPT_EXACT_SYMBOL  0 (default): the semi-implicit multiplier follows
 *                         the reference's float32 k-table; 1: it is computed
 *                         in float64 from the exact wavenumbers, a smooth
 *                         function of eps and D for finite-difference checks
 *                         and float64 parameter estimation.
 *   EVX_OPT_SLABS         0 (default): evx_ch_imex_step runs each sweep over
 *                         the whole grid; K > 1 (one-channel plans): the
 *                         right-hand side, z and y sweeps run x-slab by
 *                         x-slab so a slab's intermediates stay in the
 *                         last-level cache (trace events then mark the four
 *                         phase boundaries).
 *   EVX_OPT_FUSED_RHS     1 (default): float32 Cahn-Hilliard steps with
 *                         nz = 512 and ny a multiple of 16 compute the
 *                         right-hand side inside the z R2C sweep (five
 *                         sweeps per step); 0: separate right-hand-side
 *                         kernel and z R2C (six sweeps).
 *   EVX_OPT_STEP_SUMS     1 (default): one-channel semi-implicit steps whose
 *                         z C2R is size-specialised leave per-block records
 *                         of the state they write (compensated float64 sum,
 *                         NaN / Inf flags) in the workspace, at no extra HBM
 *                         traffic; 0: no records. */
enum evx_plan_option { EVX_OPT_SPECIALIZED = 1, EVX_OPT_EXACT_SYMBOL = 2, EVX_OPT_SLABS = 3,
                       EVX_OPT_FUSED_RHS = 4, EVX_OPT_STEP_SUMS = 5 };
int evx_fft_plan_set_option(evx_fft_plan* plan, int option, int value);

/* Mass and non-finite flags of the state the last semi-implicit step of this
 * plan wrote with workspace `workspace` (evx_ch_imex_step, evx_rd_imex_step,
 * evx_ch_etd1_step, evx_rd_etd1_step): out[0] = its compensated float64 sum
 * (the numerator of VoxelGrid.average, voxelgrid.py:137-141) folded from the
 * step's per-block records; nonfinite, if not NULL, as for evx_sum_f64
 * (set to 1 on NaN / Inf, left untouched otherwise; the per-frame NaN check
 * of solvers.py:163-165).  EVX_ERR_INVALID when that step kept no records. */
int evx_fft_plan_step_sums(evx_fft_plan* plan, const void* workspace, double* out,
                           int32_t* nonfinite, void* stream);

/* ---- reductions ----------------------------------------------------------- */

/* Bytes of device workspace evx_sum_f64 / evx_ch_energy need for `channels`. */
int evx_reduce_workspace_bytes(int64_t channels, size_t* bytes);

/* out[c] = sum over channel c of x, accumulated in compensated float64
 * (the numerator of VoxelGrid.average, voxelgrid.py:137-141).  out is a
 * device double[channels]; nonfinite, if not NULL, is a device int32[2]:
 * [0] is set to 1 when any element is NaN (the per-frame check of
 * solvers.py:163-165), [1] when any is +-Inf; both are left untouched
 * otherwise, so the caller zeroes them.  Deterministic: fixed reduction order. */
int evx_sum_f64(int32_t dtype, const void* x, int64_t channels,
                int64_t elems_per_channel, double* out, int32_t* nonfinite,
                void* workspace, void* stream);

/* evx_sum_f64 without a workspace argument and without any allocation: the
 * per-block partials live in a static device array of the library (1 MiB per
 * GPU, held until the process ends), so at most EVX_NOALLOC_MAX_CHANNELS
 * channels; more returns EVX_ERR_INVALID before any launch.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Results equal evx_sum_f64's bit for bit.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (two streams,
 * two threads) race and must be serialised by the caller. */
#define EVX_NOALLOC_MAX_CHANNELS 64
int evx_sum_f64_noalloc(int32_t dtype, const void* x, int64_t channels,
                        int64_t elems_per_channel, double* out,
                        int32_t* nonfinite, void* stream);

/* Discrete CH free energy per channel in float64:
 *   sum_voxels (eps*|grad_center c|^2 + 9/eps*c^2(1-c)^2) * dv
 * with centred gradients (fd_stencils.py:44-60) on pad_bc(c).  The reference
 * defines the functional (docs/notebooks/01-using-solvers.ipynb) but no
 * function for it; this entry point has no reference counterpart. */
int evx_ch_energy(int32_t dtype, const void* c, const evx_grid* g,
                  const evx_bc* bc, double eps, double dv, double* out,
                  void* workspace, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* EVOXELS_HIP_H */
