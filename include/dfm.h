/*
 * dfm.h -- C-ABI of libdfm.so, the MI355X (gfx950) HIP implementation of the
 * leapfrog flow-map fluid step and its adjoint.
 *
 * The reference (pearseven/differentiable-flowmap) has no native FFI on this
 * path: its kernels are Taichi @ti.kernel functions called from Python.  Each
 * entry point below replaces one of those kernels (or a fixed composition of
 * them), cited as file:line; the Python drop-in (dfm.simulator) binds them
 * with ctypes.
 *
 * Conventions
 *  - All field pointers are device pointers (e.g. torch.Tensor.data_ptr()).
 *  - Scalar fields are dense (sx, sy, sz) arrays, z fastest; 2D uses sz = 1.
 *    MAC component a lives on the face grid n + e_a; cell fields on n.
 *  - Flow-map point sets are SoA: component c of point i at [c*npts + i].
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *    asynchronous and never synchronise the host, except where noted.
 *  - Return value: 0 = OK, < 0 = error (see DFM_E_*); dfm_last_error()
 *    returns a thread-local description of the last error.
 */
#ifndef DFM_H
#define DFM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFM_OK 0
#define DFM_E_ARG (-1)        /* bad dimension, order, kernel or size */
#define DFM_E_NULL (-2)       /* required pointer is NULL */
#define DFM_E_HIP (-3)        /* HIP runtime error (launch / allocation) */
#define DFM_E_NOCONV (-4)     /* PCG reached max_iters without converging */
#define DFM_E_NONFINITE (-5)  /* PCG residual became NaN/Inf */

/* The flow-map entry points accept grids whose face grids hold fewer than
 * 2^30 values and, in 3D, whose (x, y) planes (in 2D, each extent + 1) stay
 * below 2^24 (the samplers address fields with 32-bit offsets and 24-bit
 * multiplies); larger grids give DFM_E_ARG. */
typedef struct dfm_grid {
  int32_t dim;  /* 2 or 3 */
  int32_t n[3]; /* cells per axis; n[2] = 1 in 2D */
  float dx;     /* cell size */
} dfm_grid;

/* MAC field: c[a] is the face-a component (c[2] = NULL in 2D). */
typedef struct dfm_mac {
  float* c[3];
} dfm_mac;

/* Flow map: cell-centred map psi (dim x cells, may be NULL), per face axis the
 * face map psi_f[a] (dim x faces_a) and one Jacobian column T_f[a]
 * (dim x faces_a; NULL -> march positions only). */
typedef struct dfm_flowmap {
  float* psi;
  float* psi_f[3];
  float* T_f[3];
} dfm_flowmap;

const char* dfm_last_error(void);
int dfm_version(void);

/* reset_to_identity: psi = X, psi_a = X_a, T_a = e_a.
 * Replaces 2D/vortex/flowmap.py:343-358, 3D/flowmap.py:336-356. */
int dfm_reset_identity(const dfm_grid* g, const dfm_flowmap* m, void* stream);

/* RK2/RK4 march of every point set of m through velocity u:
 * psi += -dt u(psi), T += -dt (grad u)(psi) T (quadratic MAC interpolation).
 * One launch covers RK_grid on all face sets and RK_grid_only_psi on psi.
 * Replaces RK_grid / RK_grid_only_psi, 2D/vortex/flowmap.py:40-57,125-215;
 * 3D/flowmap.py:6-25,85-179. */
int dfm_march(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m, float dt, int order, void* stream);

/* dfm_march with at most blocks_per_cu of its workgroups resident on a CU at
 * once (0: no cap; 1..16), through an LDS request the kernel does not use.
 * The results are dfm_march's, bit for bit.  No counterpart in the
 * reference: for a march queued on a second stream beside latency-bound work
 * (the projection's one-workgroup coarse levels need a whole CU free), which
 * otherwise waits for march workgroups to retire.  Caps the LDS cannot
 * enforce return DFM_E_ARG. */
int dfm_march_capped(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m, float dt, int order,
                     int blocks_per_cu, void* stream);

/* Impulse pullback with epilogue: out_a = alpha * (T_a . src(psi_a)) + beta * add_a
 * (add may be NULL when beta == 0; out may alias add).
 * alpha=1, beta=0: advect_u, 2D/vortex/flowmap.py:83-97, 3D/flowmap.py:51-69.
 * alpha=dt, beta=1, add=out=init: accumulate_init, 2D/vortex/flowmap.py:508-520. */
int dfm_pullback_u(const dfm_grid* g, const dfm_mac* src, const dfm_flowmap* m, float alpha, float beta,
                   const dfm_mac* add, const dfm_mac* out, void* stream);

/* Scalar pullback: out = alpha * src(psi) + beta * add (cells, quadratic).
 * Replaces advect_scalar, 2D/vortex/flowmap.py:99-107, 3D/flowmap.py:71-79. */
int dfm_pullback_s(const dfm_grid* g, const float* src, const float* psi, float alpha, float beta, const float* add,
                   float* out, void* stream);

/* BFECC of a MAC field (BFECC_clamp = False): out = T.init(psi);
 * err = (F.out(phi) - init)/2; out -= T.err(psi), where (T, psi) = fwd and
 * (F, phi) = bwd.  err is caller-provided scratch.  out must not alias init.
 * Replaces BFECC, 2D/vortex/flowmap.py:440-463, 3D/flowmap.py:405-483. */
int dfm_bfecc_u(const dfm_grid* g, const dfm_mac* init, const dfm_flowmap* fwd, const dfm_flowmap* bwd,
                const dfm_mac* err, const dfm_mac* out, void* stream);

/* dfm_bfecc_u without the err argument, allocating nothing: the error field
 * lives in a library-owned device array of DFM_NOALLOC_MAX_VALUES floats
 * (32 MiB per GPU, kept until the process ends), so a grid whose face values
 * (sum over a of the face-grid sizes; 6,340,608 at 128^3) exceed that gives
 * DFM_E_ARG.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (different streams or threads) race
 * and must be ordered by the caller.  The first call on each device looks
 * the array up (hipGetSymbolAddress) with that device current. */
#define DFM_NOALLOC_MAX_VALUES (1 << 23)
int dfm_bfecc_u_noalloc(const dfm_grid* g, const dfm_mac* init, const dfm_flowmap* fwd, const dfm_flowmap* bwd,
                        const dfm_mac* out, void* stream);

/* BFECC of a cell scalar.  Replaces BFECC_scalar, 2D/vortex/flowmap.py:465-481. */
int dfm_bfecc_s(const dfm_grid* g, const float* init, const float* psi, const float* phi, float* err, float* out,
                void* stream);

/* BFECC with BFECC_clamp = True: as dfm_bfecc_u, then every face value is
 * clamped to the range of the first pullback T.init(psi) at the face's 2*dim
 * axis neighbours on its own grid (clamp-to-edge indices).  tmp (caller
 * scratch) holds that first pullback; out must not alias init, err or tmp.
 * Replaces BFECC + clamp_u, 2D/vortex/flowmap.py:332-341,440-463,
 * 3D/flowmap.py:323-334,405-434. */
int dfm_bfecc_u_clamp(const dfm_grid* g, const dfm_mac* init, const dfm_flowmap* fwd, const dfm_flowmap* bwd,
                      const dfm_mac* err, const dfm_mac* tmp, const dfm_mac* out, void* stream);

/* BFECC_scalar with BFECC_clamp = True (cells).  Replaces BFECC_scalar +
 * clamp_u, 2D/vortex/flowmap.py:465-481, 3D/flowmap.py:436-452. */
int dfm_bfecc_s_clamp(const dfm_grid* g, const float* init, const float* psi, const float* phi, float* err, float* tmp,
                      float* out, void* stream);

/* accumulate_init_new (AB3 (23,-16,5)/12 with the reference's ind==1 form):
 * g = T_a . f(psi_a); init += dt * AB3(g, h1, h2); h2 = h1; h1 = g.
 * If f2 is non-NULL, additionally init += dt2 * (T_a . f2(psi_a))
 * (accumulate_init of a second source through the same map).
 * Replaces 2D/vortex/flowmap.py:508-553, 3D/flowmap.py:485-545. */
int dfm_accumulate(const dfm_grid* g, const dfm_mac* f, const dfm_flowmap* m, const dfm_mac* init,
                   const dfm_mac* h1, const dfm_mac* h2, int ind, float dt, const dfm_mac* f2, float dt2,
                   void* stream);

/* Semi-Lagrangian advection: back-trace every face node (and, if nscal > 0,
 * every cell centre) through vel over dt with RK order 2/4, then sample
 * component a of each of the nvec MAC sources (into dst_vec) and each of the
 * nscal cell sources (into dst_s) there.  nvec <= 3, nscal <= 8.
 * Destinations must not alias sources.
 * Replaces advect_u_grid / advect_scalar_grid (RK2_/RK4_advect_*_grid),
 * 2D/vortex/flowmap.py:59-81,233-329; 3D/flowmap.py:26-50,202-320. */
int dfm_sl_advect(const dfm_grid* g, const dfm_mac* vel, float dt, int order, int nvec, const dfm_mac* src_vec,
                  const dfm_mac* dst_vec, int nscal, const float* const* src_s, float* const* dst_s, void* stream);

/* ---------------------------------------- interpolation-kernel variants --
 * The pullbacks, BFECC and semi-Lagrangian advection above sample the
 * advected quantity with the quadratic B-spline N_2, as the reference's
 * simulators do.  These variants take the sampling spline as kern: 1 = the
 * linear N_1 of interp_1 / interp_u_MAC_1 (2D/vortex/taichi_utils.py:210-247,
 * 647-652; 3D/taichi_utils.py:262-325; sample coordinate clamped to
 * [0, n-1-1e-6]), 2 = quadratic (the entry points above), 3 = the cubic N_3
 * of interp_grad_3 (2D/vortex/taichi_utils.py:450-484).  The velocity that
 * moves points (the march and the semi-Lagrangian back-trace) stays
 * quadratic.  Any other kern gives DFM_E_ARG. */
int dfm_pullback_u_k(const dfm_grid* g, int kern, const dfm_mac* src, const dfm_flowmap* m, float alpha, float beta,
                     const dfm_mac* add, const dfm_mac* out, void* stream);
int dfm_pullback_s_k(const dfm_grid* g, int kern, const float* src, const float* psi, float alpha, float beta,
                     const float* add, float* out, void* stream);
int dfm_bfecc_u_k(const dfm_grid* g, int kern, const dfm_mac* init, const dfm_flowmap* fwd, const dfm_flowmap* bwd,
                  const dfm_mac* err, const dfm_mac* out, void* stream);
int dfm_bfecc_s_k(const dfm_grid* g, int kern, const float* init, const float* psi, const float* phi, float* err,
                  float* out, void* stream);
int dfm_sl_advect_k(const dfm_grid* g, int kern, const dfm_mac* vel, float dt, int order, int nvec,
                    const dfm_mac* src_vec, const dfm_mac* dst_vec, int nscal, const float* const* src_s,
                    float* const* dst_s, void* stream);

/* Library-wide kernel choices (no counterpart in the reference: A/B switches
 * of this library's kernels; the results do not depend on them).
 * DFM_FM_TILED_SOURCES: 3D adjoint sources (dfm_adj_conv_source,
 * dfm_adj_scalar_source) by the LDS-tiled fixed-stencil kernel (1, default)
 * or by per-point gathers (0); both give bitwise the same fields. */
#define DFM_FM_TILED_SOURCES 0
int dfm_flowmap_set_option(int option, int value);
int dfm_flowmap_get_option(int option, int* value);

/* Adjoint convective source r_a(X_a) = ((grad u)^T w)_a with the B-spline of
 * order kern (4 = quartic as in 2D, 3 = cubic as in 3D).
 * Replaces calculate_nabla_u_w, 2D/vortex/flowmap.py:361-377, 3D/flowmap.py:359-380. */
int dfm_adj_conv_source(const dfm_grid* g, const dfm_mac* u, const dfm_mac* w, int kern, const dfm_mac* r,
                        void* stream);

/* Adjoint scalar source summed over nscal passive pairs (cubic):
 * r_a(X_a) = sum_m d_a s_m(X_a) * lam_m(X_a).  nscal <= 8.
 * Replaces calculate_nabla_scalar_adjoint, 2D/vortex/flowmap.py:379-395,
 * 3D/flowmap.py:382-403. */
int dfm_adj_scalar_source(const dfm_grid* g, int nscal, const float* const* s, const float* const* lam,
                          const dfm_mac* r, void* stream);

/* ------------------------------------------- vector-Jacobian products ----
 * Reverse-mode derivatives of the forward stages (no counterpart in the
 * reference, which differentiates by its continuous adjoint; they back the
 * per-stage torch.autograd.Functions).  The transpose of a B-spline sample
 * scatters the incoming value with the sample's own weights to its own
 * clamped taps by FP32 atomic adds: results agree to rounding from run to
 * run.  Every output gradient is ACCUMULATED (the caller zeroes it first).
 * A sample's derivative w.r.t. its point is zero along an axis where the
 * sample coordinate is clamped at the wall band, as the forward's value is
 * constant there.  kern is the sampling spline of the forward call (1, 2 or
 * 3, as the _k variants; 2 for the plain entry points).
 *
 * dfm_pullback_u_vjp: for dfm_pullback_u's alpha (T_a . src(psi_a)) part,
 * g_src_b += alpha T_a[b] g_out_a at psi_a; with g_m non-NULL (psi_f / T_f
 * sets; src then required) also the map's gradients g_T_a[b] += alpha g_out_a
 * src_b(psi_a) and g_psi_a += alpha g_out_a sum_b T_a[b] grad src_b(psi_a). */
int dfm_pullback_u_vjp(const dfm_grid* g, int kern, const dfm_mac* g_out, const dfm_mac* src, const dfm_flowmap* m,
                       float alpha, const dfm_mac* g_src, const dfm_flowmap* g_m, void* stream);
/* dfm_pullback_s_vjp: for dfm_pullback_s's alpha src(psi) part, g_src +=
 * alpha g_out at psi; with g_psi non-NULL ((dim, *n), src then required)
 * also g_psi += alpha g_out grad src(psi). */
int dfm_pullback_s_vjp(const dfm_grid* g, int kern, const float* g_out, const float* src, const float* psi,
                       float alpha, float* g_src, float* g_psi, void* stream);
/* dfm_sl_advect_vjp: for dfm_sl_advect (dst = src at the back-traced node),
 * g_src_* += the transpose of the sample applied to g_dst_*; with g_vel
 * non-NULL (src_vec / src_s then required, the forward sources) also the
 * advecting velocity's gradient through the RK back-trace. */
int dfm_sl_advect_vjp(const dfm_grid* g, int kern, const dfm_mac* vel, float dt, int order, int nvec,
                      const dfm_mac* src_vec,
                      const dfm_mac* g_dst_vec, const dfm_mac* g_src_vec, int nscal, const float* const* src_s,
                      const float* const* g_dst_s, float* const* g_src_s, const dfm_mac* g_vel, void* stream);
/* dfm_march_vjp: the vector-Jacobian product of dfm_march.  m0 is the map
 * before the march (read only); gm holds on entry the gradients w.r.t. the
 * marched map's psi / T sets (the same sets as m0) and on return those
 * w.r.t. m0's (replaced in place, not accumulated); g_u accumulates the
 * velocity's.  The RK stages are recomputed from m0. */
int dfm_march_vjp(const dfm_grid* g, const dfm_mac* u, const dfm_flowmap* m0, float dt, int order,
                  const dfm_flowmap* gm, const dfm_mac* g_u, void* stream);

/* out_a = x_a + alpha * y_a + beta * z_a on every face grid (z may be NULL
 * when beta == 0).  Replaces add_fields / scale_field on MAC fields,
 * 2D/vortex/taichi_utils.py:105-113. */
int dfm_axpy_mac(const dfm_grid* g, const dfm_mac* x, float alpha, const dfm_mac* y, float beta, const dfm_mac* z,
                 const dfm_mac* out, void* stream);

/* ------------------------------------------------ batched 2D instances --
 * Forward + adjoint of many independent 2D problems of uneven size in one
 * engine: the reference's per-problem LFM_Simulator.step_flow_map /
 * LFM_Diff_Simulator.backtrack_flow_map (2D/vortex/lfm_midpoint_simulator.py:
 * 517-661, 2D/vortex/lfm_midpoint_diff_simulator.py:260-503; RK2, quartic
 * convective source, plain BFECC, every midpoint velocity and forward map
 * kept), with every stage one launch over all instances, a batched MGPCG
 * (per-instance FP64 scalars, convergence flags and level counts,
 * int(log2(min n)) - 2 as 2D/vortex/mgpcg.py:28) and the launches between
 * two convergence checks replayed from captured hipGraphs.  No counterpart
 * in the reference, which runs one problem per process.  Extents must be
 * 8..2048 and divisible by 2^levels.  The engine owns its device memory
 * (allocated at creation). */
typedef struct {
  int n[2];  /* cells per axis */
  float dx;
  float dt;
} dfm_batch_instance;
typedef struct dfm_batch dfm_batch;

int dfm_batch_create(const dfm_batch_instance* inst, int count, int reinit_every, int total_steps, double tol,
                     int max_iters, dfm_batch** out);
int dfm_batch_destroy(dfm_batch* b);

/* Device pointers of instance inst's field in the engine: MAC fields give
 * the x- and y-face arrays ((n0+1) x n1 and n0 x (n1+1), row-major), cell
 * fields c0 only (n0 x n1).  The caller writes u (unprojected initial
 * velocity), s (density) and target before dfm_batch_init, and reads u, s,
 * adj_u and adj_s. */
#define DFM_BATCH_U 0
#define DFM_BATCH_S 1
#define DFM_BATCH_TARGET 2
#define DFM_BATCH_ADJ_U 3
#define DFM_BATCH_ADJ_S 4
int dfm_batch_field(dfm_batch* b, int field, int inst, float** c0, float** c1);
/* Copies instance inst's field between the engine and caller device buffers
 * of the same shapes (c1: the y faces of a MAC field, unused for cells):
 * to_engine = 1 writes the engine's field, 0 reads it.  Asynchronous on
 * stream. */
int dfm_batch_copy(dfm_batch* b, int field, int inst, float* c0, float* c1, int to_engine, void* stream);

/* DFM_BATCH_GRAPH: 1 = replay captured graphs (default), 0 = launch eagerly.
 * DFM_BATCH_FIRST_ITERS: PCG iterations inside the segment that starts a
 * projection (default 8); DFM_BATCH_CHUNK: iterations per further replay
 * until every instance has converged (default 4).  DFM_BATCH_TAIL_CELLS:
 * the largest level (in cells) an instance's one-workgroup LDS tail holds
 * (default 16384; each instance's finer levels run as batched global
 * launches), the batched form's tile-shape knob; setting it re-plans the
 * levels and reallocates the level arrays (between runs only).
 * DFM_BATCH_BOTTOM_EXACT: 1 = solve the coarsest level exactly by DCT-II
 * (default, as the single-grid simulator's mg_bottom_exact), 0 = the
 * reference's 50 red-black sweeps (2D/vortex/mgpcg.py:31,180-182).  Setting
 * an option drops the captured graphs. */
#define DFM_BATCH_GRAPH 0
#define DFM_BATCH_FIRST_ITERS 1
#define DFM_BATCH_CHUNK 2
#define DFM_BATCH_TAIL_CELLS 3
#define DFM_BATCH_BOTTOM_EXACT 4
/* DFM_BATCH_TILE: edge (16 or 32, default 32) of the square tiles of the
 * LDS-tiled smoothing legs on the global multigrid levels. */
#define DFM_BATCH_TILE 5
int dfm_batch_set_option(dfm_batch* b, int option, int value);

/* LFM_Simulator.init after the fields are written: project u, init_u = u,
 * init_s = s, reset the forward map. */
int dfm_batch_init(dfm_batch* b, void* stream);
/* step_midpoint, `steps` times (at most total_steps in all). */
int dfm_batch_forward(dfm_batch* b, int steps, void* stream);
/* init_gradient after the last forward step: adj_s = 2 (s - target),
 * adj_u = 0; loss[inst] = sum (s - target)^2 (host array of count, or NULL). */
int dfm_batch_init_gradient(dfm_batch* b, double* loss, void* stream);
/* backtrack_step_midpoint, `steps` times. */
int dfm_batch_backward(dfm_batch* b, int steps, void* stream);
/* Projections run, the most PCG iterations any instance took, and how many
 * instance solves stopped without converging (max_iters or non-finite). */
int dfm_batch_stats(dfm_batch* b, int* projections, int* max_iters, int* noconv);
/* Multigrid levels of an instance and the first level its LDS tail holds. */
int dfm_batch_levels(dfm_batch* b, int inst, int* levels, int* tail_level);

/* ---------------------------------------------------------------- MGPCG --
 * Handle owning the multigrid workspaces for one grid.  Levels, smoothing and
 * bottom sweeps follow MGPCG (2D/vortex/mgpcg.py:17-31): levels <= 0 picks
 * log2(min n) - 3 + 1, lowered until every n halves exactly that often;
 * explicit levels require every n divisible by 2^(levels-1).  Creation only
 * allocates (it submits no work and waits for none); the handle's counters
 * and control block are zeroed on the stream of the first call that launches
 * work (project, solve or set_mask). */
int dfm_mgpcg_create(const dfm_grid* g, int levels, int pre_post, int bottom, void** handle);
int dfm_mgpcg_destroy(void* handle);

/* Pressure projection in place (MGPCG_2/3.Poisson: apply_bc, divergence,
 * PCG solve of A p = -div with all-Neumann A, subtract_grad_p, apply_bc;
 * 2D/vortex/mgpcg.py:358-368, 3D/mgpcg.py:441-451, 3D/fast_mgpcg.py:241-323).
 * Stops when r.r < tol^2 * b.b (relative != 0) or r.r < tol (relative == 0),
 * or after max_iters; returns DFM_E_NOCONV / DFM_E_NONFINITE in those cases
 * (the gradient is subtracted either way).  The FP32 V-cycle and residual run
 * with FP64 dot products, FP64 alpha/beta and an FP64 pressure iterate.
 * Synchronises the stream once per PCG iteration to poll the device
 * convergence flag; the pressure is kept in the handle. */
int dfm_mgpcg_project(void* handle, const dfm_mac* u, double tol, int relative, int max_iters, void* stream);

/* Solve A x = b (b recentred to mean zero first).  Same stopping rule. */
int dfm_mgpcg_solve(void* handle, const float* b, float* x, double tol, int relative, int max_iters, void* stream);

/* The projection in two halves, so that independent projections on
 * different handles and streams overlap on the device:
 *  _begin queues the divergence, the PCG set-up and as many iterations as
 *  this handle's previous solve needed (no host synchronisation);
 *  _end waits for the convergence flags, queues further iterations if
 *  needed, then the gradient subtraction, on the stream given to _begin.
 * dfm_mgpcg_project is _begin followed by _end.  One begun projection per
 * handle at a time; no other call on the handle in between.  No counterpart
 * in the reference (its solves are synchronous). */
int dfm_mgpcg_project_begin(void* handle, const dfm_mac* u, double tol, int relative, int max_iters, void* stream);
int dfm_mgpcg_project_end(void* handle);

/* Iterations and relative residual sqrt(r.r / b.b) of the last solve. */
int dfm_mgpcg_status(void* handle, int* iters, double* rel_residual);

/* Device pointer of the pressure (cells, FP64, natural cell order) from the
 * last projection.  With a colour-split finest level the pressure is first
 * reordered into a second buffer by a kernel queued on the last solve's
 * stream: read it on that stream (or after synchronising it).  DFM_E_ARG
 * before the handle's first solve. */
int dfm_mgpcg_pressure(void* handle, double** p);

/* Solver options (no counterpart in the reference; tuning of this library).
 *  DFM_MG_BF16_COARSE: 1 = store levels >= 2 in BF16 (FP32 arithmetic):
 *      in the one-workgroup LDS tail, which then holds up to 32^3, and in
 *      the global coarse levels as DFM_MG_BF16_LEVEL says; 0 = FP32.  get:
 *      whether the LDS tail is BF16.
 *  DFM_MG_TAIL: 1 = run the coarse levels that fit in LDS in one kernel
 *      (default); 0 = one launch per sweep on every level.  get: first tail level.
 *  DFM_MG_MAX_REFINE: FP64 iterative-refinement restarts (default 3).
 *  DFM_MG_CHECK_EVERY: PCG iterations between convergence polls (default 1).
 *  DFM_MG_GRAPH: 1 = replay the check_every PCG iterations between polls as
 *      one captured hipGraph (default); 0 = launch each kernel.  get: 2 when a
 *      graph is instantiated.
 *  DFM_MG_BOTTOM_EXACT: 1 = solve the coarsest level exactly (orthonormal
 *      DCT-II in LDS, mean-zero) instead of `bottom` RBGS sweeps; applies when
 *      the coarse tail runs.  Changes only the preconditioner. */
#define DFM_MG_BF16_COARSE 1
#define DFM_MG_TAIL 2
#define DFM_MG_MAX_REFINE 3
#define DFM_MG_CHECK_EVERY 4
#define DFM_MG_GRAPH 5
#define DFM_MG_BOTTOM_EXACT 6
/*  DFM_MG_TAIL_LEVEL: lowest level the LDS tail may start at (default 0: the
 *      largest suffix of levels that fits, with a first level of at most
 *      16384 cells -- one workgroup runs the tail, and a larger level is
 *      faster on the global kernels even where it fits, as 32^3 does in
 *      BF16); levels below it use global kernels.  get: the first tail level in use (L = no tail).  The batch
 *      executor's autotuner sets it per grid size. */
#define DFM_MG_TAIL_LEVEL 7
/*  DFM_MG_FUSED_SWEEPS: smoothing launches of the coarse global levels (at
 *      most 2^19 cells).  2 (default) = where the level has no obstacle mask,
 *      2 pre/post sweeps and extents that are multiples of the tile (8^3 in
 *      3D, 16^2 in 2D), its pre-smoothing with the restriction and its
 *      post-smoothing with the prolongation each run as one LDS-tiled launch
 *      (a halo of 4 cells recomputed per tile), else as 1; 1 = one launch per
 *      red-black sweep, recomputing the first colour's neighbour values; 0 =
 *      two launches per sweep.  All three give bitwise the same results. */
#define DFM_MG_FUSED_SWEEPS 8
/*  DFM_MG_SPLIT_LAYOUT: 1 (default) = store the finest level colour-split
 *  (red cells in one half of each level-0 array, black in the other), so a
 *  red-black half-sweep moves only the cache lines of the colours it uses.
 *  Used when no obstacle mask is set, the finest level is not in the LDS
 *  tail and its innermost extent is even; reading the option gives whether
 *  the next solve will use it.  0 = natural order everywhere.  The results
 *  agree to rounding (the FP64 reductions sum in another order). */
#define DFM_MG_SPLIT_LAYOUT 9
/*  DFM_MG_BF16_LEVEL (get only): the first level the next solve stores in
 *  BF16 in global memory (L = none).  With DFM_MG_BF16_COARSE set, no
 *  obstacle mask and the LDS tail starting at level >= 2, that is level 2
 *  when every global level from 2 to the tail runs the LDS-tiled pre/post
 *  kernels (e.g. 64^3 at 256^3), else the tail's first level; the
 *  restriction into it and the prolongation out of it convert, and the
 *  arithmetic stays FP32. */
#define DFM_MG_BF16_LEVEL 10
/*  DFM_MG_LEVELS (get only): the handle's number of MG levels L. */
#define DFM_MG_LEVELS 11
/*  DFM_MG_ROW_KERNELS: 1 (default) = on a 3D level whose innermost extent is
 *  a multiple of 8 (and twice the next coarser one's), the half-sweeps run one
 *  thread per 4 consecutive cells of a row and colour (16-byte row loads, the
 *  index arithmetic shared by the 4): always on a colour-split finest level,
 *  on a natural-order level when that makes at least 256 workgroups of 256
 *  threads; 2 = the same on every such level; 0 = one thread per cell pair.
 *  The smoothed values are bitwise the same; on a colour-split level the
 *  z.r sums add in another order.  Other values: DFM_E_ARG. */
#define DFM_MG_ROW_KERNELS 12
int dfm_mgpcg_set_option(void* handle, int option, int value);
int dfm_mgpcg_get_option(void* handle, int option, int* value);

/* Obstacle mask for the projection (MGPCG_2_solid, 2D/vortex/mgpcg_solid.py:15-461;
 * the all-solid coarsening downsample_bm :98-112).  mask: device FP32 cells,
 * solid where > 0; NULL removes the mask.  While a mask is set, the operator
 * couples fluid cells only, every face of a solid cell is a wall, solid cells
 * carry zero pressure, the coarse levels run as global kernels (no LDS tail)
 * and the bottom level gets bottom/2 RBGS sweeps in the order (0,1) then
 * bottom/2 in (1,0).  Recentring uses the mean over fluid cells.  Synchronises
 * the stream once (fluid-cell count). */
int dfm_mgpcg_set_mask(void* handle, const float* mask, void* stream);

/* mask_velocity_by_boundary (bvel = cell boundary velocity, SoA (dim, cells))
 * and, with bvel = NULL, mask_adj_velocity_by_boundary
 * (2D/vortex/init_conditions.py:548-556,571-578; 3D/init_conditions.py:535-545,563-572):
 * every face of a cell with mask >= 1 is set to that cell's boundary velocity
 * (0 with bvel = NULL). */
int dfm_mask_velocity(const dfm_grid* g, const float* mask, const float* bvel, const dfm_mac* u, void* stream);

/* mask_passive_by_boundary (adjoint = 0: a cell with mask >= 1 takes the mean
 * of its 3^d neighbours with mask <= 0) and mask_adj_passive_by_boundary
 * (adjoint = 1: such cells are zeroed) (2D/vortex/init_conditions.py:558-569,580-585;
 * 3D/init_conditions.py:547-561,574-578). */
int dfm_mask_scalar(const dfm_grid* g, const float* mask, float* s, int adjoint, void* stream);

/* ---------------------------------------------- 2D vortex control problem */
/* Initial velocity of K vortices on the MAC faces: the cell-centred vortex
 * sum of sixteen_vortex_vel_func_with_pos_coef (2D/vortex/init_conditions.py:368-383,
 * profile angular_vel_func :141-145) split to faces by split_central_vector
 * (2D/vortex/taichi_utils.py:158-167); the caller projects it (the
 * reference's init does, 2D/vortex/lfm_midpoint_simulator.py:420-422).
 * theta: device FP32 [w (K), cx (K), cy (K), rr (K)], radius 0.05 rr. */
int dfm_vortex_velocity_2d(const dfm_grid* g, const float* theta, int K, const dfm_mac* u, void* stream);

/* dtheta (device FP32, 4K, theta's layout) = J^T lam, J = d u_faces / d theta
 * of dfm_vortex_velocity_2d (analytic, FP64 sums, deterministic).  Replaces
 * the per-component finite differences of calculate_dtheta / _ind
 * (2D/vortex/optimize_vortex.py:193-226), which form the same inner product
 * <lam, d u / d theta_i> with eps = 1e-5. */
int dfm_vortex_velocity_2d_vjp(const dfm_grid* g, const float* theta, int K, const dfm_mac* lam, float* dtheta,
                               void* stream);

/* ------------------------------------------------ cell-centre utilities */
/* central (device FP32, SoA: component a of cell c at [a*cells + c]) =
 * 0.5 (u_a[c + e_a] + u_a[c]).  Replaces get_central_vector
 * (2D/vortex/taichi_utils.py:152-156, 3D/taichi_utils.py:163-168). */
int dfm_central_vector(const dfm_grid* g, const dfm_mac* u, float* central, void* stream);

/* u_a(face) = 0.5 (central_a(high cell) + central_a(low cell)), cells clamped
 * to the grid.  Replaces split_central_vector (2D/vortex/taichi_utils.py:158-167,
 * 3D/taichi_utils.py:170-183). */
int dfm_split_central_vector(const dfm_grid* g, const float* central, const dfm_mac* u, void* stream);

/* *max_speed (device FP32) = max(1e-3, max over cells of |central velocity|);
 * NaN propagates.  Replaces calc_max_speed (2D/vortex/lfm_midpoint_simulator.py:328-335,
 * 3D/lfm_midpoint_simulator.py:347-355). */
int dfm_max_speed(const dfm_grid* g, const dfm_mac* u, float* max_speed, void* stream);

/* *count (device, uint32) = number of NaN / Inf entries over `nfields`
 * (<= 8) FP32 device fields of sizes[m] values; asynchronous on `stream` (no
 * host synchronisation: the caller reads *count where it already waits).
 * No counterpart in the reference, which has no NaN checks on this path; the
 * simulators use it to return DFM_E_NONFINITE for a forward state or an
 * adjoint result that is not finite. */
int dfm_count_nonfinite(int nfields, const float* const* fields, const int64_t* sizes, unsigned* count,
                        void* stream);

/* ----------------------------------------- PoissonSolver-compatible entry */
/* Convert a 3D cell field between the dense (n0, n1, n2) layout and the
 * 8^3-tiled flat layout MGPCG_3 passes to its external PoissonSolver
 * (rearrange_tensor_for_cuda / _from_cuda, 3D/fast_mgpcg.py:121-158):
 * to_tiles != 0: dense -> tiled, else tiled -> dense.  Every n must be a
 * multiple of 8; src and dst must not overlap. */
int dfm_tile8_convert(const dfm_grid* g, const float* src, float* dst, int to_tiles, void* stream);

/* Count (into *mismatches, host int) the cells whose coefficients differ from
 * the all-Neumann 7-point operator MGPCG_3.init builds for the external
 * solver (construct_adiag / construct_ax, is_dof = 1; 3D/fast_mgpcg.py:48-74,
 * 162-190); arrays in the tiled layout when tiled != 0.  Synchronises the
 * stream. */
int dfm_poisson_operator_check(const dfm_grid* g, const float* a_diag, const float* a_x, const float* a_y,
                               const float* a_z, const uint8_t* is_dof, int tiled, int* mismatches, void* stream);

/* ------------------------------------------------ Gaussian control forces */
/* Force field of `count` Gaussian forces on every face (fills f; faces no
 * force reaches get 0).  Force I: centre[I*dim + b], inverse squared radius
 * inv_radius[I], strength[I*dim + a] (all device FP32).  At face node p of
 * axis a, with m = |p - centre_I| in FP32, force I adds
 * exp(-(m*m) r_I) * strength_I[a] when (m*m) r_I < 5, in increasing I.
 * Replaces apply_gausion_force / apply_gausion_force_kernel
 * (2D/vortex/lfm_midpoint_simulator.py:302-326, 3D/lfm_midpoint_simulator.py:316-344). */
int dfm_gaussian_force(const dfm_grid* g, const float* center, const float* inv_radius, const float* strength,
                       int count, const dfm_mac* f, void* stream);

/* Parameter adjoint of the force field against the face field adj_u: per
 * force (device FP32, count x dim each) adj_strength[a] = sum e adj_u_a,
 * adj_center = sum s[a] e 2 (p - c) r adj_u_a, penalty_strength[a] =
 * sum 2 (e s[a]) e, penalty_center = sum 2 (e s[a]) s[a] e 2 (p - c) r, over
 * the faces inside the cut-off (the FP32 test above; e and p - c in FP64,
 * FP64 sums, deterministic).
 * Replaces calculate_adj_gausion_force / _kernel
 * (2D/shape/lfm_midpoint_diff_simulator.py:68-103, 3D/lfm_midpoint_diff_simulator.py:63-110). */
int dfm_gaussian_force_adjoint(const dfm_grid* g, const float* center, const float* inv_radius,
                               const float* strength, int count, const dfm_mac* adj_u, float* adj_center,
                               float* adj_strength, float* penalty_center, float* penalty_strength, void* stream);

#ifdef __cplusplus
}
#endif
#endif
