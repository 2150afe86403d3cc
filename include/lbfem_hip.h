/* lbfem_hip: the D2Q9 discrete-Boltzmann time step on continuous Q_p
 * elements (ivan-pi/lbm-fem) on one MI355X (gfx950), behind a C ABI.
 *
 * The entry points mirror the reference's operator and scheme API, so that a
 * driver written against include/lbfem/ of the reference calls this library
 * in their place (each declaration names the reference interface it
 * replaces). Conventions:
 *   - every call returns an int status (LBFEM_OK, LBFEM_EINVAL for what the
 *     reference rejects with AssertThrow, LBFEM_ENOCONV for
 *     SolverControl::NoConvergence, LBFEM_EHIP for a HIP runtime error);
 *     lbfem_last_error() gives the message. No C++ exception crosses the ABI.
 *   - a population vector is Q = 9 blocks of n_nodes values, block a at
 *     offset a * stride (the reference's BlockVector, docs/method.md:165); an
 *     increment or right-hand side has 8 blocks (block a - 1 <-> population a).
 *   - nodes are numbered lexicographically on the (pN+1) x (pN+1) grid, x
 *     fastest; the periodic images of the x = L and y = L lines are kept as
 *     constrained (slave) entries, touched by the nodal operations only, as
 *     deal.II's periodicity constraints do (discretization.h:110-112).
 *   - host-buffer entry points copy; device-pointer entry points take a
 *     hipStream_t (as void*, NULL = the null stream) and do not synchronize.
 *   - handles own their device memory and are not thread-safe (the reference
 *     is single-threaded per rank, discretization.h:115).
 */
#ifndef LBFEM_HIP_H
#define LBFEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { LBFEM_OK = 0, LBFEM_EINVAL = 1, LBFEM_ENOCONV = 2, LBFEM_EHIP = 3 };
enum { LBFEM_LEELIN = 0, LBFEM_BARDOW = 1 };                      /* cgdbe.cc:56-60 */
enum { LBFEM_TG2 = 0, LBFEM_TG3 = 1, LBFEM_TG3_SPLIT = 2 };        /* streaming.h:25-30 */
enum { LBFEM_MASS_CG = 0, LBFEM_MASS_LUMPED = 1, LBFEM_MASS_RICHARDSON = 2 }; /* mass.h:26-31 */
enum { LBFEM_FP64 = 0, LBFEM_MIXED = 1 };

typedef struct lbfem_disc   lbfem_disc;
typedef struct lbfem_scheme lbfem_scheme;

/* Copies the message of the last failed call of this thread (NUL-terminated). */
size_t lbfem_last_error(char *buf, size_t len);
/* Number of visible HIP devices (0 without a GPU). */
int lbfem_device_count(int *count);

/* ---------------------------------------------------------- Discretization
 * Replaces lbfem::Discretization::reinit (include/lbfem/discretization.h:78-144)
 * for the structured box meshes of the reference: N x N Q_p cells on
 * [0, L]^2, periodic in the first n_periodic directions, vertices optionally
 * moved by the tanh clustering of LidDrivenCavity::transform_mesh
 * (include/lbfem/test_cases.h:215-225) when stretch > 0, or by the
 * TaylorGreenVortex distortion (test_cases.h:53-65, --distort) when
 * distort != 0: general quadrilaterals, periodic box only (n_periodic 2,
 * stretch 0), FP64 only, not batched. */
typedef struct
{
  int    n_cells;    /* N, elements per direction (MeshSettings::n_cells) */
  int    degree;     /* p in 1..3 (fe_degree, discretization.h:45) */
  int    n_periodic; /* 0, 1 (x) or 2 (TestCase::n_periodic) */
  double L;          /* box size (TestCase::L) */
  double stretch;    /* tanh clustering gamma, 0 = uniform */
  double distort;    /* x -> x + distort L sin(2 pi x/L) sin(2 pi y/L) (1, 1), 0 = none */
} lbfem_mesh_desc;

typedef struct
{
  int64_t n_nodes;       /* (pN+1)^2, slaves included (the reference's n_dofs) */
  int64_t n_independent; /* non-slave nodes (Discretization::independent) */
  int64_t stride;        /* block stride of device population vectors */
  int32_t n_per_dir;     /* pN + 1 */
  int32_t n_walls;
  double  h_min;         /* Discretization::h_min */
} lbfem_disc_info;

int lbfem_disc_create(const lbfem_mesh_desc *desc, lbfem_disc **disc);
int lbfem_disc_destroy(lbfem_disc *disc);
int lbfem_disc_get_info(const lbfem_disc *disc, lbfem_disc_info *info);
/* Discretization::node (support points), 2 doubles per node. */
int lbfem_disc_node_coordinates(const lbfem_disc *disc, double *xy);
/* 1 for periodic slave nodes (constraints.is_constrained), else 0. */
int lbfem_disc_constrained(const lbfem_disc *disc, int32_t *flag);
/* Discretization::node_weight (row-sum lumped mass, constrained entries 1) and
 * the diagonal of M used by the Jacobi preconditioner (discretization.h:123-125). */
int lbfem_disc_lumped_mass(const lbfem_disc *disc, double *ml);
int lbfem_disc_mass_diagonal(const lbfem_disc *disc, double *diag);
/* Discretization::walls (discretization.h:63-67, 138-142): the wall nodes and
 * the velocity of the wall each is on. Wall nodes must lie on the boundary of
 * the box in a non-periodic direction. Replaces any earlier list. */
int lbfem_disc_set_walls(lbfem_disc *disc, int32_t n, const int32_t *nodes, const double *velocity);
/* Discretization::time_step (discretization.h:158-162): cfl * h_min / p^2. */
int lbfem_disc_time_step(const lbfem_disc *disc, double cfl, double *dt);

/* ---------------------------------------------------------- operators
 * Device pointers, stream as void* (hipStream_t). lbfem_mass_solve (CG and
 * Richardson), lbfem_diagnostics and lbfem_disc_stream_function keep their
 * work vectors in the discretization (allocated on first use) and return
 * after their work on it is done; calls of these on one discretization
 * from several host threads must be serialized by the caller. */

/* lbfem::AdvectionOperator::apply (include/lbfem/advection.h:49-58) with the
 * two implementations of the reference: kind LBFEM_BARDOW is
 * TaylorGalerkinAdvection (advection.h:211-246), LBFEM_LEELIN is
 * LeeLinAdvection (advection.h:256-293; feq required). rhs: 8 blocks, f and
 * feq: 9 blocks, all with the disc's stride. e: the 9 directions of the sweep
 * (18 doubles, host). remove_mass_flux: the wall surface-term option
 * (advection.h:182-199). */
int lbfem_advection_apply(const lbfem_disc *disc, int kind, double *rhs, const double *f, const double *feq,
                          const double *e, double dt, double lambda, int remove_mass_flux, void *stream);

/* lbfem::collide_bgk, compute_equilibrium, predictor_corrector
 * (include/lbfem/collision.h:105-168), over every node, walls imposing
 * their velocity. incr: 8 blocks. */
int lbfem_collide_bgk(const lbfem_disc *disc, double *g, double omega, void *stream);
int lbfem_compute_equilibrium(const lbfem_disc *disc, const double *f, double *feq, void *stream);
int lbfem_predictor_corrector(const lbfem_disc *disc, double *f, const double *feq, const double *incr, double theta,
                              void *stream);

/* y = (M + c K_e) x (lbfem::StreamingMatrix::vmult, mass.h:78-83), constrained
 * rows the identity. */
int lbfem_streaming_vmult(const lbfem_disc *disc, double *y, const double *x, double c, const double *e,
                          void *stream);

/* lbfem::MassSettings (include/lbfem/mass.h:33-48). */
typedef struct
{
  int    type;              /* LBFEM_MASS_* */
  int    richardson;        /* passes for LBFEM_MASS_RICHARDSON */
  double cg_tolerance;      /* relative to |r| */
  int    cg_max_iterations; /* NoConvergence when hit */
  int    fused;             /* accepted for interface parity; the CG updates are always fused here */
} lbfem_mass_settings;

/* lbfem::MassSolver::solve (include/lbfem/mass.h:184-213): x <- A^{-1} r for
 * one population, A = M + tg3_coefficient K_e(e); CG starts from x.
 * *iterations receives the CG steps (0 for the lumped variants). CG and
 * Richardson synchronize the stream (the stopping test reads the residual;
 * the Richardson passes use the discretization's scratch). */
int lbfem_mass_solve(const lbfem_disc *disc, const lbfem_mass_settings *settings, double *x, const double *r,
                     const double *e, double tg3_coefficient, int *iterations, void *stream);

/* No counterpart in the reference: lbfem_mass_solve for callers that cannot
 * allocate (serving and time-stepping loops that must not call hipMalloc, and
 * harnesses that forbid allocation). It takes the same arguments and gives the
 * same x and iteration count, bit for bit, but keeps its work vectors in one
 * fixed device array per GPU (39.3 MB, reserved when the library is loaded on
 * that GPU and held until the process ends) instead of the discretization's
 * scratch. It allocates nothing on any path.
 * Limits: at most 1179648 nodes (the disc's stride) and 7168 operator tiles,
 * else LBFEM_EINVAL. Call it with the discretization's device current. All
 * calls on one GPU share that GPU's array, so calls that overlap on one GPU
 * (from several host threads or streams) race; CG and Richardson synchronize
 * the stream before returning, so consecutive calls do not. */
int lbfem_mass_solve_noalloc(const lbfem_disc *disc, const lbfem_mass_settings *settings, double *x, const double *r,
                             const double *e, double tg3_coefficient, int *iterations, void *stream);

/* ---------------------------------------------------------- schemes
 * lbfem::Scheme, LeeLin, Bardow (include/lbfem/schemes.h:31-142) with
 * lbfem::TaylorGalerkin streaming (include/lbfem/streaming.h:38-130). The
 * scheme owns its populations on the device. */
typedef struct
{
  int                 scheme;    /* LBFEM_LEELIN | LBFEM_BARDOW */
  int                 streaming; /* LBFEM_TG2 | LBFEM_TG3 | LBFEM_TG3_SPLIT */
  lbfem_mass_settings mass;
  int                 precision; /* LBFEM_FP64 | LBFEM_MIXED: FP32 populations stored as
                                    deviations f - w rho0, FP64 moments and equilibria;
                                    with the consistent mass, CG solves on FP32 increments
                                    (FP64 scalars and sums) from the extrapolated FP32
                                    increment, stopping at max(cg_tolerance, floor) |b|,
                                    floor = max(1e-6, 2^-24 n / (2 pi)) for n nodes per
                                    mesh line (the FP32 right-hand side's round-off
                                    level); lumped or CG, not Richardson */
} lbfem_scheme_settings;

int lbfem_scheme_create(const lbfem_disc *disc, const lbfem_scheme_settings *settings, lbfem_scheme **scheme);
int lbfem_scheme_destroy(lbfem_scheme *scheme);
/* Scheme::set_time_step (schemes.h:36-41): {dt, lambda = nu / c_s^2}. */
int lbfem_scheme_set_time_step(lbfem_scheme *scheme, double dt, double lambda);
/* Scheme::set_initial_populations (schemes.h:50-59, Bardow's transform
 * schemes.h:128-141 applied): f0 are the test case's populations at t = 0,
 * 9 blocks of n_nodes doubles on the host. */
int lbfem_scheme_set_initial_populations(lbfem_scheme *scheme, const double *f0);
/* Raw access to Scheme::f (f for LeeLin, g for Bardow), host, 9 x n_nodes. */
int lbfem_scheme_set_populations(lbfem_scheme *scheme, const double *f);
int lbfem_scheme_get_populations(const lbfem_scheme *scheme, double *f);
/* The scheme's device population vector (9 blocks, disc stride) for
 * device-side callers; valid until the next step. */
int lbfem_scheme_device_populations(lbfem_scheme *scheme, double **f);
/* Scheme::step (schemes.h:92-98, 118-124), n_steps times, on the scheme's
 * stream; returns after the work is queued unless CG needs host checks. */
int lbfem_scheme_step(lbfem_scheme *scheme, int n_steps);
int lbfem_scheme_synchronize(lbfem_scheme *scheme);
/* The scheme's HIP stream (hipStream_t as void*), on which step() queues its
 * kernels: for callers that order their own work or events after a step. */
int lbfem_scheme_get_stream(const lbfem_scheme *scheme, void **stream);
/* MassSolver::cg_iterations (mass.h:215), summed over all solves. */
int lbfem_scheme_cg_iterations(const lbfem_scheme *scheme, int64_t *iterations);
/* The CG steps of each solve of the last step (the SolverControl::last_step()
 * of the reference's per-population solves, mass.h:204-211): its[sw * 8 + a - 1]
 * for sweep sw (0, and 1 for split TG3) and population a, 0 where no CG ran
 * (lumped, Richardson, a population at rest in the sweep); the mixed form's
 * are its FP32 solves'. */
int lbfem_scheme_last_cg_iterations(const lbfem_scheme *scheme, int32_t *its);
/* StageTimers (include/lbfem/timers.h:17-23): seconds in collision,
 * advection, mass, and the total of the timed steps; timing adds events. */
int lbfem_scheme_enable_stage_timers(lbfem_scheme *scheme, int enable);
int lbfem_scheme_stage_times(const lbfem_scheme *scheme, double *seconds4);

/* No counterpart in the reference (its deal.II kernels have no launch shape):
 * the shape of the scheme's tile kernel, a TileVariant (0: 32 x 8 tiles of
 * 256 threads, 1: 32 x 16 tiles of 512, 2 and 3: 32 x 8 tiles with a node's
 * populations split over 2 or 4 threads, 4: 16 x 8 tiles split over 4, 5 and
 * 6: 40 x 8 and 48 x 8 tiles of 320 and 384 threads, 7: 40 x 8 tiles split over
 * 2 threads per node (640 threads); 8-11: the row-march kernel (Q1 meshes, and
 * Q2 meshes of the mixed form; wavefronts own strips of 64 - 2p columns and
 * march up segments of 3, 4, 8 or 16 node rows with the stencil rows in
 * registers; elsewhere variant 0); -1 restores the size-based default). Every
 * variant gives bitwise identical populations. */
int lbfem_scheme_set_tile_variant(lbfem_scheme *scheme, int variant);
int lbfem_scheme_get_tile_variant(const lbfem_scheme *scheme, int *variant);
/* Times every variant of the scheme's tile kernel on the scheme's own buffers
 * (its state is not changed) and keeps the fastest; the choice is cached per
 * (kernel mode, degree, mesh size, periodicity) for the process and reused
 * unless force != 0. seconds (optional, 12 doubles, one per variant): seconds
 * per launch (-1 where not timed). Synchronizes the scheme's stream. */
int lbfem_scheme_autotune(lbfem_scheme *scheme, int force, int *variant, double *seconds);

/* The diagnostics sums of cgdbe (apps/cgdbe/cgdbe.cc:353-368) over the
 * non-slave nodes with lumped-mass weights, from device populations:
 * sums[0..5] = |u-u_ex|^2, |u_ex|^2, |u|^2, rho, area, |u-u_prev|;
 * u_exact and u_prev: device arrays of 2 doubles per node (u_prev updated). */
int lbfem_diagnostics(const lbfem_disc *disc, const double *f, const double *u_exact, double *u_prev, double *sums,
                      void *stream);

/* The stream function of a nodal velocity field on a walled box mesh
 * (replaces lbfem::compute_stream_function, include/lbfem/io.h:31-93):
 * (grad phi, grad psi) = (dphi/dx, v) - (dphi/dy, u), psi = 0 on the walls,
 * SolverCG with the Jacobi preconditioner to 1e-12 |b| from zero, at most
 * 10000 iterations (LBFEM_ENOCONV beyond). u, v, psi: device arrays of
 * n_nodes (lexicographic) on the stream; the disc must be walled (n_periodic
 * 0) and undistorted (LBFEM_EINVAL otherwise). Synchronises the stream. */
int lbfem_disc_stream_function(const lbfem_disc *disc, const double *u, const double *v, double *psi,
                               int *iterations, void *stream);
/* The same from host arrays and a mesh description (the computation runs on
 * the device, on a discretization of its own). */
int lbfem_stream_function(const lbfem_mesh_desc *desc, const double *u, const double *v, double *psi,
                          int *iterations);

/* The six diagnostics sums of cgdbe (apps/cgdbe/cgdbe.cc:353-368) from the
 * scheme's own populations (FP64 or mixed), on the device: the sums of
 * lbfem_diagnostics with u_prev the velocity at the scheme's previous call
 * (zero before the first) kept by the scheme. u_exact (optional, host, 2
 * doubles per node) replaces the scheme's exact velocity field; the field used
 * is exact_scale times the last one given (zero if none was): a separable field
 * such as the Taylor-Green vortex's (test_cases.h:92-99) is given once and
 * scaled by its decay factor. Synchronizes the scheme's stream. */
int lbfem_scheme_diagnostics(lbfem_scheme *scheme, const double *u_exact, double exact_scale, double *sums);

/* The velocity-change monitor of cgdbe's steady-state test (the diagnostics
 * of apps/cgdbe/cgdbe.cc:385-393) on the device, for FP64 and mixed schemes:
 * over the non-slave nodes, the largest |u - u_prev| and the lumped-mass mean
 * of |u - u_prev|, u_prev the velocity at the scheme's previous call (kept by
 * the scheme, 16 bytes per node in FP64, allocated on the first call, which only
 * records it and returns +inf for both), and the number of nodes whose moments
 * are not finite (a diverged run). Synchronizes the scheme's stream. */
int lbfem_scheme_velocity_change(lbfem_scheme *scheme, double *max_change, double *mean_change, int64_t *nonfinite);

/* ---------------------------------------------------------- batches
 * No counterpart in the reference (one mesh per process there): many schemes
 * (any degree, scheme, sizes and flows) stepped together.
 *   - Lumped TG2 members (FP64 or mixed) are grouped by (degree, scheme,
 *     precision); a step is one launch per group of the tile kernel over the
 *     group's concatenated tile list, replayed from captured HIP graphs
 *     (graph_steps per graph).
 *   - Consistent-mass members (FP64, TG2, TG3 or split TG3) are grouped by
 *     (degree, scheme, streaming); a step is the single-mesh step's stages
 *     (advection, extrapolated start, the eight CG solves of each member,
 *     nodal update), each stage one launch over all members of the group, with
 *     one host check of all members' stopping states per CG iteration checked,
 *     on a second stream beside the lumped groups.
 * Each member's populations and CG iteration counts are bitwise those it gets
 * stepped alone. While batched, members are stepped only through the batch and
 * read after lbfem_batch_synchronize; destroy the batch before its members.
 * lbfem_batch_step returns LBFEM_ENOCONV when a member's solve reached its
 * iteration cap (the step of every member is complete; lbfem_batch_member_status
 * names the member). */
typedef struct lbfem_batch lbfem_batch;
/* n may be 0 (members join later with lbfem_batch_add) */
int lbfem_batch_create(lbfem_scheme *const *schemes, int n, lbfem_batch **batch);
/* A scheme joins a running batch, from its current populations, at any step; a
 * member leaves it (the batch's stream is synchronized first, so the member can
 * then be read, stepped alone or destroyed). The group's device table is
 * rewritten; the captured graphs are kept unless a (degree, scheme, precision)
 * group appears or disappears, or a group launched one block per tile changes
 * its tile count. */
int lbfem_batch_add(lbfem_batch *batch, lbfem_scheme *scheme);
int lbfem_batch_remove(lbfem_batch *batch, lbfem_scheme *scheme);
/* lbfem_scheme_velocity_change for the members of the batch in one launch:
 * arrays of the batch's size in member order (creation order, then joins);
 * mask (optional): nonzero for the members to measure, the others' entries and
 * u_prev untouched. */
int lbfem_batch_velocity_change(lbfem_batch *batch, const int32_t *mask, double *max_change, double *mean_change,
                                int64_t *nonfinite);
int lbfem_batch_destroy(lbfem_batch *batch);
/* steps per captured graph: even, 0 = launch every step eagerly (default 8) */
int lbfem_batch_set_graph_steps(lbfem_batch *batch, int steps);
int lbfem_batch_step(lbfem_batch *batch, int n_steps);
int lbfem_batch_synchronize(lbfem_batch *batch);
/* per member (member order): LBFEM_OK, or LBFEM_ENOCONV once its CG reached its
 * iteration cap in some step (SolverControl::NoConvergence, mass.h:38-43) */
int lbfem_batch_member_status(const lbfem_batch *batch, int32_t *status);
int lbfem_batch_get_stream(const lbfem_batch *batch, void **stream);
/* groups, nodes (slaves included) and tiles of the batch */
int lbfem_batch_info(const lbfem_batch *batch, int *n_groups, int64_t *n_nodes, int64_t *n_tiles);
/* every group's tile variant (-1: each keeps its own) and grid (0: one block
 * per tile; else that many persistent blocks, rounded up to a multiple of 8);
 * groups created later take the same shape */
int lbfem_batch_set_shape(lbfem_batch *batch, int variant, int grid);
/* times variant x grid for every group (no state change) and keeps the fastest;
 * seconds: the sum over groups of the best time per step */
int lbfem_batch_autotune(lbfem_batch *batch, double *seconds);

/* ---------------------------------------------------------- measurement
 * No counterpart in the reference: a STREAM-like device copy of `bytes`
 * (a multiple of 16, 16-byte aligned pointers) on the stream, the measured
 * HBM bandwidth a step is compared with and the known-size kernel that
 * calibrates the profiler's byte counters. */
int lbfem_hbm_copy(void *dst, const void *src, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif
