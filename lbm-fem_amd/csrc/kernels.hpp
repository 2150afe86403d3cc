// Host-side launch interface of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace lbfem_amd
{
  // modes of the tile kernel
  enum TileMode : int
  {
    kStepBardow   = 0, // collide, advect, lumped solve, g += x  (one fused step)
    kStepLeeLin   = 1, // equilibria, advect, lumped solve, predictor-corrector
    kAdvectBardow = 2, // (optional collision, writing g*) + right-hand side
    kAdvectLeeLin = 3, // right-hand side of LeeLinAdvection (feq given or computed)
  };

  struct TileParams
  {
    double e[Q][2];           // directions of the sweep
    double dt, a, b, c;       // advection coefficients (advection.h:223, 268-270)
    double omega;             // BGK rate (schemes.h:122)
    double theta;             // Lee-Lin dt / lambda (schemes.h:97)
    int    collide;           // Bardow: collide on load
    int    remove_mass_flux;  // wall surface term option (advection.h:182-199)
    int    tiles_x, tiles_y;  // set by the launcher (the row-march kernel: strips, row segments)
    int    rows;              // the row-march kernel: node rows per wavefront segment
  };

  // shapes of the tile kernel (all give bitwise identical results)
  enum TileVariant : int
  {
    kTileWide     = 0, // 32 x 8 nodes, 256 threads, all populations of a node in one thread
    kTileTall     = 1, // 32 x 16 nodes, 512 threads (less halo per node)
    kTileSplit2   = 2, // 32 x 8 nodes, 512 threads, the populations of a node split over 2 threads
    kTileSplit4   = 3, // 32 x 8 nodes, 1024 threads, split over 4
    kTileSmall4   = 4, // 16 x 8 nodes, 512 threads, split over 4
    kTileWide40   = 5, // 40 x 8 nodes, 320 threads: 231 tiles (one per CU) on a 257^2 node grid
    kTileWide48   = 6, // 48 x 8 nodes, 384 threads
    kTileSplit2W40 = 7, // 40 x 8 nodes, 640 threads, split over 2 (about 1.5 node-waves per SIMD at 257^2)
    kTileLdsVariants = 8, // the variants above stage the halo'd tile in LDS (the batch kernel's shapes)
    // the row-march kernel (Q1; elsewhere these run as kTileWide): wavefront strips of 62
    // columns marching up segments of 1, 4, 8 or 16 node rows
    kMarch3       = 8,
    kMarch4       = 9,
    kMarch8       = 10,
    kMarch16      = 11,
    kTileVariants = 12
  };

  // output columns per wavefront strip of the row-march kernel on Q_p (p lanes of halo each side)
  constexpr int
  march_out(const int p)
  {
    return 64 - 2 * p;
  }

  // The row-march kernel runs on Q1 and, where its five-row register window fits beside two
  // wavefronts per SIMD, on Q2 (the mixed form's FP32 populations); other forms run the LDS tile
  // kernel under every variant (the same results bit for bit).
  constexpr bool
  march_runs(const int p, const bool fp32)
  {
    return p == 1 || (p == 2 && fp32);
  }

  // the row-march kernel addresses a population block (Q vectors of the mesh's stride) with
  // 32-bit byte offsets: it runs on blocks under 4 GiB (the LDS tile kernel takes the rest)
  inline bool
  march_fits(const Geo &G, const size_t elem_bytes)
  {
    return uint64_t(Q) * uint64_t(G.stride) * uint64_t(elem_bytes) < (uint64_t(1) << 32);
  }

  // node rows per wavefront segment of a row-march variant
  inline int
  march_rows(const int variant)
  {
    return variant == kMarch3 ? 3 : variant == kMarch4 ? 4 : variant == kMarch8 ? 8 : 16;
  }

  template <typename T, int p>
  hipError_t tile_launch_degree(int mode, int variant, const Geo &G, const TileParams &S, const T *fin, const T *feq,
                                T *out, T *rhs, hipStream_t stream);
  extern template hipError_t tile_launch_degree<double, 1>(int, int, const Geo &, const TileParams &, const double *,
                                                           const double *, double *, double *, hipStream_t);
  extern template hipError_t tile_launch_degree<double, 2>(int, int, const Geo &, const TileParams &, const double *,
                                                           const double *, double *, double *, hipStream_t);
  extern template hipError_t tile_launch_degree<double, 3>(int, int, const Geo &, const TileParams &, const double *,
                                                           const double *, double *, double *, hipStream_t);
  extern template hipError_t tile_launch_degree<float, 1>(int, int, const Geo &, const TileParams &, const float *,
                                                          const float *, float *, float *, hipStream_t);
  extern template hipError_t tile_launch_degree<float, 2>(int, int, const Geo &, const TileParams &, const float *,
                                                          const float *, float *, float *, hipStream_t);
  extern template hipError_t tile_launch_degree<float, 3>(int, int, const Geo &, const TileParams &, const float *,
                                                          const float *, float *, float *, hipStream_t);

  // One launch of the tile kernel over the mesh: double (the FP64 form) or float
  // (the mixed form's FP32 deviations from the rest state), every mode.
  template <typename T>
  inline hipError_t
  tile_launch(const int mode, const int variant, const Geo &G, const TileParams &S, const T *fin, const T *feq,
              T *out, T *rhs, hipStream_t stream)
  {
    switch (G.p)
      {
        case 1:
          return tile_launch_degree<T, 1>(mode, variant, G, S, fin, feq, out, rhs, stream);
        case 2:
          return tile_launch_degree<T, 2>(mode, variant, G, S, fin, feq, out, rhs, stream);
        case 3:
          return tile_launch_degree<T, 3>(mode, variant, G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  // tile shape of a variant
  inline void
  tile_dims(const int variant, int &tx, int &ty)
  {
    tx = variant == kTileSmall4                                  ? 16 :
         (variant == kTileWide40 || variant == kTileSplit2W40) ? 40 :
         variant == kTileWide48                                  ? 48 :
                                                                   32;
    ty = variant == kTileTall ? 16 : 8;
  }

  // one instance of a batched fused step (device table entry)
  template <typename T>
  struct BatchEntry
  {
    Geo        G;
    TileParams S;          // tiles_x, tiles_y of the batch's variant
    T         *f[2];       // the instance's population buffers
    T         *rhs;        // the advection modes' right-hand side (8 blocks), else null
    int        par;        // buffer read at parity 0
    int        tile_begin; // first tile in the batch's tile list
  };

  // the live member and tile counts of a batch group's table (read by the kernel at run time)
  struct BatchCounts
  {
    int n_inst, n_tiles;
  };

  template <typename T, int p>
  hipError_t tile_batch_launch_degree(int mode, int variant, const BatchEntry<T> *table, const BatchCounts *counts,
                                      int parity, int grid, hipStream_t stream);
#define LBFEM_BATCH_EXTERN(T, P)                                                                                   \
  extern template hipError_t tile_batch_launch_degree<T, P>(int, int, const BatchEntry<T> *, const BatchCounts *,  \
                                                            int, int, hipStream_t);
  LBFEM_BATCH_EXTERN(double, 1)
  LBFEM_BATCH_EXTERN(double, 2)
  LBFEM_BATCH_EXTERN(double, 3)
  LBFEM_BATCH_EXTERN(float, 1)
  LBFEM_BATCH_EXTERN(float, 2)
  LBFEM_BATCH_EXTERN(float, 3)
#undef LBFEM_BATCH_EXTERN

  // grid: blocks of the launch, a multiple of 8
  template <typename T>
  inline hipError_t
  tile_batch_launch(const int p, const int mode, const int variant, const BatchEntry<T> *table,
                    const BatchCounts *counts, const int parity, const int grid, hipStream_t stream)
  {
    switch (p)
      {
        case 1:
          return tile_batch_launch_degree<T, 1>(mode, variant, table, counts, parity, grid, stream);
        case 2:
          return tile_batch_launch_degree<T, 2>(mode, variant, table, counts, parity, grid, stream);
        case 3:
          return tile_batch_launch_degree<T, 3>(mode, variant, table, counts, parity, grid, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  // The variant used until a measurement picks one, by the node count of the launch: below
  // about four wide-tile wavefronts per SIMD of the device (256 CUs x 4 SIMDs) the 40 x 8 tile
  // split over two threads per node (the fastest measured shape at 257^2 nodes, Q1 and Q2,
  // FP64 and FP32); above, the row-march kernel (Q1 Lee-Lin FP64 segment heights measured: 3
  // rows fastest at 513^2 and 1025^2 nodes (22.0 and 56.1 us against 25.1 at one row and 59.1 at
  // four), 8 at 2049^2, 16 at 4097^2; 30-42 % ahead of the 32 x 8 tile from 1025^2, FP32 28-49 %
  // ahead of 32 x 16). Where the row-march kernel does not run (Q2 in FP64, Q3) the march variants
  // run as the 32 x 8 tile, 20-30 % ahead of 32 x 16 in FP64.
  inline int
  size_default_variant(const long long nodes, const bool, const int p = 1)
  {
    if (nodes < 64LL * 4 * 1024)
      return kTileSplit2W40;
    // Q2 (FP32) below 768^2 nodes: the tile kernel (513^2: 26 us against 30-34 for the row march)
    if (p == 2 && nodes < 768LL * 768)
      return kTileWide;
    return nodes < 1536LL * 1536 ? kMarch3 : nodes < 3072LL * 3072 ? kMarch8 : kMarch16;
  }
  // The default shape of a batch group: where the row-march kernel runs (march_runs), over 4-row
  // segments (on the 96-instance seeded batch 19 % ahead of the best tile shape in FP32 on Q1,
  // 9 % in FP64); elsewhere, and for the consistent-mass groups' advection (tile lists only,
  // p = 0 here), the tile shapes above.
  inline int
  batch_default_variant(const long long nodes, const bool fp32, const int p = 0)
  {
    if (p > 0 && march_runs(p, fp32))
      return kMarch4;
    const int v = size_default_variant(nodes, fp32);
    return v < kTileLdsVariants ? v : fp32 ? kTileTall : kTileWide;
  }
  inline int
  tile_default_variant(const Geo &G, const bool fp32)
  {
    return size_default_variant((long long)(G.P + 1) * (G.P + 1), fp32, G.p);
  }
  // nodal operations over every node (collision.h:78-168)
  hipError_t collide_launch(const Geo &G, double *g, double omega, hipStream_t stream);
  hipError_t equilibrium_launch(const Geo &G, const double *f, double *feq, hipStream_t stream);
  hipError_t predictor_corrector_launch(const Geo &G, double *f, const double *feq, const double *incr, double theta,
                                        hipStream_t stream);
  // Lee-Lin nodal update after a consistent-mass solve: f <- PC(f, feq(f), x)
  hipError_t pc_from_f_launch(const Geo &G, double *f, const double *incr, double theta, hipStream_t stream);
  // f_{a+1} += x_a for the populations flagged in mask (streaming.h:94-98)
  hipError_t add_increments_launch(const Geo &G, double *f, const double *x, unsigned mask, hipStream_t stream);

  // ---- consistent-mass solves: block CG over up to 8 vectors sharing the mesh
  // (one SolverCG per vector, include/lbfem/mass.h:204-211, run side by side).
  struct CgVector
  {
    double e[2];  // direction of the K_e term
    double coef;  // c of A = M + c K_e
  };

  struct CgState // device, one per vector
  {
    double rho, rho_prev, alpha, beta, res, tol;
    int    its, status; // status: 0 active, 1 converged, 2 failed (NoConvergence)
  };

  // partial sums per block of the CG kernels (3 per vector), contiguous per block
  constexpr int kCgPartials = 3 * NM;

  struct CgWork
  {
    int       nv;         // vectors
    int       block[NM];  // block of the increment/rhs arrays vector v lives in
    int       max_its;
    double    tol_rel;
    CgVector  vec[NM];
    // device [nv]: the stopping states; with fold_rows > 0 those of even iterations, and
    // state2 [nv] those of odd ones (the direction stage writes iteration it - 1's while its
    // blocks read iteration it - 2's)
    CgState  *state;
    CgState  *state2;
    double   *partial;    // device [nv * 4 * max_blocks]
    int       max_blocks;
    // > 0: the direction stage's partial rows, which the update kernel reduces itself
    // (every block, in the order the finalisation uses) instead of a separate launch
    int       alpha_rows;
    // > 0 (with alpha_rows > 0 and state2): the update stage's partial rows (after the
    // direction stage's), which the next direction stage reduces itself (every block, in the
    // order the finalisation uses) for beta and the stopping test, instead of a separate
    // launch; the finalisation then runs only where the host reads the states
    int       fold_rows;
    // 1: the stream function's Laplacian Kx (x) My + Mx (x) Ky on a walled box instead of
    // M + c K_e: wall entries read as zero, wall rows the identity; the Jacobi diagonal
    // comes from Geo::dinv
    int       laplace;
  };

  // the most update-stage rows the direction stage reduces itself (CgWork::fold_rows)
  constexpr int kMaxCgFoldRows = 256;

  // partial rows the CG direction stage writes (its blocks per vector)
  int cg_direction_rows(const Geo &G);
  // the tensor operator's tiles per direction (64 - 2p owned columns, 16 owned rows each)
  void cg_operator_tiles(const Geo &G, int &tiles_x, int &tiles_y);
  // the update stage's grid of a CG with the partial rows alpha_rows ahead of its own
  int cg_update_blocks(const Geo &G, int max_blocks, int alpha_rows);

  // ---- the consistent-mass solves of many meshes in one launch per stage (a batch of
  // schemes of one degree, scheme and streaming, FP64): each member's blocks, partial rows,
  // states and stopping tests are its single-mesh ones, so its solves are bitwise the same
  struct CgMember
  {
    Geo           G;
    CgWork        W;          // the member's vectors, states (in the batch's state array), partial rows
    const double *b;          // right-hand sides (8 blocks)
    double       *x, *r, *p[2], *q;
    int           tiles_x, tiles_y, nt, op_begin; // operator tiles (32 x 16) and the first block of the member
    int           upd_blocks, upd_begin;          // the update grid and its first block
  };
  // start (r = A x - b, finalisation 1); direction (p, q = A p, finalisation 2 where the
  // update does not reduce it); update (x, r, finalisation 3). op: 0 mass, 1 mass + c K_e.
  hipError_t cg_batch_start_launch(int p, int op, const CgMember *T, int n, int op_blocks, hipStream_t stream);
  hipError_t cg_batch_direction_launch(int p, int op, const CgMember *T, int n, int op_blocks, int it,
                                       hipStream_t stream);
  hipError_t cg_batch_update_launch(const CgMember *T, int n, int upd_blocks, int it, hipStream_t stream);

  struct NodalMember
  {
    Geo      G;
    double  *f[2];  // population buffers, f[(par + parity) & 1] the current one
    double  *x, *x_prev;
    int      par;
    unsigned mask;  // the populations of the sweep (kinds 0, 1)
    double   theta; // Lee-Lin dt / lambda (kind 2)
    int      block_begin, nblocks;
  };
  // kind 0: extrapolated start; 1: f_{a+1} += x_a; 2: Lee-Lin predictor-corrector
  hipError_t nodal_batch_launch(int kind, const NodalMember *T, int n, int blocks, int parity, hipStream_t stream);
  // the blocks a member of the nodal batch kernels gets
  int nodal_batch_blocks(const Geo &G);

  // r = A x - b, tolerances, rho = r.D^-1 r (iteration 0 of SolverCG)
  hipError_t cg_start_launch(const Geo &G, const CgWork &W, const double *x, const double *b, double *r,
                             hipStream_t stream);
  // p_new = beta p_old - D^-1 r (p_new = -D^-1 r at it == 1); q = A p_new; alpha = rho / p.q
  hipError_t cg_direction_launch(const Geo &G, const CgWork &W, int it, const double *p_old, const double *r,
                                 double *p_new, double *q, hipStream_t stream);
  // x += alpha p; r += alpha q; res = |r|; stopping test; rho = r.D^-1 r, beta (with
  // W.fold_rows > 0 the last three are formed by the next direction stage, or by
  // cg_update_finalize_launch where the host reads the states)
  hipError_t cg_update_launch(const Geo &G, const CgWork &W, int it, double *x, double *r, const double *p,
                              const double *q, hipStream_t stream);
  hipError_t cg_update_finalize_launch(const Geo &G, const CgWork &W, int it, hipStream_t stream);
  // the same three stages on FP32 vectors (the mixed form's inner solves; scalars in FP64)
  hipError_t cg_start_launch(const Geo &G, const CgWork &W, const float *x, const float *b, float *r,
                             hipStream_t stream);
  hipError_t cg_direction_launch(const Geo &G, const CgWork &W, int it, const float *p_old, const float *r,
                                 float *p_new, float *q, hipStream_t stream);
  hipError_t cg_update_launch(const Geo &G, const CgWork &W, int it, float *x, float *r, const float *p,
                              const float *q, hipStream_t stream);
  // the mixed form's consistent mass: the extrapolated start and the nodal updates on FP32 deviations from
  // FP32 increments
  hipError_t extrapolate_launch(const Geo &G, float *x, float *x_prev, unsigned mask, hipStream_t stream);
  hipError_t add_increments_launch(const Geo &G, float *g, const float *x, unsigned mask, hipStream_t stream);
  hipError_t pc_from_f_launch(const Geo &G, float *f, const float *incr, double theta, hipStream_t stream);
  // General quadrilaterals (quad_kernels.hip): the advection right-hand side
  // and the fused lumped step (TileMode semantics, no wall faces), and the mass
  // operator stages of the CG (KIND 0 product, 1 start, 2 direction).
  hipError_t quad_advect_launch(int mode, const Geo &G, const TileParams &S, const double *fin, const double *feq,
                                double *out, double *rhs, hipStream_t stream);
  template <int KIND>
  hipError_t quad_operator_launch(const Geo &G, const CgWork &W, int it, const double *x, const double *b,
                                  const double *r, double *y, double *y2, hipStream_t stream);
  extern template hipError_t quad_operator_launch<0>(const Geo &, const CgWork &, int, const double *, const double *,
                                                     const double *, double *, double *, hipStream_t);
  extern template hipError_t quad_operator_launch<1>(const Geo &, const CgWork &, int, const double *, const double *,
                                                     const double *, double *, double *, hipStream_t);
  extern template hipError_t quad_operator_launch<2>(const Geo &, const CgWork &, int, const double *, const double *,
                                                     const double *, double *, double *, hipStream_t);
  // tiles of the general-mesh kernels (blocks of one operator launch per vector)
  int quad_tile_count(const Geo &G);

  // y = (M + c K_e) x for nv vectors (constrained rows the identity)
  hipError_t vmult_launch(const Geo &G, const CgWork &W, const double *x, double *y, hipStream_t stream);
  // x_guess = 2 x - x_prev, x_prev = x (the extrapolated start, streaming.h:76-80), masked
  hipError_t extrapolate_launch(const Geo &G, double *x, double *x_prev, unsigned mask, hipStream_t stream);
  // lumped-type solves: x = M_L^-1 r, and k Richardson passes x += M_L^-1 (r - A x)
  hipError_t lumped_solve_launch(const Geo &G, double *x, const double *r, int nv, hipStream_t stream);
  hipError_t richardson_update_launch(const Geo &G, double *x, const double *r, const double *Ax, int nv,
                                      hipStream_t stream);

  // the stream function's right-hand side b = (Dx^T (x) My) v - (Mx (x) Dy^T) u and inverse
  // Jacobi diagonal 1 / (Kx My + Mx Ky)_ii, with b = 0 and 1 on the walls (io.h:55-85)
  hipError_t psi_setup_launch(const Geo &G, const double *u, const double *v, double *b, double *dinv,
                              hipStream_t stream);

  // ---- the velocity-change monitor of the steady-state tests (cgdbe.cc:385-393) on the device: per member, over
  // its non-slave nodes, max |u - u_prev|, sum of w |u - u_prev|, sum of w (w the lumped mass) and the count of
  // nodes whose moments are not finite; u_prev updated. The members of one launch are a table; member k owns
  // blocks [block_begin, block_begin + nblocks) of the launch, so every member's sums keep a fixed order.
  struct MonitorEntry
  {
    Geo         G;
    const void *f;           // the current populations: double, or float deviations from the rest state
    double     *u_prev;      // 2 per node
    int         fp32;        // 1: float deviations
    int         init;        // 1: record u_prev only (the results are then +inf)
    int         block_begin; // first block of the member
    int         nblocks;
  };
  constexpr int kMonitorPartials = 4;
  // blocks a member of n nodes gets (the same at every call, so its sums are reproducible)
  inline int
  monitor_blocks(const long long n)
  {
    const long long b = (n + 2047) / 2048;
    return int(b < 1 ? 1 : b > 256 ? 256 : b);
  }
  // results: kMonitorPartials doubles per member (max, sum w |du|, sum w, nonfinite nodes)
  hipError_t monitor_launch(const MonitorEntry *table, int n_members, int total_blocks, double *partial,
                            double *results, hipStream_t stream);

  // diagnostics sums over non-slave nodes (cgdbe.cc:353-368)
  hipError_t stream_copy_launch(void *dst, const void *src, size_t bytes, hipStream_t s);
  hipError_t diagnostics_launch(const Geo &G, const double *f, const double *u_exact, double *u_prev, double *sums,
                                double *partial, int max_blocks, hipStream_t stream);
  // the same from FP64 populations or the mixed form's FP32 deviations, the exact velocity scaled by ex_scale
  template <typename T>
  hipError_t diagnostics_scaled_launch(const Geo &G, const T *f, const double *u_exact, double ex_scale,
                                       double *u_prev, double *sums, double *partial, int max_blocks,
                                       hipStream_t stream);
} // namespace lbfem_amd
