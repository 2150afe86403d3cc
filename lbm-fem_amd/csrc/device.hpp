// Device-side data layout and the D2Q9 nodal physics shared by the kernels.
//
// HBM layout (one discretization): populations are structure-of-arrays, Q = 9
// blocks of `stride` values (block a at a * stride), node (i, j) at j * np + i,
// x fastest, so a wavefront reads 64 consecutive nodes of one population.
// Geometry is the per-axis bands of the assembled 1-D operators (a few KB),
// read through L1/L2; nothing per node or per cell is stored.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbfem_amd
{
  constexpr int Q  = 9;
  constexpr int NM = 8;

  // include/lbfem/d2q9.h:18-22, as compile-time functions of the index
  __host__ __device__ constexpr int
  ex_of(const int a)
  {
    return (a == 1 || a == 2 || a == 8) ? 1 : (a == 4 || a == 5 || a == 6) ? -1 : 0;
  }
  __host__ __device__ constexpr int
  ey_of(const int a)
  {
    return (a == 2 || a == 3 || a == 4) ? 1 : (a == 6 || a == 7 || a == 8) ? -1 : 0;
  }
  __host__ __device__ constexpr double
  w_of(const int a)
  {
    return a == 0 ? 4. / 9. : (a % 2 == 1 ? 1. / 9. : 1. / 36.);
  }

  // One axis: bands of the assembled 1-D operators, (P+1) x (2p+1).
  struct AxisDev
  {
    const double *M, *D, *DT, *K;
    const double *ml_inv;  // 1 / row sums of M, P + 1
    const double *ml;      // row sums of M (lumped mass), P + 1
    const double *diag;    // diagonal of M, P + 1
    const double *diag_inv; // 1 / diagonal of M, P + 1
    double        b0[4], b1[4];
    int           periodic;
  };

  // Wall tables: per boundary line (0 x = 0, 1 x = L, 2 y = 0, 3 y = L) the
  // index of the node's wall velocity, or -1; null where no wall node lies.
  struct WallsDev
  {
    const int32_t *side[4];
    const double  *velocity; // 2 per wall
  };

  // General quadrilaterals (the vortex's distorted periodic mesh): the cells are
  // the bilinear images of their four vertices, so the operators are evaluated
  // cell by cell with the Jacobian per quadrature point, and the lumped mass and
  // the Jacobi diagonal are per node. All pointers null on a tensor mesh.
  struct GeneralDev
  {
    const double *vert;     // (N+1)^2 vertices, (x, y) pairs, x fastest
    const double *ml;       // lumped mass per node (constrained entries 1)
    const double *ml_inv;   // its inverse
    const double *diag_inv; // inverse diagonal of M per node (constrained entries 1)
    const double *basis;    // S[4][4], D[4][4] (value, derivative of phi_k at Gauss point q: [q][k]), qx[4], qw[4]
    int           N;        // cells per direction
  };

  struct Geo
  {
    AxisDev    ax, ay;
    WallsDev   walls;
    int        p, P, np; // degree, p N, p N + 1
    int64_t    stride;
    GeneralDev gen;
    // the Jacobi preconditioner of the CG solves per node: the inverse of diag(M) (the axes'
    // product, or the general cells' per-node diagonal; 1 on constrained rows), or the
    // stream function's Laplacian diagonal
    const double *dinv;
  };

  // a double moved across lanes by a DPP control (both halves; lanes of rows outside
  // ROW_MASK read zero)
  template <int CTRL, int ROW_MASK = 0xf>
  __device__ __forceinline__ double
  dpp_move(const double v)
  {
    const uint64_t b  = __builtin_bit_cast(uint64_t, v);
    const int      lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(b)), CTRL, ROW_MASK, 0xf, false);
    const int      hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(b >> 32)), CTRL, ROW_MASK, 0xf, false);
    return __builtin_bit_cast(double, uint64_t(uint32_t(lo)) | (uint64_t(uint32_t(hi)) << 32));
  }

  // The sum of a wavefront's 64 values, in lane 63, in a fixed order without LDS: pairs and
  // quads (quad permutations), the four quads of each 16-lane row (row rotations by 4 and
  // 8), then the rows (lane 15 into row 1 and lane 47 into row 3, then lane 31 into rows 2
  // and 3)
  __device__ __forceinline__ double
  wave_sum(double s)
  {
    s += dpp_move<0xb1>(s);       // quad_perm [1, 0, 3, 2]
    s += dpp_move<0x4e>(s);       // quad_perm [2, 3, 0, 1]
    s += dpp_move<0x124>(s);      // row_ror 4
    s += dpp_move<0x128>(s);      // row_ror 8
    s += dpp_move<0x142, 0xa>(s); // row_bcast 15
    s += dpp_move<0x143, 0xc>(s); // row_bcast 31
    return s;
  }
  constexpr int kWaveSumLane = 63;

  // sum over a block of NT threads of up to K values per thread; the result is
  // valid in thread 0 (a fixed order: wave_sum, then the wavefronts)
  template <int K, int NT>
  __device__ __forceinline__ void
  block_sum_n(double (&v)[K], double (*lds)[NT / 64])
  {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k)
      {
        const double s = wave_sum(v[k]);
        if (lane == kWaveSumLane)
          lds[k][wave] = s;
      }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < K; ++k)
        {
          double s = 0.;
#pragma unroll
          for (int w = 0; w < NT / 64; ++w)
            s += lds[k][w];
          v[k] = s;
        }
    __syncthreads();
  }

  // block_sum_n with the result in every thread (each sums the wavefronts' partials in the
  // same order: thread 0's value, bit for bit, with one barrier less)
  template <int K, int NT>
  __device__ __forceinline__ void
  block_sum_all(double (&v)[K], double (*lds)[NT / 64])
  {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k)
      {
        const double s = wave_sum(v[k]);
        if (lane == kWaveSumLane)
          lds[k][wave] = s;
      }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k)
      {
        double s = 0.;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w)
          s += lds[k][w];
        v[k] = s;
      }
    __syncthreads();
  }

  // global node line of a tile position, -1 outside a non-periodic axis. A periodic
  // position lies in [-p, P + P): positions past P + p feed no owned node and map to
  // any line inside.
  __device__ __forceinline__ int
  wrap_line(const int i, const int P, const int periodic)
  {
    if (periodic)
      {
        const int w = i < 0 ? i + P : (i >= P ? i - P : i);
        return w < P ? w : P - 1;
      }
    return (i < 0 || i > P) ? -1 : i;
  }

  __device__ __forceinline__ int64_t
  node_index(const Geo &G, const int i, const int j)
  {
    return int64_t(j) * G.np + i;
  }

  // The wall velocity a node imposes on its moments, if any (collision.h:145-165).
  __device__ __forceinline__ bool
  wall_of(const Geo &G, const int i, const int j, double &wx, double &wy)
  {
    int k = -1;
    if (j == 0 && G.walls.side[2])
      k = G.walls.side[2][i];
    else if (j == G.P && G.walls.side[3])
      k = G.walls.side[3][i];
    if (k < 0 && i == 0 && G.walls.side[0])
      k = G.walls.side[0][j];
    else if (k < 0 && i == G.P && G.walls.side[1])
      k = G.walls.side[1][j];
    if (k < 0)
      return false;
    wx = G.walls.velocity[2 * k];
    wy = G.walls.velocity[2 * k + 1];
    return true;
  }

  // D2Q9 moments (include/lbfem/d2q9.h:30-41), in the reference's summation order.
  __device__ __forceinline__ void
  moments(const double *f, double &rho, double &ux, double &uy)
  {
    rho           = ((((((((f[0] + f[1]) + f[2]) + f[3]) + f[4]) + f[5]) + f[6]) + f[7]) + f[8]);
    const double mx = ((((f[1] + f[2]) - f[4]) - f[5]) - f[6]) + f[8];
    const double my = ((((f[2] + f[3]) + f[4]) - f[6]) - f[7]) - f[8];
    const double inv = 1. / rho;
    ux               = mx * inv;
    uy               = my * inv;
  }

  // The second-order equilibria (include/lbfem/d2q9.h:44-55), w_a rho (1 + 3 e_a.u + 4.5 (e_a.u)^2 -
  // 1.5 u^2), formed per pair of opposite directions a, a': with c = e_a.u = -e_a'.u and
  // A = w rho (1 - 1.5 u^2 + 4.5 c^2), the pair is A + 3 w rho c and A - 3 w rho c (the same
  // polynomial in about half the operations of one evaluation per population). The pairs of
  // this numbering: (1, 5) c = ux, (3, 7) c = uy, (2, 6) c = ux + uy, (4, 8) c = uy - ux.
  __device__ __forceinline__ void
  equilibrium_all(const double rho, const double ux, const double uy, double *out)
  {
    const double uu = ux * ux + uy * uy;
    const double c0 = 1. - 1.5 * uu;
    const double r1 = w_of(1) * rho, r2 = w_of(2) * rho;
    const double t1 = 3. * r1, t2 = 3. * r2;
    const auto   pair = [&](const double c, const double r, const double t, const int a, const int b) {
      const double A = r * (c0 + 4.5 * c * c);
      const double B = t * c;
      out[a]         = A + B;
      out[b]         = A - B;
    };
    out[0] = w_of(0) * rho * c0;
    pair(ux, r1, t1, 1, 5);
    pair(uy, r1, t1, 3, 7);
    pair(ux + uy, r2, t2, 2, 6);
    pair(uy - ux, r2, t2, 4, 8);
  }

  // BGK collision of Bardow's transformed populations at one node
  // (include/lbfem/collision.h:152-168): g <- geq(rho, u_imposed) + (1 - omega)(g - geq(rho, u)).
  __device__ __forceinline__ void
  collide_node(double *g, const double omega, const bool wall, const double wx, const double wy)
  {
    double rho, ux, uy;
    moments(g, rho, ux, uy);
    double ge[Q], gt[Q];
    equilibrium_all(rho, ux, uy, ge);
    if (wall)
      equilibrium_all(rho, wx, wy, gt);
#pragma unroll
    for (int a = 0; a < Q; ++a)
      g[a] = (wall ? gt[a] : ge[a]) + (1. - omega) * (g[a] - ge[a]);
  }

  // Nodal equilibrium with the wall velocity imposed (collision.h:105-113).
  __device__ __forceinline__ void
  equilibrium_node(const double *f, double *out, const bool wall, const double wx, const double wy)
  {
    double rho, ux, uy;
    moments(f, rho, ux, uy);
    if (wall)
      {
        ux = wx;
        uy = wy;
      }
    equilibrium_all(rho, ux, uy, out);
  }

  // Lee & Lin predictor-corrector at one node (collision.h:119-145); x holds
  // the increments of the moving populations (x[a-1] for population a).
  __device__ __forceinline__ void
  predictor_corrector_node(double *f, const double *fe, const double *x, const double theta, const bool wall,
                           const double wx, const double wy)
  {
    const double inv = 1. / (1. + theta);
    double       fh[Q], eh[Q];
    fh[0] = inv * (f[0] + theta * fe[0]);
#pragma unroll
    for (int a = 1; a < Q; ++a)
      fh[a] = inv * (f[a] + theta * fe[a] + x[a - 1]);
    equilibrium_node(fh, eh, wall, wx, wy);
#pragma unroll
    for (int a = 0; a < Q; ++a)
      f[a] = fh[a] + theta * (eh[a] - fe[a]);
  }
} // namespace lbfem_amd

// ---------------------------------------------------------------- mixed precision
// Populations stored in FP32 as deviations d_a = f_a - w_a rho0 (rho0 = 1) from
// the rest state, so that FP32 carries the small part of f. The moments and
// the equilibria are evaluated in FP64 and rounded once; the relaxation, the
// predictor-corrector and the streaming operator run in FP32 on deviations.
// Every operator of the step is affine in f with the rest state as a fixed
// point, so the deviation form is the same scheme.
namespace lbfem_amd
{
  // rho - rho0, rho, u from deviations, in FP64
  __device__ __forceinline__ void
  moments_dev(const float *d, double &drho, double &rho, double &ux, double &uy)
  {
    drho = ((((((((double(d[0]) + double(d[1])) + double(d[2])) + double(d[3])) + double(d[4])) + double(d[5])) +
              double(d[6])) +
             double(d[7])) +
            double(d[8]));
    const double mx = ((((double(d[1]) + double(d[2])) - double(d[4])) - double(d[5])) - double(d[6])) + double(d[8]);
    const double my = ((((double(d[2]) + double(d[3])) + double(d[4])) - double(d[6])) - double(d[7])) - double(d[8]);
    rho              = 1. + drho;
    const double inv = 1. / rho;
    ux               = mx * inv;
    uy               = my * inv;
  }

  // The deviation of the equilibrium, w (rho - rho0) + w rho (3 e.u + 4.5 (e.u)^2 - 1.5 u^2),
  // evaluated in FP64 and rounded to FP32 once: evaluated in FP32 its rounding
  // errors are correlated across the nine populations and from step to step on a
  // slowly varying flow, and the step drifts coherently (1e-5 relative velocity
  // error after 1000 Taylor-Green steps, against 1e-7 with this form). The nine
  // are paired as in equilibrium_all: with c = e_a.u, A = w drho + w rho (4.5 c^2 -
  // 1.5 u^2) and the pair A +- 3 w rho c.
  __device__ __forceinline__ void
  equilibrium_dev_fp64(const double drho, const double rho, const double ux, const double uy, double *out)
  {
    const double uu = ux * ux + uy * uy;
    const double c0 = -1.5 * uu;
    const double r1 = w_of(1) * rho, r2 = w_of(2) * rho;
    const double d1 = w_of(1) * drho, d2 = w_of(2) * drho;
    const double t1 = 3. * r1, t2 = 3. * r2;
    const auto   pair = [&](const double c, const double r, const double d, const double t, const int a, const int b) {
      const double A = d + r * (c0 + 4.5 * c * c);
      const double B = t * c;
      out[a]         = A + B;
      out[b]         = A - B;
    };
    out[0] = w_of(0) * (drho + rho * c0);
    pair(ux, r1, d1, t1, 1, 5);
    pair(uy, r1, d1, t1, 3, 7);
    pair(ux + uy, r2, d2, t2, 2, 6);
    pair(uy - ux, r2, d2, t2, 4, 8);
  }

  __device__ __forceinline__ void
  equilibrium_dev_all(const double drho, const double rho, const double ux, const double uy, float *out)
  {
    double e[Q];
    equilibrium_dev_fp64(drho, rho, ux, uy, e);
#pragma unroll
    for (int a = 0; a < Q; ++a)
      out[a] = float(e[a]);
  }

  // BGK collision of deviations (collision.h:152-168): equilibria as above, the
  // relaxation in FP32
  __device__ __forceinline__ void
  collide_node(float *g, const double omega, const bool wall, const double wx, const double wy)
  {
    double drho, rho, ux, uy;
    moments_dev(g, drho, rho, ux, uy);
    float ge[Q], gt[Q];
    equilibrium_dev_all(drho, rho, ux, uy, ge);
    if (wall)
      equilibrium_dev_all(drho, rho, wx, wy, gt);
    const float om1 = float(1. - omega);
#pragma unroll
    for (int a = 0; a < Q; ++a)
      g[a] = (wall ? gt[a] : ge[a]) + om1 * (g[a] - ge[a]);
  }

  __device__ __forceinline__ void
  equilibrium_node(const float *f, float *out, const bool wall, const double wx, const double wy)
  {
    double drho, rho, ux, uy;
    moments_dev(f, drho, rho, ux, uy);
    if (wall)
      {
        ux = wx;
        uy = wy;
      }
    equilibrium_dev_all(drho, rho, ux, uy, out);
  }

  // Lee & Lin predictor-corrector on deviations (collision.h:119-145): affine in
  // f, feq and x with the rest state fixed, so the same formulas apply
  // f_a = fh_a + theta (feq_a(fh) - fe_a), the corrector in FP64, rounded once
  __device__ __forceinline__ void
  pc_correct(float *f, const double *fh, const float *fe, const double theta, const double drho, const double rho,
             const double ux, const double uy)
  {
    double eh[Q];
    equilibrium_dev_fp64(drho, rho, ux, uy, eh);
#pragma unroll
    for (int a = 0; a < Q; ++a)
      f[a] = float(fh[a] + theta * (eh[a] - double(fe[a])));
  }

  // Lee & Lin predictor-corrector on deviations (collision.h:119-145), in FP64
  // from the FP32 inputs: in FP32 the rounded 1 / (1 + theta) times (1 + theta)
  // differs from 1 and scales every deviation by that factor each step, a
  // coherent drift of 4e-8 per step at theta ~ 2
  __device__ __forceinline__ void
  predictor_corrector_node(float *f, const float *fe, const float *x, const double theta, const bool wall,
                           const double wx, const double wy)
  {
    const double inv = 1. / (1. + theta);
    double       fh[Q];
    fh[0] = inv * (double(f[0]) + theta * double(fe[0]));
#pragma unroll
    for (int a = 1; a < Q; ++a)
      fh[a] = inv * (double(f[a]) + theta * double(fe[a]) + double(x[a - 1]));
    double drho = ((((((((fh[0] + fh[1]) + fh[2]) + fh[3]) + fh[4]) + fh[5]) + fh[6]) + fh[7]) + fh[8]);
    double ux   = (((((fh[1] + fh[2]) - fh[4]) - fh[5]) - fh[6]) + fh[8]);
    double uy   = (((((fh[2] + fh[3]) + fh[4]) - fh[6]) - fh[7]) - fh[8]);
    const double rho = 1. + drho, ir = 1. / rho;
    ux *= ir;
    uy *= ir;
    if (wall)
      {
        ux = wx;
        uy = wy;
      }
    pc_correct(f, fh, fe, theta, drho, rho, ux, uy);
  }
} // namespace lbfem_amd
