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

  struct Geo
  {
    AxisDev  ax, ay;
    WallsDev walls;
    int      p, P, np; // degree, p N, p N + 1
    int64_t  stride;
  };

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
    ux            = mx / rho;
    uy            = my / rho;
  }

  // e_a . u for the D2Q9 directions
  template <int a>
  __device__ __forceinline__ double
  e_dot(const double ux, const double uy)
  {
    constexpr int sx = ex_of(a), sy = ey_of(a);
    if constexpr (sx == 0 && sy == 0)
      return 0.;
    else if constexpr (sy == 0)
      return sx > 0 ? ux : -ux;
    else if constexpr (sx == 0)
      return sy > 0 ? uy : -uy;
    else
      return (sx > 0 ? ux : -ux) + (sy > 0 ? uy : -uy);
  }

  // Second-order equilibrium of population a (include/lbfem/d2q9.h:44-55).
  template <int a>
  __device__ __forceinline__ double
  feq(const double rho, const double ux, const double uy, const double uu)
  {
    const double eu = e_dot<a>(ux, uy);
    return w_of(a) * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * uu);
  }

  template <int a = 0>
  __device__ __forceinline__ void
  equilibrium_all(const double rho, const double ux, const double uy, double *out)
  {
    const double uu = ux * ux + uy * uy;
    if constexpr (a < Q)
      {
        out[a] = feq<a>(rho, ux, uy, uu);
        if constexpr (a + 1 < Q)
          equilibrium_all<a + 1>(rho, ux, uy, out);
      }
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
