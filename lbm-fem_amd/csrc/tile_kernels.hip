// Tile kernels of the streaming step: the weak-form advection right-hand side
// of all eight moving populations (include/lbfem/advection.h:94-293), fused
// with the nodal work before it (Bardow's BGK collision, or Lee & Lin's
// equilibria, collision.h:105-168) and, for the lumped mass, with the solve
// and the nodal work after it (f += M_L^-1 r, or the predictor-corrector).
//
// One workgroup owns a TX x TY block of nodes. It loads the block plus a halo
// of p nodes (the reach of a Q_p row) for all nine populations, does the
// nodal prologue in registers, then per population stages the field in LDS
// and applies the operator as Kronecker products of assembled 1-D bands: a y
// pass over the halo columns into LDS, then an x pass at the owned nodes.
// Every output node is computed by exactly one thread from the same sequence
// of operations, so the result is deterministic and translation invariant (no
// scatter, no atomics); the periodic seam is a wrap of the halo index.
#include "device.hpp"
#include "kernels.hpp"

namespace lbfem_amd
{
  namespace
  {
    constexpr int kThreads = 256;

    __device__ __forceinline__ double
    feq_rt(const int a, const double rho, const double ux, const double uy, const double uu)
    {
      const double eu = double(ex_of(a)) * ux + double(ey_of(a)) * uy;
      return w_of(a) * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * uu);
    }

    // Band coefficient o in [-p, p] of row i.
    template <int p>
    __device__ __forceinline__ double
    band(const double *B, const int i, const int o)
    {
      return B[i * (2 * p + 1) + o + p];
    }

    // global node line for a tile position, with the periodic wrap; -1 outside
    __device__ __forceinline__ int
    wrap(const int i, const int P, const int periodic)
    {
      if (periodic)
        return ((i % P) + P) % P;
      return (i < 0 || i > P) ? -1 : i;
    }
  } // namespace

  template <int p, int TX, int TY, int MODE>
  __global__ void __launch_bounds__(kThreads)
  k_tile(const Geo G, const TileParams S, const double *__restrict__ fin, const double *__restrict__ feq_in,
         double *__restrict__ out, double *__restrict__ rhs_out)
  {
    constexpr int HX = TX + 2 * p, HY = TY + 2 * p, NH = HX * HY;
    constexpr int R  = (NH + kThreads - 1) / kThreads; // halo'd nodes per thread
    constexpr int RO = (TX * TY) / kThreads;           // owned nodes per thread
    static_assert((TX * TY) % kThreads == 0 && kThreads % TX == 0, "tile shape");
    constexpr bool LL = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
    constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);

    __shared__ double fld[HY][HX];
    __shared__ double hld[LL ? HY : 1][LL ? HX : 1];
    __shared__ double yv[4][TY][HX];

    // XCD-aware tile order: consecutive tiles of a row band go to one XCD (one L2)
    const int nt   = S.tiles_x * S.tiles_y;
    const int nb8  = (nt + 7) & ~7;
    const int bid  = blockIdx.x;
    const int tile = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    if (tile >= nt)
      return;
    const int i0 = (tile % S.tiles_x) * TX, j0 = (tile / S.tiles_x) * TY;
    const int P = G.P, tid = threadIdx.x;

    // ---------------------------------------------------------- prologue
    // halo'd block: f (LL: also the moments the equilibria are built from) or g*
    double v[R][Q];
    double mom[LL ? R : 1][3];
#pragma unroll
    for (int r = 0; r < R; ++r)
      {
        const int k  = tid + r * kThreads;
        const int kx = k % HX, ky = k / HX;
        const int gi = wrap(i0 - p + kx, P, G.ax.periodic), gj = wrap(j0 - p + ky, P, G.ay.periodic);
        const bool ok = k < NH && gi >= 0 && gj >= 0;
        if (ok)
          {
            const int64_t idx = node_index(G, gi, gj);
#pragma unroll
            for (int a = 0; a < Q; ++a)
              v[r][a] = fin[a * G.stride + idx];
            double wx = 0., wy = 0.;
            const bool wall = wall_of(G, gi, gj, wx, wy);
            if constexpr (LL)
              {
                double rho, ux, uy;
                if (feq_in == nullptr)
                  {
                    moments(v[r], rho, ux, uy);
                    if (wall)
                      {
                        ux = wx;
                        uy = wy;
                      }
                    mom[r][0] = rho;
                    mom[r][1] = ux;
                    mom[r][2] = uy;
                  }
              }
            else if (S.collide)
              collide_node(v[r], S.omega, wall, wx, wy);
          }
        else
          {
#pragma unroll
            for (int a = 0; a < Q; ++a)
              v[r][a] = 0.;
            if constexpr (LL)
              mom[r][0] = mom[r][1] = mom[r][2] = 0.;
          }
      }

    // owned nodes: their own populations after the prologue (a slave reads its
    // own entries, not its master's), and the accumulated increments
    double own[RO][Q];
    double xinc[RO][NM];
    double mflux[RO];
#pragma unroll
    for (int r = 0; r < RO; ++r)
      {
        const int ii = tid % TX, jj = tid / TX + r * (kThreads / TX);
        const int gi = i0 + ii, gj = j0 + jj;
        mflux[r]     = 0.;
#pragma unroll
        for (int a = 0; a < NM; ++a)
          xinc[r][a] = 0.;
        if (gi <= P && gj <= P)
          {
            const int64_t idx = node_index(G, gi, gj);
#pragma unroll
            for (int a = 0; a < Q; ++a)
              own[r][a] = fin[a * G.stride + idx];
            if constexpr (!LL)
              if (S.collide)
                {
                  double wx = 0., wy = 0.;
                  const bool wall = wall_of(G, gi, gj, wx, wy);
                  collide_node(own[r], S.omega, wall, wx, wy);
                }
          }
        else
#pragma unroll
          for (int a = 0; a < Q; ++a)
            own[r][a] = 0.;
      }

    const double a_c = S.a, b_c = S.b, c_c = S.c, hdt2 = 0.5 * S.dt * S.dt;

    // ---------------------------------------------------------- populations 1..8
#pragma unroll
    for (int al = 1; al < Q; ++al)
      {
        const double ex = S.e[al][0], ey = S.e[al][1];
        const bool   nx = ex != 0., ny = ey != 0.;
        if (!nx && !ny)
          continue; // at rest in this sweep: r = 0 (streaming.h:118-122)

        // stage f_al (and h_al = a f + b feq for Lee-Lin) of the halo'd block
#pragma unroll
        for (int r = 0; r < R; ++r)
          {
            const int k = tid + r * kThreads;
            if (k < NH)
              {
                const int kx = k % HX, ky = k / HX;
                fld[ky][kx] = v[r][al];
                if constexpr (LL)
                  {
                    double fe;
                    if (feq_in == nullptr)
                      {
                        const double rho = mom[r][0], ux = mom[r][1], uy = mom[r][2];
                        fe               = feq_rt(al, rho, ux, uy, ux * ux + uy * uy);
                      }
                    else
                      {
                        const int gi = wrap(i0 - p + kx, P, G.ax.periodic), gj = wrap(j0 - p + ky, P, G.ay.periodic);
                        fe = (gi >= 0 && gj >= 0) ? feq_in[al * G.stride + node_index(G, gi, gj)] : 0.;
                      }
                    hld[ky][kx] = a_c * v[r][al] + b_c * fe;
                  }
              }
          }
        __syncthreads();

        // y pass: owned rows, all halo columns
        for (int idx = tid; idx < TY * HX; idx += kThreads)
          {
            const int row = idx / HX, col = idx % HX;
            const int gj  = j0 + row;
            if (gj > P)
              continue;
            double My = 0., Dy = 0., DTy = 0., Ky = 0., Myh = 0., Dyh = 0.;
#pragma unroll
            for (int o = -p; o <= p; ++o)
              {
                const double fv = fld[row + p + o][col];
                if (nx)
                  My += band<p>(G.ay.M, gj, o) * fv;
                if (ny)
                  {
                    Dy += band<p>(G.ay.D, gj, o) * fv;
                    Ky += band<p>(G.ay.K, gj, o) * fv;
                  }
                if (nx && ny)
                  DTy += band<p>(G.ay.DT, gj, o) * fv;
                if constexpr (LL)
                  {
                    const double hv = hld[row + p + o][col];
                    if (nx)
                      Myh += band<p>(G.ay.M, gj, o) * hv;
                    if (ny)
                      Dyh += band<p>(G.ay.D, gj, o) * hv;
                  }
              }
            if constexpr (!LL)
              {
                Myh = a_c * My;
                Dyh = a_c * Dy;
              }
            // r = Dx U_D + Kx U_K + Mx U_M + Dx^T U_DT
            yv[0][row][col] = ex * Myh + c_c * ex * ey * DTy; // U_D
            yv[1][row][col] = c_c * ex * ex * My;               // U_K
            yv[2][row][col] = ey * Dyh + c_c * ey * ey * Ky;   // U_M
            yv[3][row][col] = c_c * ex * ey * Dy;               // U_DT
          }
        __syncthreads();

        // x pass at the owned nodes, plus the wall-face terms
#pragma unroll
        for (int r = 0; r < RO; ++r)
          {
            const int ii = tid % TX, jj = tid / TX + r * (kThreads / TX);
            const int gi = i0 + ii, gj = j0 + jj;
            if (gi > P || gj > P)
              continue;
            double acc = 0.;
#pragma unroll
            for (int o = -p; o <= p; ++o)
              {
                const int c = ii + p + o;
                if (nx)
                  acc += band<p>(G.ax.D, gi, o) * yv[0][jj][c] + band<p>(G.ax.K, gi, o) * yv[1][jj][c];
                if (ny)
                  acc += band<p>(G.ax.M, gi, o) * yv[2][jj][c];
                if (nx && ny)
                  acc += band<p>(G.ax.DT, gi, o) * yv[3][jj][c];
              }
            // surface term (dt^2/2) oint phi (n.e)(e.grad f) on walls (advection.h:177-199),
            // as 1-D integrals along the wall: Dx for the tangential, Mx of the normal derivative
            double face = 0.;
            const int ty = jj + p, tx = ii + p;
            if (!G.ay.periodic && ny && (gj == 0 || gj == P))
              {
                const bool    lo = gj == 0;
                const double *bn = lo ? G.ay.b0 : G.ay.b1;
                const int     r0 = lo ? ty : ty - p; // first row of the boundary cell
                double        t  = 0., nrm = 0.;
#pragma unroll
                for (int o = -p; o <= p; ++o)
                  {
                    double dn = 0.;
#pragma unroll
                    for (int l = 0; l <= p; ++l)
                      dn += bn[l] * fld[r0 + l][tx + o];
                    t += band<p>(G.ax.D, gi, o) * fld[ty][tx + o];
                    nrm += band<p>(G.ax.M, gi, o) * dn;
                  }
                face += hdt2 * (lo ? -ey : ey) * (ex * t + ey * nrm);
              }
            if (!G.ax.periodic && nx && (gi == 0 || gi == P))
              {
                const bool    lo = gi == 0;
                const double *bn = lo ? G.ax.b0 : G.ax.b1;
                const int     c0 = lo ? tx : tx - p;
                double        t  = 0., nrm = 0.;
#pragma unroll
                for (int o = -p; o <= p; ++o)
                  {
                    double dn = 0.;
#pragma unroll
                    for (int l = 0; l <= p; ++l)
                      dn += bn[l] * fld[ty + o][c0 + l];
                    t += band<p>(G.ay.D, gj, o) * fld[ty + o][tx];
                    nrm += band<p>(G.ay.M, gj, o) * dn;
                  }
                face += hdt2 * (lo ? -ex : ex) * (ex * nrm + ey * t);
              }
            mflux[r] += face;
            xinc[r][al - 1] = acc + face;
          }
        __syncthreads();
      }

    // ---------------------------------------------------------- epilogue
#pragma unroll
    for (int r = 0; r < RO; ++r)
      {
        const int ii = tid % TX, jj = tid / TX + r * (kThreads / TX);
        const int gi = i0 + ii, gj = j0 + jj;
        if (gi > P || gj > P)
          continue;
        const int64_t idx   = node_index(G, gi, gj);
        const bool    slave = (G.ax.periodic && gi == P) || (G.ay.periodic && gj == P);
        // mass flux of the wall term removed isotropically (advection.h:194-196)
        if (S.remove_mass_flux)
#pragma unroll
          for (int a = 1; a < Q; ++a)
            xinc[r][a - 1] -= (w_of(a) / (1. - w_of(0))) * mflux[r];
        if constexpr (STEP)
          {
            const double mli = slave ? 0. : G.ax.ml_inv[gi] * G.ay.ml_inv[gj];
#pragma unroll
            for (int a = 0; a < NM; ++a)
              xinc[r][a] *= mli;
            if constexpr (LL)
              {
                double wx = 0., wy = 0.;
                const bool wall = wall_of(G, gi, gj, wx, wy);
                double     fe[Q];
                equilibrium_node(own[r], fe, wall, wx, wy);
                predictor_corrector_node(own[r], fe, xinc[r], S.theta, wall, wx, wy);
              }
            else
#pragma unroll
              for (int a = 1; a < Q; ++a)
                own[r][a] += xinc[r][a - 1];
#pragma unroll
            for (int a = 0; a < Q; ++a)
              out[a * G.stride + idx] = own[r][a];
          }
        else
          {
            // right-hand side (constrained entries get nothing) and, fused, g*
#pragma unroll
            for (int a = 0; a < NM; ++a)
              rhs_out[a * G.stride + idx] = slave ? 0. : xinc[r][a];
            if (out != nullptr)
#pragma unroll
              for (int a = 0; a < Q; ++a)
                out[a * G.stride + idx] = own[r][a];
          }
      }
  }

  // ------------------------------------------------------------------ launch
  template <int p, int TX, int TY, int MODE>
  static hipError_t
  launch_tile(const Geo &G, TileParams S, const double *fin, const double *feq, double *out, double *rhs,
              hipStream_t stream)
  {
    S.tiles_x       = (G.P + 1 + TX - 1) / TX;
    S.tiles_y       = (G.P + 1 + TY - 1) / TY;
    const int nt    = S.tiles_x * S.tiles_y;
    const int nb8   = (nt + 7) & ~7;
    hipLaunchKernelGGL((k_tile<p, TX, TY, MODE>), dim3(nb8), dim3(kThreads), 0, stream, G, S, fin, feq, out, rhs);
    return hipGetLastError();
  }

  template <int MODE>
  static hipError_t
  dispatch_degree(const Geo &G, const TileParams &S, const double *fin, const double *feq, double *out, double *rhs,
                  hipStream_t stream)
  {
    switch (G.p)
      {
        case 1:
          return launch_tile<1, 32, 16, MODE>(G, S, fin, feq, out, rhs, stream);
        case 2:
          return launch_tile<2, 32, 16, MODE>(G, S, fin, feq, out, rhs, stream);
        case 3:
          return launch_tile<3, 32, 16, MODE>(G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  hipError_t
  tile_launch(const int mode, const Geo &G, const TileParams &S, const double *fin, const double *feq, double *out,
              double *rhs, hipStream_t stream)
  {
    switch (mode)
      {
        case kStepBardow:
          return dispatch_degree<kStepBardow>(G, S, fin, feq, out, rhs, stream);
        case kStepLeeLin:
          return dispatch_degree<kStepLeeLin>(G, S, fin, feq, out, rhs, stream);
        case kAdvectBardow:
          return dispatch_degree<kAdvectBardow>(G, S, fin, feq, out, rhs, stream);
        case kAdvectLeeLin:
          return dispatch_degree<kAdvectLeeLin>(G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }
} // namespace lbfem_amd
