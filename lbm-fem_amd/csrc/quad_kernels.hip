// General quadrilaterals: the advection and mass operators on a periodic Q_p
// mesh whose cells are the bilinear (MappingQ1) images of their four vertices
// (the vortex's --distort mesh, include/lbfem/test_cases.h:53-65,
// docs/method.md:216-224). Without the tensor structure the operators are
// evaluated cell by cell as the reference's cell loop does
// (include/lbfem/advection.h:94-124, mass.h:105-144): values and reference
// gradients at the (p+1)^2 Gauss points by sum factorisation, the inverse
// Jacobian and JxW of the cell's bilinear map per point (computed from the
// four vertices, 16 bytes per node in HBM rather than a stored Jacobian), the
// point operation, and the integration against the test functions.
//
// A block owns the nodes of a tile of CX x CY cells (the nodes of each cell but
// its right column and top row; the last tiles also the periodic slave lines)
// and evaluates the (CX+1) x (CY+1) cells that touch them, one cell per thread,
// from a window of nodes staged in LDS. The cell results go to LDS and each
// owned node sums the cells around it in a fixed order (lower row first, left
// cell first), so results do not depend on scheduling and no atomics are used.
#include "device.hpp"
#include "kernels.hpp"

namespace lbfem_amd
{
  namespace
  {
    constexpr int kQT = 256; // threads per block

    template <int p>
    struct QuadShape;
    template <>
    struct QuadShape<1>
    {
      static constexpr int CX = 31, CY = 7; // 32 x 8 cells evaluated, 217 nodes owned
    };
    template <>
    struct QuadShape<2>
    {
      static constexpr int CX = 15, CY = 7;
    };
    template <>
    struct QuadShape<3>
    {
      static constexpr int CX = 9, CY = 7;
    };

    template <int p>
    struct QT
    {
      static constexpr int CX = QuadShape<p>::CX, CY = QuadShape<p>::CY;
      static constexpr int n1 = p + 1, NQ = n1 * n1;
      static constexpr int WC = (CX + 1) * (CY + 1);                        // cells evaluated
      static constexpr int HX = p * (CX + 1) + 1, HY = p * (CY + 1) + 1;    // window nodes
      static constexpr int OWN = (p * CX + 1) * (p * CY + 1);               // owned nodes, at most
      static_assert(WC <= kQT, "one cell per thread");
    };

    __device__ __forceinline__ int
    pmod(const int i, const int m)
    {
      const int r = i % m;
      return r < 0 ? r + m : r;
    }

    // the tile's cells and owned nodes
    template <int p>
    struct Box
    {
      int cx0, cy0, wcx, wcy; // first owned cell, cells evaluated per row and column
      int oi0, oi1, oj0, oj1; // owned node lines [oi0, oi1) x [oj0, oj1), unwrapped
      __device__
      Box(const Geo &G, const int tile, const int tiles_x)
      {
        const int N = G.gen.N;
        cx0         = (tile % tiles_x) * QT<p>::CX;
        cy0         = (tile / tiles_x) * QT<p>::CY;
        const int cx1 = min(cx0 + QT<p>::CX, N), cy1 = min(cy0 + QT<p>::CY, N);
        wcx           = cx1 - cx0 + 1;
        wcy           = cy1 - cy0 + 1;
        oi0           = p * cx0;
        oj0           = p * cy0;
        oi1           = cx1 == N ? G.P + 1 : p * cx1;
        oj1           = cy1 == N ? G.P + 1 : p * cy1;
      }
    };

    // the four vertices of a cell
    struct Verts
    {
      double x00, y00, x10, y10, x01, y01, x11, y11;
    };

    __device__ __forceinline__ Verts
    cell_vertices(const Geo &G, const int gcx, const int gcy)
    {
      const int     N1 = G.gen.N + 1;
      const double *v  = G.gen.vert;
      const int     a = gcy * N1 + gcx, b = a + N1;
      Verts         c;
      c.x00 = v[2 * a];
      c.y00 = v[2 * a + 1];
      c.x10 = v[2 * a + 2];
      c.y10 = v[2 * a + 3];
      c.x01 = v[2 * b];
      c.y01 = v[2 * b + 1];
      c.x11 = v[2 * b + 2];
      c.y11 = v[2 * b + 3];
      return c;
    }

    // inverse Jacobian K[e][d] = dxi_e / dx_d and det J of the bilinear map at (xi, eta)
    __device__ __forceinline__ void
    inverse_jacobian(const Verts &c, const double xi, const double eta, double &k00, double &k01, double &k10,
                     double &k11, double &det)
    {
      const double xxi  = (1. - eta) * (c.x10 - c.x00) + eta * (c.x11 - c.x01);
      const double yxi  = (1. - eta) * (c.y10 - c.y00) + eta * (c.y11 - c.y01);
      const double xeta = (1. - xi) * (c.x01 - c.x00) + xi * (c.x11 - c.x10);
      const double yeta = (1. - xi) * (c.y01 - c.y00) + xi * (c.y11 - c.y10);
      det               = xxi * yeta - xeta * yxi;
      const double id   = 1. / det;
      k00               = yeta * id;
      k01               = -xeta * id;
      k10               = -yxi * id;
      k11               = xxi * id;
    }

    // the 1-D tables in LDS: S[q][k], D[q][k], qx[q], qw[q]
    struct Basis
    {
      double S[4][4], D[4][4], qx[4], qw[4];
    };

    __device__ __forceinline__ void
    stage_basis(const Geo &G, Basis &B)
    {
      double *dst = &B.S[0][0];
      for (int k = threadIdx.x; k < 40; k += blockDim.x)
        dst[k] = G.gen.basis[k];
    }

    // Reference gradients (d/dxi, d/deta) at the Gauss points of a cell field u
    // (FEEvaluation::evaluate(gradients)): sum factorisation, x direction first.
    template <int p>
    __device__ __forceinline__ void
    ref_gradients(const Basis &B, const double (&u)[p + 1][p + 1], double (&uxi)[p + 1][p + 1],
                  double (&ueta)[p + 1][p + 1])
    {
      constexpr int n = p + 1;
      if constexpr (p == 1)
        {
          // Q1: phi' = -1, +1 (exact): d/dxi is the x difference interpolated in y, and d/deta alike
          const double dx0 = u[0][1] - u[0][0], dx1 = u[1][1] - u[1][0];
          const double dy0 = u[1][0] - u[0][0], dy1 = u[1][1] - u[0][1];
#pragma unroll
          for (int q = 0; q < 2; ++q)
            {
              const double gxi = B.S[q][0] * dx0 + B.S[q][1] * dx1; // at eta_q, any xi
              const double get = B.S[q][0] * dy0 + B.S[q][1] * dy1; // at xi_q, any eta
              uxi[q][0] = uxi[q][1] = gxi;
              ueta[0][q] = ueta[1][q] = get;
            }
          return;
        }
      double        A[n][n], Bv[n][n]; // [ky][qx]
#pragma unroll
      for (int ky = 0; ky < n; ++ky)
#pragma unroll
        for (int qx = 0; qx < n; ++qx)
          {
            double a = 0., b = 0.;
#pragma unroll
            for (int kx = 0; kx < n; ++kx)
              {
                a += B.D[qx][kx] * u[ky][kx];
                b += B.S[qx][kx] * u[ky][kx];
              }
            A[ky][qx]  = a;
            Bv[ky][qx] = b;
          }
#pragma unroll
      for (int qy = 0; qy < n; ++qy)
#pragma unroll
        for (int qx = 0; qx < n; ++qx)
          {
            double a = 0., b = 0.;
#pragma unroll
            for (int ky = 0; ky < n; ++ky)
              {
                a += B.S[qy][ky] * A[ky][qx];
                b += B.D[qy][ky] * Bv[ky][qx];
              }
            uxi[qy][qx]  = a;
            ueta[qy][qx] = b;
          }
    }

    // values at the Gauss points (FEEvaluation::evaluate(values))
    template <int p>
    __device__ __forceinline__ void
    ref_values(const Basis &B, const double (&u)[p + 1][p + 1], double (&uq)[p + 1][p + 1])
    {
      constexpr int n = p + 1;
      double        T[n][n];
#pragma unroll
      for (int ky = 0; ky < n; ++ky)
#pragma unroll
        for (int qx = 0; qx < n; ++qx)
          {
            double s = 0.;
#pragma unroll
            for (int kx = 0; kx < n; ++kx)
              s += B.S[qx][kx] * u[ky][kx];
            T[ky][qx] = s;
          }
#pragma unroll
      for (int qy = 0; qy < n; ++qy)
#pragma unroll
        for (int qx = 0; qx < n; ++qx)
          {
            double s = 0.;
#pragma unroll
            for (int ky = 0; ky < n; ++ky)
              s += B.S[qy][ky] * T[ky][qx];
            uq[qy][qx] = s;
          }
    }

    // r[ky][kx] = sum_q phi v + dphi/dxi Fxi + dphi/deta Feta (FEEvaluation::integrate),
    // entry k written to R[k * ld]
    template <int p>
    __device__ __forceinline__ void
    integrate(const Basis &B, const double (&v)[p + 1][p + 1], const double (&fxi)[p + 1][p + 1],
              const double (&feta)[p + 1][p + 1], double *R, const int ld)
    {
      constexpr int n = p + 1;
      if constexpr (p == 1)
        {
          // Q1: phi' = -1, +1 on [0,1] (exact), so the flux terms need only sums over one direction
          double Vx[2][2]; // [qy][kx]
#pragma unroll
          for (int qy = 0; qy < 2; ++qy)
#pragma unroll
            for (int kx = 0; kx < 2; ++kx)
              Vx[qy][kx] = B.S[0][kx] * v[qy][0] + B.S[1][kx] * v[qy][1];
          const double sx0 = fxi[0][0] + fxi[0][1], sx1 = fxi[1][0] + fxi[1][1];     // sum over qx, per qy
          const double sy0 = feta[0][0] + feta[1][0], sy1 = feta[0][1] + feta[1][1]; // sum over qy, per qx
#pragma unroll
          for (int ky = 0; ky < 2; ++ky)
            {
              const double ax = B.S[0][ky] * sx0 + B.S[1][ky] * sx1;
#pragma unroll
              for (int kx = 0; kx < 2; ++kx)
                {
                  const double by = B.S[0][kx] * sy0 + B.S[1][kx] * sy1;
                  const double vv = B.S[0][ky] * Vx[0][kx] + B.S[1][ky] * Vx[1][kx];
                  R[(ky * 2 + kx) * ld] = vv + (kx ? ax : -ax) + (ky ? by : -by);
                }
            }
          return;
        }
      double        V[n][n], Wm[n][n]; // [qy][kx]
#pragma unroll
      for (int qy = 0; qy < n; ++qy)
#pragma unroll
        for (int kx = 0; kx < n; ++kx)
          {
            double a = 0., b = 0.;
#pragma unroll
            for (int qx = 0; qx < n; ++qx)
              {
                a += B.S[qx][kx] * v[qy][qx] + B.D[qx][kx] * fxi[qy][qx];
                b += B.S[qx][kx] * feta[qy][qx];
              }
            V[qy][kx]  = a;
            Wm[qy][kx] = b;
          }
#pragma unroll
      for (int ky = 0; ky < n; ++ky)
#pragma unroll
        for (int kx = 0; kx < n; ++kx)
          {
            double a = 0.;
#pragma unroll
            for (int qy = 0; qy < n; ++qy)
              a += B.S[qy][ky] * V[qy][kx] + B.D[qy][ky] * Wm[qy][kx];
            R[(ky * n + kx) * ld] = a;
          }
    }

    // the sum over the cells around owned node (gi, gj) (unwrapped, not a slave)
    // of their entries for it, lower cell row first, left cell first
    template <int p>
    __device__ __forceinline__ double
    gather_node(const double (*R)[QT<p>::WC], const Box<p> &bx, const int gi, const int gj)
    {
      int cxs[2], kxs[2], cys[2], kys[2], mx = 0, my = 0;
      if (gi % p == 0)
        {
          cxs[0] = gi / p - 1;
          kxs[0] = p;
          cxs[1] = gi / p;
          kxs[1] = 0;
          mx     = 2;
        }
      else
        {
          cxs[0] = gi / p;
          kxs[0] = gi % p;
          mx     = 1;
        }
      if (gj % p == 0)
        {
          cys[0] = gj / p - 1;
          kys[0] = p;
          cys[1] = gj / p;
          kys[1] = 0;
          my     = 2;
        }
      else
        {
          cys[0] = gj / p;
          kys[0] = gj % p;
          my     = 1;
        }
      double s = 0.;
      for (int b = 0; b < my; ++b)
        for (int a = 0; a < mx; ++a)
          {
            const int lc = (cys[b] - (bx.cy0 - 1)) * bx.wcx + (cxs[a] - (bx.cx0 - 1));
            s += R[kys[b] * (p + 1) + kxs[a]][lc];
          }
      return s;
    }
  } // namespace

  // ------------------------------------------------------------------ advection
  // The right-hand sides of the eight moving populations (advection.h:211-293):
  // Bardow v = a e.grad g, Lee-Lin v = e.grad(a f + b feq); flux c e (e.grad f).
  // Bardow may collide on load (writing g* of the owned nodes to out); Lee-Lin
  // takes the equilibria given or computes them from f. Constrained rows get 0.
  // STEP: the whole lumped TG2 step in one launch (schemes.h:92-98, 118-124,
  // mass.h:189-191): the increments x = M_L^-1 r of an owned node are kept in
  // LDS and the node is updated in place of writing r: Bardow f = g* + x,
  // Lee-Lin the predictor-corrector with the node's own equilibria.
  template <int p, bool LL, bool STEP>
  __global__ void __launch_bounds__(kQT)
  k_quad_advect(const Geo G, const TileParams S, const int tiles_x, const double *__restrict__ fin,
                const double *__restrict__ feq_in, double *__restrict__ out, double *__restrict__ rhs)
  {
    using C              = QT<p>;
    constexpr int n1     = C::n1, NQ = C::NQ, HX = C::HX, HY = C::HY, NF = LL ? 2 * NM : NM;
    // Lee-Lin's step keeps population 0 (plane NF) and the increments of the owned
    // nodes in LDS for the nodal epilogue; Bardow's step updates population by population
    constexpr bool    KEEP = STEP && LL;
    // Q1 owns at most one node per thread: its increments stay in registers (the population
    // loop unrolled), which keeps the LDS of three blocks per CU
    constexpr bool    XREG = KEEP && C::OWN <= kQT;
    constexpr int     NF0 = KEEP ? 1 : 0, NX = (KEEP && !XREG) ? C::OWN : 1;
    __shared__ double F[NF + NF0][HY][HX];
    __shared__ double R[NQ][C::WC]; // local node-major: a wavefront's cells are consecutive
    __shared__ double XI[(KEEP && !XREG) ? NM : 1][NX];
    double            xreg[NM];
#pragma unroll
    for (int a = 0; a < NM; ++a)
      xreg[a] = 0.;
    __shared__ Basis  B;

    const int nt = tiles_x * ((G.gen.N + C::CY - 1) / C::CY);
    const int nb8 = (nt + 7) & ~7, bid = blockIdx.x;
    const int tile = (bid & 7) * (nb8 >> 3) + (bid >> 3); // XCD-aware order
    if (tile >= nt)
      return;
    const Box<p> bx(G, tile, tiles_x);
    const int    P = G.P, tid = threadIdx.x;
    stage_basis(G, B);

    // stage the window (collided for Bardow's first sweep) and write g* of the owned nodes
    const int wi = p * bx.wcx + 1, wj = p * bx.wcy + 1;
    for (int k = tid; k < wi * wj; k += kQT)
      {
        const int     li = k % wi, lj = k / wi;
        const int     gi = bx.oi0 - p + li, gj = bx.oj0 - p + lj;
        const int64_t idx = node_index(G, pmod(gi, P), pmod(gj, P));
        double        v[Q];
#pragma unroll
        for (int a = 0; a < Q; ++a)
          v[a] = fin[a * G.stride + idx];
        if constexpr (!LL)
          if (S.collide)
            collide_node(v, S.omega, false, 0., 0.);
        const bool owned = gi >= bx.oi0 && gi < bx.oi1 && gj >= bx.oj0 && gj < bx.oj1;
        if constexpr (KEEP)
          F[NF][lj][li] = v[0];
        if (STEP && !LL && owned) // Bardow's step: g*_0 now, a slave's whole (collided) update now
          {
            if (gi == P || gj == P)
              {
                const int64_t own = node_index(G, gi, gj);
                double        w[Q];
#pragma unroll
                for (int a = 0; a < Q; ++a)
                  w[a] = fin[a * G.stride + own];
                if (S.collide)
                  collide_node(w, S.omega, false, 0., 0.);
#pragma unroll
                for (int a = 0; a < Q; ++a)
                  out[a * G.stride + own] = w[a];
              }
            else
              out[idx] = v[0];
          }
        if (!STEP && owned && out != nullptr)
          {
            if (gi == P || gj == P) // a slave: its own entries, collided
              {
                const int64_t own = node_index(G, gi, gj);
                double        w[Q];
#pragma unroll
                for (int a = 0; a < Q; ++a)
                  w[a] = fin[a * G.stride + own];
                if constexpr (!LL)
                  if (S.collide)
                    collide_node(w, S.omega, false, 0., 0.);
#pragma unroll
                for (int a = 0; a < Q; ++a)
                  out[a * G.stride + own] = w[a];
              }
            else
#pragma unroll
              for (int a = 0; a < Q; ++a)
                out[a * G.stride + idx] = v[a];
          }
#pragma unroll
        for (int a = 1; a < Q; ++a)
          F[a - 1][lj][li] = v[a];
        if constexpr (LL)
          {
            double fe[Q];
            if (feq_in != nullptr)
#pragma unroll
              for (int a = 0; a < Q; ++a)
                fe[a] = feq_in[a * G.stride + idx];
            else
              equilibrium_node(v, fe, false, 0., 0.);
#pragma unroll
            for (int a = 1; a < Q; ++a)
              F[NM + a - 1][lj][li] = S.a * v[a] + S.b * fe[a];
          }
      }

    // this thread's cell
    const bool has_cell = tid < bx.wcx * bx.wcy;
    const int  lcx = has_cell ? tid % bx.wcx : 0, lcy = has_cell ? tid / bx.wcx : 0;
    Verts      cv{};
    if (has_cell)
      cv = cell_vertices(G, pmod(bx.cx0 - 1 + lcx, G.gen.N), pmod(bx.cy0 - 1 + lcy, G.gen.N));
    __syncthreads();
    // Q1: the cell's inverse Jacobians and JxW at its four points, once for all
    // populations (40 registers); higher degrees recompute them per population
    constexpr bool GEO = p == 1;
    double         gk[GEO ? NQ : 1][5];
    if constexpr (GEO)
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        {
          const int qx = q % n1, qy = q / n1;
          double    det;
          inverse_jacobian(cv, B.qx[qx], B.qx[qy], gk[q][0], gk[q][1], gk[q][2], gk[q][3], det);
          gk[q][4] = det * B.qw[qx] * B.qw[qy];
        }

    const int ow = bx.oi1 - bx.oi0, nown = ow * (bx.oj1 - bx.oj0);
    constexpr int UNROLL_POP = XREG ? NM : 1;
#pragma unroll UNROLL_POP
    for (int al = 1; al < Q; ++al)
      {
        const double ex = S.e[al][0], ey = S.e[al][1];
        double      *ra = STEP ? nullptr : rhs + (al - 1) * G.stride; // the step writes no right-hand side
        if (ex == 0. && ey == 0.) // at rest in this sweep: r = 0 (streaming.h:118-122)
          {
            for (int k = tid; k < nown; k += kQT)
              {
                const int gi = bx.oi0 + k % ow, gj = bx.oj0 + k / ow;
                if constexpr (XREG)
                  xreg[al - 1] = 0.;
                else if constexpr (KEEP)
                  XI[al - 1][k] = 0.;
                else if constexpr (STEP)
                  {
                    if (gi < P && gj < P)
                      out[al * G.stride + node_index(G, gi, gj)] = F[al - 1][gj - bx.oj0 + p][gi - bx.oi0 + p];
                  }
                else
                  ra[node_index(G, gi, gj)] = 0.;
              }
            continue;
          }
        if (has_cell)
          {
            double u[n1][n1], uxi[n1][n1], ueta[n1][n1];
#pragma unroll
            for (int ky = 0; ky < n1; ++ky)
#pragma unroll
              for (int kx = 0; kx < n1; ++kx)
                u[ky][kx] = F[al - 1][p * lcy + ky][p * lcx + kx];
            ref_gradients<p>(B, u, uxi, ueta);
            double hxi[n1][n1], heta[n1][n1];
            if constexpr (LL)
              {
#pragma unroll
                for (int ky = 0; ky < n1; ++ky)
#pragma unroll
                  for (int kx = 0; kx < n1; ++kx)
                    u[ky][kx] = F[NM + al - 1][p * lcy + ky][p * lcx + kx];
                ref_gradients<p>(B, u, hxi, heta);
              }
            double v[n1][n1], fxi[n1][n1], feta[n1][n1];
#pragma unroll
            for (int qy = 0; qy < n1; ++qy)
#pragma unroll
              for (int qx = 0; qx < n1; ++qx)
                {
                  double k00, k01, k10, k11, jxw;
                  if constexpr (GEO)
                    {
                      const int q = qy * n1 + qx;
                      k00         = gk[q][0];
                      k01         = gk[q][1];
                      k10         = gk[q][2];
                      k11         = gk[q][3];
                      jxw         = gk[q][4];
                    }
                  else
                    {
                      double det;
                      inverse_jacobian(cv, B.qx[qx], B.qx[qy], k00, k01, k10, k11, det);
                      jxw = det * B.qw[qx] * B.qw[qy];
                    }
                  // e . grad = (K e) . grad_xi, and the flux c e (e . grad f) JxW tested with grad phi
                  // is c (e . grad f) JxW (K e) . grad_xi phi
                  const double w0 = k00 * ex + k01 * ey, w1 = k10 * ex + k11 * ey;
                  const double sj = (w0 * uxi[qy][qx] + w1 * ueta[qy][qx]) * jxw;
                  if constexpr (LL)
                    v[qy][qx] = (w0 * hxi[qy][qx] + w1 * heta[qy][qx]) * jxw;
                  else
                    v[qy][qx] = S.a * sj;
                  const double t = S.c * sj;
                  fxi[qy][qx]    = t * w0;
                  feta[qy][qx]   = t * w1;
                }
            integrate<p>(B, v, fxi, feta, &R[0][tid], C::WC);
          }
        __syncthreads();
        for (int k = tid; k < nown; k += kQT)
          {
            const int    gi = bx.oi0 + k % ow, gj = bx.oj0 + k / ow;
            const double r  = (gi < P && gj < P) ? gather_node<p>(R, bx, gi, gj) : 0.; // 0 on slaves
            if constexpr (XREG)
              xreg[al - 1] = r;
            else if constexpr (KEEP)
              XI[al - 1][k] = r;
            else if constexpr (STEP)
              {
                if (gi < P && gj < P) // g* + M_L^-1 r
                  {
                    const int64_t idx = node_index(G, gi, gj);
                    out[al * G.stride + idx] = F[al - 1][gj - bx.oj0 + p][gi - bx.oi0 + p] + r * G.gen.ml_inv[idx];
                  }
              }
            else
              ra[node_index(G, gi, gj)] = r;
          }
        __syncthreads(); // R is rewritten by the next population
      }

    if constexpr (KEEP) // Lee-Lin: x = M_L^-1 r, then the predictor-corrector with the node's own equilibria
      for (int k = tid; k < nown; k += kQT)
        {
          const int     gi = bx.oi0 + k % ow, gj = bx.oj0 + k / ow;
          const int64_t idx   = node_index(G, gi, gj);
          const bool    slave = gi == P || gj == P;
          double        own[Q], x[NM];
          if (slave) // its own entries (the window holds its master's); only nodal work applies
            {
#pragma unroll
              for (int a = 0; a < Q; ++a)
                own[a] = fin[a * G.stride + idx];
            }
          else
            {
              const int li = gi - bx.oi0 + p, lj = gj - bx.oj0 + p;
              own[0]       = F[NF][lj][li];
#pragma unroll
              for (int a = 1; a < Q; ++a)
                own[a] = F[a - 1][lj][li];
            }
          const double mli = slave ? 0. : G.gen.ml_inv[idx];
#pragma unroll
          for (int a = 0; a < NM; ++a)
            x[a] = (XREG ? xreg[a] : XI[a][k]) * mli;
          double fe[Q];
          equilibrium_node(own, fe, false, 0., 0.);
          predictor_corrector_node(own, fe, x, S.theta, false, 0., 0.);
#pragma unroll
          for (int a = 0; a < Q; ++a)
            out[a * G.stride + idx] = own[a];
        }
  }

  // ------------------------------------------------------------------ mass operator
  // KIND 0: y = A x; KIND 1 (CG start): r = A x - b and the partials |r|^2,
  // |b|^2, r.D^-1 r; KIND 2 (CG direction): p_new = beta p_old - D^-1 r, q = A
  // p_new and the partial p_new.q. A = M + c K_e with constrained rows the
  // identity (mass.h:105-144). One vector per block (blockIdx.y), the same
  // partial-sum layout as the tensor-mesh operator. STIFF: some vector has a
  // K_e term (TG3); without one the gradient arrays are not instantiated
  // (registers set the occupancy).
  template <int p, int KIND, bool STIFF>
  __global__ void __launch_bounds__(kQT)
  k_quad_operator(const Geo G, const CgWork W, const int it, const int tiles_x, const double *x, const double *b,
                  const double *r_in, double *y, double *y2)
  {
    using C          = QT<p>;
    constexpr int n1 = C::n1, NQ = C::NQ, HX = C::HX, HY = C::HY;
    __shared__ double X[HY][HX];
    __shared__ double R[NQ][C::WC]; // local node-major: a wavefront's cells are consecutive
    __shared__ double red[3][kQT / 64];
    __shared__ Basis  B;

    const int nt = tiles_x * ((G.gen.N + C::CY - 1) / C::CY);
    const int nb8 = (nt + 7) & ~7, bid = blockIdx.x;
    const int tile = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    if (tile >= nt)
      return;
    const int tid = threadIdx.x, P = G.P;
    const int vsel = blockIdx.y;
    if (KIND == 2 && W.state[vsel].status != 0)
      {
        if (tid == 0)
          W.partial[size_t(tile) * kCgPartials + 3 * vsel] = 0.;
        return;
      }
    const Box<p>  bx(G, tile, tiles_x);
    const double  beta = (KIND == 2 && it > 1) ? W.state[vsel].beta : 0.;
    const int64_t vb   = int64_t(W.block[vsel]) * G.stride;
    const double  c = W.vec[vsel].coef, ex = W.vec[vsel].e[0], ey = W.vec[vsel].e[1];
    const bool    stiff = STIFF && !(c == 0. || (ex == 0. && ey == 0.));
    stage_basis(G, B);

    const int wi = p * bx.wcx + 1, wj = p * bx.wcy + 1;
    for (int k = tid; k < wi * wj; k += kQT)
      {
        const int     li = k % wi, lj = k / wi;
        const int     gi = pmod(bx.oi0 - p + li, P), gj = pmod(bx.oj0 - p + lj, P);
        const int64_t idx = node_index(G, gi, gj);
        double        val = (KIND != 2 || it > 1) ? x[vb + idx] : 0.;
        if constexpr (KIND == 2)
          {
            const double z = G.gen.diag_inv[idx] * r_in[vb + idx];
            val            = it > 1 ? beta * val - z : -z;
          }
        X[lj][li] = val;
      }
    const bool has_cell = tid < bx.wcx * bx.wcy;
    const int  lcx = has_cell ? tid % bx.wcx : 0, lcy = has_cell ? tid / bx.wcx : 0;
    Verts      cv{};
    if (has_cell)
      cv = cell_vertices(G, pmod(bx.cx0 - 1 + lcx, G.gen.N), pmod(bx.cy0 - 1 + lcy, G.gen.N));
    __syncthreads();

    if (has_cell)
      {
        constexpr int ng = STIFF ? n1 : 1; // the gradient arrays of the K_e term
        double        u[n1][n1], uq[n1][n1], uxi[ng][ng], ueta[ng][ng];
#pragma unroll
        for (int ky = 0; ky < n1; ++ky)
#pragma unroll
          for (int kx = 0; kx < n1; ++kx)
            u[ky][kx] = X[p * lcy + ky][p * lcx + kx];
        ref_values<p>(B, u, uq);
        if constexpr (STIFF)
          if (stiff)
            ref_gradients<p>(B, u, uxi, ueta);
        double v[n1][n1], fxi[n1][n1], feta[n1][n1];
#pragma unroll
        for (int qy = 0; qy < n1; ++qy)
#pragma unroll
          for (int qx = 0; qx < n1; ++qx)
            {
              double k00, k01, k10, k11, det;
              inverse_jacobian(cv, B.qx[qx], B.qx[qy], k00, k01, k10, k11, det);
              const double jxw = det * B.qw[qx] * B.qw[qy];
              v[qy][qx]        = uq[qy][qx] * jxw;
              fxi[qy][qx] = feta[qy][qx] = 0.;
              if constexpr (STIFF)
                if (stiff)
                  {
                    const double w0 = k00 * ex + k01 * ey, w1 = k10 * ex + k11 * ey; // K e
                    const double t  = c * (w0 * uxi[qy][qx] + w1 * ueta[qy][qx]) * jxw;
                    fxi[qy][qx]     = t * w0;
                    feta[qy][qx]    = t * w1;
                  }
            }
        integrate<p>(B, v, fxi, feta, &R[0][tid], C::WC);
      }
    __syncthreads();

    double    part[3] = {0., 0., 0.};
    const int ow = bx.oi1 - bx.oi0, nown = ow * (bx.oj1 - bx.oj0);
    for (int k = tid; k < nown; k += kQT)
      {
        const int     gi = bx.oi0 + k % ow, gj = bx.oj0 + k / ow;
        const int64_t idx   = node_index(G, gi, gj);
        const bool    slave = gi == P || gj == P;
        double        own, a;
        if (slave) // constrained rows are the identity, on the slave's own entry
          {
            if constexpr (KIND == 2)
              own = it > 1 ? beta * x[vb + idx] - r_in[vb + idx] : -r_in[vb + idx]; // D^-1 = 1
            else
              own = x[vb + idx];
            a = own;
          }
        else
          {
            own = X[gj - bx.oj0 + p][gi - bx.oi0 + p];
            a   = gather_node<p>(R, bx, gi, gj);
          }
        if constexpr (KIND == 0)
          y[vb + idx] = a;
        else if constexpr (KIND == 1)
          {
            const double bb = b[vb + idx];
            const double r  = a - bb;
            y[vb + idx]     = r;
            part[0] += r * r;
            part[1] += bb * bb;
            part[2] += r * (G.gen.diag_inv[idx] * r);
          }
        else
          {
            y[vb + idx]  = own; // p_new
            y2[vb + idx] = a;   // q = A p_new
            part[0] += own * a;
          }
      }
    if constexpr (KIND != 0)
      {
        block_sum_n<3, kQT>(part, red);
        if (tid == 0)
          for (int k = 0; k < (KIND == 1 ? 3 : 1); ++k)
            W.partial[size_t(tile) * kCgPartials + 3 * vsel + k] = part[k];
      }
  }

  // ------------------------------------------------------------------ launchers
  template <int p>
  static void
  quad_tiles(const Geo &G, int &tiles_x, int &nt)
  {
    tiles_x = (G.gen.N + QT<p>::CX - 1) / QT<p>::CX;
    nt      = tiles_x * ((G.gen.N + QT<p>::CY - 1) / QT<p>::CY);
  }

  int
  quad_tile_count(const Geo &G)
  {
    int tx = 0, nt = 0;
    switch (G.p)
      {
        case 1:
          quad_tiles<1>(G, tx, nt);
          break;
        case 2:
          quad_tiles<2>(G, tx, nt);
          break;
        case 3:
          quad_tiles<3>(G, tx, nt);
          break;
        default:
          break;
      }
    return nt;
  }

  template <int p>
  static hipError_t
  quad_advect_degree(const int mode, const Geo &G, const TileParams &S, const double *fin, const double *feq,
                     double *out, double *rhs, hipStream_t s)
  {
    int tx, nt;
    quad_tiles<p>(G, tx, nt);
    const dim3 grid((nt + 7) & ~7), block(kQT);
    switch (mode)
      {
        case kStepBardow:
          hipLaunchKernelGGL((k_quad_advect<p, false, true>), grid, block, 0, s, G, S, tx, fin, feq, out, rhs);
          break;
        case kStepLeeLin:
          hipLaunchKernelGGL((k_quad_advect<p, true, true>), grid, block, 0, s, G, S, tx, fin, feq, out, rhs);
          break;
        case kAdvectBardow:
          hipLaunchKernelGGL((k_quad_advect<p, false, false>), grid, block, 0, s, G, S, tx, fin, feq, out, rhs);
          break;
        case kAdvectLeeLin:
          hipLaunchKernelGGL((k_quad_advect<p, true, false>), grid, block, 0, s, G, S, tx, fin, feq, out, rhs);
          break;
        default:
          return hipErrorInvalidValue;
      }
    return hipGetLastError();
  }

  hipError_t
  quad_advect_launch(const int mode, const Geo &G, const TileParams &S, const double *fin, const double *feq,
                     double *out, double *rhs, hipStream_t s)
  {
    if (G.gen.vert == nullptr)
      return hipErrorInvalidValue;
    switch (G.p)
      {
        case 1:
          return quad_advect_degree<1>(mode, G, S, fin, feq, out, rhs, s);
        case 2:
          return quad_advect_degree<2>(mode, G, S, fin, feq, out, rhs, s);
        case 3:
          return quad_advect_degree<3>(mode, G, S, fin, feq, out, rhs, s);
        default:
          return hipErrorInvalidValue;
      }
  }

  template <int p, int KIND>
  static hipError_t
  quad_operator_degree(const Geo &G, const CgWork &W, const int it, const double *x, const double *b, const double *r,
                       double *y, double *y2, hipStream_t s)
  {
    int tx, nt;
    quad_tiles<p>(G, tx, nt);
    if (KIND != 0 && nt > W.max_blocks)
      return hipErrorInvalidValue;
    const int nb8 = (nt + 7) & ~7;
    bool      stiff = false;
    for (int v = 0; v < W.nv; ++v)
      stiff |= W.vec[v].coef != 0. && (W.vec[v].e[0] != 0. || W.vec[v].e[1] != 0.);
    if (stiff)
      hipLaunchKernelGGL((k_quad_operator<p, KIND, true>), dim3(nb8, W.nv), dim3(kQT), 0, s, G, W, it, tx, x, b, r, y,
                         y2);
    else
      hipLaunchKernelGGL((k_quad_operator<p, KIND, false>), dim3(nb8, W.nv), dim3(kQT), 0, s, G, W, it, tx, x, b, r,
                         y, y2);
    return hipGetLastError();
  }

  template <int KIND>
  hipError_t
  quad_operator_launch(const Geo &G, const CgWork &W, const int it, const double *x, const double *b, const double *r,
                       double *y, double *y2, hipStream_t s)
  {
    if (G.gen.vert == nullptr)
      return hipErrorInvalidValue;
    switch (G.p)
      {
        case 1:
          return quad_operator_degree<1, KIND>(G, W, it, x, b, r, y, y2, s);
        case 2:
          return quad_operator_degree<2, KIND>(G, W, it, x, b, r, y, y2, s);
        case 3:
          return quad_operator_degree<3, KIND>(G, W, it, x, b, r, y, y2, s);
        default:
          return hipErrorInvalidValue;
      }
  }
  template hipError_t quad_operator_launch<0>(const Geo &, const CgWork &, int, const double *, const double *,
                                              const double *, double *, double *, hipStream_t);
  template hipError_t quad_operator_launch<1>(const Geo &, const CgWork &, int, const double *, const double *,
                                              const double *, double *, double *, hipStream_t);
  template hipError_t quad_operator_launch<2>(const Geo &, const CgWork &, int, const double *, const double *,
                                              const double *, double *, double *, hipStream_t);
} // namespace lbfem_amd
