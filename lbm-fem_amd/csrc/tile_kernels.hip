// Tile kernels of the streaming step: the weak-form advection right-hand side
// of all eight moving populations (include/lbfem/advection.h:94-293), fused
// with the nodal work before it (Bardow's BGK collision, or Lee & Lin's
// equilibria, collision.h:105-168) and, for the lumped mass, with the solve
// and the nodal work after it (f += M_L^-1 r, or the predictor-corrector).
//
// One workgroup owns a TX x TY block of nodes. It loads the block plus a halo
// of p nodes (the reach of a Q_p row) for all nine populations once, does the
// nodal prologue, and stages the results in LDS: g* (Bardow), or f and
// h = a f + b feq (Lee-Lin). The operator of each population is then applied
// at the owned nodes straight from LDS, as Kronecker products of the assembled
// 1-D bands: each thread owns RO consecutive nodes of one column, sums the
// y bands over a strip of RO + 2p rows of each of the 2p + 1 neighbouring
// columns (shared between its rows), and contracts with its column's x bands.
// The population loop has no barrier. Every output node is computed by exactly
// one thread from the same operation sequence: deterministic, translation
// invariant, no scatter, no atomics; the periodic seam is a wrap of the halo.
#include "device.hpp"
#include "kernels.hpp"

namespace lbfem_amd
{
  namespace
  {
    constexpr int kThreads = 256;

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

  // T: the storage and arithmetic type of the populations: double (the FP64
  // form) or float (the mixed form, deviations from the rest state; moment sums
  // in FP64). bands: 0 M, 1 D, 2 K, 3 D^T
  template <typename T, int p, int TX, int TY, int MODE>
  __global__ void __launch_bounds__(kThreads)
  k_tile(const Geo G, const TileParams S, const T *__restrict__ fin, const T *__restrict__ feq_in,
         T *__restrict__ out, T *__restrict__ rhs_out)
  {
    constexpr int  W = 2 * p + 1, HX = TX + 2 * p, HY = TY + 2 * p, NH = HX * HY;
    constexpr int  TR = kThreads / TX; // thread rows
    constexpr int  RO = TY / TR;       // owned nodes per thread, consecutive rows of one column
    static_assert(TY % TR == 0 && kThreads % TX == 0, "tile shape");
    constexpr bool LL   = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
    constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);
    constexpr int  NF   = LL ? Q + NM : Q; // f (or g*) and, for Lee-Lin, h of the moving populations

    __shared__ T F[NF][HY][HX];

    // XCD-aware tile order: consecutive tiles of a row band go to one XCD (one L2)
    const int nt   = S.tiles_x * S.tiles_y;
    const int nb8  = (nt + 7) & ~7;
    const int bid  = blockIdx.x;
    const int tile = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    if (tile >= nt)
      return;
    const int i0 = (tile % S.tiles_x) * TX, j0 = (tile / S.tiles_x) * TY;
    const int P = G.P, tid = threadIdx.x;
    const T a_c = T(S.a), b_c = T(S.b), c_c = T(S.c), hdt2 = T(0.5 * S.dt * S.dt);

    // ---------------------------------------------------------- prologue
    // all loads of the halo'd block first (R rounds in flight), then the nodal work
    constexpr int R = (NH + kThreads - 1) / kThreads;
    T             v[R][Q];
    int           gis[R], gjs[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      {
        const int k  = tid + r * kThreads;
        const int kx = k % HX, ky = k / HX;
        gis[r]       = k < NH ? wrap(i0 - p + kx, P, G.ax.periodic) : -1;
        gjs[r]       = k < NH ? wrap(j0 - p + ky, P, G.ay.periodic) : -1;
        if (gis[r] >= 0 && gjs[r] >= 0)
          {
            const int64_t idx = node_index(G, gis[r], gjs[r]);
#pragma unroll
            for (int a = 0; a < Q; ++a)
              v[r][a] = fin[a * G.stride + idx];
          }
        else
#pragma unroll
          for (int a = 0; a < Q; ++a)
            v[r][a] = 0.;
      }
#pragma unroll
    for (int r = 0; r < R; ++r)
      {
        const int k = tid + r * kThreads;
        if (k >= NH)
          continue;
        const int  kx = k % HX, ky = k / HX;
        const bool ok = gis[r] >= 0 && gjs[r] >= 0;
        double     wx = 0., wy = 0.;
        const bool wall = ok && wall_of(G, gis[r], gjs[r], wx, wy);
        if constexpr (LL)
          {
            T fe[Q];
            if (!ok)
#pragma unroll
              for (int a = 0; a < Q; ++a)
                fe[a] = T(0);
            else if (feq_in == nullptr)
              equilibrium_node(v[r], fe, wall, wx, wy);
            else
#pragma unroll
              for (int a = 0; a < Q; ++a)
                fe[a] = feq_in[a * G.stride + node_index(G, gis[r], gjs[r])];
#pragma unroll
            for (int a = 1; a < Q; ++a)
              F[Q + a - 1][ky][kx] = a_c * v[r][a] + b_c * fe[a];
          }
        else if (S.collide && ok)
          collide_node(v[r], S.omega, wall, wx, wy);
#pragma unroll
        for (int a = 0; a < Q; ++a)
          F[a][ky][kx] = v[r][a];
      }
    // this thread's column and rows; the x bands of the column and the y bands of
    // the rows in registers (bands: 0 M, 1 D, 2 K, 3 D^T)
    const int ii = tid % TX, jr = (tid / TX) * RO;
    const int gi = i0 + ii;
    T         xc[4][W], yc[RO][4][W];
#pragma unroll
    for (int o = 0; o < W; ++o)
      {
        const bool in = gi <= P;
        xc[0][o]      = in ? T(band<p>(G.ax.M, gi, o - p)) : T(0);
        xc[1][o]      = in ? T(band<p>(G.ax.D, gi, o - p)) : T(0);
        xc[2][o]      = in ? T(band<p>(G.ax.K, gi, o - p)) : T(0);
        xc[3][o]      = in ? T(band<p>(G.ax.DT, gi, o - p)) : T(0);
#pragma unroll
        for (int r = 0; r < RO; ++r)
          {
            const int  gj = j0 + jr + r;
            const bool iy = gj <= P;
            yc[r][0][o]   = iy ? T(band<p>(G.ay.M, gj, o - p)) : T(0);
            yc[r][1][o]   = iy ? T(band<p>(G.ay.D, gj, o - p)) : T(0);
            yc[r][2][o]   = iy ? T(band<p>(G.ay.K, gj, o - p)) : T(0);
            yc[r][3][o]   = iy ? T(band<p>(G.ay.DT, gj, o - p)) : T(0);
          }
      }
    __syncthreads();

    T xinc[RO][NM], mflux[RO];
#pragma unroll
    for (int r = 0; r < RO; ++r)
      {
        mflux[r] = 0.;
#pragma unroll
        for (int a = 0; a < NM; ++a)
          xinc[r][a] = 0.;
      }

    // ---------------------------------------------------------- populations 1..8
#pragma unroll
    for (int al = 1; al < Q; ++al)
      {
        // directions as uniform run-time values: compile-time D2Q9 constants let the
        // scheduler interleave the barrier-free populations and spill registers
        const T ex = T(S.e[al][0]), ey = T(S.e[al][1]);
        const bool nx = ex != T(0), ny = ey != T(0);
        if (!nx && !ny)
          continue; // at rest in this sweep: r = 0 (streaming.h:118-122)
        const T(*Ff)[HX] = F[al];
        const T(*Fh)[HX] = F[LL ? Q + al - 1 : al];
        T acc[RO];
#pragma unroll
        for (int r = 0; r < RO; ++r)
          acc[r] = 0.;
        // r = Dx (ex My h + c ex ey Dy^T f) + Kx (c ex^2 My f) + Mx (ey Dy h + c ey^2 Ky f) + Dx^T (c ex ey Dy f),
        // h = f for Bardow (its a folded into the x weights); one branch-free body per direction type
        const T sa = LL ? T(1) : a_c;
        // the mixed form applies the operator to f - f(node): every term annihilates
        // constants (its bands' row or column sums vanish), and FP32 then carries the
        // small differences instead of cancelling large values
        constexpr bool CEN = sizeof(T) == 4;
        T              cf[RO], ch[RO];
#pragma unroll
        for (int r = 0; r < RO; ++r)
          {
            cf[r] = CEN ? Ff[jr + r + p][ii + p] : T(0);
            ch[r] = (CEN && LL) ? Fh[jr + r + p][ii + p] : T(0);
          }
        if (nx && !ny)
          {
            const T wd = ex * sa, wk = c_c * ex * ex;
#pragma unroll
            for (int c = 0; c < W; ++c)
              {
                const T kd = xc[1][c] * wd, kk = xc[2][c] * wk;
#pragma unroll
                for (int r = 0; r < RO; ++r)
                  {
                    T My = 0, Myh = 0;
#pragma unroll
                    for (int o = 0; o < W; ++o)
                      {
                        My += yc[r][0][o] * (Ff[jr + r + o][ii + c] - cf[r]);
                        if constexpr (LL)
                          Myh += yc[r][0][o] * (Fh[jr + r + o][ii + c] - ch[r]);
                      }
                    if constexpr (!LL)
                      Myh = My;
                    acc[r] += kd * Myh + kk * My;
                  }
              }
          }
        else if (!nx && ny)
          {
            const T wd = ey * sa, wk = c_c * ey * ey;
#pragma unroll
            for (int c = 0; c < W; ++c)
              {
#pragma unroll
                for (int r = 0; r < RO; ++r)
                  {
                    T s = 0;
#pragma unroll
                    for (int o = 0; o < W; ++o)
                      {
                        const T fv = Ff[jr + r + o][ii + c] - cf[r];
                        const T hv = LL ? Fh[jr + r + o][ii + c] - ch[r] : fv;
                        s += (yc[r][1][o] * wd) * hv + (yc[r][2][o] * wk) * fv;
                      }
                    acc[r] += xc[0][c] * s;
                  }
              }
          }
        else
          {
            const T wdx = ex * sa, wdy = ey * sa, wk = c_c * ex * ex, wky = c_c * ey * ey, wx = c_c * ex * ey;
#pragma unroll
            for (int c = 0; c < W; ++c)
              {
#pragma unroll
                for (int r = 0; r < RO; ++r)
                  {
                    T My = 0, Dy = 0, Ky = 0, DTy = 0, Myh = 0, Dyh = 0;
#pragma unroll
                    for (int o = 0; o < W; ++o)
                      {
                        const T fv = Ff[jr + r + o][ii + c] - cf[r];
                        My += yc[r][0][o] * fv;
                        Dy += yc[r][1][o] * fv;
                        Ky += yc[r][2][o] * fv;
                        DTy += yc[r][3][o] * fv;
                        if constexpr (LL)
                          {
                            const T hv = Fh[jr + r + o][ii + c] - ch[r];
                            Myh += yc[r][0][o] * hv;
                            Dyh += yc[r][1][o] * hv;
                          }
                      }
                    if constexpr (!LL)
                      {
                        Myh = My;
                        Dyh = Dy;
                      }
                    acc[r] += xc[1][c] * (wdx * Myh + wx * DTy) + xc[2][c] * (wk * My) +
                              xc[0][c] * (wdy * Dyh + wky * Ky) + xc[3][c] * (wx * Dy);
                  }
              }
          }
        // surface term (dt^2/2) oint phi (n.e)(e.grad f) on walls (advection.h:177-199), as exact
        // 1-D integrals along the wall: Dx of the tangential derivative, Mx of the normal one
#pragma unroll
        for (int r = 0; r < RO; ++r)
          {
            const int jj = jr + r, gj = j0 + jj;
            const int ty = jj + p, tx = ii + p;
            T         face = 0;
            if (gi <= P && gj <= P)
              {
                if (!G.ay.periodic && ny && (gj == 0 || gj == P))
                  {
                    const bool    lo = gj == 0;
                    const double *bn = lo ? G.ay.b0 : G.ay.b1;
                    const int     r0 = lo ? ty : ty - p;
                    T             t = 0, nrm = 0;
#pragma unroll
                    for (int o = 0; o < W; ++o)
                      {
                        T dn = 0;
#pragma unroll
                        for (int l = 0; l <= p; ++l)
                          dn += T(bn[l]) * Ff[r0 + l][ii + o];
                        t += xc[1][o] * Ff[ty][ii + o];
                        nrm += xc[0][o] * dn;
                      }
                    face += hdt2 * (lo ? -ey : ey) * (ex * t + ey * nrm);
                  }
                if (!G.ax.periodic && nx && (gi == 0 || gi == P))
                  {
                    const bool    lo = gi == 0;
                    const double *bn = lo ? G.ax.b0 : G.ax.b1;
                    const int     c0 = lo ? tx : tx - p;
                    T             t = 0, nrm = 0;
#pragma unroll
                    for (int o = 0; o < W; ++o)
                      {
                        T dn = 0;
#pragma unroll
                        for (int l = 0; l <= p; ++l)
                          dn += T(bn[l]) * Ff[jj + o][c0 + l];
                        t += yc[r][1][o] * Ff[jj + o][tx];
                        nrm += yc[r][0][o] * dn;
                      }
                    face += hdt2 * (lo ? -ex : ex) * (ex * nrm + ey * t);
                  }
              }
            mflux[r] += face;
            xinc[r][al - 1] = acc[r] + face;
          }
      }

    // ---------------------------------------------------------- epilogue
#pragma unroll
    for (int r = 0; r < RO; ++r)
      {
        const int jj = jr + r, gj = j0 + jj;
        if (gi > P || gj > P)
          continue;
        const int64_t idx   = node_index(G, gi, gj);
        const bool    slave = (G.ax.periodic && gi == P) || (G.ay.periodic && gj == P);
        T             own[Q];
        if (!slave)
#pragma unroll
          for (int a = 0; a < Q; ++a)
            own[a] = F[a][jj + p][ii + p];
        else
          {
            // a slave's own entries (the staged tile holds its master's); only nodal work applies
#pragma unroll
            for (int a = 0; a < Q; ++a)
              own[a] = fin[a * G.stride + idx];
            if constexpr (!LL)
              if (S.collide)
                {
                  double     wx = 0., wy = 0.;
                  const bool wall = wall_of(G, gi, gj, wx, wy);
                  collide_node(own, S.omega, wall, wx, wy);
                }
          }
        // mass flux of the wall term removed isotropically (advection.h:194-196)
        if (S.remove_mass_flux)
#pragma unroll
          for (int a = 1; a < Q; ++a)
            xinc[r][a - 1] -= T(w_of(a) / (1. - w_of(0))) * mflux[r];
        if constexpr (STEP)
          {
            const T mli = slave ? T(0) : T(G.ax.ml_inv[gi] * G.ay.ml_inv[gj]);
#pragma unroll
            for (int a = 0; a < NM; ++a)
              xinc[r][a] *= mli;
            if constexpr (LL)
              {
                double     wx = 0., wy = 0.;
                const bool wall = wall_of(G, gi, gj, wx, wy);
                T          fe[Q];
                equilibrium_node(own, fe, wall, wx, wy);
                predictor_corrector_node(own, fe, xinc[r], S.theta, wall, wx, wy);
              }
            else
#pragma unroll
              for (int a = 1; a < Q; ++a)
                own[a] += xinc[r][a - 1];
#pragma unroll
            for (int a = 0; a < Q; ++a)
              out[a * G.stride + idx] = own[a];
          }
        else
          {
            // right-hand side (constrained entries get nothing) and, fused, g*
#pragma unroll
            for (int a = 0; a < NM; ++a)
              rhs_out[a * G.stride + idx] = slave ? T(0) : xinc[r][a];
            if (out != nullptr)
#pragma unroll
              for (int a = 0; a < Q; ++a)
                out[a * G.stride + idx] = own[a];
          }
      }
  }

  // ------------------------------------------------------------------ launch
  template <typename T, int p, int TX, int TY, int MODE>
  static hipError_t
  launch_tile(const Geo &G, TileParams S, const T *fin, const T *feq, T *out, T *rhs, hipStream_t stream)
  {
    S.tiles_x     = (G.P + 1 + TX - 1) / TX;
    S.tiles_y     = (G.P + 1 + TY - 1) / TY;
    const int nt  = S.tiles_x * S.tiles_y;
    const int nb8 = (nt + 7) & ~7;
    hipLaunchKernelGGL((k_tile<T, p, TX, TY, MODE>), dim3(nb8), dim3(kThreads), 0, stream, G, S, fin, feq, out, rhs);
    return hipGetLastError();
  }

  // tile heights: 16 rows where the staged fields fit in about 55 KB of LDS, else 8
  template <typename T, int MODE>
  static hipError_t
  dispatch_degree(const Geo &G, const TileParams &S, const T *fin, const T *feq, T *out, T *rhs, hipStream_t stream)
  {
    constexpr bool LL   = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
    constexpr bool F32  = sizeof(T) == 4;
    constexpr int  TY12 = (LL && !F32) ? 8 : 16;
    switch (G.p)
      {
        case 1:
          return launch_tile<T, 1, 32, TY12, MODE>(G, S, fin, feq, out, rhs, stream);
        case 2:
          return launch_tile<T, 2, 32, TY12, MODE>(G, S, fin, feq, out, rhs, stream);
        case 3:
          return launch_tile<T, 3, 32, 8, MODE>(G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  template <typename T>
  static hipError_t
  tile_dispatch(const int mode, const Geo &G, const TileParams &S, const T *fin, const T *feq, T *out, T *rhs,
                hipStream_t stream)
  {
    switch (mode)
      {
        case kStepBardow:
          return dispatch_degree<T, kStepBardow>(G, S, fin, feq, out, rhs, stream);
        case kStepLeeLin:
          return dispatch_degree<T, kStepLeeLin>(G, S, fin, feq, out, rhs, stream);
        case kAdvectBardow:
          return dispatch_degree<T, kAdvectBardow>(G, S, fin, feq, out, rhs, stream);
        case kAdvectLeeLin:
          return dispatch_degree<T, kAdvectLeeLin>(G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  hipError_t
  tile_launch(const int mode, const Geo &G, const TileParams &S, const double *fin, const double *feq, double *out,
              double *rhs, hipStream_t stream)
  {
    return tile_dispatch<double>(mode, G, S, fin, feq, out, rhs, stream);
  }

  hipError_t
  tile_launch(const int mode, const Geo &G, const TileParams &S, const float *fin, const float *feq, float *out,
              float *rhs, hipStream_t stream)
  {
    return tile_dispatch<float>(mode, G, S, fin, feq, out, rhs, stream);
  }
} // namespace lbfem_amd
