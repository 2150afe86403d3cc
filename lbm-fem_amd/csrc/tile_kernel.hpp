// The tile kernel of the streaming step: the weak-form advection right-hand side
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
// 1-D bands: a thread owns RO consecutive nodes of one column, sums the y bands
// over a strip of RO + 2p rows of each of the 2p + 1 neighbouring columns
// (shared between its rows), and contracts with its column's x bands.
//
// With PS > 1 the eight populations of a node are split over PS threads (PS
// wavefronts of the block hold the same nodes, each wavefront a fixed set of
// whole population pairs): the increments meet in LDS and the first part does
// the nodal epilogue. Small meshes then fill the device with PS times as many
// wavefronts, each with a PS times shorter chain of dependent operations.
//
// Every variant computes every output value by the same operation sequence
// (the wall mass flux is summed by pairs of populations, then over the pairs,
// in every variant), so that all variants give bitwise identical results:
// deterministic, translation invariant, no scatter, no atomics; the periodic
// seam is a wrap of the halo.
#pragma once

#include "device.hpp"
#include "kernels.hpp"

namespace lbfem_amd
{
  namespace tile
  {
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

    // A tile variant: T the storage and arithmetic type of the populations
    // (double for the FP64 form, float for the mixed form's deviations), Q_p,
    // tile TX x TY, RO owned rows per thread, PS threads per node.
    template <typename T_, int p_, int TX_, int TY_, int RO_, int PS_, int MODE_>
    struct Cfg
    {
      using T                    = T_;
      static constexpr int  p    = p_, TX = TX_, TY = TY_, RO = RO_, PS = PS_, MODE = MODE_;
      static constexpr int  W    = 2 * p + 1, HX = TX + 2 * p, HY = TY + 2 * p, NH = HX * HY;
      static constexpr int  TPP  = TX * TY / RO; // threads per population part
      static constexpr int  NT   = TPP * PS;     // threads of the block
      static constexpr int  PPP  = NM / PS;      // populations per part
      static constexpr bool LL   = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
      static constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);
      static constexpr int  NF   = LL ? Q + NM : Q; // f (or g*) and, for Lee-Lin, h of the moving populations
      static constexpr int  R    = (NH + NT - 1) / NT; // prologue rounds
      static_assert(TY % RO == 0 && TPP % TX == 0, "tile shape");
      static_assert(TPP % 64 == 0 && NT <= 1024, "whole wavefronts, at most 1024 threads");
      static_assert(PS == 1 || RO == 1, "population split with one row per thread");
      static_assert(NM % PS == 0 && PPP % 2 == 0, "parts of whole population pairs");
    };

    template <class C>
    struct Smem
    {
      typename C::T F[C::NF][C::HY][C::HX];
      typename C::T X[C::PS > 1 ? NM : 1][C::TY][C::TX];      // increments of the parts
      typename C::T MP[C::PS > 1 ? NM / 2 : 1][C::TY][C::TX]; // wall mass flux of population pairs
    };

    // bands: 0 M, 1 D, 2 K, 3 D^T
    template <class C>
    __device__ __forceinline__ void
    body(const Geo &G, const TileParams &S, const int tile, const typename C::T *__restrict__ fin,
         const typename C::T *__restrict__ feq_in, typename C::T *__restrict__ out,
         typename C::T *__restrict__ rhs_out, Smem<C> &sm)
    {
      using T                   = typename C::T;
      constexpr int  p          = C::p, TX = C::TX, TY = C::TY, RO = C::RO, PS = C::PS;
      constexpr int  W          = C::W, HX = C::HX, HY = C::HY, NH = C::NH, NT = C::NT, R = C::R, NF = C::NF;
      constexpr bool LL         = C::LL, STEP = C::STEP;
      T(&F)[NF][HY][HX]         = sm.F;
      const int i0 = (tile % S.tiles_x) * TX, j0 = (tile / S.tiles_x) * TY;
      const int P = G.P, tid = threadIdx.x;
      const T   a_c = T(S.a), b_c = T(S.b), c_c = T(S.c), hdt2 = T(0.5 * S.dt * S.dt);

      // ---------------------------------------------------------- prologue
      // all loads of the halo'd block first (R rounds in flight), then the nodal work
      T             v[R][Q];
      int           gis[R], gjs[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
        {
          const int k  = tid + r * NT;
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
          const int k = tid + r * NT;
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
      // population part of this thread, uniform over its wavefront
      const int q  = PS > 1 ? __builtin_amdgcn_readfirstlane(tid / C::TPP) : 0;
      const int lt = PS > 1 ? tid % C::TPP : tid;
      const int ii = lt % TX, jr = (lt / TX) * RO;
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

      // ---------------------------------------------------------- populations 1..8
      // res: the operator's value at the owned rows, fc: the wall surface term
      auto pop = [&](const int al, T(&res)[RO], T(&fc)[RO]) {
        // directions as uniform run-time values: compile-time D2Q9 constants let the
        // scheduler interleave the barrier-free populations and spill registers
        const T ex = T(S.e[al][0]), ey = T(S.e[al][1]);
        const bool nx = ex != T(0), ny = ey != T(0);
        if (!nx && !ny)
          {
            // at rest in this sweep: r = 0 (streaming.h:118-122)
#pragma unroll
            for (int r = 0; r < RO; ++r)
              res[r] = fc[r] = T(0);
            return;
          }
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
            res[r] = acc[r];
            fc[r]  = face;
          }
      };

      T xinc[RO][NM], mp[RO][NM / 2];
      if constexpr (PS == 1)
        {
#pragma unroll
          for (int al = 1; al < Q; ++al)
            {
              T res[RO], fc[RO];
              pop(al, res, fc);
#pragma unroll
              for (int r = 0; r < RO; ++r)
                {
                  xinc[r][al - 1] = res[r] + fc[r];
                  if (al & 1)
                    mp[r][(al - 1) >> 1] = fc[r];
                  else
                    mp[r][(al - 1) >> 1] += fc[r];
                }
            }
        }
      else
        {
          T pair = T(0);
#pragma unroll
          for (int k = 0; k < C::PPP; ++k)
            {
              const int al = q * C::PPP + k + 1; // uniform over the wavefront
              T         res[1], fc[1];
              pop(al, res, fc);
              sm.X[al - 1][jr][ii] = res[0] + fc[0];
              if ((k & 1) == 0)
                pair = fc[0];
              else
                {
                  pair += fc[0];
                  sm.MP[(al - 1) >> 1][jr][ii] = pair;
                }
            }
          __syncthreads();
          if (q != 0)
            return;
#pragma unroll
          for (int a = 0; a < NM; ++a)
            xinc[0][a] = sm.X[a][jr][ii];
#pragma unroll
          for (int k = 0; k < NM / 2; ++k)
            mp[0][k] = sm.MP[k][jr][ii];
        }
      T mflux[RO];
#pragma unroll
      for (int r = 0; r < RO; ++r)
        mflux[r] = ((mp[r][0] + mp[r][1]) + mp[r][2]) + mp[r][3];

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
  } // namespace tile

  template <class C>
  __global__ void __launch_bounds__(C::NT)
  k_tile(const Geo G, const TileParams S, const typename C::T *__restrict__ fin,
         const typename C::T *__restrict__ feq_in, typename C::T *__restrict__ out,
         typename C::T *__restrict__ rhs_out)
  {
    __shared__ tile::Smem<C> sm;
    // XCD-aware tile order: consecutive tiles of a row band go to one XCD (one L2)
    const int nt   = S.tiles_x * S.tiles_y;
    const int nb8  = (nt + 7) & ~7;
    const int bid  = blockIdx.x;
    const int tile = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    if (tile >= nt)
      return;
    tile::body<C>(G, S, tile, fin, feq_in, out, rhs_out, sm);
  }

  template <class C>
  hipError_t
  launch_tile(const Geo &G, TileParams S, const typename C::T *fin, const typename C::T *feq, typename C::T *out,
              typename C::T *rhs, hipStream_t stream)
  {
    S.tiles_x     = (G.P + 1 + C::TX - 1) / C::TX;
    S.tiles_y     = (G.P + 1 + C::TY - 1) / C::TY;
    const int nt  = S.tiles_x * S.tiles_y;
    const int nb8 = (nt + 7) & ~7;
    hipLaunchKernelGGL((k_tile<C>), dim3(nb8), dim3(C::NT), 0, stream, G, S, fin, feq, out, rhs);
    return hipGetLastError();
  }

  // The variants (TileVariant): kTileWide 32 x TY tiles of 256 threads, TY = 16 where
  // the staged fields fit in about 55 KB of LDS, else 8 (two or one rows a thread);
  // kTileSplit2 / kTileSplit4: 32 x 8 tiles, one node per thread, 2 or 4 threads
  // per node; kTileSmall4: 16 x 8 tiles, 4 threads per node.
  template <typename T, int p, int MODE>
  hipError_t
  launch_variant(const int variant, const Geo &G, const TileParams &S, const T *fin, const T *feq, T *out, T *rhs,
                 hipStream_t stream)
  {
    constexpr bool LL  = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int  TYW = (p == 3 || (LL && !F32)) ? 8 : 16;
    switch (variant)
      {
        case kTileWide:
          return launch_tile<tile::Cfg<T, p, 32, TYW, TYW / 8, 1, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileSplit2:
          return launch_tile<tile::Cfg<T, p, 32, 8, 1, 2, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileSplit4:
          return launch_tile<tile::Cfg<T, p, 32, 8, 1, 4, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileSmall4:
          return launch_tile<tile::Cfg<T, p, 16, 8, 1, 4, MODE>>(G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  // every mode (double) or the fused-step modes (float, the mixed form) of one degree
  template <typename T, int p>
  hipError_t
  tile_launch_degree(const int mode, const int variant, const Geo &G, const TileParams &S, const T *fin,
                     const T *feq, T *out, T *rhs, hipStream_t stream)
  {
    switch (mode)
      {
        case kStepBardow:
          return launch_variant<T, p, kStepBardow>(variant, G, S, fin, feq, out, rhs, stream);
        case kStepLeeLin:
          return launch_variant<T, p, kStepLeeLin>(variant, G, S, fin, feq, out, rhs, stream);
        case kAdvectBardow:
          if constexpr (sizeof(T) == 8)
            return launch_variant<T, p, kAdvectBardow>(variant, G, S, fin, feq, out, rhs, stream);
          else
            return hipErrorInvalidValue;
        case kAdvectLeeLin:
          if constexpr (sizeof(T) == 8)
            return launch_variant<T, p, kAdvectLeeLin>(variant, G, S, fin, feq, out, rhs, stream);
          else
            return hipErrorInvalidValue;
        default:
          return hipErrorInvalidValue;
      }
  }
} // namespace lbfem_amd
