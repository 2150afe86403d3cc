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
// at the owned nodes straight from LDS. Where the row-march kernel also runs (Q1,
// and Q2 in FP32) a thread evaluates the column form of column.hpp, the same
// operation sequence as that kernel: at each of the 2p + 1 neighbouring columns
// the y contractions grouped by the x band they meet, then the x bands; in FP64
// on Q2 and on Q3 it contracts the y bands over the neighbouring columns on the
// node-centred values and then the x bands of its column.
//
// With PS > 1 the eight populations of a node are split over PS threads (PS
// wavefronts of the block hold the same nodes, each wavefront a fixed set of
// whole population pairs): the increments meet in LDS and the first part does
// the nodal epilogue. Small meshes then fill the device with PS times as many
// wavefronts, each with a PS times shorter chain of dependent operations.
//
// The prologue loads a thread's own node first, so that the Lee-Lin epilogue
// reuses that node's equilibria. Every variant computes every output value by
// the same operation sequence (the wall mass flux is summed by pairs of
// populations, then over the pairs), so that all variants give bitwise
// identical results: deterministic, translation invariant, no scatter, no
// atomics; the periodic seam is a wrap of the halo.
#pragma once

#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"
#include "march_kernel.hpp"
#include "column.hpp"

namespace lbfem_amd
{
  namespace tile
  {
    // Band coefficient o in [-p, p] of row i.
    template <int p>
    __device__ __forceinline__ double
    band(const double *B, const int i, const int o)
    {
      // device (global) memory, also where the pointer comes from a batch table
      return ((const __attribute__((address_space(1))) double *)B)[i * (2 * p + 1) + o + p];
    }

    __device__ __forceinline__ int
    wrap(const int i, const int P, const int periodic)
    {
      return wrap_line(i, P, periodic);
    }

    // A tile variant: T the storage and arithmetic type of the populations
    // (double for the FP64 form, float for the mixed form's deviations), Q_p,
    // tile TX x TY (one node per thread and part), PS threads per node.
    template <typename T_, int p_, int TX_, int TY_, int PS_, int MODE_>
    struct Cfg
    {
      using T                    = T_;
      static constexpr int  p    = p_, TX = TX_, TY = TY_, PS = PS_, MODE = MODE_;
      static constexpr int  W    = 2 * p + 1, HX = TX + 2 * p, HY = TY + 2 * p, NH = HX * HY;
      static constexpr int  TPP  = TX * TY; // threads per population part
      static constexpr int  NT   = TPP * PS; // threads of the block
      static constexpr int  PPP  = NM / PS;  // populations per part
      static constexpr bool LL   = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
      static constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);
      static constexpr int  NF   = LL ? Q + NM : Q;      // f (or g*) and, for Lee-Lin, h of the moving populations
      static constexpr int  R    = (NH + NT - 1) / NT;   // prologue rounds
      static constexpr bool WGT  = p == 1;               // Q1: the face code reads the bands again (registers)
      // LDS row pitch of the staged planes. The LDS serves 128 bytes a cycle (LB lanes of T);
      // where a row of TX threads does not end on a multiple of LB lanes, a group of LB lanes
      // spans two rows and hits the same banks twice unless the pitch is congruent to TX
      // modulo LB (then the two rows' words follow on in the banks). Padding is skipped where
      // it would more than double the plane (16-wide FP32 tiles).
      static constexpr int  LB   = 128 / int(sizeof(T));
      static constexpr int  PXP  = HX + ((TX - HX) % LB + LB) % LB;
      static constexpr int  PX   = (TX % LB == 0 || PXP > 2 * HX) ? HX : PXP;
      static_assert(TPP % 64 == 0 && NT <= 1024, "whole wavefronts, at most 1024 threads");
      static_assert(NM % PS == 0 && PPP % 2 == 0, "parts of whole population pairs");
    };

    template <class C>
    struct Smem
    {
      typename C::T F[C::NF][C::HY][C::PX];
      typename C::T X[C::PS > 1 ? NM : 1][C::TY][C::TX];      // increments of the parts
      typename C::T MP[C::PS > 1 ? NM / 2 : 1][C::TY][C::TX]; // wall mass flux of population pairs
    };

    // Halo-tile position of prologue slot k: the TX x TY owned nodes first (slot
    // k is the node of thread k of part 0), then the ring of width p.
    template <class C>
    __device__ __forceinline__ void
    slot_position(const int k, int &kx, int &ky)
    {
      constexpr int p = C::p, TX = C::TX, TY = C::TY, HX = C::HX, HY = C::HY;
      if (k < TX * TY)
        {
          kx = p + k % TX;
          ky = p + k / TX;
          return;
        }
      const int m = k - TX * TY;
      if (m < p * HX)
        {
          kx = m % HX;
          ky = m / HX;
        }
      else if (m < 2 * p * HX)
        {
          kx = (m - p * HX) % HX;
          ky = HY - p + (m - p * HX) / HX;
        }
      else
        {
          const int m3 = m - 2 * p * HX, c = m3 % (2 * p);
          kx           = c < p ? c : TX + c;
          ky           = p + m3 / (2 * p);
        }
    }

    // bands: 0 M, 1 D, 2 K, 3 D^T
    template <class C>
    __device__ __forceinline__ void
    body(const Geo &G, const TileParams &S, const int tile, const typename C::T *__restrict__ fin,
         const typename C::T *__restrict__ feq_in, typename C::T *__restrict__ out,
         typename C::T *__restrict__ rhs_out, Smem<C> &sm)
    {
      using T                   = typename C::T;
      constexpr int  p          = C::p, TX = C::TX, PS = C::PS;
      constexpr int  W          = C::W, HX = C::HX, HY = C::HY, NH = C::NH, NT = C::NT, R = C::R, NF = C::NF;
      constexpr bool LL         = C::LL, STEP = C::STEP;
      // the mixed form applies the operator to f - f(node): every term annihilates
      // constants (its bands' row or column sums vanish), and FP32 then carries the
      // small differences instead of cancelling large values
      constexpr bool CEN        = sizeof(T) == 4;
      // the column form wherever the row-march kernel may run for this form (the same results bit
      // for bit under every variant)
      constexpr bool COL        = march_runs(p, sizeof(T) == 4);
      constexpr int  PX         = C::PX;
      T(&F)[NF][HY][PX]         = sm.F;
      // the population buffers are device (global) memory: say so, so that the loads and
      // stores are global ones even where the pointers come from a batch table
      using GT                   = __attribute__((address_space(1))) T;
      const GT *const gfin       = (const GT *)fin;
      const GT *const gfeq       = (const GT *)feq_in;
      GT *const       gout       = (GT *)out;
      GT *const       grhs       = (GT *)rhs_out;
      const int i0 = (tile % S.tiles_x) * TX, j0 = (tile / S.tiles_x) * C::TY;
      const int P = G.P, tid = threadIdx.x;
      const T   a_c = T(S.a), b_c = T(S.b), c_c = T(S.c), hdt2 = T(0.5 * S.dt * S.dt);
      const T   sa = LL ? T(1) : a_c; // Bardow's a in the first-order weights, Lee-Lin's in h

      // ---------------------------------------------------------- prologue
      // all loads of the halo'd block first (R rounds in flight), then the nodal work
      T   v[R][Q];
      int gis[R], gjs[R], kxs[R], kys[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
        {
          const int k = tid + r * NT;
          slot_position<C>(k, kxs[r], kys[r]);
          gis[r] = k < NH ? wrap(i0 - p + kxs[r], P, G.ax.periodic) : -1;
          gjs[r] = k < NH ? wrap(j0 - p + kys[r], P, G.ay.periodic) : -1;
          if (gis[r] >= 0 && gjs[r] >= 0)
            {
              const int64_t idx = node_index(G, gis[r], gjs[r]);
#pragma unroll
              for (int a = 0; a < Q; ++a)
                v[r][a] = gfin[a * G.stride + idx];
            }
          else
#pragma unroll
            for (int a = 0; a < Q; ++a)
              v[r][a] = T(0);
        }
      T fe_own[Q]; // Lee-Lin: the equilibria of slot tid, the own node of part 0 threads
#pragma unroll
      for (int r = 0; r < R; ++r)
        {
          const int k = tid + r * NT;
          if (k >= NH)
            continue;
          const int  kx = kxs[r], ky = kys[r];
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
                  fe[a] = gfeq[a * G.stride + node_index(G, gis[r], gjs[r])];
#pragma unroll
              for (int a = 1; a < Q; ++a)
                F[Q + a - 1][ky][kx] = col::fma_t(a_c, v[r][a], b_c * fe[a]);
              if (r == 0)
#pragma unroll
                for (int a = 0; a < Q; ++a)
                  fe_own[a] = fe[a];
            }
          else if (S.collide && ok)
            collide_node(v[r], S.omega, wall, wx, wy);
#pragma unroll
          for (int a = 0; a < Q; ++a)
            F[a][ky][kx] = v[r][a];
        }
      // this thread's node; the x bands of its column and the y bands of its row
      // (bands: 0 M, 1 D, 2 K, 3 D^T)
      // population part of this thread, uniform over its wavefront
      const int q  = PS > 1 ? __builtin_amdgcn_readfirstlane(tid / C::TPP) : 0;
      const int lt = PS > 1 ? tid % C::TPP : tid;
      const int ii = lt % TX, jj = lt / TX;
      const int gi = i0 + ii, gj = j0 + jj;
      T         xc[4][W], yc[4][W];
#pragma unroll
      for (int o = 0; o < W; ++o)
        {
          const bool in = gi <= P, iy = gj <= P;
          xc[0][o]      = in ? T(band<p>(G.ax.M, gi, o - p)) : T(0);
          xc[1][o]      = in ? T(band<p>(G.ax.D, gi, o - p)) : T(0);
          xc[2][o]      = in ? T(band<p>(G.ax.K, gi, o - p)) : T(0);
          xc[3][o]      = in ? T(band<p>(G.ax.DT, gi, o - p)) : T(0);
          yc[0][o]      = iy ? T(band<p>(G.ay.M, gj, o - p)) : T(0);
          yc[1][o]      = iy ? T(band<p>(G.ay.D, gj, o - p)) : T(0);
          yc[2][o]      = iy ? T(band<p>(G.ay.K, gj, o - p)) : T(0);
          yc[3][o]      = iy ? T(band<p>(G.ay.DT, gj, o - p)) : T(0);
        }
      __syncthreads();

      // ---------------------------------------------------------- populations 1..8
      // r = ex sa Dx My h + ey sa Mx Dy h + c (ex^2 Kx My + ey^2 Mx Ky + ex ey (Dx Dy^T + Dx^T Dy)) f,
      // h = f for Bardow; one branch-free body per direction class. Directions are uniform
      // run-time values: compile-time D2Q9 constants let the scheduler interleave the
      // barrier-free populations and spill registers. A population at rest in the sweep
      // gets r = 0 (streaming.h:118-122).
      // the column form shared with the row-march kernel (column.hpp), centred in FP32: the whole
      // increment of population al (the forms the row-march kernel runs on: Q1, Q2 in FP32)
      auto col_pop = [&](const int al) -> T {
        const col::Coef<T> k = col::coef<T>(T(S.e[al][0]), T(S.e[al][1]), sa, c_c);
        if (!k.nx && !k.ny)
          return T(0);
        const T(*Ff)[PX] = F[al];
        const T(*Fh)[PX] = F[LL ? Q + al - 1 : al];
        // FP32: the centre values of the node's row, and the row sums of its M and D^T bands
        const T f0c = CEN ? Ff[jj + p][ii + p] : T(0);
        const T h0c = (CEN && LL) ? Fh[jj + p][ii + p] : f0c;
        const T mM = CEN ? col::row_sum(yc[0]) : T(0), mT = CEN ? col::row_sum(yc[3]) : T(0);
        T       acc = T(0);
#pragma unroll
        for (int c = 0; c < W; ++c)
          {
            T fv[W], hv[W];
#pragma unroll
            for (int o = 0; o < W; ++o)
              {
                fv[o] = Ff[jj + o][ii + c];
                hv[o] = LL ? Fh[jj + o][ii + c] : fv[o];
              }
            col::Cols<T> u;
            if constexpr (CEN)
              u = col::assemble<T, LL>(col::ccolumns<T, LL>(yc, fv, hv, k), mM, mT, hv[p] - h0c, fv[p] - f0c, c != p,
                                       k);
            else
              u = col::columns<T, LL>(yc, fv, hv, k);
            const T w = col::xcontract(xc[0][c], xc[1][c], xc[2][c], xc[3][c], u, k);
            acc       = c == 0 ? w : acc + w; // ((w_0 + w_1) + ...) + w_2p
          }
        return acc;
      };
      // the other forms (Q2 in FP64, Q3): y bands over the columns' strips, then the column's x bands,
      // centred on the node in FP32
      auto popA = [&](const int al, T &res) {
        if constexpr (COL)
          res = col_pop(al);
        else
          {
            const T    ex = T(S.e[al][0]), ey = T(S.e[al][1]);
            const bool nx = ex != T(0), ny = ey != T(0);
            res           = T(0);
            const T(*Ff)[PX] = F[al];
            const T(*Fh)[PX] = F[LL ? Q + al - 1 : al];
            const T cf       = CEN ? Ff[jj + p][ii + p] : T(0);
            const T ch       = (CEN && LL) ? Fh[jj + p][ii + p] : T(0);
              {
                if (!nx && !ny)
                  return;
                // Q2, Q3: y bands over the columns' strips, then the column's x bands
                T acc = 0;
                if (nx && !ny)
                  {
                    const T wd = ex * sa, wk = c_c * ex * ex;
    #pragma unroll
                    for (int c = 0; c < W; ++c)
                      {
                        T My = 0, Myh = 0;
    #pragma unroll
                        for (int o = 0; o < W; ++o)
                          {
                            My += yc[0][o] * (Ff[jj + o][ii + c] - cf);
                            if constexpr (LL)
                              Myh += yc[0][o] * (Fh[jj + o][ii + c] - ch);
                          }
                        if constexpr (!LL)
                          Myh = My;
                        acc += (xc[1][c] * wd) * Myh + (xc[2][c] * wk) * My;
                      }
                  }
                else if (!nx && ny)
                  {
                    const T wd = ey * sa, wk = c_c * ey * ey;
                    T       y1[W], y2[W];
    #pragma unroll
                    for (int o = 0; o < W; ++o)
                      {
                        y1[o] = yc[1][o] * wd;
                        y2[o] = yc[2][o] * wk;
                      }
    #pragma unroll
                    for (int c = 0; c < W; ++c)
                      {
                        T s = 0;
    #pragma unroll
                        for (int o = 0; o < W; ++o)
                          {
                            const T fv = Ff[jj + o][ii + c] - cf;
                            const T hv = LL ? Fh[jj + o][ii + c] - ch : fv;
                            s += y1[o] * hv + y2[o] * fv;
                          }
                        acc += xc[0][c] * s;
                      }
                  }
                else
                  {
                    const T wdx = ex * sa, wdy = ey * sa, wk = c_c * ex * ex, wky = c_c * ey * ey, wx = c_c * ex * ey;
    #pragma unroll
                    for (int c = 0; c < W; ++c)
                      {
                        T My = 0, Dy = 0, Ky = 0, DTy = 0, Myh = 0, Dyh = 0;
    #pragma unroll
                        for (int o = 0; o < W; ++o)
                          {
                            const T fv = Ff[jj + o][ii + c] - cf;
                            My += yc[0][o] * fv;
                            Dy += yc[1][o] * fv;
                            Ky += yc[2][o] * fv;
                            DTy += yc[3][o] * fv;
                            if constexpr (LL)
                              {
                                const T hv = Fh[jj + o][ii + c] - ch;
                                Myh += yc[0][o] * hv;
                                Dyh += yc[1][o] * hv;
                              }
                          }
                        if constexpr (!LL)
                          {
                            Myh = My;
                            Dyh = Dy;
                          }
                        acc += xc[1][c] * (wdx * Myh + wx * DTy) + xc[2][c] * (wk * My) +
                               xc[0][c] * (wdy * Dyh + wky * Ky) + xc[3][c] * (wx * Dy);
                      }
                  }
                res = acc;
              }
          }
      };
      // the wall surface term of population al at the owned node
      auto face = [&](const int al) -> T {
        const T    ex = T(S.e[al][0]), ey = T(S.e[al][1]);
        const bool nx = ex != T(0), ny = ey != T(0);
        const T(*Ff)[PX] = F[al];
        T fc             = T(0);
        // surface term (dt^2/2) oint phi (n.e)(e.grad f) on walls (advection.h:177-199), as exact
        // 1-D integrals along the wall: Dx of the tangential derivative, Mx of the normal one.
        // On Q1 the bands are read again here (wall nodes only), so that they need no
        // registers next to the weights.
        auto xb = [&](const int b, const int o) -> T {
          if constexpr (C::WGT)
            return T(band<p>(b == 0 ? G.ax.M : G.ax.D, gi, o - p));
          else
            return xc[b][o];
        };
        auto yb = [&](const int b, const int o) -> T {
          if constexpr (C::WGT)
            return T(band<p>(b == 0 ? G.ay.M : G.ay.D, gj, o - p));
          else
            return yc[b][o];
        };
        const int ty = jj + p, tx = ii + p;
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
                      dn = col::fma_t(T(bn[l]), Ff[r0 + l][ii + o], dn);
                    t   = col::fma_t(xb(1, o), Ff[ty][ii + o], t);
                    nrm = col::fma_t(xb(0, o), dn, nrm);
                  }
                fc = col::fma_t(hdt2 * (lo ? -ey : ey), col::fma_t(ex, t, ey * nrm), fc);
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
                      dn = col::fma_t(T(bn[l]), Ff[jj + o][c0 + l], dn);
                    t   = col::fma_t(yb(1, o), Ff[jj + o][tx], t);
                    nrm = col::fma_t(yb(0, o), dn, nrm);
                  }
                fc = col::fma_t(hdt2 * (lo ? -ex : ex), col::fma_t(ex, nrm, ey * t), fc);
              }
          }
        return fc;
      };

      T xinc[NM], mp[NM / 2];
      if constexpr (PS == 1)
        {
#pragma unroll
          for (int al = 1; al < Q; ++al)
            popA(al, xinc[al - 1]);
#pragma unroll
          for (int al = 1; al < Q; ++al)
            {
              const T fc   = face(al);
              xinc[al - 1] = xinc[al - 1] + fc;
              if (al & 1)
                mp[(al - 1) >> 1] = fc;
              else
                mp[(al - 1) >> 1] += fc;
            }
        }
      else
        {
          T res[C::PPP], pair = T(0);
#pragma unroll
          for (int k = 0; k < C::PPP; ++k)
            popA(q * C::PPP + k + 1, res[k]); // uniform over the wavefront
#pragma unroll
          for (int k = 0; k < C::PPP; ++k)
            {
              const int al = q * C::PPP + k + 1;
              const T fc           = face(al);
              sm.X[al - 1][jj][ii] = res[k] + fc;
              if ((k & 1) == 0)
                pair = fc;
              else
                {
                  pair += fc;
                  sm.MP[(al - 1) >> 1][jj][ii] = pair;
                }
            }
          __syncthreads();
          if (q != 0)
            return;
#pragma unroll
          for (int a = 0; a < NM; ++a)
            xinc[a] = sm.X[a][jj][ii];
#pragma unroll
          for (int k = 0; k < NM / 2; ++k)
            mp[k] = sm.MP[k][jj][ii];
        }
      const T mflux = ((mp[0] + mp[1]) + mp[2]) + mp[3];

      // ---------------------------------------------------------- epilogue
      if (gi > P || gj > P)
        return;
      const int64_t idx   = node_index(G, gi, gj);
      const bool    slave = (G.ax.periodic && gi == P) || (G.ay.periodic && gj == P);
      double        wx = 0., wy = 0.;
      const bool    wall = wall_of(G, gi, gj, wx, wy);
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
            own[a] = gfin[a * G.stride + idx];
          if constexpr (!LL)
            if (S.collide)
              collide_node(own, S.omega, wall, wx, wy);
        }
      // mass flux of the wall term removed isotropically (advection.h:194-196)
      if (S.remove_mass_flux)
#pragma unroll
        for (int a = 1; a < Q; ++a)
          xinc[a - 1] -= T(w_of(a) / (1. - w_of(0))) * mflux;
      if constexpr (STEP)
        {
          const T mli = slave ? T(0) : T(G.ax.ml_inv[gi] * G.ay.ml_inv[gj]);
#pragma unroll
          for (int a = 0; a < NM; ++a)
            xinc[a] *= mli;
          if constexpr (LL)
            {
              // the own node's equilibria from the prologue (a slave's from its own entries)
              if (slave)
                equilibrium_node(own, fe_own, wall, wx, wy);
              predictor_corrector_node(own, fe_own, xinc, S.theta, wall, wx, wy);
            }
          else
#pragma unroll
            for (int a = 1; a < Q; ++a)
              own[a] += xinc[a - 1];
#pragma unroll
          for (int a = 0; a < Q; ++a)
            gout[a * G.stride + idx] = own[a];
        }
      else
        {
          // right-hand side (constrained entries get nothing) and, fused, g*
#pragma unroll
          for (int a = 0; a < NM; ++a)
            grhs[a * G.stride + idx] = slave ? T(0) : xinc[a];
          if (out != nullptr)
#pragma unroll
            for (int a = 0; a < Q; ++a)
              gout[a * G.stride + idx] = own[a];
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

  // Many meshes in one launch: the tiles of all instances of the table form one
  // list (instance k owns tiles [tile_begin_k, tile_begin_k+1)); block b takes
  // tiles b, b + gridDim.x, ... in the XCD-aware order (gridDim.x a multiple of
  // 8), so that a grid of a few blocks per CU is persistent and a grid of one
  // block per tile is not. Each instance flips its own population buffers: the
  // step reads f[(par + parity) & 1] and writes the other.
  // The member and tile counts are read from the table's header at run time, so that a
  // captured launch keeps working after members join or leave the batch.
  template <class C>
  __global__ void __launch_bounds__(C::NT)
  k_tile_batch(const BatchEntry<typename C::T> *__restrict__ E, const BatchCounts *__restrict__ counts,
               const int parity)
  {
    __shared__ tile::Smem<C> sm;
    const int n_inst = counts->n_inst, n_tiles = counts->n_tiles;
    const int nb8    = (n_tiles + 7) & ~7;
    for (int b = blockIdx.x; b < nb8; b += gridDim.x)
      {
        const int t = (b & 7) * (nb8 >> 3) + (b >> 3);
        if (t >= n_tiles)
          continue;
        int lo = 0, hi = n_inst - 1; // the instance: last tile_begin <= t
        while (lo < hi)
          {
            const int mid = (lo + hi + 1) >> 1;
            if (E[mid].tile_begin <= t)
              lo = mid;
            else
              hi = mid - 1;
          }
        const BatchEntry<typename C::T> &e   = E[lo];
        const int                        cur = (e.par + parity) & 1;
        // the fused step writes the other buffer; the advection writes the right-hand side and, with
        // the collision fused, g* into the other buffer
        typename C::T *out = (C::STEP || e.S.collide) ? e.f[cur ^ 1] : nullptr;
        tile::body<C>(e.G, e.S, t - e.tile_begin, e.f[cur], nullptr, out, e.rhs, sm);
        __syncthreads(); // the next tile reuses the LDS
      }
  }

  // grid: the blocks of the launch (a multiple of 8; blocks stride over the tiles)
  template <class C>
  hipError_t
  launch_batch(const BatchEntry<typename C::T> *table, const BatchCounts *counts, const int parity, const int grid,
               hipStream_t stream)
  {
    hipLaunchKernelGGL((k_tile_batch<C>), dim3(grid), dim3(C::NT), 0, stream, table, counts, parity);
    return hipGetLastError();
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

  // The variants (TileVariant): kTileWide 32 x 8 tiles of 256 threads, kTileTall
  // 32 x 16 tiles of 512 threads (less halo per node), kTileSplit2 / kTileSplit4
  // 32 x 8 tiles with 2 or 4 threads per node, kTileSmall4 16 x 8 tiles with 4,
  // kTileWide40 / kTileWide48 40 x 8 and 48 x 8 tiles (5 and 6 wavefronts) whose
  // tile counts fit the 256 CUs on meshes just past a power of two nodes.
  template <typename T, int p, int MODE>
  hipError_t
  launch_variant(const int variant, const Geo &G, const TileParams &S, const T *fin, const T *feq, T *out, T *rhs,
                 hipStream_t stream)
  {
    switch (variant)
      {
        case kTileWide:
          return launch_tile<tile::Cfg<T, p, 32, 8, 1, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileTall:
          return launch_tile<tile::Cfg<T, p, 32, 16, 1, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileSplit2:
          return launch_tile<tile::Cfg<T, p, 32, 8, 2, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileSplit4:
          return launch_tile<tile::Cfg<T, p, 32, 8, 4, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileSmall4:
          return launch_tile<tile::Cfg<T, p, 16, 8, 4, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileWide40:
          return launch_tile<tile::Cfg<T, p, 40, 8, 1, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileWide48:
          return launch_tile<tile::Cfg<T, p, 48, 8, 1, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kTileSplit2W40:
          return launch_tile<tile::Cfg<T, p, 40, 8, 2, MODE>>(G, S, fin, feq, out, rhs, stream);
        case kMarch3:
        case kMarch4:
        case kMarch8:
        case kMarch16:
          if constexpr (march_runs(p, sizeof(T) == 4))
            {
              if (standard_directions(S) && march_fits(G, sizeof(T)))
                return launch_march<T, MODE, p>(march_rows(variant), G, S, fin, feq, out, rhs, stream);
              return launch_tile<tile::Cfg<T, p, 32, 8, 1, MODE>>(G, S, fin, feq, out, rhs, stream);
            }
          else
            return launch_tile<tile::Cfg<T, p, 32, 8, 1, MODE>>(G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  template <typename T, int p, int MODE>
  hipError_t
  launch_batch_variant(const int variant, const BatchEntry<T> *table, const BatchCounts *counts, const int parity,
                       const int grid, hipStream_t stream)
  {
    switch (variant)
      {
        case kTileWide:
          return launch_batch<tile::Cfg<T, p, 32, 8, 1, MODE>>(table, counts, parity, grid, stream);
        case kTileTall:
          return launch_batch<tile::Cfg<T, p, 32, 16, 1, MODE>>(table, counts, parity, grid, stream);
        case kTileSplit2:
          return launch_batch<tile::Cfg<T, p, 32, 8, 2, MODE>>(table, counts, parity, grid, stream);
        case kTileSplit4:
          return launch_batch<tile::Cfg<T, p, 32, 8, 4, MODE>>(table, counts, parity, grid, stream);
        case kTileSmall4:
          return launch_batch<tile::Cfg<T, p, 16, 8, 4, MODE>>(table, counts, parity, grid, stream);
        case kTileWide40:
          return launch_batch<tile::Cfg<T, p, 40, 8, 1, MODE>>(table, counts, parity, grid, stream);
        case kTileWide48:
          return launch_batch<tile::Cfg<T, p, 48, 8, 1, MODE>>(table, counts, parity, grid, stream);
        case kTileSplit2W40:
          return launch_batch<tile::Cfg<T, p, 40, 8, 2, MODE>>(table, counts, parity, grid, stream);
        case kMarch3:
        case kMarch4:
        case kMarch8:
        case kMarch16:
          // the table holds each member's march decomposition (march_runs) or its 32 x 8 tiles
          if constexpr (march_runs(p, sizeof(T) == 4))
            return launch_march_batch<T, MODE, p>(table, counts, parity, grid, stream);
          else
            return launch_batch<tile::Cfg<T, p, 32, 8, 1, MODE>>(table, counts, parity, grid, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  // the fused lumped step of many instances of one degree and form
  template <typename T, int p>
  hipError_t
  tile_batch_launch_degree(const int mode, const int variant, const BatchEntry<T> *table, const BatchCounts *counts,
                           const int parity, const int grid, hipStream_t stream)
  {
    switch (mode)
      {
        case kStepBardow:
          return launch_batch_variant<T, p, kStepBardow>(variant, table, counts, parity, grid, stream);
        case kStepLeeLin:
          return launch_batch_variant<T, p, kStepLeeLin>(variant, table, counts, parity, grid, stream);
        case kAdvectBardow:
          return launch_batch_variant<T, p, kAdvectBardow>(variant, table, counts, parity, grid, stream);
        case kAdvectLeeLin:
          return launch_batch_variant<T, p, kAdvectLeeLin>(variant, table, counts, parity, grid, stream);
        default:
          return hipErrorInvalidValue;
      }
  }

  // every mode of one degree
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
          return launch_variant<T, p, kAdvectBardow>(variant, G, S, fin, feq, out, rhs, stream);
        case kAdvectLeeLin:
          return launch_variant<T, p, kAdvectLeeLin>(variant, G, S, fin, feq, out, rhs, stream);
        default:
          return hipErrorInvalidValue;
      }
  }
} // namespace lbfem_amd
