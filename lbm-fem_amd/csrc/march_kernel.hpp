// The row-march kernel of the Q1 streaming step: the same fused step (or
// advection right-hand side) as the tile kernel, computed without staging the
// populations in LDS.
//
// A wavefront owns a strip of 62 node columns (lanes 1..62; lanes 0 and 63 hold
// the neighbouring columns) and a segment of H node rows, and marches up the
// segment one row at a time. Each lane keeps its column's three rows j - 1, j,
// j + 1 of the moving populations (f, and h = a f + b feq for Lee-Lin) in
// registers; each new row's nodal prologue (collision or equilibria) is done once
// per node and the oldest row dropped. The operator is the column form of
// q1_column.hpp: each lane evaluates the y contractions of its own column once,
// forms the contributions w_0, w_1, w_2 of that column to the nodes x + 1, x,
// x - 1, and the lanes exchange w_0 and w_2 by wavefront DPP shifts, so every node
// gets r = (w_0 + w_1) + w_2 bit for bit as the tile kernel computes it. The wall
// surface terms take their neighbouring columns the same way. The row's own f_0
// and (Lee-Lin) equilibria wait in a small per-lane LDS slot until the row's
// epilogue, and the lane's x-band entries sit in LDS (read per row).
//
// Memory: the next row's loads are issued before a row's operator (in flight
// during the operator, the epilogue and the stores, in the registers the row's
// successor's prologue consumed). The wall surface terms run in a pass of their
// own on wall rows and wall strips, so that the operator's register peak does not
// include them. Nothing else in the row loop is a vector load on every row:
// per-lane constants are read once with unconditional loads, row-uniform lookups
// (y bands, 1 / lumped mass, side-wall tables) are scalar loads, and the per-lane
// bottom/top wall lookups run on those two rows only, so that no wait for a
// conditional load drains the row loads in flight. Nodal work per node is done
// once per segment instead of once per halo'd tile, and no LDS bandwidth is spent
// on the stencil (about half the tile kernel's VALU instructions per node).
#pragma once

#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"
#include "q1_column.hpp"

namespace lbfem_amd
{
  namespace march
  {
    constexpr int kOut   = kMarchOut; // output columns per wavefront
    constexpr int kWaves = 4;  // wavefronts per block (independent; no barrier)

    // a value from the lane below (lane l receives lane l - 1's; lane 0 its own)
    __device__ __forceinline__ double
    from_below(const double v)
    {
      const uint64_t b  = __builtin_bit_cast(uint64_t, v);
      const int      lo = int(uint32_t(b)), hi = int(uint32_t(b >> 32));
      const int      l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false); // wave_shr:1
      const int      h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
      return __builtin_bit_cast(double, uint64_t(uint32_t(l2)) | (uint64_t(uint32_t(h2)) << 32));
    }
    // a value from the lane above (lane l receives lane l + 1's; lane 63 its own)
    __device__ __forceinline__ double
    from_above(const double v)
    {
      const uint64_t b  = __builtin_bit_cast(uint64_t, v);
      const int      lo = int(uint32_t(b)), hi = int(uint32_t(b >> 32));
      const int      l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false); // wave_shl:1
      const int      h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
      return __builtin_bit_cast(double, uint64_t(uint32_t(l2)) | (uint64_t(uint32_t(h2)) << 32));
    }
    __device__ __forceinline__ float
    from_below(const float v)
    {
      const int b = __builtin_bit_cast(int, v);
      return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, 0x138, 0xf, 0xf, false));
    }
    __device__ __forceinline__ float
    from_above(const float v)
    {
      const int b = __builtin_bit_cast(int, v);
      return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(b, b, 0x130, 0xf, 0xf, false));
    }

    // one row of a lane's column: the moving populations f (g* for Bardow) and, for
    // Lee-Lin, h
    template <typename T, bool LL>
    struct Row
    {
      T f[NM];
      T h[LL ? NM : 1];
    };

    // f(integral_constant<int, A>), ..., f(integral_constant<int, B - 1>): a population loop
    // whose index is a compile-time constant
    template <int A, int B, class F>
    __device__ __forceinline__ void
    for_pops(F &&f)
    {
      if constexpr (A < B)
        {
          f(std::integral_constant<int, A>{});
          for_pops<A + 1, B>(f);
        }
    }

    // band entry k of row i of a 1-D operator (0 outside the rows 0..P)
    __device__ __forceinline__ double
    band_at(const double *B, const int i, const int k, const int P)
    {
      return (i >= 0 && i <= P) ? ((const __attribute__((address_space(1))) double *)B)[i * 3 + k] : 0.;
    }
    // the same for a row index uniform over the wavefront: a scalar load (the bands are
    // read-only during a launch)
    __device__ __forceinline__ double
    band_uniform(const double *B, const int i, const int k, const int P)
    {
      return (i >= 0 && i <= P) ? ((const __attribute__((address_space(4))) double *)B)[i * 3 + k] : 0.;
    }
    // *p for an address uniform over the wavefront: a scalar load (read-only data)
    template <typename V>
    __device__ __forceinline__ V
    sload(const V *p)
    {
      return *(const __attribute__((address_space(4))) V *)p;
    }
    // v, as a value the compiler cannot see through: reads indexed by it stay where they are
    // written (inside the row loop) instead of being hoisted into registers held across rows
    __device__ __forceinline__ int
    opaque(int v)
    {
      asm volatile("" : "+v"(v));
      return v;
    }

    // Buffer access to a block of population vectors: the lane's node byte offset (voffset) plus
    // the population's block offset a * stride (soffset, uniform), bounded by the block's size
    // (an access past it reads zero instead of faulting). The launchers run this kernel for
    // blocks under 4 GiB.
    using Rsrc = __amdgpu_buffer_rsrc_t;
    __device__ __forceinline__ Rsrc
    block_rsrc(const void *base, const uint32_t bytes)
    {
      return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, int(bytes), 0x00020000);
    }
    template <typename T>
    __device__ __forceinline__ T
    bload(const Rsrc r, const uint32_t voff, const uint32_t soff)
    {
      if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
      else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    }
    __device__ __forceinline__ void
    bstore(const double v, const Rsrc r, const uint32_t voff, const uint32_t soff)
    {
      using V2 = unsigned __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, voff, soff, 0);
    }
    __device__ __forceinline__ void
    bstore(const float v, const Rsrc r, const uint32_t voff, const uint32_t soff)
    {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    }
  } // namespace march

  namespace march
  {
    // the per-wavefront LDS of the kernel: a row's own f_0 and (Lee-Lin) equilibria by row parity,
    // and the lane's x-band entries (read per row: no registers held)
    template <typename T, bool LL>
    struct Lds
    {
      static constexpr int NX = LL ? Q + 1 : 1;
      T                    ext[kWaves][2][NX][64];
      T                    xbl[kWaves][4][3][64];
    };

    // one wavefront's march: wavefront gw of the mesh's list (strip gw mod S.tiles_x, row segment
    // gw / S.tiles_x of S.rows rows)
    template <typename T, int MODE>
    __device__ __forceinline__ void
    body(const Geo &G, const TileParams &S, const T *__restrict__ fin, const T *__restrict__ feq_in,
         T *__restrict__ out, T *__restrict__ rhs_out, const int gw, const int wid,
         Lds<T, MODE == kStepLeeLin || MODE == kAdvectLeeLin> &lds)
  {
    constexpr bool LL   = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
    constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);
    constexpr int  NX   = LL ? Q + 1 : 1; // per-lane extras of a row: f_0 and (Lee-Lin) the equilibria
    using R             = Row<T, LL>;
    auto           &ext  = lds.ext;
    auto           &xbl  = lds.xbl;

    const int lane  = threadIdx.x & 63;
    const int strip = gw % S.tiles_x, seg = gw / S.tiles_x;
    const int P = G.P, H = S.rows;
    const int jb = seg * H, je = min(jb + H, P + 1);
    const int x  = strip * kOut - 1 + lane; // this lane's column position
    const int xw = wrap_line(x, P, G.ax.periodic);
    const bool outl = lane >= 1 && lane <= kOut && x <= P;

    const T a_c = T(S.a), b_c = T(S.b), c_c = T(S.c), hdt2 = T(0.5 * S.dt * S.dt);
    const T sa  = LL ? T(1) : a_c;
    // Per-lane values that do not change over the rows, read once with unconditional loads at
    // clamped indices: a conditional (divergent) vector load in the row loop would make the
    // wavefront wait for every memory operation in flight, the next row's loads included.
    // x bands: entry k of output node x + 1 - k (the nodes this column feeds), bands 0 M, 1 D, 2 K, 3 D^T
    {
      const double *xb[4] = {G.ax.M, G.ax.D, G.ax.K, G.ax.DT};
      double        bv[4][3];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          bv[b][k] = ((const __attribute__((address_space(1))) double *)xb[b])[min(max(x + 1 - k, 0), P) * 3 + k];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          xbl[wid][b][k][lane] = (x + 1 - k >= 0 && x + 1 - k <= P) ? T(bv[b][k]) : T(0);
    }
    // 1 / lumped mass of the lane's column (output lanes: 0 <= x <= P)
    const double mlx = ((const __attribute__((address_space(1))) double *)G.ax.ml_inv)[min(max(x, 0), P)];
    // the population blocks (Q vectors of fin, feq and out, NM of rhs) as buffer resources; a
    // population's soffset is a * sb
    const uint32_t sb   = uint32_t(G.stride) * uint32_t(sizeof(T));
    const Rsrc     r_in = block_rsrc(fin, Q * sb), r_eq = block_rsrc(feq_in, Q * sb);
    const Rsrc     r_out = block_rsrc(out, Q * sb), r_rhs = block_rsrc(rhs_out, NM * sb);
    // the wall columns x = 0 and x = P (non-periodic x) and the strips that hold them
    const bool xwall_strip = !G.ax.periodic && (x == 0 || x == P);
    const bool xwall_wave  = !G.ax.periodic && (strip == 0 || strip * kOut + kOut > P);

    // ------------------------------------------------------------------ loads and prologue
    // the loads of row tr (node (0, 0) stands in outside a non-periodic axis; the prologue
    // zeroes those values): unconditional, so that no branch separates them from the code
    auto load = [&](const int tr, T (&v)[Q]) {
      const int      yw  = wrap_line(tr, P, G.ay.periodic);
      const uint32_t off = (xw >= 0 && yw >= 0) ? uint32_t(yw * G.np + xw) * uint32_t(sizeof(T)) : 0u;
#pragma unroll
      for (int a = 0; a < Q; ++a)
        v[a] = bload<T>(r_in, off, a * sb);
    };
    // wall_of(G, i, jj) (collision.h:145-165) with the row index jj uniform over the wavefront: the
    // side walls' table entries and velocities are scalar loads; the bottom and top walls' per-lane
    // lookups run on those two rows only
    auto wall_at = [&](const int i, const int jj, double &wx, double &wy) -> bool {
      int k = -1;
      if ((jj == 0 && G.walls.side[2]) || (jj == P && G.walls.side[3]))
        {
          k = (jj == 0 && G.walls.side[2]) ? G.walls.side[2][i] : G.walls.side[3][i];
          if (k >= 0)
            {
              wx = G.walls.velocity[2 * k];
              wy = G.walls.velocity[2 * k + 1];
            }
          // the values are used here, so that the wait for these loads stays on these two rows
          // instead of landing where the paths join (where it would wait for the next row's loads)
          asm volatile("" : "+v"(k), "+v"(wx), "+v"(wy));
        }
      if (G.walls.side[0] || G.walls.side[1])
        {
          const int s0 = G.walls.side[0] ? sload(G.walls.side[0] + jj) : -1;
          const int s1 = G.walls.side[1] ? sload(G.walls.side[1] + jj) : -1;
          double    v0x = 0., v0y = 0., v1x = 0., v1y = 0.;
          if (s0 >= 0)
            {
              v0x = sload(G.walls.velocity + 2 * s0);
              v0y = sload(G.walls.velocity + 2 * s0 + 1);
            }
          if (s1 >= 0)
            {
              v1x = sload(G.walls.velocity + 2 * s1);
              v1y = sload(G.walls.velocity + 2 * s1 + 1);
            }
          if (k < 0 && i == 0 && s0 >= 0)
            {
              k  = s0;
              wx = v0x;
              wy = v0y;
            }
          else if (k < 0 && i == P && s1 >= 0)
            {
              k  = s1;
              wx = v1x;
              wy = v1y;
            }
        }
      return k >= 0;
    };
    // the nodal prologue of row tr from its loaded values v into the row slot and the extras
    auto prologue = [&](const int tr, T (&v)[Q], R &row) {
      const int  yw = wrap_line(tr, P, G.ay.periodic);
      const bool ok = xw >= 0 && yw >= 0;
      double     wx = 0., wy = 0.;
      const bool wall = yw >= 0 && wall_at(max(xw, 0), yw, wx, wy) && ok;
      T(&e)[NX][64]   = ext[wid][tr & 1];
      if (!ok)
#pragma unroll
        for (int a = 0; a < Q; ++a)
          v[a] = T(0);
      if constexpr (LL)
        {
          T fe[Q];
          if (!ok)
#pragma unroll
            for (int a = 0; a < Q; ++a)
              fe[a] = T(0);
          else if (STEP || feq_in == nullptr) // the step modes compute the equilibria (given: advection only)
            equilibrium_node(v, fe, wall, wx, wy);
          else
#pragma unroll
            for (int a = 0; a < Q; ++a)
              fe[a] = bload<T>(r_eq, uint32_t(yw * G.np + xw) * uint32_t(sizeof(T)), a * sb);
#pragma unroll
          for (int a = 1; a < Q; ++a)
            row.h[a - 1] = q1::fma_t(a_c, v[a], b_c * fe[a]);
#pragma unroll
          for (int a = 0; a < Q; ++a)
            e[1 + a][lane] = fe[a];
        }
      else if (S.collide && ok)
        collide_node(v, S.omega, wall, wx, wy);
      e[0][lane] = v[0];
#pragma unroll
      for (int a = 1; a < Q; ++a)
        row.f[a - 1] = v[a];
    };

    // ------------------------------------------------------------------ one output row
    auto output = [&](const int j, const R &lo, const R &mid, const R &hi, auto &&between) {
      T yb[4][3];
      {
        const double *ybp[4] = {G.ay.M, G.ay.D, G.ay.K, G.ay.DT};
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int o = 0; o < 3; ++o)
            yb[b][o] = T(band_uniform(ybp[b], j, o, P));
      }
      const bool ywall = !G.ay.periodic && (j == 0 || j == P);
      // FP32: the row sums of the M and D^T bands (the centred form's row terms)
      const int  ln    = opaque(lane);
      const T mM = q1::row_sum(yb[0]), mT = q1::row_sum(yb[3]);
      T          xinc[NM];
      // the next row's loads go out before the operator: in flight for the whole operator and
      // epilogue (their registers are those the prologue of this row's successor consumed), and
      // ahead of this row's stores, so that waiting for them does not wait for the stores (the
      // vector memory counter counts both)
      between();
      // the D2Q9 directions as compile-time constants (the launcher runs this kernel for the
      // standard direction set only): the coefficients fold to +-sa, +-c, c, exactly the
      // products the tile kernel forms at run time, and no per-population value stays live
      // across the rows
      for_pops<1, Q>([&](auto I) {
          constexpr int     al = decltype(I)::value;
          const q1::Coef<T> k  = q1::coef<T>(T(ex_of(al)), T(ey_of(al)), sa, c_c);
          T                 r  = T(0);
          if (k.nx || k.ny)
            {
              const T fv[3] = {lo.f[al - 1], mid.f[al - 1], hi.f[al - 1]};
              T       hv[3];
#pragma unroll
              for (int o = 0; o < 3; ++o)
                hv[o] = LL ? (o == 0 ? lo.h[al - 1] : o == 1 ? mid.h[al - 1] : hi.h[al - 1]) : fv[o];
              q1::Cols<T> u0, u1, u2;
              if constexpr (sizeof(T) == 8)
                u0 = u1 = u2 = q1::columns<T, LL>(yb, fv, hv, k);
              else
                {
                  // centred (FP32): the column's contractions once, the row terms per fed node with
                  // the neighbouring columns' centre values
                  const q1::CCols<T> cc = q1::ccolumns<T, LL>(yb, fv, hv, k);
                  const T            fR = from_above(fv[1]), fL = from_below(fv[1]);
                  const T            hR = LL ? from_above(hv[1]) : fR, hL = LL ? from_below(hv[1]) : fL;
                  u0 = q1::assemble<T, LL>(cc, mM, mT, hv[1] - hR, fv[1] - fR, true, k);
                  u1 = q1::assemble<T, LL>(cc, mM, mT, T(0), T(0), false, k);
                  u2 = q1::assemble<T, LL>(cc, mM, mT, hv[1] - hL, fv[1] - fL, true, k);
                }
              const auto &xr = xbl[wid];
              const T w0 = q1::xcontract(xr[0][0][ln], xr[1][0][ln], xr[2][0][ln], xr[3][0][ln], u0, k);
              const T w1 = q1::xcontract(xr[0][1][ln], xr[1][1][ln], xr[2][1][ln], xr[3][1][ln], u1, k);
              const T w2 = q1::xcontract(xr[0][2][ln], xr[1][2][ln], xr[2][2][ln], xr[3][2][ln], u2, k);
              r                    = (from_below(w0) + w1) + from_above(w2);
            }
          xinc[al - 1] = r;
        });

      // The wall surface terms (advection.h:177-199), as tile::body's face(), in a pass of their own
      // over the wall rows and the strips that hold a wall column (uniform branch): xinc = r + fc
      // for every population (fc = 0 where the term vanishes), the same sums as one pass; the wall
      // mass flux ((m_0 + m_1) + m_2) + m_3, m_k = fc(2k + 1) + fc(2k + 2). Out of the operator loop,
      // the face code adds no registers to the operator's.
      T mflux = T(0);
      if (ywall || xwall_wave)
        {
          T mp = T(0);
          for_pops<1, Q>([&](auto I) {
            constexpr int     al = decltype(I)::value;
            const q1::Coef<T> k  = q1::coef<T>(T(ex_of(al)), T(ey_of(al)), sa, c_c);
            T                 fc = T(0);
            const T    ex = T(ex_of(al)), ey = T(ey_of(al));
            if (ywall && k.ny)
              {
                const bool    lo_w = j == 0;
                const double *bn   = lo_w ? G.ay.b0 : G.ay.b1;
                const T       r0v  = lo_w ? mid.f[al - 1] : lo.f[al - 1];
                const T       r1v  = lo_w ? hi.f[al - 1] : mid.f[al - 1];
                T             dn   = q1::fma_t(T(bn[0]), r0v, T(0));
                dn                 = q1::fma_t(T(bn[1]), r1v, dn);
                const T rv[3] = {from_below(mid.f[al - 1]), mid.f[al - 1], from_above(mid.f[al - 1])};
                const T dv[3] = {from_below(dn), dn, from_above(dn)};
                T       t = T(0), nrm = T(0);
#pragma unroll
                for (int o = 0; o < 3; ++o)
                  {
                    t   = q1::fma_t(T(band_at(G.ax.D, opaque(x), o, P)), rv[o], t);
                    nrm = q1::fma_t(T(band_at(G.ax.M, opaque(x), o, P)), dv[o], nrm);
                  }
                fc = q1::fma_t(hdt2 * (lo_w ? -ey : ey), q1::fma_t(ex, t, ey * nrm), fc);
              }
            if (xwall_wave && k.nx)
              {
                // lanes at x = 0 take column 1 from the lane above, lanes at x = P column P - 1 from below
                const bool lo_w = x == 0;
                const T    fr[3] = {lo.f[al - 1], mid.f[al - 1], hi.f[al - 1]};
                T          nb[3];
#pragma unroll
                for (int o = 0; o < 3; ++o)
                  {
                    const T up = from_above(fr[o]), dw = from_below(fr[o]);
                    nb[o]      = lo_w ? up : dw;
                  }
                if (xwall_strip && outl)
                  {
                    const double *bn = lo_w ? G.ax.b0 : G.ax.b1;
                    T             t = T(0), nrm = T(0);
#pragma unroll
                    for (int o = 0; o < 3; ++o)
                      {
                        const T c0 = lo_w ? fr[o] : nb[o], c1 = lo_w ? nb[o] : fr[o];
                        T       dn = q1::fma_t(T(bn[0]), c0, T(0));
                        dn         = q1::fma_t(T(bn[1]), c1, dn);
                        t          = q1::fma_t(yb[1][o], fr[o], t);
                        nrm        = q1::fma_t(yb[0][o], dn, nrm);
                      }
                    fc = q1::fma_t(hdt2 * (lo_w ? -ex : ex), q1::fma_t(ex, nrm, ey * t), fc);
                  }
              }
            xinc[al - 1] = xinc[al - 1] + fc;
            if constexpr (al & 1)
              mp = fc;
            else if constexpr (al == 2)
              mflux = mp + fc;
            else
              mflux += mp + fc;
          });
        }
      else
#pragma unroll
        for (int a = 0; a < NM; ++a)
          xinc[a] = xinc[a] + T(0); // r + fc with fc = 0, as above

      // ---------------------------------------------------------------- epilogue (tile::body's)
      const int      gi = x, gj = j;
      const uint32_t off   = uint32_t(gj * G.np + gi) * uint32_t(sizeof(T));
      const bool     slave = (G.ax.periodic && gi == P) || (G.ay.periodic && gj == P);
      if (outl && !slave)
        {
          double     wx = 0., wy = 0.;
          const bool wall = wall_at(gi, gj, wx, wy);
          T          own[Q];
          const T(&e)[NX][64] = ext[wid][j & 1];
          own[0]              = e[0][lane];
#pragma unroll
          for (int a = 1; a < Q; ++a)
            own[a] = mid.f[a - 1];
          if (S.remove_mass_flux)
#pragma unroll
            for (int a = 1; a < Q; ++a)
              xinc[a - 1] -= T(w_of(a) / (1. - w_of(0))) * mflux;
          if constexpr (STEP)
            {
              const T mli = T(mlx * sload(G.ay.ml_inv + gj));
#pragma unroll
              for (int a = 0; a < NM; ++a)
                xinc[a] *= mli;
              if constexpr (LL)
                {
                  T fe_own[Q];
#pragma unroll
                  for (int a = 0; a < Q; ++a)
                    fe_own[a] = e[1 + a][lane];
                  predictor_corrector_node(own, fe_own, xinc, S.theta, wall, wx, wy);
                }
              else
#pragma unroll
                for (int a = 1; a < Q; ++a)
                  own[a] += xinc[a - 1];
#pragma unroll
              for (int a = 0; a < Q; ++a)
                bstore(own[a], r_out, off, a * sb);
            }
          else
            {
#pragma unroll
              for (int a = 0; a < NM; ++a)
                bstore(xinc[a], r_rhs, off, a * sb);
              if (out != nullptr)
#pragma unroll
                for (int a = 0; a < Q; ++a)
                  bstore(own[a], r_out, off, a * sb);
            }
        }
      // Periodic slaves (the lines x = P, y = P of a periodic axis), after the row's stores and in
      // waves that hold one only: their staged values are their masters', the lumped solve gives
      // them no increment (M_L^-1 is 0 on constrained rows, so the tile kernel's increment times
      // 0 is an exact zero) and the right-hand side is 0, so only nodal work on their own entries
      // remains: the collision (Bardow), the predictor-corrector with zero increments (Lee-Lin)
      if ((G.ay.periodic && j == P) || (G.ax.periodic && strip * kOut + kOut > P - 1))
        if (outl && slave)
          {
            double     wx = 0., wy = 0.;
            const bool wall = wall_of(G, gi, gj, wx, wy);
            T          own[Q];
#pragma unroll
            for (int a = 0; a < Q; ++a)
              own[a] = bload<T>(r_in, off, a * sb);
            if constexpr (!LL)
              if (S.collide)
                collide_node(own, S.omega, wall, wx, wy);
            if constexpr (STEP)
              {
                // the tile kernel's increments of a constrained row: (r - mass flux) times 0 (exact
                // zeros, NaN where the operator's value is not finite)
                if (S.remove_mass_flux)
#pragma unroll
                  for (int a = 1; a < Q; ++a)
                    xinc[a - 1] -= T(w_of(a) / (1. - w_of(0))) * mflux;
#pragma unroll
                for (int a = 0; a < NM; ++a)
                  xinc[a] *= T(0);
                if constexpr (LL)
                  {
                    T fe_own[Q];
                    equilibrium_node(own, fe_own, wall, wx, wy);
                    predictor_corrector_node(own, fe_own, xinc, S.theta, wall, wx, wy);
                  }
                else
#pragma unroll
                  for (int a = 1; a < Q; ++a)
                    own[a] += xinc[a - 1];
#pragma unroll
                for (int a = 0; a < Q; ++a)
                  bstore(own[a], r_out, off, a * sb);
              }
            else
              {
#pragma unroll
                for (int a = 0; a < NM; ++a)
                  bstore(T(0), r_rhs, off, a * sb);
                if (out != nullptr)
#pragma unroll
                  for (int a = 0; a < Q; ++a)
                    bstore(own[a], r_out, off, a * sb);
              }
          }
    };

    // ------------------------------------------------------------------ the march
    // start-up: the first three rows' loads go out together (one memory round trip, not three)
    R r0, r1, r2;
    T v[Q];
    {
      T v0[Q], v1[Q];
      load(jb - 1, v0);
      load(jb, v1);
      load(jb + 1, v);
      prologue(jb - 1, v0, r0);
      prologue(jb, v1, r1);
    }
    // one row: the prologue of row j + 1 from its loads (issued during the previous row), the
    // operator of row j, the loads of row j + 2, the epilogue and stores of row j
    auto advance = [&](const int j, const R &lo, const R &mid, R &hi) {
      prologue(j + 1, v, hi);
      output(j, lo, mid, hi, [&]() {
        if (j + 1 < je)
          load(j + 2, v);
      });
    };
    for (int j = jb; j < je; j += 3)
      {
        advance(j, r0, r1, r2);
        if (j + 1 >= je)
          break;
        advance(j + 1, r1, r2, r0);
        if (j + 2 >= je)
          break;
        advance(j + 2, r2, r0, r1);
      }
  }
  } // namespace march

  // One mesh: XCD-aware block order (consecutive blocks of the list on one XCD); the wavefronts
  // of the list run along a band of row segments, so that a block's wavefronts are neighbouring
  // strips of one segment: they read the same rows at about the same time and share the halo
  // columns' cache lines in L2.
  template <typename T, int MODE>
  __global__ void __launch_bounds__(64 * march::kWaves)
  k_march(const Geo G, const TileParams S, const T *__restrict__ fin, const T *__restrict__ feq_in,
          T *__restrict__ out, T *__restrict__ rhs_out)
  {
    using namespace march;
    __shared__ Lds<T, MODE == kStepLeeLin || MODE == kAdvectLeeLin> lds;
    const int wid   = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwave = S.tiles_x * S.tiles_y;
    const int nb    = (nwave + kWaves - 1) / kWaves;
    const int nb8   = (nb + 7) & ~7;
    const int bid   = blockIdx.x;
    const int tb    = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    const int gw    = tb * kWaves + wid;
    if (tb >= nb || gw >= nwave)
      return;
    body<T, MODE>(G, S, fin, feq_in, out, rhs_out, gw, wid, lds);
  }

  // Many meshes of one degree and form (a batch group): the wavefronts of all members form one
  // list (member k owns waves [tile_begin_k, tile_begin_k+1), its S.tiles_x strips by S.tiles_y
  // row segments of S.rows rows); wavefront w of the grid takes list entries w, w + 4 gridDim.x,
  // ... in the XCD-aware block order, so that any grid covers the list and a captured launch keeps
  // working after members join or leave. Each member flips its own population buffers by parity.
  template <typename T, int MODE>
  __global__ void __launch_bounds__(64 * march::kWaves)
  k_march_batch(const BatchEntry<T> *__restrict__ E, const BatchCounts *__restrict__ counts, const int parity)
  {
    using namespace march;
    constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);
    __shared__ Lds<T, MODE == kStepLeeLin || MODE == kAdvectLeeLin> lds;
    const int wid    = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n_inst = counts->n_inst, n_waves = counts->n_tiles;
    const int nb8    = gridDim.x; // a multiple of 8
    const int tb     = (blockIdx.x & 7) * (nb8 >> 3) + (blockIdx.x >> 3);
    for (int gw = tb * kWaves + wid; gw < n_waves; gw += nb8 * kWaves)
      {
        int lo = 0, hi = n_inst - 1; // the member: last tile_begin <= gw
        while (lo < hi)
          {
            const int mid = (lo + hi + 1) >> 1;
            if (E[mid].tile_begin <= gw)
              lo = mid;
            else
              hi = mid - 1;
          }
        const BatchEntry<T> &e   = E[lo];
        const int            cur = (e.par + parity) & 1;
        T                   *out = (STEP || e.S.collide) ? e.f[cur ^ 1] : nullptr;
        body<T, MODE>(e.G, e.S, e.f[cur], nullptr, out, e.rhs, gw - e.tile_begin, wid, lds);
      }
  }

  // the sweep's directions are the D2Q9 set (the row-march kernel has them built in)
  inline bool
  standard_directions(const TileParams &S)
  {
    for (int a = 0; a < Q; ++a)
      if (S.e[a][0] != double(ex_of(a)) || S.e[a][1] != double(ey_of(a)))
        return false;
    return true;
  }

  // Launch of the row-march kernel over the mesh with H rows per wavefront segment.
  template <typename T, int MODE>
  hipError_t
  launch_march(const int rows, const Geo &G, TileParams S, const T *fin, const T *feq, T *out, T *rhs,
               hipStream_t stream)
  {
    S.rows         = rows;
    S.tiles_x      = (G.P + 1 + march::kOut - 1) / march::kOut; // strips
    S.tiles_y      = (G.P + 1 + rows - 1) / rows;               // row segments
    const int nw   = S.tiles_x * S.tiles_y;
    const int nb   = (nw + march::kWaves - 1) / march::kWaves;
    const int nb8  = (nb + 7) & ~7;
    hipLaunchKernelGGL((k_march<T, MODE>), dim3(nb8), dim3(64 * march::kWaves), 0, stream, G, S, fin, feq, out, rhs);
    return hipGetLastError();
  }

  // grid: blocks of the batch launch (a multiple of 8; its wavefronts stride over the list)
  template <typename T, int MODE>
  hipError_t
  launch_march_batch(const BatchEntry<T> *table, const BatchCounts *counts, const int parity, const int grid,
                     hipStream_t stream)
  {
    hipLaunchKernelGGL((k_march_batch<T, MODE>), dim3(grid), dim3(64 * march::kWaves), 0, stream, table, counts,
                       parity);
    return hipGetLastError();
  }
} // namespace lbfem_amd
