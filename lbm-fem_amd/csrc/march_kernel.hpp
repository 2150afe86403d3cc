// The row-march kernel of the streaming step on Q1 and Q2 meshes: the same fused
// step (or advection right-hand side) as the tile kernel, computed without staging
// the populations in LDS.
//
// A wavefront owns a strip of 64 - 2p node columns (lanes p .. 63 - p; the p lanes
// on either side hold the neighbouring columns) and a segment of H node rows, and
// marches up the segment one row at a time. Each lane keeps its column's 2p + 1
// rows j - p .. j + p of the moving populations (f, and h = a f + b feq for
// Lee-Lin) in registers; each new row's nodal prologue (collision or equilibria)
// is done once per node and the oldest row dropped. The operator is the column
// form of column.hpp: each lane evaluates the y contractions of its own column
// once, forms the contributions w_0 .. w_2p of that column to the nodes x + p ..
// x - p, and the lanes exchange them by wavefront DPP shifts, so every node gets
// r = ((w_0 + w_1) + ...) + w_2p bit for bit as the tile kernel computes it. The
// wall surface terms take their neighbouring columns the same way. A row's own f_0
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
#include "column.hpp"

namespace lbfem_amd
{
  namespace march
  {
    constexpr int kWaves = 4; // wavefronts per block (independent; no barrier)

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

    // a value from the lane N below (N > 0: lane l receives lane l - N's) or above (N < 0: lane
    // l - N's), one lane per step; lanes at the wavefront's edge receive what they hold
    template <int N, typename V>
    __device__ __forceinline__ V
    shifted(const V v)
    {
      if constexpr (N > 0)
        return shifted<N - 1>(from_below(v));
      else if constexpr (N < 0)
        return shifted<N + 1>(from_above(v));
      else
        return v;
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

    // band entry k of row i of a 1-D operator of W entries per row (0 outside the rows 0..P)
    template <int W>
    __device__ __forceinline__ double
    band_at(const double *B, const int i, const int k, const int P)
    {
      return (i >= 0 && i <= P) ? ((const __attribute__((address_space(1))) double *)B)[i * W + k] : 0.;
    }
    // the same for a row index uniform over the wavefront: a scalar load (the bands are
    // read-only during a launch)
    template <int W>
    __device__ __forceinline__ double
    band_uniform(const double *B, const int i, const int k, const int P)
    {
      return (i >= 0 && i <= P) ? ((const __attribute__((address_space(4))) double *)B)[i * W + k] : 0.;
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
    // the per-wavefront LDS of the kernel: a row's own f_0 and (Lee-Lin) equilibria, for the p + 1
    // rows between a row's prologue and its epilogue, and the lane's x-band entries (read per row:
    // no registers held)
    template <typename T, bool LL, int p>
    struct Lds
    {
      static constexpr int NX = LL ? Q + 1 : 1;
      T                    ext[kWaves][p + 1][NX][64];
      T                    xbl[kWaves][4][2 * p + 1][64];
    };

    // one wavefront's march: wavefront gw of the mesh's list (strip gw mod S.tiles_x, row segment
    // gw / S.tiles_x of S.rows rows)
    template <typename T, int MODE, int p>
    __device__ __forceinline__ void
    body(const Geo &G, const TileParams &S, const T *__restrict__ fin, const T *__restrict__ feq_in,
         T *__restrict__ out, T *__restrict__ rhs_out, const int gw, const int wid,
         Lds<T, MODE == kStepLeeLin || MODE == kAdvectLeeLin, p> &lds)
  {
    constexpr bool LL   = (MODE == kStepLeeLin || MODE == kAdvectLeeLin);
    constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);
    constexpr int  NX   = LL ? Q + 1 : 1; // per-lane extras of a row: f_0 and (Lee-Lin) the equilibria
    constexpr int  W    = 2 * p + 1;     // band entries, window rows
    constexpr int  KO   = march_out(p);  // output columns per strip
    using R             = Row<T, LL>;
    auto           &ext  = lds.ext;
    auto           &xbl  = lds.xbl;

    const int lane  = threadIdx.x & 63;
    const int strip = gw % S.tiles_x, seg = gw / S.tiles_x;
    const int P = G.P, H = S.rows;
    const int jb = seg * H, je = min(jb + H, P + 1);
    const int x  = strip * KO - p + lane; // this lane's column position
    const int xw = wrap_line(x, P, G.ax.periodic);
    const bool outl = lane >= p && lane < p + KO && x <= P;

    const T a_c = T(S.a), b_c = T(S.b), c_c = T(S.c), hdt2 = T(0.5 * S.dt * S.dt);
    const T sa  = LL ? T(1) : a_c;
    // Per-lane values that do not change over the rows, read once with unconditional loads at
    // clamped indices: a conditional (divergent) vector load in the row loop would make the
    // wavefront wait for every memory operation in flight, the next row's loads included.
    // x bands: entry k of output node x + p - k (the nodes this column feeds), bands 0 M, 1 D, 2 K, 3 D^T
    {
      const double *xb[4] = {G.ax.M, G.ax.D, G.ax.K, G.ax.DT};
      double        bv[4][W];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int k = 0; k < W; ++k)
          bv[b][k] = ((const __attribute__((address_space(1))) double *)xb[b])[min(max(x + p - k, 0), P) * W + k];
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int k = 0; k < W; ++k)
          xbl[wid][b][k][lane] = (x + p - k >= 0 && x + p - k <= P) ? T(bv[b][k]) : T(0);
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
    const bool xwall_wave  = !G.ax.periodic && (strip == 0 || strip * KO + KO > P);
    // the LDS slot of a row's extras: p + 1 rows live between a row's prologue and its epilogue
    auto slot = [](const int tr) { return (tr + p + 1) % (p + 1); };

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
      T(&e)[NX][64]   = ext[wid][slot(tr)];
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
            row.h[a - 1] = col::fma_t(a_c, v[a], b_c * fe[a]);
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
    // row j from the window rw, whose slot (B + o) mod W holds row j - p + o
    auto output = [&](auto Bc, const int j, const R (&rw)[W], auto &&between) {
      constexpr int B   = decltype(Bc)::value;
      const R      &mid = rw[(B + p) % W];
      T             yb[4][W];
      {
        const double *ybp[4] = {G.ay.M, G.ay.D, G.ay.K, G.ay.DT};
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int o = 0; o < W; ++o)
            yb[b][o] = T(band_uniform<W>(ybp[b], j, o, P));
      }
      const bool ywall = !G.ay.periodic && (j == 0 || j == P);
      // FP32: the row sums of the M and D^T bands (the centred form's row terms)
      const int  ln    = opaque(lane);
      const T mM = col::row_sum(yb[0]), mT = col::row_sum(yb[3]);
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
          constexpr int      al = decltype(I)::value;
          const col::Coef<T> k  = col::coef<T>(T(ex_of(al)), T(ey_of(al)), sa, c_c);
          T                  r  = T(0);
          if (k.nx || k.ny)
            {
              T fv[W], hv[W];
#pragma unroll
              for (int o = 0; o < W; ++o)
                {
                  fv[o] = rw[(B + o) % W].f[al - 1];
                  hv[o] = LL ? rw[(B + o) % W].h[al - 1] : fv[o];
                }
              const auto &xr = xbl[wid];
              // w_kk: this column's contribution to the node x + p - kk, then moved to that node's
              // lane and summed in the order of kk
              // FP64: the column values, once; FP32: the column-centred contractions, once
              col::Cols<T>  u64{};
              col::CCols<T> cc{};
              if constexpr (sizeof(T) == 8)
                u64 = col::columns<T, LL, W>(yb, fv, hv, k);
              else
                cc = col::ccolumns<T, LL, W>(yb, fv, hv, k);
              for_pops<0, W>([&](auto KK) {
                constexpr int kk = decltype(KK)::value;
                col::Cols<T>  u;
                if constexpr (sizeof(T) == 8)
                  u = u64;
                else
                  {
                    // centred (FP32): the row terms with the fed node's centre values
                    if constexpr (kk == p)
                      u = col::assemble<T, LL>(cc, mM, mT, T(0), T(0), false, k);
                    else
                      {
                        const T fC = shifted<kk - p>(fv[p]);
                        const T hC = LL ? shifted<kk - p>(hv[p]) : fC;
                        u          = col::assemble<T, LL>(cc, mM, mT, hv[p] - hC, fv[p] - fC, true, k);
                      }
                  }
                const T w  = col::xcontract(xr[0][kk][ln], xr[1][kk][ln], xr[2][kk][ln], xr[3][kk][ln], u, k);
                const T wm = shifted<p - kk>(w);
                r          = kk == 0 ? wm : r + wm;
              });
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
          // on a wall row, the node's own rows of Dx and Mx (the tangential derivative and the mass
          // along the wall), read once for all populations
          T xDr[W], xMr[W];
#pragma unroll
          for (int o = 0; o < W; ++o)
            {
              xDr[o] = ywall ? T(band_at<W>(G.ax.D, opaque(x), o, P)) : T(0);
              xMr[o] = ywall ? T(band_at<W>(G.ax.M, opaque(x), o, P)) : T(0);
            }
          for_pops<1, Q>([&](auto I) {
            constexpr int      al = decltype(I)::value;
            const col::Coef<T> k  = col::coef<T>(T(ex_of(al)), T(ey_of(al)), sa, c_c);
            T                  fc = T(0);
            const T            ex = T(ex_of(al)), ey = T(ey_of(al));
            if (ywall && k.ny)
              {
                // the wall-normal derivative over rows j .. j + p (bottom) or j - p .. j (top), at
                // each of the W columns around the node
                const bool    lo_w = j == 0;
                const double *bn   = lo_w ? G.ay.b0 : G.ay.b1;
                T             dn   = T(0);
#pragma unroll
                for (int l = 0; l <= p; ++l)
                  dn = col::fma_t(T(bn[l]), lo_w ? rw[(B + p + l) % W].f[al - 1] : rw[(B + l) % W].f[al - 1], dn);
                // column by column (few registers live): t = sum_o Dx[o] f(x - p + o, j), nrm likewise
                T t = T(0), nrm = T(0);
                for_pops<0, W>([&](auto O) {
                  constexpr int o = decltype(O)::value;
                  const T       rv = shifted<p - o>(mid.f[al - 1]), dv = shifted<p - o>(dn);
                  t                = col::fma_t(xDr[o], rv, t);
                  nrm              = col::fma_t(xMr[o], dv, nrm);
                });
                fc = col::fma_t(hdt2 * (lo_w ? -ey : ey), col::fma_t(ex, t, ey * nrm), fc);
              }
            if (xwall_wave && k.nx)
              {
                // lanes at x = 0 take columns 1 .. p from the lanes above, lanes at x = P columns
                // P - p .. P - 1 from below: column c_l = x + l (x = 0) or x - p + l (x = P); one
                // pass for each side a strip holds (uniform), with only that side's lane shifts
                T t = T(0);
                for_pops<0, W>([&](auto O) {
                  constexpr int o = decltype(O)::value;
                  t               = col::fma_t(yb[1][o], rw[(B + o) % W].f[al - 1], t);
                });
                auto side = [&](auto LO, const double *bn) -> T {
                  constexpr bool lo = decltype(LO)::value;
                  T              nrm = T(0);
                  for_pops<0, W>([&](auto O) {
                    constexpr int o  = decltype(O)::value;
                    const T       f  = rw[(B + o) % W].f[al - 1];
                    T             dn = T(0);
                    for_pops<0, p + 1>([&](auto L) {
                      constexpr int l = decltype(L)::value;
                      dn              = col::fma_t(T(bn[l]), lo ? shifted<-l>(f) : shifted<p - l>(f), dn);
                    });
                    nrm = col::fma_t(yb[0][o], dn, nrm);
                  });
                  return col::fma_t(hdt2 * (lo ? -ex : ex), col::fma_t(ex, nrm, ey * t), fc);
                };
                T fx = fc;
                if (strip == 0)
                  {
                    const T f0 = side(std::true_type{}, G.ax.b0);
                    fx         = x == 0 ? f0 : fx;
                  }
                if (strip * KO + KO > P)
                  {
                    const T fP = side(std::false_type{}, G.ax.b1);
                    fx         = x == P ? fP : fx;
                  }
                if (xwall_strip && outl)
                  fc = fx;
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
          const T(&e)[NX][64] = ext[wid][slot(j)];
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
      if ((G.ay.periodic && j == P) || (G.ax.periodic && strip * KO + KO > P - 1))
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
    // start-up: the first W rows' loads go out together (one memory round trip, not W)
    // PF2: two rows of loads in flight (v: row j + p, v2: row j + p + 1) on Q1 in FP64, where the
    // second buffer fits beside the window at two wavefronts per SIMD (Lee-Lin 2049^2 nodes 190
    // against 195 us, 4097^2 676 against 688); in FP32 one (Lee-Lin 142 against 138 us with two)
    constexpr bool PF2 = p == 1 && sizeof(T) == 8;
    R              rw[W];
    T              v[Q], v2[PF2 ? Q : 1];
    {
      T vs[W - 1][Q];
#pragma unroll
      for (int o = 0; o < W - 1; ++o)
        load(jb - p + o, vs[o]);
      load(jb + p, v);
      if constexpr (PF2)
        if (jb + 1 < je)
          load(jb + p + 1, v2);
#pragma unroll
      for (int o = 0; o < W - 1; ++o)
        prologue(jb - p + o, vs[o], rw[o]);
    }
    // one row (window base B): the prologue of row j + p from its loads (issued one or two rows
    // before), the operator of row j, the loads of the row needed next (j + p + 1, or j + p + 2
    // with two in flight), the epilogue and stores of row j
    auto advance = [&](auto Bc, const int j) {
      constexpr int B = decltype(Bc)::value;
      prologue(j + p, v, rw[(B + W - 1) % W]);
      output(Bc, j, rw, [&]() {
        if constexpr (PF2)
          {
            if (j + 2 < je)
              load(j + p + 2, v);
          }
        else if (j + 1 < je)
          load(j + p + 1, v);
      });
      if constexpr (PF2)
#pragma unroll
        for (int a = 0; a < Q; ++a)
          {
            const T t = v[a];
            v[a]      = v2[a];
            v2[a]     = t;
          }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    static_assert(W == 3 || W == 5, "the row-march kernel runs on Q1 and Q2");
    for (int j = jb; j < je; j += W)
      {
        advance(I0{}, j);
        if (j + 1 >= je)
          break;
        advance(I1{}, j + 1);
        if (j + 2 >= je)
          break;
        advance(I2{}, j + 2);
        if constexpr (W == 5)
          {
            if (j + 3 >= je)
              break;
            advance(I3{}, j + 3);
            if (j + 4 >= je)
              break;
            advance(I4{}, j + 4);
          }
      }
  }
  } // namespace march

  // One mesh: XCD-aware block order (consecutive blocks of the list on one XCD); the wavefronts
  // of the list run along a band of row segments, so that a block's wavefronts are neighbouring
  // strips of one segment: they read the same rows at about the same time and share the halo
  // columns' cache lines in L2.
  template <typename T, int MODE, int p>
  __global__ void __launch_bounds__(64 * march::kWaves)
  k_march(const Geo G, const TileParams S, const T *__restrict__ fin, const T *__restrict__ feq_in,
          T *__restrict__ out, T *__restrict__ rhs_out)
  {
    using namespace march;
    __shared__ Lds<T, MODE == kStepLeeLin || MODE == kAdvectLeeLin, p> lds;
    const int wid   = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwave = S.tiles_x * S.tiles_y;
    const int nb    = (nwave + kWaves - 1) / kWaves;
    const int nb8   = (nb + 7) & ~7;
    const int bid   = blockIdx.x;
    const int tb    = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    const int gw    = tb * kWaves + wid;
    if (tb >= nb || gw >= nwave)
      return;
    body<T, MODE, p>(G, S, fin, feq_in, out, rhs_out, gw, wid, lds);
  }

  // Many meshes of one degree and form (a batch group): the wavefronts of all members form one
  // list (member k owns waves [tile_begin_k, tile_begin_k+1), its S.tiles_x strips by S.tiles_y
  // row segments of S.rows rows); block b of the grid takes the list's blocks b, b + gridDim.x, ...
  // (four wavefronts each) in the XCD-aware order of the list, so that any grid covers the list and
  // a captured launch keeps working after members join or leave. Each member flips its own
  // population buffers by parity.
  template <typename T, int MODE, int p>
  __global__ void __launch_bounds__(64 * march::kWaves)
  k_march_batch(const BatchEntry<T> *__restrict__ E, const BatchCounts *__restrict__ counts, const int parity)
  {
    using namespace march;
    constexpr bool STEP = (MODE == kStepBardow || MODE == kStepLeeLin);
    __shared__ Lds<T, MODE == kStepLeeLin || MODE == kAdvectLeeLin, p> lds;
    const int wid    = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n_inst = counts->n_inst, n_waves = counts->n_tiles;
    // the list's blocks (four wavefronts each, a multiple of 8) in the XCD-aware order of the list
    // itself, so that a grid larger than the list spreads it over every XCD (its extra blocks leave)
    // and a smaller one strides over it; the grid is a multiple of 8, so block b stays on XCD b mod 8
    const int nbl = ((n_waves + kWaves - 1) / kWaves + 7) & ~7;
    for (int b = blockIdx.x; b < nbl; b += gridDim.x)
      {
        const int tb = (b & 7) * (nbl >> 3) + (b >> 3);
        const int gw = tb * kWaves + wid;
        if (gw >= n_waves)
          continue;
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
        body<T, MODE, p>(e.G, e.S, e.f[cur], nullptr, out, e.rhs, gw - e.tile_begin, wid, lds);
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
  template <typename T, int MODE, int p>
  hipError_t
  launch_march(const int rows, const Geo &G, TileParams S, const T *fin, const T *feq, T *out, T *rhs,
               hipStream_t stream)
  {
    S.rows         = rows;
    S.tiles_x      = (G.P + 1 + march_out(p) - 1) / march_out(p); // strips
    S.tiles_y      = (G.P + 1 + rows - 1) / rows;               // row segments
    const int nw   = S.tiles_x * S.tiles_y;
    const int nb   = (nw + march::kWaves - 1) / march::kWaves;
    const int nb8  = (nb + 7) & ~7;
    hipLaunchKernelGGL((k_march<T, MODE, p>), dim3(nb8), dim3(64 * march::kWaves), 0, stream, G, S, fin, feq, out,
                       rhs);
    return hipGetLastError();
  }

  // grid: blocks of the batch launch (a multiple of 8; its wavefronts stride over the list)
  template <typename T, int MODE, int p>
  hipError_t
  launch_march_batch(const BatchEntry<T> *table, const BatchCounts *counts, const int parity, const int grid,
                     hipStream_t stream)
  {
    hipLaunchKernelGGL((k_march_batch<T, MODE, p>), dim3(grid), dim3(64 * march::kWaves), 0, stream, table, counts,
                       parity);
    return hipGetLastError();
  }
} // namespace lbfem_amd
