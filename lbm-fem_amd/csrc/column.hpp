// The streaming operator in column form, shared by the tile kernel and the
// row-march kernel so that both compute every increment by the same operation
// sequence (bitwise identical results), for any degree p (W = 2p + 1 band entries).
//
// On a cartesian Q_p mesh the advection right-hand side of population a at node
// (i, j) is (advection.h:211-293, the Kronecker form of the cell integrals)
//   r = ex sa (Dx (x) My) h + ey sa (Mx (x) Dy) h
//     + c (ex^2 (Kx (x) My) + ey^2 (Mx (x) Ky) + ex ey (Dx (x) Dy^T + Dx^T (x) Dy)) f
// with h = f (Bardow) or h = a f + b feq (Lee-Lin, sa = 1). It is evaluated in
// two contractions. First, at each of the W columns x = i - p + k (k = 0 .. 2p),
// the y bands of row j act on the column's W taps (rows j - p .. j + p) and the
// results are grouped by the x band they meet:
//   U_D = ex sa My h + c ex ey Dy^T f,   U_M = ey sa Dy h + c ey^2 Ky f,
//   U_K = c ex^2 My f,                   U_T = c ex ey Dy f.
// Then w_k = Dx[k] U_D + Mx[k] U_M + Kx[k] U_K + Dx^T[k] U_T with the x bands of
// column i, and r = ((w_0 + w_1) + ...) + w_2p. The column values depend only on
// the column and the row, so a kernel that holds one column per lane computes
// them once and passes the w_k to the neighbouring lanes; the tile kernel computes
// all W columns per node. Terms of a population that vanish by its direction
// class (x: ey = 0, y: ex = 0, diagonal) are skipped the same way in both. The
// FP32 form (below) centres the taps to carry differences.
#pragma once

#include "device.hpp"

namespace lbfem_amd
{
  namespace col
  {
    // fused multiply-add in the storage type (explicit, so that no contraction choice of
    // the compiler can differ between the kernels)
    __device__ __forceinline__ double
    fma_t(const double a, const double b, const double c)
    {
      return __builtin_fma(a, b, c);
    }
    __device__ __forceinline__ float
    fma_t(const float a, const float b, const float c)
    {
      return __builtin_fmaf(a, b, c);
    }

    // the direction and scheme coefficients of one population (uniform values)
    template <typename T>
    struct Coef
    {
      T    exsa, eysa, cexx, ceyy, cexy;
      bool nx, ny;
    };

    template <typename T>
    __device__ __forceinline__ Coef<T>
    coef(const T ex, const T ey, const T sa, const T c)
    {
      Coef<T> k;
      k.nx   = ex != T(0);
      k.ny   = ey != T(0);
      k.exsa = ex * sa;
      k.eysa = ey * sa;
      k.cexx = c * (ex * ex);
      k.ceyy = c * (ey * ey);
      k.cexy = c * (ex * ey);
      return k;
    }

    // sum_o y[o] v[o], o = 0 .. W - 1, in that order
    template <typename T, int W>
    __device__ __forceinline__ T
    chain(const T (&y)[W], const T (&v)[W])
    {
      T s = y[0] * v[0];
#pragma unroll
      for (int o = 1; o < W; ++o)
        s = fma_t(y[o], v[o], s);
      return s;
    }

    template <typename T>
    struct Cols
    {
      T D, M, K, T_;
    };

    // The column values of one population at one column from its taps fv (f) and hv
    // (h; the same as fv for Bardow), yb[band][o] the y bands of the output row
    // (bands: 0 M, 1 D, 2 K, 3 D^T).
    template <typename T, bool LL, int W>
    __device__ __forceinline__ Cols<T>
    columns(const T (&yb)[4][W], const T (&fv)[W], const T (&hv)[W], const Coef<T> &k)
    {
      Cols<T> u{T(0), T(0), T(0), T(0)};
      if (k.nx)
        {
          const T yMh = chain(yb[0], hv);
          const T yMf = LL ? chain(yb[0], fv) : yMh;
          u.D         = k.exsa * yMh;
          u.K         = k.cexx * yMf;
        }
      if (k.ny)
        {
          const T yDh = chain(yb[1], hv);
          const T yKf = chain(yb[2], fv);
          u.M         = fma_t(k.ceyy, yKf, k.eysa * yDh);
          if (k.nx)
            {
              const T yTf = chain(yb[3], fv);
              const T yDf = LL ? chain(yb[1], fv) : yDh;
              u.D         = fma_t(k.cexy, yTf, u.D);
              u.T_        = k.cexy * yDf;
            }
        }
      return u;
    }

    // w = Dx u.D + Mx u.M + Kx u.K + Dx^T u.T with the x-band entries xD, xM, xK, xT of
    // the output node at this column
    template <typename T>
    __device__ __forceinline__ T
    xcontract(const T xM, const T xD, const T xK, const T xT, const Cols<T> &u, const Coef<T> &k)
    {
      if (!k.nx)
        return fma_t(xM, u.M, T(0)); // not a bare product: no kernel may fuse it into the sum that follows
      T w = xD * u.D;
      if (k.ny)
        w = fma_t(xM, u.M, w);
      w = fma_t(xK, u.K, w);
      if (k.ny)
        w = fma_t(xT, u.T_, w);
      return w;
    }

    // ------------------------------------------------------------------ FP32 (centred) form
    // The mixed form applies the operator to differences, so that FP32 carries the small
    // variations of f instead of cancelling large values. Centred on the output node (i, j),
    //   sum_o y[o] (v(x, j - p + o) - v(i, j)) = sum_o y[o] (v(x, j - p + o) - v(x, j)) + m (v(x, j) - v(i, j)),
    // m = sum_o y[o]: the first part depends on the column only (shared by the W nodes a column
    // feeds), the second is a row term. m is taken as 0 for the D and K bands (their rows sum to
    // zero: every term of the operator then annihilates constants exactly) and formed from the
    // entries for M and D^T (D^T rows sum to -1 and 1 on the bottom and top boundary rows, to 0
    // elsewhere).

    // sum_{o != p} y[o] (v[o] - v[p]), o in increasing order
    template <typename T, int W>
    __device__ __forceinline__ T
    chain_c(const T (&y)[W], const T (&v)[W])
    {
      constexpr int p = W / 2;
      T             s = y[0] * (v[0] - v[p]);
#pragma unroll
      for (int o = 1; o < W; ++o)
        if (o != p)
          s = fma_t(y[o], v[o] - v[p], s);
      return s;
    }

    // the column-centred contractions of one population at one column
    template <typename T>
    struct CCols
    {
      T Mh, Dh, Mf, Kf, Tf, Df;
    };

    template <typename T, bool LL, int W>
    __device__ __forceinline__ CCols<T>
    ccolumns(const T (&yb)[4][W], const T (&fv)[W], const T (&hv)[W], const Coef<T> &k)
    {
      CCols<T> c{T(0), T(0), T(0), T(0), T(0), T(0)};
      if (k.nx)
        {
          c.Mh = chain_c(yb[0], hv);
          c.Mf = LL ? chain_c(yb[0], fv) : c.Mh;
        }
      if (k.ny)
        {
          c.Dh = chain_c(yb[1], hv);
          c.Kf = chain_c(yb[2], fv);
          if (k.nx)
            {
              c.Tf = chain_c(yb[3], fv);
              c.Df = LL ? chain_c(yb[1], fv) : c.Dh;
            }
        }
      return c;
    }

    // the column values for the node this column feeds at offset k: the row terms with dh =
    // h(x, j) - h(i, j) and df = f(x, j) - f(i, j) (none at the node's own column, where both
    // vanish); mM and mT the row sums of the M and D^T bands
    template <typename T, bool LL>
    __device__ __forceinline__ Cols<T>
    assemble(const CCols<T> &c, const T mM, const T mT, const T dh, const T df, const bool row, const Coef<T> &k)
    {
      Cols<T> u{T(0), T(0), T(0), T(0)};
      if (k.nx)
        {
          const T yMh = row ? fma_t(mM, dh, c.Mh) : c.Mh;
          const T yMf = LL ? (row ? fma_t(mM, df, c.Mf) : c.Mf) : yMh;
          u.D         = k.exsa * yMh;
          u.K         = k.cexx * yMf;
        }
      if (k.ny)
        {
          u.M = fma_t(k.ceyy, c.Kf, k.eysa * c.Dh);
          if (k.nx)
            {
              const T yTf = (row && mT != T(0)) ? fma_t(mT, df, c.Tf) : c.Tf;
              u.D         = fma_t(k.cexy, yTf, u.D);
              u.T_        = k.cexy * c.Df;
            }
        }
      return u;
    }

    // the row sum of a y band, in order, as every kernel forms it
    template <typename T, int W>
    __device__ __forceinline__ T
    row_sum(const T (&y)[W])
    {
      T s = y[0];
#pragma unroll
      for (int o = 1; o < W; ++o)
        s = s + y[o];
      return s;
    }
  } // namespace col
} // namespace lbfem_amd
