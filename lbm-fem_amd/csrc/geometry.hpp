// Host-side geometry of the structured Q_p discretization: the 1-D FE_Q(p)
// basis on Gauss-Lobatto nodes with QGauss(p+1) quadrature, and per axis the
// assembled 1-D operators whose Kronecker products are the 2-D matrices of the
// reference on a cartesian (tensor-product) mesh:
//
//   M   = Mx (x) My
//   C_e = ex Dx (x) My + ey Mx (x) Dy
//   K_e = ex^2 Kx (x) My + ey^2 Mx (x) Ky + ex ey (Dx^T (x) Dy + Dx (x) Dy^T)
//
// with the 1-D assembled matrices M_ij = int phi_i phi_j, D_ij = int phi_i phi_j',
// K_ij = int phi_i' phi_j' (include/lbfem/advection.h:7-10, mass.h:57-62;
// QGauss(p+1) integrates all three exactly, discretization.h:121). Each row
// has at most 2p+1 entries, stored as a band [i][o + p], o = j - i, with the
// periodic wrap folded in (a periodic axis has p N masters; its slave row
// p N is left empty).
#pragma once

#include <cmath>
#include <stdexcept>
#include <vector>

namespace lbfem_amd
{
  constexpr int    kMaxDegree = 3;
  constexpr double kPi        = 3.14159265358979323846;

  struct Basis1D
  {
    int    p = 1;
    double node[kMaxDegree + 1]{}, qx[kMaxDegree + 1]{}, qw[kMaxDegree + 1]{};
    double S[kMaxDegree + 1][kMaxDegree + 1]{}; // S[q][k] = phi_k(x_q)
    double D[kMaxDegree + 1][kMaxDegree + 1]{}; // D[q][k] = phi_k'(x_q), reference cell [0,1]
    double d0[kMaxDegree + 1]{}, d1[kMaxDegree + 1]{}; // phi_k'(0), phi_k'(1)
    double m[kMaxDegree + 1][kMaxDegree + 1]{};  // int phi_k phi_l      (reference cell)
    double dm[kMaxDegree + 1][kMaxDegree + 1]{}; // int phi_k phi_l'
    double k[kMaxDegree + 1][kMaxDegree + 1]{};  // int phi_k' phi_l'

    explicit Basis1D(int degree);
  };

  inline double
  lagrange(const double *x, const int n, const int k, const double t)
  {
    double v = 1.;
    for (int m = 0; m < n; ++m)
      if (m != k)
        v *= (t - x[m]) / (x[k] - x[m]);
    return v;
  }

  inline double
  lagrange_derivative(const double *x, const int n, const int k, const double t)
  {
    double s = 0.;
    for (int j = 0; j < n; ++j)
      if (j != k)
        {
          double v = 1. / (x[k] - x[j]);
          for (int m = 0; m < n; ++m)
            if (m != k && m != j)
              v *= (t - x[m]) / (x[k] - x[m]);
          s += v;
        }
    return s;
  }

  inline Basis1D::Basis1D(const int degree)
    : p(degree)
  {
    if (p < 1 || p > kMaxDegree)
      throw std::invalid_argument("degree must be 1..3");
    const int n = p + 1;
    // Gauss-Lobatto nodes on [0,1] (FE_Q's support points)
    if (p == 1)
      {
        node[0] = 0.;
        node[1] = 1.;
      }
    else if (p == 2)
      {
        node[0] = 0.;
        node[1] = 0.5;
        node[2] = 1.;
      }
    else
      {
        const double s = 1. / std::sqrt(5.);
        node[0]        = 0.;
        node[1]        = 0.5 * (1. - s);
        node[2]        = 0.5 * (1. + s);
        node[3]        = 1.;
      }
    // Gauss-Legendre with n points on [0,1]: Golub-Welsch would do; Newton on P_n is enough here
    for (int i = 0; i < n; ++i)
      {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1.;
        for (int it = 0; it < 200; ++it)
          {
            double p0 = 1., p1 = z;
            for (int k2 = 2; k2 <= n; ++k2)
              {
                const double p2 = ((2. * k2 - 1.) * z * p1 - (k2 - 1.) * p0) / k2;
                p0              = p1;
                p1              = p2;
              }
            dp             = n * (z * p1 - p0) / (z * z - 1.);
            const double d = p1 / dp;
            z -= d;
            if (std::fabs(d) < 1e-16)
              break;
          }
        double p0 = 1., p1 = z;
        for (int k2 = 2; k2 <= n; ++k2)
          {
            const double p2 = ((2. * k2 - 1.) * z * p1 - (k2 - 1.) * p0) / k2;
            p0              = p1;
            p1              = p2;
          }
        dp               = n * (z * p1 - p0) / (z * z - 1.);
        qx[n - 1 - i]    = 0.5 * (z + 1.);
        qw[n - 1 - i]    = 1. / ((1. - z * z) * dp * dp);
      }
    for (int q = 0; q < n; ++q)
      for (int kk = 0; kk < n; ++kk)
        {
          S[q][kk] = lagrange(node, n, kk, qx[q]);
          D[q][kk] = lagrange_derivative(node, n, kk, qx[q]);
        }
    for (int kk = 0; kk < n; ++kk)
      {
        d0[kk] = lagrange_derivative(node, n, kk, 0.);
        d1[kk] = lagrange_derivative(node, n, kk, 1.);
      }
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b)
        {
          double sm = 0., sd = 0., sk = 0.;
          for (int q = 0; q < n; ++q)
            {
              sm += S[q][a] * S[q][b] * qw[q];
              sd += S[q][a] * D[q][b] * qw[q];
              sk += D[q][a] * D[q][b] * qw[q];
            }
          m[a][b]  = sm;
          dm[a][b] = sd;
          k[a][b]  = sk;
        }
  }

  // One axis of the mesh: vertices, node coordinates and the assembled bands.
  struct Axis1D
  {
    int                 N = 0, p = 1, P = 0; // cells, degree, p N
    bool                periodic = false;
    std::vector<double> vertex, coord;         // N + 1, P + 1
    std::vector<double> M, D, DT, K;           // (P + 1) x (2p + 1) bands
    std::vector<double> row_sum, diagonal;     // of M, (P + 1)
    double              b0[kMaxDegree + 1]{};  // normal derivative at the low end: phi_l'(0) / h_0
    double              b1[kMaxDegree + 1]{};  // at the high end: phi_l'(1) / h_{N-1}

    Axis1D(const Basis1D &B, const std::vector<double> &vertices, const bool periodic_)
      : N(int(vertices.size()) - 1)
      , p(B.p)
      , P(B.p * (int(vertices.size()) - 1))
      , periodic(periodic_)
      , vertex(vertices)
    {
      const int w = 2 * p + 1;
      coord.assign(P + 1, 0.);
      M.assign(size_t(P + 1) * w, 0.);
      D = DT = K = M;
      for (int c = 0; c < N; ++c)
        {
          const double h = vertex[c + 1] - vertex[c];
          for (int a = 0; a <= p; ++a)
            {
              coord[p * c + a] = vertex[c] + h * B.node[a];
              int row          = p * c + a;
              if (periodic && row == P)
                row = 0;
              for (int b = 0; b <= p; ++b)
                {
                  const size_t o = size_t(row) * w + (b - a + p);
                  M[o] += h * B.m[a][b];
                  D[o] += B.dm[a][b];
                  DT[o] += B.dm[b][a];
                  K[o] += B.k[a][b] / h;
                }
            }
        }
      coord[P] = vertex[N];
      row_sum.assign(P + 1, 0.);
      diagonal.assign(P + 1, 0.);
      for (int i = 0; i <= P; ++i)
        {
          double s = 0.;
          for (int o = 0; o < w; ++o)
            s += M[size_t(i) * w + o];
          row_sum[i]  = s;
          diagonal[i] = M[size_t(i) * w + p];
        }
      if (periodic) // the slave row: an identity row of the constrained system
        row_sum[P] = diagonal[P] = 1.;
      const double h0 = vertex[1] - vertex[0], h1 = vertex[N] - vertex[N - 1];
      for (int l = 0; l <= p; ++l)
        {
          b0[l] = B.d0[l] / h0;
          b1[l] = B.d1[l] / h1;
        }
    }
  };

  // Vertices of [0, L] in N cells, optionally moved by the tanh clustering
  // x -> L/2 (1 + tanh(gamma (2x/L - 1)) / tanh(gamma))
  // (include/lbfem/test_cases.h:215-225).
  inline std::vector<double>
  box_vertices(const int N, const double L, const double stretch)
  {
    std::vector<double> v(N + 1);
    for (int c = 0; c <= N; ++c)
      {
        const double x = L * c / N;
        v[c]           = stretch > 0. ? 0.5 * L * (1. + std::tanh(stretch * (2. * x / L - 1.)) / std::tanh(stretch)) : x;
      }
    return v;
  }
} // namespace lbfem_amd
