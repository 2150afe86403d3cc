// CPU oracle of the D2Q9 discrete-Boltzmann step on continuous Q_p elements.
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP path. It restates the
// reference algorithm (ivan-pi/lbm-fem, include/lbfem/*.h) without deal.II:
// FE_Q(p) on Gauss-Lobatto nodes, QGauss(p+1) cell and face quadrature with
// sum factorisation as FEEvaluation does it, the periodic constraints as
// read/distribute maps, deal.II's SolverCG iteration and stopping test, the
// two schemes and the three flows. Every function cites the reference lines
// it follows. Single-threaded; cell loop in lexicographic cell order.
//
// Restatement limits (see the oracle section of the design notes): deal.II's
// DoF order, vector-reduction summation order and p4est cell order are not
// reproduced, so agreement with a deal.II run is to round-off, not bitwise.
#include "lbfem_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <optional>
#include <vector>

namespace
{
  constexpr int    Q   = 9;
  constexpr int    NM  = 8; // moving populations
  constexpr double cs2 = 1. / 3.;
  constexpr double PI  = 3.14159265358979323846;

  // include/lbfem/d2q9.h:18-22
  constexpr double EX[Q] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
  constexpr double EY[Q] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
  constexpr double W[Q]  = {4. / 9., 1. / 9., 1. / 36., 1. / 9., 1. / 36., 1. / 9., 1. / 36., 1. / 9., 1. / 36.};

  struct Moments
  {
    double rho, ux, uy;
  };

  // include/lbfem/d2q9.h:30-41
  Moments
  moments(const double *f)
  {
    double rho = 0, mx = 0, my = 0;
    for (int a = 0; a < Q; ++a)
      {
        rho += f[a];
        mx += EX[a] * f[a];
        my += EY[a] * f[a];
      }
    return {rho, mx / rho, my / rho};
  }

  // include/lbfem/d2q9.h:44-55
  void
  equilibrium(const Moments &m, double *feq)
  {
    const double uu = m.ux * m.ux + m.uy * m.uy;
    for (int a = 0; a < Q; ++a)
      {
        const double eu = EX[a] * m.ux + EY[a] * m.uy;
        feq[a]          = W[a] * m.rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * uu);
      }
  }

  // ---------------------------------------------------------------- 1-D FE
  // Gauss-Legendre points on [0,1] (QGauss<1>(n), discretization.h:121).
  void
  gauss(const int n, std::vector<double> &x, std::vector<double> &w)
  {
    x.assign(n, 0.);
    w.assign(n, 0.);
    for (int i = 0; i < n; ++i)
      {
        double z = std::cos(PI * (i + 0.75) / (n + 0.5)), dp = 1.;
        for (int it = 0; it < 100; ++it)
          {
            double p0 = 1., p1 = z; // P_{n-1}, P_n by the three-term recurrence
            for (int k = 2; k <= n; ++k)
              {
                const double p2 = ((2. * k - 1.) * z * p1 - (k - 1.) * p0) / k;
                p0              = p1;
                p1              = p2;
              }
            dp              = n * (z * p1 - p0) / (z * z - 1.);
            const double dz = p1 / dp;
            z -= dz;
            if (std::abs(dz) < 1e-15)
              break;
          }
        double p0 = 1., p1 = z;
        for (int k = 2; k <= n; ++k)
          {
            const double p2 = ((2. * k - 1.) * z * p1 - (k - 1.) * p0) / k;
            p0              = p1;
            p1              = p2;
          }
        dp           = n * (z * p1 - p0) / (z * z - 1.);
        x[n - 1 - i] = 0.5 * (z + 1.);
        w[n - 1 - i] = 1. / ((1. - z * z) * dp * dp); // 2/((1-z^2) P_n'^2) on [-1,1], halved on [0,1]
      }
  }

  // Gauss-Lobatto support points of FE_Q(p) on [0,1] (docs/method.md:299-302).
  std::vector<double>
  gll(const int p)
  {
    switch (p)
      {
        case 1:
          return {0., 1.};
        case 2:
          return {0., 0.5, 1.};
        case 3:
          {
            const double s = 1. / std::sqrt(5.);
            return {0., 0.5 * (1. - s), 0.5 * (1. + s), 1.};
          }
        case 4:
          {
            const double s = std::sqrt(3. / 7.);
            return {0., 0.5 * (1. - s), 0.5, 0.5 * (1. + s), 1.};
          }
        default:
          return {};
      }
  }

  double
  lagrange(const std::vector<double> &xn, const int k, const double x)
  {
    double v = 1.;
    for (std::size_t m = 0; m < xn.size(); ++m)
      if ((int)m != k)
        v *= (x - xn[m]) / (xn[k] - xn[m]);
    return v;
  }

  double
  lagrange_d(const std::vector<double> &xn, const int k, const double x)
  {
    double s = 0.;
    for (std::size_t j = 0; j < xn.size(); ++j)
      {
        if ((int)j == k)
          continue;
        double t = 1. / (xn[k] - xn[j]);
        for (std::size_t m = 0; m < xn.size(); ++m)
          if ((int)m != k && m != j)
            t *= (x - xn[m]) / (xn[k] - xn[m]);
        s += t;
      }
    return s;
  }

  // The 1-D tables FEEvaluation works with: shape values and reference
  // derivatives at the Gauss points, and at the two end points (faces).
  struct Basis1D
  {
    int                 p = 1, n = 2, nq = 2;
    std::vector<double> node, qx, qw;
    std::vector<double> S, D;    // [q * n + k]
    std::vector<double> d0, d1;  // phi_k'(0), phi_k'(1)

    explicit Basis1D(const int p_)
      : p(p_)
      , n(p_ + 1)
      , nq(p_ + 1)
    {
      node = gll(p);
      gauss(nq, qx, qw);
      S.resize(nq * n);
      D.resize(nq * n);
      for (int q = 0; q < nq; ++q)
        for (int k = 0; k < n; ++k)
          {
            S[q * n + k] = lagrange(node, k, qx[q]);
            D[q * n + k] = lagrange_d(node, k, qx[q]);
          }
      d0.resize(n);
      d1.resize(n);
      for (int k = 0; k < n; ++k)
        {
          d0[k] = lagrange_d(node, k, 0.);
          d1[k] = lagrange_d(node, k, 1.);
        }
    }
  };

  // ---------------------------------------------------------------- test cases
  // include/lbfem/test_case.h:23-83 and include/lbfem/test_cases.h:21-245.
  struct TestCase
  {
    int    id    = ORA_TGV;
    double L     = 1., rho0 = 1., U0 = 0., nu = 0.;
    double n1    = 1, n2 = 4, stretch = 0.;
    double distort = 0.; // tgv: x -> x + eps L sin(2 pi x / L) sin(2 pi y / L) (1, 1)

    int
    n_periodic() const
    {
      return id == ORA_TGV ? 2 : id == ORA_COUETTE ? 1 : 0;
    }
    void
    set_reynolds(const double Re)
    {
      nu = U0 * L / Re; // test_case.h:38-41
    }
    double
    relaxation_time() const
    {
      return nu / cs2;
    }
    double
    k1() const
    {
      return 2. * PI * n1 / L;
    }
    double
    k2() const
    {
      return 2. * PI * n2 / L;
    }
    double
    reference_time() const
    {
      if (id == ORA_TGV) // test_cases.h:86-89
        return 1. / (nu * (k1() * k1() + k2() * k2()));
      if (id == ORA_COUETTE) // test_cases.h:159-162
        return L * L / nu;
      return L / U0; // test_cases.h:241-244
    }
    double
    energy_scale() const
    {
      if (id == ORA_TGV) // test_cases.h:90-95
        {
          const double r = k1() / k2();
          return 0.25 * U0 * U0 * (1. + r * r);
        }
      return U0 * U0; // test_case.h:69-73
    }
    // cavity: tanh wall clustering of the vertices, test_cases.h:215-225
    double
    cluster(const double x) const
    {
      return 0.5 * L * (1. + std::tanh(stretch * (2. * x / L - 1.)) / std::tanh(stretch));
    }
    bool
    transforms() const
    {
      return id == ORA_CAVITY && stretch > 0.;
    }
    // TaylorGreenVortex::transform_mesh (test_cases.h:53-65), applied to a vertex
    void
    distort_vertex(double &x, double &y) const
    {
      if (id != ORA_TGV || distort == 0.)
        return;
      const double d = distort * L * std::sin(2 * PI * x / L) * std::sin(2 * PI * y / L);
      x += d;
      y += d;
    }
    std::optional<std::array<double, 2>>
    wall_velocity(const double x, const double y) const
    {
      if (id == ORA_COUETTE) // test_cases.h:149-157
        {
          if (std::abs(y) < 1e-12 * L)
            return std::array<double, 2>{0., 0.};
          if (std::abs(y - L) < 1e-12 * L)
            return std::array<double, 2>{U0, 0.};
          return {};
        }
      if (id == ORA_CAVITY) // test_cases.h:228-239
        {
          const auto on = [&](const double a, const double a0) { return std::abs(a - a0) < 1e-12 * L; };
          if (on(x, 0.) || on(x, L) || on(y, 0.))
            return std::array<double, 2>{0., 0.};
          if (on(y, L))
            return std::array<double, 2>{U0, 0.};
          return {};
        }
      return {};
    }
    Moments
    exact(const double x, const double y, const double t) const
    {
      if (id == ORA_TGV) // test_cases.h:92-99
        {
          const double F = std::exp(-nu * (k1() * k1() + k2() * k2()) * t), x1 = k1() * x, y2 = k2() * y,
                       r = k1() / k2();
          const double dp = -0.25 * rho0 * U0 * U0 * F * F * (std::cos(2 * x1) + r * r * std::cos(2 * y2));
          return {rho0 + dp / cs2, -U0 * F * std::cos(x1) * std::sin(y2), U0 * r * F * std::sin(x1) * std::cos(y2)};
        }
      if (id == ORA_COUETTE) // test_cases.h:163-174
        {
          double u = U0 * y / L;
          for (int m = 1; m <= 400; ++m)
            {
              const double lm = m * PI / L;
              u += 2. * U0 * (m % 2 ? -1. : 1.) / (lm * L) * std::exp(-nu * lm * lm * t) * std::sin(lm * y);
            }
          return {rho0, u, 0.};
        }
      return {rho0, 0., 0.}; // test_case.h:64-68
    }
    void
    initial_populations(const double x, const double y, const double lambda, double *f) const
    {
      if (id == ORA_TGV) // test_cases.h:104-118
        {
          const Moments m = exact(x, y, 0.);
          equilibrium(m, f);
          const double x1 = k1() * x, y2 = k2() * y;
          const double ux = U0 * k1() * std::sin(x1) * std::sin(y2);
          const double sh = U0 * (k1() * k1() - k2() * k2()) / k2() * std::cos(x1) * std::cos(y2);
          for (int a = 0; a < Q; ++a)
            f[a] -= lambda / cs2 * W[a] * m.rho * ((EX[a] * EX[a] - EY[a] * EY[a]) * ux + EX[a] * EY[a] * sh);
          return;
        }
      equilibrium({rho0, 0., 0.}, f); // test_cases.h:175-178, test_case.h:76-80
    }
  };
} // namespace

// ------------------------------------------------------------------ discretization
// include/lbfem/discretization.h:69-189 on a structured N x N box.
struct ora_disc
{
  TestCase tc;
  int      N = 0, p = 1, np = 2; // cells per direction, degree, nodes per direction (pN + 1)
  int      n = 0;                // nodes incl. slaves
  bool     per[2] = {false, false};
  Basis1D  B{1};
  double   h_min = 0.;

  std::vector<double> xv, yv; // vertices per direction (before a distortion)
  std::vector<double> xn, yn; // node coordinates per direction (tensor meshes)
  // General quadrilaterals (a distorted mesh, or forced): the cells are the
  // MappingQ1 images of their four vertices, with the Jacobian evaluated per
  // quadrature point (discretization.h:121, 179). VX, VY: (N+1)^2 vertices.
  bool                general = false;
  std::vector<double> VX, VY;
  std::vector<double> XN, YN; // node coordinates, per node
  std::vector<double> ml, diag;
  std::vector<int>    constrained;
  std::vector<int>    wall_node;
  std::vector<double> wall_vel;
  std::vector<int>    wall_of; // node -> wall index or -1

  explicit ora_disc(const int Np, const int deg)
    : N(Np)
    , p(deg)
    , np(deg * Np + 1)
    , B(deg)
  {}

  // the value a cell reads at node line i: periodic slaves resolve to the
  // master at 0 (read_dof_values with constraints); scatter uses the same map.
  int
  mx(const int i) const
  {
    return (per[0] && i == p * N) ? 0 : i;
  }
  int
  my(const int j) const
  {
    return (per[1] && j == p * N) ? 0 : j;
  }
  int
  cell_node(const int cx, const int cy, const int kx, const int ky) const
  {
    return my(p * cy + ky) * np + mx(p * cx + kx);
  }
  double
  hx(const int cx) const
  {
    return xv[cx + 1] - xv[cx];
  }
  double
  hy(const int cy) const
  {
    return yv[cy + 1] - yv[cy];
  }
  // The Jacobian of the bilinear map of cell (cx, cy) at (xi, eta) in [0,1]^2,
  // J[d][e] = dx_d / dxi_e, and its inverse K[e][d] = dxi_e / dx_d and det J.
  void
  jacobian(const int cx, const int cy, const double xi, const double eta, double K[2][2], double &det) const
  {
    const int    v00 = cy * (N + 1) + cx, v10 = v00 + 1, v01 = v00 + N + 1, v11 = v01 + 1;
    const double xxi  = (1. - eta) * (VX[v10] - VX[v00]) + eta * (VX[v11] - VX[v01]);
    const double yxi  = (1. - eta) * (VY[v10] - VY[v00]) + eta * (VY[v11] - VY[v01]);
    const double xeta = (1. - xi) * (VX[v01] - VX[v00]) + xi * (VX[v11] - VX[v10]);
    const double yeta = (1. - xi) * (VY[v01] - VY[v00]) + xi * (VY[v11] - VY[v10]);
    det               = xxi * yeta - xeta * yxi;
    const double id   = 1. / det;
    K[0][0]           = yeta * id;
    K[0][1]           = -xeta * id;
    K[1][0]           = -yxi * id;
    K[1][1]           = xxi * id;
  }
  // the bilinear map of cell (cx, cy) (MappingQ1::transform_unit_to_real_cell)
  void
  map_point(const int cx, const int cy, const double xi, const double eta, double &x, double &y) const
  {
    const int v00 = cy * (N + 1) + cx, v10 = v00 + 1, v01 = v00 + N + 1, v11 = v01 + 1;
    x = (1. - eta) * ((1. - xi) * VX[v00] + xi * VX[v10]) + eta * ((1. - xi) * VX[v01] + xi * VX[v11]);
    y = (1. - eta) * ((1. - xi) * VY[v00] + xi * VY[v10]) + eta * ((1. - xi) * VY[v01] + xi * VY[v11]);
  }
  void
  gather(const double *v, const int cx, const int cy, double *G) const
  {
    for (int ky = 0; ky <= p; ++ky)
      for (int kx = 0; kx <= p; ++kx)
        G[ky * (p + 1) + kx] = v[cell_node(cx, cy, kx, ky)];
  }
  // The cells around node line i along one direction and the node's local
  // index in each, left cell first: a vertex line belongs to two cells (for the
  // periodic master at 0, one of them across the seam), other lines to one.
  int
  around(const int i, const bool periodic, int *c, int *k) const
  {
    const int P = p * N;
    int       m = 0;
    if (i % p == 0)
      {
        const int cl = i / p - 1;
        if (cl >= 0)
          {
            c[m]   = cl;
            k[m++] = p;
          }
        else if (periodic)
          {
            c[m]   = N - 1;
            k[m++] = p;
          }
        if (i < P)
          {
            c[m]   = i / p;
            k[m++] = 0;
          }
      }
    else
      {
        c[m]   = i / p;
        k[m++] = i % p;
      }
    return m;
  }
  // distribute_local_to_global of all cells at once: out[node] = the sum of
  // the local contributions Rc[(cy N + cx)(p+1)^2 + ky (p+1) + kx] of the
  // cells around the node, in an order fixed relative to the node (lower cell
  // row first, left cell first), so that the sum is the same at every node of a
  // translation-invariant field, the periodic seam included. Constrained
  // (slave) entries get nothing.
  void
  assemble(const double *Rc, double *out) const
  {
    const int nn = p + 1;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < np; ++j)
      {
        int       cxs[2], kxs[2], cys[2], kys[2];
        const int my_ = around(j, per[1], cys, kys);
        for (int i = 0; i < np; ++i)
          {
            const int g = j * np + i;
            if (constrained[g])
              {
                out[g] = 0.;
                continue;
              }
            const int mx_ = around(i, per[0], cxs, kxs);
            double    s   = 0.;
            for (int b = 0; b < my_; ++b)
              for (int a = 0; a < mx_; ++a)
                s += Rc[(cys[b] * N + cxs[a]) * nn * nn + kys[b] * nn + kxs[a]];
            out[g] = s;
          }
      }
  }
  // Adds the wall-face contributions, per face side (0 left x = 0, 1 right,
  // 2 bottom y = 0, 3 top) Rf[side][(cell along the wall)(p+1) + k], in the
  // order bottom, top, left, right and along each wall lower/left face first.
  void
  add_faces(const double *const *Rf, double *out) const
  {
    const int nn = p + 1;
    int       cs[2], ks[2];
    for (const int side : {2, 3, 0, 1})
      {
        if (!Rf[side])
          continue;
        for (int t = 0; t < np; ++t)
          {
            const bool along_x = side >= 2;
            const int  m       = around(t, per[along_x ? 0 : 1], cs, ks);
            const int  edge    = (side == 1 || side == 3) ? p * N : 0;
            const int  g       = along_x ? edge * np + t : t * np + edge;
            if (constrained[g])
              continue;
            double s = out[g];
            for (int a = 0; a < m; ++a)
              s += Rf[side][cs[a] * nn + ks[a]];
            out[g] = s;
          }
      }
  }

  void reinit(bool force_general);
};

void
ora_disc::reinit(const bool force_general)
{
  // discretization.h:81-100: subdivided_hyper_cube(N, 0, L), periodic faces, transform
  const double L = tc.L;
  xv.resize(N + 1);
  yv.resize(N + 1);
  for (int c = 0; c <= N; ++c)
    xv[c] = yv[c] = L * c / N;
  h_min = L / N;
  if (tc.transforms())
    {
      for (int c = 0; c <= N; ++c)
        {
          xv[c] = tc.cluster(xv[c]);
          yv[c] = tc.cluster(yv[c]);
        }
      h_min = tc.cluster(h_min);
    }
  const int npd = tc.n_periodic();
  per[0]        = npd >= 1;
  per[1]        = npd >= 2;
  n             = np * np;

  // support points through MappingQ1 (discretization.h:127-135)
  xn.resize(np);
  yn.resize(np);
  for (int c = 0; c < N; ++c)
    for (int k = 0; k <= p; ++k)
      {
        xn[p * c + k] = xv[c] + (xv[c + 1] - xv[c]) * B.node[k];
        yn[p * c + k] = yv[c] + (yv[c + 1] - yv[c]) * B.node[k];
      }
  xn[p * N] = xv[N];
  yn[p * N] = yv[N];

  // general quadrilaterals: the vertices moved by the case's map, nodes through
  // each cell's bilinear map (a node on a cell edge from the lower-left cell)
  general = force_general || (tc.id == ORA_TGV && tc.distort != 0.);
  XN.assign((size_t)np * np, 0.);
  YN.assign((size_t)np * np, 0.);
  if (general)
    {
      VX.assign((size_t)(N + 1) * (N + 1), 0.);
      VY.assign((size_t)(N + 1) * (N + 1), 0.);
      for (int cy = 0; cy <= N; ++cy)
        for (int cx = 0; cx <= N; ++cx)
          {
            double x = xv[cx], y = yv[cy];
            tc.distort_vertex(x, y);
            VX[cy * (N + 1) + cx] = x;
            VY[cy * (N + 1) + cx] = y;
          }
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i)
          {
            const int cx = std::min(i / p, N - 1), cy = std::min(j / p, N - 1);
            map_point(cx, cy, B.node[i - p * cx], B.node[j - p * cy], XN[j * np + i], YN[j * np + i]);
          }
    }
  else
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i)
        {
          XN[j * np + i] = xn[i];
          YN[j * np + i] = yn[j];
        }

  constrained.assign(n, 0);
  for (int j = 0; j < np; ++j)
    for (int i = 0; i < np; ++i)
      if ((per[0] && i == p * N) || (per[1] && j == p * N))
        constrained[j * np + i] = 1;

  // MassOperator::compute_lumped_diagonal / compute_diagonal (discretization.h:123-125):
  // row sums and diagonal of M, constrained entries 1.
  ml.assign(n, 0.);
  diag.assign(n, 0.);
  const int           nn = p + 1;
  std::vector<double> Rl((size_t)N * N * nn * nn), Rd((size_t)N * N * nn * nn);
  for (int cy = 0; cy < N; ++cy)
    for (int cx = 0; cx < N; ++cx)
      {
        const double J = general ? 0. : hx(cx) * hy(cy);
        double       detq[25];
        if (general)
          for (int qy = 0; qy < nn; ++qy)
            for (int qx = 0; qx < nn; ++qx)
              {
                double K[2][2];
                jacobian(cx, cy, B.qx[qx], B.qx[qy], K, detq[qy * nn + qx]);
              }
        for (int ky = 0; ky < nn; ++ky)
          for (int kx = 0; kx < nn; ++kx)
            {
              double sl = 0., sd = 0.;
              for (int qy = 0; qy < nn; ++qy)
                for (int qx = 0; qx < nn; ++qx)
                  {
                    const double phi = B.S[qx * nn + kx] * B.S[qy * nn + ky];
                    const double jxw = (general ? detq[qy * nn + qx] : J) * B.qw[qx] * B.qw[qy];
                    sl += phi * jxw; // sum_l phi_l = 1 at every point
                    sd += phi * phi * jxw;
                  }
              Rl[(size_t)(cy * N + cx) * nn * nn + ky * nn + kx] = sl;
              Rd[(size_t)(cy * N + cx) * nn * nn + ky * nn + kx] = sd;
            }
      }
  assemble(Rl.data(), ml.data());
  assemble(Rd.data(), diag.data());
  for (int i = 0; i < n; ++i)
    if (constrained[i])
      ml[i] = diag[i] = 1.;

  // walls (discretization.h:138-142): every owned node the case puts on a wall
  wall_node.clear();
  wall_vel.clear();
  wall_of.assign(n, -1);
  for (int j = 0; j < np; ++j)
    for (int i = 0; i < np; ++i)
      if (const auto u = tc.wall_velocity(XN[j * np + i], YN[j * np + i]))
        {
          wall_of[j * np + i] = (int)wall_node.size();
          wall_node.push_back(j * np + i);
          wall_vel.push_back((*u)[0]);
          wall_vel.push_back((*u)[1]);
        }
}

namespace
{
  // ---------------------------------------------------------------- nodal ops
  // include/lbfem/collision.h:78-102: walls impose their velocity on the
  // moments; applied node by node, walls from the inputs as they are.
  Moments
  impose(const ora_disc &d, const int i, Moments m)
  {
    const int k = d.wall_of[i];
    if (k >= 0)
      {
        m.ux = d.wall_vel[2 * k];
        m.uy = d.wall_vel[2 * k + 1];
      }
    return m;
  }

  // ---------------------------------------------------------------- cell kernels
  // Gradients of a cell field at the Gauss points (FEEvaluation::evaluate(gradients)),
  // reference derivatives divided by the cartesian Jacobian.
  void
  cell_gradients(const ora_disc &d, const double *G, const double hx, const double hy, double *gx, double *gy)
  {
    const int      n = d.p + 1;
    const Basis1D &B = d.B;
    double         Ax[25], Bx[25];
    for (int ky = 0; ky < n; ++ky)
      for (int qx = 0; qx < n; ++qx)
        {
          double a = 0., b = 0.;
          for (int kx = 0; kx < n; ++kx)
            {
              a += B.D[qx * n + kx] * G[ky * n + kx];
              b += B.S[qx * n + kx] * G[ky * n + kx];
            }
          Ax[ky * n + qx] = a;
          Bx[ky * n + qx] = b;
        }
    for (int qy = 0; qy < n; ++qy)
      for (int qx = 0; qx < n; ++qx)
        {
          double a = 0., b = 0.;
          for (int ky = 0; ky < n; ++ky)
            {
              a += B.S[qy * n + ky] * Ax[ky * n + qx];
              b += B.D[qy * n + ky] * Bx[ky * n + qx];
            }
          gx[qy * n + qx] = a / hx;
          gy[qy * n + qx] = b / hy;
        }
  }

  // r_loc = sum_q [phi v + dphi/dx Fx + dphi/dy Fy] (FEEvaluation::integrate(values|gradients));
  // v, Fx, Fy already multiplied by JxW.
  void
  cell_integrate(const ora_disc &d, const double *v, const double *Fx, const double *Fy, const double hx,
                 const double hy, double *R)
  {
    const int      n = d.p + 1;
    const Basis1D &B = d.B;
    double         Vx[25], Wx[25];
    for (int qy = 0; qy < n; ++qy)
      for (int kx = 0; kx < n; ++kx)
        {
          double a = 0., b = 0.;
          for (int qx = 0; qx < n; ++qx)
            {
              a += B.S[qx * n + kx] * v[qy * n + qx] + B.D[qx * n + kx] * Fx[qy * n + qx] / hx;
              b += B.S[qx * n + kx] * Fy[qy * n + qx] / hy;
            }
          Vx[qy * n + kx] = a;
          Wx[qy * n + kx] = b;
        }
    for (int ky = 0; ky < n; ++ky)
      for (int kx = 0; kx < n; ++kx)
        {
          double a = 0.;
          for (int qy = 0; qy < n; ++qy)
            a += B.S[qy * n + ky] * Vx[qy * n + kx] + B.D[qy * n + ky] * Wx[qy * n + kx];
          R[ky * n + kx] = a;
        }
  }

  // General quadrilaterals: per quadrature point the inverse Jacobian K[e][d]
  // = dxi_e / dx_d and JxW = det J w_qx w_qy of the cell's bilinear map
  // (MappingQ1 with the Jacobians stored per point, docs/method.md:216-224).
  struct CellGeo
  {
    double K[25][2][2];
    double jxw[25];
  };
  void
  cell_geometry(const ora_disc &d, const int cx, const int cy, CellGeo &g)
  {
    const int n = d.p + 1;
    for (int qy = 0; qy < n; ++qy)
      for (int qx = 0; qx < n; ++qx)
        {
          const int q = qy * n + qx;
          double    det;
          d.jacobian(cx, cy, d.B.qx[qx], d.B.qx[qy], g.K[q], det);
          g.jxw[q] = det * d.B.qw[qx] * d.B.qw[qy];
        }
  }
  // gradients at the Gauss points: reference derivatives, then grad = K^T grad_xi
  void
  cell_gradients_general(const ora_disc &d, const double *G, const CellGeo &g, double *gx, double *gy)
  {
    const int      n = d.p + 1;
    const Basis1D &B = d.B;
    double         Ax[25], Bx[25];
    for (int ky = 0; ky < n; ++ky)
      for (int qx = 0; qx < n; ++qx)
        {
          double a = 0., b = 0.;
          for (int kx = 0; kx < n; ++kx)
            {
              a += B.D[qx * n + kx] * G[ky * n + kx];
              b += B.S[qx * n + kx] * G[ky * n + kx];
            }
          Ax[ky * n + qx] = a;
          Bx[ky * n + qx] = b;
        }
    for (int qy = 0; qy < n; ++qy)
      for (int qx = 0; qx < n; ++qx)
        {
          double uxi = 0., ueta = 0.;
          for (int ky = 0; ky < n; ++ky)
            {
              uxi += B.S[qy * n + ky] * Ax[ky * n + qx];
              ueta += B.D[qy * n + ky] * Bx[ky * n + qx];
            }
          const int q = qy * n + qx;
          gx[q]       = g.K[q][0][0] * uxi + g.K[q][1][0] * ueta;
          gy[q]       = g.K[q][0][1] * uxi + g.K[q][1][1] * ueta;
        }
  }
  // r_loc = sum_q [phi v + grad phi . F] with grad phi . F = grad_xi phi . (K F);
  // v, Fx, Fy already multiplied by JxW
  void
  cell_integrate_general(const ora_disc &d, const double *v, const double *Fx, const double *Fy, const CellGeo &g,
                         double *R)
  {
    const int      n = d.p + 1;
    const Basis1D &B = d.B;
    double         Fxi[25], Feta[25], Vx[25], Wx[25];
    for (int q = 0; q < n * n; ++q)
      {
        Fxi[q]  = g.K[q][0][0] * Fx[q] + g.K[q][0][1] * Fy[q];
        Feta[q] = g.K[q][1][0] * Fx[q] + g.K[q][1][1] * Fy[q];
      }
    for (int qy = 0; qy < n; ++qy)
      for (int kx = 0; kx < n; ++kx)
        {
          double a = 0., b = 0.;
          for (int qx = 0; qx < n; ++qx)
            {
              a += B.S[qx * n + kx] * v[qy * n + qx] + B.D[qx * n + kx] * Fxi[qy * n + qx];
              b += B.S[qx * n + kx] * Feta[qy * n + qx];
            }
          Vx[qy * n + kx] = a;
          Wx[qy * n + kx] = b;
        }
    for (int ky = 0; ky < n; ++ky)
      for (int kx = 0; kx < n; ++kx)
        {
          double a = 0.;
          for (int qy = 0; qy < n; ++qy)
            a += B.S[qy * n + ky] * Vx[qy * n + kx] + B.D[qy * n + ky] * Wx[qy * n + kx];
          R[ky * n + kx] = a;
        }
  }

  // Gradient of a cell field at the Gauss points of one face (FEFaceEvaluation),
  // face f: 0 left (x=0 side of the cell), 1 right, 2 bottom, 3 top.
  void
  face_gradients(const ora_disc &d, const double *G, const int f, const double hx, const double hy, double *gx,
                 double *gy)
  {
    const int      n = d.p + 1;
    const Basis1D &B = d.B;
    for (int q = 0; q < n; ++q)
      {
        double a = 0., b = 0.;
        if (f >= 2) // along x, at eta = 0 or 1
          {
            const int     edge = f == 2 ? 0 : d.p;
            const double *dn   = f == 2 ? B.d0.data() : B.d1.data();
            for (int kx = 0; kx < n; ++kx)
              {
                a += B.D[q * n + kx] * G[edge * n + kx];
                double t = 0.;
                for (int ky = 0; ky < n; ++ky)
                  t += dn[ky] * G[ky * n + kx];
                b += B.S[q * n + kx] * t;
              }
          }
        else // along y, at xi = 0 or 1
          {
            const int     edge = f == 0 ? 0 : d.p;
            const double *dn   = f == 0 ? B.d0.data() : B.d1.data();
            for (int ky = 0; ky < n; ++ky)
              {
                b += B.D[q * n + ky] * G[ky * n + edge];
                double t = 0.;
                for (int kx = 0; kx < n; ++kx)
                  t += dn[kx] * G[ky * n + kx];
                a += B.S[q * n + ky] * t;
              }
          }
        gx[q] = a / hx;
        gy[q] = b / hy;
      }
  }

  // Weak-form advection (include/lbfem/advection.h:94-175, 182-293): rhs has
  // 8 blocks, zeroed; the point operation is TaylorGalerkinAdvection
  // (advection.h:219-241) or LeeLinAdvection (advection.h:263-289). Cell and
  // face integrals per cell, then distributed by ora_disc::assemble/add_faces.
  void
  advection(const ora_disc &d, double *rhs, const double *f, const double *feq, const double *e, const double dt,
            const double lambda, const bool leelin, const bool remove_mass_flux)
  {
    const int n = d.n, nn = d.p + 1, nq = nn * nn, N = d.N;
    double    a, b = 0., c;
    if (leelin)
      {
        b = -0.5 * dt * dt / lambda; // advection.h:268-270
        a = -dt - b;
        c = -0.5 * dt * dt;
      }
    else
      {
        a = -dt; // advection.h:223
        c = -0.5 * dt * dt;
      }

    // wall faces: non-periodic directions only (discretization.h:117-118);
    // side 0 left (x = 0), 1 right, 2 bottom (y = 0), 3 top
    const bool          has[4] = {!d.per[0], !d.per[0], !d.per[1], !d.per[1]};
    const double        nrm[4][2] = {{-1., 0.}, {1., 0.}, {0., -1.}, {0., 1.}};
    std::vector<double> Rf[4][NM];
    std::vector<double> G(nq), H(nq), gx(nq), gy(nq), hgx(nq), hgy(nq), v(nq), Fx(nq), Fy(nq), R(nq);
    std::vector<double> val(NM * nn), fgx(nn), fgy(nn);
    for (int side = 0; side < 4; ++side)
      {
        if (!has[side])
          continue;
        for (int al = 0; al < NM; ++al)
          Rf[side][al].assign((size_t)N * nn, 0.);
        for (int t = 0; t < N; ++t)
          {
            const int    cx = side == 0 ? 0 : side == 1 ? N - 1 : t;
            const int    cy = side == 2 ? 0 : side == 3 ? N - 1 : t;
            const double hx = d.hx(cx), hy = d.hy(cy);
            const double hf = side >= 2 ? hx : hy;
            for (int al = 0; al < NM; ++al)
              {
                const double ex = e[2 * (al + 1)], ey = e[2 * (al + 1) + 1];
                d.gather(f + (al + 1) * n, cx, cy, G.data());
                face_gradients(d, G.data(), side, hx, hy, fgx.data(), fgy.data());
                for (int q = 0; q < nn; ++q) // advection.h:185-192
                  val[al * nn + q] =
                    (0.5 * dt * dt) * (ex * nrm[side][0] + ey * nrm[side][1]) * (ex * fgx[q] + ey * fgy[q]);
              }
            if (remove_mass_flux) // advection.h:194-196
              for (int q = 0; q < nn; ++q)
                {
                  double mf = 0.;
                  for (int al = 0; al < NM; ++al)
                    mf += val[al * nn + q];
                  for (int al = 0; al < NM; ++al)
                    val[al * nn + q] -= (W[al + 1] / (1. - W[0])) * mf;
                }
            for (int al = 0; al < NM; ++al) // FEFaceEvaluation::integrate(values)
              for (int k = 0; k < nn; ++k)
                {
                  double s = 0.;
                  for (int q = 0; q < nn; ++q)
                    s += d.B.S[q * nn + k] * val[al * nn + q] * hf * d.B.qw[q];
                  Rf[side][al][(size_t)t * nn + k] = s;
                }
          }
      }

    std::vector<double> Rc((size_t)N * N * nq);
    for (int al = 0; al < NM; ++al)
      {
        const double ex = e[2 * (al + 1)], ey = e[2 * (al + 1) + 1];
        // cells are independent (each writes its own block of Rc); with OpenMP each
        // thread has its own scratch, and the result is the serial one bit for bit
#pragma omp parallel firstprivate(G, H, gx, gy, hgx, hgy, v, Fx, Fy)
#pragma omp for collapse(2) schedule(static)
        for (int cy = 0; cy < N; ++cy)
          for (int cx = 0; cx < N; ++cx)
            {
              if (d.general) // the same point operation with per-point Jacobians
                {
                  CellGeo g;
                  cell_geometry(d, cx, cy, g);
                  d.gather(f + (al + 1) * n, cx, cy, G.data());
                  cell_gradients_general(d, G.data(), g, gx.data(), gy.data());
                  if (leelin)
                    {
                      d.gather(feq + (al + 1) * n, cx, cy, H.data());
                      cell_gradients_general(d, H.data(), g, hgx.data(), hgy.data());
                    }
                  for (int q = 0; q < nq; ++q)
                    {
                      const double jxw = g.jxw[q];
                      const double s   = ex * gx[q] + ey * gy[q];
                      double       vv  = a * s;
                      if (leelin)
                        vv += b * (ex * hgx[q] + ey * hgy[q]);
                      v[q]  = vv * jxw;
                      Fx[q] = (c * ex) * s * jxw;
                      Fy[q] = (c * ey) * s * jxw;
                    }
                  cell_integrate_general(d, v.data(), Fx.data(), Fy.data(), g, Rc.data() + (size_t)(cy * N + cx) * nq);
                  continue;
                }
              const double hx = d.hx(cx), hy = d.hy(cy);
              d.gather(f + (al + 1) * n, cx, cy, G.data());
              cell_gradients(d, G.data(), hx, hy, gx.data(), gy.data());
              if (leelin)
                {
                  d.gather(feq + (al + 1) * n, cx, cy, H.data());
                  cell_gradients(d, H.data(), hx, hy, hgx.data(), hgy.data());
                }
              for (int qy = 0; qy < nn; ++qy)
                for (int qx = 0; qx < nn; ++qx)
                  {
                    const int    q   = qy * nn + qx;
                    const double jxw = hx * hy * d.B.qw[qx] * d.B.qw[qy];
                    const double s   = ex * gx[q] + ey * gy[q];
                    double       vv  = a * s;
                    if (leelin)
                      vv += b * (ex * hgx[q] + ey * hgy[q]);
                    v[q]  = vv * jxw;
                    Fx[q] = (c * ex) * s * jxw;
                    Fy[q] = (c * ey) * s * jxw;
                  }
              cell_integrate(d, v.data(), Fx.data(), Fy.data(), hx, hy, Rc.data() + (size_t)(cy * N + cx) * nq);
            }
        d.assemble(Rc.data(), rhs + al * n);
        const double *faces[4];
        for (int side = 0; side < 4; ++side)
          faces[side] = has[side] ? Rf[side][al].data() : nullptr;
        d.add_faces(faces, rhs + al * n);
      }
  }

  // y = (M + c K_e) x with constrained rows the identity
  // (include/lbfem/mass.h:105-144; MatrixFreeOperators::Base::vmult).
  void
  streaming_vmult(const ora_disc &d, double *y, const double *x, const double c, const double ex, const double ey)
  {
    const int           n = d.n, nn = d.p + 1, nq = nn * nn, N = d.N;
    std::vector<double> G(nq), u(nq), gx(nq), gy(nq), v(nq), Fx(nq), Fy(nq), T(nq);
    std::vector<double> Rc((size_t)N * N * nq);
    const Basis1D      &B = d.B;
    // cells are independent (each writes its own block of Rc); with OpenMP each
    // thread has its own scratch, and the result is the serial one bit for bit
#pragma omp parallel for collapse(2) schedule(static) firstprivate(G, u, gx, gy, v, Fx, Fy, T)
    for (int cy = 0; cy < N; ++cy)
      for (int cx = 0; cx < N; ++cx)
        {
          const double hx = d.hx(cx), hy = d.hy(cy);
          d.gather(x, cx, cy, G.data());
          // values at the quadrature points
          for (int ky = 0; ky < nn; ++ky)
            for (int qx = 0; qx < nn; ++qx)
              {
                double s = 0.;
                for (int kx = 0; kx < nn; ++kx)
                  s += B.S[qx * nn + kx] * G[ky * nn + kx];
                T[ky * nn + qx] = s;
              }
          for (int qy = 0; qy < nn; ++qy)
            for (int qx = 0; qx < nn; ++qx)
              {
                double s = 0.;
                for (int ky = 0; ky < nn; ++ky)
                  s += B.S[qy * nn + ky] * T[ky * nn + qx];
                u[qy * nn + qx] = s;
              }
          CellGeo g;
          if (d.general)
            cell_geometry(d, cx, cy, g);
          if (c != 0.)
            {
              if (d.general)
                cell_gradients_general(d, G.data(), g, gx.data(), gy.data());
              else
                cell_gradients(d, G.data(), hx, hy, gx.data(), gy.data());
            }
          for (int qy = 0; qy < nn; ++qy)
            for (int qx = 0; qx < nn; ++qx)
              {
                const int    q   = qy * nn + qx;
                const double jxw = d.general ? g.jxw[q] : hx * hy * B.qw[qx] * B.qw[qy];
                v[q]             = u[q] * jxw;
                if (c != 0.)
                  {
                    const double eg = ex * gx[q] + ey * gy[q];
                    Fx[q]           = (c * ex) * eg * jxw;
                    Fy[q]           = (c * ey) * eg * jxw;
                  }
                else
                  Fx[q] = Fy[q] = 0.;
              }
          if (d.general)
            cell_integrate_general(d, v.data(), Fx.data(), Fy.data(), g, Rc.data() + (size_t)(cy * N + cx) * nq);
          else
            cell_integrate(d, v.data(), Fx.data(), Fy.data(), hx, hy, Rc.data() + (size_t)(cy * N + cx) * nq);
        }
    d.assemble(Rc.data(), y);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      if (d.constrained[i])
        y[i] = x[i];
  }

  double
  dot(const double *a, const double *b, const int n)
  {
    double s = 0.;
    for (int i = 0; i < n; ++i)
      s += a[i] * b[i];
    return s;
  }

  // deal.II SolverCG (IterationWorker) with Jacobi preconditioner, SolverControl
  // with tolerance tol * |b| and max_its (include/lbfem/mass.h:204-211):
  //   r = A x - b (r = -b if x == 0); res = |r|; check(0)
  //   loop: v = P r; rho = r.v; p = (it > 1 ? beta p : 0) - v; v = A p;
  //         alpha = rho / p.v; x += alpha p; r += alpha v; res = |r|; check(it)
  // A(y, x) is the matrix-vector product, diag the matrix diagonal (1 on
  // constrained rows), applied as DiagonalMatrix::vmult of the inverse entries.
  template <class Vmult>
  int
  cg_core(const int n, const double *diag, Vmult &&A, double *x, const double *b, const double tol_rel,
          const int max_its, int *its)
  {
    const double        tol = tol_rel * std::sqrt(dot(b, b, n));
    std::vector<double> r(n), v(n), p(n);
    bool                zero = true;
    for (int i = 0; i < n && zero; ++i)
      zero = x[i] == 0.;
    if (!zero)
      {
        A(r.data(), x);
        for (int i = 0; i < n; ++i)
          r[i] -= b[i];
      }
    else
      for (int i = 0; i < n; ++i)
        r[i] = -b[i];
    double res = std::sqrt(dot(r.data(), r.data(), n));
    int    it  = 0;
    // SolverControl::check: success first, then failure at max steps
    if (res <= tol)
      {
        *its = 0;
        return ORA_OK;
      }
    if (it >= max_its || std::isnan(res))
      {
        *its = it;
        return ORA_ENOCONV;
      }
    double rho = 0., rho_prev = 0.;
    while (true)
      {
        ++it;
        rho_prev = rho;
        // the element-wise loops run in parallel (independent entries); the dot
        // products stay serial, in index order
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) // DiagonalMatrix::vmult with the inverse entries 1/diag
          v[i] = (1. / diag[i]) * r[i];
        rho = dot(r.data(), v.data(), n);
        if (it > 1)
          {
            const double beta = rho / rho_prev;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i)
              p[i] = beta * p[i] - v[i];
          }
        else
#pragma omp parallel for schedule(static)
          for (int i = 0; i < n; ++i)
            p[i] = -v[i];
        A(v.data(), p.data());
        const double pAp   = dot(p.data(), v.data(), n);
        const double alpha = rho / pAp;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
          {
            x[i] += alpha * p[i];
            r[i] += alpha * v[i];
          }
        const double rr = dot(r.data(), r.data(), n);
        res = std::sqrt(std::abs(rr));
        if (res <= tol)
          {
            *its = it;
            return ORA_OK;
          }
        if (it >= max_its || std::isnan(res))
          {
            *its = it;
            return ORA_ENOCONV;
          }
      }
  }

  int
  cg_solve(const ora_disc &d, double *x, const double *b, const double tol_rel, const int max_its, const double c,
           const double ex, const double ey, int *its)
  {
    return cg_core(
      d.n, d.diag.data(), [&](double *y, const double *v) { streaming_vmult(d, y, v, c, ex, ey); }, x, b, tol_rel,
      max_its, its);
  }

  // include/lbfem/mass.h:184-213
  int
  mass_solve(const ora_disc &d, double *x, const double *r, const int type, const int richardson, const double cg_tol,
             const int cg_max, const double tg3, const double ex, const double ey, int *its)
  {
    const int n = d.n;
    *its        = 0;
    if (type == ORA_MASS_LUMPED) // mass.h:189-191
      {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
          x[i] = r[i] / d.ml[i];
        return ORA_OK;
      }
    if (type == ORA_MASS_RICHARDSON) // mass.h:193-202: PreconditionRelaxation(omega 1, k+1 its, M_L^-1)
      {
        std::vector<double> Ax(n);
        for (int i = 0; i < n; ++i)
          x[i] = r[i] / d.ml[i];
        for (int k = 0; k < richardson; ++k)
          {
            streaming_vmult(d, Ax.data(), x, tg3, ex, ey);
            for (int i = 0; i < n; ++i)
              x[i] += (r[i] - Ax[i]) / d.ml[i];
          }
        return ORA_OK;
      }
    return cg_solve(d, x, r, cg_tol, cg_max, tg3, ex, ey, its);
  }

  // ---------------------------------------------------------------- stream function
  // include/lbfem/io.h:31-93 (compute_stream_function) on a walled box: the
  // zero boundary constraints make the wall nodes constrained (read as zero by
  // the cells, identity rows, nothing distributed to them).
  bool
  on_wall(const ora_disc &d, const int g)
  {
    const int i = g % d.np, j = g / d.np, P = d.p * d.N;
    return i == 0 || j == 0 || i == P || j == P;
  }

  // values of a cell field at the Gauss points (FEEvaluation::evaluate(values))
  void
  cell_values(const ora_disc &d, const double *G, double *u)
  {
    const int      nn = d.p + 1;
    const Basis1D &B  = d.B;
    double         T[25];
    for (int ky = 0; ky < nn; ++ky)
      for (int qx = 0; qx < nn; ++qx)
        {
          double s = 0.;
          for (int kx = 0; kx < nn; ++kx)
            s += B.S[qx * nn + kx] * G[ky * nn + kx];
          T[ky * nn + qx] = s;
        }
    for (int qy = 0; qy < nn; ++qy)
      for (int qx = 0; qx < nn; ++qx)
        {
          double s = 0.;
          for (int ky = 0; ky < nn; ++ky)
            s += B.S[qy * nn + ky] * T[ky * nn + qx];
          u[qy * nn + qx] = s;
        }
  }

  // the local Laplace operator of cell (cx, cy) on the cell values G
  // (MatrixFreeOperators::LaplaceOperator: integrate(gradients) of the gradients times JxW)
  void
  cell_laplace(const ora_disc &d, const int cx, const int cy, const double *G, double *R)
  {
    const int    nn = d.p + 1, nq = nn * nn;
    const double hx = d.hx(cx), hy = d.hy(cy);
    double       gx[25], gy[25], z[25];
    cell_gradients(d, G, hx, hy, gx, gy);
    for (int qy = 0; qy < nn; ++qy)
      for (int qx = 0; qx < nn; ++qx)
        {
          const int    q   = qy * nn + qx;
          const double jxw = hx * hy * d.B.qw[qx] * d.B.qw[qy];
          gx[q] *= jxw;
          gy[q] *= jxw;
        }
    for (int q = 0; q < nq; ++q)
      z[q] = 0.;
    cell_integrate(d, z, gx, gy, hx, hy, R);
  }

  // y = A x: the wall entries of x read as zero, wall rows the identity
  void
  laplace_vmult(const ora_disc &d, double *y, const double *x)
  {
    const int           N = d.N, nn = d.p + 1, nq = nn * nn;
    std::vector<double> Rc((size_t)N * N * nq);
    for (int cy = 0; cy < N; ++cy)
      for (int cx = 0; cx < N; ++cx)
        {
          double G[25];
          for (int ky = 0; ky < nn; ++ky)
            for (int kx = 0; kx < nn; ++kx)
              {
                const int g          = d.cell_node(cx, cy, kx, ky);
                G[ky * nn + kx]      = on_wall(d, g) ? 0. : x[g];
              }
          cell_laplace(d, cx, cy, G, Rc.data() + (size_t)(cy * N + cx) * nq);
        }
    d.assemble(Rc.data(), y);
    for (int g = 0; g < d.n; ++g)
      if (on_wall(d, g))
        y[g] = x[g];
  }

  int
  stream_function(const ora_disc &d, const double *u, const double *v, double *psi, int *its)
  {
    const int           N = d.N, nn = d.p + 1, nq = nn * nn, n = d.n;
    std::vector<double> Rc((size_t)N * N * nq), b(n), diag(n);
    // right-hand side (dphi/dx, v) - (dphi/dy, u) (io.h:61-81)
    for (int cy = 0; cy < N; ++cy)
      for (int cx = 0; cx < N; ++cx)
        {
          const double hx = d.hx(cx), hy = d.hy(cy);
          double       G[25], uq[25], vq[25], Fx[25], Fy[25], z[25];
          d.gather(u, cx, cy, G);
          cell_values(d, G, uq);
          d.gather(v, cx, cy, G);
          cell_values(d, G, vq);
          for (int qy = 0; qy < nn; ++qy)
            for (int qx = 0; qx < nn; ++qx)
              {
                const int    q   = qy * nn + qx;
                const double jxw = hx * hy * d.B.qw[qx] * d.B.qw[qy];
                Fx[q]            = vq[q] * jxw;
                Fy[q]            = -uq[q] * jxw;
                z[q]             = 0.;
              }
          cell_integrate(d, z, Fx, Fy, hx, hy, Rc.data() + (size_t)(cy * N + cx) * nq);
        }
    d.assemble(Rc.data(), b.data());
    // the diagonal, cell by cell on unit vectors (LaplaceOperator::compute_diagonal)
    for (int cy = 0; cy < N; ++cy)
      for (int cx = 0; cx < N; ++cx)
        for (int k = 0; k < nq; ++k)
          {
            double G[25], R[25];
            for (int m = 0; m < nq; ++m)
              G[m] = m == k ? 1. : 0.;
            cell_laplace(d, cx, cy, G, R);
            Rc[(size_t)(cy * N + cx) * nq + k] = R[k];
          }
    d.assemble(Rc.data(), diag.data());
    for (int g = 0; g < n; ++g)
      if (on_wall(d, g))
        {
          b[g]    = 0.;
          diag[g] = 1.;
        }
    // SolverCG(10000, 1e-12 |b|) from psi = 0 with the Jacobi preconditioner (io.h:83-85)
    for (int g = 0; g < n; ++g)
      psi[g] = 0.;
    return cg_core(
      n, diag.data(), [&](double *y, const double *x) { laplace_vmult(d, y, x); }, psi, b.data(), 1e-12, 10000,
      its);
  }
} // namespace

// ------------------------------------------------------------------ schemes
// include/lbfem/streaming.h:38-130 and include/lbfem/schemes.h:31-142.
struct ora_scheme
{
  const ora_disc     *d;
  int                 scheme, streaming, mass_type, richardson, cg_max;
  double              cg_tol;
  double              dt = 0., lambda = 0.;
  std::vector<double> f, feq, rhs, x[2];
  long long           cg_iterations = 0;
  int                 last_its[2][NM] = {}; // the last step's CG steps per sweep and moving population
  std::vector<std::array<double, Q * 2>> sweeps;
  double              tg3     = 0.;
  bool                extrapolate = false;
};

extern "C"
{
  // distort: the TGV mesh distortion eps (test_cases.h:53-65); general: 1 to take
  // the general-quadrilateral path on any mesh (a tensor mesh then gives the
  // tensor path's operators up to round-off)
  ora_disc *
  ora_disc_create2(const int n_cells, const int degree, const int test_case, const double mach, const int mode1,
                   const int mode2, const double stretch, const double distort, const int general)
  {
    if (n_cells < 1 || degree < 1 || degree > 4 || test_case < 0 || test_case > 2)
      return nullptr;
    if (distort != 0. && test_case != ORA_TGV) // the distortion is the vortex's (cgdbe.cc:77)
      return nullptr;
    if (general && test_case != ORA_TGV) // general cells here have no wall faces
      return nullptr;
    ora_disc *d   = new ora_disc(n_cells, degree);
    d->tc.id      = test_case;
    d->tc.U0      = mach * std::sqrt(cs2); // test_case.h:28-30
    d->tc.n1      = mode1;
    d->tc.n2      = mode2;
    d->tc.stretch = stretch;
    d->tc.distort = distort;
    d->reinit(general != 0);
    return d;
  }
  ora_disc *
  ora_disc_create(const int n_cells, const int degree, const int test_case, const double mach, const int mode1,
                  const int mode2, const double stretch)
  {
    return ora_disc_create2(n_cells, degree, test_case, mach, mode1, mode2, stretch, 0., 0);
  }
  void
  ora_disc_destroy(ora_disc *d)
  {
    delete d;
  }
  int
  ora_disc_n_nodes(const ora_disc *d)
  {
    return d->n;
  }
  int
  ora_disc_n_per_dir(const ora_disc *d)
  {
    return d->np;
  }
  int
  ora_disc_n_periodic(const ora_disc *d)
  {
    return d->tc.n_periodic();
  }
  double
  ora_disc_h_min(const ora_disc *d)
  {
    return d->h_min;
  }
  double
  ora_disc_time_step(const ora_disc *d, const double cfl)
  {
    return cfl * d->h_min / (d->p * d->p); // discretization.h:158-162
  }
  void
  ora_disc_nodes(const ora_disc *d, double *xy)
  {
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        {
          xy[2 * (j * d->np + i)]     = d->XN[j * d->np + i];
          xy[2 * (j * d->np + i) + 1] = d->YN[j * d->np + i];
        }
  }
  // the (N+1)^2 vertices of a general mesh (x fastest); 0 for a tensor mesh
  int
  ora_disc_general_vertices(const ora_disc *d, double *x, double *y)
  {
    if (!d->general)
      return 0;
    std::copy(d->VX.begin(), d->VX.end(), x);
    std::copy(d->VY.begin(), d->VY.end(), y);
    return 1;
  }
  void
  ora_disc_vertices(const ora_disc *d, double *x, double *y)
  {
    std::copy(d->xv.begin(), d->xv.end(), x);
    std::copy(d->yv.begin(), d->yv.end(), y);
  }
  void
  ora_disc_lumped_mass(const ora_disc *d, double *ml)
  {
    std::copy(d->ml.begin(), d->ml.end(), ml);
  }
  void
  ora_disc_mass_diagonal(const ora_disc *d, double *diag)
  {
    std::copy(d->diag.begin(), d->diag.end(), diag);
  }
  void
  ora_disc_constrained(const ora_disc *d, int *flag)
  {
    std::copy(d->constrained.begin(), d->constrained.end(), flag);
  }
  int
  ora_disc_n_walls(const ora_disc *d)
  {
    return (int)d->wall_node.size();
  }
  void
  ora_disc_walls(const ora_disc *d, int *nodes, double *velocity)
  {
    std::copy(d->wall_node.begin(), d->wall_node.end(), nodes);
    std::copy(d->wall_vel.begin(), d->wall_vel.end(), velocity);
  }

  void
  ora_case_set_reynolds(ora_disc *d, const double Re)
  {
    d->tc.set_reynolds(Re);
  }
  double
  ora_case_relaxation_time(const ora_disc *d)
  {
    return d->tc.relaxation_time();
  }
  double
  ora_case_reference_time(const ora_disc *d)
  {
    return d->tc.reference_time();
  }
  double
  ora_case_U0(const ora_disc *d)
  {
    return d->tc.U0;
  }
  double
  ora_case_energy_scale(const ora_disc *d)
  {
    return d->tc.energy_scale();
  }
  void
  ora_case_initial_populations(const ora_disc *d, const double lambda, double *f)
  {
    double fi[Q];
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        {
          d->tc.initial_populations(d->XN[j * d->np + i], d->YN[j * d->np + i], lambda, fi);
          for (int a = 0; a < Q; ++a)
            f[a * d->n + j * d->np + i] = fi[a];
        }
  }
  void
  ora_case_exact(const ora_disc *d, const double t, double *rho_u)
  {
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        {
          const Moments m = d->tc.exact(d->XN[j * d->np + i], d->YN[j * d->np + i], t);
          const int     k = j * d->np + i;
          rho_u[3 * k]    = m.rho;
          rho_u[3 * k + 1] = m.ux;
          rho_u[3 * k + 2] = m.uy;
        }
  }

  void
  ora_gll_nodes(const int degree, double *x)
  {
    const auto g = gll(degree);
    std::copy(g.begin(), g.end(), x);
  }

  // One h x h cell: M = int N N^T, C_e = int N (e.grad N^T), K_e = int (e.grad N)(e.grad N^T)
  // (include/lbfem/advection.h:7-10), entries [i * nd + j], lexicographic local order.
  void
  ora_element_matrices(const int degree, const double h, const double ex, const double ey, double *M, double *C,
                       double *K)
  {
    const Basis1D B(degree);
    const int     n = degree + 1, nd = n * n;
    for (int i = 0; i < nd; ++i)
      for (int j = 0; j < nd; ++j)
        {
          const int ix = i % n, iy = i / n, jx = j % n, jy = j / n;
          double    m = 0., c = 0., k = 0.;
          for (int qy = 0; qy < n; ++qy)
            for (int qx = 0; qx < n; ++qx)
              {
                const double jxw = h * h * B.qw[qx] * B.qw[qy];
                const double pi = B.S[qx * n + ix] * B.S[qy * n + iy], pj = B.S[qx * n + jx] * B.S[qy * n + jy];
                const double gix = B.D[qx * n + ix] * B.S[qy * n + iy] / h, giy = B.S[qx * n + ix] * B.D[qy * n + iy] / h;
                const double gjx = B.D[qx * n + jx] * B.S[qy * n + jy] / h, gjy = B.S[qx * n + jx] * B.D[qy * n + jy] / h;
                m += pi * pj * jxw;
                c += pi * (ex * gjx + ey * gjy) * jxw;
                k += (ex * gix + ey * giy) * (ex * gjx + ey * gjy) * jxw;
              }
          M[i * nd + j] = m;
          C[i * nd + j] = c;
          K[i * nd + j] = k;
        }
  }

  void
  ora_collide_bgk(const ora_disc *d, double *g, const double omega)
  {
    // include/lbfem/collision.h:152-168
    const int n = d->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      {
        double gi[Q], geq[Q], gt[Q];
        for (int a = 0; a < Q; ++a)
          gi[a] = g[a * n + i];
        const Moments m = moments(gi);
        equilibrium(m, geq);
        equilibrium(impose(*d, i, m), gt);
        for (int a = 0; a < Q; ++a)
          g[a * n + i] = gt[a] + (1. - omega) * (gi[a] - geq[a]);
      }
  }

  void
  ora_compute_equilibrium(const ora_disc *d, const double *f, double *feq)
  {
    // include/lbfem/collision.h:105-113
    const int n = d->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      {
        double fi[Q], e[Q];
        for (int a = 0; a < Q; ++a)
          fi[a] = f[a * n + i];
        equilibrium(impose(*d, i, moments(fi)), e);
        for (int a = 0; a < Q; ++a)
          feq[a * n + i] = e[a];
      }
  }

  void
  ora_predictor_corrector(const ora_disc *d, double *f, const double *feq, const double *incr, const double theta)
  {
    // include/lbfem/collision.h:119-145
    const int    n   = d->n;
    const double inv = 1. / (1. + theta);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
      {
        double fh[Q], eh[Q];
        fh[0] = inv * (f[i] + theta * feq[i]);
        for (int a = 1; a < Q; ++a)
          fh[a] = inv * (f[a * n + i] + theta * feq[a * n + i] + incr[(a - 1) * n + i]);
        equilibrium(impose(*d, i, moments(fh)), eh);
        for (int a = 0; a < Q; ++a)
          f[a * n + i] = fh[a] + theta * (eh[a] - feq[a * n + i]);
      }
  }

  void
  ora_advection_apply(const ora_disc *d, double *rhs, const double *f, const double *feq, const double *e,
                      const double dt, const double lambda, const int leelin, const int remove_mass_flux)
  {
    advection(*d, rhs, f, feq, e, dt, lambda, leelin != 0, remove_mass_flux != 0);
  }

  void
  ora_streaming_vmult(const ora_disc *d, double *y, const double *x, const double c, const double ex, const double ey)
  {
    streaming_vmult(*d, y, x, c, ex, ey);
  }

  int
  ora_mass_solve(const ora_disc *d, double *x, const double *r, const int mass_type, const int richardson,
                 const double cg_tol, const int cg_max_its, const double tg3_coef, const double ex, const double ey,
                 int *its)
  {
    return mass_solve(*d, x, r, mass_type, richardson, cg_tol, cg_max_its, tg3_coef, ex, ey, its);
  }

  ora_scheme *
  ora_scheme_create(const ora_disc *d, const int scheme, const int streaming, const int mass_type,
                    const int richardson, const double cg_tol, const int cg_max_its)
  {
    // streaming.h:51-52 (TG3 needs CG), schemes.h:88 (Lee-Lin TG2 only)
    if (streaming != ORA_TG2 && mass_type != ORA_MASS_CG)
      return nullptr;
    if (scheme == ORA_LEELIN && streaming != ORA_TG2)
      return nullptr;
    ora_scheme *s  = new ora_scheme;
    s->d           = d;
    s->scheme      = scheme;
    s->streaming   = streaming;
    s->mass_type   = mass_type;
    s->richardson  = richardson;
    s->cg_tol      = cg_tol;
    s->cg_max      = cg_max_its;
    s->tg3         = streaming == ORA_TG2 ? 0. : 1. / 6.;                              // streaming.h:47
    s->extrapolate = mass_type == ORA_MASS_CG && streaming != ORA_TG3_SPLIT;          // streaming.h:48
    std::array<double, Q * 2> all{};
    for (int a = 0; a < Q; ++a)
      {
        all[2 * a]     = EX[a];
        all[2 * a + 1] = EY[a];
      }
    if (streaming == ORA_TG3_SPLIT) // streaming.h:53-56, 108-115
      {
        std::array<double, Q * 2> ax{}, ay{};
        for (int a = 0; a < Q; ++a)
          {
            ax[2 * a]     = EX[a];
            ay[2 * a + 1] = EY[a];
          }
        s->sweeps = {ax, ay};
      }
    else
      s->sweeps = {all};
    const int n = d->n;
    s->f.assign(Q * n, 0.);
    if (scheme == ORA_LEELIN)
      s->feq.assign(Q * n, 0.);
    s->rhs.assign(NM * n, 0.);
    s->x[0].assign(NM * n, 0.);
    s->x[1].assign(NM * n, 0.);
    return s;
  }
  void
  ora_scheme_destroy(ora_scheme *s)
  {
    delete s;
  }
  void
  ora_scheme_set_time_step(ora_scheme *s, const double dt, const double lambda)
  {
    s->dt     = dt;
    s->lambda = lambda;
  }
  void
  ora_scheme_set_initial_populations(ora_scheme *s)
  {
    // schemes.h:50-59 and, for Bardow, schemes.h:128-141
    const ora_disc &d = *s->d;
    ora_case_initial_populations(&d, s->lambda, s->f.data());
    if (s->scheme == ORA_BARDOW)
      {
        const double scale = 0.5 * s->dt / s->lambda;
        double       fi[Q], fe[Q];
        for (int i = 0; i < d.n; ++i)
          {
            for (int a = 0; a < Q; ++a)
              fi[a] = s->f[a * d.n + i];
            equilibrium(moments(fi), fe);
            for (int a = 0; a < Q; ++a)
              s->f[a * d.n + i] = fi[a] + scale * (fi[a] - fe[a]);
          }
      }
  }
  void
  ora_scheme_set_populations(ora_scheme *s, const double *f)
  {
    std::copy(f, f + Q * s->d->n, s->f.begin());
  }
  void
  ora_scheme_get_populations(const ora_scheme *s, double *f)
  {
    std::copy(s->f.begin(), s->f.end(), f);
  }
  long long
  ora_scheme_cg_iterations(const ora_scheme *s)
  {
    return s->cg_iterations;
  }
  void
  ora_scheme_last_cg_iterations(const ora_scheme *s, int *its)
  {
    for (int sw = 0; sw < 2; ++sw)
      for (int a = 0; a < NM; ++a)
        its[sw * NM + a] = s->last_its[sw][a];
  }
  void
  ora_scheme_get_increments(const ora_scheme *s, const int which, double *x)
  {
    std::copy(s->x[which].begin(), s->x[which].end(), x);
  }
}

namespace
{
  bool
  moves(const std::array<double, Q * 2> &e, const int a) // streaming.h:118-122
  {
    return e[2 * (a + 1)] != 0. || e[2 * (a + 1) + 1] != 0.;
  }

  // TaylorGalerkin::increment (streaming.h:66-84)
  int
  increment(ora_scheme &s, const double *f, const double *feq, const int sw)
  {
    const ora_disc &d = *s.d;
    const int       n = d.n;
    const auto     &e = s.sweeps[sw];
    advection(d, s.rhs.data(), f, feq, e.data(), s.dt, s.lambda, s.scheme == ORA_LEELIN,
              s.scheme == ORA_BARDOW /* schemes.h:115 vs advection.h:288 */);
    for (int a = 0; a < NM; ++a)
      if (moves(e, a))
        {
          double *xa = s.x[sw].data() + a * n;
          if (s.extrapolate)
            {
              double *x1 = s.x[1].data() + a * n;
              for (int i = 0; i < n; ++i)
                {
                  const double guess = 2. * xa[i] - x1[i]; // sadd(-1, 2, xa)
                  x1[i]              = xa[i];                // swap
                  xa[i]              = guess;
                }
            }
          int       its = 0;
          const int st  = mass_solve(d, xa, s.rhs.data() + a * n, s.mass_type, s.richardson, s.cg_tol, s.cg_max,
                                    s.tg3 * s.dt * s.dt, e[2 * (a + 1)], e[2 * (a + 1) + 1], &its);
          s.cg_iterations += its;
          s.last_its[sw][a] = its; // SolverControl::last_step() of this solve
          if (st != ORA_OK)
            return st;
        }
    return ORA_OK;
  }
} // namespace

extern "C"
{
  int
  ora_scheme_step(ora_scheme *s, const int n_steps)
  {
    const ora_disc &d = *s->d;
    const int       n = d.n;
    for (int k = 0; k < n_steps; ++k)
      {
        for (auto &row : s->last_its)
          for (int &v : row)
            v = 0;
        if (s->scheme == ORA_LEELIN) // schemes.h:92-98
          {
            ora_compute_equilibrium(&d, s->f.data(), s->feq.data());
            const int st = increment(*s, s->f.data(), s->feq.data(), 0);
            if (st != ORA_OK)
              return st;
            ora_predictor_corrector(&d, s->f.data(), s->feq.data(), s->x[0].data(), s->dt / s->lambda);
          }
        else // schemes.h:118-124, streaming.h:88-100
          {
            ora_collide_bgk(&d, s->f.data(), s->dt / (s->lambda + 0.5 * s->dt));
            for (int sw = 0; sw < (int)s->sweeps.size(); ++sw)
              {
                const int st = increment(*s, s->f.data(), nullptr, sw);
                if (st != ORA_OK)
                  return st;
                for (int a = 0; a < NM; ++a)
                  if (moves(s->sweeps[sw], a))
                    for (int i = 0; i < n; ++i)
                      s->f[(a + 1) * n + i] += s->x[sw][a * n + i];
              }
          }
      }
    return ORA_OK;
  }

  int
  ora_stream_function(const ora_disc *d, const double *u, const double *v, double *psi, int *its)
  {
    if (d->per[0] || d->per[1] || d->general)
      return ORA_EINVAL;
    int       it = 0;
    const int st = stream_function(*d, u, v, psi, &it);
    if (its)
      *its = it;
    return st;
  }

  void
  ora_diagnostics(const ora_disc *d, const double *f, const double t, const double *u_prev, double *s)
  {
    // apps/cgdbe/cgdbe.cc:353-368, over the non-slave nodes with lumped-mass weights
    for (int k = 0; k < 6; ++k)
      s[k] = 0.;
    double fi[Q];
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        {
          const int k = j * d->np + i;
          if (d->constrained[k])
            continue;
          for (int a = 0; a < Q; ++a)
            fi[a] = f[a * d->n + k];
          const Moments m  = moments(fi);
          const Moments ex = d->tc.exact(d->XN[k], d->YN[k], t);
          const double  w  = d->ml[k];
          s[0] += w * ((m.ux - ex.ux) * (m.ux - ex.ux) + (m.uy - ex.uy) * (m.uy - ex.uy));
          s[1] += w * (ex.ux * ex.ux + ex.uy * ex.uy);
          s[2] += w * (m.ux * m.ux + m.uy * m.uy);
          s[3] += w * m.rho;
          s[4] += w;
          s[5] += w * std::hypot(m.ux - u_prev[2 * k], m.uy - u_prev[2 * k + 1]);
        }
  }
}
