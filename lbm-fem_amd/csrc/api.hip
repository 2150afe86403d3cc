// The C ABI of include/lbfem_hip.h: discretization and scheme objects, the
// sequencing of a time step (include/lbfem/schemes.h, streaming.h, mass.h)
// over the kernels of tile_kernels.hip and solver_kernels.hip.
#include "lbfem_hip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "geometry.hpp"
#include "kernels.hpp"

using namespace lbfem_amd;

namespace
{
  // the message of the thread's last failed call: a fixed buffer, so that recording an
  // error allocates nothing (the allocation-free entry point reports through it too)
  constexpr size_t         kErrorLen = 1024;
  thread_local char        g_error[kErrorLen] = {0};
  thread_local size_t      g_error_len        = 0;

  int
  fail(const int code, const char *msg)
  {
    const size_t k = std::min(kErrorLen - 1, std::strlen(msg));
    std::memcpy(g_error, msg, k);
    g_error[k]  = '\0';
    g_error_len = k;
    return code;
  }

  int
  fail(const int code, const std::string &msg)
  {
    return fail(code, msg.c_str());
  }

  // a HIP error as a status, without an exception
  int
  fail_hip(const hipError_t e, const char *what)
  {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    return fail(LBFEM_EHIP, buf);
  }
#define HIP_RETURN(x)                                                                                              \
  do                                                                                                               \
    {                                                                                                              \
      const hipError_t e_ = (x);                                                                                   \
      if (e_ != hipSuccess)                                                                                        \
        return fail_hip(e_, #x);                                                                                   \
    }                                                                                                              \
  while (0)

  struct HipError : std::runtime_error
  {
    using std::runtime_error::runtime_error;
  };

  void
  check(const hipError_t e, const char *what)
  {
    if (e != hipSuccess)
      throw HipError(std::string(what) + ": " + hipGetErrorString(e));
  }
#define HIP_CHECK(x) check((x), #x)

  // HIP runtime calls that allocate or free memory, create or destroy streams, events and
  // graphs, or copy on the null stream are serialized with the capture of the batches' step
  // graphs: a capture in one thread fails when another thread makes such calls meanwhile.
  std::recursive_mutex &
  runtime_mutex()
  {
    static std::recursive_mutex m;
    return m;
  }
  using RuntimeLock = std::lock_guard<std::recursive_mutex>;

  template <typename T>
  struct DevBuf
  {
    T     *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept
      : p(o.p)
      , n(o.n)
    {
      o.p = nullptr;
      o.n = 0;
    }
    DevBuf &
    operator=(DevBuf &&o) noexcept
    {
      if (this != &o)
        {
          RuntimeLock lock(runtime_mutex());
          if (p)
            (void)hipFree(p);
          p   = o.p;
          n   = o.n;
          o.p = nullptr;
          o.n = 0;
        }
      return *this;
    }
    ~DevBuf()
    {
      RuntimeLock lock(runtime_mutex());
      if (p)
        (void)hipFree(p);
    }
    void
    alloc(const size_t count)
    {
      RuntimeLock lock(runtime_mutex());
      if (p)
        (void)hipFree(p);
      p = nullptr;
      n = 0;
      if (count)
        HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
      n = count; // only once the memory is there
    }
    void
    upload(const std::vector<T> &v)
    {
      RuntimeLock lock(runtime_mutex());
      alloc(v.size());
      if (!v.empty())
        HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
  };

  // D2Q9 directions (include/lbfem/d2q9.h:18-19)
  constexpr double kE[Q][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};

  constexpr int kMaxCgBlocks = 1 << 17; // blocks of one CG operator launch (partial sums)
} // namespace

struct lbfem_disc
{
  lbfem_mesh_desc          desc{};
  std::unique_ptr<Basis1D> B;
  std::unique_ptr<Axis1D>  ax, ay;
  int                      p = 1, N = 1, P = 1, np = 2;
  int                      device = 0; // the HIP device its buffers live on
  int64_t                  n = 0, stride = 0;
  double                   h_min = 0.;
  DevBuf<double>           axbuf, aybuf, wall_vel;
  DevBuf<int32_t>          wall_side[4];
  std::vector<int32_t>     wall_nodes;
  std::vector<double>      wall_velocity;
  Geo                      G{};
  // general quadrilaterals (distort != 0): node coordinates, lumped mass and
  // diagonal of M per node (host), and the device copies behind G.gen
  bool                     general = false;
  std::vector<double>      gen_xy, gen_ml, gen_diag;
  DevBuf<double>           gen_buf;
  DevBuf<double>           dinv_buf; // the tensor mesh's Jacobi diagonal per node (Geo::dinv)
  // scratch of the operator-level entry points (the mass solve and the stream function),
  // allocated on first use and kept: one call at a time per discretization
  DevBuf<CgState>  cg_state; // two sets of NM: the stopping states of odd and even iterations
  DevBuf<double>   partial, s_r, s_p0, s_p1, s_q, s_ax, s_b, s_dinv, s_sums;
  CgState         *host_cg = nullptr; // pinned: the CG stopping states read by the host

  lbfem_disc() = default;
  lbfem_disc(const lbfem_disc &) = delete;
  lbfem_disc &operator=(const lbfem_disc &) = delete;
  ~lbfem_disc()
  {
    RuntimeLock lock(runtime_mutex());
    if (host_cg)
      (void)hipHostFree(host_cg);
  }

  void
  upload_axis(const Axis1D &a, DevBuf<double> &buf, AxisDev &dev)
  {
    const size_t        w = size_t(a.P + 1) * (2 * a.p + 1), m = size_t(a.P + 1);
    std::vector<double> host;
    host.reserve(4 * w + 4 * m);
    for (const auto *v : {&a.M, &a.D, &a.DT, &a.K})
      host.insert(host.end(), v->begin(), v->end());
    for (size_t i = 0; i < m; ++i)
      host.push_back(1. / a.row_sum[i]);
    host.insert(host.end(), a.row_sum.begin(), a.row_sum.end());
    host.insert(host.end(), a.diagonal.begin(), a.diagonal.end());
    for (size_t i = 0; i < m; ++i)
      host.push_back(1. / a.diagonal[i]);
    buf.upload(host);
    dev.M        = buf.p;
    dev.D        = buf.p + w;
    dev.DT       = buf.p + 2 * w;
    dev.K        = buf.p + 3 * w;
    dev.ml_inv   = buf.p + 4 * w;
    dev.ml       = buf.p + 4 * w + m;
    dev.diag     = buf.p + 4 * w + 2 * m;
    dev.diag_inv = buf.p + 4 * w + 3 * m;
    dev.periodic = a.periodic;
    for (int l = 0; l < 4; ++l)
      {
        dev.b0[l] = l <= a.p ? a.b0[l] : 0.;
        dev.b1[l] = l <= a.p ? a.b1[l] : 0.;
      }
  }

  // The distorted mesh of TaylorGreenVortex::transform_mesh (test_cases.h:53-65):
  // vertices moved by x -> x + eps L sin(2 pi x/L) sin(2 pi y/L) (1, 1), nodes
  // through each cell's bilinear map (MappingQ1 support points,
  // discretization.h:127-135; a node on a cell edge from its lower-left cell),
  // and the lumped mass (row sums of M) and diag(M) from QGauss(p+1) with the
  // Jacobian per point, summed per node over its cells (lower row first, left
  // cell first); constrained entries 1 (discretization.h:123-125).
  // Geo::dinv: the general cells' per-node inverse diagonal, or on the tensor mesh the
  // product of the axes' inverse diagonals (the same double product the kernels formed per
  // node), 1 on slave rows (MatrixFreeOperators' constrained diagonal entries)
  void
  build_jacobi()
  {
    if (general)
      {
        G.dinv = G.gen.diag_inv;
        return;
      }
    std::vector<double> h(static_cast<size_t>(n));
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i)
        {
          const bool slave       = (ax->periodic && i == P) || (ay->periodic && j == P);
          h[size_t(j) * np + i] = slave ? 1. : (1. / ax->diagonal[i]) * (1. / ay->diagonal[j]);
        }
    dinv_buf.upload(h);
    G.dinv = dinv_buf.p;
  }

  void
  build_general(const double eps)
  {
    const int           N1 = N + 1, n1 = p + 1;
    const double        L  = desc.L;
    std::vector<double> vx(size_t(N1) * N1), vy(size_t(N1) * N1);
    for (int j = 0; j <= N; ++j)
      for (int i = 0; i <= N; ++i)
        {
          const double x = L * i / N, y = L * j / N;
          const double dd = eps * L * std::sin(2 * kPi * x / L) * std::sin(2 * kPi * y / L);
          vx[size_t(j) * N1 + i] = x + dd;
          vy[size_t(j) * N1 + i] = y + dd;
        }
    const auto corner = [&](const int cx, const int cy, double c[8]) {
      const size_t a = size_t(cy) * N1 + cx, b = a + N1;
      c[0] = vx[a], c[1] = vy[a], c[2] = vx[a + 1], c[3] = vy[a + 1];
      c[4] = vx[b], c[5] = vy[b], c[6] = vx[b + 1], c[7] = vy[b + 1];
    };
    gen_xy.assign(size_t(2) * n, 0.);
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i)
        {
          const int cx = std::min(i / p, N - 1), cy = std::min(j / p, N - 1);
          double    c[8];
          corner(cx, cy, c);
          const double xi = B->node[i - p * cx], eta = B->node[j - p * cy];
          const size_t k  = size_t(j) * np + i;
          gen_xy[2 * k] = (1. - eta) * ((1. - xi) * c[0] + xi * c[2]) + eta * ((1. - xi) * c[4] + xi * c[6]);
          gen_xy[2 * k + 1] = (1. - eta) * ((1. - xi) * c[1] + xi * c[3]) + eta * ((1. - xi) * c[5] + xi * c[7]);
        }
    // per cell and local node: int phi_k and int phi_k^2
    std::vector<double> Rl(size_t(N) * N * n1 * n1), Rd(Rl.size());
    for (int cy = 0; cy < N; ++cy)
      for (int cx = 0; cx < N; ++cx)
        {
          double c[8], det[16];
          corner(cx, cy, c);
          for (int qy = 0; qy < n1; ++qy)
            for (int qx = 0; qx < n1; ++qx)
              {
                const double xi = B->qx[qx], eta = B->qx[qy];
                const double xxi  = (1. - eta) * (c[2] - c[0]) + eta * (c[6] - c[4]);
                const double yxi  = (1. - eta) * (c[3] - c[1]) + eta * (c[7] - c[5]);
                const double xeta = (1. - xi) * (c[4] - c[0]) + xi * (c[6] - c[2]);
                const double yeta = (1. - xi) * (c[5] - c[1]) + xi * (c[7] - c[3]);
                det[qy * n1 + qx] = xxi * yeta - xeta * yxi;
              }
          for (int ky = 0; ky < n1; ++ky)
            for (int kx = 0; kx < n1; ++kx)
              {
                double sl = 0., sd = 0.;
                for (int qy = 0; qy < n1; ++qy)
                  for (int qx = 0; qx < n1; ++qx)
                    {
                      const double phi = B->S[qx][kx] * B->S[qy][ky];
                      const double jxw = det[qy * n1 + qx] * B->qw[qx] * B->qw[qy];
                      sl += phi * jxw;
                      sd += phi * phi * jxw;
                    }
                const size_t o = (size_t(cy) * N + cx) * n1 * n1 + ky * n1 + kx;
                Rl[o]          = sl;
                Rd[o]          = sd;
              }
        }
    const auto around = [&](const int i, int *c, int *k) {
      int m = 0;
      if (i % p == 0)
        {
          c[m] = i == 0 ? N - 1 : i / p - 1; // periodic: the master at 0 also closes cell N - 1
          k[m++] = p;
          c[m]   = i / p;
          k[m++] = 0;
        }
      else
        {
          c[m]   = i / p;
          k[m++] = i % p;
        }
      return m;
    };
    gen_ml.assign(n, 1.);
    gen_diag.assign(n, 1.);
    for (int j = 0; j < P; ++j)
      {
        int       cys[2], kys[2];
        const int my = around(j, cys, kys);
        for (int i = 0; i < P; ++i)
          {
            int       cxs[2], kxs[2];
            const int mx = around(i, cxs, kxs);
            double    sl = 0., sd = 0.;
            for (int b = 0; b < my; ++b)
              for (int a = 0; a < mx; ++a)
                {
                  const size_t o = (size_t(cys[b]) * N + cxs[a]) * n1 * n1 + kys[b] * n1 + kxs[a];
                  sl += Rl[o];
                  sd += Rd[o];
                }
            gen_ml[size_t(j) * np + i]   = sl;
            gen_diag[size_t(j) * np + i] = sd;
          }
      }
    // device: vertices (x, y) pairs, ml, 1/ml, 1/diag, basis tables
    std::vector<double> host;
    host.reserve(size_t(2) * N1 * N1 + 3 * size_t(n) + 40);
    for (size_t k = 0; k < vx.size(); ++k)
      {
        host.push_back(vx[k]);
        host.push_back(vy[k]);
      }
    const size_t o_ml = host.size();
    host.insert(host.end(), gen_ml.begin(), gen_ml.end());
    const size_t o_mli = host.size();
    for (const double v : gen_ml)
      host.push_back(1. / v);
    const size_t o_di = host.size();
    for (const double v : gen_diag)
      host.push_back(1. / v);
    const size_t o_b = host.size();
    for (int q = 0; q < 4; ++q)
      for (int k = 0; k < 4; ++k)
        host.push_back(B->S[q][k]);
    for (int q = 0; q < 4; ++q)
      for (int k = 0; k < 4; ++k)
        host.push_back(B->D[q][k]);
    for (int q = 0; q < 4; ++q)
      host.push_back(B->qx[q]);
    for (int q = 0; q < 4; ++q)
      host.push_back(B->qw[q]);
    gen_buf.upload(host);
    G.gen.vert     = gen_buf.p;
    G.gen.ml       = gen_buf.p + o_ml;
    G.gen.ml_inv   = gen_buf.p + o_mli;
    G.gen.diag_inv = gen_buf.p + o_di;
    G.gen.basis    = gen_buf.p + o_b;
    G.gen.N        = N;
    general        = true;
  }

  void
  ensure_scratch()
  {
    RuntimeLock lock(runtime_mutex());
    if (!host_cg)
      HIP_CHECK(hipHostMalloc(&host_cg, sizeof(CgState) * NM));
    if (s_sums.p) // allocated last: the rest is there
      return;
    cg_state.alloc(2 * NM);
    partial.alloc(size_t(3 * NM) * kMaxCgBlocks);
    s_sums.alloc(6); // the diagnostics' sums
  }

  // the node vectors of the operator-level CG (r, two directions, A p, A x) and of the
  // stream function (its right-hand side and inverse diagonal)
  void
  ensure_vectors()
  {
    if (s_dinv.n >= size_t(stride)) // allocated last: the rest is there
      return;
    for (DevBuf<double> *bb : {&s_r, &s_p0, &s_p1, &s_q, &s_ax, &s_b, &s_dinv})
      bb->alloc(stride);
  }
};

namespace
{
  // the velocity-change monitor's device table, partial sums and results, and the pinned
  // host copy of the results (allocated on first use, grown as needed)
  struct Monitor
  {
    DevBuf<char>   table;
    DevBuf<double> partial, results;
    double        *host     = nullptr;
    size_t         host_cap = 0;

    Monitor() = default;
    Monitor(const Monitor &) = delete;
    Monitor &operator=(const Monitor &) = delete;
    ~Monitor()
    {
      RuntimeLock lock(runtime_mutex());
      if (host)
        (void)hipHostFree(host);
    }

    // runs the monitor over the entries on the stream; out: kMonitorPartials doubles per entry
    void
    run(std::vector<MonitorEntry> &E, hipStream_t stream, double *out)
    {
      const int n = int(E.size());
      if (n == 0)
        return;
      int blocks = 0;
      for (auto &e : E)
        {
          e.block_begin = blocks;
          e.nblocks     = monitor_blocks(int64_t(e.G.np) * e.G.np);
          blocks += e.nblocks;
        }
      if (table.n < E.size() * sizeof(MonitorEntry))
        table.alloc(E.size() * sizeof(MonitorEntry));
      if (partial.n < size_t(blocks) * kMonitorPartials)
        partial.alloc(size_t(blocks) * kMonitorPartials);
      const size_t nr = size_t(n) * kMonitorPartials;
      if (results.n < nr)
        results.alloc(nr);
      if (host_cap < nr)
        {
          RuntimeLock lock(runtime_mutex());
          if (host)
            (void)hipHostFree(host);
          host     = nullptr;
          host_cap = 0;
          HIP_CHECK(hipHostMalloc(&host, nr * sizeof(double)));
          host_cap = nr;
        }
      HIP_CHECK(hipMemcpyAsync(table.p, E.data(), E.size() * sizeof(MonitorEntry), hipMemcpyHostToDevice, stream));
      HIP_CHECK(monitor_launch(reinterpret_cast<const MonitorEntry *>(table.p), n, blocks, partial.p, results.p,
                               stream));
      HIP_CHECK(hipMemcpyAsync(host, results.p, nr * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      std::copy(host, host + nr, out);
    }
  };

} // namespace

struct lbfem_scheme
{
  const lbfem_disc     *d = nullptr;
  lbfem_scheme_settings set{};
  double                dt = 0., lambda = 0.;
  hipStream_t           stream = nullptr;
  DevBuf<double>        f[2];
  DevBuf<float>         g32[2]; // the mixed form: FP32 deviations f - w rho0
  bool                  mixed = false;
  int                   cur   = 0;
  DevBuf<double>        rhs, x[2], r, pb[2], q, ax_tmp;
  // the mixed form's consistent mass: FP32 right-hand side, increments (x32[1] the previous
  // step's, for the extrapolated start) and CG vectors
  DevBuf<float>         rhs32, x32[2], r32, p32[2], q32;
  DevBuf<CgState>       state32; // two sets of NM, as state
  CgState              *host_state32 = nullptr; // pinned
  DevBuf<CgState>       state; // two sets of NM: the stopping states of odd and even iterations
  DevBuf<double>        partial;
  CgState              *host_state = nullptr; // pinned
  int64_t               cg_iterations = 0;
  int                   last_step_its[2][NM] = {}; // the last step's CG steps per sweep and moving population
  // the velocity-change monitor: the velocity at its last call (2 per node) and its buffers
  DevBuf<double>        u_prev;
  bool                  u_prev_set = false;
  Monitor               monitor;
  // cgdbe's diagnostics on the device: the exact velocity (2 per node, scaled per call), the
  // velocity at the previous call (zero before the first), partial sums and the six sums
  DevBuf<double>        diag_exact, diag_prev, diag_partial, diag_sums;
  bool                  diag_exact_set = false;
  double               *diag_host      = nullptr; // pinned, 6
  std::vector<std::array<double, 2 * Q>> sweeps;
  double                tg3         = 0.;
  bool                  extrapolate = false;
  int                   last_its    = 1; // the largest CG count of the last solve
  int                   tile_variant = -1; // TileVariant of the tile kernel; -1: tile_default_variant
  // stage timers (include/lbfem/timers.h): collision, advection, mass, total
  bool                                   timing = false;
  std::vector<std::pair<hipEvent_t, int>> marks;
  double                                 stage[4] = {0., 0., 0., 0.};

  ~lbfem_scheme()
  {
    RuntimeLock lock(runtime_mutex());
    for (auto &m : marks)
      (void)hipEventDestroy(m.first);
    if (host_state)
      (void)hipHostFree(host_state);
    if (host_state32)
      (void)hipHostFree(host_state32);
    if (diag_host)
      (void)hipHostFree(diag_host);
    if (stream)
      (void)hipStreamDestroy(stream);
  }

  int
  variant() const
  {
    return tile_variant >= 0 ? tile_variant : tile_default_variant(d->G, set.precision == LBFEM_MIXED);
  }

  void
  mark(const int stage_id)
  {
    if (!timing)
      return;
    hipEvent_t e;
    {
      RuntimeLock lock(runtime_mutex());
      HIP_CHECK(hipEventCreate(&e));
    }
    HIP_CHECK(hipEventRecord(e, stream));
    marks.push_back({e, stage_id});
  }

  void
  resolve_marks()
  {
    if (marks.empty())
      return;
    HIP_CHECK(hipEventSynchronize(marks.back().first));
    for (size_t k = 1; k < marks.size(); ++k)
      {
        if (marks[k].second < 0) // a step boundary opens a new interval
          continue;
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, marks[k - 1].first, marks[k].first));
        stage[marks[k].second] += 1e-3 * ms;
        stage[3] += 1e-3 * ms;
      }
    for (auto &m : marks)
      (void)hipEventDestroy(m.first);
    marks.clear();
  }
};

namespace
{
  // ------------------------------------------------------------------ helpers
  TileParams
  tile_params(const lbfem_scheme &s, const std::array<double, 2 * Q> &e, const bool leelin)
  {
    TileParams S{};
    for (int a = 0; a < Q; ++a)
      {
        S.e[a][0] = e[2 * a];
        S.e[a][1] = e[2 * a + 1];
      }
    S.dt = s.dt;
    if (leelin) // advection.h:268-270
      {
        S.b = -0.5 * s.dt * s.dt / s.lambda;
        S.a = -s.dt - S.b;
      }
    else // advection.h:223
      {
        S.a = -s.dt;
        S.b = 0.;
      }
    S.c                = -0.5 * s.dt * s.dt;
    S.omega            = s.dt / (s.lambda + 0.5 * s.dt); // schemes.h:122
    S.theta            = s.dt / s.lambda;                // schemes.h:97
    S.remove_mass_flux = leelin ? 0 : 1;                 // advection.h:288, schemes.h:115
    return S;
  }

  bool
  moves(const std::array<double, 2 * Q> &e, const int a) // streaming.h:118-122
  {
    return e[2 * (a + 1)] != 0. || e[2 * (a + 1) + 1] != 0.;
  }

  CgWork
  cg_work(const lbfem_disc &d, CgState *state, CgState *state2, double *partial, const lbfem_mass_settings &m)
  {
    CgWork W{};
    W.max_its    = m.cg_max_iterations;
    W.tol_rel    = m.cg_tolerance;
    W.state      = state;
    W.state2     = state2;
    W.partial    = partial;
    W.max_blocks = kMaxCgBlocks;
    (void)d;
    return W;
  }

  // CG on the vectors of W (include/lbfem/mass.h:204-211): returns LBFEM_OK,
  // LBFEM_ENOCONV, or LBFEM_EHIP (a status, not an exception: nothing is
  // allocated on any path); adds the iterations of all vectors to *its.
  // The host reads the stopping state after iteration check_from and then every
  // check_every iterations; iterations of converged vectors are no-ops.
  template <typename V>
  int
  run_cg(const Geo &G, const CgWork &W, V *x, const V *b, V *r, V *p0, V *p1, V *q, CgState *host_state,
         hipStream_t stream, int64_t *its, const int check_from, const bool fail_at_cap = true,
         const int check_every = 1)
  {
    CgWork Wf = W;
    // small grids: the update kernel reduces the direction stage's partials itself, and the
    // direction kernel the update stage's (two launches less per iteration; with two sets of
    // states, tensor meshes); large ones keep the separate one-block finalisations
    const int rows = cg_direction_rows(G);
    Wf.alpha_rows  = int64_t(rows) * W.nv <= 4096 && rows + 1024 <= W.max_blocks ? rows : 0;
    const int upd  = cg_update_blocks(G, W.max_blocks, Wf.alpha_rows);
    Wf.fold_rows   = Wf.alpha_rows > 0 && upd <= kMaxCgFoldRows && W.state2 && G.gen.vert == nullptr ? upd : 0;
    HIP_RETURN(cg_start_launch(G, W, x, b, r, stream));
    V         *pb[2] = {p0, p1};
    // the host's read of the states after iteration it (with the fold, they are formed first)
    const auto done = [&](const int it, bool &finished) -> int {
      finished = true;
      if (Wf.fold_rows > 0 && it > 0)
        HIP_RETURN(cg_update_finalize_launch(G, Wf, it, stream));
      const CgState *src = Wf.fold_rows > 0 && (it & 1) ? W.state2 : W.state;
      HIP_RETURN(hipMemcpyAsync(host_state, src, W.nv * sizeof(CgState), hipMemcpyDeviceToHost, stream));
      HIP_RETURN(hipStreamSynchronize(stream));
      bool active = false, failed = false;
      for (int v = 0; v < W.nv; ++v)
        {
          active |= host_state[v].status == 0;
          failed |= host_state[v].status == 2;
        }
      finished = !active;
      if (!finished)
        return LBFEM_OK;
      for (int v = 0; v < W.nv; ++v)
        *its += host_state[v].its;
      if (failed && fail_at_cap)
        {
          char msg[128];
          std::snprintf(msg, sizeof(msg), "CG: no convergence within %d iterations (SolverControl::NoConvergence)",
                        W.max_its);
          return fail(LBFEM_ENOCONV, msg);
        }
      return LBFEM_OK;
    };
    // iterations of converged vectors are no-ops, so the first checks can be skipped
    for (int it = 1; it <= W.max_its; ++it)
      {
        HIP_RETURN(cg_direction_launch(G, Wf, it, pb[it % 2], r, pb[(it + 1) % 2], q, stream));
        HIP_RETURN(cg_update_launch(G, Wf, it, x, r, pb[(it + 1) % 2], q, stream));
        if (it >= check_from && (it - check_from) % check_every == 0)
          {
            bool      finished = false;
            const int st       = done(it, finished);
            if (finished)
              return st;
          }
      }
    bool finished = false;
    return done(std::max(W.max_its, 0), finished); // the update at max_its leaves no vector active
  }

  // the increments of one sweep: advection, then the mass solves (streaming.h:66-84)
  int
  increment(lbfem_scheme &s, const int sw, const bool collide)
  {
    const lbfem_disc &d  = *s.d;
    const Geo        &G  = d.G;
    const auto       &e  = s.sweeps[sw];
    const bool        ll = s.set.scheme == LBFEM_LEELIN;
    TileParams        S  = tile_params(s, e, ll);
    S.collide            = collide ? 1 : 0;
    double *fin = s.f[s.cur].p, *gstar = nullptr;
    if (collide)
      gstar = s.f[1 - s.cur].p;
    s.mark(-1);
    if (d.general)
      HIP_CHECK(quad_advect_launch(ll ? kAdvectLeeLin : kAdvectBardow, G, S, fin, nullptr, gstar, s.rhs.p, s.stream));
    else
      HIP_CHECK(tile_launch(ll ? kAdvectLeeLin : kAdvectBardow, s.variant(), G, S, fin, (const double *)nullptr,
                            gstar, s.rhs.p, s.stream));
    s.mark(1);
    if (collide)
      s.cur = 1 - s.cur;

    unsigned mask = 0;
    for (int a = 0; a < NM; ++a)
      if (moves(e, a))
        mask |= 1u << a;
    double   *xs = s.x[sw].p;
    const int type = s.set.mass.type;
    if (type == LBFEM_MASS_LUMPED)
      HIP_CHECK(lumped_solve_launch(G, xs, s.rhs.p, NM, s.stream));
    else if (type == LBFEM_MASS_RICHARDSON) // mass.h:193-202, x_0 = M_L^-1 r, x += M_L^-1 (r - A x)
      {
        HIP_CHECK(lumped_solve_launch(G, xs, s.rhs.p, NM, s.stream));
        CgWork W = cg_work(d, s.state.p, nullptr, s.partial.p, s.set.mass);
        W.nv     = NM;
        for (int a = 0; a < NM; ++a)
          {
            W.block[a]       = a;
            W.vec[a].e[0]    = e[2 * (a + 1)];
            W.vec[a].e[1]    = e[2 * (a + 1) + 1];
            W.vec[a].coef    = 0.;
          }
        for (int k = 0; k < s.set.mass.richardson; ++k)
          {
            HIP_CHECK(vmult_launch(G, W, xs, s.ax_tmp.p, s.stream));
            HIP_CHECK(richardson_update_launch(G, xs, s.rhs.p, s.ax_tmp.p, NM, s.stream));
          }
      }
    else
      {
        if (s.extrapolate)
          HIP_CHECK(extrapolate_launch(G, xs, s.x[1].p, mask, s.stream));
        CgWork W = cg_work(d, s.state.p, s.state.p + NM, s.partial.p, s.set.mass);
        W.nv     = 0;
        for (int a = 0; a < NM; ++a)
          if (mask & (1u << a))
            {
              W.block[W.nv]      = a;
              W.vec[W.nv].e[0]   = e[2 * (a + 1)];
              W.vec[W.nv].e[1]   = e[2 * (a + 1) + 1];
              W.vec[W.nv].coef   = s.tg3 * s.dt * s.dt;
              ++W.nv;
            }
        // the host first reads the stopping state at the last solve's largest count (a vector
        // that converges earlier idles through no-op iterations, cheaper than a host round trip)
        int64_t   its = 0;
        const int st  = run_cg(d.G, W, xs, s.rhs.p, s.r.p, s.pb[0].p, s.pb[1].p, s.q.p, s.host_state, s.stream, &its,
                               std::max(1, s.last_its));
        s.cg_iterations += its;
        s.last_its = 0;
        for (int v = 0; v < W.nv; ++v)
          {
            s.last_its                      = std::max(s.last_its, s.host_state[v].its);
            s.last_step_its[sw][W.block[v]] = s.host_state[v].its;
          }
        if (st != LBFEM_OK)
          return st;
      }
    s.mark(2);
    return LBFEM_OK;
  }

  // The mixed form's consistent-mass solves stop at max(cg_tolerance, mixed_cg_floor) |b|, the
  // floor max(1e-6, 2^-24 np / (2 pi)) with np nodes per mesh line: the right-hand side comes from
  // FP32 deviations and carries their round-off, about 2^-24 |g| per value while |b| is about
  // |g| 2 pi / np times smaller (a difference over one mesh spacing of values that vary over the
  // domain), so a solve below that level resolves the round-off at a cost of iterations and no
  // accuracy. Measured against FP64 solves to 1e-10 (relative L2 velocity, Taylor-Green U = 0.01,
  // Re 100, CFL 0.25, 5000 steps): Q1 1024^2 at a floor of 1e-6 6.0 iterations per solve, 0.713 ms
  // per step, 1.1e-6; at 1e-5 4.1 iterations, 0.538 ms, 6.0e-7 (FP64: 2.65, 0.667 ms); Q2 512^2
  // 4.5 / 0.630 ms / 4.9e-6 and 2.9 / 0.483 ms / 3.5e-6 (FP64: 1.68, 0.563 ms). On meshes of 20^2
  // to 48^2 cells (600 steps) fixed floors of 1e-5 and 1e-4 lost up to 9e-6 and 7e-5 where 1e-6
  // keeps 1e-6: the floor scales with the round-off, 1e-6 up to 105 nodes per line, 9.7e-6 at 1025.
  inline double
  mixed_cg_floor(const Geo &G)
  {
    return std::max(1e-6, 0x1p-24 * double(G.np) / (2. * M_PI));
  }

  // the mixed form's increments of one sweep: FP32 advection (collision fused) into an FP32
  // right-hand side, then the consistent-mass solves on FP32 vectors (FP64 scalars and sums)
  // from the extrapolated FP32 increment, as the FP64 form's solves (streaming.h:66-84)
  int
  increment_mixed(lbfem_scheme &s, const int sw, const bool collide)
  {
    const lbfem_disc &d  = *s.d;
    const Geo        &G  = d.G;
    const auto       &e  = s.sweeps[sw];
    const bool        ll = s.set.scheme == LBFEM_LEELIN;
    TileParams        S  = tile_params(s, e, ll);
    S.collide            = collide ? 1 : 0;
    float *gstar         = collide ? s.g32[1 - s.cur].p : nullptr;
    s.mark(-1);
    HIP_CHECK(tile_launch(ll ? kAdvectLeeLin : kAdvectBardow, s.variant(), G, S, s.g32[s.cur].p,
                          (const float *)nullptr, gstar, s.rhs32.p, s.stream));
    s.mark(1);
    if (collide)
      s.cur = 1 - s.cur;
    unsigned mask = 0;
    for (int a = 0; a < NM; ++a)
      if (moves(e, a))
        mask |= 1u << a;
    float *xs = s.x32[sw].p;
    if (s.extrapolate)
      HIP_CHECK(extrapolate_launch(G, xs, s.x32[1].p, mask, s.stream));
    CgWork W  = cg_work(d, s.state32.p, s.state32.p + NM, s.partial.p, s.set.mass);
    W.tol_rel = std::max(s.set.mass.cg_tolerance, mixed_cg_floor(G));
    W.nv      = 0;
    for (int a = 0; a < NM; ++a)
      if (mask & (1u << a))
        {
          W.block[W.nv]    = a;
          W.vec[W.nv].e[0] = e[2 * (a + 1)];
          W.vec[W.nv].e[1] = e[2 * (a + 1) + 1];
          W.vec[W.nv].coef = s.tg3 * s.dt * s.dt;
          ++W.nv;
        }
    int64_t   its = 0;
    const int st  = run_cg<float>(G, W, xs, s.rhs32.p, s.r32.p, s.p32[0].p, s.p32[1].p, s.q32.p, s.host_state32,
                                 s.stream, &its, std::max(1, s.last_its));
    s.cg_iterations += its;
    s.last_its = 0;
    for (int v = 0; v < W.nv; ++v)
      {
        s.last_its                      = std::max(s.last_its, s.host_state32[v].its);
        s.last_step_its[sw][W.block[v]] = s.host_state32[v].its;
      }
    if (st != LBFEM_OK)
      return st;
    s.mark(2);
    return LBFEM_OK;
  }

  // One launch of the tile kernel the scheme's step uses, with variant v, on the
  // scheme's buffers without changing its state: the fused step writes the
  // spare population buffer, the advection the right-hand side (and, for
  // Bardow, g* into the spare buffer).
  int
  tile_trial(lbfem_scheme &s, const int v)
  {
    const Geo &G  = s.d->G;
    const bool ll = s.set.scheme == LBFEM_LEELIN;
    TileParams S  = tile_params(s, s.sweeps[0], ll);
    if (s.set.mass.type == LBFEM_MASS_LUMPED && s.set.streaming == LBFEM_TG2)
      {
        S.collide = 1;
        if (s.mixed)
          HIP_CHECK(tile_launch(ll ? kStepLeeLin : kStepBardow, v, G, S, s.g32[s.cur].p, (const float *)nullptr,
                                s.g32[1 - s.cur].p, (float *)nullptr, s.stream));
        else
          HIP_CHECK(tile_launch(ll ? kStepLeeLin : kStepBardow, v, G, S, s.f[s.cur].p, (const double *)nullptr,
                                s.f[1 - s.cur].p, (double *)nullptr, s.stream));
      }
    else if (s.mixed)
      {
        S.collide = ll ? 0 : 1;
        HIP_CHECK(tile_launch(ll ? kAdvectLeeLin : kAdvectBardow, v, G, S, s.g32[s.cur].p, (const float *)nullptr,
                              ll ? nullptr : s.g32[1 - s.cur].p, s.rhs32.p, s.stream));
      }
    else
      {
        S.collide = ll ? 0 : 1;
        HIP_CHECK(tile_launch(ll ? kAdvectLeeLin : kAdvectBardow, v, G, S, s.f[s.cur].p, (const double *)nullptr,
                              ll ? nullptr : s.f[1 - s.cur].p, s.rhs.p, s.stream));
      }
    return LBFEM_OK;
  }

  struct TuneKey
  {
    int  mode, p, P, px, py;
    bool operator<(const TuneKey &o) const
    {
      return std::tie(mode, p, P, px, py) < std::tie(o.mode, o.p, o.P, o.px, o.py);
    }
  };

  std::map<TuneKey, int> &
  tune_cache()
  {
    static std::map<TuneKey, int> cache;
    return cache;
  }

  TuneKey
  tune_key(const lbfem_scheme &s)
  {
    const Geo &G     = s.d->G;
    const bool fused = s.set.mass.type == LBFEM_MASS_LUMPED && s.set.streaming == LBFEM_TG2;
    const int  mode  = (fused ? 0 : 2) + (s.set.scheme == LBFEM_LEELIN ? 1 : 0) + (s.mixed ? 4 : 0);
    return TuneKey{mode, G.p, G.P, G.ax.periodic, G.ay.periodic};
  }

  int
  step_once(lbfem_scheme &s)
  {
    for (auto &row : s.last_step_its)
      for (int &v : row)
        v = 0;
    const lbfem_disc &d = *s.d;
    const Geo        &G = d.G;
    const bool        ll = s.set.scheme == LBFEM_LEELIN;
    if (s.set.mass.type == LBFEM_MASS_LUMPED && s.set.streaming == LBFEM_TG2 && d.general)
      {
        // general quadrilaterals: one fused cell-kernel launch per step
        TileParams S = tile_params(s, s.sweeps[0], ll);
        S.collide    = 1;
        s.mark(-1);
        HIP_CHECK(quad_advect_launch(ll ? kStepLeeLin : kStepBardow, G, S, s.f[s.cur].p, nullptr, s.f[1 - s.cur].p,
                                     nullptr, s.stream));
        s.mark(1);
        s.cur = 1 - s.cur;
        return LBFEM_OK;
      }
    if (s.set.mass.type == LBFEM_MASS_LUMPED && s.set.streaming == LBFEM_TG2)
      {
        // one fused kernel: nodal prologue, advection, M_L^-1, nodal epilogue
        TileParams S = tile_params(s, s.sweeps[0], ll);
        S.collide    = 1;
        s.mark(-1);
        if (s.mixed)
          HIP_CHECK(tile_launch(ll ? kStepLeeLin : kStepBardow, s.variant(), G, S, s.g32[s.cur].p,
                                (const float *)nullptr, s.g32[1 - s.cur].p, (float *)nullptr, s.stream));
        else
          HIP_CHECK(tile_launch(ll ? kStepLeeLin : kStepBardow, s.variant(), G, S, s.f[s.cur].p,
                                (const double *)nullptr, s.f[1 - s.cur].p, (double *)nullptr, s.stream));
        s.mark(1);
        s.cur = 1 - s.cur;
        return LBFEM_OK;
      }
    if (s.mixed) // the mixed form with the consistent mass
      {
        if (ll)
          {
            const int st = increment_mixed(s, 0, false);
            if (st != LBFEM_OK)
              return st;
            HIP_CHECK(pc_from_f_launch(G, s.g32[s.cur].p, s.x32[0].p, s.dt / s.lambda, s.stream));
            s.mark(0);
            return LBFEM_OK;
          }
        for (int sw = 0; sw < int(s.sweeps.size()); ++sw)
          {
            const int st = increment_mixed(s, sw, sw == 0);
            if (st != LBFEM_OK)
              return st;
            unsigned mask = 0;
            for (int a = 0; a < NM; ++a)
              if (moves(s.sweeps[sw], a))
                mask |= 1u << a;
            HIP_CHECK(add_increments_launch(G, s.g32[s.cur].p, s.x32[sw].p, mask, s.stream));
            s.mark(0);
          }
        return LBFEM_OK;
      }
    if (ll) // schemes.h:92-98
      {
        const int st = increment(s, 0, false);
        if (st != LBFEM_OK)
          return st;
        HIP_CHECK(pc_from_f_launch(G, s.f[s.cur].p, s.x[0].p, s.dt / s.lambda, s.stream));
        s.mark(0);
        return LBFEM_OK;
      }
    // schemes.h:118-124, streaming.h:88-100: collide (fused into the first advection), stream sweep by sweep
    for (int sw = 0; sw < int(s.sweeps.size()); ++sw)
      {
        const int st = increment(s, sw, sw == 0);
        if (st != LBFEM_OK)
          return st;
        unsigned mask = 0;
        for (int a = 0; a < NM; ++a)
          if (moves(s.sweeps[sw], a))
            mask |= 1u << a;
        HIP_CHECK(add_increments_launch(G, s.f[s.cur].p, s.x[sw].p, mask, s.stream));
        s.mark(0);
      }
    return LBFEM_OK;
  }

  template <typename F>
  int
  guarded(F &&fn)
  {
    try
      {
        return fn();
      }
    catch (const HipError &e)
      {
        return fail(LBFEM_EHIP, e.what());
      }
    catch (const std::bad_alloc &)
      {
        return fail(LBFEM_EHIP, "out of host memory");
      }
    catch (const std::exception &e)
      {
        return fail(LBFEM_EINVAL, e.what());
      }
  }

  hipStream_t
  as_stream(void *s)
  {
    return reinterpret_cast<hipStream_t>(s);
  }
} // namespace

// ---------------------------------------------------------------------- batches
// Many lumped-mass TG2 schemes stepped together: the members are grouped by
// (degree, scheme, precision), and a step is one launch per group of the
// batched tile kernel over the group's concatenated tile list. Members join and
// leave a running batch: a group's device table (its live counts, then the
// entries) is rewritten, and the captured step graphs, which read the counts at
// run time, are kept unless a group appears or disappears.
namespace
{
  // the monitor entry of a scheme at its current populations (its u_prev allocated on first use)
  MonitorEntry
  monitor_entry(lbfem_scheme &s);
} // namespace

struct lbfem_batch
{
  struct Group
  {
    int                         p = 1, mode = 0, variant = kTileWide, grid = 0; // grid 0: one block per tile
    bool                        mixed = false;
    std::vector<lbfem_scheme *> members;
    int                         n_tiles = 0, launch_grid = 0, capacity = 0;
    int64_t                     n_nodes = 0;
    DevBuf<char>                table; // BatchCounts (kCountsBytes), then BatchEntry<double or float>[capacity]
  };
  static constexpr size_t kCountsBytes = 256;

  // Members with the consistent mass (FP64, tensor meshes) of one (degree, scheme,
  // streaming): each step is the single-mesh step's sequence of stages (advection,
  // extrapolated start, CG start, direction and update per iteration, nodal update), each
  // stage one launch over all members from per-sweep device tables, with one host check
  // of all members' stopping states per CG iteration checked. A member's buffers flip
  // with the group's own parity (once per Bardow step; Lee-Lin updates in place).
  struct CgGroup
  {
    int                         p = 1, scheme = 0, streaming = 0, n_sweeps = 1;
    std::vector<lbfem_scheme *> members;
    std::vector<int>            par0; // each member's buffer read at group parity 0
    std::vector<char>           failed; // NoConvergence in some step
    int                         gparity = 0, last_its = 1, variant = kTileWide;
    DevBuf<char>                adv[2];      // BatchCounts (kCountsBytes), then BatchEntry<double>[] per sweep
    DevBuf<CgMember>            cgt[2];      // per sweep
    DevBuf<NodalMember>         nod[2];      // per sweep
    DevBuf<CgState>             states;      // members x NM
    CgState                    *host_states = nullptr;
    int                         adv_tiles[2] = {0, 0}, op_blocks[2] = {0, 0}, upd_blocks[2] = {0, 0};
    int                         nod_blocks = 0;
    int64_t                     n_nodes = 0;

    CgGroup() = default;
    CgGroup(const CgGroup &) = delete;
    CgGroup &operator=(const CgGroup &) = delete;
    ~CgGroup()
    {
      RuntimeLock lock(runtime_mutex());
      if (host_states)
        (void)hipHostFree(host_states);
    }
  };

  std::vector<lbfem_scheme *> members;
  std::vector<int>            par0; // each member's buffer read at batch parity 0 (lumped members)
  std::vector<Group>          groups;
  std::vector<std::unique_ptr<CgGroup>> cg_groups;
  hipStream_t                 cg_stream = nullptr; // the CG groups' work, beside the lumped groups'
  hipStream_t                 stream      = nullptr;
  int                         parity      = 0;
  int                         graph_steps = 8; // steps per captured graph, 0: no graphs
  int                         shape_variant = -1, shape_grid = 0; // the shape of groups created from here
  hipGraph_t                  graph[2]    = {nullptr, nullptr};
  hipGraphExec_t              exec[2]     = {nullptr, nullptr};
  Monitor                     monitor;

  void
  drop_graphs()
  {
    RuntimeLock lock(runtime_mutex());
    for (int k = 0; k < 2; ++k)
      {
        if (exec[k])
          (void)hipGraphExecDestroy(exec[k]);
        if (graph[k])
          (void)hipGraphDestroy(graph[k]);
        exec[k]  = nullptr;
        graph[k] = nullptr;
      }
  }

  int
  index_of(const lbfem_scheme *s) const
  {
    for (size_t k = 0; k < members.size(); ++k)
      if (members[k] == s)
        return int(k);
    return -1;
  }

  CgGroup *
  cg_group_of(const lbfem_scheme *s) const
  {
    for (const auto &g : cg_groups)
      if (std::find(g->members.begin(), g->members.end(), s) != g->members.end())
        return g.get();
    return nullptr;
  }

  ~lbfem_batch()
  {
    RuntimeLock lock(runtime_mutex());
    drop_graphs();
    if (stream)
      (void)hipStreamDestroy(stream);
    if (cg_stream)
      (void)hipStreamDestroy(cg_stream);
  }
};

namespace
{
  // The group's device table for its current members: entries at the members' buffer
  // parities, tile lists at the group's variant, and the live counts. Rewritten on the
  // batch's stream after the work queued there (the copy is synchronized). A larger table
  // or another launch grid (one block per tile) drops the captured graphs.
  template <typename T>
  void
  write_table(lbfem_batch &b, lbfem_batch::Group &g)
  {
    std::vector<BatchEntry<T>> E(g.members.size());
    int                        begin = 0;
    int                        tx, ty;
    // Q1 and Q2 groups at a march variant list each member's wavefronts (strips of march_out(p)
    // columns by row segments), where the row-march kernel runs for the group's form; a member's
    // population block of 4 GiB or more puts the group on the LDS tile kernel
    const bool marchable = march_runs(g.p, sizeof(T) == 4);
    if (marchable && g.variant >= kTileLdsVariants)
      for (const lbfem_scheme *s : g.members)
        if (!march_fits(s->d->G, sizeof(T)))
          {
            g.variant = kTileWide;
            b.drop_graphs();
            break;
          }
    tile_dims(g.variant, tx, ty);
    const bool march = marchable && g.variant >= kTileLdsVariants;
    const int  rows  = march ? march_rows(g.variant) : 0;
    g.n_nodes = 0;
    for (size_t k = 0; k < g.members.size(); ++k)
      {
        lbfem_scheme &s  = *g.members[k];
        const Geo    &G  = s.d->G;
        const bool    ll = s.set.scheme == LBFEM_LEELIN;
        E[k].G           = G;
        E[k].S           = tile_params(s, s.sweeps[0], ll);
        E[k].S.collide   = 1;
        E[k].S.tiles_x   = march ? (G.P + 1 + march_out(g.p) - 1) / march_out(g.p) : (G.P + 1 + tx - 1) / tx;
        E[k].S.tiles_y   = march ? (G.P + 1 + rows - 1) / rows : (G.P + 1 + ty - 1) / ty;
        E[k].S.rows      = rows;
        if constexpr (sizeof(T) == 4)
          {
            E[k].f[0] = s.g32[0].p;
            E[k].f[1] = s.g32[1].p;
          }
        else
          {
            E[k].f[0] = s.f[0].p;
            E[k].f[1] = s.f[1].p;
          }
        E[k].par        = b.par0[b.index_of(&s)];
        E[k].tile_begin = begin;
        begin += E[k].S.tiles_x * E[k].S.tiles_y;
        g.n_nodes += int64_t(G.np) * G.np;
      }
    g.n_tiles            = begin;
    // one block per tile, or per four wavefronts of the march list
    const int nb8        = ((march ? (g.n_tiles + 3) / 4 : g.n_tiles) + 7) & ~7;
    const int grid       = g.grid > 0 ? (g.grid + 7) & ~7 : std::max(8, nb8);
    if (grid != g.launch_grid)
      b.drop_graphs();
    g.launch_grid = grid;
    if (int(E.size()) > g.capacity)
      {
        HIP_CHECK(hipStreamSynchronize(b.stream));
        b.drop_graphs(); // the graphs hold the table's address
        const int cap = std::max<int>(8, 2 * int(E.size()));
        g.table.alloc(lbfem_batch::kCountsBytes + size_t(cap) * sizeof(BatchEntry<T>));
        g.capacity = cap;
      }
    std::vector<char> host(lbfem_batch::kCountsBytes + E.size() * sizeof(BatchEntry<T>), 0);
    const BatchCounts c{int(E.size()), g.n_tiles};
    std::memcpy(host.data(), &c, sizeof(c));
    if (!E.empty())
      std::memcpy(host.data() + lbfem_batch::kCountsBytes, E.data(), E.size() * sizeof(BatchEntry<T>));
    HIP_CHECK(hipMemcpyAsync(g.table.p, host.data(), host.size(), hipMemcpyHostToDevice, b.stream));
    HIP_CHECK(hipStreamSynchronize(b.stream)); // the host buffer is released on return
  }

  void
  write_group(lbfem_batch &b, lbfem_batch::Group &g)
  {
    if (g.mixed)
      write_table<float>(b, g);
    else
      write_table<double>(b, g);
  }

  void
  rebuild_tables(lbfem_batch &b)
  {
    HIP_CHECK(hipStreamSynchronize(b.stream));
    for (auto &g : b.groups)
      write_group(b, g);
  }

  int
  launch_group(lbfem_batch &b, const lbfem_batch::Group &g, const int parity)
  {
    if (g.members.empty())
      return LBFEM_OK;
    const BatchCounts *counts = reinterpret_cast<const BatchCounts *>(g.table.p);
    if (g.mixed)
      HIP_CHECK(tile_batch_launch<float>(
        g.p, g.mode, g.variant,
        reinterpret_cast<const BatchEntry<float> *>(g.table.p + lbfem_batch::kCountsBytes), counts, parity,
        g.launch_grid, b.stream));
    else
      HIP_CHECK(tile_batch_launch<double>(
        g.p, g.mode, g.variant,
        reinterpret_cast<const BatchEntry<double> *>(g.table.p + lbfem_batch::kCountsBytes), counts, parity,
        g.launch_grid, b.stream));
    return LBFEM_OK;
  }

  int
  batch_step_eager(lbfem_batch &b, const int parity)
  {
    for (auto &g : b.groups)
      if (const int st = launch_group(b, g, parity); st != LBFEM_OK)
        return st;
    return LBFEM_OK;
  }

  // The device tables of a CG group for its current members (on its stream, synchronized):
  // per sweep, the advection entries (the single-mesh tile kernel's, any variant gives the
  // same bits), the CG members (each with its single-mesh operator tiles, partial rows,
  // update grid and folded alpha reduction, its states in the group's array) and the nodal
  // members.
  void
  write_cg_group(lbfem_batch &b, lbfem_batch::CgGroup &g)
  {
    HIP_CHECK(hipStreamSynchronize(b.cg_stream));
    const int n = int(g.members.size());
    if (g.states.n < size_t(n) * NM || !g.host_states)
      {
        RuntimeLock  lock(runtime_mutex());
        const size_t cap = size_t(std::max(8, 2 * n)) * NM;
        g.states.alloc(cap);
        if (g.host_states)
          (void)hipHostFree(g.host_states);
        g.host_states = nullptr;
        HIP_CHECK(hipHostMalloc(&g.host_states, cap * sizeof(CgState)));
      }
    g.n_nodes = 0;
    for (const lbfem_scheme *s : g.members)
      g.n_nodes += int64_t(s->d->G.np) * s->d->G.np;
    g.variant = batch_default_variant(g.n_nodes, false);
    int tx, ty;
    tile_dims(g.variant, tx, ty);
    const bool ll = g.scheme == LBFEM_LEELIN;
    for (int sw = 0; sw < g.n_sweeps; ++sw)
      {
        std::vector<BatchEntry<double>> E(n);
        std::vector<CgMember>           M(n);
        std::vector<NodalMember>        N(n);
        int                             tiles = 0, op_blocks = 0, upd_blocks = 0, nod_blocks = 0;
        for (int k = 0; k < n; ++k)
          {
            lbfem_scheme &s = *g.members[k];
            const Geo    &G = s.d->G;
            const auto   &e = s.sweeps[sw];
            unsigned      mask = 0;
            for (int a = 0; a < NM; ++a)
              if (moves(e, a))
                mask |= 1u << a;
            // advection (streaming.h:66-72), the collision fused into Bardow's first sweep
            E[k].G          = G;
            E[k].S          = tile_params(s, e, ll);
            E[k].S.collide  = (!ll && sw == 0) ? 1 : 0;
            E[k].S.tiles_x  = (G.P + 1 + tx - 1) / tx;
            E[k].S.tiles_y  = (G.P + 1 + ty - 1) / ty;
            E[k].f[0]       = s.f[0].p;
            E[k].f[1]       = s.f[1].p;
            E[k].rhs        = s.rhs.p;
            E[k].par        = g.par0[k];
            E[k].tile_begin = tiles;
            tiles += E[k].S.tiles_x * E[k].S.tiles_y;
            // the solves (mass.h:204-211), as increment() sets them up for the scheme alone
            CgWork W = cg_work(*s.d, g.states.p + size_t(k) * NM, nullptr, s.partial.p, s.set.mass);
            W.nv     = 0;
            for (int a = 0; a < NM; ++a)
              if (mask & (1u << a))
                {
                  W.block[W.nv]    = a;
                  W.vec[W.nv].e[0] = e[2 * (a + 1)];
                  W.vec[W.nv].e[1] = e[2 * (a + 1) + 1];
                  W.vec[W.nv].coef = s.tg3 * s.dt * s.dt;
                  ++W.nv;
                }
            const int rows = cg_direction_rows(G);
            W.alpha_rows   = int64_t(rows) * W.nv <= 4096 && rows + 1024 <= W.max_blocks ? rows : 0;
            M[k].G         = G;
            M[k].W         = W;
            M[k].b         = s.rhs.p;
            M[k].x         = s.x[sw].p;
            M[k].r         = s.r.p;
            M[k].p[0]      = s.pb[0].p;
            M[k].p[1]      = s.pb[1].p;
            M[k].q         = s.q.p;
            cg_operator_tiles(G, M[k].tiles_x, M[k].tiles_y);
            M[k].nt        = M[k].tiles_x * M[k].tiles_y;
            M[k].op_begin  = op_blocks;
            op_blocks += (M[k].nt + 7) & ~7;
            M[k].upd_blocks = cg_update_blocks(G, W.max_blocks, W.alpha_rows);
            M[k].upd_begin  = upd_blocks;
            upd_blocks += M[k].upd_blocks;
            // the nodal stages
            N[k].G           = G;
            N[k].f[0]        = s.f[0].p;
            N[k].f[1]        = s.f[1].p;
            N[k].x           = s.x[sw].p;
            N[k].x_prev      = s.x[1].p;
            N[k].par         = g.par0[k];
            N[k].mask        = mask;
            N[k].theta       = s.dt / s.lambda;
            N[k].block_begin = nod_blocks;
            N[k].nblocks     = nodal_batch_blocks(G);
            nod_blocks += N[k].nblocks;
          }
        g.adv_tiles[sw]  = tiles;
        g.op_blocks[sw]  = op_blocks;
        g.upd_blocks[sw] = upd_blocks;
        g.nod_blocks     = nod_blocks;
        std::vector<char> host(lbfem_batch::kCountsBytes + E.size() * sizeof(BatchEntry<double>), 0);
        const BatchCounts c{n, tiles};
        std::memcpy(host.data(), &c, sizeof(c));
        if (n)
          std::memcpy(host.data() + lbfem_batch::kCountsBytes, E.data(), E.size() * sizeof(BatchEntry<double>));
        if (g.adv[sw].n < host.size())
          g.adv[sw].alloc(host.size() + 16 * sizeof(BatchEntry<double>));
        if (g.cgt[sw].n < M.size())
          g.cgt[sw].alloc(M.size() + 16);
        if (g.nod[sw].n < N.size())
          g.nod[sw].alloc(N.size() + 16);
        HIP_CHECK(hipMemcpyAsync(g.adv[sw].p, host.data(), host.size(), hipMemcpyHostToDevice, b.cg_stream));
        if (n)
          {
            HIP_CHECK(hipMemcpyAsync(g.cgt[sw].p, M.data(), M.size() * sizeof(CgMember), hipMemcpyHostToDevice,
                                     b.cg_stream));
            HIP_CHECK(hipMemcpyAsync(g.nod[sw].p, N.data(), N.size() * sizeof(NodalMember), hipMemcpyHostToDevice,
                                     b.cg_stream));
          }
        HIP_CHECK(hipStreamSynchronize(b.cg_stream)); // the host tables are released on return
      }
  }

  // One step of every member of a CG group (schemes.h:92-98, 118-124; streaming.h:66-100),
  // the single-mesh step's stages each one launch over all members. Returns LBFEM_OK, or
  // LBFEM_ENOCONV when a member's solve hit its cap (that member is marked failed; the step
  // of the others is complete).
  int
  cg_group_step(lbfem_batch &b, lbfem_batch::CgGroup &g)
  {
    const int   n  = int(g.members.size());
    const bool  ll = g.scheme == LBFEM_LEELIN;
    hipStream_t st = b.cg_stream;
    if (n == 0)
      return LBFEM_OK;
    for (lbfem_scheme *s : g.members)
      for (auto &row : s->last_step_its)
        for (int &v : row)
          v = 0;
    const int  op         = g.streaming != LBFEM_TG2 ? 1 : 0;
    const bool extrapolate = g.streaming != LBFEM_TG3_SPLIT;
    int        max_its    = 0;
    for (const lbfem_scheme *s : g.members)
      max_its = std::max(max_its, s->set.mass.cg_max_iterations);
    bool failed = false;
    for (int sw = 0; sw < g.n_sweeps; ++sw)
      {
        const BatchCounts *counts = reinterpret_cast<const BatchCounts *>(g.adv[sw].p);
        const auto *E = reinterpret_cast<const BatchEntry<double> *>(g.adv[sw].p + lbfem_batch::kCountsBytes);
        HIP_CHECK(tile_batch_launch<double>(g.p, ll ? kAdvectLeeLin : kAdvectBardow, g.variant, E, counts,
                                            g.gparity, std::max(8, (g.adv_tiles[sw] + 7) & ~7), st));
        if (!ll && sw == 0)
          g.gparity ^= 1; // g* went to the other buffer
        if (extrapolate)
          HIP_CHECK(nodal_batch_launch(0, g.nod[sw].p, n, g.nod_blocks, g.gparity, st));
        HIP_CHECK(cg_batch_start_launch(g.p, op, g.cgt[sw].p, n, g.op_blocks[sw], st));
        // the vectors each member solves in this sweep (its W.nv): the states past them are unused
        std::vector<int> nvs(n);
        for (int k = 0; k < n; ++k)
          {
            const auto &e = g.members[k]->sweeps[sw];
            int         c = 0;
            for (int a = 0; a < NM; ++a)
              c += moves(e, a) ? 1 : 0;
            nvs[k] = c;
          }
        const auto active = [&]() -> bool {
          HIP_CHECK(hipMemcpyAsync(g.host_states, g.states.p, size_t(n) * NM * sizeof(CgState),
                                   hipMemcpyDeviceToHost, st));
          HIP_CHECK(hipStreamSynchronize(st));
          for (int k = 0; k < n; ++k)
            for (int v = 0; v < nvs[k]; ++v)
              if (g.host_states[size_t(k) * NM + v].status == 0)
                return true;
          return false;
        };
        // iterations of converged vectors are no-ops, so the first check waits for the last
        // step's largest count, as for a scheme alone
        bool running = true;
        for (int it = 1; it <= max_its && running; ++it)
          {
            HIP_CHECK(cg_batch_direction_launch(g.p, op, g.cgt[sw].p, n, g.op_blocks[sw], it, st));
            HIP_CHECK(cg_batch_update_launch(g.cgt[sw].p, n, g.upd_blocks[sw], it, st));
            if (it >= g.last_its)
              running = active();
          }
        if (running)
          (void)active(); // the update at the cap leaves no vector active; read the final states
        int largest = 0;
        for (int k = 0; k < n; ++k)
          {
            lbfem_scheme &s  = *g.members[k];
            const auto   &e  = s.sweeps[sw];
            int           v  = 0;
            s.last_its       = 0;
            for (int a = 0; a < NM; ++a)
              if (moves(e, a))
                {
                  const CgState &c = g.host_states[size_t(k) * NM + v];
                  s.cg_iterations += c.its;
                  s.last_step_its[sw][a] = c.its;
                  s.last_its             = std::max(s.last_its, c.its);
                  largest                = std::max(largest, c.its);
                  if (c.status == 2)
                    {
                      g.failed[k] = 1;
                      failed      = true;
                    }
                  ++v;
                }
          }
        g.last_its = std::max(1, largest);
        if (ll) // the predictor-corrector after the solves
          HIP_CHECK(nodal_batch_launch(2, g.nod[sw].p, n, g.nod_blocks, g.gparity, st));
        else // f_{a+1} += x_a (streaming.h:94-98)
          HIP_CHECK(nodal_batch_launch(1, g.nod[sw].p, n, g.nod_blocks, g.gparity, st));
      }
    return failed ? fail(LBFEM_ENOCONV, "batch: a member's CG reached its iteration cap (SolverControl::NoConvergence)")
                  : LBFEM_OK;
  }

  // graph[k]: graph_steps steps starting at parity k (graph_steps even, so parity returns to k).
  // Both graphs or neither: any failure drops what was captured.
  int
  capture_graphs(lbfem_batch &b)
  {
    RuntimeLock lock(runtime_mutex()); // no allocation in another thread while the stream captures
    b.drop_graphs();
    try
      {
        for (int k = 0; k < 2; ++k)
          {
            HIP_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
            int st = LBFEM_OK;
            try
              {
                for (int n = 0; n < b.graph_steps && st == LBFEM_OK; ++n)
                  st = batch_step_eager(b, (k + n) & 1);
              }
            catch (...)
              {
                hipGraph_t gr = nullptr;
                (void)hipStreamEndCapture(b.stream, &gr); // leave the stream out of capture mode
                if (gr)
                  (void)hipGraphDestroy(gr);
                throw;
              }
            hipGraph_t       gr = nullptr;
            const hipError_t ec = hipStreamEndCapture(b.stream, &gr);
            b.graph[k]          = gr; // owned by the batch from here (drop_graphs frees it)
            if (st != LBFEM_OK)
              {
                b.drop_graphs();
                return st;
              }
            HIP_CHECK(ec);
            HIP_CHECK(hipGraphInstantiate(&b.exec[k], gr, nullptr, nullptr, 0));
          }
      }
    catch (...)
      {
        b.drop_graphs();
        throw;
      }
    return LBFEM_OK;
  }

  void
  sync_members(lbfem_batch &b)
  {
    for (size_t k = 0; k < b.members.size(); ++k)
      b.members[k]->cur = (b.par0[k] + b.parity) & 1;
    for (const auto &g : b.cg_groups)
      for (size_t k = 0; k < g->members.size(); ++k)
        g->members[k]->cur = (g->par0[k] + g->gparity) & 1;
  }

  bool
  is_lumped_fused(const lbfem_scheme *s)
  {
    return s->set.mass.type == LBFEM_MASS_LUMPED && s->set.streaming == LBFEM_TG2;
  }

  // what a batch steps: tensor meshes, time step set; the lumped mass with TG2 streaming (one
  // fused kernel per step), or, in FP64, the consistent mass with any streaming
  int
  check_batchable(const lbfem_scheme *s)
  {
    if (!s)
      return fail(LBFEM_EINVAL, "null scheme");
    if (s->d->general)
      return fail(LBFEM_EINVAL, "a batch steps tensor-product meshes only (no distort)");
    if (!(s->dt > 0.))
      return fail(LBFEM_EINVAL, "set the time step of every scheme before batching");
    if (is_lumped_fused(s))
      return LBFEM_OK;
    if (s->set.mass.type != LBFEM_MASS_CG)
      return fail(LBFEM_EINVAL, "a batch steps the lumped or the consistent (CG) mass, not Richardson");
    if (s->mixed)
      return fail(LBFEM_EINVAL, "a batch steps consistent-mass schemes in FP64 (a mixed consistent-mass scheme is "
                                "stepped alone)");
    return LBFEM_OK;
  }

  // a consistent-mass member joins its (degree, scheme, streaming) CG group, its buffer
  // parity taken relative to the group's (the tables are written by the caller)
  void
  join_cg(lbfem_batch &b, lbfem_scheme *s)
  {
    if (!b.cg_stream)
      {
        RuntimeLock lock(runtime_mutex());
        HIP_CHECK(hipStreamCreateWithFlags(&b.cg_stream, hipStreamNonBlocking));
      }
    lbfem_batch::CgGroup *g = nullptr;
    for (const auto &c : b.cg_groups)
      if (c->p == s->d->G.p && c->scheme == s->set.scheme && c->streaming == s->set.streaming)
        g = c.get();
    if (!g)
      {
        b.cg_groups.push_back(std::make_unique<lbfem_batch::CgGroup>());
        g            = b.cg_groups.back().get();
        g->p         = s->d->G.p;
        g->scheme    = s->set.scheme;
        g->streaming = s->set.streaming;
        g->n_sweeps  = int(s->sweeps.size());
      }
    g->members.push_back(s);
    g->par0.push_back((s->cur ^ g->gparity) & 1);
    g->failed.push_back(0);
  }

  // a member joins: its buffer parity taken relative to the batch's, its group's table rewritten
  // (a new group drops the captured graphs)
  void
  join(lbfem_batch &b, lbfem_scheme *s)
  {
    HIP_CHECK(hipStreamSynchronize(s->stream)); // its own queued work first
    b.members.push_back(s);
    b.par0.push_back((s->cur ^ b.parity) & 1);
    if (!is_lumped_fused(s))
      {
        join_cg(b, s);
        return;
      }
    const int mode = s->set.scheme == LBFEM_LEELIN ? kStepLeeLin : kStepBardow;
    auto      it   = std::find_if(b.groups.begin(), b.groups.end(), [&](const auto &g) {
      return g.p == s->d->G.p && g.mode == mode && g.mixed == s->mixed;
    });
    if (it == b.groups.end())
      {
        b.drop_graphs();
        b.groups.emplace_back();
        it        = b.groups.end() - 1;
        it->p     = s->d->G.p;
        it->mode  = mode;
        it->mixed = s->mixed;
        it->grid  = b.shape_grid;
        it->variant =
          b.shape_variant >= 0 ? b.shape_variant
                             : batch_default_variant(int64_t(s->d->G.np) * s->d->G.np, s->mixed, s->d->G.p);
      }
    it->members.push_back(s);
  }
} // namespace

namespace
{
  MonitorEntry
  monitor_entry(lbfem_scheme &s)
  {
    MonitorEntry e{};
    e.G = s.d->G;
    if (s.mixed)
      {
        e.f    = s.g32[s.cur].p;
        e.fp32 = 1;
      }
    else
      e.f = s.f[s.cur].p;
    if (!s.u_prev.p)
      {
        s.u_prev.alloc(2 * size_t(s.d->n));
        s.u_prev_set = false;
      }
    e.u_prev = s.u_prev.p;
    e.init   = s.u_prev_set ? 0 : 1;
    return e;
  }
} // namespace

// ------------------------------------------------------- allocation-free mass solve
// The work space of lbfem_mass_solve_noalloc: the CG state, the partial-sum rows of the
// CG kernels and four node vectors (r, two directions, A p; Richardson's A x uses the
// first), in one fixed device array per GPU, held until the process ends. Every call on a
// GPU shares that GPU's array.
namespace
{
  constexpr int64_t kNoAllocMaxStride    = 1179648; // nodes (a multiple of 64): 1024^2 Q1 and 512^2 Q2 cells fit
  constexpr int     kNoAllocRows         = 8192;    // partial-sum rows: operator tiles, then at most 1024 update blocks
  constexpr size_t  kNoAllocStateDoubles = 16;      // the CgStates of the one vector (56 bytes each, 64 apart)
  constexpr size_t  kNoAllocDoubles =
    kNoAllocStateDoubles + size_t(kNoAllocRows) * kCgPartials + 4 * size_t(kNoAllocMaxStride);
  constexpr int kNoAllocMaxDevices = 64;
  static_assert(sizeof(CgState) <= 8 * sizeof(double), "CG state slot");

  std::mutex g_scratch_mutex;
  double    *g_scratch_ptr[kNoAllocMaxDevices] = {}; // the array's address on each device, once taken
} // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace
{
  // The device address of `scratch` on the current device, taken by hipGetSymbolAddress once
  // per device with that device current (the host-side name of the array is not a device
  // address: only this pointer is passed to kernels).
  int
  noalloc_scratch(const int dev, double **out)
  {
    if (dev < 0 || dev >= kNoAllocMaxDevices)
      return fail(LBFEM_EINVAL, "lbfem_mass_solve_noalloc: device index beyond the work-space table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev])
      {
        void            *ptr = nullptr;
        const hipError_t e   = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
          return fail_hip(e, "hipGetSymbolAddress(scratch)");
        if (!ptr)
          return fail(LBFEM_EHIP, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = static_cast<double *>(ptr);
      }
    *out = g_scratch_ptr[dev];
    return LBFEM_OK;
  }
} // namespace

extern "C"
{
  size_t
  lbfem_last_error(char *buf, const size_t len)
  {
    if (buf && len)
      {
        const size_t k = std::min(len - 1, g_error_len);
        std::memcpy(buf, g_error, k);
        buf[k] = '\0';
      }
    return g_error_len;
  }

  int
  lbfem_device_count(int *count)
  {
    return guarded([&]() -> int {
      int        c = 0;
      const auto e = hipGetDeviceCount(&c);
      *count       = e == hipSuccess ? c : 0;
      return LBFEM_OK;
    });
  }

  int
  lbfem_disc_create(const lbfem_mesh_desc *desc, lbfem_disc **out)
  {
    RuntimeLock lock(runtime_mutex());
    return guarded([&]() -> int {
      if (!desc || !out)
        return fail(LBFEM_EINVAL, "null argument");
      if (desc->n_cells < 2 || desc->degree < 1 || desc->degree > kMaxDegree || desc->n_periodic < 0 ||
          desc->n_periodic > 2 || !(desc->L > 0.) || desc->stretch < 0.)
        return fail(LBFEM_EINVAL, "invalid mesh description (N >= 2, degree 1..3, n_periodic 0..2, L > 0)");
      if (desc->distort != 0. && (desc->n_periodic != 2 || desc->stretch != 0. || !std::isfinite(desc->distort)))
        return fail(LBFEM_EINVAL, "the distortion maps the vortex's periodic box (n_periodic 2, stretch 0, "
                                  "cgdbe.cc:77, test_cases.h:53-65)");
      auto d    = std::make_unique<lbfem_disc>();
      HIP_CHECK(hipGetDevice(&d->device));
      d->desc   = *desc;
      d->p      = desc->degree;
      d->N      = desc->n_cells;
      d->P      = d->p * d->N;
      d->np     = d->P + 1;
      d->n      = int64_t(d->np) * d->np;
      d->stride = (d->n + 63) / 64 * 64;
      d->B      = std::make_unique<Basis1D>(d->p);
      const auto v = box_vertices(d->N, desc->L, desc->stretch);
      d->ax = std::make_unique<Axis1D>(*d->B, v, desc->n_periodic >= 1);
      d->ay = std::make_unique<Axis1D>(*d->B, v, desc->n_periodic >= 2);
      // discretization.h:99-100: h_min = L / N, moved by the clustering
      d->h_min = v[1] - v[0];
      d->upload_axis(*d->ax, d->axbuf, d->G.ax);
      d->upload_axis(*d->ay, d->aybuf, d->G.ay);
      d->G.p      = d->p;
      d->G.P      = d->P;
      d->G.np     = d->np;
      d->G.stride = d->stride;
      for (auto &s : d->G.walls.side)
        s = nullptr;
      d->G.walls.velocity = nullptr;
      d->G.gen            = GeneralDev{};
      if (desc->distort != 0.)
        d->build_general(desc->distort);
      d->build_jacobi();
      *out = d.release();
      return LBFEM_OK;
    });
  }

  int
  lbfem_disc_destroy(lbfem_disc *d)
  {
    RuntimeLock lock(runtime_mutex());
    delete d;
    return LBFEM_OK;
  }

  int
  lbfem_disc_get_info(const lbfem_disc *d, lbfem_disc_info *info)
  {
    if (!d || !info)
      return fail(LBFEM_EINVAL, "null argument");
    info->n_nodes   = d->n;
    const int64_t m = d->ax->periodic ? d->P : d->np, k = d->ay->periodic ? d->P : d->np;
    info->n_independent = m * k;
    info->stride        = d->stride;
    info->n_per_dir     = d->np;
    info->n_walls       = int32_t(d->wall_nodes.size());
    info->h_min         = d->h_min;
    return LBFEM_OK;
  }

  int
  lbfem_disc_node_coordinates(const lbfem_disc *d, double *xy)
  {
    if (!d || !xy)
      return fail(LBFEM_EINVAL, "null argument");
    if (d->general)
      {
        std::copy(d->gen_xy.begin(), d->gen_xy.end(), xy);
        return LBFEM_OK;
      }
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        {
          xy[2 * (int64_t(j) * d->np + i)]     = d->ax->coord[i];
          xy[2 * (int64_t(j) * d->np + i) + 1] = d->ay->coord[j];
        }
    return LBFEM_OK;
  }

  int
  lbfem_disc_constrained(const lbfem_disc *d, int32_t *flag)
  {
    if (!d || !flag)
      return fail(LBFEM_EINVAL, "null argument");
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        flag[int64_t(j) * d->np + i] = (d->ax->periodic && i == d->P) || (d->ay->periodic && j == d->P);
    return LBFEM_OK;
  }

  int
  lbfem_disc_lumped_mass(const lbfem_disc *d, double *ml)
  {
    if (!d || !ml)
      return fail(LBFEM_EINVAL, "null argument");
    if (d->general)
      {
        std::copy(d->gen_ml.begin(), d->gen_ml.end(), ml);
        return LBFEM_OK;
      }
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        {
          const bool slave = (d->ax->periodic && i == d->P) || (d->ay->periodic && j == d->P);
          ml[int64_t(j) * d->np + i] = slave ? 1. : d->ax->row_sum[i] * d->ay->row_sum[j];
        }
    return LBFEM_OK;
  }

  int
  lbfem_disc_mass_diagonal(const lbfem_disc *d, double *diag)
  {
    if (!d || !diag)
      return fail(LBFEM_EINVAL, "null argument");
    if (d->general)
      {
        std::copy(d->gen_diag.begin(), d->gen_diag.end(), diag);
        return LBFEM_OK;
      }
    for (int j = 0; j < d->np; ++j)
      for (int i = 0; i < d->np; ++i)
        {
          const bool slave = (d->ax->periodic && i == d->P) || (d->ay->periodic && j == d->P);
          diag[int64_t(j) * d->np + i] = slave ? 1. : d->ax->diagonal[i] * d->ay->diagonal[j];
        }
    return LBFEM_OK;
  }

  int
  lbfem_disc_set_walls(lbfem_disc *d, const int32_t n, const int32_t *nodes, const double *velocity)
  {
    RuntimeLock lock(runtime_mutex());
    return guarded([&]() -> int {
      if (!d || n < 0 || (n > 0 && (!nodes || !velocity)))
        return fail(LBFEM_EINVAL, "null argument");
      std::vector<int32_t> side[4];
      bool                 used[4] = {false, false, false, false};
      for (auto &s : side)
        s.assign(d->np, -1);
      for (int32_t k = 0; k < n; ++k)
        {
          const int64_t g = nodes[k];
          if (g < 0 || g >= d->n)
            return fail(LBFEM_EINVAL, "wall node index out of range");
          const int i = int(g % d->np), j = int(g / d->np);
          int       sd = -1;
          if (!d->ay->periodic && (j == 0 || j == d->P))
            sd = j == 0 ? 2 : 3;
          else if (!d->ax->periodic && (i == 0 || i == d->P))
            sd = i == 0 ? 0 : 1;
          if (sd < 0)
            return fail(LBFEM_EINVAL, "wall node " + std::to_string(g) +
                                        " is not on a non-periodic boundary of the box");
          side[sd][sd >= 2 ? i : j] = k;
          used[sd]                  = true;
        }
      d->wall_nodes.assign(nodes, nodes + n);
      d->wall_velocity.assign(velocity, velocity + 2 * size_t(n));
      d->wall_vel.upload(d->wall_velocity);
      for (int s = 0; s < 4; ++s)
        {
          if (used[s])
            d->wall_side[s].upload(side[s]);
          else
            d->wall_side[s].alloc(0);
          d->G.walls.side[s] = used[s] ? d->wall_side[s].p : nullptr;
        }
      d->G.walls.velocity = d->wall_vel.p;
      return LBFEM_OK;
    });
  }

  int
  lbfem_disc_time_step(const lbfem_disc *d, const double cfl, double *dt)
  {
    if (!d || !dt)
      return fail(LBFEM_EINVAL, "null argument");
    *dt = cfl * d->h_min / (d->p * d->p);
    return LBFEM_OK;
  }

  // ------------------------------------------------------------------ operators
  int
  lbfem_advection_apply(const lbfem_disc *d, const int kind, double *rhs, const double *f, const double *feq,
                        const double *e, const double dt, const double lambda, const int remove_mass_flux,
                        void *stream)
  {
    return guarded([&]() -> int {
      if (!d || !rhs || !f || !e)
        return fail(LBFEM_EINVAL, "null argument");
      if (kind == LBFEM_LEELIN && !feq)
        return fail(LBFEM_EINVAL, "LeeLinAdvection needs the equilibria"); // advection.h:267
      if (kind == LBFEM_LEELIN && !(lambda > 0.))
        return fail(LBFEM_EINVAL, "lambda must be positive");
      lbfem_scheme s;
      s.d      = d;
      s.dt     = dt;
      s.lambda = kind == LBFEM_LEELIN ? lambda : 1.;
      std::array<double, 2 * Q> ee{};
      std::copy(e, e + 2 * Q, ee.begin());
      TileParams S       = tile_params(s, ee, kind == LBFEM_LEELIN);
      S.collide          = 0;
      S.remove_mass_flux = remove_mass_flux ? 1 : 0;
      if (d->general) // no wall faces on the periodic box
        HIP_CHECK(quad_advect_launch(kind == LBFEM_LEELIN ? kAdvectLeeLin : kAdvectBardow, d->G, S, f,
                                     kind == LBFEM_LEELIN ? feq : nullptr, nullptr, rhs, as_stream(stream)));
      else
        HIP_CHECK(tile_launch(kind == LBFEM_LEELIN ? kAdvectLeeLin : kAdvectBardow, tile_default_variant(d->G, false), d->G,
                              S, f, kind == LBFEM_LEELIN ? feq : nullptr, (double *)nullptr, rhs,
                              as_stream(stream)));
      return LBFEM_OK;
    });
  }

  int
  lbfem_collide_bgk(const lbfem_disc *d, double *g, const double omega, void *stream)
  {
    return guarded([&]() -> int {
      if (!d || !g)
        return fail(LBFEM_EINVAL, "null argument");
      HIP_CHECK(collide_launch(d->G, g, omega, as_stream(stream)));
      return LBFEM_OK;
    });
  }

  int
  lbfem_compute_equilibrium(const lbfem_disc *d, const double *f, double *feq, void *stream)
  {
    return guarded([&]() -> int {
      if (!d || !f || !feq)
        return fail(LBFEM_EINVAL, "null argument");
      HIP_CHECK(equilibrium_launch(d->G, f, feq, as_stream(stream)));
      return LBFEM_OK;
    });
  }

  int
  lbfem_predictor_corrector(const lbfem_disc *d, double *f, const double *feq, const double *incr, const double theta,
                            void *stream)
  {
    return guarded([&]() -> int {
      if (!d || !f || !feq || !incr)
        return fail(LBFEM_EINVAL, "null argument");
      HIP_CHECK(predictor_corrector_launch(d->G, f, feq, incr, theta, as_stream(stream)));
      return LBFEM_OK;
    });
  }

  int
  lbfem_streaming_vmult(const lbfem_disc *d, double *y, const double *x, const double c, const double *e,
                        void *stream)
  {
    return guarded([&]() -> int {
      if (!d || !y || !x)
        return fail(LBFEM_EINVAL, "null argument");
      CgWork W{};
      W.nv          = 1;
      W.block[0]    = 0;
      W.vec[0].coef = c;
      W.vec[0].e[0] = e ? e[0] : 0.;
      W.vec[0].e[1] = e ? e[1] : 0.;
      HIP_CHECK(vmult_launch(d->G, W, x, y, as_stream(stream)));
      return LBFEM_OK;
    });
  }

  int
  lbfem_mass_solve(const lbfem_disc *dc, const lbfem_mass_settings *m, double *x, const double *r, const double *e,
                   const double tg3_coefficient, int *iterations, void *stream)
  {
    return guarded([&]() -> int {
      if (!dc || !m || !x || !r)
        return fail(LBFEM_EINVAL, "null argument");
      auto       *d = const_cast<lbfem_disc *>(dc); // scratch only
      hipStream_t st = as_stream(stream);
      if (iterations)
        *iterations = 0;
      if (m->type == LBFEM_MASS_LUMPED)
        {
          HIP_CHECK(lumped_solve_launch(d->G, x, r, 1, st));
          return LBFEM_OK;
        }
      if (m->type != LBFEM_MASS_CG && m->type != LBFEM_MASS_RICHARDSON)
        return fail(LBFEM_EINVAL, "unknown mass type");
      d->ensure_scratch();
      d->ensure_vectors();
      CgWork W      = cg_work(*d, d->cg_state.p, d->cg_state.p + NM, d->partial.p, *m);
      W.nv          = 1;
      W.block[0]    = 0;
      W.vec[0].coef = tg3_coefficient;
      W.vec[0].e[0] = e ? e[0] : 0.;
      W.vec[0].e[1] = e ? e[1] : 0.;
      if (m->type == LBFEM_MASS_RICHARDSON)
        {
          HIP_CHECK(lumped_solve_launch(d->G, x, r, 1, st));
          for (int k = 0; k < m->richardson; ++k)
            {
              HIP_CHECK(vmult_launch(d->G, W, x, d->s_ax.p, st));
              HIP_CHECK(richardson_update_launch(d->G, x, r, d->s_ax.p, 1, st));
            }
          // the passes use the discretization's scratch (A x): done before the next call may reuse it
          HIP_CHECK(hipStreamSynchronize(st));
          return LBFEM_OK;
        }
      int64_t   its = 0;
      const int rc  = run_cg(d->G, W, x, r, d->s_r.p, d->s_p0.p, d->s_p1.p, d->s_q.p, d->host_cg, st, &its, 1);
      if (iterations)
        *iterations = int(its);
      return rc;
    });
  }

  // lbfem_mass_solve on the fixed work space `scratch` instead of the discretization's:
  // no hipMalloc, hipFree or other allocation on any path (errors are statuses recorded in
  // a fixed buffer). The same kernels with the same launch shapes, so the result and the
  // iteration count are bitwise those of lbfem_mass_solve.
  int
  lbfem_mass_solve_noalloc(const lbfem_disc *d, const lbfem_mass_settings *m, double *x, const double *r,
                           const double *e, const double tg3_coefficient, int *iterations, void *stream)
  {
    return guarded([&]() -> int {
      if (!d || !m || !x || !r)
        return fail(LBFEM_EINVAL, "null argument");
      hipStream_t st = as_stream(stream);
      if (iterations)
        *iterations = 0;
      if (m->type == LBFEM_MASS_LUMPED)
        {
          HIP_RETURN(lumped_solve_launch(d->G, x, r, 1, st));
          return LBFEM_OK;
        }
      if (m->type != LBFEM_MASS_CG && m->type != LBFEM_MASS_RICHARDSON)
        return fail(LBFEM_EINVAL, "unknown mass type");
      if (d->stride > kNoAllocMaxStride)
        return fail(LBFEM_EINVAL, "lbfem_mass_solve_noalloc: the mesh has more nodes than its fixed work space "
                                  "holds (1179648); use lbfem_mass_solve");
      const int rows = cg_direction_rows(d->G);
      if (rows + 1024 > kNoAllocRows)
        return fail(LBFEM_EINVAL, "lbfem_mass_solve_noalloc: the mesh has more operator tiles than its fixed "
                                  "work space holds partial sums for (7168); use lbfem_mass_solve");
      int dev = 0;
      HIP_RETURN(hipGetDevice(&dev));
      if (dev != d->device)
        return fail(LBFEM_EINVAL, "lbfem_mass_solve_noalloc: call it with the discretization's device current");
      double *base = nullptr;
      if (const int rc = noalloc_scratch(dev, &base); rc != LBFEM_OK)
        return rc;
      CgState     *state   = reinterpret_cast<CgState *>(base);
      double      *partial = base + kNoAllocStateDoubles;
      double      *vec     = partial + size_t(kNoAllocRows) * kCgPartials;
      const size_t sd      = size_t(d->stride);
      CgWork W = cg_work(*d, state, reinterpret_cast<CgState *>(base + 8), partial, *m);
      W.max_blocks         = kNoAllocRows;
      W.nv                 = 1;
      W.block[0]           = 0;
      W.vec[0].coef        = tg3_coefficient;
      W.vec[0].e[0]        = e ? e[0] : 0.;
      W.vec[0].e[1]        = e ? e[1] : 0.;
      if (m->type == LBFEM_MASS_RICHARDSON)
        {
          HIP_RETURN(lumped_solve_launch(d->G, x, r, 1, st));
          for (int k = 0; k < m->richardson; ++k)
            {
              HIP_RETURN(vmult_launch(d->G, W, x, vec, st));
              HIP_RETURN(richardson_update_launch(d->G, x, r, vec, 1, st));
            }
          HIP_RETURN(hipStreamSynchronize(st)); // the work space is free for the next call
          return LBFEM_OK;
        }
      CgState   host_state[1]; // the stopping state, read into the caller's stack
      int64_t   its = 0;
      const int rc  = run_cg(d->G, W, x, r, vec, vec + sd, vec + 2 * sd, vec + 3 * sd, host_state, st, &its, 1);
      if (iterations)
        *iterations = int(its);
      return rc;
    });
  }

  // ------------------------------------------------------------------ schemes
  int
  lbfem_scheme_create(const lbfem_disc *d, const lbfem_scheme_settings *set, lbfem_scheme **out)
  {
    RuntimeLock lock(runtime_mutex());
    return guarded([&]() -> int {
      if (!d || !set || !out)
        return fail(LBFEM_EINVAL, "null argument");
      const auto &m = set->mass;
      if (set->scheme != LBFEM_LEELIN && set->scheme != LBFEM_BARDOW)
        return fail(LBFEM_EINVAL, "unknown scheme");
      if (set->streaming < LBFEM_TG2 || set->streaming > LBFEM_TG3_SPLIT)
        return fail(LBFEM_EINVAL, "unknown streaming");
      if (m.type < LBFEM_MASS_CG || m.type > LBFEM_MASS_RICHARDSON)
        return fail(LBFEM_EINVAL, "unknown mass type");
      if (set->streaming != LBFEM_TG2 && m.type != LBFEM_MASS_CG) // streaming.h:51-52
        return fail(LBFEM_EINVAL, "TG3 streaming needs the consistent mass (Mass::cg)");
      if (set->scheme == LBFEM_LEELIN && set->streaming != LBFEM_TG2) // schemes.h:88
        return fail(LBFEM_EINVAL, "LeeLin streams with TG2 only");
      if (set->precision != LBFEM_FP64 && set->precision != LBFEM_MIXED)
        return fail(LBFEM_EINVAL, "unknown precision");
      if (set->precision == LBFEM_MIXED && m.type == LBFEM_MASS_RICHARDSON)
        return fail(LBFEM_EINVAL, "the mixed-precision form is available with the lumped or the consistent mass");
      if (set->precision == LBFEM_MIXED && d->general)
        return fail(LBFEM_EINVAL, "general quadrilaterals (distort) are stepped in FP64 only");
      auto s = std::make_unique<lbfem_scheme>();
      s->d   = d;
      s->set = *set;
      HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
      const size_t st = size_t(d->stride);
      s->mixed        = set->precision == LBFEM_MIXED;
      if (s->mixed)
        for (auto &b : s->g32)
          {
            b.alloc(Q * st);
            HIP_CHECK(hipMemset(b.p, 0, Q * st * sizeof(float)));
          }
      else
        for (auto &b : s->f)
          {
            b.alloc(Q * st);
            HIP_CHECK(hipMemset(b.p, 0, Q * st * sizeof(double)));
          }
      std::array<double, 2 * Q> all{};
      for (int a = 0; a < Q; ++a)
        {
          all[2 * a]     = kE[a][0];
          all[2 * a + 1] = kE[a][1];
        }
      if (set->streaming == LBFEM_TG3_SPLIT) // streaming.h:53-56, 108-115
        {
          std::array<double, 2 * Q> ex{}, ey{};
          for (int a = 0; a < Q; ++a)
            {
              ex[2 * a]     = kE[a][0];
              ey[2 * a + 1] = kE[a][1];
            }
          s->sweeps = {ex, ey};
        }
      else
        s->sweeps = {all};
      s->tg3         = set->streaming == LBFEM_TG2 ? 0. : 1. / 6.; // streaming.h:47
      s->extrapolate = m.type == LBFEM_MASS_CG && set->streaming != LBFEM_TG3_SPLIT;
      const bool lumped_fused = m.type == LBFEM_MASS_LUMPED && set->streaming == LBFEM_TG2;
      if (!lumped_fused && !(s->mixed && m.type == LBFEM_MASS_CG)) // FP64 increments
        {
          s->rhs.alloc(NM * st);
          s->x[0].alloc(NM * st);
          s->x[1].alloc(NM * st);
          HIP_CHECK(hipMemset(s->x[0].p, 0, NM * st * sizeof(double)));
          HIP_CHECK(hipMemset(s->x[1].p, 0, NM * st * sizeof(double)));
        }
      if (m.type == LBFEM_MASS_RICHARDSON)
        s->ax_tmp.alloc(NM * st);
      if (m.type != LBFEM_MASS_LUMPED)
        {
          s->state.alloc(2 * NM);
          s->partial.alloc(size_t(3 * NM) * kMaxCgBlocks);
        }
      if (m.type == LBFEM_MASS_CG)
        {
          HIP_CHECK(hipHostMalloc(&s->host_state, sizeof(CgState) * NM));
          if (s->mixed)
            {
              for (DevBuf<float> *b : {&s->rhs32, &s->x32[0], &s->x32[1], &s->r32, &s->p32[0], &s->p32[1], &s->q32})
                b->alloc(NM * st);
              HIP_CHECK(hipMemset(s->x32[0].p, 0, NM * st * sizeof(float)));
              HIP_CHECK(hipMemset(s->x32[1].p, 0, NM * st * sizeof(float)));
              s->state32.alloc(2 * NM);
              HIP_CHECK(hipHostMalloc(&s->host_state32, sizeof(CgState) * NM));
            }
          else
            {
              s->r.alloc(NM * st);
              s->pb[0].alloc(NM * st);
              s->pb[1].alloc(NM * st);
              s->q.alloc(NM * st);
            }
        }
      *out = s.release();
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_destroy(lbfem_scheme *s)
  {
    RuntimeLock lock(runtime_mutex());
    delete s;
    return LBFEM_OK;
  }

  int
  lbfem_scheme_set_time_step(lbfem_scheme *s, const double dt, const double lambda)
  {
    if (!s)
      return fail(LBFEM_EINVAL, "null argument");
    if (!(dt > 0.) || !(lambda > 0.))
      return fail(LBFEM_EINVAL, "dt and lambda must be positive");
    s->dt     = dt;
    s->lambda = lambda;
    return LBFEM_OK;
  }

  int
  lbfem_scheme_set_populations(lbfem_scheme *s, const double *f)
  {
    return guarded([&]() -> int {
      if (!s || !f)
        return fail(LBFEM_EINVAL, "null argument");
      const int64_t n = s->d->n;
      if (s->mixed) // deviations from the rest state, rounded to FP32
        {
          // on the scheme's stream, after the steps still queued there (a step writes g32[cur])
          std::vector<float> dv(static_cast<size_t>(Q * n));
          for (int a = 0; a < Q; ++a)
            for (int64_t i = 0; i < n; ++i)
              dv[a * n + i] = float(f[a * n + i] - w_of(a));
          for (int a = 0; a < Q; ++a)
            HIP_CHECK(hipMemcpyAsync(s->g32[s->cur].p + a * s->d->stride, dv.data() + a * n, n * sizeof(float),
                                     hipMemcpyHostToDevice, s->stream));
          HIP_CHECK(hipStreamSynchronize(s->stream));
          return LBFEM_OK;
        }
      for (int a = 0; a < Q; ++a)
        HIP_CHECK(hipMemcpyAsync(s->f[s->cur].p + a * s->d->stride, f + a * n, n * sizeof(double),
                                 hipMemcpyHostToDevice, s->stream));
      HIP_CHECK(hipStreamSynchronize(s->stream));
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_set_initial_populations(lbfem_scheme *s, const double *f0)
  {
    return guarded([&]() -> int {
      if (!s || !f0)
        return fail(LBFEM_EINVAL, "null argument");
      if (s->set.scheme == LBFEM_LEELIN)
        return lbfem_scheme_set_populations(s, f0);
      if (!(s->dt > 0.) || !(s->lambda > 0.))
        return fail(LBFEM_EINVAL, "set the time step first (schemes.h:128-141 uses dt and lambda)");
      // g = f + dt/(2 lambda) (f - feq(moments(f))) (schemes.h:128-141), on the host
      const int64_t       n     = s->d->n;
      const double        scale = 0.5 * s->dt / s->lambda;
      std::vector<double> g(Q * size_t(n));
      for (int64_t i = 0; i < n; ++i)
        {
          double fi[Q];
          for (int a = 0; a < Q; ++a)
            fi[a] = f0[a * n + i];
          double rho = 0., mx = 0., my = 0.;
          for (int a = 0; a < Q; ++a)
            {
              rho += fi[a];
              mx += kE[a][0] * fi[a];
              my += kE[a][1] * fi[a];
            }
          const double ux = mx / rho, uy = my / rho, uu = ux * ux + uy * uy;
          for (int a = 0; a < Q; ++a)
            {
              const double eu = kE[a][0] * ux + kE[a][1] * uy;
              const double fe = w_of(a) * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * uu);
              g[a * n + i]    = fi[a] + scale * (fi[a] - fe);
            }
        }
      return lbfem_scheme_set_populations(s, g.data());
    });
  }

  int
  lbfem_scheme_get_populations(const lbfem_scheme *s, double *f)
  {
    return guarded([&]() -> int {
      if (!s || !f)
        return fail(LBFEM_EINVAL, "null argument");
      const int64_t n = s->d->n;
      if (s->mixed)
        {
          std::vector<float> dv(static_cast<size_t>(Q * n));
          for (int a = 0; a < Q; ++a)
            HIP_CHECK(hipMemcpyAsync(dv.data() + a * n, s->g32[s->cur].p + a * s->d->stride, n * sizeof(float),
                                     hipMemcpyDeviceToHost, s->stream));
          HIP_CHECK(hipStreamSynchronize(s->stream));
          for (int a = 0; a < Q; ++a)
            for (int64_t i = 0; i < n; ++i)
              f[a * n + i] = w_of(a) + double(dv[a * n + i]);
          return LBFEM_OK;
        }
      for (int a = 0; a < Q; ++a)
        HIP_CHECK(hipMemcpyAsync(f + a * n, s->f[s->cur].p + a * s->d->stride, n * sizeof(double),
                                 hipMemcpyDeviceToHost, s->stream));
      HIP_CHECK(hipStreamSynchronize(s->stream));
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_device_populations(lbfem_scheme *s, double **f)
  {
    if (!s || !f)
      return fail(LBFEM_EINVAL, "null argument");
    if (s->mixed)
      return fail(LBFEM_EINVAL, "the mixed form keeps FP32 deviations; use lbfem_scheme_get_populations");
    *f = s->f[s->cur].p;
    return LBFEM_OK;
  }

  int
  lbfem_scheme_step(lbfem_scheme *s, const int n_steps)
  {
    return guarded([&]() -> int {
      if (!s || n_steps < 0)
        return fail(LBFEM_EINVAL, "invalid argument");
      if (!(s->dt > 0.) || !(s->lambda > 0.))
        return fail(LBFEM_EINVAL, "set_time_step first");
      for (int k = 0; k < n_steps; ++k)
        {
          const int st = step_once(*s);
          if (st != LBFEM_OK)
            return st;
          if (s->timing && s->marks.size() > 256)
            s->resolve_marks();
        }
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_synchronize(lbfem_scheme *s)
  {
    return guarded([&]() -> int {
      if (!s)
        return fail(LBFEM_EINVAL, "null argument");
      HIP_CHECK(hipStreamSynchronize(s->stream));
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_cg_iterations(const lbfem_scheme *s, int64_t *its)
  {
    if (!s || !its)
      return fail(LBFEM_EINVAL, "null argument");
    *its = s->cg_iterations;
    return LBFEM_OK;
  }

  int
  lbfem_scheme_last_cg_iterations(const lbfem_scheme *s, int32_t *its)
  {
    if (!s || !its)
      return fail(LBFEM_EINVAL, "null argument");
    for (int sw = 0; sw < 2; ++sw)
      for (int a = 0; a < NM; ++a)
        its[sw * NM + a] = s->last_step_its[sw][a];
    return LBFEM_OK;
  }

  int
  lbfem_scheme_diagnostics(lbfem_scheme *s, const double *u_exact, const double exact_scale, double *sums)
  {
    return guarded([&]() -> int {
      if (!s || !sums)
        return fail(LBFEM_EINVAL, "null argument");
      const size_t n2 = 2 * size_t(s->d->n);
      if (!s->diag_host)
        {
          RuntimeLock lock(runtime_mutex());
          s->diag_exact.alloc(n2);
          s->diag_prev.alloc(n2);
          s->diag_partial.alloc(size_t(1024) * kCgPartials);
          s->diag_sums.alloc(8);
          HIP_CHECK(hipHostMalloc(&s->diag_host, 8 * sizeof(double)));
          HIP_CHECK(hipMemsetAsync(s->diag_prev.p, 0, n2 * sizeof(double), s->stream));
        }
      if (u_exact)
        {
          HIP_CHECK(hipMemcpyAsync(s->diag_exact.p, u_exact, n2 * sizeof(double), hipMemcpyHostToDevice, s->stream));
          s->diag_exact_set = true;
        }
      const double *ex = s->diag_exact_set ? s->diag_exact.p : nullptr;
      if (s->mixed)
        HIP_CHECK(diagnostics_scaled_launch<float>(s->d->G, s->g32[s->cur].p, ex, exact_scale, s->diag_prev.p,
                                                   s->diag_sums.p, s->diag_partial.p, 1024, s->stream));
      else
        HIP_CHECK(diagnostics_scaled_launch<double>(s->d->G, s->f[s->cur].p, ex, exact_scale, s->diag_prev.p,
                                                    s->diag_sums.p, s->diag_partial.p, 1024, s->stream));
      HIP_CHECK(hipMemcpyAsync(s->diag_host, s->diag_sums.p, 6 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
      HIP_CHECK(hipStreamSynchronize(s->stream)); // also releases the caller's u_exact
      std::copy(s->diag_host, s->diag_host + 6, sums);
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_velocity_change(lbfem_scheme *s, double *max_change, double *mean_change, int64_t *nonfinite)
  {
    return guarded([&]() -> int {
      if (!s)
        return fail(LBFEM_EINVAL, "null argument");
      std::vector<MonitorEntry> E{monitor_entry(*s)};
      double                    r[kMonitorPartials];
      s->monitor.run(E, s->stream, r);
      s->u_prev_set = true;
      if (max_change)
        *max_change = r[0];
      if (mean_change)
        *mean_change = r[1] / r[2];
      if (nonfinite)
        *nonfinite = int64_t(r[3]);
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_enable_stage_timers(lbfem_scheme *s, const int enable)
  {
    return guarded([&]() -> int {
      if (!s)
        return fail(LBFEM_EINVAL, "null argument");
      s->resolve_marks();
      s->timing = enable != 0;
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_stage_times(const lbfem_scheme *sc, double *seconds4)
  {
    return guarded([&]() -> int {
      if (!sc || !seconds4)
        return fail(LBFEM_EINVAL, "null argument");
      auto *s = const_cast<lbfem_scheme *>(sc);
      s->resolve_marks();
      for (int k = 0; k < 4; ++k)
        seconds4[k] = s->stage[k];
      return LBFEM_OK;
    });
  }

  int
  lbfem_diagnostics(const lbfem_disc *dc, const double *f, const double *u_exact, double *u_prev, double *sums,
                    void *stream)
  {
    return guarded([&]() -> int {
      if (!dc || !f || !sums)
        return fail(LBFEM_EINVAL, "null argument");
      auto *d = const_cast<lbfem_disc *>(dc);
      d->ensure_scratch();
      hipStream_t st = as_stream(stream);
      HIP_CHECK(
        diagnostics_launch(d->G, f, u_exact, u_prev, d->s_sums.p, d->partial.p, kMaxCgBlocks, st));
      HIP_CHECK(hipMemcpyAsync(sums, d->s_sums.p, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      return LBFEM_OK;
    });
  }
  int
  lbfem_scheme_get_stream(const lbfem_scheme *s, void **stream)
  {
    return guarded([&]() -> int {
      if (!s || !stream)
        return fail(LBFEM_EINVAL, "null argument");
      *stream = static_cast<void *>(s->stream);
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_set_tile_variant(lbfem_scheme *s, const int variant)
  {
    return guarded([&]() -> int {
      if (!s)
        return fail(LBFEM_EINVAL, "null argument");
      if (variant < -1 || variant >= kTileVariants)
        return fail(LBFEM_EINVAL, "tile variant out of range");
      s->tile_variant = variant;
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_get_tile_variant(const lbfem_scheme *s, int *variant)
  {
    return guarded([&]() -> int {
      if (!s || !variant)
        return fail(LBFEM_EINVAL, "null argument");
      *variant = s->variant();
      return LBFEM_OK;
    });
  }

  int
  lbfem_scheme_autotune(lbfem_scheme *s, const int force, int *variant, double *seconds)
  {
    return guarded([&]() -> int {
      if (!s)
        return fail(LBFEM_EINVAL, "null argument");
      if (!(s->dt > 0.))
        return fail(LBFEM_EINVAL, "set the time step before tuning");
      if (s->d->general) // one cell kernel shape: nothing to tune
        {
          if (seconds)
            for (int v = 0; v < kTileVariants; ++v)
              seconds[v] = -1.;
          if (variant)
            *variant = s->variant();
          return LBFEM_OK;
        }
      const TuneKey key = tune_key(*s);
      auto         &cache = tune_cache();
      auto          hit   = cache.find(key);
      if (seconds)
        for (int v = 0; v < kTileVariants; ++v)
          seconds[v] = -1.;
      if (hit != cache.end() && !force)
        s->tile_variant = hit->second;
      else
        {
          RuntimeLock lock(runtime_mutex());
          hipEvent_t  e0, e1;
          HIP_CHECK(hipEventCreate(&e0));
          HIP_CHECK(hipEventCreate(&e1));
          double best = 1e300;
          int    pick = s->variant();
          for (int v = 0; v < kTileVariants; ++v)
            {
              // two launches to load the code, three to size the sample, then about 2 ms of launches
              for (int k = 0; k < 2; ++k)
                if (const int st = tile_trial(*s, v); st != LBFEM_OK)
                  return st;
              float ms = 0.f;
              HIP_CHECK(hipEventRecord(e0, s->stream));
              for (int k = 0; k < 3; ++k)
                if (const int st = tile_trial(*s, v); st != LBFEM_OK)
                  return st;
              HIP_CHECK(hipEventRecord(e1, s->stream));
              HIP_CHECK(hipEventSynchronize(e1));
              HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
              const int n = std::max(5, std::min(200, int(2.0 / std::max(1e-4, double(ms) / 3))));
              HIP_CHECK(hipEventRecord(e0, s->stream));
              for (int k = 0; k < n; ++k)
                if (const int st = tile_trial(*s, v); st != LBFEM_OK)
                  return st;
              HIP_CHECK(hipEventRecord(e1, s->stream));
              HIP_CHECK(hipEventSynchronize(e1));
              HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
              const double t = 1e-3 * ms / n;
              if (seconds)
                seconds[v] = t;
              if (t < best)
                {
                  best = t;
                  pick = v;
                }
            }
          (void)hipEventDestroy(e0);
          (void)hipEventDestroy(e1);
          cache[key]      = pick;
          s->tile_variant = pick;
        }
      if (variant)
        *variant = s->tile_variant;
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_create(lbfem_scheme *const *schemes, const int n, lbfem_batch **out)
  {
    RuntimeLock lock(runtime_mutex());
    return guarded([&]() -> int {
      if (!out || n < 0 || (n > 0 && !schemes))
        return fail(LBFEM_EINVAL, "lbfem_batch_create: invalid arguments");
      for (int k = 0; k < n; ++k)
        {
          if (const int st = check_batchable(schemes[k]); st != LBFEM_OK)
            return st;
          for (int j = 0; j < k; ++j)
            if (schemes[j] == schemes[k])
              return fail(LBFEM_EINVAL, "a scheme appears twice in the batch");
        }
      auto b = std::make_unique<lbfem_batch>();
      HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
      for (int k = 0; k < n; ++k)
        join(*b, schemes[k]);
      // default shape: the row-march kernel on Q1 groups, else the tile shape the size-based rule
      // picks for the group's total node count
      for (auto &g : b->groups)
        {
          int64_t nodes = 0;
          for (const lbfem_scheme *m : g.members)
            nodes += int64_t(m->d->G.np) * m->d->G.np;
          g.variant = batch_default_variant(nodes, g.mixed, g.p);
        }
      rebuild_tables(*b);
      for (auto &g : b->cg_groups)
        write_cg_group(*b, *g);
      *out = b.release();
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_add(lbfem_batch *b, lbfem_scheme *s)
  {
    return guarded([&]() -> int {
      if (!b)
        return fail(LBFEM_EINVAL, "null batch");
      if (const int st = check_batchable(s); st != LBFEM_OK)
        return st;
      if (b->index_of(s) >= 0)
        return fail(LBFEM_EINVAL, "the scheme is already a member of the batch");
      HIP_CHECK(hipStreamSynchronize(b->stream));
      if (b->cg_stream)
        HIP_CHECK(hipStreamSynchronize(b->cg_stream));
      join(*b, s);
      if (!is_lumped_fused(s))
        {
          write_cg_group(*b, *b->cg_group_of(s));
          return LBFEM_OK;
        }
      const int mode = s->set.scheme == LBFEM_LEELIN ? kStepLeeLin : kStepBardow;
      for (auto &g : b->groups)
        if (g.p == s->d->G.p && g.mode == mode && g.mixed == s->mixed)
          write_group(*b, g);
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_remove(lbfem_batch *b, lbfem_scheme *s)
  {
    return guarded([&]() -> int {
      if (!b || !s)
        return fail(LBFEM_EINVAL, "null argument");
      const int k = b->index_of(s);
      if (k < 0)
        return fail(LBFEM_EINVAL, "the scheme is not a member of the batch");
      HIP_CHECK(hipStreamSynchronize(b->stream));
      if (b->cg_stream)
        HIP_CHECK(hipStreamSynchronize(b->cg_stream));
      sync_members(*b);
      b->members.erase(b->members.begin() + k);
      b->par0.erase(b->par0.begin() + k);
      if (lbfem_batch::CgGroup *cg = b->cg_group_of(s))
        {
          const size_t j = size_t(std::find(cg->members.begin(), cg->members.end(), s) - cg->members.begin());
          cg->members.erase(cg->members.begin() + j);
          cg->par0.erase(cg->par0.begin() + j);
          cg->failed.erase(cg->failed.begin() + j);
          if (cg->members.empty())
            b->cg_groups.erase(std::find_if(b->cg_groups.begin(), b->cg_groups.end(),
                                            [&](const auto &u) { return u.get() == cg; }));
          else
            write_cg_group(*b, *cg);
          return LBFEM_OK;
        }
      for (size_t gi = 0; gi < b->groups.size(); ++gi)
        {
          auto &g  = b->groups[gi];
          auto  it = std::find(g.members.begin(), g.members.end(), s);
          if (it == g.members.end())
            continue;
          g.members.erase(it);
          if (g.members.empty())
            {
              b->drop_graphs(); // the graphs launch that group
              b->groups.erase(b->groups.begin() + gi);
            }
          else
            write_group(*b, g);
          break;
        }
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_destroy(lbfem_batch *b)
  {
    return guarded([&]() -> int {
      if (b)
        {
          (void)hipStreamSynchronize(b->stream);
          if (b->cg_stream)
            (void)hipStreamSynchronize(b->cg_stream);
          delete b;
        }
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_set_graph_steps(lbfem_batch *b, const int steps)
  {
    return guarded([&]() -> int {
      if (!b || steps < 0 || (steps & 1))
        return fail(LBFEM_EINVAL, "steps per graph: an even number >= 0");
      HIP_CHECK(hipStreamSynchronize(b->stream));
      b->drop_graphs();
      b->graph_steps = steps;
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_step(lbfem_batch *b, const int n_steps)
  {
    return guarded([&]() -> int {
      if (!b || n_steps < 0)
        return fail(LBFEM_EINVAL, "invalid argument");
      if (b->members.empty())
        return LBFEM_OK;
      int left = b->groups.empty() ? 0 : n_steps;
      if (b->graph_steps > 0 && left >= b->graph_steps)
        {
          if (!b->exec[0] || !b->exec[1])
            if (const int st = capture_graphs(*b); st != LBFEM_OK)
              return st;
          for (; left >= b->graph_steps; left -= b->graph_steps)
            HIP_CHECK(hipGraphLaunch(b->exec[b->parity], b->stream));
        }
      for (; left > 0; --left)
        {
          if (const int st = batch_step_eager(*b, b->parity); st != LBFEM_OK)
            return st;
          b->parity ^= 1;
        }
      // the consistent-mass groups on their own stream, beside the lumped groups' queued work
      int status = LBFEM_OK;
      for (int k = 0; k < n_steps; ++k)
        for (auto &g : b->cg_groups)
          if (const int st = cg_group_step(*b, *g); st != LBFEM_OK)
            {
              if (st != LBFEM_ENOCONV)
                {
                  sync_members(*b);
                  return st;
                }
              status = st;
            }
      sync_members(*b);
      return status;
    });
  }

  int
  lbfem_batch_synchronize(lbfem_batch *b)
  {
    return guarded([&]() -> int {
      if (!b)
        return fail(LBFEM_EINVAL, "null argument");
      HIP_CHECK(hipStreamSynchronize(b->stream));
      if (b->cg_stream)
        HIP_CHECK(hipStreamSynchronize(b->cg_stream));
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_member_status(const lbfem_batch *b, int32_t *status)
  {
    return guarded([&]() -> int {
      if (!b || !status)
        return fail(LBFEM_EINVAL, "null argument");
      for (size_t k = 0; k < b->members.size(); ++k)
        {
          status[k] = LBFEM_OK;
          for (const auto &g : b->cg_groups)
            for (size_t j = 0; j < g->members.size(); ++j)
              if (g->members[j] == b->members[k] && g->failed[j])
                status[k] = LBFEM_ENOCONV;
        }
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_get_stream(const lbfem_batch *b, void **stream)
  {
    return guarded([&]() -> int {
      if (!b || !stream)
        return fail(LBFEM_EINVAL, "null argument");
      *stream = static_cast<void *>(b->stream);
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_info(const lbfem_batch *b, int *n_groups, int64_t *n_nodes, int64_t *n_tiles)
  {
    return guarded([&]() -> int {
      if (!b)
        return fail(LBFEM_EINVAL, "null argument");
      int64_t nodes = 0, tiles = 0;
      for (const auto &g : b->groups)
        {
          nodes += g.n_nodes;
          tiles += g.n_tiles;
        }
      for (const auto &g : b->cg_groups)
        {
          nodes += g->n_nodes;
          tiles += g->adv_tiles[0];
        }
      if (n_groups)
        *n_groups = int(b->groups.size() + b->cg_groups.size());
      if (n_nodes)
        *n_nodes = nodes;
      if (n_tiles)
        *n_tiles = tiles;
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_set_shape(lbfem_batch *b, const int variant, const int grid)
  {
    return guarded([&]() -> int {
      if (!b || variant < -1 || variant >= kTileVariants || grid < 0)
        return fail(LBFEM_EINVAL, "invalid tile variant or grid");
      HIP_CHECK(hipStreamSynchronize(b->stream));
      b->shape_variant = variant;
      b->shape_grid    = grid;
      for (auto &g : b->groups)
        {
          if (variant >= 0)
            g.variant = variant;
          g.grid = grid;
        }
      b->drop_graphs();
      rebuild_tables(*b);
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_autotune(lbfem_batch *b, double *seconds)
  {
    return guarded([&]() -> int {
      if (!b)
        return fail(LBFEM_EINVAL, "null argument");
      HIP_CHECK(hipStreamSynchronize(b->stream));
      sync_members(*b);
      RuntimeLock lock(runtime_mutex());
      hipEvent_t  e0, e1;
      HIP_CHECK(hipEventCreate(&e0));
      HIP_CHECK(hipEventCreate(&e1));
      const int grids[3] = {0, 512, 2048};
      double    total    = 0.;
      for (auto &g : b->groups)
        {
          double best = 1e300;
          int    bv = g.variant, bg = g.grid;
          for (int v = 0; v < kTileVariants; ++v)
            for (const int grid : grids)
              {
                // tile counts depend on the variant; launches at a fixed parity read one buffer and write the
                // other: no state changes
                g.variant = v;
                g.grid    = grid;
                write_group(*b, g);
                for (int k = 0; k < 2; ++k)
                  if (const int st = launch_group(*b, g, b->parity); st != LBFEM_OK)
                    return st;
                float ms = 0.f;
                HIP_CHECK(hipEventRecord(e0, b->stream));
                for (int k = 0; k < 3; ++k)
                  if (const int st = launch_group(*b, g, b->parity); st != LBFEM_OK)
                    return st;
                HIP_CHECK(hipEventRecord(e1, b->stream));
                HIP_CHECK(hipEventSynchronize(e1));
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                const int n = std::max(5, std::min(100, int(2.0 / std::max(1e-4, double(ms) / 3))));
                HIP_CHECK(hipEventRecord(e0, b->stream));
                for (int k = 0; k < n; ++k)
                  if (const int st = launch_group(*b, g, b->parity); st != LBFEM_OK)
                    return st;
                HIP_CHECK(hipEventRecord(e1, b->stream));
                HIP_CHECK(hipEventSynchronize(e1));
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                const double t = 1e-3 * ms / n;
                if (t < best)
                  {
                    best = t;
                    bv   = v;
                    bg   = grid;
                  }
              }
          g.variant = bv;
          g.grid    = bg;
          total += best;
        }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      b->drop_graphs();
      rebuild_tables(*b); // the chosen shapes
      if (seconds)
        *seconds = total;
      return LBFEM_OK;
    });
  }

  int
  lbfem_batch_velocity_change(lbfem_batch *b, const int32_t *mask, double *max_change, double *mean_change,
                              int64_t *nonfinite)
  {
    return guarded([&]() -> int {
      if (!b)
        return fail(LBFEM_EINVAL, "null argument");
      if (b->cg_stream)
        HIP_CHECK(hipStreamSynchronize(b->cg_stream));
      HIP_CHECK(hipStreamSynchronize(b->stream));
      sync_members(*b);
      std::vector<MonitorEntry> E;
      std::vector<int>          which;
      for (size_t k = 0; k < b->members.size(); ++k)
        if (!mask || mask[k])
          {
            E.push_back(monitor_entry(*b->members[k]));
            which.push_back(int(k));
          }
      std::vector<double> r(E.size() * kMonitorPartials);
      b->monitor.run(E, b->stream, r.data());
      for (size_t q = 0; q < which.size(); ++q)
        {
          const int     k = which[q];
          const double *x = r.data() + q * kMonitorPartials;
          b->members[k]->u_prev_set = true;
          if (max_change)
            max_change[k] = x[0];
          if (mean_change)
            mean_change[k] = x[1] / x[2];
          if (nonfinite)
            nonfinite[k] = int64_t(x[3]);
        }
      return LBFEM_OK;
    });
  }

  // Stream function of a nodal velocity field (include/lbfem/io.h:31-93): with
  // u = -dpsi/dy, v = dpsi/dx, (grad phi, grad psi) = (dphi/dx, v) - (dphi/dy, u)
  // and psi = 0 on the walls, solved by deal.II's SolverCG sequence with the
  // Jacobi preconditioner to 1e-12 |b| from zero, at most 10000 iterations. On
  // the tensor mesh the Laplacian is Kx (x) My + Mx (x) Ky and the right-hand
  // side (Dx^T (x) My) v - (Mx (x) Dy^T) u, from the same assembled 1-D bands as
  // the streaming operator; the CG is the mass solves' (one vector, the operator
  // tile in its Laplacian mode), its stopping state read every 16 iterations.
  int
  lbfem_disc_stream_function(const lbfem_disc *dc, const double *u, const double *v, double *psi, int *iterations,
                             void *stream)
  {
    return guarded([&]() -> int {
      if (!dc || !u || !v || !psi)
        return fail(LBFEM_EINVAL, "null argument");
      if (dc->G.ax.periodic || dc->G.ay.periodic || dc->G.gen.vert != nullptr)
        return fail(LBFEM_EINVAL, "the stream function is computed for walled boxes (n_periodic = 0, no distortion)");
      auto       *d  = const_cast<lbfem_disc *>(dc); // scratch only
      hipStream_t st = as_stream(stream);
      if (iterations)
        *iterations = 0;
      d->ensure_scratch();
      d->ensure_vectors();
      Geo G          = d->G;
      G.dinv         = d->s_dinv.p; // the Jacobi preconditioner of the Laplacian
      HIP_CHECK(psi_setup_launch(G, u, v, d->s_b.p, d->s_dinv.p, st));
      HIP_CHECK(hipMemsetAsync(psi, 0, sizeof(double) * size_t(d->n), st));
      CgWork W{};
      W.nv         = 1;
      W.block[0]   = 0;
      W.max_its    = 10000;
      W.tol_rel    = 1e-12;
      W.state      = d->cg_state.p;
      W.state2     = d->cg_state.p + NM;
      W.partial    = d->partial.p;
      W.max_blocks = kMaxCgBlocks;
      W.laplace    = 1;
      int64_t   its = 0;
      const int rc  = run_cg(G, W, psi, d->s_b.p, d->s_r.p, d->s_p0.p, d->s_p1.p, d->s_q.p, d->host_cg, st, &its, 16,
                             true, 16);
      if (iterations)
        *iterations = int(its);
      return rc;
    });
  }

  // The same from host arrays and a mesh description: a discretization of its own,
  // the fields copied to and from the device.
  int
  lbfem_stream_function(const lbfem_mesh_desc *desc, const double *u, const double *v, double *psi,
                        int *iterations)
  {
    RuntimeLock lock(runtime_mutex());
    return guarded([&]() -> int {
      if (!desc || !u || !v || !psi)
        return fail(LBFEM_EINVAL, "null argument");
      if (desc->n_periodic != 0 || desc->distort != 0.)
        return fail(LBFEM_EINVAL, "the stream function is computed for walled boxes (n_periodic = 0, no distortion)");
      lbfem_disc *raw = nullptr;
      if (const int rc = lbfem_disc_create(desc, &raw); rc != LBFEM_OK)
        return rc;
      std::unique_ptr<lbfem_disc> d(raw);
      const size_t                n = size_t(d->n);
      DevBuf<double>              du, dv, dpsi;
      du.alloc(n);
      dv.alloc(n);
      dpsi.alloc(n);
      HIP_CHECK(hipMemcpy(du.p, u, n * sizeof(double), hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(dv.p, v, n * sizeof(double), hipMemcpyHostToDevice));
      const int rc = lbfem_disc_stream_function(d.get(), du.p, dv.p, dpsi.p, iterations, nullptr);
      if (rc != LBFEM_OK)
        return rc;
      HIP_CHECK(hipMemcpy(psi, dpsi.p, n * sizeof(double), hipMemcpyDeviceToHost));
      return LBFEM_OK;
    });
  }

  int
  lbfem_hbm_copy(void *dst, const void *src, size_t bytes, void *stream)
  {
    return guarded([&]() -> int {
      if (!dst || !src)
        return fail(LBFEM_EINVAL, "null argument");
      if (bytes % 16 != 0 || reinterpret_cast<uintptr_t>(dst) % 16 != 0 || reinterpret_cast<uintptr_t>(src) % 16 != 0)
        return fail(LBFEM_EINVAL, "lbfem_hbm_copy: size and addresses must be multiples of 16 bytes");
      HIP_CHECK(stream_copy_launch(dst, src, bytes, as_stream(stream)));
      return LBFEM_OK;
    });
  }
}
