// cgdbe_hip: the reference's cgdbe driver (apps/cgdbe/cgdbe.cc:207-342, 348-472)
// over the C ABI of liblbfem_hip.so -- the same options, Reynolds continuation,
// diagnostics lines, steady-state test, checkpoint/restart and output files,
// with the time step on one MI355X. Plain C++ (no HIP headers): the library is
// the only GPU code, as it would be for the reference's own driver.
//
//   cgdbe_hip --case cavity --refine 4 --reynolds 100,400 --mass lumped
//
// Exit status: 0, 1 on an error (the message on stderr; also a run whose moments
// stop being finite under --steady-max-change), 2 if a steady-state search ran
// out of time (cgdbe.cc:583-590).
#include "lbfem_hip.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
  constexpr int    Q     = 9;
  constexpr double CS2   = 1. / 3.;
  constexpr int    EX[Q] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
  constexpr int    EY[Q] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
  constexpr double WQ[Q] = {4. / 9, 1. / 9, 1. / 36, 1. / 9, 1. / 36, 1. / 9, 1. / 36, 1. / 9, 1. / 36};
  const double     PI    = std::acos(-1.);

  void
  check(const int rc)
  {
    if (rc == LBFEM_OK)
      return;
    char msg[1024];
    lbfem_last_error(msg, sizeof msg);
    throw std::runtime_error(std::string(rc == LBFEM_ENOCONV ? "NoConvergence: " : "") + msg);
  }

  // ------------------------------------------------------------ parameters
  // apps/cgdbe/cgdbe.cc:62-200: the options, their defaults and finish()
  struct Parameters
  {
    std::string         test_case = "tgv";
    std::string         scheme = "bardow", mass = "cg", streaming = "tg2", precision = "fp64";
    int                 refine = 6, cells = 0, degree = 1, richardson = 2;
    double              mach = 0.1, cfl = 0., t_end = 0., stretch = 1.2, steady_tol = 1e-4, cg_tol = 1e-8;
    double              distort = 0.; // tgv: x += eps L sin(2 pi x/L) sin(2 pi y/L) (1, 1) (cgdbe.cc:77)
    std::array<int, 2>  modes = {{1, 4}};
    std::vector<double> reynolds;
    unsigned            max_steps = 0, n_diagnostic = 20;
    // steady test by the largest velocity change over a window of steps, max|u - u_prev|/U0 < eps
    // (0: the reference's mean-rate test, cgdbe.cc:391-393)
    double              steady_max_change = 0.;
    unsigned            steady_window     = 1000;
    bool                output = true, fused = false, stage_timers = false;
    std::string         checkpoint, restart;

    void
    finish()
    {
      const bool cavity = test_case == "cavity";
      if (reynolds.empty())
        reynolds = {test_case == "tgv" ? 100. : cavity ? 400. : 10.};
      if (cfl <= 0.)
        cfl = !cavity ? 0.25 : (scheme == "leelin" ? 0.5 : 0.4);
      if (t_end <= 0.)
        t_end = cavity ? 100. : 1.;
      if (mass.rfind("richardson", 0) == 0 && mass.size() > 10)
        richardson = std::stoi(mass.substr(10));
    }
  };

  const char *usage =
    "cgdbe_hip: characteristic Galerkin discrete Boltzmann solver, D2Q9 with Q_p elements, on one GPU\n"
    "  --case tgv|couette|cavity   --scheme leelin|bardow [bardow]   --mass cg|lumped|richardson[k] [cg]\n"
    "  --streaming tg2|tg3|tg3-split [tg2]   --refine n [6] | --cells N   --degree p [1]\n"
    "  --mach Ma [0.1]   --modes n1,n2 [1,4]   --reynolds Re[,Re...] [100 | 10 | 400]\n"
    "  --cfl c [0.25; cavity 0.4, with leelin 0.5]   --tend t [1; cavity 100]   --steps n\n"
    "  --stretch gamma [1.2]   --distort eps [0]   --steady-tol eps [1e-4]   --cg-tol eps [1e-8]   --fused\n"
    "  --steady-max-change eps [0: off]   --steady-window n [1000] (steady when max|du|/U0 < eps over n steps)\n"
    "  --checkpoint name   --restart name   --precision fp64|mixed [fp64]   --no-output\n"
    "  --stage-timers (the summary's stage breakdown; GPU events at every stage, inside the timed run)\n";

  std::vector<double>
  split_doubles(const std::string &s)
  {
    std::vector<double> v;
    std::stringstream   ss(s);
    for (std::string t; std::getline(ss, t, ',');)
      v.push_back(std::stod(t));
    return v;
  }

  Parameters
  parse(int argc, char **argv, bool &help)
  {
    Parameters prm;
    help = false;
    for (int i = 1; i < argc; ++i)
      {
        const std::string a = argv[i];
        auto              val = [&]() -> std::string {
          if (i + 1 >= argc)
            throw std::runtime_error(a + ": a value is needed");
          return argv[++i];
        };
        if (a == "--help" || a == "-h")
          help = true;
        else if (a == "--case")
          prm.test_case = val();
        else if (a == "--scheme")
          prm.scheme = val();
        else if (a == "--mass")
          prm.mass = val();
        else if (a == "--streaming")
          prm.streaming = val();
        else if (a == "--refine")
          prm.refine = std::stoi(val());
        else if (a == "--cells")
          prm.cells = std::stoi(val());
        else if (a == "--degree")
          prm.degree = std::stoi(val());
        else if (a == "--mach")
          prm.mach = std::stod(val());
        else if (a == "--modes")
          {
            const auto m = split_doubles(val());
            if (m.size() != 2)
              throw std::runtime_error("--modes: n1,n2");
            prm.modes = {{int(m[0]), int(m[1])}};
          }
        else if (a == "--reynolds")
          prm.reynolds = split_doubles(val());
        else if (a == "--cfl")
          prm.cfl = std::stod(val());
        else if (a == "--tend")
          prm.t_end = std::stod(val());
        else if (a == "--steps")
          prm.max_steps = unsigned(std::stoul(val()));
        else if (a == "--stretch")
          prm.stretch = std::stod(val());
        else if (a == "--distort")
          prm.distort = std::stod(val());
        else if (a == "--steady-tol")
          prm.steady_tol = std::stod(val());
        else if (a == "--steady-max-change")
          prm.steady_max_change = std::stod(val());
        else if (a == "--steady-window")
          prm.steady_window = unsigned(std::stoul(val()));
        else if (a == "--cg-tol")
          prm.cg_tol = std::stod(val());
        else if (a == "--fused")
          prm.fused = true;
        else if (a == "--checkpoint")
          prm.checkpoint = val();
        else if (a == "--restart")
          prm.restart = val();
        else if (a == "--precision")
          prm.precision = val();
        else if (a == "--no-output")
          prm.output = false;
        else if (a == "--stage-timers")
          prm.stage_timers = true;
        else
          throw std::runtime_error("unknown option " + a);
      }
    const auto one_of = [](const std::string &v, std::initializer_list<const char *> set, const char *opt) {
      for (const char *s : set)
        if (v == s)
          return;
      throw std::runtime_error(std::string(opt) + ": unknown value " + v);
    };
    one_of(prm.test_case, {"tgv", "couette", "cavity"}, "--case");
    one_of(prm.scheme, {"leelin", "bardow"}, "--scheme");
    one_of(prm.streaming, {"tg2", "tg3", "tg3-split"}, "--streaming");
    one_of(prm.precision, {"fp64", "mixed"}, "--precision");
    if (!(prm.mass == "cg" || prm.mass == "lumped" ||
          (prm.mass.rfind("richardson", 0) == 0 &&
           prm.mass.find_first_not_of("0123456789", 10) == std::string::npos)))
      throw std::runtime_error("--mass: not one of cg, lumped, richardson[k]: " + prm.mass);
    if (prm.steady_max_change < 0. || prm.steady_window == 0)
      throw std::runtime_error("--steady-max-change eps >= 0, --steady-window n > 0");
    prm.finish();
    return prm;
  }

  // ------------------------------------------------------------ the flows
  // include/lbfem/test_case.h:23-83, test_cases.h:21-245 (lattice units, U0 = Ma c_s, nu = U0 L / Re)
  struct Flow
  {
    std::string kind;
    double      L = 1., rho0 = 1., U0 = 0., nu = 0., stretch = 0.;
    int         n1 = 1, n2 = 4;

    int
    n_periodic() const
    {
      return kind == "tgv" ? 2 : kind == "couette" ? 1 : 0;
    }
    bool
    steady() const
    {
      return kind == "cavity";
    }
    std::string
    name() const
    {
      if (kind == "tgv")
        return "Taylor-Green vortex, periodic, modes (" + std::to_string(n1) + "," + std::to_string(n2) + ")";
      return kind == "couette" ? "start-up Couette flow" : "lid-driven cavity";
    }
    std::string
    tag(const double Re) const
    {
      return kind == "tgv" ? "vortex" : kind == "couette" ? "couette" : "cavity_Re" + std::to_string(std::lround(Re));
    }
    void
    set_reynolds(const double Re)
    {
      nu = U0 * L / Re;
    }
    double
    relaxation_time() const
    {
      return nu / CS2;
    }
    double
    k1() const
    {
      return 2 * PI * n1 / L;
    }
    double
    k2() const
    {
      return 2 * PI * n2 / L;
    }
    double
    reference_time() const
    {
      return kind == "tgv" ? 1. / (nu * (k1() * k1() + k2() * k2())) : kind == "couette" ? L * L / nu : L / U0;
    }
    double
    energy_scale() const
    {
      const double r = k1() / k2();
      return kind == "tgv" ? 0.25 * U0 * U0 * (1 + r * r) : U0 * U0;
    }
    // the wall velocity of a node, if it is on a wall (cavity: the top corners belong to the side walls)
    bool
    wall(const double x, const double y, double &ux, double &uy) const
    {
      const auto on = [&](const double a, const double a0) { return std::abs(a - a0) < 1e-12 * L; };
      ux = uy = 0.;
      if (kind == "couette")
        {
          if (on(y, L))
            ux = U0;
          return on(y, 0.) || on(y, L);
        }
      if (kind == "cavity")
        {
          if (on(x, 0.) || on(x, L) || on(y, 0.))
            return true;
          if (on(y, L))
            {
              ux = U0;
              return true;
            }
        }
      return false;
    }
    std::array<double, 3>
    exact(const double x, const double y, const double t) const
    {
      if (kind == "tgv")
        {
          const double F = std::exp(-nu * (k1() * k1() + k2() * k2()) * t), x1 = k1() * x, y2 = k2() * y,
                       r = k1() / k2();
          const double dp = -0.25 * rho0 * U0 * U0 * F * F * (std::cos(2 * x1) + r * r * std::cos(2 * y2));
          return {{rho0 + dp / CS2, -U0 * F * std::cos(x1) * std::sin(y2), U0 * r * F * std::sin(x1) * std::cos(y2)}};
        }
      if (kind == "couette")
        {
          double u = U0 * y / L;
          for (int m = 1; m <= 400; ++m)
            {
              const double lm = m * PI / L;
              u += 2 * U0 * ((m % 2) ? -1. : 1.) / (lm * L) * std::exp(-nu * lm * lm * t) * std::sin(lm * y);
            }
          return {{rho0, u, 0.}};
        }
      return {{rho0, 0., 0.}};
    }
    void
    equilibrium(const double rho, const double ux, const double uy, double *f) const
    {
      const double uu = ux * ux + uy * uy;
      for (int a = 0; a < Q; ++a)
        {
          const double eu = EX[a] * ux + EY[a] * uy;
          f[a]            = WQ[a] * rho * (1. + 3. * eu + 4.5 * eu * eu - 1.5 * uu);
        }
    }
    // schemes.h:50-59 initial populations: the equilibrium of the exact state, for the
    // vortex with the Chapman-Enskog non-equilibrium part (test_cases.h:92-118)
    void
    initial(const double x, const double y, const double lam, double *f) const
    {
      if (kind == "couette")
        return equilibrium(rho0, 0., 0., f);
      const auto e = exact(x, y, 0.);
      equilibrium(e[0], e[1], e[2], f);
      if (kind != "tgv")
        return;
      const double x1 = k1() * x, y2 = k2() * y;
      const double dudx = U0 * k1() * std::sin(x1) * std::sin(y2);
      const double sh   = U0 * (k1() * k1() - k2() * k2()) / k2() * std::cos(x1) * std::cos(y2);
      for (int a = 0; a < Q; ++a)
        f[a] -= lam / CS2 * WQ[a] * e[0] * ((EX[a] * EX[a] - EY[a] * EY[a]) * dudx + EX[a] * EY[a] * sh);
    }
  };

  // ------------------------------------------------------------ the driver
  class CGDBE
  {
  public:
    explicit CGDBE(const Parameters &p)
      : prm(p)
    {
      flow.kind    = prm.test_case;
      flow.U0      = prm.mach * std::sqrt(CS2);
      flow.stretch = prm.test_case == "cavity" ? prm.stretch : 0.;
      flow.n1      = prm.modes[0];
      flow.n2      = prm.modes[1];
      steady       = flow.steady();

      N = prm.cells > 0 ? prm.cells : (1 << prm.refine);
      lbfem_mesh_desc mesh{N, prm.degree, flow.n_periodic(), flow.L, flow.stretch,
                           prm.test_case == "tgv" ? prm.distort : 0.};
      check(lbfem_disc_create(&mesh, &disc));
      check(lbfem_disc_get_info(disc, &info));
      n = size_t(info.n_nodes);
      xy.resize(2 * n);
      check(lbfem_disc_node_coordinates(disc, xy.data()));
      slave.resize(n);
      check(lbfem_disc_constrained(disc, slave.data()));
      ml.resize(n);
      check(lbfem_disc_lumped_mass(disc, ml.data()));
      // walls: the nodes on a wall of the flow, with the wall's velocity (discretization.h:138-142)
      std::vector<int32_t> wn;
      std::vector<double>  wv;
      for (size_t i = 0; i < n; ++i)
        {
          double ux, uy;
          if (flow.wall(xy[2 * i], xy[2 * i + 1], ux, uy))
            {
              wn.push_back(int32_t(i));
              wv.push_back(ux);
              wv.push_back(uy);
            }
        }
      check(lbfem_disc_set_walls(disc, int32_t(wn.size()), wn.data(), wv.data()));
      walls.assign(n, {{NAN, NAN}});
      for (size_t k = 0; k < wn.size(); ++k)
        walls[wn[k]] = {{wv[2 * k], wv[2 * k + 1]}};

      lbfem_scheme_settings set{};
      set.scheme    = prm.scheme == "leelin" ? LBFEM_LEELIN : LBFEM_BARDOW;
      set.streaming = prm.streaming == "tg2" ? LBFEM_TG2 : prm.streaming == "tg3" ? LBFEM_TG3 : LBFEM_TG3_SPLIT;
      set.mass      = {prm.mass == "cg" ? LBFEM_MASS_CG : prm.mass == "lumped" ? LBFEM_MASS_LUMPED :
                                                                              LBFEM_MASS_RICHARDSON,
                       prm.richardson, prm.cg_tol, 200, prm.fused ? 1 : 0};
      set.precision = prm.precision == "mixed" ? LBFEM_MIXED : LBFEM_FP64;
      check(lbfem_scheme_create(disc, &set, &scheme));
      // StageTimers (timers.h): on request only, their GPU events would time themselves inside the run
      check(lbfem_scheme_enable_stage_timers(scheme, prm.stage_timers ? 1 : 0));

      // time step (cgdbe.cc:236-242): steps of the first Reynolds number; a transient hits t_end exactly
      flow.set_reynolds(prm.reynolds.front());
      const double T = prm.t_end * flow.reference_time();
      check(lbfem_disc_time_step(disc, prm.cfl, &dt));
      n_steps = prm.max_steps > 0 ? prm.max_steps : unsigned(std::ceil(T / dt));
      if (!steady && prm.max_steps == 0)
        dt = T / n_steps;
      check(lbfem_scheme_set_time_step(scheme, dt, flow.relaxation_time()));

      // initial condition, then the checkpoint if restarting
      std::vector<double> f0(Q * n), fi(Q);
      for (size_t i = 0; i < n; ++i)
        {
          flow.initial(xy[2 * i], xy[2 * i + 1], flow.relaxation_time(), fi.data());
          for (int a = 0; a < Q; ++a)
            f0[a * n + i] = fi[a];
        }
      check(lbfem_scheme_set_initial_populations(scheme, f0.data()));
      if (!prm.restart.empty())
        read_checkpoint(prm.restart);

      const char *streaming = prm.streaming == "tg2" ? "Lax-Wendroff" : prm.streaming == "tg3" ? "TG3" : "split TG3";
      std::cout << "cgdbe_hip (Lee & Lin 2001), liblbfem_hip on one GPU, " << prm.precision << "\n  case         : "
                << flow.name() << "\n  elements     : " << N * N << "  (Q" << prm.degree << ", " << n << " nodes x " << Q
                << " populations), h_min = " << info.h_min << "\n  scheme       : "
                << (prm.scheme == "leelin" ? std::string("Lee & Lin 2001 (predictor-corrector)") :
                                             std::string("Bardow et al. 2006 (collide, then ") + streaming +
                                               " streaming)")
                << "\n  mass matrix  : "
                << (prm.mass == "lumped" ? std::string("lumped") :
                    prm.mass == "cg"     ? std::string("consistent (CG + Jacobi, extrapolated start)") :
                                           "lumped + " + std::to_string(prm.richardson) + " Richardson pass(es)")
                << "\n  Ma = " << prm.mach << ", dt = " << dt << " (CFL " << dt / info.h_min << ")\n";
    }

    ~CGDBE()
    {
      lbfem_scheme_destroy(scheme);
      lbfem_disc_destroy(disc);
    }

    // cgdbe.cc:276-342: the Reynolds numbers in turn; false if a steady-state search ran out of time
    bool
    run()
    {
      bool all_steady = true;
      for (const double Re : prm.reynolds)
        {
          flow.set_reynolds(Re);
          t_ref = flow.reference_time();
          check(lbfem_scheme_set_time_step(scheme, dt, flow.relaxation_time()));
          time = time_previous = 0.;
          step_no              = 0;
          std::cout << "\n  Re = " << Re << ", nu = " << flow.nu << ", lambda = " << flow.relaxation_time()
                    << ", dt/lambda = " << dt / flow.relaxation_time() << ", at most " << n_steps
                    << " steps, t_ref = " << t_ref << "\n\n";
          const std::string tag = flow.tag(Re);
          history.close();
          if (prm.output)
            {
              history.open(tag + "_history.dat");
              history << (!steady                      ? "# t  E/E0  E/E0(exact)  rel_L2_error_u  mean_rho-rho0\n" :
                          prm.steady_max_change > 0. ? "# t  <u^2>/U0^2  -  max|du|/U0  <|du|>/U0\n" :
                                                       "# t  <u^2>/U0^2  -  <|du|>/U0/t_ref  mean_rho-rho0\n");
            }
          diagnostics();
          if (prm.output && !steady)
            write_gnuplot(tag + "_initial.dat", Re);

          const unsigned interval =
            !steady                      ? std::max(1u, n_steps / std::max(1u, prm.n_diagnostic)) :
            prm.steady_max_change > 0. ? prm.steady_window :
                                           std::max(1u, unsigned(std::lround(t_ref / dt)));
          bool converged = false;
          while (step_no < n_steps && !converged)
            {
              // all steps up to the next diagnostic line in one call (the GPU runs them back to back)
              const unsigned next = std::min(n_steps, (step_no / interval + 1) * interval);
              const unsigned k    = next - step_no;
              const auto     t0   = std::chrono::steady_clock::now();
              check(lbfem_scheme_step(scheme, int(k)));
              check(lbfem_scheme_synchronize(scheme));
              wall += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
              time += k * dt;
              step_no += k;
              total_steps += k;
              converged = diagnostics() && prm.max_steps == 0;
              if (!prm.checkpoint.empty())
                write_checkpoint(prm.checkpoint);
              if (converged)
                std::cout << "  -> steady state\n";
            }
          if (steady && !converged && prm.max_steps == 0)
            {
              std::cout << "  -> no steady state within t/t_ref = " << prm.t_end << "\n";
              all_steady = false;
            }
          if (prm.output)
            write_gnuplot(steady ? tag + ".dat" : tag + "_final.dat", Re);
          if (!steady)
            break;
        }
      print_summary();
      return all_steady;
    }

  private:
    // host copy of the populations, the periodic slave entries filled from their masters
    std::vector<double>
    populations() const
    {
      std::vector<double> f(Q * n);
      check(lbfem_scheme_get_populations(scheme, f.data()));
      const size_t np = size_t(info.n_per_dir);
      for (size_t j = 0; j < np; ++j)
        for (size_t i = 0; i < np; ++i)
          {
            const size_t k = j * np + i;
            if (!slave[k])
              continue;
            const size_t mi = (flow.n_periodic() >= 1 && i == np - 1) ? 0 : i;
            const size_t mj = (flow.n_periodic() >= 2 && j == np - 1) ? 0 : j;
            for (int a = 0; a < Q; ++a)
              f[a * n + k] = f[a * n + mj * np + mi];
          }
      return f;
    }

    static std::array<double, 3>
    moments(const std::vector<double> &f, const size_t n, const size_t i)
    {
      double rho = 0., mx = 0., my = 0.;
      for (int a = 0; a < Q; ++a)
        {
          rho += f[a * n + i];
          mx += EX[a] * f[a * n + i];
          my += EY[a] * f[a * n + i];
        }
      return {{rho, mx / rho, my / rho}};
    }

    // the exact velocity of the flow at time t, 2 per node (host)
    std::vector<double>
    exact_velocity(const double t) const
    {
      std::vector<double> u(2 * n);
      for (size_t i = 0; i < n; ++i)
        {
          const auto ex = flow.exact(xy[2 * i], xy[2 * i + 1], t);
          u[2 * i]      = ex[1];
          u[2 * i + 1]  = ex[2];
        }
      return u;
    }

    // cgdbe.cc:348-397: nodal quadrature over the non-slave nodes with lumped-mass weights, on the
    // device (lbfem_scheme_diagnostics): the vortex's exact field is the t = 0 field times its decay
    // factor, given once; the Couette series is given at each line; the cavity has none
    bool
    diagnostics()
    {
      double s[6] = {0., 0., 0., 0., 0., 0.};
      if (flow.kind == "tgv")
        {
          const double F = std::exp(-flow.nu * (flow.k1() * flow.k1() + flow.k2() * flow.k2()) * time);
          if (!exact_given)
            {
              const auto u0 = exact_velocity(0.);
              check(lbfem_scheme_diagnostics(scheme, u0.data(), F, s));
              exact_given = true;
            }
          else
            check(lbfem_scheme_diagnostics(scheme, nullptr, F, s));
        }
      else if (flow.kind == "couette")
        {
          const auto u = exact_velocity(time);
          check(lbfem_scheme_diagnostics(scheme, u.data(), 1., s));
        }
      else
        check(lbfem_scheme_diagnostics(scheme, nullptr, 0., s));
      const double mass_drift = s[3] / s[4] - flow.rho0;
      const double E0         = flow.energy_scale();
      auto         line       = [&](const std::array<double, 4> &v) {
        std::cout << std::setw(8) << step_no << std::setw(12) << time / t_ref << std::fixed << std::setw(12) << v[0]
                  << std::setw(12) << v[1] << std::scientific << std::setw(16) << v[2] << std::setw(14) << v[3]
                  << std::defaultfloat << std::endl;
        if (history.is_open())
          history << time << ' ' << v[0] << ' ' << v[1] << ' ' << v[2] << ' ' << v[3] << std::endl;
      };
      bool converged = false;
      if (!steady)
        {
          if (step_no == 0)
            std::cout << "    step     t/t_ref        E/E0  E/E0 exact   rel. L2 err(u)    <rho>-rho0\n";
          line({{s[2] / s[4] / E0, s[1] / s[4] / E0, std::sqrt(s[0] / std::max(s[1], 1e-300)), mass_drift}});
        }
      else
        {
          if (prm.steady_max_change > 0.)
            {
              // the largest velocity change since the previous line (the device monitor; a new Reynolds
              // number starts from the last line of the previous one, so its first line only records)
              double  dmax = 0., dmean = 0.;
              int64_t bad  = 0;
              check(lbfem_scheme_velocity_change(scheme, &dmax, &dmean, &bad));
              if (bad > 0)
                throw std::runtime_error("diverged at step " + std::to_string(step_no) + ": " +
                                         std::to_string(bad) + " nodes with non-finite moments");
              if (step_no == 0)
                std::cout << "    step     t/t_ref  <u^2>/U0^2           -      max|du|/U0     <|du|>/U0\n";
              const double m = step_no == 0 ? 1. : dmax / flow.U0;
              line({{s[2] / s[4] / E0, 0., m, step_no == 0 ? 1. : dmean / flow.U0}});
              converged = step_no > 0 && m < prm.steady_max_change;
            }
          else
            {
              if (step_no == 0)
                std::cout << "    step     t/t_ref  <u^2>/U0^2           -  <|du|>/U0/t_ref    <rho>-rho0\n";
              const double rate = step_no == 0 ? 1. : s[5] / s[4] / flow.U0 / ((time - time_previous) / t_ref);
              line({{s[2] / s[4] / E0, 0., rate, mass_drift}});
              converged = rate < prm.steady_tol;
            }
        }
      time_previous = time;
      return converged;
    }

    // cgdbe.cc:401-472: nodal rows sorted by y then x (already lexicographic), blank-line separated;
    // tgv/couette "x y u/U0 v/U0 |u|/U0", cavity "x y u/U0 v/U0 psi/(U0 L) rho-rho0" and the vortex extrema
    void
    write_gnuplot(const std::string &name, const double Re)
    {
      const auto                         f  = populations();
      const size_t                       np = size_t(info.n_per_dir);
      std::vector<std::array<double, 6>> rows(n);
      for (size_t i = 0; i < n; ++i)
        {
          const auto m = moments(f, n, i);
          rows[i]      = {{xy[2 * i], xy[2 * i + 1], m[1] / flow.U0, m[2] / flow.U0, std::hypot(m[1], m[2]) / flow.U0,
                           m[0] - flow.rho0}};
          if (!std::isnan(walls[i][0])) // the wall velocity on the wall nodes
            {
              rows[i][2] = walls[i][0] / flow.U0;
              rows[i][3] = walls[i][1] / flow.U0;
              rows[i][4] = std::hypot(walls[i][0], walls[i][1]) / flow.U0;
            }
        }
      if (steady)
        {
          std::vector<double> u(n), v(n), psi(n);
          for (size_t i = 0; i < n; ++i)
            {
              u[i] = rows[i][2];
              v[i] = rows[i][3];
            }
          lbfem_mesh_desc mesh{N, prm.degree, 0, flow.L, flow.stretch, 0.};
          int             its = 0;
          check(lbfem_stream_function(&mesh, u.data(), v.data(), psi.data(), &its));
          for (size_t i = 0; i < n; ++i)
            rows[i][4] = psi[i] / flow.L;
        }
      std::ofstream out(name);
      out << "# t/t_ref = " << time / t_ref << " (step " << step_no << "), Re = " << Re << ", Ma = " << prm.mach
          << ", " << N << "^2 Q" << prm.degree << " elements\n"
          << (steady ? "# x y u/U0 v/U0 psi/(U0 L) rho-rho0\n" : "# x y u/U0 v/U0 |u|/U0\n");
      const int cols = steady ? 6 : 5;
      for (size_t i = 0; i < n; ++i)
        {
          if (i > 0 && i % np == 0)
            out << '\n';
          for (int c = 0; c < cols; ++c)
            out << rows[i][c] << ' ';
          out << '\n';
        }
      if (!steady)
        return;
      // extrema of psi: primary, lower-left and lower-right vortices (Ghia et al., Table I)
      const auto extremum = [&](const char *label, const double sign, auto inside) {
        const std::array<double, 6> *best = nullptr;
        for (const auto &p : rows)
          if (inside(p) && (!best || sign * p[4] > sign * (*best)[4]))
            best = &p;
        std::cout << "  " << std::left << std::setw(20) << label << std::right << " psi = " << std::setw(12)
                  << (*best)[4] << "  at (" << (*best)[0] << ", " << (*best)[1] << ")\n";
        out << "# " << label << ": psi = " << (*best)[4] << " at " << (*best)[0] << ' ' << (*best)[1] << '\n';
      };
      std::cout << "\n";
      extremum("primary vortex", +1., [](const auto &) { return true; });
      extremum("lower left vortex", -1., [](const auto &p) { return p[0] < 0.5 && p[1] < 0.5; });
      extremum("lower right vortex", -1., [](const auto &p) { return p[0] > 0.5 && p[1] < 0.5; });
    }

    // include/lbfem/io.h:138-165: a raw dump of the populations, 9 blocks of n doubles, file name.0
    void
    write_checkpoint(const std::string &name) const
    {
      std::vector<double> f(Q * n);
      check(lbfem_scheme_get_populations(scheme, f.data()));
      const std::string file = name + ".0";
      {
        std::ofstream out(file + ".tmp", std::ios::binary);
        out.write(reinterpret_cast<const char *>(f.data()), std::streamsize(f.size() * sizeof(double)));
        if (!out.good())
          throw std::runtime_error("cannot write checkpoint " + file);
      }
      std::rename((file + ".tmp").c_str(), file.c_str()); // never leave a torn file
    }

    void
    read_checkpoint(const std::string &name)
    {
      std::vector<double> f(Q * n);
      const std::string   file = name + ".0";
      std::ifstream       in(file, std::ios::binary);
      in.read(reinterpret_cast<char *>(f.data()), std::streamsize(f.size() * sizeof(double)));
      if (!in.good())
        throw std::runtime_error("cannot read checkpoint " + file);
      check(lbfem_scheme_set_populations(scheme, f.data()));
    }

    // cgdbe.cc:479-540 (the throughput lines; the stage breakdown needs stage timers)
    void
    print_summary() const
    {
      int64_t its = 0;
      check(lbfem_scheme_cg_iterations(scheme, &its));
      const double solves = prm.streaming == "tg3-split" ? 12. : 8.;
      std::cout << "\n  time stepping : " << wall << " s for " << total_steps << " steps  ("
                << 1e3 * wall / std::max(1u, total_steps) << " ms/step)\n"
                << "  throughput    : " << 1e-6 * double(n) * total_steps / wall
                << " million node updates/s (MNUPS = MDoF/s, one DoF = all " << Q << " populations)\n";
      const double it_per_solve = double(its) / (std::max(1u, total_steps) * solves);
      if (prm.mass == "cg")
        std::cout << "  CG iterations : " << it_per_solve << " per mass solve\n";

      // cgdbe.cc:479-540: the stage times with the reference's single-pass traffic model (a vector pass is
      // 8 n bytes) and flop model (a multiply-add counts 2), per node and step. The fused lumped step runs
      // collision, advection and the lumped solve in one kernel: its time is under advection.
      double t[4] = {0., 0., 0., 0.};
      check(lbfem_scheme_stage_times(scheme, t));
      if (!prm.stage_timers)
        {
          std::cout << "  (the stage breakdown: run with --stage-timers)\n";
          return;
        }
      const bool   leelin = prm.scheme == "leelin";
      const double sweeps = prm.streaming == "tg3-split" ? 2. : 1.;
      const double gb     = 1e-9 * double(n) * 8. * total_steps;
      struct Work
      {
        double passes, flops;
      };
      const Work   collision = leelin ? Work{53., 303.} : Work{18. + 3. * solves, 149. + solves};
      const double cells_per_node = double(N) * N / double(n);
      const double flops_cell     = 184. * (prm.degree == 1 ? 1. : 3.5 * prm.degree) * 8. * cells_per_node;
      const Work   advection      = {sweeps * (leelin ? 24. : 16.), sweeps * flops_cell * (leelin ? 1.4 : 1.)};
      const bool   fused_cg       = prm.fused;
      const Work   per_solve      = prm.mass == "cg" ? Work{(fused_cg ? 8. : 10.) * it_per_solve + 4.,
                                                             (68. + (fused_cg ? 21. : 11.)) * it_per_solve} :
                                    prm.mass == "lumped" ? Work{2., 1.} :
                                                           Work{3. + (fused_cg ? 5. : 7.) * prm.richardson,
                                                                1. + 71. * prm.richardson};
      const Work   solve_work     = {per_solve.passes * solves, per_solve.flops * solves};
      const double seconds        = std::max(wall, 1e-300);
      const auto   stage          = [&](const char *name, const double sec, const Work &w) {
        std::cout << "  " << std::left << std::setw(12) << name << std::right << std::setw(9) << std::fixed
                  << std::setprecision(3) << sec << " s " << std::setw(5) << std::setprecision(1)
                  << 100 * sec / seconds << " % ";
        if (sec > 0.) // a stage fused into another has no time of its own
          std::cout << std::setw(6) << std::setprecision(2) << w.passes * gb / sec << " GB/s " << std::setw(6)
                    << std::setprecision(2) << 1e-9 * w.flops * double(n) * total_steps / sec << " GFlop/s  (";
        else
          std::cout << "(fused; ";
        std::cout << std::setprecision(0) << w.passes << " vector passes, " << w.flops
                  << " flops per node and step, intensity " << std::setprecision(2) << w.flops / (w.passes * 8.)
                  << " flop/byte)" << std::defaultfloat << "\n";
      };
      std::cout << "  breakdown of the time stepping (GPU stage timers; single-pass traffic model, flop model"
                << (prm.mass == "lumped" && prm.streaming == "tg2" ? "; fused step: all under advection" : "")
                << "):\n";
      stage("collision", t[0], collision);
      stage("advection", t[1], advection);
      stage("mass solves", t[2], solve_work);
      std::cout << "  other (timers, loop overhead): " << std::fixed << std::setprecision(3)
                << std::max(0., wall - t[3]) << " s" << std::defaultfloat << "\n";
    }

    const Parameters             prm;
    Flow                         flow;
    bool                         steady = false;
    int                          N      = 0;
    lbfem_disc                  *disc   = nullptr;
    lbfem_scheme                *scheme = nullptr;
    lbfem_disc_info              info{};
    size_t                       n = 0;
    std::vector<double>          xy, ml;
    std::vector<int32_t>         slave;
    std::vector<std::array<double, 2>> walls;
    bool                               exact_given = false; // the vortex's t = 0 field is on the device
    std::ofstream                history;
    double                       dt = 0., t_ref = 1., time = 0., time_previous = 0., wall = 0.;
    unsigned                     n_steps = 0, step_no = 0, total_steps = 0;
  };
} // namespace

int
main(int argc, char **argv)
{
  try
    {
      bool             help = false;
      const Parameters prm  = parse(argc, argv, help);
      if (help)
        {
          std::cout << usage;
          return 0;
        }
      if (!CGDBE(prm).run())
        return 2;
    }
  catch (const std::exception &exc)
    {
      std::cerr << exc.what() << std::endl;
      return 1;
    }
  return 0;
}
