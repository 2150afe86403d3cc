"""Benchmark of the D2Q9 Q_p time step on one MI355X.

Workload (the plain-method GPU configuration): lid-driven cavity on the unit
square, U0 = 0.1 (Ma = 0.1732), tanh clustering 1.2, Q1 on 256 x 256 cells,
Lee-Lin predictor-corrector, Re 1000, CFL 0.5 (dt / lambda = 2.88), FP64. Of the
configuration's two Reynolds numbers, Lee-Lin holds on this mesh at Re 1000 only
(at Re 400, dt / lambda = 1.15, it diverges within a few hundred steps with either
mass, in the oracle as on the GPU). The headline value is the lumped-mass step
(one fused kernel per step, 66 049 nodes: a cache-resident, latency-bound launch);
the consistent-mass (CG, tol 1e-10) step, Bardow on Q1 256^2 and Q2 128^2 at
Re 400, the mixed-precision form and the HBM-resident Q1 2048^2 step (0.6 GB of
populations per step, beyond the 256 MB Infinity Cache; Re 8000 keeps
dt / lambda = 2.88) with its own roofline are reported beside it; --sweep adds
4096^2. A STREAM-like copy measures the achievable HBM bandwidth of the device.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--sweep]
       python bench.py --profile-headline   # the headline steps and the copy only (profiler passes)
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lbm-fem_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0  # MI355X datasheet HBM3E bandwidth
Q = 9
# kernel names of the tile variants (Cfg<T, p, TX, TY, PS, mode>), fused Lee-Lin step, Q1; variants 8-11
# are the row-march kernel (Q1; elsewhere they run as variant 0)
TILE_SHAPES = ((32, 8, 1), (32, 16, 1), (32, 8, 2), (32, 8, 4), (16, 8, 4), (40, 8, 1), (48, 8, 1), (40, 8, 2))
N_VARIANTS = 12


def tile_kernel(variant, t="double"):
    if variant >= len(TILE_SHAPES):
        return f"k_march<{t}, 1, 1>"  # Lee-Lin step, Q1
    tx, ty, ps = TILE_SHAPES[variant]
    return f"k_tile<Cfg<{t}, 1, {tx}, {ty}, {ps}, 1> >"


TILE_KERNEL = {v: tile_kernel(v) for v in range(N_VARIANTS)}
PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_k_tile_cavity_q1_256.txt")
PMC_SUMMARY_HBM = os.path.join(ROOT, "profiles", "pmc_k_march_cavity_q1_2048.txt")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_scheme(A, N, p, leelin=True, mass=None, cg_tol=1e-10, Re=1000.0, cfl=0.5, mixed=False, tune=True):
    tc = A.test_cases.make("cavity", mach=0.1 / (1 / 3) ** 0.5, stretch=1.2)  # U0 = Ma c_s = 0.1
    tc.set_reynolds(Re)
    disc = A.Discretization(tc, N, p)
    cls = A.LeeLin if leelin else A.Bardow
    s = cls(disc, mass=A.MASS_LUMPED if mass is None else mass, cg_tolerance=cg_tol,
            precision=A.MIXED if mixed else A.FP64)
    dt = disc.time_step(cfl)
    s.set_time_step(dt, tc.relaxation_time())
    s.set_initial_populations(tc)
    if tune:  # the tile shape is measured once per kernel and mesh, before any timing
        s.autotune()
    return disc, s


def timed(torch, s, steps, warmup):
    s.step(warmup)
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_time(torch, s, steps):
    """Average time per step from two HIP events on the scheme's own stream around `steps` queued steps.
    A lumped TG2 step is one launch of the fused tile kernel, so this is that kernel's average launch
    duration (back-to-back launches on one stream)."""
    st = torch.cuda.ExternalStream(s.stream_handle)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.synchronize()
    e0.record(st)
    s.step(steps)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / steps


def hbm_probe(torch, A, nbytes=1 << 30, reps=10):
    """Copy bandwidth (read + write bytes / time) of lbfem_hbm_copy over nbytes, HIP events on torch's stream."""
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    st = torch.cuda.current_stream()
    A.hbm_copy(dst, src, st.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        A.hbm_copy(dst, src, st.cuda_stream)
    e1.record(st)
    e1.synchronize()
    gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    ok = bool(torch.equal(dst[:1024], src[:1024]))
    del src, dst
    return gbs, ok


def pmc_traffic(kernel, path=PMC_SUMMARY):
    """HBM bytes per launch of a tile kernel (variant) read from a committed profiler summary (FETCH_SIZE +
    WRITE_SIZE per dispatch, scaled by the copy kernel's reading of its known size), or None. A lookup of an
    earlier profiler run of the same command, not a measurement of this run."""
    key = f"traffic_bytes_per_launch[{kernel}]:"
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


CPU_CHILD = r"""
import json, os, sys, time
sys.path.insert(0, os.path.join(sys.argv[1], "oracle"))
import oracle as O
steps = int(sys.argv[2])
od = O.Disc(256, 1, O.CAVITY, mach=0.1 / (1 / 3) ** 0.5, stretch=1.2, reynolds=1000.0)
s = O.Scheme(od, O.LEELIN, O.TG2, O.MASS_LUMPED)
s.set_time_step(od.time_step(0.5), od.relaxation_time)
s.set_initial_populations()
s.step(2)
t0 = time.perf_counter()
s.step(steps)
dt = time.perf_counter() - t0
print(json.dumps({"MLUPS": round(9 * od.n * steps / dt / 1e6, 3), "steps": steps, "seconds": dt}))
"""


def host_cpu():
    """The host CPU model (lscpu's model name), for the baseline's record."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_oracle_rate(threads):
    """The oracle's headline-workload rate on `threads` host threads, in a child process (no GPU there)."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), LBFEM_ORACLE_OMP="1" if threads > 1 else "0")
    steps = 200 if threads == 1 else 600  # about 7 s on one core, 30 core-seconds on 16 threads
    r = subprocess.run([sys.executable, "-c", CPU_CHILD, ROOT, str(steps)], env=env, capture_output=True,
                       text=True, timeout=300, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep", action="store_true")
    ap.add_argument("--no-hbm", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-general", action="store_true")
    ap.add_argument("--precision-steps", type=int, default=5000)
    ap.add_argument("--profile-headline", action="store_true")
    ap.add_argument("--profile-cells", type=int, default=256)
    ap.add_argument("--profile-general", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(the sharding design is a domain decomposition, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    import lbfem_amd as A

    assert torch.cuda.is_available(), "bench needs a GPU"
    K, W = args.steps, args.warmup

    if args.profile_headline:
        # profiler passes: the headline steps, then the copy probe over 256 MiB (known size)
        # the size-based default shape (the one the autotuner picks at 257^2 and 2049^2 nodes), untuned, so
        # that the profile holds that kernel's launches only
        N = args.profile_cells
        disc, s = make_scheme(A, N, 1, Re=1000.0 * N / 256, tune=False)  # dt / lambda 2.88 on every mesh size
        log(f"profile-headline: tile variant {s.tile_variant} ({TILE_KERNEL[s.tile_variant]})")
        timed(torch, s, K, W)
        tk = kernel_time(torch, s, K)
        gbs, ok = hbm_probe(torch, A, nbytes=256 << 20, reps=3)
        log(f"profile-headline: {2 * K + W} steps of cavity Q1 {N}^2 (HIP-event kernel time {1e6 * tk:.2f} us over the "
            f"last {K}), copy probe 3 x 256 MiB at {gbs:.0f} GB/s (ok={ok})")
        return

    if args.profile_general:
        # profiler passes of the general-mesh step kernel: distorted vortex, Q1 1024^2, Bardow lumped, then the
        # copy probe over 256 MiB (known size)
        tc = A.test_cases.TaylorGreenVortex(0.1, 1, 4, distort=0.04)
        tc.set_reynolds(100.0)
        dg = A.Discretization(tc, args.profile_cells, 1)
        sg = A.Bardow(dg, mass=A.MASS_LUMPED)
        sg.set_time_step(dg.time_step(0.25), tc.relaxation_time())
        sg.set_initial_populations(tc)
        timed(torch, sg, K, W)
        gbs, ok = hbm_probe(torch, A, nbytes=256 << 20, reps=3)
        log(f"profile-general: {K + W} steps of the distorted vortex Q1 {args.profile_cells}^2 (Bardow, lumped), "
            f"copy probe 3 x 256 MiB at {gbs:.0f} GB/s (ok={ok})")
        return

    hbm_gbs, hbm_ok = hbm_probe(torch, A)
    log(f"copy probe: {hbm_gbs:.0f} GB/s (read + write), checked={hbm_ok}")

    # ---- headline: cavity Q1 256^2, Lee-Lin, lumped (fused kernel)
    disc, s = make_scheme(A, 256, 1, tune=False)
    variant, tune_s = s.autotune(force=True)
    log("autotune (s per launch by variant):", [round(1e6 * t, 2) for t in tune_s], "us -> variant", variant)
    n = disc.n_nodes
    sec = timed(torch, s, K, W)
    value = 9 * n * K / sec / 1e6
    t_kernel = kernel_time(torch, s, K)
    alg_bytes = 18 * 8 * n
    achieved = alg_bytes / t_kernel / 1e9
    head_finite = bool(np.isfinite(s.f).all())
    log(f"cavity Q1 256^2 leelin lumped Re 1000: {1e3 * sec / K:.4f} ms/step, {value:.1f} MLUPS, kernel "
        f"{1e6 * t_kernel:.2f} us, {achieved:.0f} GB/s algorithmic, finite={head_finite}")

    # ---- the stream function of the headline flow's velocity (include/lbfem/io.h:31-93): Laplace CG on the device
    _, u_h, v_h = A.moments(s.f)
    u_d = torch.tensor(u_h, device="cuda", dtype=torch.float64)
    v_d = torch.tensor(v_h, device="cuda", dtype=torch.float64)
    psi_d = torch.empty_like(u_d)
    sf_its = A.disc_stream_function(disc, u_d, v_d, psi_d)
    t0 = time.perf_counter()
    for _ in range(3):
        A.disc_stream_function(disc, u_d, v_d, psi_d)
    sf_ms = 1e3 * (time.perf_counter() - t0) / 3
    stream_fn = {"workload": f"cavity Q1 256^2 velocity after {K + W} headline steps, CG 1e-12 |b|",
                 "cg_iterations": sf_its, "ms": round(sf_ms, 3), "us_per_iteration": round(1e3 * sf_ms / sf_its, 2)}
    log("stream function", stream_fn)
    del u_d, v_d, psi_d

    variants = {}
    # ---- the consistent mass (CG tol 1e-10), and the points where Lee-Lin diverges (Re 400, and Q2 128^2 at
    # either Reynolds number) with Bardow, the reference's default scheme, which holds there
    for name, N, p, mass, Re in (("q1_256_leelin_cg_re1000", 256, 1, A.MASS_CG, 1000.0),
                                 ("q1_256_bardow_lumped_re400", 256, 1, A.MASS_LUMPED, 400.0),
                                 ("q1_256_bardow_cg_re400", 256, 1, A.MASS_CG, 400.0),
                                 ("q2_128_bardow_lumped_re400", 128, 2, A.MASS_LUMPED, 400.0),
                                 ("q2_128_bardow_cg_re400", 128, 2, A.MASS_CG, 400.0),
                                 ("q1_256_leelin_lumped_re1000_mixed", 256, 1, A.MASS_LUMPED, 1000.0),
                                 ("q2_128_bardow_lumped_re400_mixed", 128, 2, A.MASS_LUMPED, 400.0),
                                 ("q1_256_leelin_cg_re1000_mixed", 256, 1, A.MASS_CG, 1000.0)):
        leelin = "leelin" in name
        d2, s2 = make_scheme(A, N, p, leelin=leelin, mass=mass, mixed=name.endswith("mixed"), Re=Re)
        steps = max(10, K // 4) if mass == A.MASS_CG else K
        its0 = s2.cg_iterations
        sec2 = timed(torch, s2, steps, W)
        its = (s2.cg_iterations - its0) / (8 * (steps + W)) if mass == A.MASS_CG else 0.0
        variants[name] = {"ms_per_step": round(1e3 * sec2 / steps, 5),
                          "MLUPS": round(9 * d2.n_nodes * steps / sec2 / 1e6, 1),
                          "cg_its_per_solve": round(its, 2), "tile_variant": s2.tile_variant,
                          "finite": bool(np.isfinite(s2.f).all())}
        log(name, variants[name])
        del s2, d2

    # ---- general quadrilaterals: the vortex (1,4) on the distorted mesh (eps 0.04, cells +-25 %), Q1 1024^2,
    # Re 100, CFL 0.25: the fused lumped step (one cell-kernel launch) and the consistent mass (CG 1e-8)
    general = None
    if not args.no_general:
        general = {"workload": "Taylor-Green (1,4) on the distorted mesh x += 0.04 L sin(2 pi x) sin(2 pi y) (1,1), "
                               "Q1 1024^2 cells, Re 100, Ma 0.1, CFL 0.25"}
        for name, leelin, mass in (("bardow_lumped", False, A.MASS_LUMPED), ("leelin_lumped", True, A.MASS_LUMPED),
                                   ("bardow_cg", False, A.MASS_CG)):
            tc = A.test_cases.TaylorGreenVortex(0.1, 1, 4, distort=0.04)
            tc.set_reynolds(100.0)
            dg = A.Discretization(tc, 1024, 1)
            cls = A.LeeLin if leelin else A.Bardow
            sg = cls(dg, mass=mass, cg_tolerance=1e-8)
            sg.set_time_step(dg.time_step(0.25), tc.relaxation_time())
            sg.set_initial_populations(tc)
            steps = max(10, K // 4) if mass == A.MASS_CG else K
            its0 = sg.cg_iterations
            secg = timed(torch, sg, steps, W)
            row = {"ms_per_step": round(1e3 * secg / steps, 4), "MLUPS": round(9 * dg.n_nodes * steps / secg / 1e6, 1)}
            if mass == A.MASS_CG:
                row["cg_its_per_solve"] = round((sg.cg_iterations - its0) / (8 * (steps + W)), 2)
            else:
                tk = kernel_time(torch, sg, steps)
                # populations in and out (144 B per node) and the vertices (16 B per cell corner, about per node)
                gbs = (144 + 16) * dg.n_nodes / tk / 1e9
                row.update({"kernel_us": round(1e6 * tk, 2), "achieved_GBs": round(gbs, 1),
                            "frac_hbm": round(gbs / HBM_PEAK_GBS, 3)})
            general[name] = row
            log("general", name, row)
            del sg, dg

    # ---- HBM-resident step (working set > 256 MB Infinity Cache): Q1 2048^2 (0.6 GB of FP64 populations read
    # and written per step), FP64 and mixed; with --sweep also 4096^2. Re scaled with N (Re 1000 N / 256) keeps
    # dt / lambda = 2.88, the headline's stable Lee-Lin point.
    sweep, roofline_hbm = [], None
    legs = [(2048, 40, False), (2048, 40, True)]
    if args.sweep:
        legs += [(4096, 12, False), (4096, 12, True)]
    if not args.no_hbm:
        for N, steps, mixed in legs:
            d3, s3 = make_scheme(A, N, 1, mixed=mixed, Re=1000.0 * N / 256)
            sec3 = timed(torch, s3, steps, 3)
            tk = kernel_time(torch, s3, steps)
            alg3 = 18 * (4 if mixed else 8) * d3.n_nodes
            gbs = alg3 / tk / 1e9
            kern = tile_kernel(s3.tile_variant, "float" if mixed else "double")
            sweep.append({"cells": f"{N}x{N}", "degree": 1, "form": "mixed" if mixed else "fp64",
                          "Re": 1000.0 * N / 256, "n_nodes": d3.n_nodes, "tile_variant": s3.tile_variant,
                          "ms_per_step": round(1e3 * sec3 / steps, 4),
                          "MLUPS": round(9 * d3.n_nodes * steps / sec3 / 1e6, 1), "kernel": kern,
                          "kernel_us": round(1e6 * tk, 2), "achieved_GBs": round(gbs, 1),
                          "frac_hbm": round(gbs / HBM_PEAK_GBS, 4), "finite": bool(np.isfinite(s3.f).all())})
            log("hbm", sweep[-1])
            if N == 2048 and not mixed:
                roofline_hbm = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(gbs / HBM_PEAK_GBS, 4),
                                "traffic": pmc_traffic(kern, PMC_SUMMARY_HBM),
                                "traffic_source": "lookup: profiles/pmc_k_march_cavity_q1_2048.txt (an earlier "
                                                  "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE run of this leg)",
                                "kernel": kern, "kernel_us": round(1e6 * tk, 2), "algorithmic_bytes_per_launch": alg3,
                                "workload": "cavity Q1 2048^2 cells (4 198 401 nodes), Lee-Lin, lumped, Re 8000, FP64"}
            del s3, d3

    # ---- precision: TGV (1,1) on Q1 1024^2 and Q2 512^2, Bardow, consistent mass (CG 1e-10), FP64 and mixed (FP32
    # populations and CG vectors, FP64 scalars and sums); L2 velocity error of the mixed run vs FP64 and both vs the
    # analytic decay
    precision = None
    if not args.no_precision:
        steps_p = args.precision_steps
        precision = {"workload": "Taylor-Green (1,1), U0 0.01, Bardow, consistent mass CG 1e-10, CFL 0.25, Re 100, "
                                 f"{steps_p} steps after 5; Q1 1024^2 and Q2 512^2 cells"}
        for label, N, p in (("q1_1024", 1024, 1), ("q2_512", 512, 2)):
            runs = {}
            # the mixed form stops its FP32 solves at the round-off level of its right-hand side (a floor that
            # scales with the nodes per mesh line: 9.7e-6 |b| on these meshes) whatever the tolerance asked
            for prec, name, tol in ((A.FP64, "fp64", 1e-10), (A.MIXED, "mixed", 1e-10)):
                tc = A.test_cases.make("tgv", mach=0.01 * 3 ** 0.5, modes=(1, 1))  # U0 = 0.01 (lattice units)
                tc.set_reynolds(100.0)
                dp = A.Discretization(tc, N, p)
                sp = A.Bardow(dp, mass=A.MASS_CG, cg_tolerance=tol, precision=prec)
                dtp = dp.time_step(0.25)
                sp.set_time_step(dtp, tc.relaxation_time())
                sp.set_initial_populations(tc)
                sp.step(5)
                sp.synchronize()
                its0 = sp.cg_iterations
                t0 = time.perf_counter()
                sp.step(steps_p)
                sp.synchronize()
                tp = time.perf_counter() - t0
                _, u, v = A.moments(sp.f)
                runs[name] = {"ms_per_step": round(1e3 * tp / steps_p, 3),
                              "MLUPS": round(9 * dp.n_nodes * steps_p / tp / 1e6, 1),
                              "cg_its_per_solve": round((sp.cg_iterations - its0) / (8 * steps_p), 2), "u": u,
                              "v": v}
                xy, indep, w = dp.node, ~dp.constrained(), dp.lumped_mass()
                log("precision", label, name, {k: runs[name][k] for k in ("ms_per_step", "cg_its_per_solve")})
                del sp
            _, ue, ve = tc.exact(xy[:, 0], xy[:, 1], (steps_p + 5) * dtp)
            l2 = lambda u, v, ur, vr: float(np.sqrt((w[indep] * ((u - ur) ** 2 + (v - vr) ** 2)[indep]).sum()
                                                    / (w[indep] * (ur ** 2 + vr ** 2)[indep]).sum()))
            row = {"l2_velocity_mixed_vs_fp64": l2(runs["mixed"]["u"], runs["mixed"]["v"], runs["fp64"]["u"],
                                                   runs["fp64"]["v"]),
                   "l2_velocity_vs_analytic": {k: l2(r["u"], r["v"], ue, ve) for k, r in runs.items()}}
            for k, r in runs.items():
                row[k] = {x: r[x] for x in ("ms_per_step", "MLUPS", "cg_its_per_solve")}
            precision[label] = row
            log("precision", label, row)

    # ---- batched: 96 seeded uneven instances (flow, N 32..512, Q1/Q2, Re, U0), Bardow lumped, one launch per
    # (degree, scheme, precision) group per step, graph-replayed, autotuned shapes
    batched = None
    if not args.no_batch:
        from lbfem_amd import workloads as WL
        insts = WL.batch()
        tot = sum(WL.nodes(i) for i in insts)
        batched = {"workload": "96 instances seeded 20261203 (splitmix64): tgv/couette/cavity, N = 32 16^u, "
                               "Q1 p=0.75 else Q2, Re = 50 64^u, U0 in [0.02, 0.1]; Bardow, lumped, CFL 0.5",
                   "instances": len(insts), "n_nodes": tot}
        for prec, name in ((A.FP64, "fp64"), (A.MIXED, "mixed")):
            ss = [WL.make_scheme(A, i, prec) for i in insts]
            b = A.Batch(ss)
            b.autotune()
            b.step(16)
            b.synchronize()
            kb = 160
            t0 = time.perf_counter()
            b.step(kb)
            b.synchronize()
            tb = time.perf_counter() - t0
            b.close()
            for x in ss:
                x.synchronize()
            k1 = 40
            t0 = time.perf_counter()
            for _ in range(k1):
                for x in ss:
                    x.step(1)
            for x in ss:
                x.synchronize()
            t1 = time.perf_counter() - t0
            nonfinite = sum(0 if np.isfinite(x.f).all() else 1 for x in ss)
            mlups_b = 9 * tot * kb / tb / 1e6
            batched[name] = {"ms_per_step": round(1e3 * tb / kb, 4), "MLUPS": round(mlups_b, 1),
                             "one_by_one_MLUPS": round(9 * tot * k1 / t1 / 1e6, 1),
                             "vs_plain_headline": round(mlups_b / value, 2),
                             # against the plain FP64 step where it too is throughput-bound (Q1 2048^2, from HBM)
                             "vs_plain_hbm_fp64": round(mlups_b / sweep[0]["MLUPS"], 2) if sweep else None,
                             "nonfinite_members": nonfinite}
            log("batched", name, batched[name])
            del ss, b
        # the same instances with the consistent mass (FP64, CG 1e-8): each stage of the step one launch over
        # the members of a (degree, scheme, streaming) group, one host check of all stopping states per
        # iteration; against the same schemes stepped one by one
        ss = [WL.make_scheme(A, i, A.FP64, mass=A.MASS_CG) for i in insts]
        b = A.Batch(ss)
        b.step(2)
        b.synchronize()
        its0 = sum(x.cg_iterations for x in ss)
        kb = 10
        t0 = time.perf_counter()
        b.step(kb)
        b.synchronize()
        tb = time.perf_counter() - t0
        its = (sum(x.cg_iterations for x in ss) - its0) / (8 * len(ss) * kb)
        b.close()
        k1 = 3
        t0 = time.perf_counter()
        for _ in range(k1):
            for x in ss:
                x.step(1)
        for x in ss:
            x.synchronize()
        t1 = time.perf_counter() - t0
        batched["fp64_cg"] = {"ms_per_step": round(1e3 * tb / kb, 3), "MLUPS": round(9 * tot * kb / tb / 1e6, 1),
                              "one_by_one_MLUPS": round(9 * tot * k1 / t1 / 1e6, 1),
                              "cg_its_per_solve": round(its, 2),
                              "nonfinite_members": sum(0 if np.isfinite(x.f).all() else 1 for x in ss)}
        log("batched", "fp64_cg", batched["fp64_cg"])
        del ss

    # ---- served: open-loop Poisson streams of seeded cavity/Couette requests, mixed precision, continuous
    # batching (one running batch; requests join and leave it); time-to-solution percentiles per arrival rate
    served = None
    if not args.no_serve:
        from lbfem_amd import serving, workloads as WL
        served = {"workload": "cavity/Couette, N in {32, 64, 128, 256, 512, 1024}, Q2 p=0.25, Re log-uniform "
                              "[100, 5000], U0 in [0.02, 0.1]; half to the steady tolerance 1e-6 (max|du|/U0 over "
                              "1000 steps, capped at the step draw), half for their 2000-20000 steps; diverged "
                              "requests retired with that status; Bardow, lumped, CFL 0.5; mixed precision; "
                              "budget 240 GB, at most 96 resident; 32 requests per rate", "rates": []}
        ref = batched["mixed"]["MLUPS"] if batched else None
        for rate in (10, 20, 50, 100, 200):
            reqs = WL.request_stream(20261203 + rate, 32, rate=float(rate), sizes=(32, 64, 128, 256, 512, 1024))
            done, summ = serving.Server(A).run(reqs)
            row = {"rate_per_s": rate, "requests": summ["requests"], "statuses": summ["statuses"],
                   "tts_p50_s": round(summ["tts_p50_s"], 4), "tts_p99_s": round(summ["tts_p99_s"], 4),
                   "MLUPS": round(summ["MLUPS"], 1), "wall_s": round(summ["wall_s"], 3),
                   "vs_batched_mixed": round(summ["MLUPS"] / ref, 2) if ref else None,
                   "vs_plain_headline": round(summ["MLUPS"] / value, 2)}
            served["rates"].append(row)
            log("served", row)

    # ---- CPU baseline: the oracle (C++ restatement of the reference algorithm) on the host cores, a bounded
    # sample of the same workload; one core in this process, all cores (OpenMP build, bitwise the same results)
    # in a child process
    cpu = None
    if not args.no_cpu:
        one = cpu_oracle_rate(threads=1)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        many = cpu_oracle_rate(threads=threads)
        cpu = {"value": many["MLUPS"], "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": f"{many['steps']} steps of the same workload (cavity Q1 256^2, Lee-Lin, lumped, Re 1000) on "
                         f"{threads} host threads (OpenMP build of the oracle), {many['seconds']:.1f} s; one core: "
                         f"{one['MLUPS']} MLUPS ({one['steps']} steps, {one['seconds']:.1f} s)",
               "single_core_MLUPS": one["MLUPS"], "host_cpu": host_cpu(), "host_nproc": os.cpu_count()}
        log("cpu", cpu)

    out = {
        "metric": "MLUPS (node·pop updates/s)",
        "value": round(value, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * sec / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (lid-driven cavity from rest, seeded by the flow definition)",
        "config": {"workload": "lid-driven cavity, Q1 256x256 cells (66049 nodes), Lee-Lin, lumped mass, "
                               "Re 1000, CFL 0.5 (dt/lambda 2.88), U0 0.1, stretch 1.2",
                   "n_nodes": n, "mnups": round(value / 9, 3), "finite": head_finite},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(TILE_KERNEL[variant]),
                     "traffic_source": "lookup: profiles/pmc_k_tile_cavity_q1_256.txt (an earlier rocprofv3 --pmc "
                                       "FETCH_SIZE / WRITE_SIZE run of this workload)",
                     "regime": "latency-bound: 9.5 MB per step is cache-resident and 66 049 nodes give about one "
                               "wavefront per SIMD; roofline_hbm is the same kernel from HBM",
                     "kernel": TILE_KERNEL[variant], "kernel_us": round(1e6 * t_kernel, 3),
                     "algorithmic_bytes_per_launch": alg_bytes, "measured_copy_GBs": round(hbm_gbs, 1)},
        "roofline_hbm": roofline_hbm,
        "cpu_baseline": cpu,
        "tile_variant": variant, "autotune_us": [round(1e6 * t, 3) for t in tune_s],
        "variants": variants,
        "stream_function": stream_fn,
        "hbm_sweep": sweep,
        "precision": precision,
        "general_quadrilaterals": general,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
