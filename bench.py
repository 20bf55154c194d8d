#!/usr/bin/env python3
"""Benchmark of the integral + k-point eigensolve hot path on one MI355X.

Workload (one "step"): the Si64 supercell (2x2x2 conventional diamond Si,
64 atoms, seeded 0.05 A displacements, s,s,p,s,p,d = 14 functions/atom,
n = 896, 2x2x2 Monkhorst-Pack mesh = 4 irreducible k-points, lattice
cutoff 16 bohr, potDim = 1024), taken through one full pass of the
reference's hot path:
  * imago_set_system   pair/image screening tables, Bloch phases
  * imago_assemble     OL, KE, NP and the 1024 EP terms, all k-points,
                       into HBM-resident packed panels (26.3 GB of EP)
  * imago_secular      H(k) = NP + KE + sum_j c_j V_j for 4 k-points and
                       4 generalized Hermitian eigensolves of order 896
  * populate (host)    + imago_density_traces: rho(k), Tr(rho S/NP/KE) and
                       Tr(rho V_j) for all 1024 terms
i.e. the integral setup plus one SCF iteration of the secular stage.  The
host potential update (makeSCFPot) is not part of the path built here.

Prints one JSON line.  --gpus N with N > 1 is refused: this build runs on
one GPU (the path would shard by k-point and by potential term).

--stage elecstat measures instead the potential-side electrostatics set-up
(neutralAndNuclearQPot + residualQ of elecStat.f90: imago_elecstat and
imago_residual_q) on the chosen workload, one set-up per step, and prints
its own JSON line of the same shape; --stage xcmesh the exchange-
correlation mesh and its overlap (getECMeshParameters +
makeECMeshAndOverlap of exchCorr.f90: imago_xc_mesh and imago_xc_overlap,
makeinput's default mesh of 100 angular vectors).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SCF wall time to 1e-6 Ha & iterations: learned-guess batched vs plain GPU vs CPU")
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-terms", type=int, default=48,
                   help="EP terms the CPU baseline evaluates (sample, extrapolated)")
    p.add_argument("--small", action="store_true", help="tiny workload for smoke runs")
    p.add_argument("--workload", choices=("si64", "sio2"), default="si64",
                   help="si64: the 64-atom Si supercell (the benchmark line); sio2: the "
                        "216-atom amorphous SiO2 cell, Gamma only")
    p.add_argument("--ep-mode", choices=("direct", "panel"), default="direct",
                   help="direct: EP terms from the stored primitive blocks each iteration "
                        "(IMAGO_EP_DIRECT); panel: all potDim EP panels stored (IMAGO_EP)")
    p.add_argument("--stage", choices=("scf", "step", "elecstat", "xcmesh"), default="scf",
                   help="scf: SCF to convergence (the benchmark line); step: one hot-path "
                        "pass (set-up + one secular/density iteration); elecstat: the "
                        "electrostatics set-up (imago_elecstat + imago_residual_q); xcmesh: "
                        "the exchange-correlation mesh and overlap (imago_xc_mesh + "
                        "imago_xc_overlap)")
    p.add_argument("--max-iter", type=int, default=100,
                   help="LAST_ITERATION of the SCF (makeinputrc's default is 50; 100 here)")
    p.add_argument("--guess", choices=("superposition", "seeded"), default="superposition",
                   help="initial potential of the plain GPU method: the atomic-superposition "
                        "guess (isolated-atom potentials, the reference's scfV.dat) or the "
                        "workload's seeded coefficients")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run only a plain 4 GiB device-to-device copy kernel (scale factor "
                        "for the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return p.parse_args()


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of kernel from its committed rocprofv3 PMC
    summary for this workload (FETCH_SIZE and WRITE_SIZE scaled by a plain
    copy's counts), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "%s_%s_pmc.md" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def calibrate_copy():
    import torch
    n = 1 << 29   # 2^29 doubles = 4 GiB
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                      "launches": 3}), flush=True)


def electrons(system):
    """All-electron count: Z of every atom's element."""
    from imago_amd import configs
    return int(round(sum(configs.Z_OF[system.atom_types[t].element]
                         for t in system.atom_type_of)))


def eigensolve_kernels(system, st, ncols):
    """Rooflines of the eigensolve's stages from their HIP-event times (ms in
    st) at order n, nk matrices, ncols finished eigenvector columns.

    k_hetrd_coop, the Householder tridiagonalisation, is half BLAS2: LAPACK's
    zhetrd reads the trailing matrix once per column for the Hermitian
    matrix-vector product (zlatrd), so its algorithmic traffic is
    sum_{m<n} m^2 elements per matrix (16 B complex, 8 B real), and its
    roofline is HBM bandwidth; its LAWN-41 flops (16/3 n^3 complex, 4/3 n^3
    real) are reported beside it against the FP64 peak.  The dense stages on
    the matrix cores are priced with LAWN-41 flops against the FP64 matrix
    peak: the Cholesky transform as zhegst / dsygst (4 n^3 / n^3), the
    solve X = U^-1 Y as ztrsm (4 n^2 ncols / n^2 ncols), the back-transform
    as zunmtr (8 n^2 ncols / 2 n^2 ncols)."""
    n, nk = system.vale_dim, (1 if system.gamma else system.num_kpoints)
    cx = 1.0 if system.gamma else 4.0
    es = 8.0 if system.gamma else 16.0
    out = {}
    if st.get("k_hetrd_coop", 0.0) > 0.0:
        ms = st["k_hetrd_coop"]
        nbytes = es * (n - 1) * n * (2 * n - 1) / 6.0 * nk
        flop = (4.0 / 3.0 if system.gamma else 16.0 / 3.0) * n ** 3 * nk
        out["k_hetrd_coop"] = {
            "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "ms": ms, "bytes": nbytes,
            "achieved": nbytes / (ms * 1e-3) / 1e9,
            "fp64_flop": flop, "fp64_tflops": flop / (ms * 1e-3) / 1e12,
            "fp64_frac": flop / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS}
    for name, key, flop in (
            ("cholesky_transform", "cholesky_transform", cx * n ** 3 * nk),
            ("vector_solve", "vector_solve", cx * n ** 2 * ncols * nk),
            ("back_transform", "back_transform", 2.0 * cx * n ** 2 * ncols * nk)):
        ms = st.get(key, 0.0)
        if ms > 0.0:
            out[name] = {"bound": "mfma", "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS, "ms": ms,
                         "flop": flop, "achieved": flop / (ms * 1e-3) / 1e12}
    return out


def gpu_step(ctx, system, coeffs, num_states, oracle_populate, ep_kind):
    t0 = time.perf_counter()
    ctx.set_system(system)
    t1 = time.perf_counter()
    stage = {"set_system_lib": ctx.timing(6), "geometry": ctx.timing(7)}
    for kind in ("ol", "ke", "np", ep_kind):
        ctx.assemble(kind)
        stage["assemble_" + kind] = ctx.timing(0)
        if kind.startswith("ep"):
            stage["ep_screen"] = ctx.timing(5)
            stage["ep_store"] = ctx.timing(4)
    t2 = time.perf_counter()
    ev = ctx.secular(coeffs, num_states)
    stage["hbuild"] = ctx.timing(1)
    stage["eigensolve"] = ctx.timing(2)
    stage["k_hetrd_coop"] = ctx.timing(10)
    stage["cholesky_transform"] = ctx.timing(18)
    if ep_kind == "ep_direct":
        stage["k_epd_hsum"] = ctx.timing(8)
    t3 = time.perf_counter()
    occ = oracle_populate(ev, system.kpoint_weights, electrons(system))
    tr, prho = ctx.density_traces(occ)
    stage["traces"] = ctx.timing(3)
    stage["back_transform"] = ctx.timing(20)
    stage["vector_solve"] = ctx.timing(21)
    if ep_kind == "ep_direct":
        stage["k_epd_trace"] = ctx.timing(9)
    t4 = time.perf_counter()
    stage["set_system_wall"] = (t1 - t0) * 1e3
    stage["assemble_wall"] = (t2 - t1) * 1e3
    stage["secular_wall"] = (t3 - t2) * 1e3
    stage["density_wall"] = (t4 - t3) * 1e3
    return (t4 - t0), stage, tr


def populate(eigenvalues, kweights, num_electrons):
    """populateStandard (populate.F90:190-485), spin 1: lowest states first,
    kPointWeight electrons per state, degenerate Fermi-level states share."""
    nk, ns = eigenvalues.shape
    flat = eigenvalues.reshape(-1)
    order = np.argsort(flat, kind="stable")
    w = np.repeat(np.asarray(kweights, dtype=np.float64), ns)
    pop = np.zeros(nk * ns)
    filled = 0.0
    occ_i = len(order) - 1
    for i, idx in enumerate(order):
        pop[idx] = w[idx]
        filled += pop[idx]
        if filled > num_electrons:
            pop[idx] = w[idx] + num_electrons - filled
            filled = float(num_electrons)
        if abs(num_electrons - filled) < 1e-8:
            occ_i = i
            break
    ef = flat[order[occ_i]]
    lo = hi = occ_i
    degen, possible = pop[order[occ_i]], w[order[occ_i]]
    while hi + 1 < len(order) and abs(ef - flat[order[hi + 1]]) <= 1e-8:
        hi += 1
        possible += w[order[hi]]
    while lo > 0 and abs(ef - flat[order[lo - 1]]) <= 1e-8:
        lo -= 1
        possible += w[order[lo]]
        degen += w[order[lo]]
    for i in range(lo, hi + 1):
        pop[order[i]] = degen / possible * w[order[i]]
    return pop.reshape(nk, ns)


def cpu_baseline(system, coeffs, num_states, cpu_terms):
    """Oracle (CPU restatement of the reference path) on a bounded sample of
    the same workload, single thread, extrapolated to one full step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(1)
    except Exception:  # pragma: no cover
        limiter = None
    import oracle
    n = system.vale_dim
    npack = n * (n + 1) // 2
    nk = system.num_kpoints
    pot_dim = system.pot_dim
    t = {}
    t0 = time.perf_counter()
    S = oracle.assemble(system, "ol")
    K = oracle.assemble(system, "ke")
    N = oracle.assemble(system, "np")
    t["ol_ke_np"] = time.perf_counter() - t0
    nt = min(cpu_terms, pot_dim)
    t0 = time.perf_counter()
    V = oracle.assemble(system, "ep", 0, nt, raw=True)
    t["ep_sample"] = time.perf_counter() - t0
    t["ep_full_est"] = t["ep_sample"] * pot_dim / nt
    Sr = oracle.assemble(system, "ol", raw=True)
    Kr = oracle.assemble(system, "ke", raw=True)
    Nr = oracle.assemble(system, "np", raw=True)
    # H(k) build: the reference's readPackedMatrixAccum loop over potDim
    # terms per k (C loop); time nt terms at one k-point and scale
    t0 = time.perf_counter()
    h = oracle.build_hamiltonian_raw(Nr[0], Kr[0], [V[j, 0] for j in range(nt)],
                                     coeffs[:nt])
    t["hbuild_sample"] = time.perf_counter() - t0
    t["hbuild_full_est"] = t["hbuild_sample"] * (pot_dim / nt) * nk
    # generalized eigensolve (LAPACK zhegv, one k-point) and scale by nk
    H = oracle.unpack(h if system.gamma else h[0::2] + 1j * h[1::2], n)
    Sf = oracle.unpack(S[0], n)
    t0 = time.perf_counter()
    w, vec, info = oracle.hegv(H, Sf)
    t["zhegv_one_k"] = time.perf_counter() - t0
    t["zhegv_full_est"] = t["zhegv_one_k"] * nk
    # density matrix (zher over occupied states) + potDim traces
    # (matrixElementMult, C loop over the packed panel)
    t0 = time.perf_counter()
    occ = np.zeros(n)
    occ[: electrons(system) // 2] = 2.0 / nk
    rho = (vec * occ) @ vec.conj().T
    t["rho_one_k"] = time.perf_counter() - t0
    cols = np.concatenate([np.full(c + 1, c) for c in range(n)])
    rows = np.concatenate([np.arange(c + 1) for c in range(n)])
    rp = rho[rows, cols]
    rpr = np.ascontiguousarray(rp.real if system.gamma else np.stack([rp.real, rp.imag], -1).ravel())
    t0 = time.perf_counter()
    for j in range(nt):
        oracle.matrix_element_mult_raw(V[j, 0], rpr, n)
    t["traces_sample"] = time.perf_counter() - t0
    t["traces_full_est"] = (t["traces_sample"] * (pot_dim + 3) / nt + t["rho_one_k"]) * nk
    total = (t["ol_ke_np"] + t["ep_full_est"] + t["hbuild_full_est"] + t["zhegv_full_est"]
             + t["traces_full_est"])
    if limiter is not None:
        limiter.unregister() if hasattr(limiter, "unregister") else None
    del npack
    return total, t, nt


def recip_flop(system, dims):
    """Algorithmic FP64 flops of the reciprocal-space sums of residualQ
    (elecStat.f90:986-990) as summed here: +G and -G add the same term, so
    one cell of each pair is summed (twice weighted); for every site pair
    and summed cell the phase cos(G.(r_i - r_j)) as cos_i cos_j + sin_i
    sin_j (3 flops) and one multiply-add per alpha of site i (2 flops)."""
    keys = {tuple(d) for d in dims[1:]}
    pairs = sum(1 for d in dims[1:] if tuple(-d) in keys) // 2
    nsum = len(dims) - 1 - pairs
    nsites = len(system.pot_site_pos)
    nalpha_i = sum(len(p.alphas) * len(p.sites) for p in system.pot_types)
    return nsum * (3.0 * nsites * nsites + 2.0 * nalpha_i * nsites)


def elecstat_cpu_baseline(system, budget_s=10.0):
    """Oracle neutralAndNuclearQPot in full and residualQ over its outer-loop
    sites in order until about budget_s seconds have passed, scaled to all
    sites; one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t0 = time.perf_counter()
    oracle.elecstat(system)
    t_es = time.perf_counter() - t0
    nsites = len(system.pot_site_pos)
    dims, sizes = system.recip_cells()
    ns, spent = 0, 0.0
    while ns < nsites and spent < budget_s:
        nxt = min(nsites, ns + max(1, nsites // 64))
        t0 = time.perf_counter()
        oracle.residual_q(system, recip_cells=(dims, sizes), sites=(ns, nxt))
        spent += time.perf_counter() - t0
        ns = nxt
    return t_es, spent * nsites / ns, ns


def bench_elecstat(args, system, ctx, log):
    dims, sizes = system.recip_cells()
    ncr = len(sizes)
    P = system.pot_dim
    run = lambda: (ctx.elecstat(), ctx.residual_q(recip_cells=(dims, sizes)))
    for i in range(args.warmup):
        run()
    ms_es, ms_rq, ms_rc, walls = [], [], [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        run()
        walls.append(time.perf_counter() - t0)
        ms_es.append(ctx.timing(11))
        ms_rq.append(ctx.timing(12))
        ms_rc.append(ctx.timing(13))
        log("step %d: elecstat %.3f ms, residualQ %.3f ms (reciprocal sum %.3f ms), wall with "
            "host copies %.1f ms" % (i, ms_es[-1], ms_rq[-1], ms_rc[-1], walls[-1] * 1e3))
    es, rq, rc = (float(np.mean(x)) for x in (ms_es, ms_rq, ms_rc))
    flop = recip_flop(system, dims)
    achieved = flop / (rc * 1e-3) / 1e12
    roofline = {"kernel": "k_es_recip", "bound": "mfma", "achieved": round(achieved, 3),
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 5), "traffic": None,
                "flop_per_launch": flop}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline leg (oracle)")
        t_es, t_rq, ns = elecstat_cpu_baseline(system)
        cpu = {"value": round((t_es + t_rq) * 1e3, 1), "unit": "ms per set-up", "cores": 1,
               "kind": "port",
               "sample": ("oracle on %s: neutralAndNuclearQPot in full (%.2f s), residualQ "
                          "over %d of %d outer-loop sites (%.2f s scaled to all)")
                         % (system.name, t_es, ns, len(system.pot_site_pos), t_rq),
               "host": platform.processor() or platform.machine()}
    line = {
        "metric": "potential-side electrostatics set-up time (neutralAndNuclearQPot + "
                  "residualQ)",
        "value": round(es + rq, 4),
        "unit": "ms per set-up (kernels, results resident in HBM)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(es + rq, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded %s)" % system.name,
        "config": {"workload": "%s: %d potential sites, potDim=%d, %d real-space and %d "
                               "reciprocal cells, electroConvg 1e-16"
                               % (system.name, len(system.pot_site_pos), P,
                                  len(system.cell_sizes), ncr),
                   "parallelism": "single GPU"},
        "stages_ms": {"elecstat": round(es, 4), "residual_q": round(rq, 4),
                      "k_es_recip": round(rc, 4),
                      "wall_with_host_copies": round(float(np.mean(walls)) * 1e3, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def xc_cpu_baseline(system, budget_s=10.0):
    """Oracle getECMeshParameters in full and the overlap sum over its sites
    in order until about budget_s seconds have passed, scaled to all sites;
    one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    t0 = time.perf_counter()
    om = oracle.XcOracleMesh(system)
    t_mesh = time.perf_counter() - t0
    nsites = len(system.pot_site_pos)
    ns, spent = 0, 0.0
    while ns < nsites and spent < budget_s:
        t0 = time.perf_counter()
        om.overlap(sites=(ns, ns + 1))
        spent += time.perf_counter() - t0
        ns += 1
    return t_mesh, spent * nsites / ns, ns, om


def bench_xcmesh(args, system, ctx, log):
    P = system.pot_dim
    for i in range(args.warmup):
        ctx.xc_mesh()
        ctx.xc_overlap()
    walls_mesh, ms_ov, walls_ov, ms_gemm, ms_pot = [], [], [], [], []
    # charge coefficients for the potential terms: a neutral-atom-like
    # profile on every potential type
    rho_fit = np.concatenate([0.05 * np.exp(-0.3 * np.log(p.alphas)) for p in system.pot_types])
    for i in range(args.steps):
        t0 = time.perf_counter()
        mx = ctx.xc_mesh()
        walls_mesh.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        ctx.xc_overlap()
        walls_ov.append(time.perf_counter() - t0)
        ms_ov.append(ctx.timing(14))
        ms_gemm.append(ctx.timing(15))
        ctx.xc_potential(rho_fit)
        ms_pot.append(ctx.timing(16))
        log("step %d: mesh %.1f ms (wall), overlap %.3f ms (device), %.1f ms with the copy"
            % (i, walls_mesh[-1] * 1e3, ms_ov[-1], walls_ov[-1] * 1e3))
    mesh_ms = float(np.mean(walls_mesh)) * 1e3
    ov_ms = float(np.mean(ms_ov))
    gemm_ms = float(np.mean(ms_gemm))
    npoints, gflop = ctx.xc_counts()
    # the overlap's products B_b^T B_b, one strided-batched rocBLAS dgemm per
    # batch of sites: its flops as executed (full squares, zero padding to
    # the batch's largest row and point counts) over its HIP-event time
    achieved = gflop / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    roofline = {"kernel": "rocblas_dgemm_strided_batched (B_b^T B_b)", "bound": "mfma",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 5), "traffic": None,
                "flop_per_launch": gflop}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline leg (oracle)")
        t_mesh, t_ov, ns, _ = xc_cpu_baseline(system)
        cpu = {"value": round((t_mesh + t_ov) * 1e3, 1), "unit": "ms per set-up", "cores": 1,
               "kind": "port",
               "sample": ("oracle on %s: getECMeshParameters in full (%.2f s), the overlap sum "
                          "over %d of %d sites (%.2f s scaled to all)")
                         % (system.name, t_mesh, ns, len(system.pot_site_pos), t_ov),
               "host": platform.processor() or platform.machine()}
    line = {
        "metric": "exchange-correlation mesh set-up time (getECMeshParameters + "
                  "exchCorrOverlap)",
        "value": round(mesh_ms + ov_ms, 3),
        "unit": "ms per set-up (mesh call wall time + overlap device time, result in HBM)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mesh_ms + ov_ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded %s)" % system.name,
        "config": {"workload": "%s: %d potential sites, potDim=%d, 100 angular vectors, "
                               "%d ray points per site at most"
                               % (system.name, len(system.pot_site_pos), P, mx),
                   "parallelism": "single GPU"},
        "stages_ms": {"xc_mesh_wall": round(mesh_ms, 3), "xc_overlap": round(ov_ms, 3),
                      "xc_overlap_gemm": round(gemm_ms, 3),
                      "xc_potential": round(float(np.mean(ms_pot)), 3),
                      "xc_overlap_wall_with_copy": round(float(np.mean(walls_ov)) * 1e3, 3)},
        "ray_points": npoints,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def host_info():
    """CPU model, core count, compiler and LAPACK of the host (for the CPU
    baseline's record)."""
    info = {"machine": platform.machine()}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        info["cores_available"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["cores_available"] = os.cpu_count()
    try:
        import subprocess
        info["oracle_cc"] = subprocess.run(["gcc", "--version"], capture_output=True,
                                           text=True).stdout.splitlines()[0]
    except Exception:
        pass
    info["oracle_cflags"] = "-O2 -fPIC -std=c11 -ffp-contract=off -fopenmp"
    try:
        from threadpoolctl import threadpool_info
        libs = [(d.get("internal_api"), d.get("version"), d.get("num_threads"))
                for d in threadpool_info() if d.get("user_api") == "blas"]
        info["lapack"] = "scipy %s, %s" % (__import__("scipy").__version__,
                                           ", ".join("%s %s" % (a, b) for a, b, _ in libs))
    except Exception:
        pass
    return info


def scf_cpu_baseline(system, iterations, first_inputs, threads, log, ep_chunk=64):
    """The oracle's SCF on this workload, timed on this host with the given
    thread count (assembly loops over atoms with OpenMP, LAPACK threads):
    set-up in full (OL, KE, NP, every EP term in chunks of ep_chunk terms,
    the electrostatics and the exchange-correlation mesh of makeSCFPot),
    one iteration's secular equation in full (zhegv at every k-point), its
    H build and traces over one chunk of EP terms scaled to all terms, and
    one makeSCFPot call on the GPU run's first-iteration traces; the SCF
    time is set-up + iterations x iteration (the GPU run's iteration count)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import scf as oscf
    from threadpoolctl import threadpool_limits
    L = oracle.lib()
    L.orc_set_threads(threads)
    t = {}
    with threadpool_limits(threads):
        n, nk, P = system.vale_dim, system.num_kpoints, system.pot_dim
        t0 = time.perf_counter()
        S = oracle.assemble(system, "ol", raw=True)       # [nk][packed, (re, im)]
        K = oracle.assemble(system, "ke", raw=True)
        N = oracle.assemble(system, "np", raw=True)
        t["ol_ke_np"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        for j0 in range(0, P, ep_chunk):
            V = oracle.assemble(system, "ep", j0, min(P, j0 + ep_chunk), raw=True)
            if j0 == 0:
                V0 = V
            del V
        t["ep_all_terms"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        pot = oscf.ScfPotential(system)
        t["scfpot_setup"] = time.perf_counter() - t0
        nc = V0.shape[0]
        coeffs = first_inputs[2]
        t0 = time.perf_counter()
        H = [oracle.build_hamiltonian_raw(N[k], K[k], [V0[j, k] for j in range(nc)], coeffs[:nc])
             for k in range(nk)]
        t["hbuild_chunk"] = time.perf_counter() - t0

        def full(a):
            return oracle.unpack(a if system.gamma else a[0::2] + 1j * a[1::2], n)
        Hf = [full(H[k]) for k in range(nk)]
        Sf = [full(S[k]) for k in range(nk)]
        t0 = time.perf_counter()
        vecs, evals = [], []
        for k in range(nk):
            w, v, info = oracle.hegv(Hf[k], Sf[k])
            evals.append(w)
            vecs.append(v)
        t["zhegv_all_k"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        occ = oracle.populate(np.array(evals), system.kpoint_weights, oscf.num_electrons(system))
        t["populate"] = time.perf_counter() - t0
        cols = np.concatenate([np.full(c + 1, c) for c in range(n)])
        rows = np.concatenate([np.arange(c + 1) for c in range(n)])
        t0 = time.perf_counter()
        for k in range(nk):
            X = vecs[k]
            rho = (X * occ[k]) @ X.conj().T
            rp = rho[rows, cols]
            rpr = np.ascontiguousarray(rp.real if system.gamma
                                       else np.stack([rp.real, rp.imag], -1).ravel())
            dim1 = 1 if system.gamma else 2
            for M in (S[k], N[k], K[k]):
                oracle.matrix_element_mult_raw(M, rpr, n, dim1)
            for j in range(nc):
                oracle.matrix_element_mult_raw(V0[j, k], rpr, n, dim1)
        t["traces_chunk"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        pot.update(first_inputs[0], first_inputs[1], first_inputs[2])
        t["make_scf_pot"] = time.perf_counter() - t0
    L.orc_set_threads(1)
    scale = P / nc
    setup = t["ol_ke_np"] + t["ep_all_terms"] + t["scfpot_setup"]
    per_iter = (t["hbuild_chunk"] * scale + t["zhegv_all_k"] + t["populate"]
                + t["traces_chunk"] * scale + t["make_scf_pot"])
    total = setup + iterations * per_iter
    log("cpu baseline (%d thread(s)): set-up %.2f s, %.3f s per iteration, SCF %.1f s"
        % (threads, setup, per_iter, total))
    return total, setup, per_iter, t


def bench_scf(args, system, coeffs, ctx, log):
    """SCF wall time to convergence: imago_set_system, the four integral
    kinds, makeSCFPot's set-up, then secular / populate / traces /
    makeSCFPot iterations until testableDelta < 1e-4 and |dE| < 1e-6 Ha."""
    from imago_amd import guess, scf
    ep_kind = "ep_direct" if args.ep_mode == "direct" else "ep"
    guess_info = {"kind": args.guess}
    if args.guess == "superposition":
        # the atomic database entries: isolated-atom SCFs on the GPU, made
        # once per element like the reference's atomicPDB files (not timed)
        db = guess.AtomicPotentials()
        coeffs = guess.superposition_guess(ctx, system, db)
        guess_info.update(atom_scf_s=round(db.seconds, 3),
                          atom_iterations=list(db.iterations.values()))
        log("superposition guess: %d isolated-atom SCF(s), %s iterations, %.2f s"
            % (len(db.entries), list(db.iterations.values()), db.seconds))
    def leg(name, start):
        """warmup + steps SCFs from start() (called inside the timed region)."""
        out = []
        for i in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            c0 = start()
            t_guess = time.perf_counter() - t0
            res = scf.run(ctx, system, c0, max_iter=args.max_iter, ep_kind=ep_kind)
            res.total_s += t_guess
            res.guess_s = t_guess
            tag = "warmup" if i < args.warmup else "step"
            log("%s %s %d: %d iterations, converged %s, E = %.10f Ha, guess %.1f ms, set-up "
                "%.1f ms, loop %.1f ms, total %.1f ms"
                % (name, tag, i, res.iterations, res.converged, res.total_energy, t_guess * 1e3,
                   res.setup_s * 1e3, res.loop_s * 1e3, res.total_s * 1e3))
            if i >= args.warmup:
                out.append(res)
        return out

    plain = leg("plain", lambda: coeffs)
    runs = plain
    learned_runs = None
    from imago_amd import learned
    if args.guess == "superposition" and os.path.exists(learned.MODEL_PATH):
        model = learned.LearnedGuess.load()
        learned_runs = leg("learned", lambda: model.predict(ctx, system, coeffs))
        runs = learned_runs
    last = runs[-1]
    for r in last.records:
        log("  iter %2d  E %.10f  dE %s  delta %.3e  %.2f ms" % (
            r["iteration"], r["total_energy"],
            "-" if r["energy_change"] is None else "%.2e" % r["energy_change"],
            r["testable_delta"], r["wall_s"] * 1e3))
    wall = float(np.mean([r.total_s for r in runs]))
    wall_plain = float(np.mean([r.total_s for r in plain]))
    n, nk = system.vale_dim, system.num_kpoints
    # per-iteration stages of the last iteration (HIP events)
    st = {"hbuild": ctx.timing(1), "eigensolve": ctx.timing(2),
          "cholesky_transform": ctx.timing(18), "k_hetrd_coop": ctx.timing(10),
          "tridiagonal_dc": ctx.timing(19), "wy_factors": ctx.timing(22),
          "back_transform": ctx.timing(20), "vector_solve": ctx.timing(21),
          "traces": ctx.timing(3), "xc_potential": ctx.timing(16)}
    # occupied columns the traces finish each iteration (spin 1, no smearing)
    kernels = eigensolve_kernels(system, st, (electrons(system) + 1) // 2)
    roofline = None
    if "k_hetrd_coop" in kernels:   # absent when rocSOLVER's eigensolver ran
        dom = "k_hetrd_coop"
        d = kernels[dom]
        roofline = {"kernel": dom, "bound": d["bound"], "achieved": round(d["achieved"], 3),
                    "peak": d["peak"], "unit": d["unit"],
                    "frac": round(d["achieved"] / d["peak"], 5),
                    "traffic": profiled_traffic(dom, args.workload),
                    # the same launch priced by its LAWN-41 flops
                    "fp64_tflops": round(d["fp64_tflops"], 3),
                    "fp64_frac": round(d["fp64_frac"], 5)}
    line = {
        "metric": METRIC,
        "value": round(wall, 6),
        "unit": "s per SCF to 1e-6 Ha (%s; set-up + %d iterations)" % (
            "learned initial potential, the guess's inference included" if learned_runs
            else "plain GPU method, %s initial potential" % args.guess, last.iterations),
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded %s, random-normalised contractions)" % system.name,
        "config": {"workload": "%s: %d atoms, n=%d, %d k-point(s)%s, potDim=%d, SCF to "
                               "testableDelta < 1e-4 and |dE| < 1e-6 Ha"
                               % (system.name, system.num_atoms, n, nk,
                                  " (Gamma)" if system.gamma else "", system.pot_dim),
                   "ep_mode": args.ep_mode, "parallelism": "single GPU"},
        "scf_plain_gpu": {"seconds": round(wall_plain, 6), "iterations": plain[-1].iterations,
                          "converged": plain[-1].converged,
                          "total_energy": plain[-1].total_energy, "guess": guess_info},
        "scf_learned_gpu": None if not learned_runs else {
            "seconds": round(wall, 6), "iterations": last.iterations,
            "converged": last.converged, "total_energy": last.total_energy,
            "guess_ms": round(float(np.mean([r.guess_s for r in learned_runs])) * 1e3, 3),
            "iteration_reduction": round(1.0 - last.iterations / plain[-1].iterations, 4),
            "energy_difference": last.total_energy - plain[-1].total_energy},
        "iteration_wall_ms": {k: round(1e3 * float(np.mean([r["stage_s"][k] for r in
                                                             last.records[1:]])), 3)
                              for k in last.records[-1]["stage_s"]},
        "scf": {"iterations": last.iterations, "converged": last.converged, "guess": guess_info,
                "total_energy": last.total_energy,
                "setup_ms": round(last.setup_s * 1e3, 3),
                "loop_ms": round(last.loop_s * 1e3, 3),
                "ms_per_iteration": round(last.loop_s * 1e3 / max(1, last.iterations), 3)},
        "stages_ms": {k: round(v, 3) for k, v in st.items()},
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in kernels.items()},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline leg: the oracle's SCF on %s" % system.name)
        info = host_info()
        cores = int(info.get("cores_available") or 1)
        base = plain[-1]     # the reference's CPU path starts from the superposition
        c1, s1, p1, t1 = scf_cpu_baseline(system, base.iterations, base.first_inputs, 1, log)
        call, sall, pall, tall = scf_cpu_baseline(system, base.iterations, base.first_inputs,
                                                  cores, log)
        line["cpu_baseline"] = {
            "value": round(c1, 3), "unit": "s per SCF to 1e-6 Ha (%d iterations)"
                                           % base.iterations,
            "cores": 1, "kind": "port",
            "sample": ("oracle (C restatement + SciPy LAPACK) on %s: set-up in full (OL/KE/NP, "
                       "all %d EP terms, electrostatics, XC mesh), one iteration's zhegv at all "
                       "%d k-points, H build and traces over %d of %d EP terms scaled, one "
                       "makeSCFPot; SCF = set-up + %d x iteration"
                       % (system.name, system.pot_dim, system.num_kpoints, 64,
                          system.pot_dim, base.iterations)),
            "setup_s": round(s1, 3), "per_iteration_s": round(p1, 4),
            "parts_s": {k: round(v, 4) for k, v in t1.items()},
            "all_cores": {"value": round(call, 3), "cores": cores, "setup_s": round(sall, 3),
                          "per_iteration_s": round(pall, 4),
                          "parts_s": {k: round(v, 4) for k, v in tall.items()}},
            "host": info,
        }
        line["vs_cpu"] = {"speedup_1core": round(c1 / wall, 1),
                          "speedup_all_cores": round(call / wall, 1),
                          "plain_gpu_speedup_1core": round(c1 / wall_plain, 1),
                          "plain_gpu_speedup_all_cores": round(call / wall_plain, 1)}
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(it would shard the integrals by atom pair / potential term and the "
              "eigensolves by k-point).", file=sys.stderr)
        sys.exit(2)
    from imago_amd import configs
    from imago_amd.lib import Context
    if args.small:
        system = configs.si64_supercell(reps=(1, 1, 1), cutoff_bohr=16.0)
    elif args.workload == "sio2":
        system = configs.sio2_amorphous()
    else:
        system = configs.si64_supercell()
    coeffs = configs.pot_coefficients(system, seed=7)
    num_states = system.vale_dim
    ep_kind = "ep_direct" if args.ep_mode == "direct" else "ep"
    # the context is closed before the interpreter exits, so its streams,
    # library handles and device buffers never go through exit handlers
    with Context(0) as ctx:
        run(args, system, coeffs, num_states, ep_kind, ctx)


def run(args, system, coeffs, num_states, ep_kind, ctx):
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    log("system %s n=%d nk=%d potDim=%d" % (system.name, system.vale_dim, system.num_kpoints,
                                          system.pot_dim))
    if args.stage == "elecstat":
        ctx.set_system(system)
        bench_elecstat(args, system, ctx, log)
        return
    if args.stage == "xcmesh":
        ctx.set_system(system)
        bench_xcmesh(args, system, ctx, log)
        return
    if args.stage == "scf":
        bench_scf(args, system, coeffs, ctx, log)
        return
    for i in range(args.warmup):
        dt, st, _ = gpu_step(ctx, system, coeffs, num_states, populate, ep_kind)
        log("warmup %d: %.3f s %s" % (i, dt, {k: round(v, 2) for k, v in st.items()}))
    times, stages = [], []
    t_all = time.perf_counter()
    for i in range(args.steps):
        dt, st, tr = gpu_step(ctx, system, coeffs, num_states, populate, ep_kind)
        times.append(dt)
        stages.append(st)
        log("step %d: %.3f s" % (i, dt))
    wall = time.perf_counter() - t_all
    step_s = wall / args.steps
    st = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    n = system.vale_dim
    nk = system.num_kpoints
    ns = np.array([len(system.atom_types[t].states()) for t in system.atom_type_of])
    nbe = int((np.sum(ns) ** 2 + np.sum(ns ** 2)) // 2)   # block-layout entries per k
    per = (1 if system.gamma else 2) * nbe * nk  # doubles of one panel
    ntup, nitems, bsz, npart = ctx.ep_counts()
    kernels = {}
    if ep_kind == "ep_direct":
        # every primitive block and tuple read once, plus each item's sorted
        # index and tuple index (4 + 4 bytes)
        item_bytes = nitems * bsz * 8.0 + ntup * 24.0 + nitems * 8.0
        kernels["k_epd_hsum"] = {
            "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "ms": st["k_epd_hsum"],
            # ... and the runs' partial blocks written
            "bytes": item_bytes + npart * 8.0}
        kernels["k_epd_trace"] = {
            "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "ms": st["k_epd_trace"],
            # ... the density (folded per image for k-points) read once, one
            # value per item written
            "bytes": item_bytes + per * 8.0 + nitems * 8.0}
    else:
        kernels["k_ep_store"] = {
            "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "ms": st["ep_store"],
            "bytes": system.pot_dim * per * 8.0 + nitems * bsz * 8.0 + ntup * 24.0}
        kernels["k_hbuild"] = {
            "bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "ms": st["hbuild"],
            "bytes": (system.pot_dim + 3) * per * 8.0}
    for v in kernels.values():
        v["achieved"] = v["bytes"] / (v["ms"] * 1e-3) / 1e9
    kernels.update(eigensolve_kernels(system, st, (electrons(system) + 1) // 2))
    # the roofline line is the kernel of this library that takes the most
    # time per step; the eigensolve stage (the Cholesky transform, the
    # reduction, the divide and conquer; its back-transform of the occupied
    # columns runs with the traces) is reported in stages_ms
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
    roofline = {"kernel": dom, "bound": d["bound"], "achieved": round(d["achieved"], 3),
                "peak": d["peak"], "unit": d["unit"], "frac": round(d["achieved"] / d["peak"], 5),
                "traffic": profiled_traffic(dom, args.workload)}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline leg (%d EP terms)" % args.cpu_terms)
        total, parts, nt = cpu_baseline(system, coeffs, num_states, args.cpu_terms)
        cpu = {"value": round(total, 3), "unit": "s per step (extrapolated)", "cores": 1,
               "kind": "port",
               "sample": ("oracle on %s: OL/KE/NP in full, %d of %d EP terms, H build of %d "
                          "terms at 1 k-point, zhegv at 1 of %d k-points, %d traces at 1 "
                          "k-point; each part scaled to the full step (%s)") %
                         (system.name, nt, system.pot_dim, nt, system.num_kpoints, nt,
                          ", ".join("%s=%.3fs" % kv for kv in parts.items())),
               "host": platform.processor() or platform.machine()}
    line = {
        "metric": METRIC,
        "value": round(step_s, 6),
        "unit": "s per step (integral setup + 1 SCF secular/density iteration, %s n=%d nk=%d)"
                % (system.name, n, nk),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded %s, random-normalised contractions)" % system.name,
        "config": {"workload": "%s: %d atoms, n=%d, %d k-point(s)%s, potDim=%d, cutoff 16 bohr, "
                               "one hot-path pass per step"
                               % (system.name, system.num_atoms, n, nk,
                                  " (Gamma)" if system.gamma else "", system.pot_dim),
                   "ep_mode": args.ep_mode, "parallelism": "single GPU"},
        "stages_ms": {k: round(v, 3) for k, v in st.items()},
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in kernels.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "trace_check": {"tr_rho_S": float(tr[0]), "electrons": electrons(system)},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
