"""Benchmark: domain-wall propagator solve on one MI355X.

Workload (the metric's configuration): 24^3 x 64 lattice, links
U = exp(i 0.3 H) (seeded, antiperiodic in time), Moebius Ls = 12, M5 = 1.8,
b = 1.5, c = 0.5, mass = 0.01; a 12-column point-source propagator solved
to a 1e-8 FP64 true residual of the preconditioned odd system.  One step is
one full propagator solve with the mixed-precision multi-RHS solver (the 12
columns in autotuned multi-RHS blocks; Krylov vectors stored in FP16 with a
per-site scale, FP16 links and FP16 products with FP32 sums in the hop
phase, FP32 s-matrix and CG arithmetic, FP64 reliable updates by default);
the FP64 solver is timed once on the same inputs.
Inputs (gauge field, source propagator) are resident in device memory
before the timed region; solutions stay in device memory.

Prints one JSON line (value = mixed-precision time-to-solution in seconds).
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CG time-to-1e-8 & dslash HBM GB/s, 24^3x64 Ls=12: mixed multi-RHS vs FP64"
PEAK_HBM_GBS = 8000.0

CONFIGS = {
    "B": dict(site=(24, 24, 24, 64), ls=12, sigma=0.3, mass=0.01, m5=1.8, scale=2.0, seed=2024),
    # the production action's fifth dimension (Ls 24, mobius_scale 4, as in the RBC/UKQCD 24D
    # ensembles) on the same lattice: kernel timings and profiles (not a bench line)
    "B24": dict(site=(24, 24, 24, 64), ls=24, sigma=0.3, mass=0.01, m5=1.8, scale=4.0, seed=2024),
    # small configuration for quick local runs (not a bench line)
    "S": dict(site=(8, 8, 8, 16), ls=12, sigma=0.3, mass=0.01, m5=1.8, scale=2.0, seed=2024),
    # batched multi-RHS: 16 Gaussian sources x 4 masses (its own JSON line)
    "C": dict(site=(16, 16, 16, 32), ls=8, sigma=0.3, masses=(0.005, 0.01, 0.03, 0.1), nsrc=16, m5=1.8,
              scale=2.0, seed=3031),
    # large single-GPU share: 12-column propagator and one force evaluation (its own JSON line)
    "E": dict(site=(32, 32, 32, 64), ls=16, sigma=0.25, mass=0.004, m5=1.8, scale=2.0, seed=4041),
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def hop_bytes(site, ls, real_bytes, link_reals=18):
    """Algorithmic bytes of one hop application (one output parity):
    read the input half spinor once, write the output once, read every link
    once: N*24*w*2 + V*4*r*w with N = V/2*Ls."""
    V = 1
    for n in site:
        V *= n
    N = V // 2 * ls
    return N * 24 * real_bytes * 2 + V * 4 * link_reals * real_bytes


def hop_flops(site, ls):
    V = 1
    for n in site:
        V *= n
    return 1320 * (V // 2) * ls


def hop_m5s_bytes(site, ls, ncol, site_bytes, link_real_bytes):
    """Algorithmic bytes of one multi-RHS hop + s-matrix launch over ncol
    columns: read and write every column's 5D half field once (site_bytes
    per 5D site and field: 52 = 48 FP16 data + 4 scale, or 96 FP32), read
    the 8 links (18 reals each, FP16 or FP32) of every output site once."""
    V = 1
    for n in site:
        V *= n
    N = V // 2 * ls
    return ncol * N * site_bytes * 2 + V * 4 * 18 * link_real_bytes


def read_traffic(kernel_tag):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE, calibrated against a copy kernel) and kept in profiles/."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel_tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline(cfg, U, iters_fp64, sample_hermops):
    """The oracle (C++ restatement of the reference CPU path, FP64, OpenMP)
    on this host: time `sample_hermops` applications of Mpc^+ Mpc at the
    full lattice, then extrapolate the 12-column time-to-solution with the
    FP64 GPU iteration count (an extrapolation: the full CPU solve would take hours)."""
    import numpy as np

    from oracle import oracle as O

    site, ls = cfg["site"], cfg["ls"]
    orc = O.OracleInverter(site, U, mass=cfg["mass"], ls=ls, m5=cfg["m5"], mobius_scale=cfg["scale"])
    V = int(np.prod(site))
    x = np.ones((V // 2, ls, 12), dtype=np.complex128)
    t0 = time.perf_counter()
    for _ in range(sample_hermops):
        x = orc.apply_half("hermop_sym2", x, 1)
    dt = (time.perf_counter() - t0) / sample_hermops
    threads = O.num_threads()
    return dict(
        value=dt * iters_fp64,
        unit="s (12-column time-to-solution, extrapolated: FP64 GPU iteration count x oracle hermop time)",
        cores=threads,
        kind="port",
        sample=f"{sample_hermops} applications of Mpc^+Mpc (multiply_hermop_sym2) at 24^3x64 Ls=12, "
               f"{dt:.3f} s each on {threads} threads ({platform.processor() or platform.machine()})",
        hermop_seconds=dt,
    )


def bench_config_c(args):
    """64 systems (16 seeded Gaussian 5D sources x 4 masses) at 16^3x32 Ls8 in
    one batched call with the mixed16 solver (autotuned multi-RHS tiles with
    per-column masses and retirement), then the same systems with the FP64
    solver one by one; prints one JSON line."""
    import numpy as np
    import torch

    import qlattice_amd as q
    from qlattice_amd import _lib
    from qlattice_amd.gauge import rand_gauge_links, twist_boundary_at_boundary

    cfg = CONFIGS["C"]
    site, ls = cfg["site"], cfg["ls"]
    V = int(np.prod(site))
    U = rand_gauge_links(site, cfg["sigma"], cfg["seed"])
    twist_boundary_at_boundary(U, site, -0.5, 3)
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    invs = []
    for m in cfg["masses"]:
        inv = q.InverterDomainWall(gf=gf, fa=q.FermionAction(mass=m, ls=ls, m5=cfg["m5"], mobius_scale=cfg["scale"]),
                                   solver="mixed16", reliable_delta=args.reliable_delta)
        inv.set_stop_rsd(1e-8)
        invs.append(inv)
    rng = np.random.default_rng(cfg["seed"] + 1)
    srcs = [torch.from_numpy(rng.standard_normal((V, ls, 12)) + 1j * rng.standard_normal((V, ls, 12))).cuda()
            for _ in range(cfg["nsrc"])]
    pairs = [(mi, k) for k in range(cfg["nsrc"]) for mi in range(len(invs))]
    sols = [torch.zeros_like(srcs[0]) for _ in pairs]
    L = _lib.load()
    n = len(pairs)

    def run(solver):
        for inv in invs:
            inv.set_solver(solver, args.reliable_delta)
        hv = [invs[mi].cdata for mi, _ in pairs]
        sp = [_lib.dev_ptr(srcs[k]) for _, k in pairs]
        op = [_lib.dev_ptr(x) for x in sols]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it = np.array(L.invert_5d_batch(hv, op, sp, device=True), dtype=np.int64)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, it, invs[0].stats()

    log("config C: warm-up batch (tile autotuning)")
    run("mixed16")
    t16, it16, st16 = run("mixed16")
    log(f"mixed16 batch: {t16:.2f} s, tile {st16['rhs_tile']}, iterations per mass "
        f"{[int(np.mean(it16[[j for j, (mi, _) in enumerate(pairs) if mi == m]])) for m in range(len(invs))]}, "
        f"true residual {st16['true_residual']:.2e}")
    t64, diff = None, None
    if not args.skip_fp64:
        x16 = [x.clone() for x in sols]
        t64, it64, st64 = run("fp64")
        diff = max((torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item() for a, b in zip(x16, sols))
        log(f"fp64 one by one: {t64:.2f} s, max rel diff mixed16 vs fp64 {diff:.2e}")
    line = {
        "metric": "batched multi-RHS time-to-1e-8, 16^3x32 Ls8, 16 sources x 4 masses",
        "value": t16, "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "fp16 hop products with fp32 sums, fp32 CG, fp64 reliable updates" if st16["hop_arith_bits"] == 16
        else "fp32",
        "data": "synthetic: seeded exp(i*0.3*H) links, antiperiodic t, Gaussian 5D sources",
        "config": {"workload": "64 systems (16 sources x masses 0.005, 0.01, 0.03, 0.1) in one batched call, "
                               "Moebius b=1.5 c=0.5, M5=1.8, Ls=8, stop 1e-8 (FP64 true residual)",
                   "rhs_tile": st16["rhs_tile"], "sites_per_block": st16["sites_per_block"]},
        "iters_per_mass": {str(cfg["masses"][m]): int(np.mean(it16[[j for j, (mi, _) in enumerate(pairs) if mi == m]]))
                           for m in range(len(invs))},
        "fp64_seconds": t64, "speedup_vs_fp64": None if t64 is None else t64 / t16, "max_rel_diff_vs_fp64": diff,
        "true_residual": st16["true_residual"], "true_residual_5d": st16["true_residual_5d"],
    }
    print(json.dumps(line), flush=True)


def bench_config_e(args):
    """32^3x64 Ls16 mass 0.004: 12-column point-source propagator with the
    mixed16 multi-RHS solver, then one fermion-force evaluation (FP64 CG to
    1e-10 plus the adjoint kernel) on a seeded pseudofermion; one JSON line."""
    import numpy as np
    import torch

    import qlattice_amd as q
    from qlattice_amd import _lib
    from qlattice_amd.gauge import rand_gauge_links, twist_boundary_at_boundary

    cfg = CONFIGS["E"]
    site, ls = cfg["site"], cfg["ls"]
    V = int(np.prod(site))
    log("config E: generating gauge field")
    U = rand_gauge_links(site, cfg["sigma"], cfg["seed"])
    twist_boundary_at_boundary(U, site, -0.5, 3)
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=cfg["mass"], ls=ls, m5=cfg["m5"], mobius_scale=cfg["scale"])
    inv = q.InverterDomainWall(gf=gf, fa=fa, solver="mixed16", reliable_delta=args.reliable_delta)
    inv.set_stop_rsd(1e-8)
    inv.set_max_num_iter(1000)
    del U, gf
    L = _lib.load()
    t_prop, st = None, None
    if not args.force_only:
        src_h = np.zeros((V, 12, 12), dtype=np.complex128)
        src_h[0] = np.eye(12)
        src = torch.from_numpy(src_h).cuda()
        del src_h
        sol = torch.zeros_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.invert_prop(inv.cdata, _lib.dev_ptr(sol), _lib.dev_ptr(src), device=True)
        torch.cuda.synchronize()
        t_prop = time.perf_counter() - t0
        st = inv.stats()
        log(f"mixed16 propagator: {t_prop:.2f} s, iters {st['total_iter']}, reliable {st['reliable_updates']}, "
            f"true residual {st['true_residual']:.2e}, tile {st['rhs_tile']}")
        del src, sol
        torch.cuda.empty_cache()
    if args.skip_force:
        print(json.dumps({"propagator": {"seconds": t_prop, "iters": st["total_iter"], "rhs_tile": st["rhs_tile"],
                                         "sites_per_block": st["sites_per_block"],
                                         "true_residual": st["true_residual"]}}), flush=True)
        return
    rng = np.random.default_rng(cfg["seed"] + 2)
    phi = torch.from_numpy(rng.standard_normal((V // 2, ls, 12)) + 1j * rng.standard_normal((V // 2, ls, 12))).cuda()
    force = torch.zeros((V, 4, 3, 3), dtype=torch.complex128, device="cuda")
    inv.set_solver("fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S, S_lo, itf = L.fermion_force(inv.cdata, _lib.dev_ptr(phi), args.force_stop_rsd, _lib.dev_ptr(force), 0,
                                   device=True)
    torch.cuda.synchronize()
    t_force = time.perf_counter() - t0
    herm = (force + force.conj().transpose(-1, -2)).abs().max().item()
    trace = force.diagonal(dim1=-2, dim2=-1).sum(-1).abs().max().item()
    log(f"force: {t_force:.2f} s, CG iters {itf}, S {S:.6e}, anti-hermiticity {herm:.1e}, trace {trace:.1e}")
    if args.force_only:
        print(json.dumps({"force": {"seconds": t_force, "cg_iters": itf, "action": S}}), flush=True)
        return
    line = {
        "metric": "32^3x64 Ls16 single-GPU share: propagator time-to-1e-8 and force time",
        "value": t_prop, "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "dtype": "fp16 hop products with fp32 sums, fp32 CG, fp64 reliable updates" if st["hop_arith_bits"] == 16
        else "fp32",
        "data": "synthetic: seeded exp(i*0.25*H) links, antiperiodic t, point sources, Gaussian pseudofermion",
        "config": {"workload": "12-column point-source propagator, 32^3x64, Ls=16, Moebius b=1.5 c=0.5, M5=1.8, "
                               "mass=0.004, mixed16 multi-RHS, stop 1e-8; then one fermion force (FP64 CG 1e-10)",
                   "rhs_tile": st["rhs_tile"]},
        "propagator": {"seconds": t_prop, "iters": st["total_iter"], "reliable_updates": st["reliable_updates"],
                       "true_residual": st["true_residual"], "true_residual_5d": st["true_residual_5d"]},
        "force": {"seconds": t_force, "cg_iters": itf, "action": S},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--config", default="B", choices=sorted(CONFIGS))
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-single-column", action="store_true",
                    help="skip the FP32 one-column-at-a-time mixed solve (the configuration's literal form)")
    ap.add_argument("--solver", default="mixed16", choices=["mixed16", "mixed"],
                    help="mixed16: FP16 Krylov storage with per-site scale; mixed: FP32 storage")
    ap.add_argument("--reliable-delta", type=float, default=0.01,
                    help="mixed solvers: residual drop between FP64 reliable updates")
    ap.add_argument("--cpu-hermops", type=int, default=3, help="oracle hermop applications (0 = skip)")
    ap.add_argument("--kernel-iters", type=int, default=50)
    ap.add_argument("--force-only", action="store_true",
                    help="config E: one fermion-force evaluation without the propagator (for profiler runs)")
    ap.add_argument("--skip-force", action="store_true", help="config E: the propagator only")
    ap.add_argument("--force-stop-rsd", type=float, default=1e-10, help="config E: stop_rsd of the force's CG")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the kernels and stop (for profiler runs)")
    ap.add_argument("--profile-column", action="store_true",
                    help="solve one propagator column with each solver and stop (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the solve would shard by propagator column or by 4D domain).", file=sys.stderr)
        sys.exit(2)

    if args.config == "C":
        return bench_config_c(args)
    if args.config == "E":
        return bench_config_e(args)

    import numpy as np
    import torch

    import qlattice_amd as q
    from qlattice_amd import _lib
    from qlattice_amd.gauge import rand_gauge_links, twist_boundary_at_boundary

    cfg = CONFIGS[args.config]
    site, ls = cfg["site"], cfg["ls"]
    V = int(np.prod(site))
    log(f"config {args.config}: site {site} ls {ls}; generating gauge field")
    U = rand_gauge_links(site, cfg["sigma"], cfg["seed"])
    twist_boundary_at_boundary(U, site, -0.5, 3)
    geo = q.Geometry(site)
    gf = q.GaugeField(geo)
    gf.data = U
    fa = q.FermionAction(mass=cfg["mass"], ls=ls, m5=cfg["m5"], mobius_scale=cfg["scale"])
    inv = q.InverterDomainWall(gf=gf, fa=fa, solver="mixed")
    inv.set_stop_rsd(1e-8)
    L = _lib.load()

    # kernel timing (HIP events on the inverter's stream)
    log("timing hop kernels")
    hop32_ms = q.bench_kernel(inv, 0, 32, args.kernel_iters)
    hop64_ms = q.bench_kernel(inv, 0, 64, args.kernel_iters)
    m5_32_ms = q.bench_kernel(inv, 2, 32, args.kernel_iters)
    herm32_ms = q.bench_kernel(inv, 3, 32, max(1, args.kernel_iters // 5))
    herm64_ms = q.bench_kernel(inv, 3, 64, max(1, args.kernel_iters // 5))
    prec_tp = 16 if args.solver == "mixed16" else 32
    hm5s_ms = q.bench_kernel(inv, 5, prec_tp, max(1, args.kernel_iters // 5))
    tile = inv.stats()["rhs_tile"]  # the solver's tile (autotuned unless QLAT_DWF_RHS_TILE is set)
    sb = inv.stats()["sites_per_block"]
    arith = inv.stats()["hop_arith_bits"]  # 16: FP16 links and FP16 dot products in the hop phase
    step_ms = q.bench_kernel(inv, 6, prec_tp, max(1, args.kernel_iters // 5))
    hermtp_ms = q.bench_kernel(inv, 7, prec_tp, max(1, args.kernel_iters // 10))
    site_bytes = 52 if prec_tp == 16 else 96
    Vb = 1
    for n_ in site:
        Vb *= n_
    link_bytes = 144 if (prec_tp == 16 and inv.stats()["hop_arith_bits"] == 16) else 288
    iter_bytes = tile * (Vb // 2) * ls * (11 * site_bytes + 96 + 4 * site_bytes + 384) + 4 * Vb * link_bytes
    bm5s = hop_m5s_bytes(site, ls, tile, site_bytes, 2 if arith == 16 else 4)
    gbs_m5s = bm5s / (hm5s_ms * 1e-3) / 1e9
    log(f"three-phase ({prec_tp}-bit storage, FP{arith} hop, {tile} columns, {sb} sites/block): hop+s-matrix {hm5s_ms:.3f} ms "
        f"{gbs_m5s:.0f} GB/s | cg step {step_ms:.3f} ms | hermop+dot {hermtp_ms:.3f} ms "
        f"({hermtp_ms / tile:.3f} ms per column)")
    b32 = hop_bytes(site, ls, 4)
    b64 = hop_bytes(site, ls, 8)
    gbs32 = b32 / (hop32_ms * 1e-3) / 1e9
    gbs64 = b64 / (hop64_ms * 1e-3) / 1e9
    log(f"hop fp32 {hop32_ms:.3f} ms {gbs32:.0f} GB/s | hop fp64 {hop64_ms:.3f} ms {gbs64:.0f} GB/s | "
        f"m5 fp32 {m5_32_ms:.3f} ms | hermop fp32 {herm32_ms:.3f} ms fp64 {herm64_ms:.3f} ms")

    if args.kernels_only:
        # plain copy of one FP32 field: the calibration kernel for counter runs
        copy_ms = q.bench_kernel(inv, 4, 32, args.kernel_iters)
        copy_bytes = 2 * (V // 2) * ls * 96
        log(f"copy fp32 {copy_ms:.3f} ms {copy_bytes / (copy_ms * 1e-3) / 1e9:.0f} GB/s ({copy_bytes} bytes)")
        print(json.dumps({"hop_fp32_ms": hop32_ms, "hop_fp32_gbps": gbs32, "hop_fp64_ms": hop64_ms,
                          "hop_fp64_gbps": gbs64, "m5_fp32_ms": m5_32_ms, "hermop_fp32_ms": herm32_ms,
                          "hermop_fp64_ms": herm64_ms, "hop_m5s_ms": hm5s_ms, "hop_m5s_bytes": bm5s,
                          "rhs_tile": tile, "copy_fp32_ms": copy_ms, "copy_bytes": copy_bytes}), flush=True)
        return

    if args.profile_column:
        src4 = q.FermionField4d(geo)
        src4.data[0, 0] = 1.0
        for solver in (args.solver, "fp64"):
            inv.set_solver(solver)
            t0 = time.perf_counter()
            inv * src4
            st = inv.stats()
            log(f"{solver}: one column {time.perf_counter() - t0:.2f} s, iters {st['total_iter']}")
        return

    # point-source propagator, resident on the device
    ncol = 12
    src_h = np.zeros((V, 12, 12), dtype=np.complex128)
    src_h[0] = np.eye(12)
    src = torch.from_numpy(src_h).cuda()
    sol = torch.zeros_like(src)

    def solve(solver):
        inv.set_solver(solver, args.reliable_delta)
        torch.cuda.synchronize()
        L.invert_prop(inv.cdata, _lib.dev_ptr(sol), _lib.dev_ptr(src), device=True)
        torch.cuda.synchronize()
        return inv.stats()

    for w in range(args.warmup):
        log(f"warmup {w}")
        solve(args.solver)
    log(f"timed {args.solver} solves: {args.steps}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        st_mixed = solve(args.solver)
        log(f"step {s}: iters {st_mixed['total_iter']} reliable {st_mixed['reliable_updates']} "
            f"true_res {st_mixed['true_residual']:.2e} solve {st_mixed['seconds']:.2f} s")
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    sol_mixed = sol.clone()

    t_fp64 = None
    st64 = None
    if not args.skip_fp64:
        log("FP64 solve")
        t0 = time.perf_counter()
        st64 = solve("fp64")
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t0
        log(f"fp64: iters {st64['total_iter']} cycles {st64['cycles']} true_res {st64['true_residual']:.2e} "
            f"{t_fp64:.2f} s")
        diff = (torch.linalg.vector_norm(sol_mixed - sol) / torch.linalg.vector_norm(sol)).item()
    else:
        diff = None

    # The configuration as stated: the 12 columns solved one at a time with
    # FP32 dslash/CG and FP64 reliable updates (multi-RHS tile 1, FP32 storage).
    single = None
    if not args.skip_single_column:
        log("FP32 mixed solve, one column at a time")
        os.environ["QLAT_DWF_RHS_TILE"] = "1"
        try:
            inv1 = q.InverterDomainWall(gf=gf, fa=fa, solver="mixed", reliable_delta=args.reliable_delta)
        finally:
            del os.environ["QLAT_DWF_RHS_TILE"]
        inv1.set_stop_rsd(1e-8)
        sol1 = torch.zeros_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.invert_prop(inv1.cdata, _lib.dev_ptr(sol1), _lib.dev_ptr(src), device=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter() - t0
        st1 = inv1.stats()
        d1 = None if st64 is None else (torch.linalg.vector_norm(sol1 - sol) / torch.linalg.vector_norm(sol)).item()
        single = {"seconds": t1, "iters": st1["total_iter"], "reliable_updates": st1["reliable_updates"],
                  "true_residual": st1["true_residual"], "true_residual_5d": st1["true_residual_5d"],
                  "spinor_bits": st1["spinor_bits"], "rhs_tile": 1, "solution_rel_diff_vs_fp64": d1}
        log(f"fp32 one column at a time: {t1:.2f} s, iters {st1['total_iter']}, true_res {st1['true_residual']:.2e}")
        del inv1, sol1
        torch.cuda.empty_cache()

    cpu = None
    if args.cpu_hermops > 0:
        iters_ref = st64["total_iter"] if st64 else st_mixed["total_iter"]
        log("cpu baseline (oracle)")
        cpu = cpu_baseline(cfg, U, iters_ref, args.cpu_hermops)

    N = V // 2 * ls
    tr = read_traffic(f"hop_m5s_{'fp16' if prec_tp == 16 else 'fp32'}_arith{arith}_{args.config}_tile{tile}_sb{sb}")
    traffic_bytes = None if tr is None else tr["hbm_bytes_per_launch"]
    line = {
        "metric": METRIC,
        "value": t_mixed,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 hop products with fp32 sums, fp32 CG, fp64 reliable updates" if arith == 16 else "fp32",
        "data": "synthetic: seeded exp(i*0.3*H) SU(3) links, antiperiodic t, point sources",
        "config": {
            "workload": f"12-column point-source propagator {site[0]}^3x{site[3]} Ls={ls} Moebius "
                        f"b={0.5 * cfg['scale'] + 0.5:g} c={0.5 * cfg['scale'] - 0.5:g} "
                        f"M5={cfg['m5']:g} mass={cfg['mass']:g} to a 1e-8 FP64 true residual, columns solved in multi-RHS "
                        f"blocks of {tile}, "
                        + ("FP16 links and hop products with FP32 sums, FP32 s-matrix and CG, " if arith == 16
                           else "FP32 arithmetic, ")
                        + f"{prec_tp}-bit Krylov storage"
                        + (" with per-site scale" if prec_tp == 16 else "") + ", FP64 reliable updates",
            "lattice": list(site), "ls": ls, "columns": ncol, "stop_rsd": 1e-8, "rhs_tile": tile,
            "solver": args.solver, "hop_arith_bits": arith,
        },
        "mixed": {"seconds": t_mixed, "iters": st_mixed["total_iter"],
                  "reliable_updates": st_mixed["reliable_updates"],
                  "true_residual": st_mixed["true_residual"], "true_residual_5d": st_mixed["true_residual_5d"],
                  "spinor_bits": st_mixed["spinor_bits"]},
        "fp64": None if st64 is None else {"seconds": t_fp64, "iters": st64["total_iter"],
                                           "cycles": st64["cycles"], "true_residual": st64["true_residual"],
                                           "true_residual_5d": st64["true_residual_5d"]},
        "fp32_single_column": single,
        "speedup_vs_fp64": None if t_fp64 is None else t_fp64 / t_mixed,
        "solution_rel_diff_mixed_vs_fp64": diff,
        "dslash": {
            "hop_fp32_ms": hop32_ms, "hop_fp32_gbps": gbs32, "hop_fp32_gflops": hop_flops(site, ls) / hop32_ms / 1e6,
            "hop_fp64_ms": hop64_ms, "hop_fp64_gbps": gbs64,
            "m5_fp32_ms": m5_32_ms, "hermop_fp32_ms": herm32_ms, "hermop_fp64_ms": herm64_ms,
            "hermop_ref_flops_convention": 5500 * ls * V,
            "three_phase": {"storage_bits": prec_tp, "columns": tile, "sites_per_block": sb, "hop_m5_ms": hm5s_ms,
                            "hop_m5_gbps": gbs_m5s, "hop_arith_bits": arith, "cg_step_ms": step_ms,
                            "hermop_dot_ms": hermtp_ms},
            # one CG iteration of the tile: Mpc^+ Mpc with its dots (4 hop kernels: read
            # t, b, p, y, z, p and r once each, write b, y, z, ap; links 4 times) and the
            # step with x every step (read p, ap, r, x; write r, x, p', t)
            "iteration": {
                "algorithmic_bytes": iter_bytes, "ms": hermtp_ms + step_ms,
                "tbps": iter_bytes / ((hermtp_ms + step_ms) * 1e-3) / 1e12,
                "frac_of_hbm_peak": iter_bytes / ((hermtp_ms + step_ms) * 1e-3) / 1e9 / PEAK_HBM_GBS,
            },
        },
        "roofline": {
            "kernel": f"k_hop_m5s<{'Half16' if prec_tp == 16 else 'float'}, {sb}, {ls}, HM={1 if arith == 16 else 0}> "
                      f"(hop + s-matrix, {tile} columns per launch, {prec_tp}-bit spinor storage, "
                      f"FP{arith} links and hop arithmetic, FP32 s-matrix)",
            "bound": "hbm",
            "achieved": gbs_m5s,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": gbs_m5s / PEAK_HBM_GBS,
            "traffic": traffic_bytes,
            "algorithmic_bytes_per_launch": bm5s,
            "units_per_launch": tile * N,
            "launch_ms": hm5s_ms,
        },
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
