"""Headline benchmark: energetic space-time Galerkin BEM single-layer hot path on one MI355X.

Workload (the FP64 plain method at scale): unit icosphere of level 3
(sphereUniform_3.mesh, 1280 triangles, 3840 DOF), lambda = mu = rho = 1
(c_P = sqrt 3, c_S = 1), t in [0, 4] with 128 steps, L = 66 lag blocks,
7-point Gauss-Hammer outer rule, Dirichlet datum n(x) t^2 e^{-4t}.

One step = the whole hot path on device-resident inputs: assembly of
E_0..E_65 (5.61e7 culled element-pair.lag integrals), right-hand side, E_0
factorisation, the 128-step marching-on-in-time solve and the post-processed
displacement at 4096 points of the shell 1.2 <= |x| <= 2 (seeded) at every step.
value = culled pair.lag integrals / step time.

Also reported: a roofline object for the dominant kernel (the assembly's
per-level integral stage, whose k_node_eval evaluates one outer node per lane)
and the history-sum kernel, and the CPU oracle
(FP64 C/OpenMP restatement of the reference SA path) timed on the host on a
bounded sample of the same assembly.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FP64 operations per (pair, level) item of the per-level stage on
# this workload (7 outer nodes): mean over the items outside the kP and
# between-front ranges (which k_levels_kp and k_levels_bf evaluate) in a
# uniform sample of 200000 of the 43938820 items of the pruned lag windows,
# counted by the op-counting host build of the device math, with the static
# outer nodes whose value an earlier level of the pair holds left out as the
# node tasks leave them out (python tests/opcount.py 200000 -> 6866.8 +- 16.4;
# 8030.1 +- 14.2 with every node evaluated; kP share 0.2585, between-front
# share 0.2634); each + - * / (a fused multiply-add as 2), sqrt and libm call
# counts once.
ALG_FLOP_PER_ITEM = 6866.8
# HBM bytes per call of the per-level integral stage on this workload (the
# kernels k_node_keys, k_node_scatter, k_node_eval, k_node_reduce and k_levels
# for the coplanar pairs) from the rocprofv3 --pmc passes FETCH_SIZE (x2.0:
# the counter reads 512 MiB for a 1 GiB device copy) and WRITE_SIZE (x1.0):
# 17.0 GB + 7.2 GB, of which k_node_eval 2.3 GB + 5.5 GB (the node slots) and
# k_node_reduce 14.2 GB + 1.1 GB (the slots read back).
K_LEVELS_TRAFFIC_BYTES = 24.1e9
K_LEVELS_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes of this bench's per-level stage "
                           "(profiles/pmc_counters_sphere3_fp64.md), not measured in this run")
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# Algorithmic FP64 operations of the double-layer assembly, from the
# op-counting host build of its math (tests/native/k_opcount.cpp).  Regular
# part (k_ksf, by field triangle, level sums): per evaluation point (one outer
# node, one inner node of the field triangle F, F's three hat functions) and
# chunk of nL <= 2 levels, 242 for its geometry and basis values (powers of
# 1/r once) and 488 per level (R_j and KL + KT once, combined for the three
# vertices, 27 sums).  The
# gather's combination -1, 3, -3, 1 of the level sums (at most 81 per slot
# and triangle, ~1e-4 of the total) is not counted.  Singular part (k_ksing):
# 585 per level and evaluation point (three children x nDiag nodes per outer
# node).
K_SF_POINT_BASE, K_SF_LEVEL = 242, 488
# FP64 flops the hardware executes for the regular part (k_ksf) on the dd
# workload (barH1 level 2, FN 19 64 3 256): 64 lanes x (2 FMA + ADD + MUL +
# TRANS) wave instructions from the rocprofv3 --pmc pass, lane activity 1.00
# (profiles/pmc_counters_barH1_2_dd.md); the source-level count above gives
# 1.90e13 for the same work.
K_SF_EXEC_FLOPS = 1.673e13
K_SING_POINT_LEVEL = 585
K_SF_CHUNK = 2


def k_sf_point_flops(mA, nL):
    """Operations of one field-triangle evaluation point for the chunk of nL
    levels from mA."""
    return K_SF_POINT_BASE + K_SF_LEVEL * np.asarray(nL)


def workload():
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import read_mesh
    m = read_mesh(os.path.join(ROOT, "tests", "golden", "meshes", "sphereUniform_3.mesh"))
    pb = api.PbParam(rho=1.0, mu=1.0, lambda_=1.0, Tfin=4.0, nT=128, domainType="sphereUniform", lev=3)
    return m, pb


def shell_points(n=4096, seed=0):
    """Points uniform in volume in the shell 1.2 <= |x| <= 2.0 (numpy PCG64, seed 0)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = (1.2 ** 3 + rng.random(n) * (2.0 ** 3 - 1.2 ** 3)) ** (1 / 3)
    return d * r[:, None]


def history_bytes(per_lag, nT):
    """Bytes of the nT history sums if every step streamed its blocks: 48 B of
    values + 4 B column index per stored block with 1 <= lag <= min(n, L-1)
    (per_lag: stored blocks per lag)."""
    L = len(per_lag)
    blocks = sum(int(per_lag[1:min(n, L - 1) + 1].sum()) for n in range(nT))
    return blocks * 52


def history_bytes_blocked(per_lag, nT, B=8):
    """Store bytes the blocked history sum actually streams: per block of B
    steps starting at n0 > 0 one pass over lags 1..min(L-1, n0+B-1)
    (k_history_far), plus lags 1..min(b, L-1) for step n0+b (k_history)."""
    L = len(per_lag)
    blocks = 0
    for n0 in range(0, nT, B):
        nb = min(B, nT - n0)
        if n0 > 0:
            blocks += int(per_lag[1:min(L - 1, n0 + nb - 1) + 1].sum())
        for b in range(nb):
            blocks += int(per_lag[1:min(b, L - 1) + 1].sum())
    return blocks * 52


def _cpu_info():
    """Host cores this process may use and the CPU model (lscpu's model name)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": os.cpu_count(), "affinity": avail, "omp_num_threads": env, "model": model}


def _median3(fn):
    """Median wall time of three runs of fn (and its last result)."""
    ts, r = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), r


def cpu_baseline(m, pb):
    """The oracle (FP64 C/OpenMP restatement of the reference's SA path) on the
    host cores this process may use (the GPU box allots a share of the
    machine; OMP_NUM_THREADS states it), median of 3 runs per part:
    * this workload's assembly on a sample of test rows (all lags), in the
      metric's unit (pair.lag integrals/s) -- the reported value;
    * the analytic post-processing of the solved density on this workload at a
      sample of the shell points and times (single-threaded per evaluation, as
      the reference's per-point loop);
    * the whole plain method on the CPU-runnable 8x8 screen problem (128
      triangles, 16 steps): assembly of E_0..E_15, the marching (LU of E_0 +
      16 steps) and the post-processing at 16 points x 16 times."""
    from oracle import Oracle
    from elastodynamicebem_amd.mesh import read_mesh
    threads, info = _cpu_info()
    o = Oracle(m.coordinates, m.triangles, Tfin=pb.Tfin, nT=pb.nT)
    L = o.num_blocks()
    rows = (0, 32)
    c, mx = m.center, m.maxL
    dt = pb.Tfin / pb.nT
    d = np.sqrt(((c[rows[0]:rows[1], None, :] - c[None, :, :]) ** 2).sum(-1))
    dmin = d - mx[None, :] - mx[rows[0]:rows[1], None]
    dmax = d + mx[None, :] + mx[rows[0]:rows[1], None]
    units = sum(np.count_nonzero(((l - 1) * pb.velS * dt < dmax) & ((l + 1) * pb.velP * dt >= dmin))
                for l in range(L))
    t_asm, _ = _median3(lambda: [o.matrix_block(l, rows=rows, nthreads=threads) for l in range(L)])
    # post-processing sample on this workload: a smooth random density (the
    # evaluation's cost does not depend on the values), 4 points x 8 times
    rng = np.random.default_rng(0)
    dens = rng.normal(size=(3 * m.numberTriangles, pb.nT))
    xs = shell_points()[:4]
    ts = pb.Tfin * np.arange(1, 9) / 8
    t_pp1, _ = _median3(lambda: [o.pp_point(dens, x, t) for x in xs for t in ts])
    # the whole plain method on the screen problem
    sm = read_mesh(os.path.join(ROOT, "tests", "golden", "meshes", "screenUniform_2.mesh"))
    so = Oracle(sm.coordinates, sm.triangles, Tfin=1.0, nT=16)
    sL = so.num_blocks()
    t0_asm, E0 = _median3(lambda: np.stack([so.matrix_block(l, nthreads=threads)[0] for l in range(sL)]))
    beta0 = so.rhs(0)
    t0_mot, a0 = _median3(lambda: so.mot(E0, beta0))
    x0 = np.stack([sm.center[j] + np.array([0.0, 0.0, 0.25]) for j in range(0, 128, 8)])
    t0s = np.arange(1, 17) / 16
    t0_pp, _ = _median3(lambda: [so.pp_point(a0, x, t) for x in x0 for t in t0s])
    return {"value": units / t_asm, "unit": "pair·lag integrals/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C/OpenMP FP64 restatement of the reference SA path), median of 3: this "
                      f"workload's assembly of test rows {rows[0]}-{rows[1] - 1} for E_0..E_{L - 1} "
                      f"({units} culled pair.lags, {t_asm:.2f} s)",
            "host": info,
            "parts": {
                "workload_assembly_s": t_asm, "workload_assembly_pair_lags": int(units),
                "workload_pp_s_per_point_time": t_pp1 / (len(xs) * len(ts)),
                "workload_pp_sample": f"{len(xs)} shell points x {len(ts)} times, one thread per evaluation",
                "screen8x8_assembly_s": t0_asm, "screen8x8_mot_s": t0_mot,
                "screen8x8_pp_s": t0_pp, "screen8x8_pp_sample": "16 points x 16 times",
            }}


def _solve(s, api, xpts, tpts, precision):
    """Assemble (timed by the library's events), solve, post-process; host copies."""
    s.assemble(precision)
    t = s.timings()
    s.rhs(api.DATUM_SPHERE_T2EXP)
    a = s.mot_solve(None)
    u = s.pp_eval(xpts, tpts)
    return t, a, u


def chosen_forms(s, m, pb, xpts, tpts, train_samples=2_000_000, epochs=20):
    """The two chosen forms on the same problem, each against the plain FP64
    form of the same run: (1) far field in FP32, (2) near field from the learned
    surrogate (trained here, on the device, on seeded near-field pairs of
    generated icospheres and screens), and both together."""
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd import surrogate as SG
    rel = lambda p, q: float(np.linalg.norm(p - q) / np.linalg.norm(q))
    t64, a64, u64 = _solve(s, api, xpts, tpts, api.PREC_FP64)
    out = {"fp64": {"assembly_ms": t64["assembly_ms"], "integral_kernel_ms": t64["levels_ms"]}}
    tm, am, um = _solve(s, api, xpts, tpts, api.PREC_MIXED)
    st = s.assembly_stats()
    out["mixed"] = {"assembly_ms": tm["assembly_ms"], "integral_kernel_ms": tm["levels_ms"],
                    "speedup_vs_fp64": t64["assembly_ms"] / tm["assembly_ms"],
                    "fp32_items": st["far_items"], "fp64_items": st["fp64_items"],
                    "density_rel_l2_vs_fp64": rel(am, a64), "field_rel_l2_vs_fp64": rel(um, u64)}
    print("bench: forms: mixed done; generating surrogate training samples", file=sys.stderr, flush=True)
    sur = SG.Surrogate()
    t0 = time.perf_counter()
    ntr, nhe = SG.build_training_set(sur, train_samples, seed=0, held_frac=0.02)
    tgen = time.perf_counter() - t0
    sur.init(ntr, seed=0)
    print(f"bench: forms: {ntr} samples in {tgen:.1f} s; training", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    losses = sur.train(ntr, epochs, batch=2048, lr=2e-3, seed=0)
    ttrain = time.perf_counter() - t0
    print(f"bench: forms: trained in {ttrain:.1f} s", file=sys.stderr, flush=True)
    held = sur.loss(ntr, nhe)
    x, y = sur.get_samples(ntr, nhe)
    lab_err = rel(sur.predict(x), y)
    sur.attach(s)
    s.set_option(api.OPT_SURROGATE_SOLVE, 1)  # measured here against the FP64 solve
    res = {}
    for name, prec in (("surrogate", api.PREC_SURROGATE), ("mixed_surrogate", api.PREC_MIXED | api.PREC_SURROGATE)):
        ts, as_, us = _solve(s, api, xpts, tpts, prec)
        ss = s.surrogate_stats()
        res[name] = {"assembly_ms": ts["assembly_ms"], "integral_kernel_ms": ts["levels_ms"],
                     "surrogate_ms": ss["ms"], "surrogate_blocks": ss["blocks"],
                     "speedup_vs_fp64": t64["assembly_ms"] / ts["assembly_ms"],
                     "density_rel_l2_vs_fp64": rel(as_, a64), "field_rel_l2_vs_fp64": rel(us, u64)}
    sur.detach(s)
    sur.close()
    out.update(res)
    out["surrogate_training"] = {"samples": ntr, "held_out": nhe, "sample_generation_s": tgen, "epochs": epochs,
                                 "train_s": ttrain, "final_train_loss": float(losses[-1]),
                                 "held_out_loss": held, "held_out_label_rel_l2": lab_err}
    return out


def dd_workload(level=2):
    """The reference's barH1 problems (inputFiles/input_barH1_*.txt): rho 0.25,
    mu 0.5, lambda 0, t in [0, 6]; level 2 is 'large' (768 triangles, 96 steps)."""
    from elastodynamicebem_amd import api
    from elastodynamicebem_amd.mesh import read_mesh
    m = read_mesh(os.path.join(ROOT, "tests", "golden", "meshes", f"barH1_{level}.mesh"))
    nT = {1: 48, 2: 96}[level]
    pb = api.PbParam(rho=0.25, mu=0.5, lambda_=0.0, Tfin=6.0, nT=nT, domainType="barH1", lev=level)
    return m, pb


def k_windows_and_flops(m, pb, LK, nGH, nInt, nDiag, rows=None):
    """Slots of the K store (pair windows of isBlockNull and the singular
    cut-off, kernelK.cu:226-260, calcSubBlockSingK.m:12) and the algorithmic
    FP64 operations of the assembly: the regular part by (test triangle S,
    field triangle F != S) over the positive levels of the union of the
    windows of F's vertices, each evaluation point's geometry counted once per
    K_SF_CHUNK = 2 levels (k_ksf recomputes it per level, which runs faster at
    two waves per SIMD; the count does not charge that) and its kernel once
    per level; singular parts per lag and level.
    rows = (r0, r1) restricts the test triangles."""
    X, T0 = m.coordinates, m.triangles
    r0, r1 = (0, T0.shape[0]) if rows is None else rows
    T = T0[r0:r1]
    N, Nv = T.shape[0], X.shape[0]
    dt, cP, cS = pb.Tfin / pb.nT, pb.velP, pb.velS
    maxLen = float(m.maxL.max())
    d = np.sqrt(((X[T][:, :, None, :] - X[None, None, :, :]) ** 2).sum(-1))  # N x 3 x Nv
    dmax, dmin = d.max(1), d.min(1)
    own = np.zeros((N, Nv), bool)
    own[np.arange(N)[:, None], T] = True
    cut = np.ceil(m.maxL[r0:r1] / (cS * dt)).astype(int) + 2
    big = 1 << 30
    rlo, rhi = np.full((N, Nv), big), np.full((N, Nv), -1)
    slots = 0
    sing_flops = 0.0
    for lag in range(LK):
        reg = ~(((lag - 2) * cS * dt > dmax + maxLen) | ((lag + 1) * cP * dt < dmin - 2 * maxLen))
        sing = own & (lag < cut[:, None])
        slots += int(np.count_nonzero(reg | sing))
        rlo = np.where(reg, np.minimum(rlo, lag), rlo)
        rhi = np.where(reg, np.maximum(rhi, lag + 1), rhi)
        nlev = sum(1 for c in (-2, -1, 0, 1) if lag + c >= 1)
        sing_flops += float(sing.sum()) * nGH * 3 * nDiag * K_SING_POINT_LEVEL * nlev
    # window of (S, F): lags where some vertex of F is not null against S
    a = rlo[:, T0].min(-1)  # N x N_all
    b = rhi[:, T0].max(-1)
    S_glob = np.arange(r0, r1)
    a[np.arange(N), S_glob] = 0
    b[np.arange(N), S_glob] = 0
    nlev = np.where(b > a, b - np.maximum(a - 2, 1) + 1, 0)
    nchunk = (nlev + K_SF_CHUNK - 1) // K_SF_CHUNK
    reg_flops = float((nchunk * K_SF_POINT_BASE + nlev * K_SF_LEVEL).sum())
    return slots, reg_flops * nGH * nInt, sing_flops


def dd_main(args):
    """Direct Dirichlet step on the level-2 bar with the reference's default
    method FN 19 64 3 256 and post-processing "GPU 64 3" (main.m:16-20):
    V (numeric GHC form), K_0..K_{LK-1}, rhs / 2 + beta_K, the MOT and the
    direct post-processing u = G alpha - P g at the reference's barH1 point
    set (BEMenerg_postProc_loadPoints: 3 points x 600 times), inputs resident
    on the device; value = K blocks / K assembly time."""
    from elastodynamicebem_amd import api
    m, pb = dd_workload(2)
    nGH, nSub, nSing, nDiag = 19, 64, 3, 256
    s = api.Solver(m, pb, nGH, device=0)
    s.set_inner_rule(nSub, nSing, nDiag)
    s.set_pp_rule(64, 3)
    xpp, tpp, _, npp, _ = api.BEMenerg_postProc_loadPoints(pb)
    LK = s.k_num_blocks()

    def step():
        s.assemble(api.PREC_GHC)
        s.assemble_k()
        s.rhs_dirichlet(api.DATUM_REF_BAR_H1)
        s.mot_solve(None)
        return s.pp_eval_direct(xpp, tpp)

    for w in range(args.warmup):
        step()
        print(f"bench dd: warmup step {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    times, tims = [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
        tims.append(s.timings())
        print(f"bench dd: timed step {k + 1}/{args.steps}: {1e3 * times[-1]:.1f} ms", file=sys.stderr, flush=True)
    info = s.k_store_info()
    # the Neumann driver on the same V and K (timeMarchingN.m:170-296), once
    t0 = time.perf_counter()
    s.mot_solve_neumann(api.DATUM_REF_BAR_H1)
    neu_wall = 1e3 * (time.perf_counter() - t0)
    neu_ms = s.timings()["mot_neumann_ms"]
    s.close()
    slots, reg_flops, sing_flops = k_windows_and_flops(m, pb, LK, nGH, nSub * nSing, nDiag)
    assert slots == info["slots"], (slots, info["slots"])
    k_ms = float(np.mean([t["k_assembly_ms"] for t in tims]))
    # the regular part at the FP64 work the hardware executes for it (the
    # source-level count repeats level-independent products the compiler
    # evaluates once per point), the singular part at its source-level count
    ach = (K_SF_EXEC_FLOPS + sing_flops) / (k_ms * 1e-3) / 1e12
    out = {
        "metric": "double-layer K blocks (test triangle, vertex, lag) per second, direct Dirichlet driver",
        "value": slots / (k_ms * 1e-3), "unit": "K blocks/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.mean(times)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "barH1 level 2 (768 triangles, 386 vertices), rho 0.25, mu 0.5, lambda 0, "
                               "t in [0, 6], 96 steps, FN 19 64 3 256 + PP GPU 64 3: V (GHC) + K + rhs/2 + "
                               "beta_K + MOT + direct post-processing at 3 points x 600 times",
                   "pp_points": int(npp),
                   "k_lags": LK, "k_slots": slots, "k_pairs": info["pairs"], "k_store_bytes": info["bytes"]},
        "breakdown_ms": {"v_assembly": float(np.mean([t["assembly_ms"] for t in tims])), "k_assembly": k_ms,
                         "beta_k": float(np.mean([t["beta_k_ms"] for t in tims])),
                         "mot_total": float(np.mean([t["mot_ms"] for t in tims])),
                         "post_processing_direct": float(np.mean([t["pp_direct_ms"] for t in tims]))},
        "neumann_driver_ms": {"mot_neumann": neu_ms, "wall": neu_wall,
                              "note": "datum, beta_V, QR of K_0 + I_Gamma (rocSOLVER), 96 least-squares steps; "
                                      "outside the timed step"},
        "roofline": {"bound": "valu", "pipe": "fp64-valu", "kernel": "K assembly (k_ksf + k_kgather + k_ksing)",
                     "achieved": ach,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS,
                     "traffic": None,
                     "flops": {"regular_executed": K_SF_EXEC_FLOPS, "regular_source_count": reg_flops,
                               "singular_source_count": sing_flops,
                               "source": "regular: rocprofv3 --pmc FP64 instruction counts of k_ksf "
                                         "(profiles/pmc_counters_barH1_2_dd.md)"}},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import Oracle
        threads, _ = _cpu_info()
        o = Oracle(m.coordinates, m.triangles, rho=pb.rho, mu=pb.mu, lam=pb.lambda_, Tfin=pb.Tfin, nT=pb.nT,
                   nGH=nGH)
        o.set_ghc(nSub, nSing, nDiag)
        rows = (0, 1)
        t0 = time.perf_counter()
        for lag in range(LK):
            o.k_matrix(lag, rows, nthreads=threads)
        el = time.perf_counter() - t0
        sub_slots, _, _ = k_windows_and_flops(m, pb, LK, nGH, nSub * nSing, nDiag, rows)
        out["cpu_baseline"] = {"value": sub_slots / el, "unit": "K blocks/s", "cores": threads, "kind": "port",
                               "sample": f"oracle (C/OpenMP FP64) K blocks of test row {rows[0]} for all "
                                         f"{LK} lags ({sub_slots} blocks, {el:.1f} s)"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-overlap", action="store_true",
                    help="post-process after the marching without ebem_pp_prepare (no overlap)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the mixed-precision and surrogate forms (their timing and error vs FP64)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy kernel of MIB MiB (traffic-counter calibration)")
    ap.add_argument("--workload", choices=("step", "dd"), default="step",
                    help="step: the single-layer hot path (default); dd: the direct Dirichlet driver with K")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "dd":
        import torch  # noqa: F401  (HIP runtime present)
        dd_main(args)
        return

    import torch  # noqa: F401  (HIP runtime present)
    from elastodynamicebem_amd import api

    if args.calib_copy:
        import torch
        a = torch.ones(args.calib_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    m, pb = workload()
    s = api.Solver(m, pb, 7, device=0)
    L = s.num_blocks()

    xpts = shell_points()
    tpts = pb.Tfin / pb.nT * np.arange(1, pb.nT + 1)
    upts = np.zeros((pb.nT, xpts.shape[0], 3))
    from elastodynamicebem_amd import _lib

    overlap = not args.no_overlap

    def step():
        s.assemble(api.PREC_FP64)
        s.rhs(api.DATUM_SPHERE_T2EXP)
        if overlap:  # the density-independent part of the post-processing runs beside the marching
            _lib.check(s.lib.ebem_pp_prepare(s.h, xpts.shape[0], _lib.dptr(xpts), tpts.size, _lib.dptr(tpts)))
        _lib.check(s.lib.ebem_mot_solve(s.h, None, None))
        _lib.check(s.lib.ebem_pp_eval(s.h, None, xpts.shape[0], _lib.dptr(xpts), tpts.size, _lib.dptr(tpts),
                                      _lib.dptr(upts)))

    for w in range(args.warmup):
        step()
        print(f"bench: warmup step {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    times, tims = [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
        tims.append(s.timings())
        print(f"bench: timed step {k + 1}/{args.steps}: {1e3 * times[-1]:.1f} ms", file=sys.stderr, flush=True)
    iso = None
    if overlap:  # one more step, untimed, without the overlap: the marching's own figures
        overlap = False
        step()
        iso = s.timings()
        overlap = True
    info = s.store_info()
    ms_step = 1e3 * float(np.mean(times))
    units = info["pair_lags"]
    lev_ms = float(np.mean([t["levels_ms"] for t in tims]))
    kp_ms = float(np.mean([t["levels_kp_ms"] for t in tims]))
    ppkp_ms = float(np.mean([t["pp_kp_ms"] for t in tims]))
    bf_ms = float(np.mean([t["levels_bf_ms"] for t in tims]))
    ppbf_ms = float(np.mean([t["pp_bf_ms"] for t in tims]))
    kp = s.assembly_kp_stats()
    level_items = info["items"] - kp["levels"] - kp["bf_levels"]
    asm_ms = float(np.mean([t["assembly_ms"] for t in tims]))
    mot_ms = float(np.mean([t["mot_ms"] for t in tims]))
    hist_ms = float(np.mean([t["history_ms"] for t in tims]))
    fac_ms = float(np.mean([t["factor_ms"] for t in tims]))
    pp_ms = float(np.mean([t["pp_ms"] for t in tims]))
    prep_ms = float(np.mean([t["pp_prepare_ms"] for t in tims]))
    per_lag = s.lag_blocks()
    hb = history_bytes(per_lag, pb.nT)
    hb_streamed = history_bytes_blocked(per_lag, pb.nT)

    roof = None
    if ALG_FLOP_PER_ITEM:
        ach = ALG_FLOP_PER_ITEM * level_items / (lev_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "pipe": "fp64-valu",
                "kernel": "per-level integral stage (k_node_keys, k_node_scatter, k_node_eval, k_node_reduce; "
                          "k_levels for coplanar pairs)", "achieved": ach,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS,
                "traffic": K_LEVELS_TRAFFIC_BYTES, "traffic_source": K_LEVELS_TRAFFIC_SOURCE}
    hist_alone_ms = iso["history_ms"] if iso else hist_ms
    hist_gbs = hb_streamed / (hist_alone_ms * 1e-3) / 1e9
    out = {
        "metric": "element-pair·lag integrals/s + solve time at ≤1e-6 rel err vs FP64, one MI355X",
        "value": units / (ms_step * 1e-3),
        "unit": "pair·lag integrals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "sphereUniform_3 icosphere (1280 triangles, 3840 DOF), T=4, nT=128, L=66, "
                               "SA 7-point outer rule, plain FP64: assembly + RHS + E0 factorisation + MOT "
                               "+ post-processing at 4096 shell points x 128 steps"
                               + ("; the post-processing's density-independent part (M integrals) on a second "
                                  "stream beside the MOT (ebem_pp_prepare)" if overlap else ""),
                   "pair_lags": units, "items": info["items"], "items_per_level": level_items,
                   "kp_ranges": kp["ranges"], "kp_levels": kp["levels"], "bf_ranges": kp["bf_ranges"],
                   "bf_levels": kp["bf_levels"], "store_bytes": info["bytes"]},
        "breakdown_ms": {"assembly": asm_ms, "assembly_integral_kernel": lev_ms,
                         "assembly_kp_kernel": kp_ms, "assembly_bf_kernel": bf_ms, "mot_total": mot_ms,
                         "e0_factor": fac_ms, "history_sum": hist_ms, "post_processing": pp_ms,
                         "post_processing_prepared_beside_mot": prep_ms,
                         "mot_total_alone": iso["mot_ms"] if iso else mot_ms,
                         "history_sum_alone": hist_alone_ms,
                         "post_processing_kp_kernel": ppkp_ms, "post_processing_bf_kernel": ppbf_ms},
        "history_sum": {"bytes_streamed": hb_streamed, "bytes_if_per_step": hb, "achieved_GBs": hist_gbs,
                        "peak_GBs": PEAK_HBM_GBS, "frac": hist_gbs / PEAK_HBM_GBS,
                        "note": "blocked over 8 steps: one store pass per block plus the short lags per step"
                                + ("; timed in an untimed extra step without the post-processing beside it "
                                   "(in the timed steps the two share the GPU: history_sum in breakdown_ms)"
                                   if iso else "")},
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_forms:
        out["forms"] = chosen_forms(s, m, pb, xpts, tpts)
    s.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, pb)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
