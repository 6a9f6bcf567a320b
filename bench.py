#!/usr/bin/env python3
"""Benchmark of the HFPx3D hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): an elliptical
crack with semi-axes 1 and 0.5, rotated 30 degrees about x, meshed with 4,000
six-node triangles (72,000 DoF, fix_tip = 0), E = 1, nu = 0.3, and 64 load
cases (uniform pressure p = 1 plus 63 random symmetric far-field stresses with
entries N(0, 0.2^2), seed 1).

One step = the reference's timed "Assembly" and "Solution" stages on the GPU:
element geometry + the dense influence matrix (FP64, plus its FP32 rounding)
+ the 64 right-hand sides, then the FP32-MFMA LU with partial pivoting and
FP64 GMRES iterative refinement of all 64 solutions against the FP64 matrix.
The mesh is uploaded to HBM before the timed region.

value = assembly throughput in matrix entries per second (n^2 / t_assembly);
solve seconds, the plain FP64 GPU path, the error against it, the roofline of
the dominant kernel and a CPU baseline are reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"hbm_GBs": 8000.0, "fp64_mfma_TFs": 78.6, "fp64_valu_TFs": 78.6, "fp32_mfma_TFs": 157.3,
         "bf16_mfma_TFs": 2500.0}
# FP64 VALU flops per (point, element) pair evaluation of the element
# integral: the kernel's SQ_INSTS_VALU_FLOPS_FP64 + SQ_INSTS_VALU_FLOPS_FP64_TRANS
# counts x 64 (the counters count per wave instruction; calibrated on a kernel
# of known flop count) / pairs, from rocprofv3 --pmc of
# `bench.py --workload assemble72k` (k_assemble: 9.6e7 pairs) and
# `--workload stress1m` (k_stress: 6e9 pairs); profiles/valu_pmc_assemble72k_stress1m.md
VALU_FLOPS_PER_PAIR = {"assemble72k": 6992.5, "stress1m": 7140.2}
# The FP32 trailing updates run as six bf16 MFMA products of three-way operand
# splits: their FP32-equivalent ceiling is the dense bf16 peak / 6.
X3_PRODUCTS = 6
X3_PEAK = PEAKS["bf16_mfma_TFs"] / X3_PRODUCTS
# bytes per call of the first trailing update (bench.py --workload gemm72k),
# from the FETCH_SIZE / WRITE_SIZE passes summarised in profiles/
GEMM72K_TRAFFIC = 1.378e+11
GEMM72K_ALG_BYTES = 4.0882e+10
# PMC bytes of one truncation gather of the default volume-control workload
# (72,001^2 resident, 58,285^2 truncated; k_vc_gather1): FETCH_SIZE x 1.99995
# + WRITE_SIZE x 1.0, the factors from a 2 GiB device copy in the same runs
VC72K_TRAFFIC = 7.249e+10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-el", type=int, default=4000)
    ap.add_argument("--nrhs", type=int, default=64)
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-el", type=int, default=1400)
    ap.add_argument("--gemm-dtype", default="f32", choices=["f32", "f64"], help="gemm72k: FP32 or FP64 product")
    ap.add_argument("--workload", default="ellipse72k",
                    choices=["ellipse72k", "wavy108k", "batch256", "gemm72k", "vc72k", "vctrunc72k", "cf72k", "assemble72k",
                             "stress1m"],
                    help="ellipse72k: the headline configuration (default); wavy108k: the non-planar "
                         "6,000-element crack with stresses at 10^6 points; batch256: 256 independent cracks; "
                         "vc72k: volume-control steps (truncation + FP32 LU + FP64 refinement) on the 72k ellipse; "
                         "vctrunc72k: the truncation (mod_3dbem_system_vc) alone, same system; "
                         "cf72k: steps of the cohesion/friction volume-control iteration on the 72k ellipse")
    ap.add_argument("--vc-active", type=float, default=0.9,
                    help="vc72k: DoF of elements whose centroid lies inside this fraction of the ellipse stay active")
    ap.add_argument("--mixed-min-n", type=int, default=1024,
                    help="batch256: cracks with at least this many DoF use FP32 LU + FP64 refinement")
    ap.add_argument("--n-pts", type=int, default=1_000_000)
    ap.add_argument("--fp64-compare", action="store_true", help="wavy108k: also run the plain FP64 path")
    ap.add_argument("--no-lib-prof", action="store_true",
                    help="do not time kernel launches with HIP events (for runs under an external profiler)")
    ap.add_argument("--lu-nb", type=int, default=0, help="outer block of the LU (0: library default)")
    ap.add_argument("--gemm-variant", type=int, default=0, help="FP32 GEMM variant (0: library default)")
    ap.add_argument("--panel", type=int, default=0, help="LU panel kernel knob (0: library default)")
    ap.add_argument("--lookahead", type=int, default=0, help="LU look-ahead knob (0: library default)")
    ap.add_argument("--autotune", default="", help="apply the knob settings of an autotune table "
                                                   "(python -m hfpx3d_amd.autotune) for the LU's n; "
                                                   "batch256: the table's batch row (update kernel and "
                                                   "FP32 threshold, overriding --mixed-min-n)")
    ap.add_argument("--calib-copy-gib", type=float, default=0.0,
                    help="before the workload, copy this many GiB device-to-device (a transfer of known size "
                         "for scaling hardware byte counters)")
    ap.add_argument("--tol", type=float, default=1e-14,
                    help="normwise backward error at which the refinement stops")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def c3_problem(n_el, nrhs, seed=1):
    import numpy as np
    from hfpx3d_amd import meshgen
    R = meshgen.rotation([1.0, 0.0, 0.0], np.pi / 6)
    nods, conn = meshgen.ellipse(n_el, 1.0, 0.5, R)
    rng = np.random.default_rng(seed)
    S = np.zeros((6, nrhs))
    S[:3, 0] = 1.0  # pressure p = 1 on the faces <=> s_inf = p I
    if nrhs > 1:
        S[:, 1:] = rng.normal(0.0, 0.2, size=(6, nrhs - 1))
    return nods, conn, S


def host_cpu():
    """(model name, cores this process may run on, cores used): the GPU box
    gives one GPU's share of its cores (16), whatever the machine has."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, visible, max(1, min(16, visible))


def cpu_baseline(n_el_sample, G, nu, n_target, nrhs):
    """The reference's own assembly (oracle/_ref, else the oracle
    restatement) on the same geometry with fewer elements, and LAPACK
    dgetrf + dgetrs (the reference's il::LU, LU.h:79-138) through scipy at
    two sizes, extrapolated to n_target by n^3 from the larger one.  Every
    leg runs on the same number of host threads (OpenMP for the assembly,
    the BLAS thread pool pinned with threadpoolctl).  Returns a dict."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from hfpx3d_amd import meshgen
    model, visible, threads = host_cpu()
    R = meshgen.rotation([1.0, 0.0, 0.0], np.pi / 6)
    nods, conn = meshgen.ellipse(n_el_sample, 1.0, 0.5, R)
    ref = oracle_lib.reference_or_none()
    kind = "reference" if ref is not None else "port"
    t = time.perf_counter()
    if ref is not None:
        A, _ = ref.assemble(nods, conn, G, nu, 0.125, 0, nthreads=threads)
    else:
        A, _ = oracle_lib.oracle().assemble(nods, conn, G, nu, 0.125, 0, nthreads=threads)
    dt = time.perf_counter() - t
    n = A.shape[0]
    del A
    out = {"value": n * n / dt / 1e9, "unit": "Gentries/s", "cores": threads, "kind": kind,
           "cpu_model": model, "host_cores_visible": visible,
           "sample": f"{'reference make_3dbem_matrix_s (oracle/_ref)' if ref else 'oracle C restatement'} "
                     f"on the same rotated ellipse meshed with {n_el_sample} elements ({n} DoF), "
                     f"{threads} OpenMP threads, {dt:.1f} s; entries/s is the per-entry cost"}
    try:
        import scipy.linalg as sl
        from threadpoolctl import threadpool_limits
        lapack = {}
        with threadpool_limits(limits=threads, user_api="blas"):
            for m in (8000, 16000):
                M = np.random.default_rng(0).normal(size=(m, m)) + m ** 0.5 * np.eye(m)
                Bm = np.ones((m, nrhs))
                t = time.perf_counter()
                lu = sl.lu_factor(M, overwrite_a=True, check_finite=False)
                sl.lu_solve(lu, Bm, check_finite=False)
                lapack[m] = time.perf_counter() - t
                del M, lu
        out["lapack_getrf_getrs_s"] = {str(m): round(v, 3) for m, v in lapack.items()}
        out["lapack_threads"] = threads
        out["lapack_n3_ratio_16k_8k"] = round(lapack[16000] / lapack[8000], 2)
        out["solve_s_extrapolated"] = round(lapack[16000] * (n_target / 16000.0) ** 3, 1)
        out["solve_extrapolation"] = (f"LAPACK dgetrf + dgetrs ({nrhs} RHS) through scipy at n = 16,000, "
                                      f"scaled by (n / 16,000)^3 to n = {n_target}")
    except Exception as ex:  # pragma: no cover
        out["solve_error"] = str(ex)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "hfpx3d_amd runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is design-only)"}))
        sys.exit(2)
    if args.workload == "wavy108k":
        return run_wavy(args)
    if args.workload == "batch256":
        return run_batch(args)
    if args.workload == "gemm72k":
        return run_gemm(args)
    if args.workload in ("vc72k", "vctrunc72k"):
        return run_vc(args)
    if args.workload == "cf72k":
        return run_cf(args)
    if args.workload in ("assemble72k", "stress1m"):
        return run_valu(args)
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, _lib

    E_mod, nu = 1.0, 0.3
    G = E_mod / (2 * (1 + nu))
    nods, conn, S = c3_problem(args.n_el, args.nrhs)
    n_el = conn.shape[1]
    n = 18 * n_el
    ctx = hx.get_context(0)
    lib = ctx.lib
    if args.autotune:
        from hfpx3d_amd import autotune
        log(f"bench: autotune settings {autotune.apply(autotune.load(args.autotune), n, 'f32', ctx)}")
    if args.lu_nb:
        hx.tune_set(_lib.KNOB_LU_NB, args.lu_nb, ctx)
    if args.gemm_variant:
        hx.tune_set(_lib.KNOB_GEMM_F32, args.gemm_variant, ctx)
    if args.panel:
        hx.tune_set(_lib.KNOB_PANEL, args.panel, ctx)
    if args.lookahead:
        hx.tune_set(_lib.KNOB_LU_LOOKAHEAD, args.lookahead, ctx)
    calib_copy(args, ctx)
    log(f"bench: n_el={n_el} n={n} nrhs={args.nrhs}")

    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    dm = hx.DeviceMesh(mesh, None, 0.125, ctx)
    A64, lda = hx.alloc_colmajor(n, n, torch.float64, ctx)
    A32, lda32 = hx.alloc_colmajor(n, n, torch.float32, ctx)
    B, ldb = hx.alloc_colmajor(n, args.nrhs, torch.float64, ctx)
    X, ldx = hx.alloc_colmajor(n, args.nrhs, torch.float64, ctx)
    with torch.cuda.stream(ctx.stream):
        ipiv = torch.empty(n, dtype=torch.int32, device=ctx.device)
    S_f = np.asfortranarray(S)
    ptr = hx._ptr
    rep = _lib.IRReport()
    info = __import__("ctypes").c_int64(0)

    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def step(record=None):
        e0, e1, e2 = ev(), ev(), ev()
        e0.record(ctx.stream)
        ctx.check(lib.hfpx_assemble_mixed(ctx.handle, dm.handle, G, nu, ptr(A64), lda, ptr(A32), lda32))
        with torch.cuda.stream(ctx.stream):
            B.zero_()
        ctx.check(lib.hfpx_rhs_sinf(ctx.handle, dm.handle, hx._np_ptr(S_f), args.nrhs, ptr(B), ldb))
        e1.record(ctx.stream)
        rc = lib.hfpx_getrf(ctx.handle, _lib.HFPX_F32, ptr(A32), lda32, n, ptr(ipiv), __import__("ctypes").byref(info))
        ctx.check(rc)
        rc = lib.hfpx_solve_ir(ctx.handle, ptr(A64), lda, ptr(A32), lda32, ptr(ipiv), n, ptr(B), ldb, ptr(X), ldx,
                               args.nrhs, args.tol, 10, 20, __import__("ctypes").byref(rep))
        if rc not in (_lib.HFPX_OK, _lib.HFPX_ERR_NOT_CONVERGED):
            ctx.check(rc)
        e2.record(ctx.stream)
        if record is not None:
            record.append((e0, e1, e2, rep.iterations, rep.inner_iterations, rep.backward_error, rep.converged))

    for w in range(args.warmup):
        step()
        ctx.stream.synchronize()
        log(f"bench: warmup {w + 1}/{args.warmup} done")

    # timed steps: HIP events only around the dominant kernel class (the
    # trailing-update GEMM), for its live roofline
    ctx.check(lib.hfpx_profile_classes(ctx.handle, 1 << _lib.KERNEL_CLASSES["gemm_f32"]))
    ctx.check(lib.hfpx_profile(ctx.handle, 0 if args.no_lib_prof else 1))
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(recs)
        ctx.stream.synchronize()
        log(f"bench: step {s + 1}/{args.steps} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    g = read_profile(ctx, args.steps)["gemm_f32"]
    ctx.check(lib.hfpx_profile(ctx.handle, 0))
    # per-class breakdown from one more, untimed step with every class timed
    prof = {}
    if not args.no_lib_prof:
        ctx.check(lib.hfpx_profile_classes(ctx.handle, 0xFFFFFFFF))
        ctx.check(lib.hfpx_profile(ctx.handle, 1))
        step(None)
        ctx.stream.synchronize()
        prof = read_profile(ctx, 1)
        ctx.check(lib.hfpx_profile(ctx.handle, 0))
    t_asm = np.mean([a.elapsed_time(b) for a, b, *_ in recs]) / 1e3
    t_solve = np.mean([b.elapsed_time(c) for _, b, c, *_ in recs]) / 1e3
    last = recs[-1]
    log(f"bench: assembly {t_asm * 1e3:.1f} ms, solve {t_solve:.3f} s, IR {last[3]} outer / {last[4]} inner, "
        f"backward error {last[5]:.2e}")

    # dominant kernel by time: the FP32 trailing-update GEMM of the LU
    gemm_tf = g["flops"] / (g["ms"] / 1e3) / 1e12 if g["ms"] > 0 else 0.0
    a = prof.get("assemble", dict(ms=0.0, bytes=0.0))
    asm_gbs = a["bytes"] / (a["ms"] / 1e3) / 1e9 if a["ms"] > 0 else 0.0
    roofline = {"kernel": "FP32 trailing-update GEMM of the LU: k_split_a + k_split_b + k_x3gemm (bf16 MFMA "
                          "32x32x16, six products of three-way operand splits, FP32 accumulation) on the large "
                          "updates, k_x3gemm_f (the split fused into the product) on the mid-size ones; products "
                          "below 256 x 128 x 64 on FP32 MFMA (k_sgemm32 / k_gemm_sub<float>)",
                "bound": "mfma", "achieved": round(gemm_tf, 3), "peak": round(X3_PEAK, 1), "unit": "TFLOP/s",
                "frac": round(gemm_tf / X3_PEAK, 4),
                "peak_basis": "FP32-equivalent: dense bf16 MFMA 2500 TFLOP/s / 6 products (FP32 MFMA peak 157.3)",
                "achieved_bf16_mfma_TFs": round(gemm_tf * X3_PRODUCTS, 1),
                # PMC-measured bytes of one representative call (profiles/gemm72k_pmc.md)
                "traffic": GEMM72K_TRAFFIC, "traffic_algorithmic": GEMM72K_ALG_BYTES,
                "traffic_for": "first trailing update 70976 x 70976 x 1024 (bench.py --workload gemm72k): "
                               "FETCH_SIZE x2.0 + WRITE_SIZE x1.0, scales from a 2 GiB device copy",
                "launches_per_step": g["launches"], "avg_launch_ms": g["ms"] / max(1, g["launches"]),
                "kernel_ms_per_step": g["ms"]}
    pairs = 6.0 * n_el * n_el
    asm_tf = VALU_FLOPS_PER_PAIR["assemble72k"] * pairs / (a["ms"] / 1e3) / 1e12 if a["ms"] > 0 else 0.0
    assembly_roofline = {"kernel": "k_assemble (FP64 element integrals, FP64 + FP32 matrix write)", "bound": "valu",
                         "achieved": round(asm_tf, 2), "peak": PEAKS["fp64_valu_TFs"], "unit": "TFLOP/s",
                         "frac": round(asm_tf / PEAKS["fp64_valu_TFs"], 4), "kernel_ms": a["ms"],
                         "pairs": pairs, "flops_per_pair": VALU_FLOPS_PER_PAIR["assemble72k"],
                         "hbm_write_GBs": round(asm_gbs, 2), "hbm_frac": round(asm_gbs / PEAKS["hbm_GBs"], 4),
                         "note": "FP64 VALU flops per pair from hardware counters (VALU_FLOPS_PER_PAIR); the "
                                 "matrix write is the secondary HBM bound"}

    result = {
        "metric": "assembly Gentries/s; solve s at 72k DOF, FP32-IR vs FP64 (rel err<=1e-10)",
        "value": round(n * n / t_asm / 1e9, 3),
        "unit": "Gentries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 assembly; f32 LU (trailing updates as FP32-accurate products on bf16 MFMA, three-way "
                 "splits) + f64 refinement",
        "data": "synthetic",
        "config": {"workload": f"rotated elliptical crack 1x0.5 (30 deg about x), {n_el} quadratic triangles, "
                               f"{n} DoF, {args.nrhs} load cases, E=1 nu=0.3",
                   "n_dof": n, "nrhs": args.nrhs, "solver": "FP32-MFMA LU + FP64 GMRES-IR"},
        "assembly_s": round(t_asm, 5),
        "solve_s": round(t_solve, 4),
        "ir": {"outer": last[3], "inner": last[4], "backward_error": last[5], "converged": bool(last[6])},
        "kernel_profile_note": "per kernel class, from one untimed step after the timed ones (HIP events around "
                               "every launch); the timed steps carry events around the GEMM class only",
        "kernel_profile_per_step": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                                    for k, v in prof.items()},
        "roofline": roofline,
        "assembly_roofline": assembly_roofline,
    }

    if not args.no_fp64_compare:
        # plain FP64 GPU path on a copy: FP64 LU + getrs; error of the mixed solution against it
        log("bench: FP64 comparison path")
        A2, _ = hx.alloc_colmajor(n, n, torch.float64, ctx)
        X64, _ = hx.alloc_colmajor(n, args.nrhs, torch.float64, ctx)
        with torch.cuda.stream(ctx.stream):
            A2.copy_(A64)
            X64.copy_(B)
        e0, e1 = ev(), ev()
        # HIP events around the FP64 trailing-update GEMM launches only
        ctx.check(lib.hfpx_profile_classes(ctx.handle, 1 << _lib.KERNEL_CLASSES["gemm_f64"]))
        ctx.check(lib.hfpx_profile(ctx.handle, 0 if args.no_lib_prof else 1))
        e0.record(ctx.stream)
        ctx.check(lib.hfpx_getrf(ctx.handle, _lib.HFPX_F64, ptr(A2), n, n, ptr(ipiv), __import__("ctypes").byref(info)))
        ctx.check(lib.hfpx_getrs(ctx.handle, _lib.HFPX_F64, ptr(A2), n, ptr(ipiv), n, ptr(X64), n, args.nrhs))
        e1.record(ctx.stream)
        ctx.stream.synchronize()
        t64 = e0.elapsed_time(e1) / 1e3
        g64 = read_profile(ctx, 1)["gemm_f64"]
        ctx.check(lib.hfpx_profile(ctx.handle, 0))
        with torch.cuda.stream(ctx.stream):
            err = ((X - X64).abs().amax(dim=0) / X64.abs().amax(dim=0)).max().item()
        ctx.stream.synchronize()
        result["fp64_path"] = {"solve_s": round(t64, 4), "assembly_plus_solve_s": round(t64 + t_asm, 4),
                               "max_rel_err_mixed_vs_fp64": err,
                               "speedup_mixed_vs_fp64": round((t64 + t_asm) / (t_solve + t_asm), 3),
                               "getrf_TFs": round(2.0 / 3.0 * n ** 3 / t64 / 1e12, 2)}
        if g64["ms"] > 0:
            tf64 = g64["flops"] / (g64["ms"] / 1e3) / 1e12
            result["fp64_path"]["gemm_f64"] = {
                "kernel": "k_dgemm128 (v_mfma_f64_16x16x4f64, 128 x 128 tiles, two waves per SIMD) + small "
                          "products on k_gemm_sub<double>",
                "launches": g64["launches"], "ms": g64["ms"], "achieved_TFs": round(tf64, 2),
                "peak_TFs": PEAKS["fp64_mfma_TFs"], "frac": round(tf64 / PEAKS["fp64_mfma_TFs"], 4)}
        log(f"bench: FP64 path solve {t64:.3f} s; mixed vs FP64 max rel err {err:.2e}")
        del A2, X64

    if not args.no_cpu_baseline:
        log("bench: CPU baseline")
        try:
            cb = cpu_baseline(args.cpu_sample_el, G, nu, n, args.nrhs)
            result["cpu_baseline"] = cb
            if "solve_s_extrapolated" in cb:
                cpu_total = n * n / (cb["value"] * 1e9) + cb["solve_s_extrapolated"]
                cb["assembly_plus_solve_s_extrapolated"] = round(cpu_total, 1)
                result["speedup_vs_cpu_baseline"] = round(cpu_total / (t_asm + t_solve), 1)
        except Exception as ex:
            result["cpu_baseline"] = {"error": str(ex)}
    print(json.dumps(result))


def read_profile(ctx, steps):
    import ctypes
    from hfpx3d_amd import _lib
    prof = {}
    for name, kc in _lib.KERNEL_CLASSES.items():
        L, ms, fl, by = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ctx.check(ctx.lib.hfpx_profile_read(ctx.handle, kc, L, ms, fl, by))
        prof[name] = dict(launches=L.value // steps, ms=round(ms.value / steps, 4), flops=fl.value / steps,
                          bytes=by.value / steps)
    return prof


def run_gemm(args):
    """The dominant kernel alone: the first trailing update of the 72,000-DoF
    FP32 LU, C (70,976 x 70,976) -= A (70,976 x 1,024) B (1,024 x 70,976), on
    random data (for hardware counters and roofline)."""
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, _lib
    M = N = 72000 - 1024
    K = 1024
    ctx = hx.get_context(0)
    lib = ctx.lib
    f64 = args.gemm_dtype == "f64"
    if args.gemm_variant:
        hx.tune_set(_lib.KNOB_GEMM_F64 if f64 else _lib.KNOB_GEMM_F32, args.gemm_variant, ctx)
    calib_copy(args, ctx)
    dt = torch.float64 if f64 else torch.float32
    with torch.cuda.stream(ctx.stream):
        C = torch.randn(M * N, dtype=dt, device=ctx.device)
        A = torch.randn(M * K, dtype=dt, device=ctx.device)
        B = torch.randn(K * N, dtype=dt, device=ctx.device)
    ctx.stream.synchronize()
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    code = _lib.HFPX_F64 if f64 else _lib.HFPX_F32
    call = lambda: ctx.check(lib.hfpx_gemm_sub(ctx.handle, code, hx._ptr(C), M, hx._ptr(A), M,  # noqa: E731
                                               hx._ptr(B), K, M, N, K))
    for _ in range(args.warmup):
        call()
    ctx.stream.synchronize()
    times = []
    for _ in range(args.steps):
        e0, e1 = ev(), ev()
        e0.record(ctx.stream)
        call()
        e1.record(ctx.stream)
        ctx.stream.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
    t = float(np.mean(times))
    flops = 2.0 * M * N * K
    alg_bytes = (8.0 if f64 else 4.0) * (2.0 * M * N + M * K + K * N)
    if f64:
        peak, basis = PEAKS["fp64_mfma_TFs"], "FP64 MFMA"
    elif not args.gemm_variant or args.gemm_variant in (13, 14):
        peak, basis = X3_PEAK, "FP32-equivalent: dense bf16 MFMA 2500 TFLOP/s / 6 products"
    else:
        peak, basis = PEAKS["fp32_mfma_TFs"], "FP32 MFMA"
    name = "FP64" if f64 else "FP32"
    result = {"metric": f"{name} trailing-update GEMM TFLOP/s (72k LU, first update)", "value": round(flops / t / 1e12, 3),
              "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "ms_per_step": round(t * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
              "dtype": "f64" if f64 else ("f32 (bf16 MFMA on three-way splits)"
                                          if not args.gemm_variant or args.gemm_variant in (13, 14) else "f32"),
              "data": "synthetic",
              "config": {"workload": f"C ({M} x {N}) -= A ({M} x {K}) B ({K} x {N}), {name}"},
              "algorithmic_bytes": alg_bytes,
              "roofline": {"bound": "mfma", "achieved": round(flops / t / 1e12, 3), "peak": round(peak, 1),
                           "unit": "TFLOP/s", "frac": round(flops / t / 1e12 / peak, 4), "peak_basis": basis}}
    print(json.dumps(result))


def run_valu(args):
    """The FP64-VALU-bound kernels alone (no LU, for hardware counters):
    assemble72k = the mixed assembly (FP64 + FP32 matrix) and the 64
    right-hand sides of the headline ellipse; stress1m = stresses of the
    6,000-element wavy crack at 10^6 points of [-2, 2]^3 (seed 3) from a
    smooth synthetic opening (the solve is not timed here)."""
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, meshgen
    ctx = hx.get_context(0)
    lib = ctx.lib
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    G, nu = 1.0 / 2.6, 0.3
    if args.workload == "assemble72k":
        nods, conn, S = c3_problem(args.n_el, args.nrhs)
        dm = hx.DeviceMesh(hx.Mesh_Geom_T(nods=nods, conn=conn), None, 0.125, ctx)
        n = dm.n_dof
        A64, lda = hx.alloc_colmajor(n, n, torch.float64, ctx)
        A32, lda32 = hx.alloc_colmajor(n, n, torch.float32, ctx)
        B, ldb = hx.alloc_colmajor(n, args.nrhs, torch.float64, ctx)
        Sf = np.asfortranarray(S)

        def call():
            ctx.check(lib.hfpx_assemble_mixed(ctx.handle, dm.handle, G, nu, hx._ptr(A64), lda, hx._ptr(A32), lda32))
            with torch.cuda.stream(ctx.stream):
                B.zero_()
            ctx.check(lib.hfpx_rhs_sinf(ctx.handle, dm.handle, hx._np_ptr(Sf), args.nrhs, hx._ptr(B), ldb))
        units, unit, pairs = n * n / 1e9, "Gentries/s", 6.0 * conn.shape[1] ** 2
        wl = (f"assembly of the rotated elliptical crack 1x0.5, {conn.shape[1]} quadratic triangles, {n} DoF "
              f"(FP64 + FP32 matrices) and {args.nrhs} right-hand sides")
    else:
        nods, conn = meshgen.wavy_disk(6000)
        n_el = conn.shape[1]
        dm = hx.DeviceMesh(hx.Mesh_Geom_T(nods=nods, conn=conn), None, 0.0, ctx)
        rng = np.random.default_rng(3)
        pts = rng.uniform(-2, 2, size=(3, args.n_pts))
        xyz = np.zeros((6 * n_el, 3))
        for q in range(3):  # nodal positions of the DD rows (vertices, then mid-edge nodes)
            v = nods[:3, conn[q]].T
            xyz[q::6] = v
            xyz[3 + q::6] = 0.5 * (nods[:3, conn[(q + 1) % 3]].T + nods[:3, conn[(q + 2) % 3]].T)
        dd = np.zeros((6 * n_el, 3))
        dd[:, 2] = np.sqrt(np.clip(1 - (xyz[:, 0] ** 2 + xyz[:, 1] ** 2), 0, None))
        with torch.cuda.stream(ctx.stream):
            d_dd = torch.from_numpy(np.ascontiguousarray(dd.T)).to(ctx.device)
            d_pts = torch.from_numpy(np.ascontiguousarray(pts.T)).to(ctx.device)
            out = torch.empty((6, args.n_pts), dtype=torch.float64, device=ctx.device)
        s_inf = np.array([1.0, 1.0, 0.5, 0, 0, 0])

        def call():
            ctx.check(lib.hfpx_stress(ctx.handle, dm.handle, G, nu, hx._ptr(d_dd), 6 * n_el, hx._ptr(d_pts),
                                      args.n_pts, hx._np_ptr(s_inf), hx._ptr(out), args.n_pts))
        units, unit, pairs = args.n_pts * n_el / 1e9, "G pair evaluations/s", float(args.n_pts) * n_el
        wl = (f"stresses of the wavy crack z=0.1 sin(2 pi x) sin(2 pi y), {n_el} quadratic triangles, at "
              f"{args.n_pts} points uniform in [-2, 2]^3 (seed 3)")
    for _ in range(args.warmup):
        call()
    ctx.stream.synchronize()
    times = []
    for _ in range(args.steps):
        e0, e1 = ev(), ev()
        e0.record(ctx.stream)
        call()
        e1.record(ctx.stream)
        ctx.stream.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
    t = float(np.mean(times))
    result = {"metric": f"{args.workload} {unit}", "value": round(units / t, 3), "unit": unit, "n_gpus": 1,
              "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
              "data": "synthetic", "config": {"workload": wl}, "pair_evaluations": pairs}
    if VALU_FLOPS_PER_PAIR.get(args.workload):
        fl = VALU_FLOPS_PER_PAIR[args.workload] * pairs
        result["roofline"] = {"bound": "valu", "achieved": round(fl / t / 1e12, 3), "peak": PEAKS["fp64_valu_TFs"],
                              "unit": "TFLOP/s", "frac": round(fl / t / 1e12 / PEAKS["fp64_valu_TFs"], 4),
                              "flops_per_pair": VALU_FLOPS_PER_PAIR[args.workload],
                              "basis": "FP64 VALU flops per pair evaluation from SQ_INSTS_VALU_FLOPS_FP64 "
                                       "(+ _TRANS) of the kernel, the step's time (includes geometry / RHS)"}
    print(json.dumps(result))


def run_batch(args):
    """256 independent planar elliptical cracks, 40-500 elements each (seed 2),
    semi-axes U[0.5, 2], random orientation, E = 1, nu U[0.2, 0.4], pressure
    p U[0.5, 2] (s_inf = p I).  One step = batched assembly + right-hand
    sides + batched LU (FP64 below --mixed-min-n DoF, FP32 + 3 FP64
    refinement steps above) + solves + backward errors, replayed as one
    hipGraph."""
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, meshgen
    cracks = []
    for c in meshgen.batch_cracks(256, 2, 40, 500):
        p = c["p"]
        cracks.append(dict(nods=c["nods"], conn=c["conn"], G=1.0 / (2.0 * (1.0 + c["nu"])), nu=c["nu"],
                           s_inf=np.array([p, p, p, 0.0, 0.0, 0.0])))
    ns = np.array([18 * c["conn"].shape[1] for c in cracks], dtype=np.float64)
    entries, flops = float((ns ** 2).sum()), float((2.0 / 3.0 * ns ** 3).sum())
    ctx = hx.get_context(0)
    if args.autotune:
        from hfpx3d_amd import autotune
        m = autotune.apply_batch(autotune.load(args.autotune), ctx)
        if m is None:
            sys.exit(f"bench: {args.autotune} has no batch row (python -m hfpx3d_amd.autotune --batch 256)")
        args.mixed_min_n = m
        log(f"bench: autotune batch row: BGEMM_F32 {hx.tune_get(autotune.BATCH_KNOBS['BGEMM_F32'], ctx)}, "
            f"mixed_min_n {m}")
    n_mixed = int((ns >= args.mixed_min_n).sum())
    log(f"bench: batch of {len(cracks)} cracks, n {int(ns.min())}..{int(ns.max())}, {entries:.3e} entries, "
        f"{n_mixed} mixed")
    bt = hx.Batch(cracks, mixed_min_n=args.mixed_min_n, ctx=ctx)
    for w in range(max(1, args.warmup)):
        bt.solve(n_refine=3, use_graph=True)  # the first call also captures the graph
        log(f"bench: warmup {w + 1} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        bt.solve(n_refine=3, use_graph=True)
        log(f"bench: step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    berrs, infos = [], []
    for i in range(len(cracks)):
        _, be, inf = bt.solution(i)
        berrs.append(be)
        infos.append(inf)
    # spot check: three cracks against the single-crack FP64 GPU path
    errs = {}
    for i in (int(np.argmin(ns)), int(np.argmax(ns)), len(cracks) // 2):
        c = cracks[i]
        mesh = hx.Mesh_Geom_T(nods=c["nods"], conn=c["conn"])
        md = hx.Mesh_Data_T(mesh=mesh)
        sae = hx.SAE_T()
        sae.matrix = hx.make_3dbem_matrix_s(c["G"], c["nu"], mesh, hx.Num_Param_T(beta=0.125, tip_type=0),
                                            md.dof_h_dd)
        hx.add_s_inf_to_3dbem_rhs(md, hx.Load_T(s_inf=c["s_inf"]), sae)
        xr = hx.LU(sae.matrix).solve(sae.rhs_v).cpu().numpy()
        x = bt.solution(i)[0]
        errs[f"crack{i}_n{int(ns[i])}"] = float(np.abs(x - xr).max() / np.abs(xr).max())
        del sae
    result = {
        "metric": "batch of 256 cracks: matrix entries/s over assembly + LU + solve",
        "value": round(entries / t / 1e9, 3), "unit": "Gentries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 assembly; f64 LU (small cracks), f32 LU + f64 refinement (large)", "data": "synthetic",
        "config": {"workload": f"256 planar elliptical cracks, 40-500 quadratic triangles each (seed 2), random "
                               f"orientation, nu U[0.2,0.4], p U[0.5,2]; {n_mixed} cracks with n >= "
                               f"{args.mixed_min_n} in FP32 + refinement", "n_cracks": len(cracks),
                   "sum_n2": entries, "sum_2n3_3": flops},
        "batch_s": round(t, 4), "lu_flop_rate_TFs": round(flops / t / 1e12, 2),
        "max_backward_error": max(berrs), "singular_cracks": int(sum(1 for v in infos if v)),
        "max_rel_err_vs_single_fp64": errs,
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (reference assembly + LAPACK on three cracks, scaled)")
        import scipy.linalg as sl
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        ref = oracle_lib.reference_or_none()
        threads = max(1, min(16, os.cpu_count() or 1))
        ta = tl = e_s = f_s = 0.0
        for i in np.argsort(ns)[[0, len(ns) // 2, -1]]:
            c = cracks[int(i)]
            tt = time.perf_counter()
            if ref is not None:
                A, _ = ref.assemble(c["nods"], c["conn"], c["G"], c["nu"], 0.125, 0, nthreads=threads)
            else:
                A, _ = oracle_lib.oracle().assemble(c["nods"], c["conn"], c["G"], c["nu"], 0.125, 0, nthreads=threads)
            ta += time.perf_counter() - tt
            tt = time.perf_counter()
            sl.lu_factor(A, overwrite_a=True, check_finite=False)
            tl += time.perf_counter() - tt
            e_s += float(ns[i]) ** 2
            f_s += 2.0 / 3.0 * float(ns[i]) ** 3
        cpu_t = ta * entries / e_s + tl * flops / f_s
        result["cpu_baseline"] = {"value": round(entries / cpu_t / 1e9, 4), "unit": "Gentries/s", "cores": threads,
                                  "kind": "reference" if ref is not None else "port",
                                  "sample": f"reference assembly + LAPACK LU of the smallest, median and largest "
                                            f"crack ({ta:.1f} s + {tl:.1f} s), scaled by sum n^2 and sum n^3 to "
                                            f"{cpu_t:.0f} s for the batch"}
    print(json.dumps(result))


def vc_handles(nods, conn, R, active):
    """Original handle: every DoF of the ellipse free (18 e + 3 node + comp).
    Used handle: the DoF of elements whose centroid (in the ellipse's own
    frame, semi-axes 1 and 0.5) lies inside `active` of the ellipse, numbered
    in element order -- the shape of the truncated systems of a volume-control
    time step, whose front region is fixed."""
    import numpy as np
    from hfpx3d_amd import hfp3d as hx
    n_el = conn.shape[1]
    orig = hx.DoF_Handle_T(n_dof=18 * n_el, dof_h=np.arange(18 * n_el, dtype=np.int64).reshape(n_el, 18))
    c = (R.T @ nods[:3]).T[conn[:3].T].mean(axis=1)
    inside = (c[:, 0] / 1.0) ** 2 + (c[:, 1] / 0.5) ** 2 < active ** 2
    dh = -np.ones((n_el, 18), dtype=np.int64)
    k = int(inside.sum())
    dh[inside] = np.arange(18 * k, dtype=np.int64).reshape(k, 18)
    return orig, hx.DoF_Handle_T(n_dof=18 * k, dof_h=dh)


def calib_copy(args, ctx):
    """--calib-copy-gib: one device-to-device copy of that size on the
    library stream, the known-size kernel that scales the PMC byte
    counters of the same run."""
    if args.calib_copy_gib <= 0:
        return
    import torch
    cnt = int(args.calib_copy_gib * 2 ** 30) // 4
    with torch.cuda.stream(ctx.stream):
        src = torch.ones(cnt, dtype=torch.float32, device=ctx.device)
        dst = torch.empty_like(src)
        dst.copy_(src)
    ctx.stream.synchronize()
    log(f"bench: calibration copy of {cnt * 4} bytes done")
    del src, dst


def run_vc(args):
    """Volume-control steps on the 72k ellipse (4,000 elements, E = 1,
    nu = 0.3): the (72,001)^2 volume-control matrix is assembled once and
    stays resident; one step gathers the truncated system of the active DoF
    (FP64 + FP32 in one pass), factorises it in FP32 and solves for a volume
    increment with FP64 GMRES-IR.  The gather is HBM-bound: per entry of the
    truncated system 8 bytes read, 8 + 4 bytes written.  vctrunc72k: one step
    is the truncation call alone (hfpx_vc_truncate, the reference's
    mod_3dbem_system_vc) on the same system."""
    import ctypes
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, _lib, meshgen
    E_mod, nu = 1.0, 0.3
    G = E_mod / (2 * (1 + nu))
    R = meshgen.rotation([1.0, 0.0, 0.0], np.pi / 6)
    nods, conn = meshgen.ellipse(args.n_el, 1.0, 0.5, R)
    n_el = conn.shape[1]
    orig, used = vc_handles(nods, conn, R, args.vc_active)
    n0, n1 = orig.n_dof, used.n_dof
    ctx = hx.get_context(0)
    lib, ptr = ctx.lib, hx._ptr
    log(f"bench: vc n_el={n_el} orig n={n0} active n={n1}")
    calib_copy(args, ctx)
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    dm = hx.DeviceMesh(mesh, None, 0.125, ctx)
    A, lda = hx.alloc_colmajor(n0 + 1, n0 + 1, torch.float64, ctx)
    nu1 = n1 + 1
    M, ldm = hx.alloc_colmajor(nu1, nu1, torch.float64, ctx)
    M32, ldm32 = hx.alloc_colmajor(nu1, nu1, torch.float32, ctx)
    with torch.cuda.stream(ctx.stream):
        rhs = torch.empty(nu1, dtype=torch.float64, device=ctx.device)
        X = torch.empty(nu1, dtype=torch.float64, device=ctx.device)
        ipiv = torch.empty(nu1, dtype=torch.int32, device=ctx.device)
    od = np.asfortranarray(orig.dof_h)
    ud = np.asfortranarray(used.dof_h)
    dt = np.zeros(n1)
    rep = _lib.IRReport()
    info = ctypes.c_int64(0)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731
    e0, e1 = ev(), ev()
    e0.record(ctx.stream)
    ctx.check(lib.hfpx_assemble_vc(ctx.handle, dm.handle, G, nu, ptr(A), lda, None, 0))
    e1.record(ctx.stream)
    ctx.stream.synchronize()
    t_asm = e0.elapsed_time(e1) / 1e3
    log(f"bench: vc matrix assembled in {t_asm * 1e3:.1f} ms")

    solve = args.workload == "vc72k"

    def step(rec=None):
        e = [ev() for _ in range(3)]
        e[0].record(ctx.stream)
        ctx.check(lib.hfpx_vc_truncate(ctx.handle, ptr(A), lda, n0, hx._np_ptr(od), hx._np_ptr(ud), n_el, 18, n1,
                                       hx._np_ptr(dt), n1, 1e-3, ptr(M), ldm, ptr(M32), ldm32, ptr(rhs)))
        e[1].record(ctx.stream)
        if not solve:
            e[2].record(ctx.stream)
            if rec is not None:
                rec.append((e, 0, 0, None, None))
            return
        ctx.check(lib.hfpx_getrf(ctx.handle, _lib.HFPX_F32, ptr(M32), ldm32, nu1, ptr(ipiv), ctypes.byref(info)))
        rc = lib.hfpx_solve_ir(ctx.handle, ptr(M), ldm, ptr(M32), ldm32, ptr(ipiv), nu1, ptr(rhs), nu1, ptr(X), nu1,
                               1, args.tol, 10, 20, ctypes.byref(rep))
        if rc not in (_lib.HFPX_OK, _lib.HFPX_ERR_NOT_CONVERGED):
            ctx.check(rc)
        e[2].record(ctx.stream)
        if rec is not None:
            rec.append((e, rep.iterations, rep.inner_iterations, rep.backward_error, bool(rep.converged)))

    for w in range(args.warmup):
        step()
        ctx.stream.synchronize()
        log(f"bench: warmup {w + 1}/{args.warmup} done")
    ctx.check(lib.hfpx_profile(ctx.handle, 0 if args.no_lib_prof else 1))
    ctx.check(lib.hfpx_profile_classes(ctx.handle, 1 << _lib.KERNEL_CLASSES["vc"]))
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(recs)
        ctx.stream.synchronize()
        log(f"bench: step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launches, ms, _, nbytes = (ctypes.c_int64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double())
    ctx.check(lib.hfpx_profile_read(ctx.handle, _lib.KERNEL_CLASSES["vc"], ctypes.byref(launches), ctypes.byref(ms),
                                    ctypes.byref(_), ctypes.byref(nbytes)))
    ctx.check(lib.hfpx_profile(ctx.handle, 0))
    t_trunc = np.mean([r[0][0].elapsed_time(r[0][1]) for r in recs]) / 1e3
    if launches.value > 0:
        t_gather = ms.value / launches.value / 1e3
        gbps = nbytes.value / launches.value / t_gather / 1e9
    else:  # --no-lib-prof: the whole truncation call, algorithmic bytes
        t_gather = t_trunc
        gbps = 20.0 * (n1 + 1) ** 2 / t_gather / 1e9
    t_solve = np.mean([r[0][1].elapsed_time(r[0][2]) for r in recs]) / 1e3
    vol = None
    if solve:
        with torch.cuda.stream(ctx.stream):
            vol = torch.dot(M[n1, :n1], X[:n1]).item()
    last = recs[-1]
    log(f"bench: gather {t_gather * 1e3:.2f} ms ({gbps:.0f} GB/s), truncate call {t_trunc * 1e3:.1f} ms, "
        f"solve {t_solve:.3f} s, volume {vol}")
    entries = float(nu1) * nu1
    result = {
        "metric": ("vc 72k: truncated-system gather Gentries/s; step s" if solve else
                   "vc 72k truncation (mod_3dbem_system_vc): gather Gentries/s"),
        "value": round(entries / t_gather / 1e9, 3), "unit": "Gentries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 gather (+ f32 rounding); f32 LU + f64 refinement", "data": "synthetic",
        "config": {"workload": f"volume control on the rotated ellipse (a=1, b=0.5), {n_el} quadratic triangles, "
                               f"{n0 + 1} x {n0 + 1} resident matrix, {nu1} x {nu1} truncated system "
                               f"(elements inside {args.vc_active} of the ellipse active), dV = 1e-3",
                   "n_dof": n0, "n_active": n1},
        "vc_assembly_s": round(t_asm, 5), "truncate_call_s": round(t_trunc, 5), "gather_kernel_s": round(t_gather, 5),
        "solve_s": round(t_solve, 4) if solve else None, "step_s": round(t_trunc + t_solve, 4),
        "ir": {"outer": last[1], "inner": last[2], "backward_error": last[3], "converged": last[4]} if solve else None,
        "volume_check": {"v_dot_d": vol, "prescribed": 1e-3} if solve else None,
        "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbps / 8000.0, 4),
                     "traffic": VC72K_TRAFFIC if (args.n_el, args.vc_active) == (4000, 0.9) else None,
                     "traffic_for": "one gather of the default workload (--n-el 4000 --vc-active 0.9): "
                                    "FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, scales from a 2 GiB device copy",
                     "kernel": "k_vc_gather (FP64 + FP32 outputs)",
                     "achieved_basis": "algorithmic 20 bytes per truncated-system entry (8 read, 8 + 4 written) "
                                       "/ gather duration from HIP events on the library stream"},
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (reference mod_3dbem_system_vc on a smaller crack)")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        ref = oracle_lib.reference_or_none()
        threads = max(1, min(16, os.cpu_count() or 1))
        ns = 1000
        nods_s, conn_s = meshgen.ellipse(ns, 1.0, 0.5, R)
        o_s, u_s = vc_handles(nods_s, conn_s, R, args.vc_active)
        As = np.asfortranarray(np.random.default_rng(5).normal(size=(o_s.n_dof + 1, o_s.n_dof + 1)))
        dts = np.zeros(u_s.n_dof)
        t = time.perf_counter()
        if ref is not None:
            ref.mod_system_vc(As, o_s.dof_h, u_s.dof_h, u_s.n_dof, dts, 1e-3, nthreads=threads)
            tc = ref.last_seconds   # the reference call alone, without the wrapper's copies
        else:
            oracle_lib.oracle().mod_system_vc(As, o_s.dof_h, u_s.dof_h, u_s.n_dof, dts, 1e-3)
            tc = time.perf_counter() - t
        result["cpu_baseline"] = {"value": round((u_s.n_dof + 1) ** 2 / tc / 1e9, 4), "unit": "Gentries/s",
                                  "cores": threads if ref is not None else 1,
                                  "kind": "reference" if ref is not None else "port",
                                  "sample": f"mod_3dbem_system_vc on the same ellipse meshed with {ns} elements "
                                            f"({o_s.n_dof + 1}^2 -> {u_s.n_dof + 1}^2), {tc:.1f} s"}
    print(json.dumps(result))


def cf_slip_problem(R):
    """Far field on the rotated ellipse: compression 1 across the crack plane
    and shear 0.3 along the ellipse's own x axis (sigma = -n n + 0.3 (n t +
    t n), n the crack normal, t the in-plane axis); constant friction 0.2,
    no cohesion, slip already started everywhere (mr_slip 0.5): every shear
    DoF slips and every normal DoF stays in the system, so the active set of
    every step is that of the first."""
    import numpy as np
    n, t = R @ np.array([0.0, 0.0, 1.0]), R @ np.array([1.0, 0.0, 0.0])
    S = -np.outer(n, n) + 0.3 * (np.outer(n, t) + np.outer(t, n))
    s_inf = np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])
    par = np.array([[1e-9, 0.0, 1.0, 0.0, 0.2, 0.2]])
    return s_inf, par


def run_cf(args):
    """Steps of the cohesion/friction volume-control iteration
    (hfpx_vc_cf_iteration) on the 72k ellipse (4,000 elements, E = 1,
    nu = 0.3) with the (72,001)^2 volume-control matrix resident.  The first
    (warm-up) step gathers the truncated system and factorises it in FP32;
    the timed steps keep the active set and reuse those factors: per step
    the traction product of the current DD against the resident matrix (HBM:
    8 bytes per entry), the node laws, the renumbering, and FP64 GMRES-IR on
    the truncated system."""
    import ctypes
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, _lib, meshgen
    E_mod, nu = 1.0, 0.3
    G = E_mod / (2 * (1 + nu))
    R = meshgen.rotation([1.0, 0.0, 0.0], np.pi / 6)
    nods, conn = meshgen.ellipse(args.n_el, 1.0, 0.5, R)
    n_el = conn.shape[1]
    n_nod = 6 * n_el
    n0 = 18 * n_el
    ctx = hx.get_context(0)
    lib, ptr = ctx.lib, hx._ptr
    s_inf, par = cf_slip_problem(R)
    log(f"bench: cf n_el={n_el} n={n0}")
    mesh = hx.Mesh_Geom_T(nods=nods, conn=conn)
    # DD in element frames: the laws read the opening and slip from the DD
    # components as given, which are those of the crack plane only in
    # element frames for a crack that is not in z = 0
    dm = hx.DeviceMesh(mesh, None, 0.125, ctx, dd_local=True)
    A, lda = hx.alloc_colmajor(n0 + 1, n0 + 1, torch.float64, ctx)
    ctx.check(lib.hfpx_assemble_vc(ctx.handle, dm.handle, G, nu, ptr(A), lda, None, 0))
    h = ctypes.c_void_p()
    ctx.check(lib.hfpx_cf_create(ctx.handle, dm.handle, ptr(A), lda, ctypes.byref(h)))
    with torch.cuda.stream(ctx.stream):
        dd = torch.zeros(3 * n_nod, dtype=torch.float64, device=ctx.device)
        pp = torch.zeros(n_nod, dtype=torch.float64, device=ctx.device)
        mid = torch.zeros(n_nod, dtype=torch.int32, device=ctx.device)
        mo = torch.zeros(n_nod, dtype=torch.float64, device=ctx.device)
        ms = torch.full((n_nod,), 0.5, dtype=torch.float64, device=ctx.device)
        fc, sc, oc = (torch.zeros(n_nod, dtype=torch.float64, device=ctx.device) for _ in range(3))
    ctx.stream.synchronize()
    fpar = (_lib.FcParam * 1)(_lib.FcParam(*par[0]))
    cm = np.zeros((1, 2), dtype=np.int32)
    dh = np.asfortranarray(np.arange(n0, dtype=np.int64).reshape(n_el, 18))
    nd = ctypes.c_int64(n0)
    si = np.ascontiguousarray(s_inf)
    rep = _lib.CfReport()

    def step():
        t = time.perf_counter()
        rc = lib.hfpx_vc_cf_iteration(ctx.handle, h, ptr(mid), hx._np_ptr(cm), fpar, 1, hx._np_ptr(si), 0, 0, 0.0,
                                      ptr(dd), n_nod, ptr(pp), ptr(mo), ptr(ms), ptr(fc), ptr(sc), ptr(oc),
                                      hx._np_ptr(dh), ctypes.byref(nd), args.tol, 10, ctypes.byref(rep))
        ctx.check(rc)
        return time.perf_counter() - t, rep.refactored, rep.residual, rep.ir.iterations, rep.ir.backward_error

    first = step()
    log(f"bench: first step (gather + FP32 LU + refinement) {first[0]:.3f} s, residual {first[2]:.4g}")
    for w in range(max(0, args.warmup - 1)):
        step()
        log(f"bench: warmup {w + 2}/{args.warmup} done")
    ctx.check(lib.hfpx_profile(ctx.handle, 0 if args.no_lib_prof else 1))
    ctx.check(lib.hfpx_profile_classes(ctx.handle, 1 << _lib.KERNEL_CLASSES["vc"]))
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        recs.append(step())
        log(f"bench: step {k + 1}/{args.steps} done ({recs[-1][0]:.3f} s, residual {recs[-1][2]:.3g})")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launches, ms_, fl, nbytes = (ctypes.c_int64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double())
    ctx.check(lib.hfpx_profile_read(ctx.handle, _lib.KERNEL_CLASSES["vc"], ctypes.byref(launches),
                                    ctypes.byref(ms_), ctypes.byref(fl), ctypes.byref(nbytes)))
    ctx.check(lib.hfpx_profile(ctx.handle, 0))
    lib.hfpx_cf_destroy(h)
    t_prod = ms_.value / max(1, launches.value) / 1e3
    gbps = nbytes.value / max(1, launches.value) / t_prod / 1e9 if t_prod > 0 else 0.0
    t_step = wall / args.steps
    reused = all(r[1] == 0 for r in recs)
    result = {
        "metric": "cf 72k: cohesion/friction iteration step s (cached FP32 factors)",
        "value": round(t_step, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 tractions and refinement; f32 LU of the truncated system", "data": "synthetic",
        "config": {"workload": f"cohesion/friction volume-control iteration on the rotated ellipse (a=1, b=0.5), "
                               f"{n_el} quadratic triangles, {n0 + 1} x {n0 + 1} resident matrix, DD in element "
                               f"frames; compression 1 and shear 0.3 on the crack plane, friction 0.2, every node "
                               f"slipping",
                   "n_dof": n0},
        "first_step_s": round(first[0], 4), "factors_reused": reused,
        "n_active": int(nd.value), "residual_per_step": [r[2] for r in recs],
        "ir": {"outer": recs[-1][3], "backward_error": recs[-1][4]},
        "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbps / 8000.0, 4), "traffic": None,
                     "kernel": "traction product of the step (current DD against the resident volume-control "
                               "matrix, k_dgemm_stream)",
                     "achieved_basis": "algorithmic 8 bytes per matrix entry / product duration from HIP events on "
                                       "the library stream", "product_ms": round(t_prod * 1e3, 3)},
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle step on a smaller crack)")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        O = oracle_lib.oracle()
        ns = 200
        nods_s, conn_s = meshgen.ellipse(ns, 1.0, 0.5, R)
        A_s, dof_s = O.assemble(nods_s, conn_s, G, nu, tip_type=0, dd_local=True)
        row, col = O.vc_border(nods_s, conn_s, dof_s, dd_local=True)
        m = A_s.shape[0]
        Avc = np.zeros((m + 1, m + 1), order="F")
        Avc[:m, :m], Avc[m, :m], Avc[:m, m] = A_s, row, col
        z = np.zeros(6 * ns)
        t = time.perf_counter()
        O.vc_cf_iteration(nods_s, conn_s, dof_s, Avc, np.zeros(6 * ns, np.int32), cm, par, z, np.full(6 * ns, 0.5),
                          s_inf, (0, 0), 0.0, np.zeros((6 * ns, 3)), z, dof_s.copy(), dd_local=True)
        tc = time.perf_counter() - t
        result["cpu_baseline"] = {"value": round(tc, 3), "unit": "s", "cores": 1, "kind": "port",
                                  "sample": f"one oracle step (FP64 LU of the truncated system, one thread) on the "
                                            f"same ellipse with {ns} elements ({m + 1}^2 system)"}
    print(json.dumps(result))


def run_wavy(args):
    """Non-planar crack z = 0.1 sin(2 pi x) sin(2 pi y) over the unit disk,
    6,000 elements (108,000 DoF), E = 1, nu = 0.3, p = 1 plus s_zz = -0.5
    (tension-positive s_inf = diag(1, 1, 0.5)); FP32-MFMA LU + FP64 GMRES-IR;
    stresses at 10^6 points uniform in [-2, 2]^3 (seed 3).  One step =
    assembly + right-hand side + factorisation + refinement + stresses."""
    import ctypes
    import numpy as np
    import torch
    from hfpx3d_amd import hfp3d as hx, _lib, meshgen
    G, nu = 1.0 / 2.6, 0.3
    nods, conn = meshgen.wavy_disk(6000)
    n_el = conn.shape[1]
    n = 18 * n_el
    s_inf = np.array([1.0, 1.0, 0.5, 0.0, 0.0, 0.0])
    ctx = hx.get_context(0)
    lib, ptr = ctx.lib, hx._ptr
    log(f"bench: wavy crack n_el={n_el} n={n} points={args.n_pts}")
    dm = hx.DeviceMesh(hx.Mesh_Geom_T(nods=nods, conn=conn), None, 0.125, ctx)
    A64, lda = hx.alloc_colmajor(n, n, torch.float64, ctx)
    A32, lda32 = hx.alloc_colmajor(n, n, torch.float32, ctx)
    B, _ = hx.alloc_colmajor(n, 1, torch.float64, ctx)
    X, _ = hx.alloc_colmajor(n, 1, torch.float64, ctx)
    pts_h = np.asfortranarray(np.random.default_rng(3).uniform(-2, 2, size=(3, args.n_pts)))
    with torch.cuda.stream(ctx.stream):
        ipiv = torch.empty(n, dtype=torch.int32, device=ctx.device)
        pts = torch.from_numpy(pts_h.T.copy()).to(ctx.device).t()  # 3 x P, column-major
        S = torch.empty((6, args.n_pts), dtype=torch.float64, device=ctx.device).t()  # P x 6, column-major
        dd = torch.zeros((3, 6 * n_el), dtype=torch.float64, device=ctx.device).t()
    s_f = np.asfortranarray(s_inf.reshape(6, 1))
    rep = _lib.IRReport()
    info = ctypes.c_int64(0)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def step(rec=None):
        e = [ev() for _ in range(5)]
        e[0].record(ctx.stream)
        ctx.check(lib.hfpx_assemble_mixed(ctx.handle, dm.handle, G, nu, ptr(A64), lda, ptr(A32), lda32))
        with torch.cuda.stream(ctx.stream):
            B.zero_()
        ctx.check(lib.hfpx_rhs_sinf(ctx.handle, dm.handle, hx._np_ptr(s_f), 1, ptr(B), n))
        e[1].record(ctx.stream)
        ctx.check(lib.hfpx_getrf(ctx.handle, _lib.HFPX_F32, ptr(A32), lda32, n, ptr(ipiv), ctypes.byref(info)))
        rc = lib.hfpx_solve_ir(ctx.handle, ptr(A64), lda, ptr(A32), lda32, ptr(ipiv), n, ptr(B), n, ptr(X), n, 1,
                               args.tol, 10, 20, ctypes.byref(rep))
        if rc not in (_lib.HFPX_OK, _lib.HFPX_ERR_NOT_CONVERGED):
            ctx.check(rc)
        e[2].record(ctx.stream)
        # identity DoF map: the solution vector is the nodal DD array row by row
        with torch.cuda.stream(ctx.stream):
            dd.copy_(X.reshape(6 * n_el, 3))
        e[3].record(ctx.stream)
        ctx.check(lib.hfpx_stress(ctx.handle, dm.handle, G, nu, ptr(dd), 6 * n_el, ptr(pts), args.n_pts,
                                  hx._np_ptr(s_inf), ptr(S), args.n_pts))
        e[4].record(ctx.stream)
        if rec is not None:
            rec.append((e, rep.iterations, rep.inner_iterations, rep.backward_error, bool(rep.converged)))

    for w in range(args.warmup):
        step()
        ctx.stream.synchronize()
        log(f"bench: warmup {w + 1}/{args.warmup} done")
    ctx.check(lib.hfpx_profile(ctx.handle, 0 if args.no_lib_prof else 1))
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(recs)
        ctx.stream.synchronize()
        log(f"bench: step {k + 1}/{args.steps} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = read_profile(ctx, args.steps)
    ctx.check(lib.hfpx_profile(ctx.handle, 0))
    t_asm = np.mean([r[0][0].elapsed_time(r[0][1]) for r in recs]) / 1e3
    t_solve = np.mean([r[0][1].elapsed_time(r[0][2]) for r in recs]) / 1e3
    t_str = np.mean([r[0][3].elapsed_time(r[0][4]) for r in recs]) / 1e3
    last = recs[-1]
    pairs = float(args.n_pts) * n_el
    log(f"bench: assembly {t_asm * 1e3:.1f} ms, solve {t_solve:.3f} s, stress {t_str:.3f} s")
    st = prof["stress"]
    result = {
        "metric": "wavy 108k: assembly Gentries/s; solve s; stress points/s",
        "value": round(n * n / t_asm / 1e9, 3), "unit": "Gentries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 assembly and stress; f32 LU + f64 refinement", "data": "synthetic",
        "config": {"workload": f"wavy crack z=0.1 sin(2 pi x) sin(2 pi y) over the unit disk, {n_el} quadratic "
                               f"triangles, {n} DoF, p=1 + s_zz=-0.5, {args.n_pts} stress points", "n_dof": n},
        "assembly_s": round(t_asm, 5), "solve_s": round(t_solve, 4),
        "ir": {"outer": last[1], "inner": last[2], "backward_error": last[3], "converged": last[4]},
        "stress_s": round(t_str, 4), "stress_points_per_s": round(args.n_pts / t_str, 1),
        "stress_pair_evals_per_s": round(pairs / t_str, 1),
        "kernel_profile_per_step": prof,
    }
    if args.fp64_compare:
        log("bench: FP64 comparison path")
        with torch.cuda.stream(ctx.stream):
            X64 = B.clone()
        del A32
        torch.cuda.empty_cache()
        e0, e1 = ev(), ev()
        e0.record(ctx.stream)
        ctx.check(lib.hfpx_getrf(ctx.handle, _lib.HFPX_F64, ptr(A64), lda, n, ptr(ipiv), ctypes.byref(info)))
        ctx.check(lib.hfpx_getrs(ctx.handle, _lib.HFPX_F64, ptr(A64), lda, ptr(ipiv), n, ptr(X64), n, 1))
        e1.record(ctx.stream)
        ctx.stream.synchronize()
        with torch.cuda.stream(ctx.stream):
            err = ((X - X64).abs().max() / X64.abs().max()).item()
        result["fp64_path"] = {"solve_s": round(e0.elapsed_time(e1) / 1e3, 4), "max_rel_err_mixed_vs_fp64": err}
        log(f"bench: FP64 path {result['fp64_path']}")
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (reference stress on a point subset)")
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        ref = oracle_lib.reference_or_none()
        sub = 2000
        dd_h = dd.cpu().numpy()
        threads = max(1, min(16, os.cpu_count() or 1))
        t = time.perf_counter()
        if ref is not None:
            ref.stress(nods, conn, G, nu, dd_h, pts_h[:, :sub].copy(), s_inf, nthreads=threads)
        else:
            oracle_lib.oracle().stress(nods, conn, G, nu, dd_h, pts_h[:, :sub].copy(), s_inf, nthreads=threads)
        dt = time.perf_counter() - t
        result["cpu_baseline"] = {"value": round(sub / dt, 1), "unit": "stress points/s", "cores": threads,
                                  "kind": "reference" if ref is not None else "port",
                                  "sample": f"{sub} of the points against all {n_el} elements, {dt:.1f} s"}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
    # free the library's contexts while torch and the HIP runtime are up
    if "hfpx3d_amd.hfp3d" in sys.modules:
        sys.modules["hfpx3d_amd.hfp3d"].close_all()
