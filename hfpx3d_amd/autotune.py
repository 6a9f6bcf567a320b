"""Autotuner for the LU's kernel knobs (no reference counterpart).

`sweep` times the factorisation on the device for each matrix size and
precision over a grid of knob settings (outer block of the LU, FP32 trailing-
update variant, panel kernel) and keeps the fastest; the result is a table
keyed by (n-range, precision), stored as JSON.  `use` attaches a table to a
context: `hfp3d.LU` then applies the matching row before it factorises.

    python -m hfpx3d_amd.autotune --out autotune.json --sizes 4096,16384,65536

A size s_k in the sweep covers the n-range between the geometric means of its
neighbours: [sqrt(s_{k-1} s_k), sqrt(s_k s_{k+1})), open at both ends.

`sweep_batch` tunes the batched form (hfp3d.Batch) on a given set of cracks:
the FP32 Schur-complement kernel of the batched LU (HFPX_KNOB_BGEMM_F32) and
the size from which a crack takes FP32 LU + FP64 refinement instead of an
FP64 LU (`mixed_min_n`).  Each candidate times whole replays of the batch's
captured graph.  The best setting is kept as the table's "batch" row;
`apply_batch` sets its knobs and returns its `mixed_min_n`.

    python -m hfpx3d_amd.autotune --out autotune.json --batch 256
"""
import argparse
import json
import math
import time

from . import _lib

KNOBS = {"GEMM_F32": _lib.KNOB_GEMM_F32, "LU_NB": _lib.KNOB_LU_NB, "PANEL": _lib.KNOB_PANEL,
         "X3_BAND": _lib.KNOB_X3_BAND}

# default grids: the FP32 trailing update on bf16 splits (0 = automatic) or
# on FP32 MFMA (6), and the outer block of the LU
DEFAULT_CANDIDATES = {
    "f32": [{"GEMM_F32": g, "LU_NB": nb} for g in (0, 6) for nb in (512, 1024, 2048)],
    "f64": [{"LU_NB": nb} for nb in (256, 512, 1024, 2048)],
}


# batched form: FP32 update kernel (1 = FP32 MFMA read-modify-write, 2 = FP32
# MFMA with C preloaded, 3 = bf16-split outer updates) x FP32 threshold
BATCH_KNOBS = {"BGEMM_F32": _lib.KNOB_BGEMM_F32}
DEFAULT_BATCH_CANDIDATES = [{"BGEMM_F32": b, "mixed_min_n": m} for b in (2, 3) for m in (512, 1024, 2048, 4096)]


def ranges(sizes):
    """(lo, hi) n-range of each sorted size: geometric midpoints between
    neighbours, 0 and infinity at the ends (hi None = unbounded)."""
    s = sorted(sizes)
    out = []
    for k, v in enumerate(s):
        lo = 0 if k == 0 else int(math.ceil(math.sqrt(s[k - 1] * v)))
        hi = None if k == len(s) - 1 else int(math.ceil(math.sqrt(v * s[k + 1])))
        out.append((lo, hi))
    return out


def lookup(table, n, precision):
    """Knob settings of the table row covering (n, precision), or None."""
    for e in table.get("entries", []):
        if e["precision"] == precision and n >= e["n_min"] and (e["n_max"] is None or n < e["n_max"]):
            return e["knobs"]
    return None


def apply(table, n, precision, ctx):
    """Set the knobs of the matching row on the context; returns them (or
    None when no row matches, leaving the context as it was)."""
    knobs = lookup(table, n, precision)
    if knobs is None:
        return None
    for name, value in knobs.items():
        ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, KNOBS[name], int(value)))
    return knobs


def save(table, path):
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def load(path):
    with open(path) as f:
        table = json.load(f)
    if "entries" not in table and "batch" not in table:
        raise ValueError(f"{path}: not an autotune table")
    return table


def apply_batch(table, ctx):
    """Set the knobs of the table's batch row on the context and return its
    mixed_min_n (None when the table has no batch row; the context is then
    left as it was)."""
    row = table.get("batch")
    if row is None:
        return None
    for name, value in row["knobs"].items():
        ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, BATCH_KNOBS[name], int(value)))
    return int(row["mixed_min_n"])


def use(table_or_path, ctx=None):
    """Attach a table (or the table in a JSON file) to the context."""
    from . import hfp3d
    ctx = ctx or hfp3d.get_context()
    ctx.autotune = load(table_or_path) if isinstance(table_or_path, str) else table_or_path
    return ctx.autotune


def _time_getrf(ctx, A0, code, reps):
    """Best time (ms) of hfpx_getrf on copies of A0, HIP events on the
    context's stream."""
    import ctypes
    from . import hfp3d
    torch = ctx.torch
    n = A0.shape[0]
    best = math.inf
    for _ in range(reps):
        with torch.cuda.stream(ctx.stream):
            A, lda = hfp3d.alloc_colmajor(n, n, A0.dtype, ctx)
            A.copy_(A0)
            ipiv = torch.empty(n, dtype=torch.int32, device=ctx.device)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(ctx.stream)
        info = ctypes.c_int64(0)
        ctx.check(ctx.lib.hfpx_getrf(ctx.handle, code, hfp3d._ptr(A), lda, n, hfp3d._ptr(ipiv), ctypes.byref(info)))
        e1.record(ctx.stream)
        ctx.stream.synchronize()
        best = min(best, e0.elapsed_time(e1))
        del A, ipiv
    return best


def sweep(sizes=(4096, 16384, 65536), precisions=("f32", "f64"), candidates=None, reps=2, ctx=None, log=None):
    """Time the LU for every size, precision and candidate setting; return
    the table of the fastest settings."""
    from . import hfp3d
    ctx = ctx or hfp3d.get_context()
    torch = ctx.torch
    candidates = candidates or DEFAULT_CANDIDATES
    saved = {name: hfp3d.tune_get(k, ctx) for name, k in KNOBS.items()}
    entries = []
    try:
        for n, (lo, hi) in zip(sorted(sizes), ranges(sizes)):
            for prec in precisions:
                dtype = torch.float32 if prec == "f32" else torch.float64
                code = _lib.HFPX_F32 if prec == "f32" else _lib.HFPX_F64
                g = torch.Generator(device=ctx.device)
                g.manual_seed(n)
                with torch.cuda.stream(ctx.stream):
                    A0, _ = hfp3d.alloc_colmajor(n, n, dtype, ctx)
                    A0.normal_(generator=g)
                    A0.diagonal().add_(math.sqrt(n))
                ctx.stream.synchronize()
                results = []
                for cand in candidates[prec]:
                    for name, k in KNOBS.items():  # every knob from its default, then the candidate's
                        ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, k, 0))
                    for name, value in cand.items():
                        ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, KNOBS[name], int(value)))
                    ms = _time_getrf(ctx, A0, code, reps)
                    results.append((ms, cand))
                    if log:
                        log(f"autotune: n={n} {prec} {cand}: {ms:.3f} ms")
                ms, best = min(results, key=lambda r: r[0])
                entries.append({"precision": prec, "n": n, "n_min": lo, "n_max": hi, "knobs": dict(best),
                                "ms": ms, "tried": [{"knobs": c, "ms": m} for m, c in results]})
                del A0
                torch.cuda.empty_cache()
    finally:
        for name, k in KNOBS.items():
            ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, k, saved[name]))
    return {"device": torch.cuda.get_device_name(ctx.device), "created": time.strftime("%Y-%m-%d %H:%M:%S"),
            "entries": entries}


def _time_batch(ctx, cracks, mixed_min_n, n_refine, reps):
    """Best time (ms) of one graph replay of the whole batched solve; the
    first call (which captures the graph) is not timed."""
    from . import hfp3d
    torch = ctx.torch
    bt = hfp3d.Batch(cracks, mixed_min_n=mixed_min_n, ctx=ctx)
    try:
        bt.solve(n_refine=n_refine, use_graph=True)
        best = math.inf
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(ctx.stream)
            bt.solve(n_refine=n_refine, use_graph=True)
            e1.record(ctx.stream)
            ctx.stream.synchronize()
            best = min(best, e0.elapsed_time(e1))
    finally:
        bt.close()
    return best


def sweep_batch(cracks, candidates=None, reps=2, n_refine=3, ctx=None, log=None):
    """Time the batched solve of `cracks` (the crack dicts hfp3d.Batch takes)
    for every candidate ({"BGEMM_F32": k, "mixed_min_n": m}); return a table
    whose "batch" row holds the fastest."""
    from . import hfp3d
    ctx = ctx or hfp3d.get_context()
    torch = ctx.torch
    candidates = candidates or DEFAULT_BATCH_CANDIDATES
    saved = {name: hfp3d.tune_get(k, ctx) for name, k in BATCH_KNOBS.items()}
    results = []
    try:
        for cand in candidates:
            for name, k in BATCH_KNOBS.items():
                ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, k, int(cand.get(name, 0))))
            ms = _time_batch(ctx, cracks, int(cand["mixed_min_n"]), n_refine, reps)
            results.append((ms, cand))
            if log:
                log(f"autotune: batch of {len(cracks)} {cand}: {ms:.3f} ms")
    finally:
        for name, k in BATCH_KNOBS.items():
            ctx.check(ctx.lib.hfpx_tune_set(ctx.handle, k, saved[name]))
    ms, best = min(results, key=lambda r: r[0])
    ns = [18 * int(c["conn"].shape[1]) for c in cracks]
    row = {"n_cracks": len(cracks), "n_min": min(ns), "n_max": max(ns), "n_refine": n_refine,
           "knobs": {k: int(best.get(k, 0)) for k in BATCH_KNOBS}, "mixed_min_n": int(best["mixed_min_n"]),
           "ms": ms, "tried": [{"knobs": c, "ms": m} for m, c in results]}
    return {"device": torch.cuda.get_device_name(ctx.device), "created": time.strftime("%Y-%m-%d %H:%M:%S"),
            "entries": [], "batch": row}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default="autotune.json")
    ap.add_argument("--sizes", default="4096,16384,65536")
    ap.add_argument("--precisions", default="f32,f64")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--batch", type=int, default=0,
                    help="also tune the batched form on this many synthetic elliptical cracks of 40-500 "
                         "elements (seed 2; 0 = skip)")
    ap.add_argument("--no-lu", action="store_true", help="skip the single-matrix LU sweep")
    args = ap.parse_args(argv)
    say = lambda m: print(m, flush=True)  # noqa: E731
    table = {"entries": []}
    if not args.no_lu:
        table = sweep([int(s) for s in args.sizes.split(",")], args.precisions.split(","), reps=args.reps, log=say)
    if args.batch:
        from . import meshgen
        cracks = []
        for c in meshgen.batch_cracks(args.batch, 2, 40, 500):
            p = c["p"]
            cracks.append(dict(nods=c["nods"], conn=c["conn"], G=1.0 / (2.0 * (1.0 + c["nu"])), nu=c["nu"],
                               s_inf=[p, p, p, 0.0, 0.0, 0.0]))
        bt = sweep_batch(cracks, reps=args.reps, log=say)
        table["batch"] = bt["batch"]
        table.setdefault("device", bt["device"])
        table.setdefault("created", bt["created"])
    save(table, args.out)
    print(f"autotune: wrote {args.out} ({len(table['entries'])} rows)", flush=True)


if __name__ == "__main__":
    main()
