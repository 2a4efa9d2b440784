"""Benchmark: closed-shell (T) time-to-solution on one MI355X.

One step = one full solve of E(T) = E[T] + E_ST (and v1) from the reference-layout input
arrays already resident in HBM: packing of the operand slabs, the persistent FP64 MFMA
launch over all i>=j>=k tuples, and the deterministic reductions.  Default workload:
no=24 occupied, nu=216 virtual (2024 distinct + 552 degenerate tuples), the size the
time-to-solution and matrix-peak targets are quoted at.  Inputs are the seeded synthetic
closed-shell generator (libcc_amd.synth, seed 1004).

Prints one JSON line.  `roofline` is for the persistent kernel (HIP events on the
library's stream): achieved = algorithmic GEMM FLOPs / kernel time, against MI355X's
published dense FP64 matrix peak.  `cpu_baseline` times the reference's own compiled
per-tuple routines (oracle/_ref, threaded MKL) on a bounded sample of the same tuples.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "(T) time-to-solution & matrix-peak %; split-prec vs FP64 at |ΔE(T)|<1 µEh"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP8_TFLOPS = 5000.0
# (peak of the matrix cores a form runs on, executed MFMA FLOPs per algorithmic FLOP): the split
# forms run one MFMA per slice product (BF16 x 3: 6, BF16 x 4 and FP8 x 4: 10)
FORM_PEAK = {"fp64": (PEAK_FP64_TFLOPS, 1), "fp32": (PEAK_FP32_TFLOPS, 1), "bf16x3": (PEAK_BF16_TFLOPS, 6),
             "bf16x4": (PEAK_BF16_TFLOPS, 10), "fp8x4": (PEAK_FP8_TFLOPS, 10)}
DTYPE = {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3/f32acc", "bf16x4": "bf16x4/f32acc", "fp8x4": "fp8e4m3x4/f32acc"}
REDUCED_FORMS = ("fp32", "bf16x3", "bf16x4", "fp8x4")
CONFIGS = {"c1": (4, 36, 1001), "c2": (4, 87, 1002), "c3": (8, 120, 1003), "c4": (24, 216, 1004),
           "c5": (30, 400, 1005)}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(config: str, form: str):
    """Bytes per launch between L2 and the data fabric of the persistent kernel at no=24,
    nu=216 in this arithmetic form, as the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    records it (copy-calibrated), else None."""
    prof = os.path.join(ROOT, "profiles", f"triples_kernel_{form}_no24_nu216.md")
    if config != "c4" or not os.path.exists(prof):
        return None
    for line in open(prof):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def rocpd_summary(paths: list) -> None:
    """Print, as markdown, per-kernel dispatch statistics and counter sums of rocprofv3
    result databases (rocpd SQLite): the figures the committed profile summaries quote."""
    import sqlite3

    for path in paths:
        c = sqlite3.connect(path)
        print(f"### {os.path.basename(path)}\n")
        print("| kernel | dispatches | total s | average ms | % |\n|---|---|---|---|---|")
        for name, calls, total, avg, pct in c.execute("select * from top_kernels"):   # microseconds
            print(f"| `{name[:90]}` | {calls} | {total / 1e6:.3f} | {avg / 1e3:.3f} | {pct:.1f} |")
        try:
            rows = c.execute("select kernel_name, counter_name, sum(value), count(*), sum(duration) "
                             "from counters_collection group by kernel_name, counter_name").fetchall()
        except sqlite3.OperationalError:
            rows = []
        if rows:
            print("\n| kernel | counter | sum over dispatches | dispatches | sum of durations ms |\n|---|---|---|---|---|")
            for k, cn, val, n, dur in rows:
                print(f"| `{k[:90]}` | {cn} | {val:.6g} | {n} | {dur / 1e6:.3f} |")
        print()


def host_threads() -> tuple:
    """(threads used, CPUs this process may run on): every CPU of the affinity mask, capped by
    OMP_NUM_THREADS when it is set (the GPU box sets it to one GPU's share of its cores)."""
    ncpu = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return (min(ncpu, int(cap)) if cap.isdigit() and int(cap) > 0 else ncpu), ncpu


def sampled_tuples(no: int, n_dist: int, n_pairs: int) -> list:
    """The first n_dist distinct tuples (DDCC_T_TASK order, src/triples-ijk/common_functions.f:62-86)
    and the (i,j,j), (i,i,j) tuples of the first n_pairs pairs i>j, 1-based."""
    out = []
    for i in range(no):
        for j in range(i):
            for k in range(j):
                if len(out) < n_dist:
                    out.append((i + 1, j + 1, k + 1))
    p = 0
    for i in range(no):
        for j in range(i):
            if p < n_pairs:
                out += [(i + 1, j + 1, j + 1), (i + 1, i + 1, j + 1)]
            p += 1
    return out


def cpu_baseline(inp_dev: dict, ctx, no: int, nu: int, budget_s: float, threads: int, ncpu: int) -> dict | None:
    """The reference's own compiled per-tuple routines (oracle/_ref: ijk_tuple + t1wt3_ijk for
    distinct tuples, the ijj/iij formers + T3SQUA for degenerate ones, MKL DGEMM) on the host.
    The whole tuple set runs when the calibrated estimate fits the budget; otherwise a prefix of
    the distinct tuples and one of the degenerate pairs, each sized to its share of the FLOPs, are
    timed and extrapolated per kind.  The sampled tuples' E[T] and v1 are compared with the GPU's
    (ccx_energy_tuples over the same tuples)."""
    import numpy as np

    from oracle import ref

    if not ref.available():
        log("cpu_baseline: oracle/_ref/ref_triples not built; skipped")
        return None
    host = {k: inp_dev[k].cpu().numpy() for k in ("eh", "ep", "t1", "t2", "vm", "voe")}
    host["no"], host["nu"] = no, nu
    vv = inp_dev["vvvo"].cpu().numpy().reshape(nu, nu, nu, no, order="F")
    p, q = np.tril_indices(nu)
    packed = vv[p, q, :, :].reshape(len(p), nu * no, order="F")
    del vv
    host["vvvo"] = None
    n_dist = no * (no - 1) * (no - 2) // 6
    n_pairs = no * (no - 1) // 2
    # calibration: two distinct tuples, then one degenerate pair (two tuples)
    cal_d = min(2, n_dist)
    per_d = ref.triples(host, 0, cal_d, degenerate=False, threads=threads, packed=packed)["seconds"] / max(cal_d, 1)
    per_p = ref.triples(host, 0, 0, deg_pairs=min(1, n_pairs), threads=threads, packed=packed)["seconds"]
    full_est = per_d * n_dist + per_p * n_pairs
    if full_est <= 1.5 * budget_s:
        s_d, s_p = n_dist, n_pairs
    else:
        share_d = per_d * n_dist / full_est
        s_d = int(max(min(2, n_dist), min(n_dist, budget_s * share_d / max(per_d, 1e-9))))
        s_p = int(max(min(1, n_pairs), min(n_pairs, budget_s * (1 - share_d) / max(per_p, 1e-9))))
    r_d = ref.triples(host, 0, s_d, degenerate=False, threads=threads, packed=packed)
    r_p = ref.triples(host, 0, 0, deg_pairs=s_p, threads=threads, packed=packed)
    per_d = r_d["seconds"] / max(s_d, 1)
    per_p = r_p["seconds"] / max(s_p, 1)
    full = per_d * n_dist + per_p * n_pairs
    whole = s_d == n_dist and s_p == n_pairs
    # parity of the sampled tuples: the reference's partial sums against the GPU's over the same set
    gpu = ctx.energy_tuples(sampled_tuples(no, s_d, s_p))
    etd_ref = r_d["etd"] + r_p["etd"]
    v1_ref = r_d["v1"] + r_p["v1"]
    rel_de = abs(gpu["etd"] - etd_ref) / abs(etd_ref)
    rel_dv1 = float(np.abs(gpu["v1"] - v1_ref).max() / np.abs(v1_ref).max())
    log(f"cpu_baseline sample: {s_d} distinct + {2 * s_p} degenerate tuples, reference {r_d['seconds'] + r_p['seconds']:.1f} s; "
        f"GPU vs reference on them: E[T] rel {rel_de:.2e}, v1 rel {rel_dv1:.2e}")
    return {"value": full, "unit": "s", "cores": threads, "kind": "reference",
            "sample": ("the whole tuple set" if whole else
                       f"{s_d} of {n_dist} distinct tuples ({r_d['seconds']:.1f} s) and {2 * s_p} of {2 * n_pairs} "
                       f"degenerate tuples ({r_p['seconds']:.1f} s) timed in the reference's own routines (MKL, "
                       f"{threads} threads), extrapolated per kind to all tuples"),
            "host_cpus": ncpu, "tuples_per_s": (n_dist + 2 * n_pairs) / full,
            "sample_parity": {"tuples": s_d + 2 * s_p, "etd_rel_diff": rel_de, "v1_rel_diff": rel_dv1}}


def relabelled(inp: dict, no: int, nu: int, p_occ, p_vir) -> dict:
    """The same problem with the occupied orbitals relabelled by p_occ and the virtual ones by
    p_vir (1-D Fortran-ordered device buffers in and out): E(T) is invariant, so the FP64 |dE|
    between the two is the form's own summation-order noise."""
    import torch

    po = torch.as_tensor(p_occ, device="cuda:0")
    pv = torch.as_tensor(p_vir, device="cuda:0")

    def arr(k, shape):   # Fortran-ordered buffer -> tensor indexed in reversed axis order
        return inp[k].reshape(tuple(reversed(shape)))

    out = {"eh": inp["eh"][po].contiguous(), "ep": inp["ep"][pv].contiguous()}
    out["t1"] = arr("t1", (no, nu))[pv][:, po]
    out["t2"] = arr("t2", (nu, nu, no, no))[po][:, po][:, :, pv][:, :, :, pv]
    out["vm"] = arr("vm", (no, nu, no, no))[po][:, po][:, :, pv][:, :, :, po]
    out["voe"] = arr("voe", (nu, nu, no, no))[po][:, po][:, :, pv][:, :, :, pv]
    out["vvvo"] = arr("vvvo", (nu, nu, nu, no))[po][:, pv][:, :, pv][:, :, :, pv]
    return {k: v.contiguous().reshape(-1) for k, v in out.items()}


def reorder_floor(ctx, inp: dict, no: int, nu: int, e64: float) -> dict:
    """|dE| of the FP64 form under a relabelling of the occupied and of the virtual orbitals."""
    import numpy as np
    import torch

    rng = np.random.default_rng(7)
    res = {}
    for which in ("occupied", "virtual"):
        p_occ = rng.permutation(no) if which == "occupied" else np.arange(no)
        p_vir = rng.permutation(nu) if which == "virtual" else np.arange(nu)
        rel = relabelled(inp, no, nu, p_occ, p_vir)
        ctx.load(rel)
        res[which] = abs(ctx.energy("fp64")["e"] - e64)
        del rel
        torch.cuda.empty_cache()
    ctx.load(inp)
    return {"abs_de_relabel_occupied": res["occupied"], "abs_de_relabel_virtual": res["virtual"],
            "rel": max(res.values()) / abs(e64)}


def legacy_tuple_times(inp_dev: dict, no: int, nu: int, ncall: int) -> dict:
    """Per-call time of the reference-ABI per-tuple entry ddcc_t_ijk_acc_ (host arrays, as a
    Fortran driver passes them, src/triples-ijk/ddcc_t_ijk_big.f:1-21) over the first ncall + 1
    distinct tuples; the first call, which creates the per-tuple workspace, is reported apart."""
    import ctypes

    import numpy as np

    from libcc_amd import lib

    L = lib()
    i64 = ctypes.c_int64

    def I(v):
        return ctypes.byref(i64(v))

    def D(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    host = {k: np.asfortranarray(inp_dev[k].cpu().numpy()) for k in ("eh", "ep", "t2", "vm", "voe")}
    vv = inp_dev["vvvo"].reshape(no, nu, nu, nu)     # [x][p][q][d] = VVVO(d,q,p,x)
    tups = [(i, j, k) for i in range(no) for j in range(i) for k in range(j)][:ncall + 1]
    ve = {}
    for t in tups:
        for x in t:
            if x not in ve:   # VE_x(d,p,q) = VVVO(d,q,p,x): Fortran buffer = C-order [q][p][d]
                ve[x] = np.ascontiguousarray(vv[x].permute(1, 0, 2).cpu().numpy()).reshape(-1)
    v1 = np.zeros(nu * no)
    v3 = np.zeros(nu ** 3)
    L.common_cc_init_(I(no), I(nu))
    times = []
    for (i, j, k) in tups:
        t0 = time.perf_counter()
        L.ddcc_t_ijk_acc_(I(no), I(nu), I(i + 1), I(j + 1), I(k + 1), D(v1), D(host["t2"]), D(host["vm"]), D(v3),
                          D(host["voe"]), D(host["eh"]), D(host["ep"]), D(ve[i]), D(ve[j]), D(ve[k]))
        times.append(time.perf_counter() - t0)
        if L.ccx_last_status() != 0:
            raise RuntimeError(f"ddcc_t_ijk_acc_ status {L.ccx_last_status()}")
    return {"entry": "ddcc_t_ijk_acc_", "no": no, "nu": nu, "first_call_ms": times[0] * 1e3,
            "calls": len(times) - 1, "ms_per_call": float(np.mean(times[1:])) * 1e3 if len(times) > 1 else None}


def reduced_adjoint(ctx, g64: dict, form: str) -> dict:
    """The adjoint with W recomputed in a reduced form (fp32, bf16x3): time and, per gradient
    array, the error against the FP64 adjoint in the relative Frobenius norm."""
    import torch

    g = ctx.gradient(device=True, form=form)
    ms = ctx.timing()[0]
    err = {k: float(torch.linalg.norm(g[k] - g64[k]) / torch.linalg.norm(g64[k]))
           for k in ("t1", "t2", "vm", "voe", "vvvo")}
    return {"ms": ms, "e_t": g["e"], "abs_de": abs(g["e"] - g64["e"]), "rel_grad_err": err}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every allowed CPU, capped by OMP_NUM_THREADS")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--rocpd-summary", nargs="+", default=None, metavar="DB",
                    help="only summarize rocprofv3 result databases as markdown")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run one adjoint (gradient) evaluation and print its timing (with --form fp32 or "
                         "bf16x3 also that form's adjoint and its gradient error against FP64)")
    ap.add_argument("--legacy-tuples", type=int, default=0, metavar="N",
                    help="only time N calls of the per-tuple entry ddcc_t_ijk_acc_ on host arrays")
    ap.add_argument("--form", default="fp64", choices=sorted(FORM_PEAK),
                    help="arithmetic form of the timed steps (the headline metric is fp64)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other arithmetic forms and the adjoint in the JSON line")
    args = ap.parse_args()
    if args.rocpd_summary:
        rocpd_summary(args.rocpd_summary)
        return
    if args.copy_probe_mib:
        import torch

        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the tuples would shard "
              f"over GPUs with one final energy reduction, but no multi-GPU path is implemented.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from libcc_amd import synth
    from libcc_amd.triples import TriplesContext, algorithmic_flops

    no, nu, seed = CONFIGS[args.config]
    log(f"generating no={no} nu={nu} seed={seed} on cuda:0")
    inp = synth.make_inputs_torch(no, nu, seed, "cuda:0")
    torch.cuda.synchronize()
    if args.legacy_tuples:
        log(f"legacy per-tuple entry, {args.legacy_tuples} calls")
        print(json.dumps(legacy_tuple_times(inp, no, nu, args.legacy_tuples)), flush=True)
        return
    ctx = TriplesContext(no, nu)
    if args.adjoint_only:
        ctx.load(inp)
        log("adjoint fp64")
        g = ctx.gradient(device=True)
        out = {"adjoint_ms": ctx.timing()[0], "e_t": g["e"], "no": no, "nu": nu}
        log(f"adjoint fp64 done: {out['adjoint_ms']:.0f} ms")
        if args.form != "fp64":
            log(f"adjoint {args.form}")
            out[args.form] = reduced_adjoint(ctx, g, args.form)
            log(f"adjoint {args.form} done: {out[args.form]['ms']:.0f} ms")
        print(json.dumps(out), flush=True)
        ctx.close()
        return
    log("warmup")
    ref = None
    for _ in range(args.warmup):
        ctx.load(inp)
        ref = ctx.energy(args.form)
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.load(inp)
        r = ctx.energy(args.form)
        kms.append(ctx.timing()[0])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    if ref is not None and (r["etd"] != ref["etd"] or r["ets"] != ref["ets"]):
        raise RuntimeError("non-deterministic energy between steps")
    kernel_s = sum(kms) / len(kms) / 1e3
    flops = algorithmic_flops(no, nu)
    items, exec_flops = ctx.work_stats()
    n_tup = no * (no - 1) * (no - 2) // 6 + no * (no - 1)
    achieved = flops / kernel_s / 1e12
    peak = FORM_PEAK[args.form][0]
    out = {
        "metric": METRIC,
        "value": ms_step / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": DTYPE[args.form],
        "data": "synthetic",
        "config": {"workload": f"closed-shell (T) E(T)=E[T]+E_ST and v1, no={no} nu={nu}, "
                               f"{no * (no - 1) * (no - 2) // 6} distinct + {no * (no - 1)} degenerate tuples, "
                               f"{args.form.upper()} form, inputs resident in HBM", "no": no, "nu": nu,
                   "seed": seed, "form": args.form},
        "e_t": r["e"], "etd": r["etd"], "ets": r["ets"],
        "tuples_per_s": n_tup / (ms_step / 1e3),
        "kernel_ms": kernel_s * 1e3,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": measured_traffic(args.config, args.form),
                     "flops_per_launch": flops, "executed_flops_per_launch": exec_flops, "work_items": items},
        "cpu_baseline": None,
    }
    if not args.no_extra and args.form == "fp64":
        # other arithmetic forms of the same solve: kernel time and |dE| against FP64
        out["forms"] = {"fp64": {"kernel_ms": kernel_s * 1e3, "e_t": r["e"], "abs_de": 0.0,
                                 "frac_of_peak": achieved / PEAK_FP64_TFLOPS,
                                 "peak_tflops": PEAK_FP64_TFLOPS}}
        for form in REDUCED_FORMS:
            peak, mult = FORM_PEAK[form]
            log(form)
            ctx.energy(form)
            rf = ctx.energy(form)
            kf = ctx.timing()[0] / 1e3
            out["forms"][form] = {"kernel_ms": kf * 1e3, "e_t": rf["e"], "abs_de": abs(rf["e"] - r["e"]),
                                  "rel_de": abs(rf["e"] - r["e"]) / abs(r["e"]),
                                  "speedup_vs_fp64": kernel_s / kf,
                                  "frac_of_peak": mult * flops / kf / 1e12 / peak, "peak_tflops": peak,
                                  "executed_flops_per_algorithmic": mult}
        # the FP64 form's own summation-order noise, the floor any |dE| is read against
        log("fp64 reorder floor")
        out["forms"]["fp64"]["reorder_floor"] = reorder_floor(ctx, inp, no, nu, r["e"])
        # the adjoint: dE/d(T1, T2, VM, VOE, VVVO) with per-item recomputation of W
        log("adjoint")
        g = ctx.gradient(device=True)
        adj_ms = ctx.timing()[0]
        out["adjoint"] = {"ms": adj_ms, "ratio_to_forward": adj_ms / (kernel_s * 1e3),
                          "flops": 3.0 * flops,
                          "frac_of_fp64_peak": 3.0 * flops / (adj_ms / 1e3) / 1e12 / PEAK_FP64_TFLOPS}
        for form in ("fp32", "bf16x3"):
            log(f"adjoint {form}")
            sa = reduced_adjoint(ctx, g, form)
            sa["ratio_to_forward"] = sa["ms"] / (kernel_s * 1e3)
            out["adjoint"][form] = sa
        del g
    if not args.no_cpu:
        log("cpu baseline")
        threads, ncpu = host_threads()
        if args.cpu_threads > 0:
            threads = args.cpu_threads
        out["cpu_baseline"] = cpu_baseline(inp, ctx, no, nu, args.cpu_budget, threads, ncpu)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
