#!/usr/bin/env python3
"""Benchmark of the PPI evaluation hot path on one MI355X.

A step is one evaluation of the whole population over the whole dataset (one
pass of the hot path: every program on every row, per-program error out).  The
headline workload is the plain drop-in at scale ("regression": 30,000 seeded
programs, depth <= 10, <= 128 symbols, over 100,000 x 16 N(0,1) rows, squared
error, sum reduction).  Throughput is GPop/s = sum(size) * rows / seconds
(ppi.cc:652), with programs and data already resident in HBM.

Prints ONE JSON line.  Extra fields carry the other forms (PDP, PP) and the
classification workload.  The CPU baseline is the oracle's restatement of the
reference's sequential loop, timed on this host over a bounded sample.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ppi-auto_amd"))

import ppiamd  # noqa: E402

VALU_PEAK = 256 * 64 * 2.4e9      # non-packed FP32 VALU lane-ops/s: one op per node-evaluation
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    model = platform.processor()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def bench_engine(eng, w, form, steps, warmup, rows_per_lane=0, programs_per_tile=0, strategy=ppiamd.SEQ):
    eng.configure(w.grammar.table, strategy=strategy, error_fn=w.error_fn, reducemax=w.reducemax,
                  protected=w.protected, cost=w.cost, form=form, rows_per_lane=rows_per_lane,
                  programs_per_tile=programs_per_tile)
    eng.set_dataset(w.data)
    t_up = time.perf_counter()
    eng.upload(w.pop)
    upload_s = time.perf_counter() - t_up
    st0 = eng.stats()
    if warmup > 0:
        eng.time(warmup)
    t0 = time.perf_counter()
    ms = eng.time(steps)
    wall = time.perf_counter() - t0
    st = eng.stats()
    return {"ms_per_step": ms, "wall_s": wall, "main_kernel_ms": st.main_kernel_ms,
            "node_evals": st.node_evals, "executed_nodes": st.executed_nodes,
            "rows_per_lane": st.rows_per_lane, "programs_per_tile": st.programs_per_tile,
            "form": st.form, "derived_columns": st0.derived_columns, "folded_subtrees": st0.folded_subtrees,
            "folded_nodes": st0.folded_nodes, "upload_ms": round(upload_s * 1e3, 2)}


def bench_generations(eng, w, gens):
    """What a GE loop calling the drop-in sees per generation: a new population each time
    (alternating between two seeded populations of the regression workload), compiled,
    uploaded, evaluated and its error vector and best index copied back (ppi.cc:604-634).
    Wall-clock time on the host, FP form; derived columns stay cached with the dataset."""
    pops = [w.pop, ppiamd.workload("regression", seed=2).pop]
    eng.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=w.protected,
                  form=ppiamd.FORM_FP)
    eng.set_dataset(w.data)
    for p in pops:
        eng.evaluate(p, alpha=1e-6, best_size=1)
    up = ker = 0.0
    t0 = time.perf_counter()
    for g in range(gens):
        eng.evaluate(pops[g % 2], alpha=1e-6, best_size=1)
        st = eng.stats()
        up += st.upload_ms
        ker += st.kernel_ms
    wall = (time.perf_counter() - t0) * 1e3 / gens
    return {"ms_per_generation": round(wall, 3), "upload_ms": round(up / gens, 3), "device_ms": round(ker / gens, 3)}


def bench_batched(steps, warmup):
    """64 uneven (population, dataset) pairs: one graph-captured persistent launch vs the plain
    engine evaluating the pairs one after another (device time, programs resident)."""
    w = ppiamd.batched_workload()
    b = ppiamd.Batch(0)
    b.configure(w.grammars[0].table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    for p, d in zip(w.pops, w.datas):
        b.add_population(b.add_dataset(d), p)
    b.build()
    for _ in range(warmup):
        b.run()
    ms = b.time(steps)
    st = b.stats()
    b.close()
    eng = ppiamd.Engine(0)
    plain_ms = 0.0
    for gr, p, d in zip(w.grammars, w.pops, w.datas):
        eng.configure(gr.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
        eng.set_dataset(d)
        eng.upload(p)
        eng.time(1)
        plain_ms += eng.time(1)
    eng.close()
    return {"node_evals": w.node_evals, "graph_ms": ms, "kernel_ms": st.kernel_ms, "plain_ms": plain_ms,
            "rows_per_lane": st.rows_per_lane, "items": st.n_items, "grid": st.grid,
            "tune_ms": [round(x, 3) for x in st.tune_ms[:3]], "pairs": st.n_pairs,
            "programs": int(sum(p.n for p in w.pops))}


def bench_adjoint(steps, warmup):
    """Adjoint workload (30,000 programs with <= 8 T_CONST each, 1,000,000 x 8 rows): device time
    of one backward pass (forward recomputation + reverse sweep + epilogue), and 20 on-device
    descent steps on the first 100,000 rows."""
    w = ppiamd.workload("adjoint")
    a = ppiamd.Adjoint(0)
    a.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    a.set_dataset(w.data)
    fit0 = a.forward(w.pop, alpha=1e-6)
    a.backward()  # picks the partial store by timing (AUTO) and compiles the adjoint tape
    if warmup > 0:
        a.time(warmup)
    ms = a.time(steps)
    st = a.stats()
    sub = 100000
    d = ppiamd.Adjoint(0)
    d.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    d.set_dataset(w.data[:sub])
    f0 = d.forward(w.pop, alpha=1e-6)
    _, f1 = d.descend(20, step=0.1)
    dst = d.stats()
    finite = f0 < 3.0e38 / sub
    has_c = np.array([(w.pop.phenotype[i * w.pop.stride:i * w.pop.stride + w.pop.size[i]] == w.grammar.t_const).any()
                      for i in range(w.pop.n)])
    sel = finite & has_c
    rel = (f0[sel] - f1[sel]) / f0[sel]
    d.close()
    a.close()
    return {"ms": ms, "kernel_ms": st.kernel_ms, "n": w.pop.n, "node_evals": st.node_evals,
            "fwd_ins": st.forward_ins, "rev_ins": st.reverse_ins, "partials": st.partials,
            "tune_ms": [round(x, 3) for x in st.tune_ms], "consts": st.n_consts, "max_segments": st.max_segments,
            "descend_ms": dst.descend_ms, "descend_rows": sub, "improved": float((rel > 0).mean()),
            "median_rel_gain": float(np.median(rel)), "worse": int((f1[finite] > f0[finite]).sum()),
            "reverted": dst.descend_reverted}


def cpu_baseline(w, budget_s, threads):
    """Oracle (restated sequential path) on the first k programs x all rows."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle  # cpu_baseline leg only
    cfg = oracle.make_cfg(ppiamd.SEQ, w.error_fn, w.reducemax, w.protected, w.cost)
    k = 50
    t, _ = oracle.time_seq(cfg, w.grammar.table, w.pop.subset(k), w.data, threads=threads)
    rate = w.pop.subset(k).total_size * w.data.shape[0] / max(t, 1e-9)
    k2 = int(min(w.pop.n, max(k, k * budget_s / max(t, 1e-6))))
    sub = w.pop.subset(k2)
    t2, _ = oracle.time_seq(cfg, w.grammar.table, sub, w.data, threads=threads)
    rate = sub.total_size * w.data.shape[0] / t2
    return rate, t2, k2


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fp_tile_regression_traffic.md")
VALU_PROFILE = os.path.join(ROOT, "profiles", "fp_tile_regression_valu_counters.md")


def profile_block(path):
    """The JSON block of a committed rocprofv3 summary, or None."""
    try:
        text = open(path).read()
        return json.loads(text.split("```json", 1)[1].split("```", 1)[0])
    except (OSError, IndexError, ValueError):
        return None


def measured_traffic():
    """HBM bytes per launch of the headline kernel (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the
    copy-kernel calibration), from the committed counter summary, or None."""
    return profile_block(TRAFFIC_PROFILE)


def calibrate_copy():
    """A plain device-to-device copy of known size for scaling the memory counters: 5 copies of
    1 GiB (torch's elementwise copy kernel), 1 GiB read + 1 GiB written per dispatch."""
    import torch
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "dispatches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds per CPU-baseline leg")
    ap.add_argument("--quick", action="store_true", help="headline workload and CPU baseline only")
    ap.add_argument("--only", choices=["fp", "batched", "adjoint"],
                    help="one leg, no CPU baseline (for profiler runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run the 1 GiB copy kernel used to scale the memory counters, then exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy()
        return
    if ppiamd.device_count() == 0:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    legs = {"fp", "batched", "adjoint"} if a.only is None else {a.only}
    extras = a.only is None and not a.quick

    line = {
        "metric": "GPop/s (program nodes x rows per second), plain drop-in evaluator",
        "unit": "GPop/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded programs and dataset)",
    }
    w = None
    cpu_rate = None
    if "fp" in legs:
        log("building workload: regression (30,000 programs x 100,000 rows x 16 features)")
        w = ppiamd.workload("regression")
        eng = ppiamd.Engine(0)
        log("timing FP form")
        fp = bench_engine(eng, w, ppiamd.FORM_FP, a.steps, a.warmup)
        value = fp["node_evals"] / (fp["ms_per_step"] * 1e-3) / 1e9
        achieved = fp["node_evals"] / (fp["main_kernel_ms"] * 1e-3)
        traffic = measured_traffic()
        valu = profile_block(VALU_PROFILE)
        line.update({
            "value": round(value, 3),
            "ms_per_step": round(fp["ms_per_step"], 4),
            "config": {"workload": "regression: 30,000 programs (depth<=10, <=128 symbols) x 100,000 rows x 16 "
                                   "features, EF=(X-Y)^2, sum, PROTECTED, sequential-path semantics",
                       "form": "FP", "rows_per_lane": fp["rows_per_lane"],
                       "programs_per_tile": fp["programs_per_tile"],
                       "node_evals_per_step": fp["node_evals"],
                       "executed_tape_ins_x_rows_per_step": fp["executed_nodes"],
                       "derived_columns": fp["derived_columns"],
                       "folded_subtrees": fp["folded_subtrees"], "folded_nodes": fp["folded_nodes"],
                       "upload_ms": fp["upload_ms"],
                       "upload_note": "one-time per population, outside the timed steps: tape compile, constant "
                                      "subtrees evaluated once on the device, H2D copy, and the derived columns "
                                      "(chains of unary operators over input columns, cached per dataset) it "
                                      "adds; the generation leg times all of it per generation"},
            "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 4), "peak": VALU_PEAK / 1e12,
                         "unit": "Tnode/s (1 node-evaluation = 1 FP32 VALU lane-op)",
                         "frac": round(achieved / VALU_PEAK, 5),
                         "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
                         "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                         "algorithmic_bytes": traffic["algorithmic_bytes_per_launch"] if traffic else None,
                         "valu_lane_ops_per_node": valu["valu_lane_ops_per_node"] if valu else None,
                         "valu_issue_frac": valu["valu_issue_frac"] if valu else None,
                         "note": "frac counts one FP32 VALU lane-op per node-evaluation; on this workload the "
                                 "exact (double-rounded) transcendentals and the per-instruction epilogues take "
                                 "valu_lane_ops_per_node lane-ops per node, and rocprofv3 SQ counters put VALU "
                                 "issue at valu_issue_frac of the SIMDs' capacity (the rest is scalar dispatch)"},
        })
        if extras:
            log("timing PDP form")
            pdp = bench_engine(eng, w, ppiamd.FORM_PDP, max(1, a.steps // 5), 1)
            line["pdp_gpops"] = round(pdp["node_evals"] / (pdp["ms_per_step"] * 1e-3) / 1e9, 3)
            log("timing ordered form (sums bit-identical to the sequential path)")
            od = bench_engine(eng, w, ppiamd.FORM_ORDERED, max(1, a.steps // 2), 1)
            line["ordered_gpops"] = round(od["node_evals"] / (od["ms_per_step"] * 1e-3) / 1e9, 3)
            line["ordered_note"] = ("FP row tiles store per-row errors and each program's errors are summed in row "
                                    "order: SEQ/PP results bit-identical to the sequential path")
            log("timing PP form (500-program sample)")
            wpp = ppiamd.Workload(w.name, w.grammar, w.pop.subset(500), w.data, w.error_fn)
            pp = bench_engine(eng, wpp, ppiamd.FORM_PP, 1, 0)
            line["pp_gpops"] = round(pp["node_evals"] / (pp["ms_per_step"] * 1e-3) / 1e9, 3)
            line["pp_note"] = ("the reference's PP geometry, one lane per program with rows in order: 500 programs "
                               "fill 2 workgroups of 256 CUs; kept for fidelity, the ordered form gives the same "
                               "bits at ordered_gpops")
            log("timing whole generations (compile + upload + evaluate + results)")
            gen = bench_generations(eng, w, 10)
            gen["gpops_end_to_end"] = round(fp["node_evals"] / (gen["ms_per_generation"] * 1e-3) / 1e9, 3)
            gen["host_threads"] = min(16, os.cpu_count() or 1)
            gen["note"] = ("a new 30,000-program population per generation through upload/launch/wait: host "
                           "compile of the tape + H2D + kernels + D2H of the error vector and the best index")
            line["generation"] = gen
            log("timing classification (COST, 30,000 programs x 150,000 rows)")
            wc = ppiamd.workload("classification")
            cl = bench_engine(eng, wc, ppiamd.FORM_FP, max(1, a.steps // 2), 1)
            line["classification_gpops"] = round(cl["node_evals"] / (cl["ms_per_step"] * 1e-3) / 1e9, 3)
        eng.close()
        if a.only is None:
            log("CPU baseline (oracle restatement of seq_interpret)")
            model, ncpu = host_info()
            threads = min(ncpu or 1, 16)
            cpu_rate, cpu_t, cpu_k = cpu_baseline(w, a.cpu_budget, threads)
            seq_rate, seq_t, seq_k = cpu_baseline(w, a.cpu_budget, 1)
            line["cpu_baseline"] = {"value": round(cpu_rate / 1e9, 4), "unit": "GPop/s", "cores": threads,
                                    "kind": "port",
                                    "sample": f"first {cpu_k} programs x all 100,000 rows, {cpu_t:.1f} s, OpenMP "
                                              f"over programs, host {model}"}
            line["cpu_baseline_1thread"] = {"value": round(seq_rate / 1e9, 4), "unit": "GPop/s", "cores": 1,
                                            "sample": f"first {seq_k} programs x all rows, {seq_t:.1f} s"}
            line["vs_cpu"] = round(value / (cpu_rate / 1e9), 2)
    if "batched" in legs and (a.only == "batched" or extras):
        log("batched: 64 uneven pairs (populations 500-20,000, lengths 8-256, rows 1e3-1e6, 2-32 features)")
        batched = bench_batched(max(1, a.steps // 5), 1)
        bg = batched["node_evals"] / (batched["graph_ms"] * 1e-3) / 1e9
        pg = batched["node_evals"] / (batched["plain_ms"] * 1e-3) / 1e9
        line["batched"] = {
            "workload": f"{batched['pairs']} pairs, {batched['programs']} programs, "
                        f"{batched['node_evals']:.3e} node-evaluations per step",
            "gpops": round(bg, 3), "ms_per_step": round(batched["graph_ms"], 3),
            "plain_gpops": round(pg, 3), "plain_ms": round(batched["plain_ms"], 3),
            "vs_plain": round(bg / pg, 3),
            "kernel_frac_of_valu_peak": round(batched["node_evals"] / (batched["kernel_ms"] * 1e-3) / VALU_PEAK, 5),
            "rows_per_lane": batched["rows_per_lane"], "items": batched["items"], "grid": batched["grid"],
            "autotune_ms_per_Gnode_R2_R4_R8": batched["tune_ms"]}
        if cpu_rate:
            line["batched"]["vs_cpu"] = round(bg / (cpu_rate / 1e9), 2)
    if "adjoint" in legs and (a.only == "adjoint" or extras):
        log("adjoint: 30,000 programs (<= 8 constants) x 1,000,000 rows x 8 features, backward pass + 20 descent steps")
        adjoint = bench_adjoint(2, 1)
        ms = adjoint["ms"]
        line["adjoint"] = {
            "workload": "30,000 programs (depth <= 10, <= 128 symbols, <= 8 T_CONST) x 1,000,000 rows x 8 "
                        "N(0,1) features, y = exp(0.5*x0) + sin(2*x1) + x2*x3*x4, EF = (X-Y)^2",
            "grads_per_s": round(adjoint["n"] / (ms * 1e-3), 1),
            "ms_per_backward": round(ms, 2),
            "fwd_gpops": round(adjoint["node_evals"] / (ms * 1e-3) / 1e9, 3),
            "fwd_plus_rev_ins_per_s": round((adjoint["fwd_ins"] + adjoint["rev_ins"]) / (ms * 1e-3) / 1e9, 3),
            "partials": {1: "recompute (registers)", 2: "store (LDS)"}.get(adjoint["partials"], "?"),
            "autotune_ms_recompute_store": adjoint["tune_ms"],
            "constants": adjoint["consts"], "max_segments": adjoint["max_segments"],
            "descend": {"steps": 20, "rows": adjoint["descend_rows"], "ms": round(adjoint["descend_ms"], 2),
                        "programs_improved": round(adjoint["improved"], 4),
                        "median_relative_fitness_gain": round(adjoint["median_rel_gain"], 5),
                        "programs_worse": adjoint["worse"],
                        "programs_kept_original": adjoint["reverted"]},
            "tolerance": "gradients within 1e-4*(|g|+m) of the float computation's exact derivative and "
                         "1e-3*(|g|+m) of FP64 (m = sum |EF' dpred/dc| / nlin); central differences h=1e-3"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
