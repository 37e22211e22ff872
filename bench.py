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


def host_record():
    """The host the CPU legs ran on: lscpu's model name, nproc, the CPUs this process may use, and the
    threads the CPU legs use (this process's CPUs, at most OMP_NUM_THREADS: a GPU box allots 16 CPUs to
    one GPU)."""
    model, ncpu = host_info()
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    aff = len(os.sched_getaffinity(0))
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        omp = 0
    threads = max(1, min(aff, omp) if omp > 0 else aff)
    return {"model": model, "nproc": ncpu, "affinity_cpus": aff, "omp_num_threads": omp or None,
            "threads": threads}


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
    # A fresh population every generation (never seen before: new chains for the derived-column
    # table, which is refilled when it covers under half of a population's chain occurrences).
    fresh = [ppiamd.workload("regression", seed=100 + g).pop for g in range(gens)]
    up2 = ker2 = 0.0
    t0 = time.perf_counter()
    for p in fresh:
        eng.evaluate(p, alpha=1e-6, best_size=1)
        st = eng.stats()
        up2 += st.upload_ms
        ker2 += st.kernel_ms
    wall2 = (time.perf_counter() - t0) * 1e3 / gens
    return {"ms_per_generation": round(wall, 3), "upload_ms": round(up / gens, 3), "device_ms": round(ker / gens, 3),
            "fresh_populations": {"ms_per_generation": round(wall2, 3), "upload_ms": round(up2 / gens, 3),
                                  "device_ms": round(ker2 / gens, 3),
                                  "derived_refills": eng.stats().derived_refills}}


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
    shapes = shape_summary(b, len(w.pops), w)
    # FP64 error on three pairs (the first 100 programs of the largest, a middle and the smallest dataset)
    rows = [d.shape[0] for d in w.datas]
    picks = sorted({int(np.argmax(rows)), int(np.argsort(rows)[len(rows) // 2]), int(np.argmin(rows))})
    b.run()
    errs = {}
    threads = host_record()["threads"]
    for k in picks:
        v, _ = b.results(k, alpha=0.0)
        sub = w.pops[k].subset(min(100, w.pops[k].n))
        errs[f"pair{k}_{rows[k]}rows"] = fp64_error(w.grammars[k].table, sub, w.datas[k], w.error_fn, True, False,
                                                    None, threads, v[:sub.n])
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
    gen = batch_generations(w, 2)
    return {"fp64_error": errs, "node_evals": w.node_evals, "graph_ms": ms, "kernel_ms": st.kernel_ms, "plain_ms": plain_ms,
            "rows_per_lane": st.rows_per_lane, "items": st.n_items, "grid": st.grid,
            "tune_ms": [round(x, 3) for x in st.tune_ms[:3]], "pairs": st.n_pairs, "shapes": shapes,
            "programs": int(sum(p.n for p in w.pops)), "generation": gen}


def shape_summary(b, n_pairs, w):
    """The autotuner's outcome: kernel launches per replay, keys timed, and for each shape
    (rows per lane, feature slots, stack slots) the pairs and node-evaluations it runs."""
    st = b.stats()
    per = {}
    for k in range(n_pairs):
        r, f, sl = b.pair_shape(k)
        key = f"R{r}_F{f}_S{sl}"
        e = per.setdefault(key, {"pairs": 0, "node_evals": 0.0})
        e["pairs"] += 1
        e["node_evals"] += float(w.pops[k].total_size) * w.datas[k].shape[0]
    for e in per.values():
        e["node_evals"] = float(f"{e['node_evals']:.4g}")
    return {"launches": st.n_shapes, "direct_pairs": st.direct_pairs, "tuned_keys": st.tuned_keys, "by_shape": per}


def batch_generations(w, gens):
    """Whole generations, wall clock, datasets resident: the batch re-adds every pair's population
    (compile), rebuilds (plan + capture; the autotuner's per-key shapes from the first, untimed,
    generation are kept), replays and reads all 64 result vectors; the plain engine evaluates each
    pair's population through evaluate() (compile + upload + launch + wait + download), one
    resident engine per pair, each pair's time taken after a warm-up evaluation."""
    b = ppiamd.Batch(0)
    b.configure(w.grammars[0].table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    ids = [b.add_dataset(d) for d in w.datas]
    t_b = []
    for g in range(gens + 1):
        t0 = time.perf_counter()
        b.clear_populations()
        for k, p in enumerate(w.pops):
            b.add_population(ids[k], p)
        b.build()
        b.run()
        for k in range(len(w.pops)):
            b.results(k)
        t_b.append(time.perf_counter() - t0)
    b.close()
    plain = 0.0
    eng = ppiamd.Engine(0)
    for gr, p, d in zip(w.grammars, w.pops, w.datas):
        eng.configure(gr.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
        eng.set_dataset(d)
        eng.evaluate(p)
        t0 = time.perf_counter()
        for _ in range(gens):
            eng.evaluate(p)
        plain += (time.perf_counter() - t0) / gens
    eng.close()
    batched = min(t_b[1:])
    return {"batched_ms": round(batched * 1e3, 3), "plain_ms": round(plain * 1e3, 3),
            "vs_plain": round(plain / batched, 3)}


def bench_batched_small(steps):
    """Many small pairs (the island / many-problems case): 64 pairs of 500-2,000 programs (8-64 symbols)
    on 1,000-4,000 rows x 2-16 features.  Batched graph replay against the plain engine serving the
    same pairs from 64 resident engines (launch + wait each), device time and wall clock; then whole
    generations as in batch_generations."""
    w = ppiamd.batched_workload(n_pairs=64, seed=7, p_range=(500, 2000), len_range=(8, 64),
                                rows_range=(1000, 4000), feat_range=(2, 16))
    b = ppiamd.Batch(0)
    b.configure(w.grammars[0].table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    ids = [b.add_dataset(d) for d in w.datas]
    for k, p in enumerate(w.pops):
        b.add_population(ids[k], p)
    b.build()
    b.run()
    dev_ms = b.time(steps)
    st = b.stats()
    shapes = shape_summary(b, len(w.pops), w)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
        for k in range(len(w.pops)):
            b.results(k)
    wall_ms = (time.perf_counter() - t0) / steps * 1e3
    vb = [b.results(k)[0] for k in range(len(w.pops))]
    b.close()
    engs = []
    for gr, p, d in zip(w.grammars, w.pops, w.datas):
        e = ppiamd.Engine(0)
        e.configure(gr.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
        e.set_dataset(d)
        e.upload(p)
        e.time(1)
        engs.append(e)
    plain_dev = sum(e.time(5) for e in engs)  # 5 launches per engine, averaged
    t0 = time.perf_counter()
    for _ in range(steps):
        outs = []
        for e, p in zip(engs, w.pops):
            e.launch()
            outs.append(e.wait(p.n)[0])
    plain_wall = (time.perf_counter() - t0) / steps * 1e3
    same = all(np.array_equal(x, y, equal_nan=True) for x, y in zip(vb, outs))
    for e in engs:
        e.close()
    gen = batch_generations(w, 3)
    ne = w.node_evals
    return {"workload": f"64 pairs, {int(sum(p.n for p in w.pops))} programs (8-64 symbols), 1,000-4,000 rows x "
                        f"2-16 features, {ne:.3e} node-evaluations per step",
            "device_ms": round(dev_ms, 3), "plain_device_ms": round(plain_dev, 3),
            "vs_plain_device": round(plain_dev / dev_ms, 3),
            "wall_ms": round(wall_ms, 3), "plain_wall_ms": round(plain_wall, 3),
            "vs_plain_wall": round(plain_wall / wall_ms, 3),
            "gpops_wall": round(ne / (wall_ms * 1e-3) / 1e9, 3), "plain_gpops_wall": round(ne / (plain_wall * 1e-3) / 1e9, 3),
            "results_identical_to_plain": bool(same), "rows_per_lane": st.rows_per_lane, "shapes": shapes,
            "generation": gen}


def bench_adjoint_pass(steps, warmup):
    """The adjoint workload's backward pass alone (forward, autotuned backward, `steps` timed passes):
    the leg profiler runs use."""
    w = ppiamd.workload("adjoint")
    a = ppiamd.Adjoint(0)
    a.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    a.set_dataset(w.data)
    a.forward(w.pop, alpha=1e-6)
    a.backward()
    if warmup > 0:
        a.time(warmup)
    ms = a.time(steps)
    st = a.stats()
    a.close()
    return {"ms_per_backward": round(ms, 3), "grads_per_s": round(w.pop.n / (ms * 1e-3), 1),
            "partials": st.partials, "rows_per_lane": st.rows_per_lane, "feature_slots": st.feature_slots,
            "tune_ms": [round(x, 3) for x in st.tune_ms], "fwd_ins": st.forward_ins, "rev_ins": st.reverse_ins}


def bench_adjoint(steps, warmup, threads):
    """Adjoint workload (30,000 programs with <= 8 T_CONST each, 1,000,000 x 8 rows): device time
    of one backward pass (forward recomputation + reverse sweep + epilogue), and 20 on-device
    descent steps on all 1,000,000 rows."""
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
    sub = w.data.shape[0]
    d = ppiamd.Adjoint(0)
    d.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True)
    d.set_dataset(w.data)
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
    k = 100
    e = ppiamd.Engine(0)
    e.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True, form=ppiamd.FORM_FP)
    e.set_dataset(w.data)
    gv, _ = e.evaluate(w.pop.subset(k), alpha=0.0)
    e.close()
    err = fp64_error(w.grammar.table, w.pop.subset(k), w.data, w.error_fn, True, False, None, threads, gv)
    return {"fp64_error": err, "ms": ms, "kernel_ms": st.kernel_ms, "n": w.pop.n, "node_evals": st.node_evals,
            "fwd_ins": st.forward_ins, "rev_ins": st.reverse_ins, "partials": st.partials,
            "tune_ms": [round(x, 3) for x in st.tune_ms], "consts": st.n_consts, "max_segments": st.max_segments,
            "rows_per_lane": st.rows_per_lane, "feature_slots": st.feature_slots,
            "descend_ms": dst.descend_ms, "descend_rows": sub, "improved": float((rel > 0).mean()),
            "median_rel_gain": float(np.median(rel)), "worse": int((f1[finite] > f0[finite]).sum()),
            "reverted": dst.descend_reverted}


def bench_ge(gens):
    """Grammatical evolution at the regression configuration's scale: 30,000 individuals (PPI's
    defaults -nb 2000 -bg 8 -bc 16 -cp 0.9 -mr 0.01 -ts 3, elitism) over the 100,000 x 16 rows, the
    config grammar (+ - * / sin cos exp log sqrt, attr0..15, const).  Device: every generation on the
    GPU (breeding plan with 1 or 16 random streams = the reference's -t, offspring, decode, tape
    compile, evaluation, epilogue).  Beside it the drop-in path of the reference's own loop: breeding
    and decoding on the host (the oracle's restatement of ppi.cc, one thread) and the engine's
    evaluate (host compile + upload + kernels + results)."""
    w = ppiamd.workload("regression")
    gr = w.grammar
    eng = ppiamd.Engine(0)
    eng.configure(gr.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=True, form=ppiamd.FORM_FP)
    eng.set_dataset(w.data)
    out = {}
    for T in (1, 16):
        setup = ppiamd.GeSetup(gr.g, population_size=30000, threads=T, seed=12345, elitism=True)
        ge = ppiamd.GE(eng, setup)
        t0 = time.perf_counter()
        ge.initialize()
        init_s = time.perf_counter() - t0
        ge.evolve(2)
        t0 = time.perf_counter()
        ge.evolve(gens)
        ms = (time.perf_counter() - t0) * 1e3 / gens
        st = ge.stats()
        out[f"threads_{T}"] = {"ms_per_generation": round(ms, 3), "init_ms": round(init_s * 1e3, 1),
                               "gpops_end_to_end": round(st.node_evals / (ms * 1e-3) / 1e9, 3),
                               "last_generation_device_ms": {"breed": round(st.breed_ms, 3),
                                                             "decode": round(st.decode_ms, 3),
                                                             "compile": round(st.compile_ms, 3),
                                                             "evaluate": round(st.evaluate_ms, 3),
                                                             "rank": round(st.rank_ms, 3)},
                               "node_evals": st.node_evals, "max_size_phenotype": st.max_size_phenotype,
                               "best_fitness": st.best_fitness}
        if T == 1:
            genomes, fit = ge.population()
        ge.close()
    # the drop-in path: host breeding + decode (one thread), engine evaluate
    oracle = _oracle()
    setup = ppiamd.GeSetup(gr.g, population_size=30000, threads=1, seed=12345, elitism=True)
    host = {"breed_ms": 0.0, "decode_ms": 0.0, "evaluate_ms": 0.0}
    for _ in range(2):
        t0 = time.perf_counter()
        kids = oracle.ge_breed(setup, genomes, fit, n_immigrants=1)
        t1 = time.perf_counter()
        ph, ep, sz, mx = oracle.ge_decode(setup, kids)
        t2 = time.perf_counter()
        pop = ppiamd.Population(ph.ravel(), ep.ravel(), sz, mx)
        fit, _ = eng.evaluate(pop, alpha=1e-6, best_size=1)
        t3 = time.perf_counter()
        genomes = kids
        host["breed_ms"] += (t1 - t0) * 500
        host["decode_ms"] += (t2 - t1) * 500
        host["evaluate_ms"] += (t3 - t2) * 500
    eng.close()
    host = {k: round(v, 3) for k, v in host.items()}
    host["ms_per_generation"] = round(sum(host.values()), 3)
    out["host_loop_with_drop_in"] = host
    return out


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle  # cpu_baseline and FP64-error legs only: the oracle is the checker, never the thing measured
    return oracle


def cpu_baseline(w, budget_s, threads):
    """Oracle (restated sequential path) on the first k programs x all rows."""
    oracle = _oracle()
    cfg = oracle.make_cfg(ppiamd.SEQ, w.error_fn, w.reducemax, w.protected, w.cost)
    k = min(50, w.pop.n)
    t, _ = oracle.time_seq(cfg, w.grammar.table, w.pop.subset(k), w.data, threads=threads)
    k2 = int(min(w.pop.n, max(k, k * budget_s / max(t, 1e-6))))
    sub = w.pop.subset(k2)
    t2, _ = oracle.time_seq(cfg, w.grammar.table, sub, w.data, threads=threads)
    rate = sub.total_size * w.data.shape[0] / t2
    return rate, t2, k2


def _err_stats(v, f64):
    """Relative (and absolute) error of float results v against FP64 results f64."""
    v = np.asarray(v, np.float64)
    a = np.abs(v - f64)
    r = a / np.maximum(np.abs(f64), 1e-30)
    if not r.size:
        return None
    return {"median_rel": float(np.median(r)), "p90_rel": float(np.quantile(r, 0.90)),
            "within_1e-5": float((r <= 1e-5).mean()), "within_1e-3": float((r <= 1e-3).mean()),
            "differs": float((a > 0).mean())}


def fp64_error(table, pop, data, error_fn, protected, reducemax, cost, threads, gpu_vec):
    """Error of the GPU's and of the reference's (the oracle's restated sequential path) per-program
    error against an FP64 CPU evaluation of the same programs (every operator and the error function
    in double, no alpha); programs whose float result is a non-finite sentinel are counted apart."""
    oracle = _oracle()
    cfg = oracle.make_cfg(ppiamd.SEQ, error_fn, reducemax, protected, cost)
    _, ref = oracle.time_seq(cfg, table, pop, data, alpha=0.0, threads=threads)
    f64 = oracle.fitness_f64(cfg, table, pop, data)
    g = np.asarray(gpu_vec, np.float32)
    ok = np.isfinite(f64) & (g < 3.0e38) & (ref < 3.0e38)
    return {"programs": int(pop.n), "compared": int(ok.sum()),
            "sentinel_gpu": int((g >= 3.0e38).sum()), "sentinel_reference": int((ref >= 3.0e38).sum()),
            "gpu": _err_stats(g[ok], f64[ok]), "reference": _err_stats(ref[ok], f64[ok])}


def config_leg(eng, w, steps, threads, cpu_budget, sample, note, reducemax=None):
    """One configuration: GPU GPop/s (FP form, programs resident), the oracle's CPU GPop/s on a bounded
    sample, and the FP64 error of both on the first `sample` programs."""
    rm = w.reducemax if reducemax is None else reducemax
    wl = ppiamd.Workload(w.name, w.grammar, w.pop, w.data, w.error_fn, w.cost, rm, w.protected)
    r = bench_engine(eng, wl, ppiamd.FORM_FP, steps, 1)
    gp = r["node_evals"] / (r["ms_per_step"] * 1e-3) / 1e9
    cpu, t, k = cpu_baseline(wl, cpu_budget, threads)
    sub = w.pop.subset(min(sample, w.pop.n))
    eng.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, reducemax=rm, protected=w.protected,
                  cost=w.cost, form=ppiamd.FORM_FP)
    eng.set_dataset(w.data)
    gv, _ = eng.evaluate(sub, alpha=0.0)
    return {"workload": note, "gpu_gpops": round(gp, 3), "gpu_ms_per_step": round(r["ms_per_step"], 4),
            "cpu_gpops": round(cpu / 1e9, 4), "cpu_threads": threads,
            "cpu_sample": f"first {k} of {w.pop.n} programs x all {w.data.shape[0]} rows, {t:.2f} s",
            "fp64_error": fp64_error(w.grammar.table, sub, w.data, w.error_fn, w.protected, rm, w.cost, threads,
                                     gv)}


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
    ap.add_argument("--only", choices=["fp", "ordered", "batched", "adjoint", "adjoint_pass", "ge"],
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
    legs = {"fp", "batched", "adjoint", "ge"} if a.only is None else {a.only}
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
            log("timing evaluate_pp semantics (ordered form, the whole population)")
            pp = bench_engine(eng, w, ppiamd.FORM_ORDERED, max(1, a.steps // 2), 1, strategy=ppiamd.PP)
            line["pp_gpops"] = round(pp["node_evals"] / (pp["ms_per_step"] * 1e-3) / 1e9, 3)
            line["pp_note"] = ("evaluate_pp's results (row-order sums, raw-error ranking) for all 30,000 programs, "
                               "bit for bit, by the ordered form: the default of -strategy PP; the reference's "
                               "own lane-per-program geometry (one program per lane, rows in order) is kept as "
                               "FORM_PP for fidelity tests and not timed here")
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
            host = host_record()
            threads = host["threads"]
            cpu_rate, cpu_t, cpu_k = cpu_baseline(w, a.cpu_budget, threads)
            seq_rate, seq_t, seq_k = cpu_baseline(w, a.cpu_budget, 1)
            line["cpu_baseline"] = {"value": round(cpu_rate / 1e9, 4), "unit": "GPop/s", "cores": threads,
                                    "kind": "port",
                                    "sample": f"first {cpu_k} programs x all 100,000 rows, {cpu_t:.1f} s, OpenMP "
                                              f"over programs on {threads} threads, host {host['model']}"}
            line["cpu_baseline_1thread"] = {"value": round(seq_rate / 1e9, 4), "unit": "GPop/s", "cores": 1,
                                            "sample": f"first {seq_k} programs x all rows, {seq_t:.1f} s"}
            line["host"] = host
            line["vs_cpu"] = round(value / (cpu_rate / 1e9), 2)
            if extras:
                log("per configuration: GPU and CPU GPop/s, FP64 error of both")
                eng = ppiamd.Engine(0)
                cfgs = {}
                ws = ppiamd.workload("small")
                cfgs["cfg1_small"] = config_leg(
                    eng, ws, a.steps, threads, 2.0, ws.pop.n,
                    "1,000 programs (depth <= 6, <= 64 symbols) x 1,000 rows x 4 U[-1,1], EF = |X-Y|, sum")
                cfgs["cfg1_small"]["cpu_1thread_gpops"] = round(cpu_baseline(ws, 2.0, 1)[0] / 1e9, 4)
                rg = {"workload": "the headline line above", "gpu_gpops": line["value"],
                      "cpu_gpops": line["cpu_baseline"]["value"], "cpu_threads": threads}
                eng.configure(w.grammar.table, strategy=ppiamd.SEQ, error_fn=w.error_fn, protected=w.protected,
                              form=ppiamd.FORM_FP)
                eng.set_dataset(w.data)
                sub = w.pop.subset(300)
                gv, _ = eng.evaluate(sub, alpha=0.0)
                rg["fp64_error"] = fp64_error(w.grammar.table, sub, w.data, w.error_fn, w.protected, False, None,
                                              threads, gv)
                cfgs["cfg2_regression"] = rg
                wc = ppiamd.workload("classification")
                for rm in (False, True):
                    cfgs["cfg3_classification_" + ("max" if rm else "sum")] = config_leg(
                        eng, wc, max(1, a.steps // 2), threads, 3.0, 300,
                        "30,000 programs x 150,000 rows x 4 (4 clusters), EF = COST[(int)clamp(X,0,3)][(int)Y], "
                        + ("REDUCEMAX" if rm else "sum"), reducemax=rm)
                eng.close()
                line["configs"] = cfgs
                line["configs_note"] = ("fp64_error: |float result - FP64 result| / |FP64 result| per program on a "
                                        "sample (the first programs, all rows), for the GPU and for the reference's "
                                        "sequential path (the oracle), with every operator and the error function "
                                        "in double for FP64; sentinel (non-finite) programs are counted apart")
    if a.only == "ordered":
        log("ordered form (sums bit-identical to the sequential path): regression workload")
        w = ppiamd.workload("regression")
        eng = ppiamd.Engine(0)
        od = bench_engine(eng, w, ppiamd.FORM_ORDERED, a.steps, a.warmup)
        eng.close()
        line.update({"value": round(od["node_evals"] / (od["ms_per_step"] * 1e-3) / 1e9, 3),
                     "ms_per_step": round(od["ms_per_step"], 4),
                     "config": {"workload": "regression: 30,000 programs x 100,000 rows x 16 features, ordered form "
                                            "(per-row errors stored, row-order sums)", "form": "ORDERED"}})
    if a.only == "adjoint_pass":
        log("adjoint backward pass: 30,000 programs x 1,000,000 rows")
        line["adjoint_pass"] = bench_adjoint_pass(a.steps, a.warmup)
    if "ge" in legs and (a.only == "ge" or extras):
        log("grammatical evolution on the device: 30,000 individuals x 100,000 rows, 10 generations")
        ge = bench_ge(10)
        ge["workload"] = ("30,000 individuals, -nb 2000 -bg 8 -bc 16 -cp 0.9 -mr 0.01 -ts 3 -e, config grammar over "
                          "100,000 x 16 rows, EF = (X-Y)^2, sequential-path semantics (FP form)")
        ge["note"] = ("ms per generation of ppi_evolve's loop: device = breeding plan + offspring + decode + "
                      "tape compile + evaluation + epilogue, all in HBM (threads = the reference's -t random "
                      "streams); host_loop_with_drop_in = the reference's loop with this engine as its "
                      "interpreter: breeding and decode on one host thread, then evaluate")
        line["ge"] = ge
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
            "autotune_ms_per_Gnode_R2_R4_R8": batched["tune_ms"], "shapes": batched["shapes"],
            "fp64_error_sample": batched["fp64_error"]}
        line["batched"]["generation_inclusive"] = dict(batched["generation"], note=(
            "wall clock per generation with datasets resident: batch = re-add 64 populations (compile) + build + "
            "replay + 64 result reads; plain = evaluate() per pair (compile + upload + launch + wait + download)"))
        if cpu_rate:
            line["batched"]["vs_cpu"] = round(bg / (cpu_rate / 1e9), 2)
        log("batched: 64 small pairs (500-2,000 programs, 1,000-4,000 rows)")
        line["batched_small"] = bench_batched_small(max(2, a.steps // 2))
    if "adjoint" in legs and (a.only == "adjoint" or extras):
        log("adjoint: 30,000 programs (<= 8 constants) x 1,000,000 rows x 8 features, backward pass + 20 descent steps")
        adjoint = bench_adjoint(2, 1, host_record()["threads"])
        ms = adjoint["ms"]
        line["adjoint"] = {
            "workload": "30,000 programs (depth <= 10, <= 128 symbols, <= 8 T_CONST) x 1,000,000 rows x 8 "
                        "N(0,1) features, y = exp(0.5*x0) + sin(2*x1) + x2*x3*x4, EF = (X-Y)^2",
            "grads_per_s": round(adjoint["n"] / (ms * 1e-3), 1),
            "ms_per_backward": round(ms, 2),
            "fwd_gpops": round(adjoint["node_evals"] / (ms * 1e-3) / 1e9, 3),
            "fwd_plus_rev_ins_per_s": round((adjoint["fwd_ins"] + adjoint["rev_ins"]) / (ms * 1e-3) / 1e9, 3),
            "partials": {1: "recompute (registers)", 2: "store (LDS)"}.get(adjoint["partials"], "?"),
            "rows_per_lane": adjoint["rows_per_lane"],
            "feature_slots": adjoint["feature_slots"],
            "autotune_ms": {"order": "R1 recompute (F12, F0), R1 store (F12, F0), R2 recompute (F12, F0), "
                                     "R2 store (F12, F0); -2 = shape does not hold the programs",
                            "ms": adjoint["tune_ms"]},
            "constants": adjoint["consts"], "max_segments": adjoint["max_segments"],
            "descend": {"steps": 20, "rows": adjoint["descend_rows"], "ms": round(adjoint["descend_ms"], 2),
                        "programs_improved": round(adjoint["improved"], 4),
                        "median_relative_fitness_gain": round(adjoint["median_rel_gain"], 5),
                        "programs_worse": adjoint["worse"],
                        "programs_kept_original": adjoint["reverted"]},
            "fp64_error_forward": adjoint["fp64_error"],
            "tolerance": "gradients within 1e-4*(|g|+m) of the float computation's exact derivative and "
                         "1e-3*(|g|+m) of FP64 (m = sum |EF' dpred/dc| / nlin); central differences h=1e-3"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
