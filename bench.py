#!/usr/bin/env python3
"""Benchmark of the alignment hot path on one MI355X.

A step = one pass of the hot path over the whole batch with inputs resident
in HBM: packed-int16 fill (+ int32 rerun of saturating pairs) and traceback
to CIGAR, for the 65,536-pair protein workload (log-normal lengths, median
350, local Smith-Waterman, BLOSUM62 -11/-1).  Prints ONE JSON line.

Extra fields: the plain form (one launch per pair) on a prefix of the same
workload, the all-vs-all guide-tree batch (512 proteins, global, score
only), one 300 x 320 pair (latency), the long-DNA batch (256 pairs of
10,000-100,000 bases, global, linear-memory traceback), the fill kernel's
roofline, and the CPU oracle timed on bounded samples of the same workloads
on this host's cores.  The fill kernel's HBM traffic per launch comes from
the rocprofv3 counter runs summarised in profiles/.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_GOPS = 256 * 64 * 2.4  # int32 lane-ops/s of MI355X in Gop/s (256 CU x 64 lanes x 2.4 GHz)
OPS_PER_CELL = {0: 9, 1: 10}     # recurrence ops per cell: global 9, local 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(batch, steps, warmup):
    for _ in range(warmup):
        batch.run()
        batch.wait()
    fill_ms, run_ms = [], []
    batch.wait()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
        batch.wait()
        t = batch.timings()
        fill_ms.append(t[0])
        run_ms.append(t[3])
    wall = time.perf_counter() - t0
    return wall, fill_ms, run_ms


def host_cpu():
    """The host's CPU: cores this process may use (the affinity mask, capped by
    OMP_NUM_THREADS when set: a GPU's share of its host), nproc and the model."""
    import subprocess
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    threads = aff
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if not model:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return {"nproc": nproc, "affinity": aff, "threads": threads, "omp_num_threads": omp or None, "model": model}


def oracle_params(w, traceback=None):
    from oracle import pyoracle as po
    p = w.params
    return po.params(mode=p.mode, alphabet=p.alphabet, gap_open=p.gap_open, gap_extend=p.gap_extend,
                     match=p.match, mismatch=p.mismatch, traceback=p.traceback if traceback is None else traceback)


def oracle_run(workload, ks, threads, with_cigar=True):
    """The oracle (oracle/gotoh.c, OpenMP) over pairs ks of the workload:
    (results, cigars, cells, seconds)."""
    from oracle import pyoracle as po
    codes, a_off, la, b_off, lb, cells = oracle_inputs(workload, ks)
    t0 = time.perf_counter()
    res, cig = po.align_batch(oracle_params(workload), codes, a_off, la, b_off, lb, nthreads=threads,
                              with_cigar=with_cigar)
    return res, cig, cells, time.perf_counter() - t0


def oracle_inputs(workload, ks):
    from oracle import pyoracle as po
    import numpy as np
    p = workload.params
    seqs, a_off, la, b_off, lb = [], [], [], [], []
    pos = 0
    for k in ks:
        a, b = workload.pair_bytes(k)
        ea, eb = po.encode(a, p.alphabet), po.encode(b, p.alphabet)
        seqs += [ea, eb]
        a_off.append(pos); la.append(len(ea)); pos += len(ea)
        b_off.append(pos); lb.append(len(eb)); pos += len(eb)
    codes = np.concatenate(seqs + [np.zeros(1, np.uint8)])
    return codes, a_off, la, b_off, lb, sum(x * y for x, y in zip(la, lb))


def parity(gpu, res, cigars, traceback):
    """Pairs whose GPU result (score, coordinates, CIGAR) differs from the
    oracle's, or whose walk gave up."""
    bad = []
    for k, g in enumerate(gpu):
        r = res[k]
        want = (r.score, r.a_begin, r.a_end, r.b_begin, r.b_end)
        have = (g.score, g.a_begin, g.a_end, g.b_begin, g.b_end)
        if have != want or (traceback and g.cigar != cigars[k]) or g.flags & 0x100:
            bad.append(k)
    return bad


def cpu_full_and_parity(workload, gpu_results, threads, label):
    """The oracle over every pair of the workload on `threads` host threads
    (the CPU baseline), its results compared with the GPU's pair by pair, and
    one thread over a seeded 1% sample (every 100th pair), extrapolated to the
    whole workload by cells."""
    n = workload.npairs
    tb = workload.params.traceback != 0
    res, cig, cells, dt = oracle_run(workload, range(n), threads, with_cigar=tb)
    bad = parity(gpu_results, res, cig, tb)
    ks = range(0, n, 100)
    _, _, scells, sdt = oracle_run(workload, ks, 1, with_cigar=tb)
    base = {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": "all %d pairs (%.4g cells), oracle/gotoh.c%s, OpenMP %d threads, %.2f s"
                      % (n, cells, " with traceback" if tb else ", score only", threads, dt),
            "single_thread_1pct": {"value": round(scells / sdt / 1e9, 4), "unit": "GCUPS", "cores": 1,
                                   "sample": "every 100th pair (%d pairs, %.3g cells), one thread, %.2f s"
                                             % (len(ks), scells, sdt),
                                   "extrapolated_s": round(sdt * cells / scells, 1),
                                   "note": "whole-workload time extrapolated from the sample by cells"}}
    par = {"pairs": n, "mismatches": len(bad), "first_bad": bad[:5], "checked": "score, coordinates"
           + (", CIGAR" if tb else "") + " against oracle/gotoh.c", "workload": label}
    return base, par


def cpu_single_pair(workload, repeats=1000):
    """The 300 x 320 global pair (workload 1) on one core: median of `repeats`
    alignments with traceback, timed in C around or_align."""
    from oracle import pyoracle as po
    a, b = workload.pair_bytes(0)
    p = oracle_params(workload, traceback=1)
    ea, eb = po.encode(a, p.alphabet), po.encode(b, p.alphabet)
    med = po.time_align(p, ea, eb, repeats)
    return {"value": round(len(ea) * len(eb) / med / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port",
            "median_ms": round(med * 1e3, 4),
            "sample": "one %dx%d pair with traceback, median of %d runs timed in C (no call overhead)"
                      % (len(ea), len(eb), repeats)}


def cpu_long_dna(workload, npairs=4):
    """Long-DNA fill on the CPU: score only (a full-matrix traceback of a
    100,000-base pair needs 10 GB), the first `npairs` pairs, one thread
    each, extrapolated by cells to all 256 pairs."""
    from oracle import pyoracle as po
    codes, a_off, la, b_off, lb, cells = oracle_inputs(workload, range(npairs))
    t0 = time.perf_counter()
    res, _ = po.align_batch(oracle_params(workload, traceback=0), codes, a_off, la, b_off, lb, nthreads=npairs,
                            with_cigar=False)
    dt = time.perf_counter() - t0
    total = workload.cells()
    return {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": npairs, "kind": "port",
            "sample": "first %d of %d pairs (%.3g cells), score only, one thread per pair, %.1f s"
                      % (npairs, workload.npairs, cells, dt),
            "extrapolated_s_all_pairs": round(dt * total / cells / npairs, 1),
            "note": "extrapolated: all %d pairs' cells at this sample's per-thread rate, on %d threads"
                    % (workload.npairs, npairs)}, res


def cigar_consumes(results, workload):
    """Pairs whose CIGAR does not consume both sequences exactly (M and I
    take query residues, M and D target residues), or that carry the
    walk-failure flag."""
    import re
    bad = []
    for k, g in enumerate(results):
        ops = re.findall(r"(\d+)([MID])", g.cigar)
        qa = sum(int(n) for n, o in ops if o != "D")
        qb = sum(int(n) for n, o in ops if o != "I")
        if qa != workload.arr[k].la or qb != workload.arr[k].lb or g.flags & 0x100:
            bad.append(k)
    return bad


def fill_traffic():
    """HBM bytes per launch of the fill kernel on the 65,536-pair local batch
    (workload 3), from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (None when absent)."""
    path = os.path.join(ROOT, "profiles", "fill_traffic_local_65536.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-pairs", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch shapes)")
    ap.add_argument("--served-requests", type=int, default=120000,
                    help="requests of the served run (120,000 = the 60 s Poisson schedule at 2,000/s)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs on a single MI355X and has no multi-GPU path "
              "(pairs are independent: sharding them is a host-side split of the pair list, one context "
              "per GPU)", file=sys.stderr)
        sys.exit(2)

    import tweakseq_amd as ta

    eng = ta.Engine(0)
    w3 = ta.Workload(3)
    tuned = {}

    def tune(name, w):
        if args.no_tune:
            return
        t0 = time.perf_counter()
        t, ms = eng.autotune(w.pairs_struct(), w.params, repeats=2)
        tuned[name] = {"waves_per_cu": t.waves_per_cu, "lanes": {16: "int16x2", 32: "int32"}.get(t.lanes, t.lanes),
                       "wave_tasks": t.wave_tasks, "ms": round(ms, 3), "tuning_s": round(time.perf_counter() - t0, 1)}
        log("autotune %s: %s" % (name, tuned[name]))

    tune("local_65536", w3)
    t0 = time.perf_counter()
    b3 = eng.batch(w3.pairs_struct(), w3.params)
    info = b3.info()
    log("local batch: %d pairs, %.4g cells, %d tasks, %.3g GB direction planes, setup %.1f s"
        % (w3.npairs, info["cells"], info["tasks"], info["tb_bytes"] / 1e9, time.perf_counter() - t0))
    wall, fill_ms, run_ms = timed_steps(b3, args.steps, args.warmup)
    cells = info["cells"]
    ms_step = wall * 1e3 / args.steps
    gcups = cells / (wall / args.steps) / 1e9
    fill_avg = sum(fill_ms) / len(fill_ms)
    fill_gcups = cells / (fill_avg / 1e3) / 1e9
    log("batched: %.2f ms/step (%.1f GCUPS), fill kernel %.2f ms (%.1f GCUPS), run %.2f ms"
        % (ms_step, gcups, fill_avg, fill_gcups, sum(run_ms) / len(run_ms)))
    res3 = b3.results()  # every pair, checked against the oracle below (parity_full)
    n_i32 = sum(1 for r in res3 if r.flags & ta.F_I32)
    b3.close()

    out = {
        "metric": "GCUPS on 1 MI355X (batched/served vs plain per-pair); p50/p99 request latency",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("int32" if tuned.get("local_65536", {}).get("lanes") == "int32"
                  else "int16x2 (int32 rerun on saturation)"),
        "data": "synthetic (seeded generator, tweakseq_amd/csrc/workload.cpp)",
        "config": {"workload": "65,536 protein pairs, lengths log-normal median 350 sigma 0.6 in [30,2000], "
                               "b = a mutated 25% sub + 4% indel; local Smith-Waterman BLOSUM62 -11/-1; "
                               "traceback + CIGAR", "pairs": w3.npairs, "cells": cells,
                   "parallelism": "1 GPU"},
        "fill_kernel_ms": round(fill_avg, 3),
        "fill_gcups": round(fill_gcups, 2),
        "pairs_int32": n_i32,
    }
    ach = fill_gcups * OPS_PER_CELL[1]
    tr = fill_traffic()
    out["roofline"] = {"bound": "valu-int32", "achieved": round(ach, 1), "peak": VALU_PEAK_GOPS,
                       "unit": "Gop/s", "frac": round(ach / VALU_PEAK_GOPS, 4),
                       "traffic": tr["bytes_per_launch"] if tr else None,
                       "ops_per_cell": OPS_PER_CELL[1],
                       "note": "achieved = cells x 10 recurrence ops per local cell over the HIP-event time of "
                               "the fill launch, which also walks the tracebacks of the pairs it fills (walker "
                               "waves beside the fill waves); traffic = HBM bytes per fill launch (rocprofv3 "
                               "FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated, profiles/)"}
    if tr:
        out["roofline"]["hbm_gbs"] = round(tr["bytes_per_launch"] / (fill_avg / 1e3) / 1e9, 1)

    bad_parity = False  # set by any check of GPU results against the oracle or their own invariants
    res5 = None  # the long DNA pairs' results (extras only)
    if not args.no_extra:
        # The same batch with the walks in their own kernel after the fill:
        # the fill kernel alone against the roofline.
        os.environ["TWEAKALIGN_WALK_IN_FILL"] = "0"
        bf = eng.batch(w3.pairs_struct(), w3.params)
        os.environ.pop("TWEAKALIGN_WALK_IN_FILL")
        _, ff, fr = timed_steps(bf, max(2, args.steps // 2), 1)
        bf.close()
        fo = sum(ff) / len(ff)
        out["fill_only"] = {"fill_kernel_ms": round(fo, 3), "traceback_after_ms": round(sum(fr) / len(fr) - fo, 3),
                            "fill_gcups": round(cells / (fo / 1e3) / 1e9, 2),
                            "frac": round(cells / (fo / 1e3) / 1e9 * OPS_PER_CELL[1] / VALU_PEAK_GOPS, 4)}
        log("fill only (walks after the fill): fill %.2f ms, %.1f GCUPS, frac %.3f"
            % (fo, out["fill_only"]["fill_gcups"], out["fill_only"]["frac"]))
        # PCIe-inclusive: ta_align_batch from host strings to host results
        # (encode, upload, run, download, CIGAR text), never the headline value.
        import ctypes
        ps = w3.pairs_struct()
        res = (ta.engine.CResult * w3.npairs)()
        cap = 4 * sum(ps.arr[k].la + ps.arr[k].lb for k in range(w3.npairs)) + 16
        buf = ctypes.create_string_buffer(cap)
        used = ctypes.c_int64()
        e2e = []
        for _ in range(3):
            t0 = time.perf_counter()
            rc = ta.lib().ta_align_batch(eng.ctx, ctypes.byref(w3.params), ps.arr, w3.npairs, res, buf, cap,
                                         ctypes.byref(used))
            e2e.append(time.perf_counter() - t0)
            if rc:
                raise RuntimeError("ta_align_batch failed: %d" % rc)
        out["align_batch_e2e_gcups"] = round(cells / min(e2e) / 1e9, 1)
        out["align_batch_e2e_ms"] = round(min(e2e) * 1e3, 1)
        log("ta_align_batch (host strings -> host CIGARs): %.1f ms, %.1f GCUPS"
            % (out["align_batch_e2e_ms"], out["align_batch_e2e_gcups"]))
        # The same, as a stream of batches from one host thread: batch k + 1
        # is built (letters, planning, uploads) while batch k runs, and
        # batch k's CIGAR text is formatted while k + 1 runs -- two contexts,
        # so the two batches sit on two streams and a creation's final drain
        # does not wait for the other's kernels.  Letters in, CIGAR text out,
        # every batch's results checked against the one-call form.
        ctxs = [eng.ctx, ta.Engine(0)]
        nb = 8
        want_used = used.value
        bufs = [ctypes.create_string_buffer(cap) for _ in range(2)]
        ress = [(ta.engine.CResult * w3.npairs)() for _ in range(2)]
        L = ta.lib()

        def cx(k):
            c = ctxs[k % 2]
            return c if isinstance(c, ctypes.c_void_p) else c.ctx

        def create(k):
            h = ctypes.c_void_p()
            rc = L.ta_batch_create(cx(k), ctypes.byref(w3.params), ps.arr, w3.npairs, ctypes.byref(h))
            if rc:
                raise RuntimeError("ta_batch_create failed: %d" % rc)
            return h

        pipe = []
        for rep in range(2):
            t0 = time.perf_counter()
            prev = None
            for k in range(nb):
                h = create(k)
                if L.ta_batch_run(cx(k), h):
                    raise RuntimeError("ta_batch_run failed")
                if prev is not None:
                    kp, hp = prev
                    u = ctypes.c_int64()
                    if L.ta_batch_wait(cx(kp), hp) or L.ta_batch_results(cx(kp), hp, ress[kp % 2], bufs[kp % 2], cap,
                                                                         ctypes.byref(u)):
                        raise RuntimeError("pipelined batch failed")
                    if u.value != want_used:
                        raise RuntimeError("pipelined batch gave other CIGARs")
                    L.ta_batch_destroy(hp)
                prev = (k, h)
            kp, hp = prev
            u = ctypes.c_int64()
            if L.ta_batch_wait(cx(kp), hp) or L.ta_batch_results(cx(kp), hp, ress[kp % 2], bufs[kp % 2], cap,
                                                                 ctypes.byref(u)):
                raise RuntimeError("pipelined batch failed")
            L.ta_batch_destroy(hp)
            pipe.append(time.perf_counter() - t0)
        same = bufs[0].raw[:want_used] == buf.raw[:want_used] and bufs[1].raw[:want_used] == buf.raw[:want_used]
        ctxs[1].close()
        out["align_batch_pipelined"] = {
            "gcups": round(nb * cells / min(pipe) / 1e9, 1), "ms_per_batch": round(min(pipe) / nb * 1e3, 2),
            "batches": nb, "cigar_text_equals_one_call": same,
            "form": "one host thread, batch k+1 created while batch k runs (two contexts / streams); "
                    "letters in, CIGAR text out"}
        log("pipelined batches (host strings -> host CIGARs): %.2f ms per batch, %.1f GCUPS"
            % (out["align_batch_pipelined"]["ms_per_batch"], out["align_batch_pipelined"]["gcups"]))
        # Resident sequences: the workload's sequences added once
        # (ta_seq_add), then the batch aligned by handle -- its creation skips
        # deduplication, letter staging, upload and encoding.  CIGAR text
        # checked against the one-call form.
        t0 = time.perf_counter()
        if L.ta_seq_reset(eng.ctx):
            raise RuntimeError("ta_seq_reset failed")
        hs = (ta.engine.SeqPair * w3.npairs)()
        hv = ctypes.c_int64()
        for k in range(w3.npairs):
            pk = ps.arr[k]
            for side, addr, ln in (("a", pk.a, pk.la), ("b", pk.b, pk.lb)):
                if L.ta_seq_add(eng.ctx, w3.params.alphabet, ctypes.cast(addr, ctypes.c_char_p), ln, ctypes.byref(hv)):
                    raise RuntimeError("ta_seq_add failed")
                setattr(hs[k], side, hv.value)
        add_s = time.perf_counter() - t0
        rbuf = ctypes.create_string_buffer(cap)
        ru = ctypes.c_int64()
        resd = []
        for _ in range(3):
            t0 = time.perf_counter()
            rc = L.ta_align_resident(eng.ctx, ctypes.byref(w3.params), hs, w3.npairs, res, rbuf, cap, ctypes.byref(ru))
            resd.append(time.perf_counter() - t0)
            if rc:
                raise RuntimeError("ta_align_resident failed: %d" % rc)
        L.ta_seq_reset(eng.ctx)
        out["align_resident_e2e"] = {
            "ms": round(min(resd) * 1e3, 1), "gcups": round(cells / min(resd) / 1e9, 1),
            "sequences": 2 * w3.npairs, "seq_add_s": round(add_s, 2),
            "first_call_ms": round(resd[0] * 1e3, 1),  # includes the one upload of the added codes
            "cigar_text_equals_one_call": ru.value == want_used and rbuf.raw[:want_used] == buf.raw[:want_used],
            "form": "sequences added once (ta_seq_add), one ta_align_resident call: handles in, CIGAR text out"}
        log("resident sequences (handles -> host CIGARs): %.1f ms, %.1f GCUPS (adding %d sequences: %.2f s)"
            % (out["align_resident_e2e"]["ms"], out["align_resident_e2e"]["gcups"], 2 * w3.npairs, add_s))
        # Plain form: one fill + one traceback launch per pair, on a prefix.
        npl = min(args.plain_pairs, w3.npairs)
        bp = eng.batch(w3.pairs_struct(0, npl), w3.params)
        pcells = bp.info()["cells"]
        bp.run_plain(); bp.wait()
        t0 = time.perf_counter()
        bp.run_plain(); bp.wait()
        pdt = time.perf_counter() - t0
        bp.close()
        plain = pcells / pdt / 1e9
        out["plain_gcups"] = round(plain, 3)
        out["plain_frac_of_roofline"] = round(plain * OPS_PER_CELL[1] / VALU_PEAK_GOPS, 6)
        out["batched_vs_plain"] = round(gcups / plain, 1)
        out["plain_sample"] = "first %d pairs, one fill launch + one traceback launch per pair" % npl
        log("plain: %.3f GCUPS on %d pairs -> batched/plain %.1fx" % (plain, npl, gcups / plain))
        # All-vs-all guide-tree scoring (global, score only).
        w2 = ta.Workload(2)
        tune("allvsall_130816", w2)
        b2 = eng.batch(w2.pairs_struct(), w2.params)
        c2 = b2.info()["cells"]
        wall2, f2, _ = timed_steps(b2, max(2, args.steps // 2), 1)
        res2 = b2.results()
        b2.close()
        out["allvsall_gcups"] = round(c2 / (wall2 / max(2, args.steps // 2)) / 1e9, 2)
        out["allvsall_fill_gcups"] = round(c2 / (sum(f2) / len(f2) / 1e3) / 1e9, 2)
        log("all-vs-all (130,816 global pairs, score only): %.1f GCUPS" % out["allvsall_gcups"])
        # One 300 x 320 global pair with traceback: launch-to-result latency.
        w1 = ta.Workload(1)
        b1 = eng.batch(w1.pairs_struct(), w1.params)
        lat = []
        for k in range(220):
            t0 = time.perf_counter()
            b1.run(); b1.wait()
            if k >= 20:
                lat.append(time.perf_counter() - t0)
        b1.close()
        lat.sort()
        out["single_pair_ms"] = round(lat[len(lat) // 2] * 1e3, 4)
        out["single_pair_gcups"] = round(w1.cells() / lat[len(lat) // 2] / 1e9, 3)
        log("single 300x320 pair: median %.3f ms" % out["single_pair_ms"])
        # Call -> CIGAR of the same pair through the C ABI, timed in C: the
        # single-pair path (ta_align_pair) and ta_align_batch with n = 1.
        import subprocess
        pb = os.path.join(ROOT, "tweakseq_amd", "tweakalign-pair-bench")
        if os.path.exists(pb):
            r = subprocess.run([pb, "2000"], capture_output=True, text=True, timeout=300)
            if r.returncode == 0:
                out["single_pair_call"] = json.loads(r.stdout.strip().splitlines()[-1])
                log("single pair call -> CIGAR:", out["single_pair_call"])
        # Long DNA pairs: strips pipelined across CUs + checkpointed traceback.
        w5 = ta.Workload(5)
        b5 = eng.batch(w5.pairs_struct(), w5.params)
        c5 = b5.info()["cells"]
        wall5, f5, _ = timed_steps(b5, 2, 1)
        t5 = b5.timings()
        res5 = b5.results()
        b5.close()
        bad5 = cigar_consumes(res5, w5)
        out["long_dna_cigars"] = {"pairs": w5.npairs, "bad": len(bad5), "first_bad": bad5[:5],
                                  "checked": "every CIGAR consumes both sequences exactly, no walk-failure flag"}
        bad_parity = bad_parity or len(bad5) > 0
        out["long_dna_gcups"] = round(c5 / (wall5 / 2) / 1e9, 1)
        out["long_dna_ms"] = round(wall5 / 2 * 1e3, 1)
        out["long_dna_fill_gcups"] = round(c5 / (sum(f5) / len(f5) / 1e3) / 1e9, 1)
        out["long_dna_traceback_ms"] = round(t5[2], 1)
        log("long DNA (256 pairs, %.3g cells): %.1f ms/run, %.1f GCUPS (forward %.1f GCUPS, traceback %.1f ms)"
            % (c5, out["long_dna_ms"], out["long_dna_gcups"], out["long_dna_fill_gcups"], t5[2]))

    if not args.no_extra:
        # Column consensus (Consensus::calculate) of a 512 x 2,000 alignment:
        # ta_consensus end to end (rows from host memory, consensus back),
        # and the oracle's O(cols x nseq^2) restatement on one core.
        import random as _r
        rng = _r.Random(5)
        base = [rng.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(2000)]
        rows = []
        for _ in range(512):
            r = [c if rng.random() > 0.25 else rng.choice("ARNDCQEGHILKMFPSTWYV-") for c in base]
            rows.append("".join(r).encode())
        eng.consensus(rows)
        ct = []
        for _ in range(5):
            t0 = time.perf_counter()
            got = eng.consensus(rows)
            ct.append(time.perf_counter() - t0)
        out["consensus_512x2000_ms"] = round(sorted(ct)[2] * 1e3, 3)
        if not args.no_cpu:
            from oracle import pyoracle as po
            t0 = time.perf_counter()
            want = po.consensus(rows)
            out["consensus_512x2000_cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
            out["consensus_matches_oracle"] = got == want
        log("consensus 512 x 2000: %.3f ms on the GPU (end to end)" % out["consensus_512x2000_ms"])
        # Progressive multiple alignment (ta_msa) of the all-vs-all workload's
        # 512 sequences (100-600 residues, 30% mutated copies): all-pair
        # scores, UPGMA, profile joins; end to end from host strings.  The
        # oracle's restatement runs on one core on the first 48 of them, and
        # the GPU on the same 48 for the like-for-like ratio.
        import ctypes as _c
        w2 = ta.Workload(2)
        seqs = [_c.string_at(w2.arr[0].a, w2.arr[0].la).decode()]
        for k in range(w2.npairs):
            if w2.arr[k].a != w2.arr[0].a:
                break
            seqs.append(_c.string_at(w2.arr[k].b, w2.arr[k].lb).decode())
        eng.msa(seqs)
        mt = []
        for _ in range(3):
            t0 = time.perf_counter()
            rows_m, joins_m = eng.msa(seqs)
            mt.append(time.perf_counter() - t0)
        out["msa_512_ms"] = round(min(mt) * 1e3, 1)
        out["msa_512_columns"] = len(rows_m[0])
        sub = seqs[:48]
        mt = []
        for _ in range(3):
            t0 = time.perf_counter()
            rows_s, joins_s = eng.msa(sub)
            mt.append(time.perf_counter() - t0)
        out["msa_48_ms"] = round(min(mt) * 1e3, 1)
        if not args.no_cpu:
            from oracle import pyoracle as po
            t0 = time.perf_counter()
            alen, cols, wj = po.msa([po.encode(x) for x in sub])
            out["msa_48_cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
            out["msa_48_matches_oracle"] = rows_s == po.msa_rows(sub, alen, cols) and joins_s == wj
            # the whole 512-sequence alignment against the oracle (rows and joins)
            t0 = time.perf_counter()
            alen5, cols5, wj5 = po.msa([po.encode(x) for x in seqs])
            out["msa_512_cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
            out["msa_512_matches_oracle"] = rows_m == po.msa_rows(seqs, alen5, cols5) and joins_m == wj5
            bad_parity = bad_parity or not (out["msa_48_matches_oracle"] and out["msa_512_matches_oracle"])
        log("msa 512 sequences: %.1f ms on the GPU (end to end, %d columns)"
            % (out["msa_512_ms"], out["msa_512_columns"]))
        # Served form: open-loop Poisson arrivals (2,000 requests/s, 1-256
        # pairs each, lengths 50-5,000, global/local 50/50) through
        # ta_submit/ta_poll in a child process; latency = completion minus
        # scheduled arrival.  The default admission path is the device queue
        # (requests pushed into the running persistent kernel); the batch
        # path (TWEAKALIGN_DEVICE_QUEUE=0) is measured beside it on the
        # first 20,000 requests of the same schedule.
        import subprocess
        sb = os.path.join(ROOT, "tweakseq_amd", "tweakalign-serve-bench")
        if os.path.exists(sb) and args.served_requests > 0:
            log("served: %d requests at 2,000/s (%.0f s of arrivals)..." % (args.served_requests,
                                                                           args.served_requests / 2000))
            r = subprocess.run([sb, str(args.served_requests)], capture_output=True, text=True,
                               timeout=300 + args.served_requests / 1000)
            if r.returncode == 0:
                out["served"] = json.loads(r.stdout.strip().splitlines()[-1])
                # served GCUPS follow the offered load (2,000 requests/s); the
                # fraction is of the GPU's int32 roofline at 9.5 ops per cell
                # (half global, half local)
                out["served"]["frac_of_roofline"] = round(out["served"]["served_gcups"] * 9.5 / VALU_PEAK_GOPS, 4)
                out["served"]["admission"] = "device queue"
                bad_parity = bad_parity or out["served"].get("bad_cigars", 0) > 0
                log("served: p50 %.1f ms, p99 %.1f ms, %.1f GCUPS served of %.1f offered"
                    % (out["served"]["p50_ms"], out["served"]["p99_ms"], out["served"]["served_gcups"],
                       out["served"]["offered_gcups"]))
            nb = min(args.served_requests, 20000)
            env = dict(os.environ, TWEAKALIGN_DEVICE_QUEUE="0")
            r = subprocess.run([sb, str(nb)], capture_output=True, text=True, timeout=300 + nb / 1000, env=env)
            if r.returncode == 0:
                out["served_batches"] = json.loads(r.stdout.strip().splitlines()[-1])
                out["served_batches"]["admission"] = "batches (TWEAKALIGN_DEVICE_QUEUE=0)"
                bad_parity = bad_parity or out["served_batches"].get("bad_cigars", 0) > 0
                log("served through batches (%d requests): p50 %.1f ms, p99 %.1f ms"
                    % (nb, out["served_batches"]["p50_ms"], out["served_batches"]["p99_ms"]))
        # Measured VALU throughput (instruction-kind microbenchmark) beside the theoretical peak.
        tool = os.path.join(ROOT, "tweakseq_amd", "tweakalign-valu-peak")
        if os.path.exists(tool):
            r = subprocess.run([tool], capture_output=True, text=True, timeout=120)
            if r.returncode == 0:
                out["valu_peak_measured"] = json.loads(r.stdout.strip().splitlines()[-1])

    if tuned:
        out["autotune"] = tuned
    if not args.no_cpu:
        hc = host_cpu()
        out["cpu_host"] = hc
        log("cpu host:", hc)
        base, par = cpu_full_and_parity(w3, res3, hc["threads"], "65,536 local pairs (the timed batch)")
        base["host"] = hc["model"]
        out["cpu_baseline"] = base
        out["parity_full"] = {"local_65536": par}
        bad_parity = bad_parity or par["mismatches"] > 0
        log("cpu baseline:", base)
        log("parity, 65,536 local pairs:", par)
        if not args.no_extra:
            b2base, par2 = cpu_full_and_parity(ta.Workload(2), res2, hc["threads"],
                                               "130,816 all-vs-all global pairs, score only")
            out["cpu_allvsall"] = b2base
            out["parity_full"]["allvsall_130816"] = par2
            bad_parity = bad_parity or par2["mismatches"] > 0
            log("cpu all-vs-all:", b2base)
            log("parity, all-vs-all:", par2)
            out["cpu_single_pair"] = cpu_single_pair(ta.Workload(1))
            out["cpu_long_dna"], resl = cpu_long_dna(ta.Workload(5))
            if res5 is not None:
                bad = parity(res5[:len(resl)], resl, None, False)
                out["parity_full"]["long_dna_first4"] = {
                    "pairs": len(resl), "mismatches": len(bad), "first_bad": bad,
                    "checked": "score, coordinates against oracle/gotoh.c (the CPU sample's pairs, 10k-100k bases)",
                    "workload": "the first 4 of the 256 long DNA pairs"}
                bad_parity = bad_parity or len(bad) > 0
            log("cpu single pair:", out["cpu_single_pair"])
            log("cpu long dna:", out["cpu_long_dna"])
    eng.close()
    print(json.dumps(out), flush=True)
    if bad_parity:
        log("PARITY FAILURE: GPU results differ from the oracle (parity_full) or fail their invariants "
            "(long_dna_cigars, served bad_cigars, msa_*_matches_oracle)")
        sys.exit(3)


if __name__ == "__main__":
    main()
