#!/usr/bin/env python3
"""Mapped reads/s of the MI355X mapping path, learned-index form against the plain FMD form,
with bit-exact SAM.

Default workload (the largest configuration, combined form): 1000 genomes with sizes
log-uniform in 1-8 Mb (3.40 Gbp; BD text 6.80e9 symbols, so the index takes the wide layout),
GC ~ U(0.3, 0.7), 30% near-duplicate strains at 0.5-5% divergence; read pairs 2x150 at 1% error
(0.9% substitutions, 0.1% 1-3 bp indels), insert N(400, 80), 10% off-target pairs, Phred
U(20, 40), in 15 resident batches of 128k pairs (3.9M reads).  `--workload strains` runs the
learned-index configuration (256 genomes of 4 Mb with 20% strains, single-end 150 bp reads),
`--workload pairs` the 64 x 4 Mb paired configuration.

A step maps one batch end to end (SMEMs, seeds, chains, DP + backtrace and pairing on the GPU;
MAPQ and SAM records on the host) with the batch's reads already in HBM; batches are pipelined
(several in flight on the worker streams).  Prints one JSON line.  The CPU baseline is timed on
the host cores over a bounded sample of the same reads (median of 3 runs): the reference's own
seeding and chaining code (RRR wavelet-tree ranks, LF-walk suffix-array lookups; compiled from
its sources into oracle/_ref) with the oracle's DP and records, and the oracle restatement with
a plain occ table beside it; on the large workload the oracle restatement alone (building the
reference's RRR wavelet tree over 6.8e9 symbols takes minutes).
"""
import argparse
import collections
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mapped reads/s on 1 MI355X, learned-index vs plain FMD, bit-exact SAM; DP GCUPS"
HBM_PEAK_GBS = 8000.0
# integer VALU peak of one MI355X: 256 CUs x 64 lanes x 2.4 GHz (lane operations per second)
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9
# integer operations one DP cell needs as written (src/Alignment.cpp:81-100): the substitution
# score (compare + select, 2), M = max of three diagonal predecessors + score (3), I and D = max
# of a gap open and a gap extension (2 subtractions + 1 max each, 6), the -inf clamps of those
# four steps (4), the backtrace decisions (matchMax 4 compares/selects, insMax 1, delMax 1, the
# M > 0 test 1, packing 4: 11), the row/column indexing of the anti-diagonal step (2)
DP_OPS_PER_CELL = 28

WORKLOADS = {
    "strains": "SE 150 bp at 1% error: 256 genomes x 4 Mb (1 Gbp, GC U(0.3,0.7)), 20% near-duplicate strains "
               "at 1-5% divergence; 15 batches of 256k reads resident in HBM",
    "pairs": "PE 2x150: 64 genomes x 4 Mb (256 Mbp, GC U(0.3,0.7)), insert N(350,50), 0.5% sub, 0.05% indel, "
             "5% off-target pairs; 1M pairs in 256k-read batches resident in HBM",
    "large": "PE 2x150 at 1% error: 1000 genomes log-uniform 1-8 Mb (3.40 Gbp, BD text 6.80e9 symbols, wide "
             "index layout), GC U(0.3,0.7), 30% near-duplicate strains at 0.5-5% divergence, insert N(400,80), "
             "10% off-target pairs; 15 batches of 128k pairs (256k reads) resident in HBM",
}
# insert-size model of each paired workload (mean, sd): msgseqtk-align -m/-d
INSERT = {"pairs": (350, 50), "large": (400, 80)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="large")
    ap.add_argument("--reads", type=int, default=0,
                    help="reads of the workload (0: 4M single-end / 2M paired / 20M paired for large)")
    ap.add_argument("--batch-reads", type=int, default=262144)
    ap.add_argument("--cpu-reads", type=int, default=20000, help="reads in the CPU-baseline sample")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU-baseline runs (the median is reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-kind", choices=("both", "port"), default="",
                    help="CPU baselines: the reference's seeding + the oracle ('both'), or the oracle alone "
                         "(default: both, except on the large workload)")
    ap.add_argument("--no-clean", action="store_true", help="skip the msgseqtk-clean seeding measurement")
    ap.add_argument("--no-compare", action="store_true", help="skip the plain-FMD comparison run")
    ap.add_argument("--no-tune", action="store_true", help="keep the default lanes per CU")
    ap.add_argument("--no-torch", action="store_true", help="no torch device synchronisation (profiler runs)")
    ap.add_argument("--no-pipeline", action="store_true", help="finish each batch before submitting the next")
    ap.add_argument("--no-ingest", action="store_true", help="skip the run with the per-batch host-to-device copy")
    ap.add_argument("--host-threads", type=int, default=0, help="host finishing threads (0: the library default)")
    ap.add_argument("--form", type=int, default=2,
                    help="0 plain FMD extensions, 1 unique-interval text path, 2 form 1 + learned index")
    ap.add_argument("--lix-segments", type=int, default=1 << 22, help="learned index: root segments")
    ap.add_argument("--lix-steps", type=int, default=1000, help="learned index: training steps")
    ap.add_argument("--no-gather", action="store_true", help="skip the random-gather ceiling measurement")
    ap.add_argument("--depth", type=int, default=8, help="batches in flight")
    ap.add_argument("--streams", type=int, default=4, help="batches whose device stages run at once")
    ap.add_argument("--lanes-per-cu", type=int, default=0, help="fix the SMEM kernel's lanes per CU (no autotuning)")
    ap.add_argument("--group-lanes", type=int, default=-1, help="fix the form-2 lanes per read: 0, 8 or 16")
    ap.add_argument("--text-path", type=int, default=-1, help="fix the form-2 one-lane kernel's text path: 0 or 1")
    ap.add_argument("--read-filter", type=int, default=-1, help="fix the form-2 read filter: 0 or 1")
    ap.add_argument("--pair-chains", type=int, default=-1, help="fix the one-lane kernels' paired chains: 0 or 1")
    ap.add_argument("--split-lanes", type=int, default=-1,
                    help="fix the form-2 split launch's lanes per read for likely-long reads: 0, 4, 8 or 16")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 1 GiB device copy (counter calibration under rocprofv3 --pmc)")
    return ap.parse_args()


def make_data(a, synth):
    """(chroms, names, batches, paired); a batch is (reads[, mates]) of (seqs, quals, names)"""
    if a.workload == "strains":
        chroms, names = synth.config_genomes(4)
        nreads = a.reads or 4_000_000
        per = a.batch_reads
        nb = max(1, min((nreads + per - 1) // per, 15))   # slot 15 is the CPU-baseline check's
        batches = [(synth.simulate_reads_fast(chroms, per, 150, sub=0.009, indel=0.001, seed=0xC0FFEE + 1004 + b,
                                              prefix="b%dr" % b),) for b in range(nb)]
        return chroms, names, batches, False
    if a.workload == "large":
        chroms, names = synth.config_genomes(5)
        nreads = a.reads or 20_000_000
        err, off, seed = (0.009, 0.001), 0.1, 0xC0FFEE + 1005
    else:
        chroms, names = synth.config_genomes(2)
        nreads = a.reads or 2_000_000
        err, off, seed = (0.005, 0.0005), 0.05, 0xC0FFEE + 1002
    per = a.batch_reads // 2
    nb = max(1, min((nreads // 2 + per - 1) // per, 15))
    batches = [synth.simulate_pairs_fast(chroms, per, 150, insert=INSERT[a.workload], sub=err[0], indel=err[1],
                                         seed=seed + b, offtarget=off, prefix="b%dp" % b)
               for b in range(nb)]
    return chroms, names, batches, True


def main():
    a = parse()
    if not a.cpu_kind:
        a.cpu_kind = "port" if a.workload == "large" else "both"
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the mapping path shards by read batch with a replicated "
              "index, and this build measures the single-GPU kernel; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import msgseqtk_amd as M
    from msgseqtk_amd import synth

    t0 = time.time()
    chroms, names, batches, paired = make_data(a, synth)
    n_batches = len(batches)
    log("data: %d genomes, %d batches of %d reads in %.1fs" % (len(chroms), n_batches, a.batch_reads,
                                                              time.time() - t0))
    t1 = time.time()
    m = M.Mapper(0).build(chroms, names)
    build_s = time.time() - t1
    log("index built on the GPU in %.1fs (BD length %d)" % (build_s, m.n))
    lix = None
    if a.form == 2:
        t3 = time.time()
        lix = m.lix_train(segments=a.lix_segments, samples=1 << 22, steps=a.lix_steps, batch=4096, lr=1e-3, seed=7)
        lix["train_s"] = round(time.time() - t3, 2)
        log("learned index trained in %.1fs: %s" % (lix["train_s"], lix))
    m.set("form", a.form)
    for b, bt in enumerate(batches):
        m.upload(b, *bt)
    if paired:
        mu, sd = INSERT[a.workload]
        p = M.default_params(insMean=mu, insSD=sd, insMin=mu - 10 * sd, insMax=mu + 10 * sd)
    else:
        p = M.default_params()
    sync = lambda: None   # msgs_batch_wait returns after the batch's stream has drained
    if not a.no_torch:
        try:
            import torch
            if torch.cuda.is_available():
                sync = torch.cuda.synchronize
        except Exception:
            pass

    tuned = None
    for key, v in (("lanes_per_cu", a.lanes_per_cu), ("group_lanes", a.group_lanes), ("text_path", a.text_path),
                   ("read_filter", a.read_filter), ("split_lanes", a.split_lanes), ("pair_chains", a.pair_chains)):
        if (key == "lanes_per_cu" and v) or (key != "lanes_per_cu" and v >= 0):
            m.set(key, v)
    if not a.no_tune and not a.lanes_per_cu:
        lanes, tus = m.tune(0, p)
        tuned = {k: m.get(k) for k in ("lanes_per_cu", "group_lanes", "text_path", "read_filter", "split_lanes",
                                       "pair_chains")}
        tuned["smem_us"] = round(tus, 1)
        log("autotuner: %s" % tuned)
    m.set("streams", a.streams)
    run_steps.serial = a.no_pipeline
    run_steps.depth = a.depth
    if a.host_threads:
        m.set("host_threads", a.host_threads)
    m.set("ingest", 0)
    elapsed, S = run_steps(m, p, n_batches, a.steps, a.warmup, sync, log)
    reads, mapped = S[:, 0].sum(), S[:, 1].sum()
    value = mapped / elapsed
    ext_calls, sa_lookups, cells = S[:, 2].sum(), S[:, 3].sum(), S[:, 4].sum()
    smem_us, dp_us = S[:, 5].sum(), S[:, 6].sum()
    # dominant kernel: the SMEM search; algorithmic bytes = one 64-B occ block per rank side
    # (2 per extension with an unambiguous base) + one 64-B line per SA/ISA lookup + 1 B per
    # text step of the unique-interval path
    launches = a.steps
    fast_steps, lookups = S[:, 10].sum(), S[:, 11].sum()
    bytes_per_launch = (64.0 * 2.0 * ext_calls + 64.0 * lookups + fast_steps) / launches
    avg_us = smem_us / launches
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    kname = smem_kernel_name(m, a.form)
    log("timed %d steps in %.2fs: %.0f mapped reads/s (%.1f%% mapped)" % (a.steps, elapsed, value,
                                                                          100 * mapped / reads))
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "mapped reads/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1000 * elapsed / a.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32+fp64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[a.workload], "batch_reads": a.batch_reads, "resident_batches": n_batches,
                   "bd_length": int(m.n),
                   "form": ["plain FMD", "unique-interval text path",
                            "learned-index exact-read path and read filter + unique-interval text path"][a.form]},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_launch": bytes_per_launch, "avg_launch_us": round(avg_us, 1)},
        "dp_gcups": round(cells / (dp_us * 1e-6) / 1e9, 3) if dp_us > 0 else None,
        "smem_iterations": {"lane_per_read": S[:, 12].sum() / reads,
                            "simd_utilisation": S[:, 12].sum() / max(1.0, 64.0 * S[:, 13].sum()),
                            "max_wavefront": int(S[:, 14].max())},
        "per_read": {"occ_extensions": ext_calls / reads, "text_steps": fast_steps / reads,
                     "smem_sa_isa_lookups": lookups / reads, "seed_sa_lookups": sa_lookups / reads,
                     "dp_cells": cells / reads},
        "stage_ms_per_step": {"read_filter": S[:, 15].sum() / a.steps / 1000,
                              "smem_kernel": smem_us / a.steps / 1000, "seed_stage": S[:, 8].sum() / a.steps / 1000,
                              "dp_kernel": dp_us / a.steps / 1000, "host_finish": S[:, 9].sum() / a.steps / 1000,
                              "whole": S[:, 7].sum() / a.steps / 1000},
        "dp_jobs_per_step": {"all": float(S[:, 21].mean()), "wavefront_fill": float(S[:, 22].mean()),
                             "serial_fill": float(S[:, 32].mean())},
        # rectangles of the jobs filled 16 lanes each: the step loop's lane use is their cells over
        # 16 x (C + rows - 1) lane-steps per strip of at most 16 rows
        "dp_fill16_rects": {"per_step": float(S[:, 38].mean()),
                            "cells_per_rect": float(S[:, 39].sum() / max(1.0, S[:, 38].sum())),
                            "strips_per_rect": float(S[:, 41].sum() / max(1.0, S[:, 38].sum())),
                            "step_lane_use": float(S[:, 39].sum() / max(1.0, S[:, 40].sum()))},
        "index_build_s": round(build_s, 2),
        "learned_index": dict(lix or {}, **({"exact_path_reads_frac": float(S[:, 26].sum() / max(1.0, reads)),
                                              "filtered_reads_frac": float(S[:, 27].sum() / max(1.0, reads))}
                                             if lix else {})),
        "autotune": tuned,
        "mapped_frac": round(mapped / reads, 4),
    }
    if not a.no_ingest:
        # the PCIe-inclusive rate: every run first copies its batch's bases and qualities to HBM
        # from pinned host memory (overlapped with other batches' stages)
        m.set("ingest", 1)
        el1, S1 = run_steps(m, p, n_batches, a.steps, a.warmup, sync, None)
        m.set("ingest", 0)
        out["with_ingest"] = {"value": round(S1[:, 1].sum() / el1, 1), "ms_per_step": round(1000 * el1 / a.steps, 3),
                              "ingest_bytes_per_step": float(S1[:, 36].mean()),
                              "ingest_ms_per_step": float(S1[:, 37].mean() / 1000)}
        log("with per-batch ingest: %s" % out["with_ingest"])
    if a.form != 0 and not a.no_compare:
        # the plain-FMD form on the same batches, same process, same clock
        saved = {k: m.get(k) for k in ("lanes_per_cu", "group_lanes", "text_path", "read_filter", "split_lanes",
                                       "pair_chains")}
        m.set("form", 0)
        lanes0 = None
        if not a.no_tune:
            lanes0, _ = m.tune(0, p)   # the plain form gets its own best setting
        pair0 = m.get("pair_chains")
        el0, S0 = run_steps(m, p, n_batches, a.steps, a.warmup, sync, None)
        v0 = S0[:, 1].sum() / el0
        out["plain_fmd"] = {"value": round(v0, 1), "ms_per_step": round(1000 * el0 / a.steps, 3),
                            "smem_kernel_ms_per_step": S0[:, 5].sum() / a.steps / 1000,
                            "stage_ms_per_step": {"seed_stage": S0[:, 8].sum() / a.steps / 1000,
                                                  "dp_kernel": S0[:, 6].sum() / a.steps / 1000,
                                                  "host_finish": S0[:, 9].sum() / a.steps / 1000},
                            "occ_extensions_per_read": S0[:, 2].sum() / S0[:, 0].sum(), "lanes_per_cu": lanes0,
                            "pair_chains": pair0}
        out["speedup_vs_plain_fmd"] = round(value / v0, 3)
        out["smem_stage_speedup_vs_plain_fmd"] = round(S0[:, 5].sum() / max(1.0, smem_us + S[:, 15].sum()), 3)
        log("plain FMD form: %.0f mapped reads/s; learned-index form / plain = %.3f" % (v0, value / v0))
        m.set("form", a.form)
        for k, v in saved.items():
            m.set(k, v)
    # HBM traffic and L2 hit rate of the same kernel on the same workload, from the committed
    # rocprofv3 --pmc summaries (profiles/*.md, lines '<key>[<kernel>@<workload>]: <value>')
    tag = "%s@%s" % (kname, a.workload)
    tr, tr_src = profile_value("traffic_bytes_per_launch", tag)
    if tr is not None:
        out["roofline"]["traffic"] = tr
        out["roofline"]["traffic_source"] = tr_src
        out["roofline"]["traffic_per_algorithmic_byte"] = round(tr / bytes_per_launch, 3)
    tg, _ = profile_value("traffic_bytes_per_launch_gather_scaled", tag)
    if tg is not None:
        # FETCH_SIZE scaled as for random 64-B gathers (factor 1.0) instead of the copy's 128-B
        # requests (factor 2.0): the lower end of the kernel's traffic
        out["roofline"]["traffic_gather_scaled"] = tg
        out["roofline"]["traffic_gather_scaled_per_algorithmic_byte"] = round(tg / bytes_per_launch, 3)
    hr, hr_src = profile_value("l2_hit_rate", tag)
    if hr is not None:
        out["roofline"]["l2_hit_rate"] = hr
        out["roofline"]["l2_hit_rate_source"] = hr_src
    if out["dp_gcups"]:
        ceil = VALU_PEAK_LANE_OPS / DP_OPS_PER_CELL / 1e9
        out["dp_roofline"] = {"bound": "valu", "achieved": out["dp_gcups"], "unit": "GCUPS",
                              "peak": round(ceil, 1), "frac": round(out["dp_gcups"] / ceil, 4),
                              "valu_peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
                              "algorithmic_int_ops_per_cell": DP_OPS_PER_CELL}
        ops, ops_src = profile_value("valu_lane_ops_per_cell", "dp@" + a.workload)
        if ops:
            out["dp_roofline"]["issued_lane_ops_per_cell"] = ops
            out["dp_roofline"]["issued_source"] = ops_src
    if not a.no_gather:
        # random 64-B line gathers over a buffer of the size the chosen kernel reads: the occ
        # table alone (occ-block extensions only) or occ + SA + ISA + BD text (unique-interval
        # text path, SA/ISA lookups)
        occ_bytes = 64 * (int(m.n) // 128 + 1)
        text_path = lookups > 0 or fast_steps > 0
        nb = occ_bytes + (9 * int(m.n) if text_path else 0)
        nb = (nb + 63) // 64 * 64
        us = m.random_gather(nb, 1 << 27)
        g = {"buffer": "occ + SA + ISA + BD" if text_path else "occ table", "buffer_bytes": nb,
             "gbs": round(64.0 * (1 << 27) / us / 1e3, 1)}
        out["roofline"]["random_gather_ceiling"] = g
        out["roofline"]["frac_of_gather_ceiling"] = round(achieved / g["gbs"], 4)
        log("random 64-B gather ceiling: %s" % g)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, chroms, names, batches[0], p, INSERT.get(a.workload), a.cpu_reads,
                                           a.cpu_runs, a.cpu_kind)
    if not a.no_clean:
        out["clean_seeding"] = clean_seeding(m, batches[0][0][0])
        log("clean seeding: %s" % out["clean_seeding"])
    if a.calib_copy:
        # a plain device copy of known size (k_calib_copy) to scale FETCH_SIZE / WRITE_SIZE
        out["calib_copy_bytes"] = 1 << 30
        out["calib_copy_us"] = m.calib_copy(1 << 30)
    print(json.dumps(out), flush=True)


def profile_value(key, tag):
    """A figure recorded in a committed rocprofv3 summary as a line '<key>[<tag>]: <value>'
    in profiles/*.md; (value, file) or (None, None)"""
    import glob
    import re
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(f):
            mt = re.match(re.escape(key) + r"\[(.+?)\]:\s*([0-9.eE+-]+)", line.strip())
            if mt and mt.group(1) == tag:
                return float(mt.group(2)), os.path.relpath(f, ROOT)
    return None, None


def smem_kernel_name(m, form):
    g = m.get("group_lanes") if form == 2 else 0
    if g:
        return "smem_group_kernel<%d>" % g
    fast = form == 1 or (form == 2 and m.get("text_path") == 1)
    name = "smem_%s<%s>" % ("pair_kernel" if m.get("pair_chains") else "kernel_t", "true" if fast else "false")
    s = m.get("split_lanes") if form == 2 else 0
    # split: a lane-group launch for the likely-long reads, then the one-lane launch
    return "smem_group_kernel<%d> + %s" % (s, name) if s else name


def run_steps(m, p, nb, steps, warmup, sync, log):
    # (run_steps.serial: submit each batch only after the previous one is finished;
    #  run_steps.depth: batches in flight at once)
    """Map `steps` batches (batch k = resident slot k % nb).  With two or more resident
    batches the runs are pipelined: up to `depth` submitted batches are in flight, so that
    the device stages of several batches overlap (worker streams) and their host
    finishing overlaps both.  Returns (seconds, per-step counters)."""
    depth = 1 if (nb == 1 or run_steps.serial) else max(1, min(nb, run_steps.depth))

    def pipelined(n, timed):
        out, inflight = [], collections.deque()
        for k in range(n):
            if len(inflight) >= depth:
                out.append(m.wait(inflight.popleft()))
            m.submit(k % nb, p, True)
            inflight.append(k % nb)
            if not timed and log:
                log("warmup step %d submitted" % k)
        while inflight:
            out.append(m.wait(inflight.popleft()))
        return out

    # warm-up at the timed depth over every resident batch, so that every buffer the
    # pipeline cycles through (device pools of the worker streams, pinned result buffers,
    # each slot's SAM text) exists before timing
    if warmup > 0:
        pipelined(max(warmup, depth + 1, min(nb, steps)), False)
    sync()
    t = time.time()
    stats = pipelined(steps, True)
    sync()
    return time.time() - t, np.array(stats)


def clean_seeding(m, seqs):
    """msgseqtk-clean's seeding (SMEM_LIST::findMEMS per read and the list's loglik,
    msgs_mems_batch) over the reads of batch 0 against the same database: reads/s on the host
    clock with the packing and upload of the reads and the copy back of the MEM records
    included (best of 3)"""
    import msgseqtk_amd as M
    p = M.default_params()
    best, cnt = 1e30, None
    for _ in range(3):
        t = time.time()
        cnt, _, _, _ = m.mems(seqs, p)
        best = min(best, time.time() - t)
    return {"value": round(len(seqs) / best, 1), "unit": "reads/s", "reads": len(seqs),
            "seconds": round(best, 4), "mems_per_read": round(float(cnt.mean()), 3),
            "timing": "host clock, reads packed and uploaded and MEM records copied back inside"}


def cpu_baseline(m, chroms, names, batch, p, insert, nreads, runs, kinds):
    """The CPU baseline over the first reads (pairs) of batch 0, OpenMP over all host cores
    (at most 16), median of `runs`: the reference's own seeding and chaining (oracle/_ref) with
    the oracle's DP and records ("reference"), and the oracle restatement with a plain occ table
    ("port").  Both SAM texts are compared with the GPU's; the oracle's count of FMD extensions
    on the sample is compared with the plain-form kernel's counter.  insert: the paired
    workload's (mean, sd), None for single-end reads; kinds: "both" or "port" (the oracle
    alone: the reference's RRR wavelet tree over a 6.8e9-symbol text takes minutes to build)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    import msgseqtk_amd as M
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(16, cores))
    bd, _, sa = m.arrays()
    t = time.time()
    odb = O.OracleDb(chroms, names, sa=sa)
    log("oracle index from the device suffix array in %.1fs" % (time.time() - t))
    paired = insert is not None
    rix = None
    if kinds == "both" and O.ref_available():
        t = time.time()
        rix = O.RefIndex(bd, sa)
        log("reference FMDIndex (RRR wavelet tree, sampled SA) in %.1fs" % (time.time() - t))
    del sa, bd
    reads = batch[0]
    mates = batch[1] if paired else None
    k = min(nreads // (2 if paired else 1), len(reads[0]))
    r1 = tuple(x[:k] for x in reads)
    r2 = tuple(x[:k] for x in mates) if mates else None
    po = O.pe_params(*insert) if paired else O.default_params()
    sam_gpu = m.map(r1, p, mates=r2)
    nr = k * (2 if paired else 1)
    res = {}
    for kind in (("reference", "port") if rix is not None else ("port",)):
        secs = []
        for _ in range(runs):
            if kind == "reference":
                sam, s = O.ref_baseline_map(rix, odb, po, *r1, mates=r2, nthreads=cores)
            else:
                sam, s = O.orc_map(odb, po, *r1, mates=r2, nthreads=cores)
            secs.append(s)
        med = float(np.median(secs))
        mapped = sum(1 for rec in sam for line in rec.splitlines() if not (int(line.split("\t")[1]) & 4))
        res[kind] = {"value": round(mapped / med, 1), "seconds": [round(x, 3) for x in secs],
                     "sam_identical_to_gpu": sam == sam_gpu}
        log("cpu baseline (%s): %d reads, median %.2fs on %d threads" % (kind, nr, med, cores))
    # FMD extensions: oracle count on the sample (first reads) against the plain kernel's counter
    seqs = r1[0][:2000]
    n_orc = O.orc_count_ext(odb, seqs, int(p.minLen), float(p.maxEvalue))
    saved = {kk: m.get(kk) for kk in ("form", "group_lanes", "pair_chains")}
    m.set("form", 0).set("group_lanes", 0).set("pair_chains", 0)
    m.upload(15, (seqs, [np.full(len(x), 30, np.uint8) for x in seqs], ["x%d" % i for i in range(len(seqs))]))
    m.run(15, M.default_params(minLen=p.minLen, maxEvalue=p.maxEvalue), False)
    n_gpu = int(m.stats()[2])
    for kk, v in saved.items():
        m.set(kk, v)
    main_kind = "reference" if "reference" in res else "port"
    out = {"value": res[main_kind]["value"], "unit": "mapped reads/s", "cores": cores, "kind": main_kind,
           "sample": "first %d %s of batch 0 (%d reads), median of %d runs, OpenMP per %s; %s" % (
               k, "pairs" if paired else "reads", nr, runs, "pair" if paired else "read",
               "the reference's own SMEM search, FMDIndex (RRR wavelet tree, LF-walk SA), SeedPair and SeedChain "
               "compiled from its sources, with the oracle's dense double DP and records (Alignment.cpp needs "
               "Eigen/htslib, absent)" if main_kind == "reference" else "oracle restatement, plain occ table, full SA"),
           "sam_identical_to_gpu": res[main_kind]["sam_identical_to_gpu"], "runs": res,
           "fmd_extensions_check": {"reads": len(seqs), "oracle": n_orc, "plain_kernel_counter": n_gpu}}
    return out


if __name__ == "__main__":
    main()
