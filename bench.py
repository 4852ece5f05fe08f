#!/usr/bin/env python3
"""Mapped reads/s of the MI355X mapping path on the paired-end workload of 64 genomes of
4 Mb (256 Mbp, GC ~ U(0.3, 0.7)), 2x150 bp pairs with insert N(350, 50), 0.5%
substitutions, 0.05% indels and 5% off-target pairs, in batches of 256k reads
(128k pairs) resident in HBM.  A step maps one batch end to end (SMEMs, seeds,
chains, DP + backtrace on the GPU, pairing/MAPQ/SAM records on the host).

Prints one JSON line (see --help).  The CPU baseline is the oracle restatement of the
reference's path timed on the host cores over a bounded sample of the same pairs.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--genomes", type=int, default=64)
    ap.add_argument("--genome-len", type=int, default=4_000_000)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--batch-reads", type=int, default=262144)
    ap.add_argument("--cpu-pairs", type=int, default=40000, help="pairs in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-clean", action="store_true", help="skip the msgseqtk-clean seeding measurement")
    ap.add_argument("--no-compare", action="store_true", help="skip the plain-FMD comparison run")
    ap.add_argument("--no-tune", action="store_true", help="keep the default lanes per CU")
    ap.add_argument("--no-torch", action="store_true", help="no torch device synchronisation (profiler runs)")
    ap.add_argument("--no-pipeline", action="store_true", help="finish each batch before submitting the next")
    ap.add_argument("--host-threads", type=int, default=0, help="host finishing threads (0: the library default)")
    ap.add_argument("--form", type=int, default=2,
                    help="0 plain FMD extensions, 1 unique-interval text path, 2 form 1 + learned-index read filter")
    ap.add_argument("--lix-segments", type=int, default=1 << 20, help="learned index: root segments")
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


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the mapping path shards by read batch with a replicated "
              "index, and this build measures the single-GPU kernel; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import msgseqtk_amd as M
    from msgseqtk_amd import synth

    t0 = time.time()
    chroms, names = synth.metagenome(a.genomes, a.genome_len, gc=(0.3, 0.7), seed=0xC0FFEE + 2)
    pairs_per_batch = a.batch_reads // 2
    n_batches = max(1, min((a.pairs + pairs_per_batch - 1) // pairs_per_batch, 16))
    batches = []
    for b in range(n_batches):
        batches.append(synth.simulate_pairs_fast(chroms, pairs_per_batch, 150, insert=(350, 50), sub=0.005,
                                                 indel=0.0005, seed=0xC0FFEE + 1002 + b, offtarget=0.05,
                                                 prefix="b%dp" % b))
    log("data: %d genomes, %d batches of %d pairs in %.1fs" % (len(chroms), n_batches, pairs_per_batch,
                                                              time.time() - t0))
    t1 = time.time()
    m = M.Mapper(0).build(chroms, names)
    build_s = time.time() - t1
    lix = None
    if a.form == 2:
        t3 = time.time()
        lix = m.lix_train(segments=a.lix_segments, samples=1 << 22, steps=a.lix_steps, batch=4096, lr=1e-3, seed=7)
        lix["train_s"] = round(time.time() - t3, 2)
        log("learned index trained in %.1fs: %s" % (lix["train_s"], lix))
    m.set("form", a.form)
    log("index built on the GPU in %.1fs (BD length %d)" % (build_s, m.n))
    for b, (r1, r2) in enumerate(batches):
        m.upload(b, r1, r2)
    p = M.default_params(insMean=350, insSD=50, insMin=350 - 10 * 50, insMax=350 + 10 * 50)
    sync = lambda: None   # msgs_batch_run returns after its stream has drained
    if not a.no_torch:
        try:
            import torch
            if torch.cuda.is_available():
                sync = torch.cuda.synchronize
        except Exception:
            pass

    tuned = None
    if a.lanes_per_cu:
        m.set("lanes_per_cu", a.lanes_per_cu)
    if a.group_lanes >= 0:
        m.set("group_lanes", a.group_lanes)
    if a.text_path >= 0:
        m.set("text_path", a.text_path)
    if a.read_filter >= 0:
        m.set("read_filter", a.read_filter)
    if a.split_lanes >= 0:
        m.set("split_lanes", a.split_lanes)
    if a.pair_chains >= 0:
        m.set("pair_chains", a.pair_chains)
    if not a.no_tune and not a.lanes_per_cu:
        lanes, tus = m.tune(0, p)
        tuned = {"lanes_per_cu": lanes, "group_lanes": m.get("group_lanes"), "text_path": m.get("text_path"),
                 "read_filter": m.get("read_filter"), "split_lanes": m.get("split_lanes"),
                 "pair_chains": m.get("pair_chains"), "smem_us": round(tus, 1)}
        log("autotuner: %d lanes per CU, %d lanes per read (%.0f us for the SMEM stage of batch 0)" % (
            lanes, tuned["group_lanes"], tus))
    m.set("streams", a.streams)
    run_steps.serial = a.no_pipeline
    run_steps.depth = a.depth
    if a.host_threads:
        m.set("host_threads", a.host_threads)
    elapsed, S = run_steps(m, p, n_batches, a.steps, a.warmup, sync, log)
    reads, mapped = S[:, 0].sum(), S[:, 1].sum()
    value = mapped / elapsed
    ext_calls, sa_lookups, cells = S[:, 2].sum(), S[:, 3].sum(), S[:, 4].sum()
    smem_us, dp_us = S[:, 5].sum(), S[:, 6].sum()
    # dominant kernel: SMEM search; algorithmic bytes = one 64-B occ block per rank side
    # (2 per extension with an unambiguous base)
    launches = a.steps  # one SMEM launch per step: the mates follow the reads in one batch
    fast_steps, lookups = S[:, 10].sum(), S[:, 11].sum()
    bytes_per_launch = (64.0 * 2.0 * ext_calls + 64.0 * lookups + fast_steps) / launches
    avg_us = smem_us / launches
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    log("timed %d steps in %.2fs: %.0f mapped reads/s (%.1f%% mapped)" % (a.steps, elapsed, value,
                                                                          100 * mapped / reads))
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "mapped reads/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1000 * elapsed / a.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "PE 2x150: 64 genomes x 4 Mb (256 Mbp, GC U(0.3,0.7)), insert N(350,50), "
                               "0.5% sub, 0.05% indel, 5% off-target pairs; 256k-read batches in HBM",
                   "batch_reads": 2 * pairs_per_batch, "bd_length": int(m.n),
                   "form": ["plain FMD", "unique-interval text path",
                            "learned-index exact-read path and read filter + unique-interval text path"][a.form]},
        "roofline": {"bound": "hbm", "kernel": smem_kernel_name(m, a.form), "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_launch": bytes_per_launch, "avg_launch_us": round(avg_us, 1)},
        "dp_gcups": round(cells / (dp_us * 1e-6) / 1e9, 3) if dp_us > 0 else None,
        "smem_iterations": {"lane_per_read": S[:, 12].sum() / reads,
                            "simd_utilisation": S[:, 12].sum() / max(1.0, 64.0 * S[:, 13].sum()),
                            "max_wavefront": int(S[:, 14].max()),
                            "us_per_iteration_of_slowest_wavefront": smem_us / a.steps / max(1.0, S[:, 14].mean())},
        "per_read": {"occ_extensions": ext_calls / reads, "text_steps": fast_steps / reads,
                     "smem_sa_isa_lookups": lookups / reads, "seed_sa_lookups": sa_lookups / reads,
                     "dp_cells": cells / reads},
        "stage_ms_per_step": {"read_filter": S[:, 15].sum() / a.steps / 1000,
                              "smem_kernel": smem_us / a.steps / 1000, "seed_stage": S[:, 8].sum() / a.steps / 1000,
                              "dp_kernel": dp_us / a.steps / 1000, "host_finish": S[:, 9].sum() / a.steps / 1000,
                              "whole": S[:, 7].sum() / a.steps / 1000},
        "host_side_ms_per_step": dict(run_steps.split, **{
            "gpu_stages_wall": S[:, 16].sum() / a.steps / 1000, "smem_wall": S[:, 17].sum() / a.steps / 1000,
            "seeds_wall": S[:, 18].sum() / a.steps / 1000, "chains_wall": S[:, 19].sum() / a.steps / 1000,
            "dp_wall": S[:, 20].sum() / a.steps / 1000,
            "dp_pools_ready": S[:, 23].sum() / a.steps / 1000, "dp_records_back": S[:, 24].sum() / a.steps / 1000,
            "dp_cigar_md_back": S[:, 25].sum() / a.steps / 1000}),
        "dp_jobs_per_step": {"all": float(S[:, 21].mean()), "wavefront_fill": float(S[:, 22].mean()),
                             "serial_fill": float(S[:, 32].mean()), "serial_fill_cells": float(S[:, 33].mean()),
                             "largest_serial_fill_cells": float(S[:, 34].max()),
                             "full_last_row_scans": float(S[:, 35].mean())},
        "host_finish_thread_ms_per_step": {k: S[:, 28 + i].sum() / a.steps / 1000 for i, k in
                                           enumerate(("sam_text", "pairing", "mapq", "alignment_lists"))},
        "index_build_s": round(build_s, 2),
        "learned_index": dict(lix or {}, **({"exact_path_reads_frac": float(S[:, 26].sum() / max(1.0, S[:, 0].sum())),
                                              "filtered_reads_frac": float(S[:, 27].sum() / max(1.0, S[:, 0].sum()))}
                                             if lix else {})),
        "autotune": tuned,
        "mapped_frac": round(mapped / reads, 4),
    }
    if a.form != 0 and not a.no_compare:
        # the plain-FMD form on the same batches, same process, same clock
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
                            "host_side_ms_per_step": {
                                "gpu_stages_wall": S0[:, 16].sum() / a.steps / 1000,
                                "smem_wall": S0[:, 17].sum() / a.steps / 1000,
                                "seeds_wall": S0[:, 18].sum() / a.steps / 1000,
                                "chains_wall": S0[:, 19].sum() / a.steps / 1000,
                                "dp_wall": S0[:, 20].sum() / a.steps / 1000},
                            "occ_extensions_per_read": S0[:, 2].sum() / S0[:, 0].sum(), "lanes_per_cu": lanes0,
                            "pair_chains": pair0}
        out["speedup_vs_plain_fmd"] = round(value / v0, 3)
        out["smem_stage_speedup_vs_plain_fmd"] = round((S0[:, 5].sum() + 0.0) / max(1.0, smem_us + S[:, 15].sum()), 3)
        log("plain FMD form: %.0f mapped reads/s; learned-index form / plain = %.3f" % (v0, value / v0))
        m.set("form", a.form)
        if tuned:
            m.set("lanes_per_cu", tuned["lanes_per_cu"]).set("group_lanes", tuned["group_lanes"])
            m.set("text_path", tuned["text_path"]).set("read_filter", tuned["read_filter"])
            m.set("split_lanes", tuned["split_lanes"]).set("pair_chains", tuned["pair_chains"])
    tr, tr_src = measured_traffic(out["roofline"]["kernel"])
    if tr is not None:
        out["roofline"]["traffic"] = tr
        out["roofline"]["traffic_source"] = tr_src
    hr, hr_src = profile_value("l2_hit_rate", out["roofline"]["kernel"])
    if hr is not None:
        out["roofline"]["l2_hit_rate"] = hr
        out["roofline"]["l2_hit_rate_source"] = hr_src
    # DP: GCUPS against the integer VALU peak divided by the VALU lane operations our
    # kernels issue per cell (SQ_INSTS_VALU x 64 over dp_fill_kernel + dp_trace_kernel,
    # from a committed rocprofv3 --pmc summary)
    ops, ops_src = profile_value("valu_lane_ops_per_cell", "dp")
    if ops and out["dp_gcups"]:
        ceil = VALU_PEAK_LANE_OPS / ops / 1e9
        out["dp_roofline"] = {"bound": "valu", "achieved": out["dp_gcups"], "unit": "GCUPS",
                              "peak": round(ceil, 1), "frac": round(out["dp_gcups"] / ceil, 4),
                              "valu_peak_lane_ops_per_s": VALU_PEAK_LANE_OPS, "lane_ops_per_cell": ops,
                              "source": ops_src}
    if not a.no_gather:
        # random 64-B line gathers over buffers the size of the occ table and of occ + SA + ISA:
        # the random-access ceiling the SMEM kernel's lookups are held against
        occ_bytes = 64 * (int(m.n) // 128 + 1)
        g = {}
        for name, nb in (("occ_table", occ_bytes), ("occ_sa_isa", occ_bytes + 8 * int(m.n))):
            nb = (nb + 63) // 64 * 64
            us = m.random_gather(nb, 1 << 27)
            g[name] = {"buffer_bytes": nb, "gbs": round(64.0 * (1 << 27) / us / 1e3, 1)}
        out["roofline"]["random_gather_ceiling"] = g
        out["roofline"]["frac_of_gather_ceiling"] = round(achieved / g["occ_sa_isa"]["gbs"], 4)
        log("random 64-B gather ceiling: %s" % g)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, chroms, names, batches[0], p, a.cpu_pairs)
    if not a.no_clean:
        out["clean_seeding"] = clean_seeding(m, batches[0])
        log("clean seeding: %s" % out["clean_seeding"])
    if a.calib_copy:
        # a plain device copy of known size (k_calib_copy) to scale FETCH_SIZE / WRITE_SIZE
        out["calib_copy_bytes"] = 1 << 30
        out["calib_copy_us"] = m.calib_copy(1 << 30)
    print(json.dumps(out), flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` as recorded in a committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, scaled by the same counters on a copy of known size):
    a line 'traffic_bytes_per_launch[<kernel>]: <bytes>' in profiles/*.md"""
    return profile_value("traffic_bytes_per_launch", kernel)


def profile_value(key, kernel):
    """A figure recorded in a committed rocprofv3 summary as a line '<key>[<kernel>]: <value>'
    in profiles/*.md; (value, file) or (None, None)"""
    import glob
    import re
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(f):
            mt = re.match(re.escape(key) + r"\[(.+?)\]:\s*([0-9.eE+-]+)", line.strip())
            if mt and mt.group(1) == kernel:
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
        ts = tw = 0.0
        for k in range(n):
            if len(inflight) >= depth:
                t1 = time.time()
                out.append(m.wait(inflight.popleft()))
                tw += time.time() - t1
            t0 = time.time()
            m.submit(k % nb, p, True)
            ts += time.time() - t0
            inflight.append(k % nb)
            if not timed and log:
                log("warmup step %d submitted" % k)
        while inflight:
            t1 = time.time()
            out.append(m.wait(inflight.popleft()))
            tw += time.time() - t1
        return out, ts, tw

    # warm-up at the timed depth over every resident batch, so that every buffer the
    # pipeline cycles through (device pools of both worker streams, pinned result buffers,
    # each slot's SAM text) exists before timing
    if warmup > 0:
        pipelined(max(warmup, depth + 1, min(nb, steps)), False)
    sync()
    t = time.time()
    stats, tsub, twait = pipelined(steps, True)
    sync()
    run_steps.split = {"submit_ms_per_step": 1000 * tsub / steps, "wait_ms_per_step": 1000 * twait / steps}
    return time.time() - t, np.array(stats)


def clean_seeding(m, batch):
    """msgseqtk-clean's seeding (SMEM_LIST::findMEMS per read and the list's loglik,
    msgs_mems_batch) over the first reads of the pairs of batch 0 against the same database:
    reads/s on the host clock with the packing and upload of the reads and the copy back of
    the MEM records included (best of 3)"""
    import msgseqtk_amd as M
    seqs = batch[0][0]
    p = M.default_params()
    best, cnt = 1e30, None
    for _ in range(3):
        t = time.time()
        cnt, _, _, _ = m.mems(seqs, p)
        best = min(best, time.time() - t)
    return {"value": round(len(seqs) / best, 1), "unit": "reads/s", "reads": len(seqs),
            "seconds": round(best, 4), "mems_per_read": round(float(cnt.mean()), 3),
            "timing": "host clock, reads packed and uploaded and MEM records copied back inside"}


def cpu_baseline(m, chroms, names, batch, p, npairs):
    """The oracle (CPU restatement of the reference path) over the first npairs pairs of
    batch 0, OpenMP over all host cores; its SAM text is also compared with the GPU's."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    _, _, sa = m.arrays()
    odb = O.OracleDb(chroms, names, sa=sa)
    del sa
    (s1, q1, n1), (s2, q2, n2) = batch
    k = min(npairs, len(s1))
    po = O.pe_params(350, 50)
    sam_cpu, secs = O.orc_map(odb, po, s1[:k], q1[:k], n1[:k], mates=(s2[:k], q2[:k], n2[:k]), nthreads=cores)
    sam_gpu = m.map((s1[:k], q1[:k], n1[:k]), p, mates=(s2[:k], q2[:k], n2[:k]))
    mapped = sum(1 for rec in sam_cpu for line in rec.splitlines() if not (int(line.split("\t")[1]) & 4))
    log("cpu baseline: %d pairs in %.2fs on %d threads" % (k, secs, cores))
    return {"value": round(mapped / secs, 1), "unit": "mapped reads/s", "cores": cores, "kind": "port",
            "sample": "first %d pairs (%d reads) of batch 0, oracle restatement with full SA, OpenMP per pair" % (k, 2 * k),
            "seconds": round(secs, 2), "sam_identical_to_gpu": sam_cpu == sam_gpu}


if __name__ == "__main__":
    main()
