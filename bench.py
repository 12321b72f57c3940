#!/usr/bin/env python3
"""Benchmark of the MI355X uplink LDPC decode path (one GPU).

Workload (the metric's configuration): the heterogeneous slot -- 16 UEs,
each TB with BG in {1,2}, Zc uniform over the lifting sizes 64..384, code
rate U[0.2, 0.93], 1-24 code blocks, Es/N0 U[-1, 8] dB, BPSK-AWGN LLRs
stored binary16, 10 layered iterations, no early termination.  Synthetic
data generated on the device from fixed seeds.

A step decodes the whole slot.  Forms:
  plain   -- one cuphyErrorCorrectionLDPCTransportBlockDecode launch per TB
  batched -- the whole slot in one persistent launch (when available)
value = decoded info bits of the slot / step time, in Gbps, for the best
form; every form's number is reported.  The CPU baseline decodes the same
slot's code blocks with the oracle (binary32 mode) on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "aerial-cuda-accelerated-ran_amd"))

METRIC = "decoded info-bit Gbps @10 iters (batched vs plain); p99 TB latency vs 500µs"
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 32-bit VALU lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
OPS_PER_EDGE = 8                           # VALU lane-ops per edge update (algorithmic count)
ITERS = 10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU work budget of the baseline leg")
    return ap.parse_args()


def cpu_baseline(tbs, budget_s):
    """Oracle (binary32 layered decoder, C) over the slot's code blocks, on
    the host cores, one thread per core up to 16.  Bounded: as many whole
    passes over the slot as fit in the CPU-work budget (at least one)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ncores = min(16, os.cpu_count() or 1)
    jobs = []
    for tb in tbs:
        cfg = O.dec_cfg(tb.bg, tb.Z, tb.mb, n_iter=ITERS, alpha=0.8125, clamp=32.0, mode=O.FP32)
        llr = tb.llr.float().cpu().numpy()[:, :tb.num_llrs].copy()
        jobs.append((cfg, llr, tb.ncb * tb.info_bits_per_cb))
    bits_per_pass = sum(j[2] for j in jobs)
    passes, t_total = 0, 0.0
    while True:
        t0 = time.perf_counter()
        for cfg, llr, _ in jobs:
            O.decode_cbs(cfg, llr, nthreads=ncores)
        t_total += time.perf_counter() - t0
        passes += 1
        if t_total * ncores >= budget_s or passes >= 50:
            break
    gbps = bits_per_pass * passes / t_total / 1e9
    return {"value": round(gbps, 4), "unit": "Gbps", "cores": ncores, "kind": "port",
            "sample": "%d pass(es) over the slot's %d code blocks (%d info bits/pass), oracle binary32 "
                      "layered NMS, 10 iterations, one thread per core" %
                      (passes, sum(tb.ncb for tb in tbs), bits_per_pass)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the decode path shards by cell with no collective -- see the design notes)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch
    from aerial_amd import _lib as L
    from aerial_amd import workloads as W
    from aerial_amd.decoder import LdpcDecodeConfig, LdpcDecoder

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    dec = LdpcDecoder()

    tbs = W.heterogeneous_slot(seed=2)
    descs = []
    for tb in tbs:
        cfg = LdpcDecodeConfig(tb.bg, tb.Z, tb.mb, llr_type=L.CUPHY_R_16F, max_iterations=ITERS, alpha=0.8125,
                               clamp_value=32.0)
        descs.append(dec.build_desc(cfg, [tb.llr], [tb.out]))
    info_bits = sum(tb.ncb * tb.info_bits_per_cb for tb in tbs)
    edge_updates = sum(tb.edge_updates(ITERS) for tb in tbs)
    ncb = sum(tb.ncb for tb in tbs)
    torch.cuda.synchronize()

    # ---- plain form: one launch per TB, per-launch events -----------------
    def plain_step(events=None):
        for i, d in enumerate(descs):
            if events is not None:
                events[i][0].record(stream)
            dec.launch(d, stream)
            if events is not None:
                events[i][1].record(stream)

    for _ in range(a.warmup):
        plain_step()
    torch.cuda.synchronize()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in descs]
          for _ in range(a.steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for s in range(a.steps):
        plain_step(ev[s])
    t1.record(stream)
    torch.cuda.synchronize()
    plain_ms = t0.elapsed_time(t1) / a.steps
    kernel_ms = sum(e0.elapsed_time(e1) for step in ev for e0, e1 in step) / a.steps  # per step
    plain_gbps = info_bits / (plain_ms * 1e-3) / 1e9

    # correctness spot check of the decoded bits against the transmitted ones
    bit_errs, blk_errs = 0, 0
    for tb in tbs:
        bits = tb.out.view(torch.int32)
        got = ((bits.unsqueeze(-1) >> torch.arange(32, device="cuda", dtype=torch.int32)) & 1).reshape(tb.ncb, -1)
        got = got[:, :tb.info_bits_per_cb].to(torch.uint8)
        diff = (got != tb.info)
        bit_errs += int(diff.sum().item())
        blk_errs += int(diff.any(dim=1).sum().item())

    # ---- batched form: one persistent launch per slot, graph-replayed -------
    from aerial_amd.batch import BatchDecoder
    for tb in tbs:
        tb.out.zero_()
    bdec = BatchDecoder(dec).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=ITERS,
                                   alpha=0.8125, clamp_value=32.0)
    bdec.capture()
    for _ in range(a.warmup):
        bdec.replay()
    torch.cuda.synchronize()
    bev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    gstream = torch.cuda.current_stream()
    for s_ in range(a.steps):
        bev[s_][0].record(gstream)
        bdec.replay()
        bev[s_][1].record(gstream)
    torch.cuda.synchronize()
    batched_ms = sum(e0.elapsed_time(e1) for e0, e1 in bev) / a.steps
    batched_gbps = info_bits / (batched_ms * 1e-3) / 1e9
    bbit_errs = 0
    for tb in tbs:
        bits = tb.out.view(torch.int32)
        got = ((bits.unsqueeze(-1) >> torch.arange(32, device="cuda", dtype=torch.int32)) & 1).reshape(tb.ncb, -1)
        bbit_errs += int((got[:, :tb.info_bits_per_cb].to(torch.uint8) != tb.info).sum().item())

    # ---- roofline of the decoder kernel (VALU-bound) ----------------------
    # dominant kernel = the batched persistent decoder: one launch per step
    achieved_tops = edge_updates * OPS_PER_EDGE / (batched_ms * 1e-3) / 1e12
    plain_tops = edge_updates * OPS_PER_EDGE / (kernel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved_tops, 3), "peak": round(VALU_PEAK_TOPS, 2),
                "unit": "TOP/s", "frac": round(achieved_tops / VALU_PEAK_TOPS, 4), "traffic": None,
                "kernel": "nrldpc::ldpc_layered_persistent<_Float16, _Float16>",
                "work": "%d edge updates x %d VALU ops per launch" % (edge_updates, OPS_PER_EDGE),
                "avg_launch_us": round(batched_ms * 1e3, 2),
                "plain_kernel": {"achieved": round(plain_tops, 3), "avg_launch_us": round(kernel_ms * 1e3 / len(descs), 2)}}

    # ---- configuration 1 (plain, binary32): GPU time -----------------------
    c1 = W.config1_tb(seed=1)
    c1cfg = LdpcDecodeConfig(1, 384, 46, llr_type=L.CUPHY_R_32F, max_iterations=ITERS, alpha=0.8125)
    c1desc = dec.build_desc(c1cfg, [c1.llr], [c1.out])
    for _ in range(3):
        dec.launch(c1desc, stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    n_c1 = 20
    for _ in range(n_c1):
        dec.launch(c1desc, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    c1_us = e0.elapsed_time(e1) * 1e3 / n_c1
    c1_gbps = 8 * 8448 / (c1_us * 1e-6) / 1e9

    # ---- 20-cell full-load slot (2800 CBs, BG1 Z=384, R=0.93): both forms ---
    c3 = W.full_load_slot(seed=3)
    c3_bits = sum(tb.ncb * tb.info_bits_per_cb for tb in c3)
    c3_descs = [dec.build_desc(LdpcDecodeConfig(tb.bg, tb.Z, tb.mb, llr_type=L.CUPHY_R_16F, max_iterations=ITERS,
                                                alpha=0.8125, clamp_value=32.0), [tb.llr], [tb.out]) for tb in c3]
    c3b = BatchDecoder(dec).build(c3, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=ITERS, alpha=0.8125)
    c3b.capture()

    def timed(fn, n):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    c3_plain_ms = timed(lambda: [dec.launch(d, stream) for d in c3_descs], max(2, a.steps // 4))
    c3_batch_ms = timed(c3b.replay, max(2, a.steps // 2))
    c3_edges = sum(tb.edge_updates(ITERS) for tb in c3)
    config3 = {"code_blocks": sum(tb.ncb for tb in c3), "info_bits": c3_bits,
               "plain_ms": round(c3_plain_ms, 4), "plain_gbps": round(c3_bits / (c3_plain_ms * 1e-3) / 1e9, 3),
               "batched_ms": round(c3_batch_ms, 4), "batched_gbps": round(c3_bits / (c3_batch_ms * 1e-3) / 1e9, 3),
               "batched_valu_frac": round(c3_edges * OPS_PER_EDGE / (c3_batch_ms * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
               "slot_budget_ms": 0.5}
    del c3, c3b

    cpu = None if a.no_cpu else cpu_baseline(tbs, a.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(batched_gbps, 4),
        "unit": "Gbps",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(batched_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic (seeded device RNG, HIP-encoded code blocks, BPSK-AWGN)",
        "config": {"workload": "heterogeneous slot: 16 UEs, BG1/BG2, Zc 64..384, R 0.2..0.93, 1-24 CBs/TB, "
                               "Es/N0 -1..8 dB, fp16 LLRs, 10 iterations",
                   "form": "batched (one persistent launch per slot, graph replay)", "code_blocks": ncb,
                   "info_bits": info_bits,
                   "tbs": len(tbs), "parallelism": "1 GPU"},
        "forms": {"plain_gbps": round(plain_gbps, 4), "plain_ms": round(plain_ms, 4),
                  "batched_gbps": round(batched_gbps, 4), "batched_ms": round(batched_ms, 4),
                  "batched_over_plain": round(batched_gbps / plain_gbps, 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "check": {"plain_bit_errors": bit_errs, "plain_block_errors": blk_errs, "batched_bit_errors": bbit_errs,
                  "code_blocks": ncb},
        "config1_plain_fp32": {"us_per_tb": round(c1_us, 2), "gbps": round(c1_gbps, 4)},
        "config3_full_load_20_cells": config3,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
