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
slot's code blocks, and configuration 1 (8 CBs BG1 Zc=384), with the oracle
(binary32 mode, built -O3 -march=native on the host) on pinned host threads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "aerial-cuda-accelerated-ran_amd"))

METRIC = "decoded info-bit Gbps @10 iters (batched vs plain); p99 TB latency vs 500µs"
# 32-bit VALU lane-ops/s: 256 CUs x 128 lane-ops per clock x 2.4 GHz (the published 157.3 TFLOP/s FP32
# vector rate counts an FMA as 2 flops; a v_add_u32 issue probe, decode_bench valupeak, measures 62.6 T)
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
OPS_PER_EDGE = 8                           # VALU lane-ops per edge update (algorithmic count)
# LDS bytes per edge update with binary16 APP values: one read and one write of 2 B
LDS_BYTES_PER_EDGE = 4
# LDS bandwidth: 256 CUs x 128 B per clock x 2.4 GHz (nominal)
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12
ITERS = 10


def progress(msg):
    """One progress line on stderr (the JSON result stays alone on stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU work budget of the baseline leg")
    ap.add_argument("--sections", default="main,config1,config3,throughput,fused,serving",
                    help="comma-separated subset of: main, config1, config3, throughput, fused, serving; or 'copyprobe' "
                         "alone: three 1 GiB device copies (calibrates HBM byte counters)")
    ap.add_argument("--serve-ms", type=float, default=40.0, help="arrival window of each serving load point")
    ap.add_argument("--serve-inflight", type=int, default=3, help="fused launches in flight in the serving engine")
    ap.add_argument("--serve-loads", default="0.6,0.9,1.2", help="offered loads as fractions of capacity")
    return ap.parse_args()


def fused_slot_section(dec, steps):
    """64-cell fused receive-side slot (int8 LLRs, one persistent graph per
    slot: de-rate-matching + HARQ write + decode with early termination + CB
    and TB CRC), first transmission of every TB."""
    import numpy as np
    import torch
    from aerial_amd import pusch_workloads as PW
    from aerial_amd.slot import SlotDecoder, SlotTb
    from aerial_amd.workloads import num_edges

    specs = PW.fused_slot_specs(seed=5)
    tbs = PW.build_tbs(specs, seed=5)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    slot = []
    for tb in tbs:
        llr = PW.tx_llrs(tb, 0, gen, mode="cg", dtype=torch.int8)
        p, mb = tb.params(0, 1)
        harq = torch.zeros(tb.C * p.Ncb_padded, dtype=torch.float16, device="cuda")
        slot.append(SlotTb(p, llr, harq, mb, max_iterations=8, early_term=True, llr_scale=0.25))
    stream = torch.cuda.current_stream()

    def time_slot(concurrent):
        sd = SlotDecoder(dec, concurrent=concurrent).build(slot)
        sd.launch()
        torch.cuda.synchronize()
        sd.capture()
        for _ in range(2):
            sd.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        n = max(3, steps // 2)
        e0.record(stream)
        for _ in range(n):
            sd.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    ms_conc = time_slot(True)
    ms_seq = time_slot(False)
    ms = min(ms_conc, ms_seq)
    info_bits = sum(tb.info_bits for tb in tbs)
    ncb = sum(tb.C for tb in tbs)
    cb_fail = sum(int((st.cb_crc != 0).sum().item()) for st in slot)
    tb_fail = sum(int((st.tb_crc != 0).sum().item()) for st in slot)
    payload_err = 0
    for tb, st in zip(tbs, slot):
        got = np.unpackbits(st.tb_bytes.cpu().numpy())[:tb.A]
        payload_err += int(not np.array_equal(got, tb.payload))
    iters = torch.cat([st.cb_iterations.to(torch.int64) for st in slot])
    edges = 0
    for tb, st in zip(tbs, slot):
        edges += tb.Zc * num_edges(tb.bg, st.mb) * int(st.cb_iterations.to(torch.int64).sum().item())
    hbm = 0
    for tb, st in zip(tbs, slot):
        p = st.params
        hbm += tb.G // tb.Qm * 8 + tb.C * (2 * (2 * p.Zc + p.Ncb) + 4 * ((tb.Kp + 31) // 32) + (tb.Kp + 7) // 8)
    return {"workload": "64 cells x (140 CBs BG1 Zc=384 R=0.93 + 28 BG2 CBs Zc 104/208 R=0.5), int8 LLRs "
                        "(4 steps/unit, +-127), consistent-Gaussian mu = SNR/5 at SNR U[18,24] dB, 8 iterations "
                        "with early termination, first transmission",
            "code_blocks": ncb, "tbs": len(tbs), "info_bits": info_bits, "ms_per_slot": round(ms, 4),
            "ms_variants_concurrent_sequential": [round(ms_conc, 4), round(ms_seq, 4)],
            "gbps": round(info_bits / (ms * 1e-3) / 1e9, 3), "slot_budget_ms": 0.5,
            "mean_iterations": round(float(iters.float().mean().item()), 3),
            "cb_crc_failures": cb_fail, "tb_crc_failures": tb_fail, "tb_payload_mismatches": payload_err,
            "valu_frac": round(edges * OPS_PER_EDGE / (ms * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
            "hbm_gbs": round(hbm / (ms * 1e-3) / 1e9, 1), "hbm_frac": round(hbm / (ms * 1e-3) / 8e12, 4)}


def serving_section(dec, serve_ms, config3_gbps=None, inflight=3, loads=(0.6, 0.9, 1.2), resident=True,
                    templates=None, config3_cbs_per_s=None):
    """Serving loop: Poisson TB arrivals from 32 cells (TB shapes as in the
    heterogeneous slot, binary16 BPSK-AWGN LLRs), HARQ with rv 0, 2, 3, 1
    and up to 4 transmissions, retransmissions released one HARQ round trip
    (2 ms) after a failing TB CRC and soft-combined into device-resident
    HARQ buffers (8 GiB pool, admission by that budget), syndrome early
    termination within 10 iterations.  resident: one resident launch takes
    code blocks as they are admitted; otherwise continuous batching over
    `inflight` fused launches.  Saturation capacity first (all TBs at
    t = 0), then offered loads as fractions of it; latency = completion seen
    by the engine - release of the transmission."""
    from aerial_amd.server import HarqServer, TraceTemplates, build_trace, poisson_arrivals, serving_specs, trace_report
    n_templates = 256
    if templates is None:
        templates = TraceTemplates(serving_specs(n_templates, seed=41), seed=41)
    srv = HarqServer(dec, pool_bytes=8 << 30, max_inflight=inflight, max_tbs=2048, max_cbs=16384, resident=resident)
    rtt = 2000.0
    mean_bits = float(np.mean([tb.info_bits for tb in templates.tbs]))
    mean_cbs = float(np.mean([tb.C for tb in templates.tbs]))
    n_sat = 1500
    sat = [(0.0, i % 32, i % n_templates) for i in range(n_sat)]
    entries, outs, meta = build_trace(srv, templates, sat)
    done = srv.serve_trace(entries, rtt)
    rep = trace_report(done, meta, templates, sat)
    first_tx_span_us = max(done[4 * a].t_done_us for a in range(n_sat))
    cap_tbs = n_sat / (first_tx_span_us * 1e-6)
    out = {"saturation": rep, "capacity_first_tx_tbs_per_s": round(cap_tbs, 1),
           "capacity_first_tx_gbps": round(cap_tbs * mean_bits / 1e9, 3),
           "capacity_first_tx_cbs_per_s": round(cap_tbs * mean_cbs, 1), "loads": []}
    loads = list(loads)
    extra = bool(config3_cbs_per_s and resident)
    if extra:  # one more point: 60 % of the 20-cell full-load slot's code blocks per second
        loads.append(0.6 * config3_cbs_per_s / (cap_tbs * mean_cbs))
    for i, x in enumerate(loads):
        arr = poisson_arrivals(32, x * cap_tbs, serve_ms * 1e3, n_templates, seed=int(100 * x))
        entries, outs, meta = build_trace(srv, templates, arr)
        done = srv.serve_trace(entries, rtt)
        r = trace_report(done, meta, templates, arr)
        r["offered_load"] = round(x, 4)
        r["offered_gbps"] = round(x * cap_tbs * mean_bits / 1e9, 3)
        if config3_gbps:
            r["offered_vs_config3"] = round(x * cap_tbs * mean_bits / 1e9 / config3_gbps, 4)
        if config3_cbs_per_s:
            r["offered_cbs_vs_config3"] = round(x * cap_tbs * mean_cbs / config3_cbs_per_s, 4)
        r["engine"] = srv.stats()
        out["loads"].append(r)
        if extra and i == len(loads) - 1:
            out["at_60pct_of_config3_cbs"] = r
        print("serving (%s) load %.3f: %s" % ("resident" if resident else "batched", x, json.dumps(r)),
              file=sys.stderr, flush=True)
    srv.close()
    out["workload"] = ("32 cells, Poisson TB arrivals, TB shapes as the heterogeneous slot (256 templates), "
                       "rv 0,2,3,1 up to 4 transmissions, HARQ RTT 2 ms, fp16 HARQ buffers in an 8 GiB pool, "
                       "early termination <= 10 iterations, %s; %.0f ms per load"
                       % ("one resident launch taking code blocks as admitted" if resident else
                          "%d in-flight batched fused launches" % inflight, serve_ms))
    out["templates"] = templates
    return out


def native_oracle():
    """The oracle built -O3 -march=native for the host this runs on (make
    -C oracle native, into a temporary directory): its path and the flags,
    or (None, reason) when the build fails."""
    import subprocess
    import tempfile
    d = tempfile.mkdtemp(prefix="nr_oracle_native_")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", "NATIVE_DIR=" + d],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return None, (r.stderr or r.stdout)[-300:]
    return os.path.join(d, "libnr_oracle.so"), "gcc -O3 -march=native -ffp-contract=off"


def host_cpu():
    """(lscpu model name, logical CPUs this process may use)."""
    import subprocess
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), "")
    except (OSError, subprocess.SubprocessError):
        pass
    if not model:
        try:
            with open("/proc/cpuinfo") as f:
                model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(tbs, c1, c1_gpu, budget_s):
    """The oracle (binary32 layered normalized min-sum, C, built -O3
    -march=native on this host) on the host cores, threads pinned one per
    logical CPU, 2 untimed warm-up passes then >= 20 timed ones, median.

    value: the slot's code blocks (all TBs' code blocks from one queue) on
    min(16, CPUs) threads, in decoded info-bit Gbps.  config1: 1 TB x 8
    CBs BG1 Zc=384 on 8 threads and on 1 thread, gated on its hard decisions
    being bit-identical to the GPU's binary32 plain-form decode (c1.out)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    path, flags = native_oracle()
    if path is None:
        O.lib()
        flags = "portable build (-O3 -march=x86-64-v2); native build failed: " + flags
    else:
        O.use_library(path)
    model, nproc = host_cpu()
    ncores = min(16, nproc)

    cfgs, llrs = [], []
    for tb in tbs:
        cfgs.append(O.dec_cfg(tb.bg, tb.Z, tb.mb, n_iter=ITERS, alpha=0.8125, clamp=32.0, mode=O.FP32))
        llrs.append(tb.llr.float().cpu().numpy()[:, :tb.num_llrs].copy())
    bits_per_pass = sum(tb.ncb * tb.info_bits_per_cb for tb in tbs)
    t_one, _ = O.time_slot(cfgs, llrs, ncores, warmup=1, reps=1)
    reps = int(min(200, max(20, budget_s / max(1e-6, t_one[0] * ncores))))
    times, _ = O.time_slot(cfgs, llrs, ncores, warmup=2, reps=reps)
    t_med = float(np.median(times))
    times_1, _ = O.time_slot(cfgs, llrs, 1, warmup=1, reps=3)
    out = {"value": round(bits_per_pass / t_med / 1e9, 4), "unit": "Gbps", "cores": ncores, "kind": "port",
           "value_1_core": round(bits_per_pass / float(np.median(times_1)) / 1e9, 5),
           "cpu_model": model, "nproc": nproc, "build": flags,
           "sample": "%d timed passes (median; 2 warm-up) over the slot's %d code blocks (%d info bits/pass), "
                     "oracle binary32 layered NMS, 10 iterations, %d pinned threads; 1-core: median of 3 passes"
                     % (reps, sum(tb.ncb for tb in tbs), bits_per_pass, ncores)}
    if c1 is not None:
        c1cfg = O.dec_cfg(1, 384, 46, n_iter=ITERS, alpha=0.8125, clamp=32.0, mode=O.FP32)
        c1llr = c1.llr.float().cpu().numpy()[:, :c1.num_llrs].copy()
        t8, hard8 = O.time_decode(c1cfg, c1llr, 8, warmup=2, reps=20)
        t1, hard1 = O.time_decode(c1cfg, c1llr, 1, warmup=2, reps=20)
        gpu_words = c1.out.cpu().numpy().view(np.uint32)[:, :hard8.shape[1]]
        identical = bool((hard8 == gpu_words).all() and (hard1 == gpu_words).all())
        g8 = 8 * 8448 / float(np.median(t8)) / 1e9
        g1 = 8 * 8448 / float(np.median(t1)) / 1e9
        out["config1"] = {"gbps_8_threads": round(g8, 5), "gbps_1_thread": round(g1, 5),
                          "median_ms_8_threads": round(float(np.median(t8)) * 1e3, 3),
                          "median_ms_1_thread": round(float(np.median(t1)) * 1e3, 3),
                          "gpu_plain_gbps": c1_gpu["gbps"] if c1_gpu else None,
                          "gpu_over_cpu_8_threads": round(c1_gpu["gbps"] / g8, 1) if c1_gpu else None,
                          "hard_decisions_identical_to_gpu": identical,
                          "sample": "1 TB x 8 CBs BG1 Zc=384 (67584 info bits), binary32, 10 iterations; "
                                    "2 warm-up + 20 timed decodes each, median"}
        if not identical:
            raise SystemExit("bench.py: CPU and GPU binary32 decodes of configuration 1 differ")
    return out


def oracle_slot_check(tbs):
    """Number of the slot's code blocks whose batched-form output words equal
    the oracle's binary16 decode (10 iterations, alpha 0.8125, clamp 32) of
    the same binary16 LLRs."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ncores = min(16, os.cpu_count() or 1)
    same = 0
    for tb in tbs:
        cfg = O.dec_cfg(tb.bg, tb.Z, tb.mb, n_iter=ITERS, alpha=0.8125, clamp=32.0, mode=O.FP16)
        ref = O.decode_cbs(cfg, tb.llr.float().cpu().numpy()[:, :tb.num_llrs], nthreads=ncores)
        got = tb.out.cpu().numpy().view(np.uint32)
        same += int((got == ref).all(axis=1).sum())
    return same


def lds_fraction(edges, ms):
    """Algorithmic LDS bandwidth of a binary16 decode (APP read + write per
    edge update) against the nominal LDS peak."""
    tbs = edges * LDS_BYTES_PER_EDGE / (ms * 1e-3) / 1e12
    return {"achieved": round(tbs, 3), "peak": round(LDS_PEAK_TBS, 2), "unit": "TB/s",
            "frac": round(tbs / LDS_PEAK_TBS, 4)}


def measured_traffic():
    """HBM bytes per launch of the batched decoder from the committed
    counter summary (FETCH_SIZE and WRITE_SIZE passes, copy-calibrated), or
    None when it is absent."""
    import re
    path = os.path.join(ROOT, "profiles", "ldpc_decoder_hetero_slot_hbm_traffic.md")
    try:
        m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", open(path).read(), re.M)
    except OSError:
        return None
    return int(m.group(1)) if m else None


def copy_probe():
    """Three device-to-device copies of 1 GiB each: a kernel of known HBM
    traffic (1 GiB read + 1 GiB written per dispatch) for calibrating the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 30
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"copy_probe_bytes_per_dispatch": n, "dispatches": 3}


def main():
    a = parse()
    sections = set(a.sections.split(","))
    if sections == {"copyprobe"}:
        return copy_probe()
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

    # decoded bits against the transmitted ones (channel errors left after 10 iterations)
    bit_errs, blk_errs = 0, 0
    for tb in tbs:
        bits = tb.out.view(torch.int32)
        got = ((bits.unsqueeze(-1) >> torch.arange(32, device="cuda", dtype=torch.int32)) & 1).reshape(tb.ncb, -1)
        got = got[:, :tb.info_bits_per_cb].to(torch.uint8)
        diff = (got != tb.info)
        bit_errs += int(diff.sum().item())
        blk_errs += int(diff.any(dim=1).sum().item())
    plain_words = [tb.out.clone() for tb in tbs]

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
    # the same slot as one launch of the full-capacity variant only
    one = BatchDecoder(dec, split_by_mb=False).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True,
                                                     max_iterations=ITERS, alpha=0.8125, clamp_value=32.0)
    one.capture()
    for _ in range(a.warmup):
        one.replay()
    torch.cuda.synchronize()
    oev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    for s_ in range(a.steps):
        oev[s_][0].record(gstream)
        one.replay()
        oev[s_][1].record(gstream)
    torch.cuda.synchronize()
    single_ms = sum(e0.elapsed_time(e1) for e0, e1 in oev) / a.steps
    del one
    batched_gbps = info_bits / (batched_ms * 1e-3) / 1e9
    # identical hard decisions: every output word of the batched form equals the plain form's,
    # and both equal the oracle's binary16 decode of the same LLRs (outside any timed region)
    batched_equals_plain = all(torch.equal(tb.out, w) for tb, w in zip(tbs, plain_words))
    oracle_cbs = oracle_slot_check(tbs) if not a.no_cpu else None

    # ---- roofline of the decoder kernel (VALU-bound) ----------------------
    # dominant kernel = the batched persistent decoder: one launch per step
    # (row-pair workgroups of 768 threads in the latency regime)
    kernel_name = ("nrldpc::ldpc_layered_persistent_pair<_Float16, _Float16, 46>"
                   if any(sb.block == 768 for sb in bdec.subs) else "nrldpc::ldpc_layered_persistent<_Float16, _Float16>")
    achieved_tops = edge_updates * OPS_PER_EDGE / (batched_ms * 1e-3) / 1e12
    plain_tops = edge_updates * OPS_PER_EDGE / (kernel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved_tops, 3), "peak": round(VALU_PEAK_TOPS, 2),
                "unit": "TOP/s", "frac": round(achieved_tops / VALU_PEAK_TOPS, 4), "traffic": measured_traffic(),
                "kernel": kernel_name,
                "work": "%d edge updates x %d VALU ops per launch" % (edge_updates, OPS_PER_EDGE),
                "avg_launch_us": round(batched_ms * 1e3, 2),
                "lds": lds_fraction(edge_updates, batched_ms),
                "plain_kernel": {"achieved": round(plain_tops, 3), "avg_launch_us": round(kernel_ms * 1e3 / len(descs), 2)}}

    progress("main slot done: batched %.3f ms, plain %.3f ms" % (batched_ms, plain_ms))
    config1, c1 = None, None
    if "config1" in sections:
        config1, c1 = config1_section(dec, stream)
    config3 = None
    if "config3" in sections:
        progress("config1 done")
        config3 = config3_section(dec, stream, a.steps, with_oracle=not a.no_cpu)
    throughput = None
    if "throughput" in sections:
        progress("config3 done")
        throughput = throughput_section(dec, stream, a.steps)
    fused = None
    if "fused" in sections:
        progress("throughput done")
        fused = fused_slot_section(dec, a.steps)
    serving = None
    if "serving" in sections:
        progress("fused slot done")
        c3g = config3["batched_gbps"] if config3 else None
        c3c = config3["code_blocks"] / (config3["batched_ms"] * 1e-3) if config3 else None
        serving = serving_section(dec, a.serve_ms, c3g, loads=[float(x) for x in a.serve_loads.split(",")],
                                  resident=True, config3_cbs_per_s=c3c)
        batched = serving_section(dec, a.serve_ms, c3g, inflight=a.serve_inflight, loads=(0.6,), resident=False,
                                  templates=serving.pop("templates"), config3_cbs_per_s=c3c)
        batched.pop("templates")
        serving["batched_launches_for_comparison"] = batched

    if not a.no_cpu:
        progress("GPU sections done; CPU baseline")
    cpu = None if a.no_cpu else cpu_baseline(tbs, c1, config1, a.cpu_seconds)
    return line_of(a, locals())


def config1_section(dec, stream):
    """Configuration 1 on the GPU: BG1 Z=384 mb=46, 8 CBs, binary32, plain."""
    import torch
    from aerial_amd import _lib as L
    from aerial_amd import workloads as W
    from aerial_amd.decoder import LdpcDecodeConfig
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
    return {"us_per_tb": round(c1_us, 2), "gbps": round(c1_gbps, 4)}, c1


def config3_section(dec, stream, steps, with_oracle=True):
    """20-cell full-load slot (2800 CBs, BG1 Z=384, R=0.93): both forms."""
    import torch
    from aerial_amd import _lib as L
    from aerial_amd import workloads as W
    from aerial_amd.batch import BatchDecoder
    from aerial_amd.decoder import LdpcDecodeConfig
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

    c3_plain_ms = timed(lambda: [dec.launch(d, stream) for d in c3_descs], max(2, steps // 4))
    plain_words = [tb.out.clone() for tb in c3]
    for tb in c3:
        tb.out.fill_(-1)
    c3_batch_ms = timed(c3b.replay, max(2, steps // 2))
    # identical hard decisions in both forms; the first two code blocks of every cell
    # against the oracle's binary16 decode
    same = all(torch.equal(tb.out, w) for tb, w in zip(c3, plain_words))
    oracle_same = None
    if with_oracle:
        import numpy as np
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        ocfg = O.dec_cfg(1, 384, c3[0].mb, n_iter=ITERS, alpha=0.8125, clamp=32.0, mode=O.FP16)
        oracle_same = 0
        for tb in c3:
            ref = O.decode_cbs(ocfg, tb.llr[:2].float().cpu().numpy()[:, :tb.num_llrs], nthreads=2)
            oracle_same += int((tb.out[:2].cpu().numpy().view(np.uint32) == ref).all(axis=1).sum())
    c3_edges = sum(tb.edge_updates(ITERS) for tb in c3)
    config3 = {"code_blocks": sum(tb.ncb for tb in c3), "info_bits": c3_bits,
               "plain_ms": round(c3_plain_ms, 4), "plain_gbps": round(c3_bits / (c3_plain_ms * 1e-3) / 1e9, 3),
               "batched_ms": round(c3_batch_ms, 4), "batched_gbps": round(c3_bits / (c3_batch_ms * 1e-3) / 1e9, 3),
               "batched_valu_frac": round(c3_edges * OPS_PER_EDGE / (c3_batch_ms * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
               "batched_lds_frac": lds_fraction(c3_edges, c3_batch_ms)["frac"],
               "batched_over_plain": round(c3_plain_ms / c3_batch_ms, 3),
               "slot_budget_ms": 0.5, "batched_equals_plain": same,
               "oracle_equal_code_blocks": oracle_same, "oracle_checked_code_blocks": 2 * len(c3) if with_oracle else 0}
    return config3


def throughput_section(dec, stream, steps, n_slots=16):
    """Batched throughput with many slots in flight: n_slots heterogeneous
    slots decoded as one batch, (a) Zc over 64..384 as in the heterogeneous
    slot, (b) Zc over all 51 lifting sizes 2..384.  Each batch is timed with
    one code block per 384-thread workgroup and with the tuned packing
    (several code blocks per workgroup, workgroup size by Zc); the decoded
    bits of the two must be identical."""
    import torch
    from aerial_amd import _lib as L
    from aerial_amd import workloads as W
    from aerial_amd.batch import BatchDecoder
    out = {}
    for name, z_min in (("zc_64_384", 64), ("zc_2_384", 2)):
        tbs = []
        for k in range(n_slots):
            tbs += W.heterogeneous_slot(seed=100 + k, z_min=z_min)
        bits = sum(tb.ncb * tb.info_bits_per_cb for tb in tbs)
        res = {}
        outs = {}
        for pack in (False, True):
            for tb in tbs:
                tb.out.zero_()
            b = BatchDecoder(dec, pack=pack).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True,
                                                   max_iterations=ITERS, alpha=0.8125, clamp_value=32.0)
            b.capture()
            for _ in range(2):
                b.replay()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(max(2, steps // 2)):
                b.replay()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / max(2, steps // 2)
            res["packed" if pack else "one_per_workgroup"] = {
                "ms": round(ms, 4), "gbps": round(bits / (ms * 1e-3) / 1e9, 3), "launches": len(b.subs)}
            outs[pack] = [tb.out.clone() for tb in tbs]
            del b
        res["packed_over_one_per_workgroup"] = round(res["one_per_workgroup"]["ms"] / res["packed"]["ms"], 3)
        res["bit_identical"] = all(torch.equal(x, y) for x, y in zip(outs[False], outs[True]))
        res["code_blocks"] = sum(tb.ncb for tb in tbs)
        res["info_bits"] = bits
        out[name] = res
        del tbs, outs
    out["workload"] = ("%d heterogeneous slots (16 UEs each, BG1/BG2, R 0.2..0.93, 1-24 CBs/TB, fp16 LLRs, "
                       "10 iterations) in one batch; packing from aerial_amd/tuning_gfx950.json" % n_slots)
    return out


def line_of(a, v):
    """The one JSON line of the run (v: main()'s locals)."""
    info_bits, ncb, tbs = v["info_bits"], v["ncb"], v["tbs"]
    batched_gbps, batched_ms, plain_gbps, plain_ms = v["batched_gbps"], v["batched_ms"], v["plain_gbps"], v["plain_ms"]
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
                  "batched_over_plain": round(batched_gbps / plain_gbps, 3),
                  "batched_single_launch_ms": round(v["single_ms"], 4)},
        "roofline": v["roofline"],
        "cpu_baseline": v["cpu"],
        "check": {"batched_equals_plain": v["batched_equals_plain"],
                  "oracle_equal_code_blocks": v["oracle_cbs"], "code_blocks": ncb,
                  "bit_errors_vs_transmitted": v["bit_errs"], "block_errors_vs_transmitted": v["blk_errs"]},
        "config1_plain_fp32": v["config1"],
        "config3_full_load_20_cells": v["config3"],
        "fused_slot_64_cells_int8": v["fused"],
        "batched_throughput_16_slots": v["throughput"],
        "serving": v["serving"],
    }
    return line


if __name__ == "__main__":
    print(json.dumps(main()))
