"""HARQ serving engine (cuphyHipHarqServer*): transmissions submitted as they
arrive are continuously batched into fused slot launches, with each HARQ
process's binary16 soft buffer kept in a device pool of a fixed byte budget
(new TBs are admitted only when their buffer fits).  Also builds load-test
traces: Poisson TB arrivals from many cells, each TB with up to four
transmissions at rv 0, 2, 3, 1, a retransmission released one HARQ round
trip after its predecessor failed the TB CRC.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .slot import SlotTb
from .workloads import nsys

RV_SEQUENCE = (0, 2, 3, 1)


def norm_f16(decoder, bg, Z, mb):
    """binary16 normalisation of the reference's table (SetNormalization)."""
    cfg = L.LDPCDecodeConfigDesc()
    cfg.llr_type = L.CUPHY_R_16F
    cfg.num_parity_nodes = mb
    cfg.Z = Z
    cfg.max_iterations = 1
    cfg.Kb = nsys(bg)
    cfg.BG = bg
    L.check("cuphyErrorCorrectionLDPCDecodeSetNormalization",
            L.lib().cuphyErrorCorrectionLDPCDecodeSetNormalization(decoder.dec, C.byref(cfg)))
    return cfg.norm.f16x2 & 0xFFFF


def harq_key(process):
    """(cell, rnti, HARQ process id) of a process given as that tuple or as
    an int n (process id n mod 16, RNTI n / 16 mod 65536, cell n / 2^20)."""
    if isinstance(process, tuple):
        cell, rnti, pid = process
        return int(cell), int(rnti), int(pid)
    n = int(process)
    return n >> 20, (n >> 4) & 0xFFFF, n & 0xF


class HarqServer:
    """Owner of a cuphyHipHarqServer_t."""

    def __init__(self, decoder, pool_bytes=2 << 30, max_inflight=3, max_tbs=1024, max_cbs=8192,
                 llr_type=L.CUPHY_R_16F, resident=False, harq_timeout_us=0.0, resident_idle_us=0.0):
        """resident: one resident launch takes code blocks as they are
        admitted (max_tbs / max_cbs then bound the TBs / code blocks in
        flight); otherwise admitted transmissions are batched into up to
        max_inflight launches.  harq_timeout_us > 0: a HARQ buffer not used
        for that long is freed (a later retransmission of it is rejected).
        resident_idle_us > 0: the resident launch leaves that long after its
        last work (default 20 ms)."""
        self.dec = decoder
        self._lib = L.lib()
        cfg = L.HarqServerConfig(harq_pool_bytes=pool_bytes, max_inflight=max_inflight, max_tbs_per_launch=max_tbs,
                                 max_cbs_per_launch=max_cbs, llr_type=llr_type,
                                 flags=L.CUPHY_HIP_HARQ_SERVER_RESIDENT if resident else 0,
                                 harq_timeout_us=harq_timeout_us, resident_idle_us=resident_idle_us)
        self.h = C.c_void_p()
        L.check("cuphyHipHarqServerCreate", self._lib.cuphyHipHarqServerCreate(decoder.dec, C.byref(cfg),
                                                                               C.byref(self.h)))
        self._norm = {}

    def close(self):
        if self.h:
            self._lib.cuphyHipHarqServerDestroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def request(self, process, params, llr, outputs, mb, max_iterations=10, early_term=True, llr_scale=1.0,
                clamp_value=32.0, last=False, tag=0, descramble=False):
        """A cuphyHipHarqRequest_t; process: the HARQ process, (cell, rnti,
        process id) or an int (harq_key); outputs = (cb_bits, cb_crc,
        tb_bytes, tb_crc, cb_iterations) device tensors as
        SlotTb.alloc_outputs makes them (the engine reports the TB CRC in the
        completion; tb_crc and cb_iterations may be None); descramble: with
        params.cinit."""
        r = L.HarqRequest()
        r.cell_id, r.rnti, r.harq_pid = harq_key(process)
        r.tag = tag
        r.prm = params
        r.llr = llr.data_ptr()
        cb_bits, cb_crc, tb_bytes = outputs[:3]
        cb_iters = outputs[4] if len(outputs) > 4 else None
        r.cb_bits = cb_bits.data_ptr()
        r.cb_crc = cb_crc.data_ptr()
        r.tb_bytes = tb_bytes.data_ptr()
        r.cb_iterations = cb_iters.data_ptr() if cb_iters is not None else None
        r.num_parity_nodes = mb
        r.max_iterations = max_iterations
        r.flags = ((L.CUPHY_LDPC_DECODE_EARLY_TERM if early_term else 0) | (L.CUPHY_HIP_HARQ_LAST_TX if last else 0) |
                   (L.CUPHY_HIP_SLOT_TB_DESCRAMBLE if descramble else 0))
        r.llr_scale = llr_scale
        key = (params.bg, params.Zc, mb)
        if key not in self._norm:
            self._norm[key] = norm_f16(self.dec, *key)
        h = self._norm[key]
        r.norm.f16x2 = h | (h << 16)
        r.clamp_value = clamp_value
        return r

    def submit(self, req):
        L.check("cuphyHipHarqServerSubmit", self._lib.cuphyHipHarqServerSubmit(self.h, C.byref(req)))

    def poll(self, max_out=256):
        out = (L.HarqCompletion * max_out)()
        n = C.c_int()
        L.check("cuphyHipHarqServerPoll", self._lib.cuphyHipHarqServerPoll(self.h, out, max_out, C.byref(n)))
        return [out[i] for i in range(n.value)]

    def release(self, process):
        L.check("cuphyHipHarqServerRelease", self._lib.cuphyHipHarqServerRelease(self.h, *harq_key(process)))

    def stats(self):
        s = L.HarqServerStats()
        L.check("cuphyHipHarqServerGetStats", self._lib.cuphyHipHarqServerGetStats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def serve_trace(self, entries, harq_rtt_us):
        """entries: ctypes array of HarqTraceEntry -> completion array."""
        n = len(entries)
        done = (L.HarqCompletion * n)()
        nd = C.c_int()
        L.check("cuphyHipHarqServeTrace",
                self._lib.cuphyHipHarqServeTrace(self.h, entries, n, harq_rtt_us, done, C.byref(nd)))
        return done


# ---------------------------------------------------------------------------
# load-test traces
# ---------------------------------------------------------------------------
LIFTING_64_384 = sorted({a << j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if 64 <= (a << j) <= 384})


def serving_specs(n, seed):
    """TB shapes drawn as in the heterogeneous slot: BG in {1,2}, Zc uniform
    over the lifting sizes 64..384, code rate U[0.2, 0.93], 1-24 CBs, Es/N0
    U[-1, 8] dB; Qm in {2, 4, 6, 8}, one layer, no filler bits."""
    rng = np.random.RandomState(seed)
    specs = []
    for _ in range(n):
        bg = int(rng.randint(1, 3))
        Zc = int(LIFTING_64_384[rng.randint(len(LIFTING_64_384))])
        R = float(rng.uniform(0.2, 0.93))
        C_ = int(rng.randint(1, 25))
        snr = float(rng.uniform(-1.0, 8.0))
        Qm = int((2, 4, 6, 8)[rng.randint(4)])
        Kp = nsys(bg) * Zc
        E = int(np.ceil(Kp / R / Qm)) * Qm
        specs.append((bg, Zc, C_, Kp, Qm, 1, E, snr))
    return specs


class TraceTemplates:
    """TB templates with device code words and the LLRs of all four
    transmissions (binary16, BPSK-AWGN at the TB's Es/N0)."""

    def __init__(self, specs, seed):
        from . import pusch_workloads as PW
        self.tbs = PW.build_tbs(specs, seed)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        for tb in self.tbs:
            for rv in RV_SEQUENCE:
                tb.llrs[rv] = PW.tx_llrs(tb, rv, gen, mode="bpsk", dtype=torch.float16)
        self.mbs = []
        for tb in self.tbs:
            mbs, m = [], 4
            for k, rv in enumerate(RV_SEQUENCE):
                _, mb = tb.params(rv, 1 if k == 0 else 0)
                m = mb if k == 0 else max(m, mb)
                mbs.append(m)
            self.mbs.append(mbs)


def poisson_arrivals(n_cells, rate_per_s, duration_us, n_templates, seed):
    """Merged per-cell Poisson processes: sorted (t_us, cell, template)."""
    rng = np.random.RandomState(seed)
    out = []
    per_cell = rate_per_s / n_cells
    for cell in range(n_cells):
        t = 0.0
        while True:
            t += rng.exponential(1e6 / per_cell)
            if t >= duration_us:
                break
            out.append((t, cell, int(rng.randint(n_templates))))
    out.sort()
    return out


def build_trace(server, templates, arrivals, max_iterations=10, outputs=None):
    """Trace entries for the arrivals (four per TB, chained), with per-TB
    output buffers.  Returns (entries, outputs, meta) where meta[i] =
    (arrival index, transmission index, template index)."""
    n = len(arrivals) * 4
    entries = (L.HarqTraceEntry * n)()
    outs, meta = [], []
    for a, (t, cell, ti) in enumerate(arrivals):
        tb = templates.tbs[ti]
        p0, _ = tb.params(0, 1)
        o = SlotTb.alloc_outputs(p0) if outputs is None else outputs[a]
        outs.append(o)
        for k, rv in enumerate(RV_SEQUENCE):
            i = 4 * a + k
            p, _ = tb.params(rv, 1 if k == 0 else 0)
            e = entries[i]
            # the arrival's HARQ process: its cell, an RNTI and process id unique within the trace
            e.req = server.request((cell, 1 + (a >> 4) % 65519, a & 15), p, tb.llrs[rv], o, templates.mbs[ti][k],
                                   max_iterations=max_iterations, early_term=True, last=(k == 3), tag=i)
            e.arrival_us = t if k == 0 else -1.0
            e.next = i + 1 if k < 3 else -1
            meta.append((a, k, ti))
    return entries, outs, meta


def trace_report(done, meta, templates, arrivals):
    """Latency percentiles over released transmissions, TB outcomes and
    decoded-info-bit goodput (bits of TBs that passed / trace duration)."""
    lat, t_end, t_start = [], 0.0, None
    passed_bits, passed, failed = 0, 0, 0
    ntx = 0
    tb_state = {}
    for i, (a, k, ti) in enumerate(meta):
        c = done[i]
        if c.status != L.CUPHY_STATUS_SUCCESS:
            continue
        ntx += 1
        lat.append(c.t_done_us - c.t_submit_us)
        t_end = max(t_end, c.t_done_us)
        if c.tb_crc == 0:
            tb_state[a] = ("pass", k)
        elif k == 3:
            tb_state[a] = ("fail", k)
    for a, (st, k) in tb_state.items():
        if st == "pass":
            passed += 1
            passed_bits += templates.tbs[arrivals[a][2]].info_bits
        else:
            failed += 1
    t_start = arrivals[0][0] if arrivals else 0.0
    lat = np.array(lat) if lat else np.zeros(1)
    span_s = max(1e-9, (t_end - t_start) * 1e-6)
    return {"transmissions": ntx, "tbs": len(arrivals), "tbs_passed": passed, "tbs_failed_after_4": failed,
            "latency_us": {"p50": round(float(np.percentile(lat, 50)), 1),
                           "p99": round(float(np.percentile(lat, 99)), 1),
                           "p99_9": round(float(np.percentile(lat, 99.9)), 1),
                           "max": round(float(lat.max()), 1)},
            "goodput_gbps": round(passed_bits / span_s / 1e9, 3), "span_ms": round(span_s * 1e3, 2)}
