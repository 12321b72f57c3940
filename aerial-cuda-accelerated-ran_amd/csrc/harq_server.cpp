// HARQ serving engine: continuous batching of PUSCH transmissions into fused
// slot launches (cuphyHipPuschSlotDecode) with device-resident HARQ soft
// buffers under a byte budget.
//
// The reference has no such layer (cuPHY is driven slot by slot by
// cuPHY-CP); its per-slot pipeline (pusch_rx.cpp) is what one launch here
// corresponds to.  Design:
//   * one device pool of harq_pool_bytes (harq_pool.hpp: size classes over a
//     coalescing first-fit arena, HARQ processes by (cell, RNTI, process id)
//     in a bucket table) hands each HARQ process a buffer of (first CB + C)
//     * Ncb_padded binary16 values when its first transmission (ndi = 1) is
//     admitted; the buffer is freed when a transmission passes the TB CRC,
//     carries CUPHY_HIP_HARQ_LAST_TX, on Release, or after harq_timeout_us
//     unused.  A new TB whose buffer does not fit waits (admission by HBM
//     budget); retransmissions always have theirs.
//   * max_inflight launch slots, each with its own stream, event, pinned
//     staging area and device descriptor area.  Poll() retires finished
//     slots, then fills every free slot with the admitted requests in
//     arrival order (a HARQ process's next transmission waits for its
//     previous one): requests that arrive while launches run are batched
//     into the next free slot, so launches overlap on the device.
//   * per launch: one host->device copy of the PerTbParams, job and item
//     arrays, then one or two persistent fused launches (the register-lean
//     variant for mb <= 8 when at least one CU's worth of such CBs is
//     present; the two on two streams, concurrently), an event per stream.
//   * resident mode (CUPHY_HIP_HARQ_SERVER_RESIDENT): instead of batched
//     launches, one resident launch (pusch_serve_resident) takes code
//     blocks from a ring in host-mapped memory as admitted TBs are
//     published; TBs occupy descriptor slots until their CRC appears.  The
//     launch leaves after 20 ms without work (its workgroups hand their
//     tickets to the next launch) and the engine starts it again when work
//     is open; no item is published into a ring slot whose item an open TB
//     may not have had read yet.
//   * per-TB completion: each TB's CRC is written by the kernel straight
//     into the slot's host-mapped pinned array (CUPHY_HIP_SLOT_TB_NOTIFY:
//     after all of the TB's outputs are in device memory) as soon as the
//     TB's own code blocks are decoded (pre-set to a value no CRC takes);
//     Poll() completes a TB when its value appears, so a TB does not wait
//     for the longest code block of its launch.  The slot, and the HARQ
//     buffers its TBs used, are released once its launches have ended.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <vector>

#include "trace.hpp"
#include "api_objects.hpp"
#include "cuphy_hip.h"
#include "harq_pool.hpp"
#include "nr_bg_tables.h"
#include "pusch_fused.hpp"

namespace {

using clk = std::chrono::steady_clock;

constexpr uint32_t CB_REJECTED = 0xFFFFFFFFu;
constexpr uint32_t CRC_PENDING = 0xFFFFFFFEu;  // TB CRC slot before the kernel writes it (CRCs are < 2^24)
constexpr int SMALL_MB = 8;

using nrharq::round_up;

nrharq::key key_of(const cuphyHipHarqRequest_t& r) { return nrharq::key{r.cell_id, r.rnti, r.harq_pid}; }

struct queued {
    cuphyHipHarqRequest_t req;
    double t_submit;
};

struct launch_slot {
    hipStream_t stream[2] = {nullptr, nullptr};  // high-rate class, full class
    hipEvent_t done[2] = {nullptr, nullptr};
    hipEvent_t staged = nullptr;       // descriptors copied (second stream waits on it)
    bool busy = false;
    unsigned char* h_stage = nullptr;  // pinned: params | jobs | items
    unsigned char* d_stage = nullptr;
    uint32_t* h_crc = nullptr;         // pinned, device-mapped: TB CRCs written by the kernel
    uint32_t* dv_crc = nullptr;        // device view of h_crc
    uint32_t* d_ws[2] = {nullptr, nullptr};
    std::vector<queued> reqs;
    std::vector<uint8_t> taken;        // TB already completed
    int n_open = 0;                    // TBs not yet completed
    double t_launch = 0.0;
};

int num_edges(int bg, int mb)
{
    return bg == 1 ? nr_bg1_row_start[mb] : nr_bg2_row_start[mb];
}

int nsys_of(int bg) { return bg == 1 ? 22 : 10; }

bool is_lifting(int Z)
{
    for (int a : {2, 3, 5, 7, 9, 11, 13, 15})
        for (int v = a; v <= 384; v <<= 1)
            if (v == Z) return true;
    return false;
}

}  // namespace

struct cuphyHipHarqServer {
    int device = 0;
    int cu_count = 1;
    cuphyHipHarqServerConfig_t cfg{};
    unsigned char* d_pool = nullptr;
    nrharq::pool pool;                 // size classes over the device pool's byte range
    nrharq::process_table procs;       // HARQ processes by (cell, RNTI, process id)
    std::deque<queued> pending;
    std::vector<launch_slot> slots;
    std::deque<cuphyHipHarqCompletion_t> ready;
    clk::time_point t0;
    size_t ws_bytes = 0, stage_bytes = 0;
    uint64_t submitted = 0, completed = 0, launches = 0, admission_waits = 0, evictions = 0;
    double t_last_evict = 0.0;

    // resident mode
    bool resident = false;
    nrldpc::serve_ring_hdr* h_ring = nullptr;  // pinned, device-mapped
    nrldpc::serve_ring_hdr* dv_ring = nullptr;
    uint32_t* h_items = nullptr;
    uint32_t* dv_items = nullptr;
    uint32_t ring_size = 0, published = 0, items_open = 0;
    cuphyHipPuschSlotTb_t* h_jobs = nullptr;
    cuphyHipPuschSlotTb_t* dv_jobs = nullptr;
    PerTbParams* h_params = nullptr;
    PerTbParams* dv_params = nullptr;
    uint32_t* h_tbcrc = nullptr;
    uint32_t* dv_tbcrc = nullptr;
    nrldpc::serve_state* d_state = nullptr;
    uint32_t* d_tbw = nullptr;
    hipStream_t r_stream = nullptr;
    hipEvent_t r_done = nullptr;
    bool r_running = false;
    uint64_t idle_ticks = 0, life_ticks = 0;
    uint32_t* d_abandoned = nullptr;  // one ticket word per workgroup of the resident grid
    int r_grid = 0;
    std::vector<int> free_tb_slots;
    std::vector<queued> slot_req;
    std::vector<uint8_t> slot_open;
    std::vector<double> slot_t_publish;
    std::vector<uint32_t> slot_first_item;  // ring index of the slot's first item
    int n_open_slots = 0;

    double now_us() const { return std::chrono::duration<double, std::micro>(clk::now() - t0).count(); }
    size_t params_off() const { return 0; }
    size_t jobs_off() const { return round_up((size_t)cfg.max_tbs_per_launch * sizeof(PerTbParams), 256); }
    size_t items_off() const
    {
        return jobs_off() + round_up((size_t)cfg.max_tbs_per_launch * sizeof(cuphyHipPuschSlotTb_t), 256);
    }

    static size_t harq_bytes(const PerTbParams& p)
    {
        return (size_t)(p.firstCodeBlockIndex + p.num_CBs) * p.Ncb_padded * 2u;
    }
    void complete(const queued& q, uint32_t crc, int32_t status, double t_launch, double t_done)
    {
        cuphyHipHarqCompletion_t c;
        c.tag = q.req.tag;
        c.cell_id = q.req.cell_id;
        c.rnti = q.req.rnti;
        c.harq_pid = q.req.harq_pid;
        c.reserved0 = 0;
        c.tb_crc = crc;
        c.status = status;
        c.t_submit_us = q.t_submit;
        c.t_launch_us = t_launch;
        c.t_done_us = t_done;
        ready.push_back(c);
        ++completed;
    }
    // Free every idle HARQ buffer unused for harq_timeout_us (checked at
    // most once per millisecond).
    void evict_stale()
    {
        if (cfg.harq_timeout_us <= 0.0) return;
        const double t = now_us();
        if (t - t_last_evict < 1000.0) return;
        t_last_evict = t;
        evictions += nrharq::evict_idle(procs, pool, t, cfg.harq_timeout_us);
    }
    void free_proc(const nrharq::key& k)
    {
        nrharq::process* pr = procs.find(k);
        if (!pr) return;
        pool.release(pr->off, pr->bytes);
        procs.erase(k);
    }
    // The process of a request, with a buffer for a new TB (ndi = 1): an
    // existing one if it is large enough, else a new one from the pool.
    // Returns null (and sets *wait) when the pool has no room now.
    nrharq::process* buffer_for(const queued& q, bool& wait)
    {
        wait = false;
        const nrharq::key k = key_of(q.req);
        nrharq::process* pr = procs.find(k);
        if (!q.req.prm.ndi) return pr;
        const size_t need = harq_bytes(q.req.prm);
        if (pr && pr->bytes < need) {
            free_proc(k);
            pr = nullptr;
        }
        if (pr) return pr;
        size_t off = 0, got = 0;
        if (!pool.alloc(need, off, got)) {
            wait = true;
            return nullptr;
        }
        nrharq::process np;
        np.off = off;
        np.bytes = got;
        np.need = need;
        return procs.insert(k, np);
    }
    cuphyStatus_t retire();
    cuphyStatus_t admit_and_launch();
    cuphyStatus_t launch(launch_slot& s);
    cuphyStatus_t retire_resident();
    cuphyStatus_t admit_resident();
    cuphyStatus_t start_resident();
    bool idle() const;
    void destroy();
};

bool cuphyHipHarqServer::idle() const
{
    if (resident) return n_open_slots == 0;
    for (auto& sl : slots)
        if (sl.busy) return false;
    return true;
}

void cuphyHipHarqServer::destroy()
{
    // teardown: errors are not actionable here, the handles are dropped either way
    for (auto& s : slots) {
        for (int k = 0; k < 2; ++k) {
            if (s.stream[k]) {
                (void)hipStreamSynchronize(s.stream[k]);
                (void)hipStreamDestroy(s.stream[k]);
            }
            if (s.done[k]) (void)hipEventDestroy(s.done[k]);
        }
        if (s.staged) (void)hipEventDestroy(s.staged);
        if (s.h_stage) (void)hipHostFree(s.h_stage);
        if (s.h_crc) (void)hipHostFree(s.h_crc);
        if (s.d_stage) (void)hipFree(s.d_stage);
        for (auto w : s.d_ws)
            if (w) (void)hipFree(w);
    }
    slots.clear();
    if (r_stream) {
        // resident launch: leave once the ring is drained (bounded by its
        // idle and lifetime limits in any case)
        if (h_ring) __atomic_store_n(&h_ring->stop, 1u, __ATOMIC_RELEASE);
        (void)hipStreamSynchronize(r_stream);
        (void)hipStreamDestroy(r_stream);
    }
    if (r_done) (void)hipEventDestroy(r_done);
    for (void* h : {(void*)h_ring, (void*)h_items, (void*)h_jobs, (void*)h_params, (void*)h_tbcrc})
        if (h) (void)hipHostFree(h);
    if (d_state) (void)hipFree(d_state);
    if (d_abandoned) (void)hipFree(d_abandoned);
    if (d_tbw) (void)hipFree(d_tbw);
    if (d_pool) (void)hipFree(d_pool);
    d_pool = nullptr;
}

// Resident mode: start the resident launch (none running).  The device
// state (tickets, tail seen, abandoned tickets) carries over from the
// previous launch.
cuphyStatus_t cuphyHipHarqServer::start_resident()
{
    CUPHY_HIP_RANGE("harq.start_resident");
    // the launch's number marks it when its workgroups leave (collective idle exit)
    const uint32_t epoch = (uint32_t)(launches + 1);
    if (nrldpc::serve_resident_launch(dv_ring, dv_items, ring_size - 1, dv_jobs, (int)slot_open.size(), d_state, d_tbw,
                                      d_abandoned, r_grid, cfg.llr_type, idle_ticks, life_ticks, epoch, r_stream) !=
            CUPHY_STATUS_SUCCESS ||
        hipEventRecord(r_done, r_stream) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    r_running = true;
    ++launches;
    return CUPHY_STATUS_SUCCESS;
}

// Resident mode: complete every TB whose CRC has appeared; start the
// resident launch again if it left (idle) while work is still open.
cuphyStatus_t cuphyHipHarqServer::retire_resident()
{
    if (r_running) {
        const hipError_t q = hipEventQuery(r_done);
        if (q == hipSuccess)
            r_running = false;
        else if (q != hipErrorNotReady)
            return CUPHY_STATUS_INTERNAL_ERROR;
    }
    const double t = now_us();
    bool open_work = false;
    for (int k = 0; k < (int)slot_open.size() && n_open_slots > 0; ++k) {
        if (!slot_open[k]) continue;
        const uint32_t crc = __atomic_load_n(&h_tbcrc[k], __ATOMIC_ACQUIRE);
        if (crc == CRC_PENDING) {
            open_work = true;
            continue;
        }
        const queued& r = slot_req[k];
        if (nrharq::process* pr = procs.find(key_of(r.req))) {
            pr->busy = false;
            pr->t_last = t;
        }
        // the kernel released every output of the TB (HARQ buffer included)
        // before its CRC: the buffer may go to another TB now
        if (crc == 0 || crc == CB_REJECTED || (r.req.flags & CUPHY_HIP_HARQ_LAST_TX)) free_proc(key_of(r.req));
        complete(r, crc, crc == CB_REJECTED ? CUPHY_STATUS_UNSUPPORTED_CONFIG : CUPHY_STATUS_SUCCESS,
                 slot_t_publish[k], t);
        items_open -= r.req.prm.num_CBs;
        slot_open[k] = 0;
        free_tb_slots.push_back(k);
        --n_open_slots;
    }
    if (open_work && !r_running) return start_resident();
    return CUPHY_STATUS_SUCCESS;
}

// Resident mode: admit pending requests in arrival order (same rules as the
// batched mode) into free descriptor slots and publish their code blocks.
cuphyStatus_t cuphyHipHarqServer::admit_resident()
{
    if (pending.empty()) return CUPHY_STATUS_SUCCESS;
    std::deque<queued> rest;
    bool new_blocked = false;
    const uint32_t first = published;
    // ring back-pressure: every item not yet read belongs to an open TB, so
    // no item at or after the oldest open TB's first item may be overwritten
    uint32_t oldest = published;
    for (int k = 0; k < (int)slot_open.size() && n_open_slots > 0; ++k)
        if (slot_open[k] && (int32_t)(slot_first_item[k] - oldest) < 0) oldest = slot_first_item[k];
    while (!pending.empty()) {
        queued q = pending.front();
        pending.pop_front();
        const PerTbParams& p = q.req.prm;
        const uint32_t C = p.num_CBs;
        // groups of consecutive CBs decoded side by side in one workgroup
        const uint32_t P = (uint32_t)nrldpc::fused_group_size(
            p, q.req.num_parity_nodes, (q.req.flags & CUPHY_HIP_SLOT_TB_DESCRAMBLE) != 0, nrldpc::SERVE_LDS_VALUES);
        const uint32_t n_items = (C + P - 1) / P;
        if (free_tb_slots.empty() || items_open + C > ring_size || items_open + C > (uint32_t)cfg.max_cbs_per_launch ||
            published + n_items - oldest > ring_size) {
            rest.push_back(q);
            continue;
        }
        {
            const nrharq::process* cur = procs.find(key_of(q.req));
            if (cur && cur->busy) {  // previous transmission still running
                rest.push_back(q);
                continue;
            }
            if (p.ndi && new_blocked && !(cur && cur->bytes >= harq_bytes(p))) {  // behind a waiting new TB
                rest.push_back(q);
                continue;
            }
        }
        bool wait = false;
        nrharq::process* pr = buffer_for(q, wait);
        if (wait) {
            ++admission_waits;
            new_blocked = true;
            rest.push_back(q);
            continue;
        }
        if (!pr) {  // retransmission without a HARQ buffer
            complete(q, CB_REJECTED, CUPHY_STATUS_INVALID_ARGUMENT, now_us(), now_us());
            continue;
        }
        pr->busy = true;
        pr->t_last = now_us();
        const int k = free_tb_slots.back();
        free_tb_slots.pop_back();
        const cuphyHipHarqRequest_t& r = q.req;
        h_params[k] = r.prm;
        cuphyHipPuschSlotTb_t& j = h_jobs[k];
        j.prm = dv_params + k;
        j.llr = r.llr;
        j.harq = d_pool + pr->off;
        j.cb_bits = r.cb_bits;
        j.cb_crc = r.cb_crc;
        j.tb_bytes = r.tb_bytes;
        j.tb_crc = dv_tbcrc + k;
        j.cb_iterations = r.cb_iterations;
        j.num_parity_nodes = r.num_parity_nodes;
        j.max_iterations = r.max_iterations;
        j.flags = (uint16_t)((r.flags & (CUPHY_LDPC_DECODE_EARLY_TERM | CUPHY_HIP_SLOT_TB_DESCRAMBLE)) |
                             CUPHY_HIP_SLOT_TB_NOTIFY);
        j.reserved = 0;
        j.llr_scale = r.llr_scale;
        j.norm = r.norm;
        j.clamp_value = r.clamp_value;
        j.share_offset = (uint32_t)k;  // accumulator pair of slot k
        h_tbcrc[k] = CRC_PENDING;
        slot_first_item[k] = published;
        if (n_open_slots == 0) oldest = published;
        uint32_t n = 0;
        for (uint32_t c = 0; c < C; c += P)
            h_items[(published + n++) & (ring_size - 1)] = nrldpc::make_item(k, (int)c, (int)std::min(P, C - c));
        published += n;
        items_open += C;
        slot_req[k] = q;
        slot_open[k] = 1;
        slot_t_publish[k] = -1.0;
        ++n_open_slots;
    }
    pending.swap(rest);
    if (published == first) return CUPHY_STATUS_SUCCESS;
    const double t = now_us();
    for (int k = 0; k < (int)slot_open.size(); ++k)
        if (slot_open[k] && slot_t_publish[k] < 0.0) slot_t_publish[k] = t;
    // descriptors and items before the tail that publishes them
    __atomic_store_n(&h_ring->tail, published, __ATOMIC_RELEASE);
    if (!r_running) return start_resident();
    return CUPHY_STATUS_SUCCESS;
}

// Complete every TB whose CRC the kernel has written (HARQ buffers of
// decoded or abandoned TBs back to the pool), and free every slot whose
// launches have ended.
cuphyStatus_t cuphyHipHarqServer::retire()
{
    for (auto& s : slots) {
        if (!s.busy) continue;
        bool ended = true;
        for (int k = 0; k < 2; ++k) {
            const hipError_t q = hipEventQuery(s.done[k]);
            if (q == hipErrorNotReady) {
                ended = false;
                continue;
            }
            if (q != hipSuccess) return CUPHY_STATUS_INTERNAL_ERROR;
        }
        const double t = now_us();
        for (size_t i = 0; i < s.reqs.size() && s.n_open > 0; ++i) {
            if (s.taken[i]) continue;
            const uint32_t crc = __atomic_load_n(&s.h_crc[i], __ATOMIC_ACQUIRE);
            if (crc == CRC_PENDING && !ended) continue;
            int32_t status = CUPHY_STATUS_SUCCESS;
            if (crc == CRC_PENDING)
                status = CUPHY_STATUS_INTERNAL_ERROR;  // launches ended without this TB's result
            else if (crc == CB_REJECTED)
                status = CUPHY_STATUS_UNSUPPORTED_CONFIG;
            complete(s.reqs[i], crc, status, s.t_launch, t);
            s.taken[i] = 1;
            --s.n_open;
        }
        if (!ended) continue;
        // HARQ buffers: another launch may use them only once this one has
        // ended (its writes are then out of every L2)
        for (size_t i = 0; i < s.reqs.size(); ++i) {
            const queued& r = s.reqs[i];
            const uint32_t crc = s.h_crc[i];
            if (nrharq::process* pr = procs.find(key_of(r.req))) {
                pr->busy = false;
                pr->t_last = now_us();
            }
            if (crc == 0 || crc == CB_REJECTED || crc == CRC_PENDING || (r.req.flags & CUPHY_HIP_HARQ_LAST_TX))
                free_proc(key_of(r.req));
        }
        s.reqs.clear();
        s.taken.clear();
        s.n_open = 0;
        s.busy = false;
    }
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t cuphyHipHarqServer::admit_and_launch()
{
    for (auto& s : slots) {
        if (s.busy) continue;
        if (pending.empty()) break;
        std::deque<queued> rest;
        bool new_blocked = false;
        int cbs = 0;
        while (!pending.empty()) {
            queued q = pending.front();
            pending.pop_front();
            const PerTbParams& p = q.req.prm;
            const int C = (int)p.num_CBs;
            if ((int)s.reqs.size() >= cfg.max_tbs_per_launch || cbs + C > cfg.max_cbs_per_launch) {
                rest.push_back(q);
                continue;
            }
            {
                const nrharq::process* cur = procs.find(key_of(q.req));
                if (cur && cur->busy) {  // previous transmission still running
                    rest.push_back(q);
                    continue;
                }
                if (p.ndi && new_blocked && !(cur && cur->bytes >= harq_bytes(p))) {  // later new TBs wait
                    rest.push_back(q);
                    continue;
                }
            }
            bool wait = false;
            nrharq::process* pr = buffer_for(q, wait);
            if (wait) {
                ++admission_waits;
                new_blocked = true;  // later new TBs wait behind this one
                rest.push_back(q);
                continue;
            }
            if (!pr) {  // retransmission without a HARQ buffer
                complete(q, CB_REJECTED, CUPHY_STATUS_INVALID_ARGUMENT, now_us(), now_us());
                continue;
            }
            pr->busy = true;
            pr->t_last = now_us();
            s.reqs.push_back(q);
            cbs += C;
        }
        pending.swap(rest);
        if (s.reqs.empty()) break;
        const cuphyStatus_t st = launch(s);
        if (st != CUPHY_STATUS_SUCCESS) return st;
    }
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t cuphyHipHarqServer::launch(launch_slot& s)
{
    CUPHY_HIP_RANGE("harq.launch");
    const int n = (int)s.reqs.size();
    PerTbParams* h_params = reinterpret_cast<PerTbParams*>(s.h_stage + params_off());
    cuphyHipPuschSlotTb_t* h_jobs = reinterpret_cast<cuphyHipPuschSlotTb_t*>(s.h_stage + jobs_off());
    uint32_t* h_items = reinterpret_cast<uint32_t*>(s.h_stage + items_off());
    const PerTbParams* d_params = reinterpret_cast<const PerTbParams*>(s.d_stage + params_off());
    struct cost_item {
        uint64_t cost;
        uint32_t id;
    };
    std::vector<cost_item> small, big;
    int lds_small = 1, lds_big = 1, mb_small = 4, mb_big = 4, g_small = 1, g_big = 1;
    uint32_t share_offset = 0;
    for (int t = 0; t < n; ++t) {
        const cuphyHipHarqRequest_t& r = s.reqs[t].req;
        h_params[t] = r.prm;
        cuphyHipPuschSlotTb_t& j = h_jobs[t];
        j.prm = d_params + t;
        j.llr = r.llr;
        j.harq = d_pool + procs.find(key_of(r))->off;
        j.cb_bits = r.cb_bits;
        j.cb_crc = r.cb_crc;
        j.tb_bytes = r.tb_bytes;
        s.h_crc[t] = CRC_PENDING;
        j.tb_crc = s.dv_crc + t;
        j.cb_iterations = r.cb_iterations;
        j.num_parity_nodes = r.num_parity_nodes;
        j.max_iterations = r.max_iterations;
        j.flags = (uint16_t)((r.flags & (CUPHY_LDPC_DECODE_EARLY_TERM | CUPHY_HIP_SLOT_TB_DESCRAMBLE)) |
                             CUPHY_HIP_SLOT_TB_NOTIFY);
        j.reserved = 0;
        j.llr_scale = r.llr_scale;
        j.norm = r.norm;
        j.clamp_value = r.clamp_value;
        j.share_offset = share_offset;
        share_offset += r.prm.num_CBs;
        const PerTbParams& p = r.prm;
        const int mb = r.num_parity_nodes;
        const uint64_t cost = (uint64_t)p.Zc * (uint64_t)num_edges((int)p.bg, mb) * (uint64_t)std::max(1, (int)r.max_iterations);
        const bool descr = (r.flags & CUPHY_HIP_SLOT_TB_DESCRAMBLE) != 0;
        const int P = nrldpc::fused_group_size(p, mb, descr, nrldpc::SLOT_MAX_LDS_VALUES);
        const int lds = nrldpc::slot_lds_values(p, mb, descr, std::min<int>(P, (int)p.num_CBs));
        auto& dst = (mb <= SMALL_MB) ? small : big;
        for (int c = 0; c < (int)p.num_CBs; c += P) {
            const int n = std::min<int>(P, (int)p.num_CBs - c);
            dst.push_back({cost * (uint64_t)n, nrldpc::make_item(t, c, n)});
        }
        const int g = std::min<int>(P, (int)p.num_CBs);
        if (mb <= SMALL_MB) {
            lds_small = std::max(lds_small, lds);
            mb_small = std::max(mb_small, mb);
            g_small = std::max(g_small, g);
        } else {
            lds_big = std::max(lds_big, lds);
            mb_big = std::max(mb_big, mb);
            g_big = std::max(g_big, g);
        }
    }
    // high-rate CBs run on the register-lean variant only when they fill the GPU
    if (!small.empty() && !big.empty() && (int)small.size() < cu_count) {
        big.insert(big.end(), small.begin(), small.end());
        small.clear();
        lds_big = std::max(lds_big, lds_small);
        mb_big = std::max(mb_big, mb_small);
        g_big = std::max(g_big, g_small);
    }
    auto by_cost = [](const cost_item& a, const cost_item& b) { return a.cost > b.cost; };
    std::stable_sort(small.begin(), small.end(), by_cost);
    std::stable_sort(big.begin(), big.end(), by_cost);
    size_t k = 0;
    for (auto& c : small) h_items[k++] = c.id;
    for (auto& c : big) h_items[k++] = c.id;
    const size_t bytes = items_off() + k * sizeof(uint32_t);
    if (hipMemcpyAsync(s.d_stage, s.h_stage, bytes, hipMemcpyHostToDevice, s.stream[0]) != hipSuccess ||
        hipEventRecord(s.staged, s.stream[0]) != hipSuccess || hipStreamWaitEvent(s.stream[1], s.staged, 0) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    const cuphyHipPuschSlotTb_t* d_jobs = reinterpret_cast<const cuphyHipPuschSlotTb_t*>(s.d_stage + jobs_off());
    const uint32_t* d_items = reinterpret_cast<const uint32_t*>(s.d_stage + items_off());
    cuphyStatus_t st = CUPHY_STATUS_SUCCESS;
    if (!small.empty())
        st = nrldpc::slot_decode(d_jobs, n, d_items, (int)small.size(), cfg.llr_type, lds_small, mb_small, g_small,
                                 s.d_ws[0], cu_count, s.stream[0]);
    if (st == CUPHY_STATUS_SUCCESS && !big.empty())
        st = nrldpc::slot_decode(d_jobs, n, d_items + small.size(), (int)big.size(), cfg.llr_type, lds_big, mb_big, g_big,
                                 s.d_ws[1], cu_count, s.stream[1]);
    if (st != CUPHY_STATUS_SUCCESS) return st;
    for (int q = 0; q < 2; ++q)
        if (hipEventRecord(s.done[q], s.stream[q]) != hipSuccess) return CUPHY_STATUS_INTERNAL_ERROR;
    s.taken.assign((size_t)n, 0);
    s.n_open = n;
    s.t_launch = now_us();
    s.busy = true;
    ++launches;
    return CUPHY_STATUS_SUCCESS;
}

extern "C" {

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerCreate(cuphyLDPCDecoder_t decoder, const cuphyHipHarqServerConfig_t* cfg,
                                                   cuphyHipHarqServer_t* server)
{
    if (!decoder || !cfg || !server) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (cfg->max_inflight < 1 || cfg->max_inflight > 8 || cfg->max_tbs_per_launch < 1 ||
        cfg->max_tbs_per_launch > 65536 || cfg->max_cbs_per_launch < 1 || cfg->harq_pool_bytes == 0 ||
        (cfg->flags & ~CUPHY_HIP_HARQ_SERVER_RESIDENT) != 0 || !(cfg->harq_timeout_us >= 0.0) ||
        !(cfg->resident_idle_us >= 0.0))
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if ((cfg->flags & CUPHY_HIP_HARQ_SERVER_RESIDENT) && cfg->max_tbs_per_launch > 65535)
        return CUPHY_STATUS_INVALID_ARGUMENT;  // descriptor slot in 16 bits of an item
    if (cfg->llr_type != CUPHY_R_8I && cfg->llr_type != CUPHY_R_16F && cfg->llr_type != CUPHY_R_32F)
        return CUPHY_STATUS_UNSUPPORTED_TYPE;
    cuphyHipHarqServer* s = new (std::nothrow) cuphyHipHarqServer();
    if (!s) return CUPHY_STATUS_ALLOC_FAILED;
    s->device = decoder->device;
    s->cu_count = std::max(1, decoder->cu_count);
    s->cfg = *cfg;
    s->t0 = clk::now();
    size_t ws = 0;
    cuphyHipPuschSlotDecodeGetWorkspaceSize(cfg->max_cbs_per_launch, &ws);
    s->ws_bytes = ws;
    s->stage_bytes = s->items_off() + (size_t)cfg->max_cbs_per_launch * sizeof(uint32_t);
    bool ok = hipMalloc(&s->d_pool, cfg->harq_pool_bytes) == hipSuccess;
    s->pool.init(cfg->harq_pool_bytes);
    s->resident = (cfg->flags & CUPHY_HIP_HARQ_SERVER_RESIDENT) != 0;
    if (s->resident) {
        const int ns = cfg->max_tbs_per_launch;
        uint32_t rs = 4096;
        while (rs < 4u * (uint32_t)cfg->max_cbs_per_launch && rs < (1u << 30)) rs <<= 1;  // slack: no slot is rewritten before it is read
        s->ring_size = rs;
        auto host = [&](void** h, void** dv, size_t bytes) {
            return hipHostMalloc(h, bytes, hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess &&
                   (std::memset(*h, 0, bytes), hipHostGetDevicePointer(dv, *h, 0) == hipSuccess);
        };
        int khz = 0;
        s->r_grid = nrldpc::serve_resident_grid(s->cu_count);
        ok = ok && s->r_grid > 0 && host((void**)&s->h_ring, (void**)&s->dv_ring, sizeof(nrldpc::serve_ring_hdr)) &&
             hipMalloc((void**)&s->d_abandoned, (size_t)s->r_grid * sizeof(uint32_t)) == hipSuccess &&
             hipMemset(s->d_abandoned, 0xFF, (size_t)s->r_grid * sizeof(uint32_t)) == hipSuccess &&
             host((void**)&s->h_items, (void**)&s->dv_items, (size_t)rs * sizeof(uint32_t)) &&
             host((void**)&s->h_jobs, (void**)&s->dv_jobs, (size_t)ns * sizeof(cuphyHipPuschSlotTb_t)) &&
             host((void**)&s->h_params, (void**)&s->dv_params, (size_t)ns * sizeof(PerTbParams)) &&
             host((void**)&s->h_tbcrc, (void**)&s->dv_tbcrc, (size_t)ns * sizeof(uint32_t)) &&
             hipMalloc((void**)&s->d_state, sizeof(nrldpc::serve_state)) == hipSuccess &&
             hipMemset(s->d_state, 0, sizeof(nrldpc::serve_state)) == hipSuccess &&
             hipMalloc((void**)&s->d_tbw, (size_t)ns * 2 * sizeof(uint32_t)) == hipSuccess &&
             hipMemset(s->d_tbw, 0, (size_t)ns * 2 * sizeof(uint32_t)) == hipSuccess &&
             hipStreamCreateWithFlags(&s->r_stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&s->r_done, hipEventDisableTiming) == hipSuccess &&
             hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) == hipSuccess && khz > 0 &&
             hipDeviceSynchronize() == hipSuccess;
        // a resident launch leaves resident_idle_us (default 20 ms) after its
        // last work, and never runs longer than 60 s
        const double idle_us = cfg->resident_idle_us > 0.0 ? cfg->resident_idle_us : 20000.0;
        s->idle_ticks = (uint64_t)((double)khz * idle_us / 1000.0);
        s->life_ticks = (uint64_t)khz * 60000u;
        s->slot_req.resize(ns);
        s->slot_open.assign(ns, 0);
        s->slot_t_publish.assign(ns, -1.0);
        s->slot_first_item.assign(ns, 0u);
        for (int k = ns - 1; k >= 0; --k) s->free_tb_slots.push_back(k);
    }
    s->slots.resize(s->resident ? 0 : cfg->max_inflight);
    for (auto& sl : s->slots) {
        if (!ok) break;
        for (int k = 0; k < 2 && ok; ++k)
            ok = hipStreamCreateWithFlags(&sl.stream[k], hipStreamNonBlocking) == hipSuccess &&
                 hipEventCreateWithFlags(&sl.done[k], hipEventDisableTiming) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&sl.staged, hipEventDisableTiming) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void**>(&sl.h_stage), s->stage_bytes, hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc(reinterpret_cast<void**>(&sl.h_crc), (size_t)cfg->max_tbs_per_launch * 4,
                           hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess &&
             hipHostGetDevicePointer(reinterpret_cast<void**>(&sl.dv_crc), sl.h_crc, 0) == hipSuccess &&
             hipMalloc(reinterpret_cast<void**>(&sl.d_stage), s->stage_bytes) == hipSuccess;
        for (auto& w : sl.d_ws)
            ok = ok && hipMalloc(reinterpret_cast<void**>(&w), ws) == hipSuccess && hipMemset(w, 0, ws) == hipSuccess;
    }
    if (!ok) {
        s->destroy();
        delete s;
        return CUPHY_STATUS_ALLOC_FAILED;
    }
    *server = s;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerDestroy(cuphyHipHarqServer_t server)
{
    if (!server) return CUPHY_STATUS_INVALID_ARGUMENT;
    server->destroy();
    delete server;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerSubmit(cuphyHipHarqServer_t server, const cuphyHipHarqRequest_t* req)
{
    CUPHY_HIP_RANGE("cuphyHipHarqServerSubmit");
    if (!server || !req) return CUPHY_STATUS_INVALID_ARGUMENT;
    const PerTbParams& p = req->prm;
    const int bg = (int)p.bg;
    if ((bg != 1 && bg != 2) || !is_lifting((int)p.Zc) || p.num_CBs < 1 || p.num_CBs > 4096 || p.Qm < 1 ||
        p.Qm > 8 || p.Nl < 1 || p.Ncb < 1 || p.Ncb > (uint32_t)(bg == 1 ? 66 : 50) * p.Zc ||
        p.Ncb_padded < 2 * p.Zc + p.Ncb || p.K != (uint32_t)nsys_of(bg) * p.Zc || p.F >= p.K)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if (!req->llr || !req->cb_bits || !req->cb_crc || !req->tb_bytes) return CUPHY_STATUS_INVALID_ARGUMENT;
    const int mb = req->num_parity_nodes;
    if (mb < 4 || mb > (bg == 1 ? 46 : 42) || req->max_iterations < 0) return CUPHY_STATUS_INVALID_ARGUMENT;
    // descrambled CBs: sequence words within the launch's LDS
    const int lds = nrldpc::slot_lds_values(p, mb, (req->flags & CUPHY_HIP_SLOT_TB_DESCRAMBLE) != 0);
    if (lds > (server->resident ? nrldpc::SERVE_LDS_VALUES : nrldpc::SLOT_MAX_LDS_VALUES))
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if ((int)p.num_CBs > server->cfg.max_cbs_per_launch) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (server->harq_bytes(p) > server->cfg.harq_pool_bytes) return CUPHY_STATUS_INVALID_ARGUMENT;
    queued q;
    q.req = *req;
    q.t_submit = server->now_us();
    server->pending.push_back(q);
    ++server->submitted;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerPoll(cuphyHipHarqServer_t server, cuphyHipHarqCompletion_t* out,
                                                 int max_out, int* n_out)
{
    if (!server || !n_out || max_out < 0 || (max_out > 0 && !out)) return CUPHY_STATUS_INVALID_ARGUMENT;
    cuphyStatus_t st = server->resident ? server->retire_resident() : server->retire();
    server->evict_stale();
    if (st == CUPHY_STATUS_SUCCESS) st = server->resident ? server->admit_resident() : server->admit_and_launch();
    int n = 0;
    while (n < max_out && !server->ready.empty()) {
        out[n++] = server->ready.front();
        server->ready.pop_front();
    }
    *n_out = n;
    return st;
}

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerRelease(cuphyHipHarqServer_t server, uint32_t cell_id, uint16_t rnti,
                                                    uint8_t harq_pid)
{
    if (!server) return CUPHY_STATUS_INVALID_ARGUMENT;
    const nrharq::key k{cell_id, rnti, harq_pid};
    const nrharq::process* pr = server->procs.find(k);
    if (!pr) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (pr->busy) return CUPHY_STATUS_NOT_SUPPORTED;  // a launch is using it
    server->free_proc(k);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerGetStats(cuphyHipHarqServer_t server, cuphyHipHarqServerStats_t* stats)
{
    if (!server || !stats) return CUPHY_STATUS_INVALID_ARGUMENT;
    stats->submitted = server->submitted;
    stats->completed = server->completed;
    stats->launches = server->launches;
    stats->pool_bytes_in_use = server->pool.in_use();
    stats->pool_bytes_peak = server->pool.peak();
    stats->pending = server->pending.size();
    stats->admission_waits = server->admission_waits;
    stats->evictions = server->evictions;
    stats->pool_bytes_cached = server->pool.cached();
    stats->pool_reclaims = server->pool.reclaims();
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServeTrace(cuphyHipHarqServer_t server, const cuphyHipHarqTraceEntry_t* trace,
                                                 int n, double harq_rtt_us, cuphyHipHarqCompletion_t* done,
                                                 int* n_done)
{
    CUPHY_HIP_RANGE("cuphyHipHarqServeTrace");
    if (!server || n < 0 || (n > 0 && (!trace || !done)) || !n_done || harq_rtt_us < 0.0)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    for (int i = 0; i < n; ++i)
        if (trace[i].next >= n || trace[i].next == i) return CUPHY_STATUS_INVALID_ARGUMENT;
    for (int i = 0; i < n; ++i) {
        std::memset(&done[i], 0, sizeof(done[i]));
        done[i].tag = (uint64_t)i;
        done[i].cell_id = trace[i].req.cell_id;
        done[i].rnti = trace[i].req.rnti;
        done[i].harq_pid = trace[i].req.harq_pid;
        done[i].status = CUPHY_STATUS_NOT_SUPPORTED;
    }
    // release schedule: (time, entry)
    std::multimap<double, int> due;
    for (int i = 0; i < n; ++i)
        if (trace[i].arrival_us >= 0.0) due.emplace(trace[i].arrival_us, i);
    const double base = server->now_us();
    int outstanding = 0, got = 0;
    std::vector<cuphyHipHarqCompletion_t> buf(256);
    while (!due.empty() || outstanding > 0) {
        const double now = server->now_us() - base;
        while (!due.empty() && due.begin()->first <= now) {
            const int i = due.begin()->second;
            due.erase(due.begin());
            cuphyHipHarqRequest_t r = trace[i].req;
            r.tag = (uint64_t)i;
            const cuphyStatus_t st = cuphyHipHarqServerSubmit(server, &r);
            if (st != CUPHY_STATUS_SUCCESS) {
                done[i].status = st;
                continue;
            }
            ++outstanding;
        }
        int k = 0;
        const cuphyStatus_t st = cuphyHipHarqServerPoll(server, buf.data(), (int)buf.size(), &k);
        if (st != CUPHY_STATUS_SUCCESS) return st;
        for (int c = 0; c < k; ++c) {
            cuphyHipHarqCompletion_t e = buf[c];
            const int i = (int)e.tag;
            if (i < 0 || i >= n) continue;
            e.t_submit_us -= base;
            e.t_launch_us -= base;
            e.t_done_us -= base;
            done[i] = e;
            --outstanding;
            ++got;
            const int nx = trace[i].next;
            if (e.status == CUPHY_STATUS_SUCCESS && e.tb_crc != 0 && nx >= 0) due.emplace(e.t_done_us + harq_rtt_us, nx);
        }
    }
    // let the last launches end: their HARQ buffers go back to the pool then
    while (!server->idle()) {
        const cuphyStatus_t st = server->resident ? server->retire_resident() : server->retire();
        if (st != CUPHY_STATUS_SUCCESS) return st;
    }
    *n_done = got;
    return CUPHY_STATUS_SUCCESS;
}

}  // extern "C"
