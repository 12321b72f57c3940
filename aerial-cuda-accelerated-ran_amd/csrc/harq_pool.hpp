// HARQ soft-buffer pool of the serving engine: host-side bookkeeping of one
// device allocation (offsets only, no device call), so that it can be unit
// tested on the CPU.
//
// Counterpart of the reference's HARQ buffer management in cuPHY-CP
// (cuphydriver/include/harq_pool.hpp:49-120 _HarqBuffer, :389-470
// HarqBucket, :472-545 HarqPool, :545+ HarqPoolManager;
// src/uplink/harq_pool.cpp:333-525), laid out for one device pool:
//   * size classes: a buffer's bytes are rounded up to a class (steps of
//     2^(1/4), at most 19 % larger); a released buffer goes to its class's
//     free list and the next buffer of that class reuses it without a search
//     (the reference keeps one pool of fixed-size buffers per size,
//     HarqPoolManager::poolAlloc picking the first large enough);
//   * the classes draw their buffers from a first-fit arena with coalescing;
//     when it has no room, the cached free buffers of every class go back to
//     the arena (merging with their neighbours) and the allocation is tried
//     again (the reference frees buffers under pool depletion,
//     HarqPoolManager::checkPoolDepletion);
//   * HARQ processes are found by (cell, RNTI, HARQ process id) in a bucket
//     table indexed by RNTI mod 16 and process id mod 16, each bucket holding
//     every colliding process (HarqBucket's bucket_map, whose entries hold
//     several RNTIs);
//   * an idle process whose buffer has not been used for a timeout is evicted
//     (the reference's cleanup of buffers unused for MAX_HARQB_NS_BUCKET).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <map>
#include <vector>

namespace nrharq {

constexpr size_t POOL_ALIGN = 256;

inline size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// First-fit allocator over [0, size) with a coalescing free list (offset ->
// length), every block a multiple of POOL_ALIGN.
class arena {
public:
    void init(size_t size)
    {
        free_.clear();
        size_ = size / POOL_ALIGN * POOL_ALIGN;
        if (size_) free_[0] = size_;
    }
    bool alloc(size_t bytes, size_t& off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t free_blocks() const { return free_.size(); }
    size_t largest_free() const
    {
        size_t m = 0;
        for (auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }
    // (offset, length) of every free block, in offset order
    const std::map<size_t, size_t>& free_map() const { return free_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0;
};

// Size classes over an arena.
class pool {
public:
    static constexpr int NUM_CLASSES = 96;  // 4 per power of two: 256 B .. 2^32 B

    void init(size_t bytes)
    {
        arena_.init(bytes);
        for (auto& f : cached_) f.clear();
        in_use_ = peak_ = cached_bytes_ = 0;
        reclaims_ = 0;
    }
    // Class of a request and its byte size: the smallest 256 * 2^(k/4),
    // rounded up to POOL_ALIGN, that holds `bytes`.
    static int class_of(size_t bytes)
    {
        for (int k = 0; k < NUM_CLASSES; ++k)
            if (class_bytes(k) >= bytes) return k;
        return -1;
    }
    static size_t class_bytes(int k)
    {
        static const double steps[4] = {1.0, 1.189207115002721, 1.414213562373095, 1.681792830507429};
        const double v = (double)POOL_ALIGN * (double)(1ull << (k / 4)) * steps[k % 4];
        return round_up((size_t)v, POOL_ALIGN);
    }
    // A buffer of at least `bytes`: its offset and its (class-rounded) size.
    bool alloc(size_t bytes, size_t& off, size_t& got)
    {
        const int k = class_of(bytes);
        if (k < 0) return false;
        got = class_bytes(k);
        if (!cached_[k].empty()) {
            off = cached_[k].back();
            cached_[k].pop_back();
            cached_bytes_ -= got;
        } else if (!arena_.alloc(got, off)) {
            if (cached_bytes_ == 0) return false;
            reclaim();
            if (!arena_.alloc(got, off)) return false;
        }
        in_use_ += got;
        peak_ = std::max(peak_, in_use_);
        return true;
    }
    // Return a buffer (offset and size as alloc gave them) to its class.
    void release(size_t off, size_t got)
    {
        const int k = class_of(got);
        in_use_ -= got;
        cached_[k].push_back(off);
        cached_bytes_ += got;
    }
    // Every cached buffer back to the arena.
    void reclaim()
    {
        for (int k = 0; k < NUM_CLASSES; ++k) {
            for (size_t off : cached_[k]) arena_.release(off, class_bytes(k));
            cached_[k].clear();
        }
        cached_bytes_ = 0;
        ++reclaims_;
    }
    size_t in_use() const { return in_use_; }
    size_t peak() const { return peak_; }
    size_t cached() const { return cached_bytes_; }
    uint64_t reclaims() const { return reclaims_; }
    const arena& arena_view() const { return arena_; }

private:
    arena arena_;
    std::vector<size_t> cached_[NUM_CLASSES];
    size_t in_use_ = 0, peak_ = 0, cached_bytes_ = 0;
    uint64_t reclaims_ = 0;
};

// A HARQ process's state: its buffer and its use.
struct process {
    size_t off = 0, bytes = 0;  // buffer in the pool (class-rounded bytes)
    size_t need = 0;            // bytes its TB needs
    bool busy = false;          // a transmission is queued or decoding
    double t_last = 0.0;        // last admission or completion of a transmission using it (us)
};

struct key {
    uint32_t cell = 0;
    uint16_t rnti = 0;
    uint8_t pid = 0;
    bool operator==(const key& o) const { return cell == o.cell && rnti == o.rnti && pid == o.pid; }
};

// HARQ processes by (cell, RNTI, process id): 16 x 16 buckets (RNTI mod 16,
// process id mod 16), each a list of the processes that map to it.
class process_table {
public:
    static constexpr int RNTI_BUCKETS = 16, PID_BUCKETS = 16;

    process* find(const key& k)
    {
        auto& b = bucket(k);
        for (auto& e : b)
            if (e.first == k) return &e.second;
        return nullptr;
    }
    process* insert(const key& k, const process& p)
    {
        auto& b = bucket(k);
        for (auto& e : b)
            if (e.first == k) {
                e.second = p;
                return &e.second;
            }
        b.emplace_back(k, p);
        ++size_;
        return &b.back().second;
    }
    bool erase(const key& k)
    {
        auto& b = bucket(k);
        for (size_t i = 0; i < b.size(); ++i)
            if (b[i].first == k) {
                b[i] = b.back();
                b.pop_back();
                --size_;
                return true;
            }
        return false;
    }
    template <typename F> void for_each(F&& f)
    {
        for (auto& row : buckets_)
            for (auto& b : row)
                for (auto& e : b) f(e.first, e.second);
    }
    size_t size() const { return size_; }
    size_t bucket_size(const key& k) const
    {
        return buckets_[k.rnti % RNTI_BUCKETS][k.pid % PID_BUCKETS].size();
    }

private:
    std::vector<std::pair<key, process>>& bucket(const key& k)
    {
        return buckets_[k.rnti % RNTI_BUCKETS][k.pid % PID_BUCKETS];
    }
    std::vector<std::pair<key, process>> buckets_[RNTI_BUCKETS][PID_BUCKETS];
    size_t size_ = 0;
};

// Free the buffer of every idle process not used for longer than
// timeout_us and forget the process; returns how many.
inline size_t evict_idle(process_table& table, pool& p, double now_us, double timeout_us)
{
    std::vector<key> stale;
    table.for_each([&](const key& k, const process& pr) {
        if (!pr.busy && now_us - pr.t_last > timeout_us) stale.push_back(k);
    });
    for (const key& k : stale) {
        const process* pr = table.find(k);
        p.release(pr->off, pr->bytes);
        table.erase(k);
    }
    return stale.size();
}

}  // namespace nrharq
