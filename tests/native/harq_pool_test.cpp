// CPU unit test of the serving engine's HARQ buffer pool
// (aerial-cuda-accelerated-ran_amd/csrc/harq_pool.hpp), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_harq_pool_cpu.py.
// Random allocate / release / evict / reclaim sequences against a plain
// model of the live buffers; exits nonzero on the first broken invariant.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "harq_pool.hpp"

using namespace nrharq;

#define CHECK(c)                                                                       \
    do {                                                                               \
        if (!(c)) {                                                                    \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

struct live_buf {
    size_t off, bytes, need;
};

// Live buffers do not overlap, lie in the pool, are aligned, and add up to in_use.
static void check_live(const pool& p, const std::map<int, live_buf>& live, size_t size)
{
    std::map<size_t, size_t> by_off;
    size_t sum = 0;
    for (auto& kv : live) {
        const live_buf& b = kv.second;
        CHECK(b.off % POOL_ALIGN == 0 && b.bytes % POOL_ALIGN == 0);
        CHECK(b.off + b.bytes <= size);
        CHECK(b.bytes >= b.need);
        CHECK((double)b.bytes <= 1.19 * (double)b.need + POOL_ALIGN);  // class steps of 2^(1/4)
        by_off[b.off] = b.bytes;
        sum += b.bytes;
    }
    size_t end = 0;
    for (auto& kv : by_off) {
        CHECK(kv.first >= end);
        end = kv.first + kv.second;
    }
    CHECK(sum == p.in_use());
    CHECK(p.in_use() + p.cached() <= size);
}

static void test_classes()
{
    for (size_t b = 1; b < (size_t)1 << 28; b = b * 3 / 2 + 1) {
        const int k = pool::class_of(b);
        CHECK(k >= 0 && pool::class_bytes(k) >= b);
        CHECK(k == 0 || pool::class_bytes(k - 1) < b);
        CHECK(pool::class_of(pool::class_bytes(k)) == k);  // release finds the class alloc used
    }
}

static void test_random(unsigned seed, size_t size)
{
    std::mt19937_64 rng(seed);
    pool p;
    p.init(size);
    std::map<int, live_buf> live;
    int next_id = 0;
    for (int step = 0; step < 20000; ++step) {
        const int op = (int)(rng() % 10);
        if (op < 6) {  // a TB's buffer: (first CB + C) * Ncb_padded * 2 bytes
            static const int zs[] = {2, 15, 36, 64, 104, 208, 288, 384};
            const size_t Zc = zs[rng() % 8], C = 1 + rng() % 24;
            const size_t need = C * ((66 * Zc + 2 * Zc + 7) / 8 * 8) * 2;
            size_t off = 0, got = 0;
            if (p.alloc(need, off, got)) {
                live[next_id++] = {off, got, need};
            } else {
                // no room: after giving back every cached buffer, no free block holds the class
                CHECK(p.cached() == 0);
                CHECK(p.arena_view().largest_free() < pool::class_bytes(pool::class_of(need)));
            }
        } else if (op < 9 && !live.empty()) {
            auto it = live.begin();
            std::advance(it, rng() % live.size());
            p.release(it->second.off, it->second.bytes);
            live.erase(it);
        } else if (op == 9) {
            p.reclaim();
            CHECK(p.cached() == 0);
        }
        if (step % 97 == 0) check_live(p, live, size);
    }
    check_live(p, live, size);
    for (auto& kv : live) p.release(kv.second.off, kv.second.bytes);
    live.clear();
    p.reclaim();
    // everything coalesced back into one block
    CHECK(p.in_use() == 0 && p.arena_view().free_blocks() == 1);
    CHECK(p.arena_view().largest_free() == p.arena_view().size());
}

static void test_reuse()
{
    pool p;
    p.init(1 << 20);
    size_t a = 0, ga = 0, b = 0, gb = 0, c = 0, gc = 0;
    CHECK(p.alloc(100000, a, ga));
    CHECK(p.alloc(50000, b, gb));
    p.release(a, ga);
    CHECK(p.cached() == ga);
    CHECK(p.alloc(99000, c, gc));  // same class: the cached buffer, no search
    CHECK(c == a && gc == ga && p.cached() == 0);
    CHECK(p.peak() == ga + gb);
}

static void test_table()
{
    std::mt19937 rng(7);
    process_table t;
    std::map<std::tuple<uint32_t, uint16_t, uint8_t>, size_t> model;
    for (int i = 0; i < 50000; ++i) {
        // few RNTI/process buckets, many cells: heavy collisions
        key k{(uint32_t)(rng() % 40), (uint16_t)(16 * (rng() % 8) + 3), (uint8_t)(rng() % 3)};
        const auto mk = std::make_tuple(k.cell, k.rnti, k.pid);
        const int op = (int)(rng() % 3);
        if (op == 0) {
            process pr;
            pr.off = (size_t)i;
            t.insert(k, pr);
            model[mk] = (size_t)i;
        } else if (op == 1) {
            const process* pr = t.find(k);
            auto it = model.find(mk);
            CHECK((pr != nullptr) == (it != model.end()));
            if (pr) CHECK(pr->off == it->second);
        } else {
            CHECK(t.erase(k) == (model.erase(mk) == 1));
        }
        CHECK(t.size() == model.size());
    }
    size_t n = 0;
    t.for_each([&](const key& k, const process& pr) {
        auto it = model.find(std::make_tuple(k.cell, k.rnti, k.pid));
        CHECK(it != model.end() && it->second == pr.off);
        ++n;
    });
    CHECK(n == model.size());
}

static void test_evict()
{
    pool p;
    p.init(64 << 20);
    process_table t;
    for (int i = 0; i < 100; ++i) {
        process pr;
        size_t got = 0;
        CHECK(p.alloc(100000 + 1000 * i, pr.off, got));
        pr.bytes = got;
        pr.busy = (i % 10) == 0;
        pr.t_last = (double)i;  // us
        t.insert(key{(uint32_t)i, (uint16_t)(i * 7), (uint8_t)(i % 16)}, pr);
    }
    // at t = 150 us with a 100 us timeout: idle processes last used before 50 us go
    const size_t n = evict_idle(t, p, 150.0, 100.0);
    size_t expect = 0;
    for (int i = 0; i < 50; ++i) expect += (i % 10) != 0;
    CHECK(n == expect && t.size() == 100 - expect);
    t.for_each([&](const key& k, const process& pr) { CHECK(pr.busy || pr.t_last >= 50.0); });
}

int main()
{
    test_classes();
    test_reuse();
    test_table();
    test_evict();
    for (unsigned s = 1; s <= 6; ++s) test_random(s, (size_t)(s * 3 + 1) << 20);
    test_random(99, (size_t)256 << 20);
    std::printf("harq pool: all checks passed\n");
    return 0;
}
