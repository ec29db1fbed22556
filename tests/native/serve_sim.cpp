// serve_sim.cpp — CPU test harness of the serving loop's host side.
//
// TEST CODE.  Runs the library's own scheduler (pps::ServerCore from
// photonprop.jl_amd/csrc/pp_serve_core.hpp, the code pp_server_* runs) against
// a simulated persistent kernel: host threads that follow the serving kernel's
// protocol (tickets, gate word raised by the head block from `published`,
// range entries mirrored and searched between first_live and the mirrored
// count, slot descriptors, per-slot tile counts, completion words then the
// sequence tag, claim reports, host stop and the kernel's own stop after
// waiting with nothing published).  A tile's "photons" give deterministic
// results: every photon runs max_steps steps, photon i of a request hits when
// (i * 2654435761 + seed) % 97 == 0, with weight 1 and time 2 (exact sums).
// Built plain, with AddressSanitizer + UndefinedBehaviorSanitizer and with
// ThreadSanitizer by tests/test_native_host.py; `serve_sim <scenario>` exits
// 0 and prints PASS when every check of the scenario holds.
#include <atomic>
#include <cassert>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../photonprop.jl_amd/csrc/pp_serve_core.hpp"

using namespace pps;
using clk = std::chrono::steady_clock;

namespace {

int g_failures = 0;
#define CHECK(cond, ...)                                                             \
    do {                                                                             \
        if (!(cond)) {                                                               \
            ++g_failures;                                                            \
            fprintf(stderr, "CHECK failed %s:%d: %s: ", __FILE__, __LINE__, #cond);  \
            fprintf(stderr, __VA_ARGS__);                                            \
            fprintf(stderr, "\n");                                                   \
        }                                                                            \
    } while (0)

constexpr uint64_t kStop = 1ull << 63;
constexpr uint64_t kExit = ~0ull;

bool photon_hits(uint64_t seed, uint64_t i) { return (i * 2654435761ull + seed) % 97ull == 0; }

// The simulated device side of a server.
struct SimBackend {
    uint32_t n_slots = 0;
    uint64_t pool_records = 0;
    int n_blocks = 4;
    std::chrono::microseconds wait{1000000};
    // shared with the host (pinned host memory in the library)
    std::vector<SrvControl> ctl_mem{1};
    std::vector<SrvEntry> entries_mem;
    std::vector<SrvSlot> slots_mem;
    std::vector<SrvDone> done_mem;
    // device memory
    std::atomic<uint64_t> ticket{0}, gate{0};
    std::atomic<uint64_t> mirrored{0}, mirrored_first_live{0};
    std::vector<SrvEntry> dentries;
    std::mutex dentries_mu;  // the device ring's copies (the kernel uses agent-scope stores)
    struct SlotAcc {
        std::atomic<uint64_t> tiles{0}, steps{0}, hits{0}, cursor{0};
    };
    std::vector<SlotAcc> acc;
    std::vector<pp_hit> pool;
    std::vector<std::thread> blocks;
    std::atomic<int> live{0};
    // observations for the checks
    std::mutex log_mu;
    std::vector<SrvEntry> entry_log;       // every entry as mirrored, in order
    std::atomic<uint64_t> search_misses{0};
    std::atomic<bool> pause{false};        // blocks take no tickets while set

    void init(const pp_server_config& cfg, int blocks_, std::chrono::microseconds wait_) {
        n_slots = cfg.max_slots;
        pool_records = cfg.hit_pool_records;
        n_blocks = blocks_;
        wait = wait_;
        memset(ctl_mem.data(), 0, sizeof(SrvControl));
        entries_mem.assign(kSrvRing, SrvEntry{});
        slots_mem.assign(n_slots, SrvSlot{});
        done_mem.assign(n_slots, SrvDone{});
        dentries.assign(kSrvRing, SrvEntry{});
        acc = std::vector<SlotAcc>(n_slots);
        pool.assign(std::max<uint64_t>(pool_records, 1), pp_hit{});
    }

    ~SimBackend() { finish(); }

    SrvShared shared() { return {ctl_mem.data(), entries_mem.data(), slots_mem.data(), done_mem.data()}; }
    int thread_init() { return PP_OK; }
    uint64_t grid() const { return (uint64_t)n_blocks; }
    bool running() { return live.load(std::memory_order_acquire) > 0; }

    void finish() {
        for (auto& t : blocks)
            if (t.joinable()) t.join();
        blocks.clear();
    }

    int launch(uint64_t resume) {
        finish();
        ticket.store(resume);
        gate.store(resume);
        live.store(n_blocks, std::memory_order_release);
        for (int b = 0; b < n_blocks; ++b) blocks.emplace_back([this] { block_main(); });
        return PP_OK;
    }

    int copy_hits(uint64_t base, uint64_t n, pp_hit* out) {
        if (base + n > pool.size()) return fail(PP_ERR_INVALID_ARG, "copy beyond the pool");
        memcpy(out, pool.data() + base, n * sizeof(pp_hit));
        return PP_OK;
    }

    SrvControl* ctl() { return ctl_mem.data(); }

    // The head's copy of new entries into the device ring (mirror_entries).
    void mirror() {
        const uint64_t ne = __atomic_load_n(&ctl()->n_entries, __ATOMIC_ACQUIRE);
        const uint64_t fl = __atomic_load_n(&ctl()->first_live, __ATOMIC_ACQUIRE);
        const uint64_t m = mirrored.load();
        {
            std::lock_guard<std::mutex> g(dentries_mu);
            std::lock_guard<std::mutex> lg(log_mu);
            for (uint64_t e = m; e < ne; ++e) {
                dentries[e % kSrvRing] = entries_mem[e % kSrvRing];
                entry_log.push_back(entries_mem[e % kSrvRing]);
            }
        }
        mirrored_first_live.store(fl, std::memory_order_release);
        mirrored.store(ne, std::memory_order_release);
    }

    // Entry holding tile t: the highest mirrored entry in [first_live, count)
    // with t0 <= t (the kernel's 64-ary search, here linear from the top).
    bool find(uint64_t t, SrvEntry* out) {
        for (int attempt = 0; attempt < 1000000; ++attempt) {
            const uint64_t hi = mirrored.load(std::memory_order_acquire);
            const uint64_t lo = mirrored_first_live.load(std::memory_order_acquire);
            std::lock_guard<std::mutex> g(dentries_mu);
            for (uint64_t e = hi; e > lo; --e) {
                const SrvEntry& en = dentries[(e - 1) % kSrvRing];
                if (en.t0 <= t) {
                    if (t < en.t1) {
                        *out = en;
                        return true;
                    }
                    break;
                }
            }
            search_misses.fetch_add(1);
            std::this_thread::yield();
        }
        return false;
    }

    void run_tile(uint64_t t, const SrvEntry& e) {
        const uint32_t slot = e.slot;
        SrvSlot sd;
        memcpy(&sd, &slots_mem[slot], sizeof sd);
        if ((uint32_t)sd.seq != e.tag) {
            fprintf(stderr, "tile %llu: slot %u holds seq %llu, entry tag %u\n",
                    (unsigned long long)t, slot, (unsigned long long)sd.seq, e.tag);
            std::abort();
        }
        const uint64_t local = t - e.t0 + e.local0;
        const uint64_t first = local * kTilePhotons;
        const uint64_t count = std::min<uint64_t>(kTilePhotons, sd.req.n_photons - first);
        const uint64_t seed = ((uint64_t)sd.req.seed_hi << 32) | sd.req.seed_lo;
        SlotAcc& a = acc[slot];
        uint64_t hits = 0;
        for (uint64_t i = first; i < first + count; ++i) {
            if (!photon_hits(seed, i)) continue;
            ++hits;
            const uint64_t c = a.cursor.fetch_add(1);
            if (c < sd.hit_cap) {
                pp_hit h{};
                h.photon_id = i;
                h.request = slot;
                h.weight = 1.0f;
                h.time = 2.0f;
                pool[sd.hit_base + c] = h;
            }
        }
        a.steps.fetch_add(count * sd.req.max_steps);
        a.hits.fetch_add(hits);
        const uint64_t done = a.tiles.fetch_add(1, std::memory_order_acq_rel) + 1;
        if (done == sd.n_tiles) {
            pp_summary s{};
            s.n_sim = sd.req.n_photons;
            s.n_steps = a.steps.load();
            s.n_hits = a.hits.load();
            const uint64_t cur = a.cursor.load();
            s.n_stored = cur < sd.hit_cap ? cur : sd.hit_cap;
            s.sum_w = (double)s.n_hits;
            s.sum_wt = 2.0 * (double)s.n_hits;
            s.sum_wt2 = 4.0 * (double)s.n_hits;
            s.status = s.n_hits > s.n_stored ? PP_ERR_OVERFLOW : PP_OK;
            a.tiles.store(0);
            a.steps.store(0);
            a.hits.store(0);
            a.cursor.store(0);
            memcpy(done_mem[slot].words, &s, sizeof s);
            __atomic_store_n(&done_mem[slot].seq, sd.seq, __ATOMIC_RELEASE);
        }
    }

    void block_main() {
        for (;;) {
            while (pause.load(std::memory_order_acquire)) std::this_thread::sleep_for(std::chrono::microseconds(50));
            const uint64_t t = ticket.fetch_add(1);
            uint64_t got = kExit;
            clk::time_point wait_start{};
            bool waiting = false;
            for (;;) {
                const uint64_t g = gate.load(std::memory_order_acquire);
                if (t < (g & ~kStop)) {
                    got = t;
                    break;
                }
                if (g & kStop) break;
                if (t == g) {  // head
                    const uint64_t hp = __atomic_load_n(&ctl()->published, __ATOMIC_ACQUIRE);
                    if (hp > t) {
                        mirror();
                        uint64_t cur = gate.load();
                        while ((cur & ~kStop) < hp && !gate.compare_exchange_weak(cur, (cur & kStop) | hp)) {
                        }
                        got = t;
                        break;
                    }
                    if (__atomic_load_n(&ctl()->stop, __ATOMIC_ACQUIRE)) {
                        gate.fetch_or(kStop);
                        break;
                    }
                    const auto now = clk::now();
                    if (!waiting) {
                        waiting = true;
                        wait_start = now;
                    } else if (now - wait_start > wait) {
                        __atomic_store_n(&ctl()->self_stop, t + 1, __ATOMIC_RELEASE);
                        gate.fetch_or(kStop);
                        break;
                    }
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                } else {
                    std::this_thread::sleep_for(std::chrono::microseconds(50));
                }
            }
            if (got == kExit) break;
            if ((got & 63u) == 0) __atomic_store_n(&ctl()->claimed, got, __ATOMIC_RELAXED);
            SrvEntry e;
            if (!find(got, &e)) {
                fprintf(stderr, "tile %llu: no range entry found\n", (unsigned long long)got);
                std::abort();
            }
            run_tile(got, e);
        }
        live.fetch_sub(1, std::memory_order_acq_rel);
    }
};

using Server = ServerCore<SimBackend>;

struct Expect {
    uint64_t n_photons, max_steps, seed;
};

pp_request make_request(uint64_t n_photons, uint64_t seed, uint32_t max_steps = 10) {
    pp_request q;
    memset(&q, 0, sizeof q);
    q.emitter.axis[2] = 1.0f;
    q.emitter.beta = 1.0f;
    q.emitter.n_photons = n_photons;
    q.emitter.profile = PP_EMIT_CHERENKOV_CONE;
    q.target.center[0] = 3.0f;
    q.target.radius = 0.3f;
    q.seed = seed;
    q.max_steps = max_steps;
    return q;
}

uint64_t expected_hits(uint64_t n, uint64_t seed) {
    uint64_t h = 0;
    for (uint64_t i = 0; i < n; ++i) h += photon_hits(seed, i);
    return h;
}

// Submit, then check one finished ticket's summary and records.
void check_result(Server& s, uint64_t tk, const pp_request& q, uint64_t cap, uint64_t hits_cap_out) {
    pp_summary sum;
    std::vector<pp_hit> out(std::max<uint64_t>(hits_cap_out, 1));
    uint64_t n = 0;
    double lat = 0;
    const int rc = s.result(tk, &sum, out.data(), hits_cap_out, &n, &lat);
    CHECK(rc == PP_OK, "result(%llu) = %d: %s", (unsigned long long)tk, rc, pph::last_error().c_str());
    if (rc) return;
    const uint64_t N = q.emitter.n_photons;
    const uint64_t H = expected_hits(N, q.seed);
    CHECK(sum.n_sim == N, "n_sim %llu != %llu", (unsigned long long)sum.n_sim, (unsigned long long)N);
    CHECK(sum.n_steps == N * q.max_steps, "n_steps");
    CHECK(sum.n_hits == H, "n_hits %llu != %llu", (unsigned long long)sum.n_hits, (unsigned long long)H);
    CHECK(sum.n_stored == std::min(H, N ? cap : 0), "n_stored %llu", (unsigned long long)sum.n_stored);
    CHECK(sum.sum_w == (double)H && sum.sum_wt2 == 4.0 * (double)H, "sums");
    CHECK(n == std::min<uint64_t>(sum.n_stored, hits_cap_out), "records returned");
    std::set<uint64_t> ids;
    for (uint64_t i = 0; i < n; ++i) {
        CHECK(out[i].photon_id < N && photon_hits(q.seed, out[i].photon_id), "record %llu not a hit",
              (unsigned long long)i);
        ids.insert(out[i].photon_id);
    }
    CHECK(ids.size() == n, "duplicate records");
    CHECK(lat >= 0.0, "latency");
}

void wait_done(Server& s, uint64_t tk) {
    int done = 0;
    const int rc = s.wait(tk, 60000, &done);
    CHECK(rc == PP_OK && done, "wait(%llu): rc %d done %d %s", (unsigned long long)tk, rc, done,
          pph::last_error().c_str());
}

pp_server_config config(uint32_t slots, uint32_t tiles, uint64_t pool, uint32_t idle_us) {
    pp_server_config c;
    memset(&c, 0, sizeof c);
    c.max_slots = slots;
    c.max_tiles_per_request = tiles;
    c.hit_pool_records = pool;
    c.idle_exit_us = idle_us;
    return c;
}

void start(Server& s, const pp_server_config& cfg, int blocks, uint32_t wait_us = 1000000) {
    s.be.init(cfg, blocks, std::chrono::microseconds(wait_us));
    s.start();
}

uint64_t submit(Server& s, const pp_request& q, uint64_t cap) {
    DevReq dr;
    int rc = pph::make_devreq(&q, 0, &dr);
    CHECK(rc == PP_OK, "make_devreq");
    uint64_t tk = 0;
    rc = s.submit(dr, cap, &tk);
    CHECK(rc == PP_OK, "submit: %d %s", rc, pph::last_error().c_str());
    return tk;
}

// ---- scenarios ---------------------------------------------------------------

// More one-tile requests than twice the 16384-entry ring, few slots: entries
// wrap the ring (host copy and device copy) and retire in order.
void scenario_ring_wrap(uint64_t n_req) {
    Server s(config(256, 64, 1 << 16, 20000));
    start(s, config(256, 64, 1 << 16, 20000), 8);
    std::vector<uint64_t> tks;
    std::vector<pp_request> qs;
    for (uint64_t i = 0; i < n_req; ++i) {
        qs.push_back(make_request(1 + (i * 7919) % 8000, 1000 + i));
        tks.push_back(submit(s, qs.back(), 4));
        if (tks.size() >= 2000) {  // collect in order, as a client would
            for (size_t k = 0; k < tks.size(); ++k) {
                wait_done(s, tks[k]);
                check_result(s, tks[k], qs[k], 4, 4);
            }
            tks.clear();
            qs.clear();
        }
    }
    for (size_t k = 0; k < tks.size(); ++k) {
        wait_done(s, tks[k]);
        check_result(s, tks[k], qs[k], 4, 4);
    }
    pp_server_stats st;
    s.stats(&st);
    CHECK(st.submitted == n_req && st.completed == n_req && st.in_flight == 0, "stats");
    CHECK(st.pool_records_free == (1u << 16), "pool returned");
    CHECK(st.tiles_published == n_req, "one tile each");
    s.shutdown();
    const SrvControl* c = s.be.ctl();
    CHECK(c->first_live == c->n_entries && c->n_entries == n_req, "every entry retired");
    CHECK(s.be.entry_log.size() == n_req, "entries mirrored once each");
}

// Large requests are published in chunks of max(64, n/128) tiles, first come
// first served; small ones (<= 512 tiles) whole and ahead of the large ones'
// remaining chunks; every request's entries cover its tiles once, in order.
void scenario_chunking() {
    const pp_server_config cfg = config(16, 20000, 1 << 20, 20000);
    Server s(cfg);
    start(s, cfg, 16);
    s.be.pause = true;  // no claims yet: only the backlog rule lets chunks out
    const uint64_t big = 9000ull * kTilePhotons;  // 9000 tiles: chunks of max(64, 71) = 71
    std::vector<pp_request> qs = {make_request(big, 1), make_request(600 * kTilePhotons, 2),
                                  make_request(300 * kTilePhotons + 5, 3), make_request(40, 4)};
    std::vector<uint64_t> tks;
    for (auto& q : qs) tks.push_back(submit(s, q, 64));
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    s.be.pause = false;
    for (size_t k = 0; k < tks.size(); ++k) {
        wait_done(s, tks[k]);
        check_result(s, tks[k], qs[k], 64, 64);
    }
    s.shutdown();
    // reconstruct each request's entries (slot + tag identify the request)
    std::map<uint64_t, std::vector<SrvEntry>> by_req;
    std::map<uint64_t, size_t> first_index;
    const auto& log = s.be.entry_log;
    for (size_t i = 0; i < log.size(); ++i) {
        by_req[log[i].tag].push_back(log[i]);
        first_index.emplace(log[i].tag, i);
    }
    CHECK(by_req.size() == 4, "four requests published (%zu)", by_req.size());
    uint64_t t_expect = 0;
    for (size_t i = 0; i < log.size(); ++i) {
        CHECK(log[i].t0 == t_expect && log[i].t1 > log[i].t0, "entries contiguous");
        t_expect = log[i].t1;
    }
    const uint64_t n_tiles[4] = {9000, 600, 301, 1};
    const bool small[4] = {false, false, true, true};
    uint64_t tags[4];
    for (int r = 0; r < 4; ++r) tags[r] = (uint32_t)(r + 1);  // admitted in order: seq 1..4
    for (int r = 0; r < 4; ++r) {
        const auto& es = by_req[tags[r]];
        uint64_t local = 0;
        for (size_t k = 0; k < es.size(); ++k) {
            CHECK(es[k].local0 == local, "request %d: chunks in order", r);
            const uint64_t len = es[k].t1 - es[k].t0;
            const uint64_t chunk = std::max<uint64_t>(64, (n_tiles[r] + 127) / 128);
            if (small[r]) CHECK(es.size() == 1 && len == n_tiles[r], "request %d published whole", r);
            else CHECK(len == std::min(chunk, n_tiles[r] - local), "request %d chunk size %llu", r,
                       (unsigned long long)len);
            local += len;
        }
        CHECK(local == n_tiles[r], "request %d: all tiles once", r);
    }
    // the small ones went out before the first large one's last chunk
    CHECK(first_index[tags[2]] < first_index[tags[0]] + by_req[tags[0]].size() - 1, "small first");
    CHECK(first_index[tags[3]] < first_index[tags[0]] + by_req[tags[0]].size() - 1, "small first");
}

// Zero-photon requests take no slot and no pool records; requests beyond
// the slots and the pool wait in FIFO order; the pool comes back whole.
void scenario_queueing() {
    const pp_server_config cfg = config(3, 64, 3000, 20000);
    Server s(cfg);
    start(s, cfg, 4);
    // first request without photons on a fresh server, collected at once
    pp_request z = make_request(0, 9);
    uint64_t tz = submit(s, z, 1000);
    wait_done(s, tz);
    check_result(s, tz, z, 1000, 1000);
    std::this_thread::sleep_for(std::chrono::milliseconds(15));
    pp_server_stats st;
    s.stats(&st);
    CHECK(st.submitted == 1 && st.completed == 1 && st.in_flight == 0 && st.pending == 0,
          "zero-photon request: submitted %llu completed %llu", (unsigned long long)st.submitted,
          (unsigned long long)st.completed);
    int done = 0;
    CHECK(s.wait(tz, 0, &done) == PP_ERR_NOT_FOUND, "collected ticket is unknown");
    std::vector<pp_request> qs;
    std::vector<uint64_t> tks, caps;
    for (int i = 0; i < 40; ++i) {
        const uint64_t n = (i % 5 == 0) ? 0 : (uint64_t)(500 + 1000 * (i % 7)) * 10;
        qs.push_back(make_request(n, 100 + i));
        caps.push_back(i % 4 == 0 ? 3000 : 700);  // some need the whole pool
        tks.push_back(submit(s, qs.back(), caps.back()));
    }
    for (size_t k = 0; k < tks.size(); ++k) {
        wait_done(s, tks[k]);
        check_result(s, tks[k], qs[k], caps[k], caps[k]);
    }
    s.stats(&st);
    CHECK(st.completed == st.submitted && st.in_flight == 0 && st.pool_records_free == 3000,
          "all done, pool whole");
}

// The host stops the kernel after idle_exit_us without work; the next
// request relaunches it at the published count.
void scenario_idle_relaunch() {
    const pp_server_config cfg = config(4, 256, 1 << 16, 2000);
    Server s(cfg);
    start(s, cfg, 6);
    for (int wave = 0; wave < 3; ++wave) {
        std::vector<pp_request> qs;
        std::vector<uint64_t> tks;
        for (int i = 0; i < 5; ++i) {
            qs.push_back(make_request(20000 + 9000 * i + wave, 40 + 10 * wave + i));
            tks.push_back(submit(s, qs.back(), 800));
        }
        for (size_t k = 0; k < tks.size(); ++k) {
            wait_done(s, tks[k]);
            check_result(s, tks[k], qs[k], 800, 800);
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(100));
        pp_server_stats st;
        s.stats(&st);
        CHECK(st.kernel_running == 0, "kernel stopped after the idle time");
    }
    pp_server_stats st;
    s.stats(&st);
    CHECK(st.kernel_launches >= 3 && st.device_stops == 0, "launches %llu",
          (unsigned long long)st.kernel_launches);
}

// With the host's idle stop far away, the kernel stops by itself after
// waiting with nothing published and reports where to resume.
void scenario_device_stop() {
    const pp_server_config cfg = config(4, 256, 1 << 16, 10000000);
    Server s(cfg);
    start(s, cfg, 6, 2000);
    for (int wave = 0; wave < 3; ++wave) {
        pp_request q = make_request(70000 + wave, 7 + wave);
        uint64_t tk = submit(s, q, 900);
        wait_done(s, tk);
        check_result(s, tk, q, 900, 900);
        std::this_thread::sleep_for(std::chrono::milliseconds(100));
        pp_server_stats st;
        s.stats(&st);
        CHECK(st.kernel_running == 0 && st.device_stops == (uint64_t)wave + 1,
              "device stop %d: running %llu stops %llu", wave, (unsigned long long)st.kernel_running,
              (unsigned long long)st.device_stops);
    }
}

// Several submitting and collecting threads; one ticket collected by two
// threads at once is returned to exactly one of them.
void scenario_threads(int n_threads, int per_thread) {
    const pp_server_config cfg = config(32, 64, 1 << 15, 5000);
    Server s(cfg);
    start(s, cfg, 8);
    std::atomic<int> bad{0};
    std::vector<std::thread> th;
    for (int w = 0; w < n_threads; ++w) {
        th.emplace_back([&, w] {
            std::mt19937_64 rng(w);
            for (int i = 0; i < per_thread; ++i) {
                pp_request q = make_request(rng() % 60000, 1000 * w + i);
                DevReq dr;
                pph::make_devreq(&q, 0, &dr);
                uint64_t tk;
                if (s.submit(dr, 256, &tk) != PP_OK) {
                    ++bad;
                    continue;
                }
                int done = 0;
                if (s.wait(tk, 60000, &done) != PP_OK || !done) {
                    ++bad;
                    continue;
                }
                // two collectors race for the same ticket
                std::atomic<int> ok{0};
                std::vector<pp_hit> a(256), b(256);
                pp_summary sa, sb;
                uint64_t na = 0, nb = 0;
                std::thread other([&] {
                    if (s.result(tk, &sb, b.data(), 256, &nb, nullptr) == PP_OK) ++ok;
                });
                if (s.result(tk, &sa, a.data(), 256, &na, nullptr) == PP_OK) ++ok;
                other.join();
                if (ok.load() != 1) ++bad;
            }
        });
    }
    for (auto& t : th) t.join();
    CHECK(bad.load() == 0, "%d failures in the threaded run", bad.load());
    pp_server_stats st;
    s.stats(&st);
    CHECK(st.completed == st.submitted && st.submitted == (uint64_t)n_threads * per_thread &&
              st.pool_records_free == (1u << 15),
          "threaded: submitted %llu completed %llu pool %llu", (unsigned long long)st.submitted,
          (unsigned long long)st.completed, (unsigned long long)st.pool_records_free);
}

// The host-side input checks on edge values (they run in the sanitizer builds).
void scenario_validate() {
    float n[3] = {1.34f, 1.35f, 1.36f}, ab[3] = {10, 20, 30}, sc[3] = {5, 6, 7};
    pp_medium_tables m;
    memset(&m, 0, sizeof m);
    m.lambda_min = 300;
    m.lambda_max = 600;
    m.n_points = 3;
    m.n_phase = n;
    m.abs_len = ab;
    m.sca_len = sc;
    m.g = 0.9f;
    CHECK(pph::check_medium(&m) == PP_OK, "medium");
    ab[1] = std::nanf("");
    CHECK(pph::check_medium(&m) == PP_ERR_INVALID_ARG, "NaN absorption length");
    ab[1] = 20;
    m.n_points = 1;
    CHECK(pph::check_medium(&m) == PP_ERR_INVALID_ARG, "one point");
    m.n_points = 3;
    m.g = 1.0f;
    CHECK(pph::check_medium(&m) == PP_ERR_INVALID_ARG, "g = 1");
    CHECK(pph::check_medium(nullptr) == PP_ERR_INVALID_ARG, "NULL medium");
    pph::SpecParams sp;
    pp_spectrum spec;
    memset(&spec, 0, sizeof spec);
    spec.kind = 7;
    CHECK(pph::check_spectrum(&spec, &sp) == PP_ERR_UNSUPPORTED, "unknown kind");
    spec.kind = PP_SPEC_KNOTS;
    spec.n_knots = PP_MAX_KNOTS + 1;
    CHECK(pph::check_spectrum(&spec, &sp) == PP_ERR_INVALID_ARG, "too many knots");
    pp_request q = make_request(10, 1);
    DevReq d;
    CHECK(pph::make_devreq(&q, 0, &d) == PP_OK, "request");
    q.max_steps = PP_MAX_STEPS + 1;
    CHECK(pph::make_devreq(&q, 0, &d) == PP_ERR_INVALID_ARG, "max_steps");
    q = make_request(10, 1);
    q.emitter.axis[2] = 0.0f;
    CHECK(pph::make_devreq(&q, 0, &d) == PP_ERR_INVALID_ARG, "zero axis");
    q = make_request(10, 1);
    q.emitter.beta = 0.5f;
    CHECK(pph::check_cone(&q, 1, 1.34f) == PP_ERR_INVALID_ARG, "no Cherenkov light");
    CHECK(pph::out_records(100, 7, 20) == 100 && pph::out_records(1000, 7, 20) == 140, "out_records");
    CHECK(pph::out_records(~0ull, ~0ull, 4096) == ~0ull, "out_records without wrap-around");
    CHECK(pph::check_records(PP_MAX_HIT_RECORDS + 1, "x") == PP_ERR_INVALID_ARG, "records limit");
}

}  // namespace

// ---- the load generator against the simulated server ---------------------------
// pp_loadgen.cpp is linked in; it drives the server through these three
// functions, which have the signatures of pp_server_submit/wait/result.
typedef int (*submit_fn)(pp_server*, const pp_request*, uint64_t, uint64_t*);
typedef int (*wait_fn)(pp_server*, uint64_t, uint32_t, int*);
typedef int (*result_fn)(pp_server*, uint64_t, pp_summary*, pp_hit*, uint64_t, uint64_t*,
                         double*);
extern "C" int ppl_open_loop(pp_server* srv, submit_fn submit, wait_fn wait, result_fn result,
                             const pp_request* reqs, const double* arrivals_s, uint64_t n,
                             uint64_t hits_cap, double timeout_s, double* latency_us,
                             double* lag_us, uint64_t* steps, double* makespan_s);

extern "C" {
static int sim_submit(pp_server* srv, const pp_request* q, uint64_t cap, uint64_t* tk) {
    DevReq dr;
    int rc = pph::make_devreq(q, 0, &dr);
    if (rc) return rc;
    return reinterpret_cast<Server*>(srv)->submit(dr, cap, tk);
}
static int sim_wait(pp_server* srv, uint64_t tk, uint32_t ms, int* done) {
    return reinterpret_cast<Server*>(srv)->wait(tk, ms, done);
}
static int sim_result(pp_server* srv, uint64_t tk, pp_summary* s, pp_hit* h, uint64_t cap,
                      uint64_t* n, double* lat) {
    return reinterpret_cast<Server*>(srv)->result(tk, s, h, cap, n, lat);
}
}

namespace {

void scenario_loadgen(uint64_t n_req) {
    const pp_server_config cfg = config(64, 128, 1 << 16, 5000);
    Server s(cfg);
    start(s, cfg, 8);
    std::vector<pp_request> qs;
    std::vector<double> arr;
    std::mt19937_64 rng(3);
    std::exponential_distribution<double> gap(20000.0);
    double t = 0;
    uint64_t steps_expect = 0;
    for (uint64_t i = 0; i < n_req; ++i) {
        qs.push_back(make_request(1 + rng() % 200000, 77 + i, 1 + i % 12));
        steps_expect += qs.back().emitter.n_photons * qs.back().max_steps;
        t += gap(rng);
        arr.push_back(t);
    }
    std::vector<double> lat(n_req), lag(n_req);
    std::vector<uint64_t> steps(n_req);
    double mk = 0;
    const int rc = ppl_open_loop(reinterpret_cast<pp_server*>(&s), &sim_submit, &sim_wait, &sim_result, qs.data(), arr.data(), n_req, 0, 120.0,
                                 lat.data(), lag.data(), steps.data(), &mk);
    CHECK(rc == 0, "ppl_open_loop = %d", rc);
    uint64_t total = 0;
    for (uint64_t i = 0; i < n_req; ++i) {
        CHECK(steps[i] == qs[i].emitter.n_photons * qs[i].max_steps, "request %llu steps",
              (unsigned long long)i);
        CHECK(lat[i] >= 0, "latency");
        total += steps[i];
    }
    CHECK(total == steps_expect, "total steps");
}

}  // namespace

// Deliberate faults, run only in the sanitizer builds to show that the
// instrumentation is live: an unsynchronised counter (ThreadSanitizer) and a
// read one past a heap array (AddressSanitizer).
int canary(const std::string& which) {
    if (which == "canary_race") {
        static int counter = 0;
        std::thread a([] { for (int i = 0; i < 1000; ++i) ++counter; });
        std::thread b([] { for (int i = 0; i < 1000; ++i) ++counter; });
        a.join();
        b.join();
        return counter > 0 ? 0 : 1;
    }
    std::vector<int> v(8, 1);
    volatile size_t i = v.size();
    return v.data()[i];
}

int main(int argc, char** argv) {
    const std::string sc = argc > 1 ? argv[1] : "all";
    if (sc.rfind("canary_", 0) == 0) return canary(sc);
    const uint64_t scale = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;  // 1 = full size
    const auto t0 = clk::now();
    auto run = [&](const char* name, auto fn) {
        if (sc != "all" && sc != name) return;
        const int before = g_failures;
        fn();
        printf("%s: %s (%.2f s)\n", name, g_failures == before ? "ok" : "FAILED",
               std::chrono::duration<double>(clk::now() - t0).count());
        fflush(stdout);
    };
    run("validate", [] { scenario_validate(); });
    run("queueing", [] { scenario_queueing(); });
    run("chunking", [] { scenario_chunking(); });
    run("idle_relaunch", [] { scenario_idle_relaunch(); });
    run("device_stop", [] { scenario_device_stop(); });
    run("ring_wrap", [&] { scenario_ring_wrap(scale == 1 ? 40000 : 20000); });
    run("threads", [&] { scenario_threads(6, scale == 1 ? 60 : 20); });
    run("loadgen", [&] { scenario_loadgen(scale == 1 ? 3000 : 600); });
    if (g_failures) {
        printf("FAIL (%d checks)\n", g_failures);
        return 1;
    }
    printf("PASS\n");
    return 0;
}
