// pp_serve_core.hpp — host side of the serving loop, independent of HIP.
//
// Requests queue FIFO and are admitted when a device slot and room in the
// device hit pool (the HBM budget) are free.  Admission writes the request's
// slot descriptor to memory the persistent kernel reads; publishing appends
// range entries (global tiles [t0, t1) = the request's tiles local0...) to a
// ring and advances `published`:
//   - small requests (<= kSrvSmallTiles tiles) go out whole, ahead of the rest;
//   - large requests go out first come first served, in chunks of
//     max(kSrvChunkTiles, n / kSrvMaxChunks) tiles, while fewer than a grid's
//     worth of published tiles waits to be claimed (the kernel reports every
//     64th claimed ticket).
// Entries retire in order (first_live stops at the oldest unfinished request);
// once the ring is half full only that request may publish, and it needs at
// most kSrvMaxChunks entries, so the ring cannot fill with it unfinished.
// The kernel posts each request's summary and a sequence tag; the worker
// reaps them, stops the kernel when no admitted request is in flight for
// idle_exit_us, and relaunches it when work arrives.  The kernel also stops by
// itself after waiting device_wait_us (device clock) with nothing published, so a
// stalled host never leaves it spinning; the worker then resumes at the tile
// the kernel reports.
//
// The device side is a Backend (the library's: HIP; the tests': a simulated
// kernel on host threads) that provides
//   SrvShared shared()                 the control block, entry ring, slot
//                                      descriptors and completions
//   int thread_init()                  worker-thread set-up
//   int launch(uint64_t resume_tile)   start the kernel at that ticket
//   bool running()                     the kernel has not exited yet
//   void finish()                      wait for the kernel to exit
//   uint64_t grid() const              persistent blocks (publishing pace)
//   int copy_hits(uint64_t base, uint64_t n, pp_hit* out)   pool -> caller
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "pp_types.hpp"

namespace pps {

using pph::DevReq;
using pph::fail;
using pph::kTilePhotons;

constexpr uint32_t kSrvRing = 16384;        // range-entry ring
constexpr uint64_t kSrvChunkTiles = 64;     // smallest range entry of a large request
constexpr uint64_t kSrvMaxChunks = 128;     // range entries per large request at most
constexpr uint64_t kSrvSmallTiles = 512;    // requests up to this many tiles are small
constexpr uint64_t kSrvBacklogGrids = 1;    // published, unclaimed tiles kept: grids' worth

struct SrvEntry {       // global tiles [t0, t1) are the request's tiles local0..
    uint64_t t0, t1;
    uint32_t slot;
    uint32_t tag;       // low word of the request's completion tag (which request holds the slot)
    uint64_t local0;
};

struct SrvSlot {       // written by the host before the tiles are published
    DevReq req;
    uint64_t n_tiles;
    uint64_t hit_base;  // record offset into the hit pool
    uint64_t hit_cap;
    uint64_t seq;       // completion tag
};

struct SrvDone {       // written by the device
    uint64_t words[8];  // pp_summary
    uint64_t seq;
    uint64_t pad[7];
};

struct SrvControl {
    uint64_t published;   // end of the published global tile range (host)
    uint64_t n_entries;   // range entries published (host)
    uint64_t first_live;  // oldest entry that can still be searched (host)
    uint64_t stop;        // stop request (host)
    uint64_t claimed;     // a recently claimed tile, every 64th (device)
    uint64_t self_stop;   // resume tile + 1 when the kernel stopped by itself (device)
    uint64_t pad[2];
};

static_assert(sizeof(SrvEntry) % 8 == 0 && 64 % (sizeof(SrvEntry) / 8) == 0,
              "a range entry is copied as whole 8-byte words, a whole number per wave");
static_assert(sizeof(SrvSlot) <= 64 * 4, "one wave loads the slot descriptor, a word a lane");

// Memory shared with the kernel (host memory the device can read and write).
struct SrvShared {
    SrvControl* ctl;
    SrvEntry* entries;   // [kSrvRing]
    SrvSlot* slots;      // [max_slots]
    SrvDone* done;       // [max_slots]
};

struct Job {
    uint64_t ticket = 0;
    DevReq req;
    uint64_t hits_cap = 0;    // pool records reserved (0 for a request without photons)
    uint64_t n_tiles = 0;
    uint32_t slot = 0;
    uint64_t hit_base = 0;
    uint64_t seq = 0;
    uint64_t next_local = 0;  // tiles published so far
    bool done = false;
    pp_summary summary;
    std::chrono::steady_clock::time_point t_submit, t_done;
};

template <class Backend>
class ServerCore {
public:
    Backend be;

    // cfg validated by the caller; call after the backend's shared memory is
    // set up (zeroed).
    explicit ServerCore(const pp_server_config& cfg) : cfg_(cfg) {}

    ~ServerCore() { shutdown(); }

    int start() {
        const uint32_t S = cfg_.max_slots;
        sh_ = be.shared();
        free_slots_.reserve(S);
        for (uint32_t i = 0; i < S; ++i) free_slots_.push_back(S - 1 - i);
        slot_ticket_.assign(S, 0);
        entry_ticket_.assign(kSrvRing, 0);
        pool_free_total_ = cfg_.hit_pool_records;
        if (cfg_.hit_pool_records) pool_free_[0] = cfg_.hit_pool_records;
        worker_ = std::thread([this] { worker_main(); });
        return PP_OK;
    }

    // Stop admitting, let the kernel finish and exit, join the worker.
    void shutdown() {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (shut_) return;
            shut_ = true;
            cv_.notify_all();
        }
        if (worker_.joinable()) worker_.join();
        if (sh_.ctl) __atomic_store_n(&sh_.ctl->stop, 1ull, __ATOMIC_RELEASE);
        be.finish();
    }

    // Queue a checked request (dr its kernel record).  Never blocks.
    int submit(const DevReq& dr, uint64_t hits_cap, uint64_t* ticket_out) {
        const uint64_t n_tiles = (dr.n_photons + kTilePhotons - 1) / kTilePhotons;
        if (n_tiles > cfg_.max_tiles_per_request)
            return fail(PP_ERR_INVALID_ARG, "request of %llu photons exceeds %u tiles of %d",
                        (unsigned long long)dr.n_photons, cfg_.max_tiles_per_request,
                        PP_TILE_PHOTONS);
        if (hits_cap > cfg_.hit_pool_records)
            return fail(PP_ERR_INVALID_ARG, "hits_cap %llu exceeds the hit pool (%llu records)",
                        (unsigned long long)hits_cap, (unsigned long long)cfg_.hit_pool_records);
        std::lock_guard<std::mutex> g(mu_);
        if (!worker_error_.empty()) return fail(PP_ERR_DEVICE, "%s", worker_error_.c_str());
        Job j;
        j.ticket = next_ticket_++;
        j.req = dr;
        j.hits_cap = n_tiles ? hits_cap : 0;  // a job without photons reserves no records
        j.n_tiles = n_tiles;
        j.t_submit = std::chrono::steady_clock::now();
        jobs_[j.ticket] = j;
        pending_.push_back(j.ticket);
        ++submitted_;
        admit_locked();
        cv_.notify_all();
        *ticket_out = j.ticket;
        return PP_OK;
    }

    int wait(uint64_t ticket, uint32_t timeout_ms, int* done_out) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = jobs_.find(ticket);
        if (it == jobs_.end())
            return fail(PP_ERR_NOT_FOUND, "unknown ticket %llu", (unsigned long long)ticket);
        const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms);
        while (!it->second.done && worker_error_.empty()) {
            if (cv_.wait_until(lk, deadline) == std::cv_status::timeout) break;
            it = jobs_.find(ticket);
            if (it == jobs_.end())
                return fail(PP_ERR_NOT_FOUND, "ticket %llu collected meanwhile",
                            (unsigned long long)ticket);
        }
        if (!worker_error_.empty()) return fail(PP_ERR_DEVICE, "%s", worker_error_.c_str());
        *done_out = it->second.done ? 1 : 0;
        return PP_OK;
    }

    int result(uint64_t ticket, pp_summary* summary, pp_hit* hits_out, uint64_t hits_cap,
               uint64_t* n_hits_out, double* latency_us) {
        Job j;
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = jobs_.find(ticket);
            if (it == jobs_.end())
                return fail(PP_ERR_NOT_FOUND, "unknown ticket %llu", (unsigned long long)ticket);
            if (!it->second.done)
                return fail(PP_ERR_BUSY, "ticket %llu not finished", (unsigned long long)ticket);
            if (collecting_.count(ticket))
                return fail(PP_ERR_BUSY, "ticket %llu is being collected", (unsigned long long)ticket);
            collecting_.insert(ticket);
            j = it->second;
        }
        const uint64_t n = std::min<uint64_t>(j.summary.n_stored, hits_cap);
        int rc = n ? be.copy_hits(j.hit_base, n, hits_out) : PP_OK;
        std::lock_guard<std::mutex> g(mu_);
        collecting_.erase(ticket);
        if (rc) return rc;  // the ticket stays; the caller may try again
        if (summary) *summary = j.summary;
        if (n_hits_out) *n_hits_out = n;
        if (latency_us)
            *latency_us = std::chrono::duration<double, std::micro>(j.t_done - j.t_submit).count();
        pool_release(j.hit_base, j.hits_cap);
        jobs_.erase(ticket);
        admit_locked();
        cv_.notify_all();
        return PP_OK;
    }

    void stats(pp_server_stats* out) {
        std::lock_guard<std::mutex> g(mu_);
        out->submitted = submitted_;
        out->completed = completed_;
        out->pending = pending_.size();
        out->in_flight = submitted_ - completed_ - pending_.size();
        out->kernel_launches = launches_;
        out->pool_records_free = pool_free_total_;
        out->tiles_published = published_;
        out->grid_blocks = be.grid();
        out->device_stops = device_stops_;
        out->kernel_running = kernel_live_ ? 1 : 0;
    }

private:
    // ---- hit pool: first fit over free ranges, coalesced on release ----
    bool pool_alloc(uint64_t n, uint64_t* off) {
        if (n == 0) {
            *off = 0;
            return true;
        }
        for (auto it = pool_free_.begin(); it != pool_free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, at = it->first + n;
                pool_free_.erase(it);
                if (rest) pool_free_[at] = rest;
                pool_free_total_ -= n;
                return true;
            }
        }
        return false;
    }

    void pool_release(uint64_t off, uint64_t n) {
        if (n == 0) return;
        pool_free_total_ += n;
        auto it = pool_free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != pool_free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            pool_free_.erase(nx);
        }
        if (it != pool_free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                pool_free_.erase(it);
            }
        }
    }

    // ---- publishing ----
    // Append one range entry: tiles local0 .. local0 + c of job j.
    void publish_entry(Job& j, uint64_t c) {
        const uint64_t e = n_entries_;
        SrvEntry en;
        memset(&en, 0, sizeof en);
        en.t0 = published_;
        en.t1 = published_ + c;
        en.slot = j.slot;
        en.tag = (uint32_t)j.seq;
        en.local0 = j.next_local;
        memcpy(&sh_.entries[e % kSrvRing], &en, sizeof en);
        entry_ticket_[e % kSrvRing] = j.ticket;
        published_ = en.t1;
        n_entries_ = e + 1;
        j.next_local += c;
    }

    void publish_locked() {
        if (stopping_) return;  // tickets restart at resume_tile_ on the next launch
        // Claim reports are plain stores from many CUs and may land out of
        // order; the highest seen is kept.  Should the estimate stay full for
        // 1 ms with neither a new report nor a new chunk (an older report
        // landing last), everything published is taken as claimed: that can
        // only publish early, never stall the kernel.
        const auto now = std::chrono::steady_clock::now();
        const uint64_t target =
            std::max<uint64_t>(kSrvBacklogGrids * be.grid(), 4 * kSrvChunkTiles);
        const uint64_t seen = __atomic_load_n(&sh_.ctl->claimed, __ATOMIC_ACQUIRE);
        if (seen > claimed_seen_ && seen <= published_) {
            claimed_seen_ = seen;
            claim_change_ = now;
        } else if (published_ - claimed_seen_ >= target &&
                   now - claim_change_ > std::chrono::milliseconds(1)) {
            claimed_seen_ = published_;
            claim_change_ = now;
        }
        const uint64_t n_before = n_entries_;
        auto in_use = [this] { return n_entries_ - first_live_; };
        const uint64_t pin =
            first_live_ < n_entries_ ? entry_ticket_[first_live_ % kSrvRing] : 0;
        while (!small_q_.empty() && in_use() < kSrvRing / 2) {
            Job& j = jobs_.find(small_q_.front())->second;
            small_q_.pop_front();
            publish_entry(j, j.n_tiles);
        }
        while (!large_q_.empty() && published_ - claimed_seen_ < target) {
            Job& j = jobs_.find(large_q_.front())->second;
            if (in_use() >= kSrvRing / 2 && j.ticket != pin && in_use() > 0) break;
            if (in_use() >= kSrvRing) break;
            const uint64_t chunk =
                std::max(kSrvChunkTiles, (j.n_tiles + kSrvMaxChunks - 1) / kSrvMaxChunks);
            publish_entry(j, std::min(chunk, j.n_tiles - j.next_local));
            if (j.next_local == j.n_tiles) large_q_.pop_front();
            claim_change_ = now;  // the 1 ms without reports counts from the last chunk
        }
        if (n_entries_ != n_before) {
            std::atomic_thread_fence(std::memory_order_seq_cst);
            __atomic_store_n(&sh_.ctl->n_entries, n_entries_, __ATOMIC_RELEASE);
            __atomic_store_n(&sh_.ctl->published, published_, __ATOMIC_RELEASE);
        }
    }

    // Admit pending jobs (FIFO) while a slot and pool room exist, then publish.
    // A job without photons completes at once and takes neither a slot nor
    // pool records, so no slot ever carries the ticket of a job that the
    // device will not complete.
    void admit_locked() {
        if (stopping_) return;
        while (!pending_.empty()) {
            auto it = jobs_.find(pending_.front());
            if (it == jobs_.end()) {  // cannot happen: pending jobs are never erased
                pending_.pop_front();
                continue;
            }
            Job& j = it->second;
            if (j.n_tiles == 0) {
                pending_.pop_front();
                memset(&j.summary, 0, sizeof j.summary);
                j.done = true;
                j.t_done = std::chrono::steady_clock::now();
                ++completed_;
                continue;
            }
            if (free_slots_.empty()) break;
            uint64_t off = 0;
            if (!pool_alloc(j.hits_cap, &off)) break;
            pending_.pop_front();
            j.slot = free_slots_.back();
            free_slots_.pop_back();
            j.hit_base = off;
            j.seq = next_seq_++;
            slot_ticket_[j.slot] = j.ticket;
            SrvSlot sd;
            memset(&sd, 0, sizeof sd);
            sd.req = j.req;
            sd.n_tiles = j.n_tiles;
            sd.hit_base = off;
            sd.hit_cap = j.hits_cap;
            sd.seq = j.seq;
            memcpy(&sh_.slots[j.slot], &sd, sizeof sd);
            j.next_local = 0;
            (j.n_tiles <= kSrvSmallTiles ? small_q_ : large_q_).push_back(j.ticket);
        }
        publish_locked();
    }

    // Completions posted by the device; advance first_live past finished entries.
    bool reap_locked() {
        bool any = false;
        for (uint32_t slot = 0; slot < cfg_.max_slots; ++slot) {
            const uint64_t tk = slot_ticket_[slot];
            if (!tk) continue;
            // a slot's ticket is cleared when its job completes, before the
            // job can be collected and erased
            auto it = jobs_.find(tk);
            if (it == jobs_.end()) continue;
            Job& j = it->second;
            if (j.done) continue;
            if (__atomic_load_n(&sh_.done[slot].seq, __ATOMIC_ACQUIRE) != j.seq) continue;
            memcpy(&j.summary, sh_.done[slot].words, sizeof(pp_summary));
            j.done = true;
            j.t_done = std::chrono::steady_clock::now();
            slot_ticket_[slot] = 0;
            free_slots_.push_back(slot);
            ++completed_;
            any = true;
        }
        while (first_live_ < n_entries_) {
            const uint64_t tk = entry_ticket_[first_live_ % kSrvRing];
            auto it = jobs_.find(tk);
            if (it != jobs_.end() && !it->second.done) break;
            ++first_live_;
        }
        if (any) __atomic_store_n(&sh_.ctl->first_live, first_live_, __ATOMIC_RELEASE);
        return any;
    }

    void worker_main() {
        const int trc = be.thread_init();
        std::unique_lock<std::mutex> lk(mu_);
        if (trc != PP_OK) {
            worker_error_ = pph::last_error();
            cv_.notify_all();
            return;
        }
        while (!shut_) {
            const bool any = reap_locked();
            admit_locked();
            if (any) cv_.notify_all();
            if (kernel_live_ && !be.running()) {
                kernel_live_ = false;
                const uint64_t self = __atomic_load_n(&sh_.ctl->self_stop, __ATOMIC_ACQUIRE);
                __atomic_store_n(&sh_.ctl->self_stop, 0ull, __ATOMIC_RELEASE);
                if (stopping_) {
                    stopping_ = false;
                    __atomic_store_n(&sh_.ctl->stop, 0ull, __ATOMIC_RELEASE);
                    admit_locked();
                } else if (self) {
                    // stopped by itself after waiting with nothing published:
                    // every tile below the reported one is done, none above it
                    // was started
                    resume_tile_ = self - 1;
                    claimed_seen_ = resume_tile_;
                    ++device_stops_;
                } else {
                    worker_error_ = "the persistent kernel exited without a stop";
                    cv_.notify_all();
                    break;
                }
            }
            // Busy while admitted requests are in flight.  Requests that wait
            // for a slot or pool room do not count: only collecting results
            // frees those, and a caller blocked on another kernel could not
            // collect while this kernel held the compute units.
            const bool busy = submitted_ - completed_ > pending_.size();
            const auto now = std::chrono::steady_clock::now();
            if (busy) idle_since_ = now;
            if (kernel_live_ && !stopping_ && !busy &&
                now - idle_since_ > std::chrono::microseconds(cfg_.idle_exit_us)) {
                // every published tile is done: the blocks all hold
                // unpublished tickets, and the next launch resumes from here
                stopping_ = true;
                resume_tile_ = published_;
                claimed_seen_ = published_;
                __atomic_store_n(&sh_.ctl->stop, 1ull, __ATOMIC_RELEASE);
            }
            if (busy && !kernel_live_ && published_ > resume_tile_) {
                if (be.launch(resume_tile_) != PP_OK) {
                    worker_error_ = pph::last_error();
                    cv_.notify_all();
                    break;
                }
                kernel_live_ = true;
                ++launches_;
            }
            if (busy) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            } else {
                cv_.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }

    pp_server_config cfg_;
    SrvShared sh_{};
    std::mutex mu_;
    std::condition_variable cv_;
    std::thread worker_;
    bool shut_ = false;
    bool kernel_live_ = false;
    bool stopping_ = false;       // stop raised, kernel not yet finished: admit nothing
    uint64_t resume_tile_ = 0;    // tiles processed when the kernel last stopped
    std::chrono::steady_clock::time_point idle_since_;
    uint64_t launches_ = 0, device_stops_ = 0;
    std::deque<uint64_t> pending_;
    std::deque<uint64_t> small_q_;             // admitted small jobs not yet published
    std::deque<uint64_t> large_q_;             // admitted large jobs with tiles left (FIFO)
    uint64_t claimed_seen_ = 0;                // highest claimed tile seen
    std::chrono::steady_clock::time_point claim_change_;
    std::map<uint64_t, Job> jobs_;
    std::set<uint64_t> collecting_;            // tickets whose hits are being copied out
    std::vector<uint32_t> free_slots_;
    std::vector<uint64_t> slot_ticket_;
    std::vector<uint64_t> entry_ticket_;       // ring: ticket of each live entry
    std::map<uint64_t, uint64_t> pool_free_;   // offset -> length (records)
    uint64_t pool_free_total_ = 0;
    uint64_t next_ticket_ = 1, next_seq_ = 1;
    uint64_t published_ = 0, n_entries_ = 0, first_live_ = 0;
    uint64_t submitted_ = 0, completed_ = 0;
    std::string worker_error_;
};

}  // namespace pps
