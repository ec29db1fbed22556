// pp_loadgen.cpp — native open-loop load generator for the serving loop.
//
// Drives a pp_server through the library's own C ABI: one thread submits
// request i at arrival time i (Poisson schedule), the calling thread collects
// the results in submission order.  It exists so that measurements of the
// server are not bounded by the caller's language runtime (a Python driver
// tops out near 60,000 small requests/s: one interpreter lock for the
// submitting and the collecting thread, and a sleep per request).  The same
// schedule can be served by one pp_propagate call per request instead
// (ppl_per_request_loop), so that both serving paths are measured from native
// code.
//
// The serving entry points come in as function pointers, taken from the
// library instance that created the server, so this file links against
// nothing of it and cannot bind to another copy of it.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "photonprop.h"

namespace {

typedef int (*submit_fn)(pp_server*, const pp_request*, uint64_t, uint64_t*);
typedef int (*wait_fn)(pp_server*, uint64_t, uint32_t, int*);
typedef int (*result_fn)(pp_server*, uint64_t, pp_summary*, pp_hit*, uint64_t, uint64_t*,
                         double*);
typedef int (*propagate_fn)(pp_ctx*, const pp_emitter*, const pp_spectrum*, const pp_sphere*,
                            uint64_t, uint32_t, uint32_t, pp_hit*, uint64_t, pp_summary*);

using clk = std::chrono::steady_clock;

double since(clk::time_point t0) {
    return std::chrono::duration<double>(clk::now() - t0).count();
}

}  // namespace

extern "C" {

// Status codes of the driver itself (the library's are passed through).
#define PPL_ERR_TIMEOUT (-100)
#define PPL_ERR_ARG (-101)

// Submit reqs[i] at arrivals_s[i] seconds after the start and collect every
// result.  Per request: latency_us[i] = the library's submit-to-completion
// latency + the submitting thread's lag behind the schedule (lag_us[i]);
// steps[i] = photon steps.  *makespan_s = start to the last result.  Returns
// 0, the first negative library status, or PPL_ERR_TIMEOUT if a request is
// not finished timeout_s after the start.
int ppl_open_loop(pp_server* srv, submit_fn submit, wait_fn wait, result_fn result,
                  const pp_request* reqs, const double* arrivals_s, uint64_t n, uint64_t hits_cap,
                  double timeout_s, double* latency_us, double* lag_us, uint64_t* steps,
                  double* makespan_s) {
    if (!srv || !submit || !wait || !result || (n && (!reqs || !arrivals_s || !latency_us ||
                                                      !lag_us || !steps)) || !makespan_s)
        return PPL_ERR_ARG;
    std::vector<uint64_t> tickets(n, 0);
    std::atomic<uint64_t> n_submitted{0};
    std::atomic<int> submit_rc{0};
    std::atomic<bool> abort{false};
    const clk::time_point t0 = clk::now();

    std::thread producer([&] {
        for (uint64_t i = 0; i < n && !abort.load(std::memory_order_relaxed); ++i) {
            // sleep while more than 200 µs early, then spin to the arrival
            for (;;) {
                const double early = arrivals_s[i] - since(t0);
                if (early <= 0.0) break;
                if (early > 200e-6)
                    std::this_thread::sleep_for(std::chrono::duration<double>(early - 150e-6));
            }
            lag_us[i] = 1e6 * std::max(0.0, since(t0) - arrivals_s[i]);
            const int rc = submit(srv, &reqs[i], hits_cap, &tickets[i]);
            if (rc < 0) {
                submit_rc.store(rc);
                break;
            }
            n_submitted.store(i + 1, std::memory_order_release);
        }
    });

    int rc = 0;
    std::vector<pp_hit> hits(hits_cap ? hits_cap : 0);
    for (uint64_t i = 0; i < n && rc == 0; ++i) {
        while (n_submitted.load(std::memory_order_acquire) <= i) {
            if (submit_rc.load()) {
                rc = submit_rc.load();
                break;
            }
            if (since(t0) > timeout_s) {
                rc = PPL_ERR_TIMEOUT;
                break;
            }
            std::this_thread::yield();
        }
        if (rc) break;
        int done = 0;
        while (!done) {
            const int w = wait(srv, tickets[i], 100, &done);
            if (w < 0) {
                rc = w;
                break;
            }
            if (!done && since(t0) > timeout_s) {
                rc = PPL_ERR_TIMEOUT;
                break;
            }
        }
        if (rc) break;
        pp_summary s;
        uint64_t n_hits = 0;
        double lat = 0.0;
        const int r = result(srv, tickets[i], &s, hits_cap ? hits.data() : nullptr, hits_cap,
                             &n_hits, &lat);
        if (r < 0 && r != PP_ERR_OVERFLOW) {
            rc = r;
            break;
        }
        latency_us[i] = lat + lag_us[i];
        steps[i] = s.n_steps;
    }
    abort.store(true);
    producer.join();
    *makespan_s = since(t0);
    return rc;
}

// The same schedule served by one propagate call (pp_propagate: kernels,
// reduction, summary copy and stream synchronisation) per request, first in
// first out on the calling thread: request i starts at max(arrivals_s[i], the
// end of request i - 1).  latency_us[i] = its completion - arrivals_s[i];
// steps[i] = photon steps; *makespan_s = start to the last completion.
// Returns 0, the first negative library status other than PP_ERR_OVERFLOW
// (no hits are kept, so every request with a hit overflows), or
// PPL_ERR_TIMEOUT once timeout_s has passed.
int ppl_per_request_loop(pp_ctx* ctx, propagate_fn propagate, const pp_spectrum* spectrum,
                         const pp_request* reqs, const double* arrivals_s, uint64_t n,
                         double timeout_s, double* latency_us, uint64_t* steps,
                         double* makespan_s) {
    if (!ctx || !propagate || !spectrum || (n && (!reqs || !arrivals_s || !latency_us || !steps)) ||
        !makespan_s)
        return PPL_ERR_ARG;
    int rc = 0;
    const clk::time_point t0 = clk::now();
    for (uint64_t i = 0; i < n; ++i) {
        for (;;) {
            const double early = arrivals_s[i] - since(t0);
            if (early <= 0.0) break;
            if (early > 200e-6)
                std::this_thread::sleep_for(std::chrono::duration<double>(early - 150e-6));
        }
        pp_summary s;
        const pp_request& r = reqs[i];
        const int p = propagate(ctx, &r.emitter, spectrum, &r.target, r.seed, r.stream,
                                r.max_steps, nullptr, 0, &s);
        if (p < 0 && p != PP_ERR_OVERFLOW) {
            rc = p;
            break;
        }
        const double now = since(t0);
        latency_us[i] = 1e6 * (now - arrivals_s[i]);
        steps[i] = s.n_steps;
        if (now > timeout_s) {
            rc = PPL_ERR_TIMEOUT;
            break;
        }
    }
    *makespan_s = since(t0);
    return rc;
}

}  // extern "C"
