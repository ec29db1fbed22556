// pp_kernels.hip — gfx950 kernels and the C ABI (include/photonprop.h).
//
// Design (one MI355X):
//   * Work unit = a tile of kTilePhotons consecutive photon ids of one request.
//     A persistent grid (resident blocks only) pulls tiles from one atomic
//     counter, so uneven batches balance themselves and a launch covers any
//     number of requests.
//   * One lane per photon.  A 256-lane block walks its tile in rounds of 256
//     photons; every lane runs its photon to the hit or to max_steps.
//   * Random numbers: Philox4x32-10 keyed by the request seed, counter
//     (block, photon_lo, photon_hi, stream).  Each block yields four 23-bit
//     uniforms and two 16-bit azimuths; block 0 holds the emission draws
//     {λ, θ, φ} and step 0, block b >= 1 holds steps 2b-1 and 2b.  Steps run
//     in pairs that consume one block, so the bits a step reads are fixed at
//     compile time.
//   * Medium tables {n, λ_sca, λ_abs} and the spectrum knots are read from
//     device memory once per photon (they stay in L2).
//   * Hits: each wave stages its records in LDS (ballot + mbcnt rank) and
//     appends a batch of up to 64 with one global atomic and coalesced 16-byte
//     stores.
//   * Sums: steps in per-lane counters and a fixed-order butterfly; the hits'
//     FP64 sums formed at each flush from the staged records by a fixed-order
//     wave reduction; the 4 waves combined in order into one partial per tile;
//     a second kernel adds a request's tile partials in tile order.  The
//     grouping depends only on the request, so a request's sums are bitwise
//     the same alone or inside any batch.
#include <hip/hip_runtime.h>
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/photonprop.h"
#include "pp_device.hpp"
#include "pp_host.hpp"
#include "pp_serve_core.hpp"

using namespace pph;

namespace {

constexpr int kBlock = 256;                       // lanes per block (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kRoundsPerTile = 32;            // photons per lane per tile
constexpr int kKeyRounds = 10;                     // Philox round keys per request (LDS)
static_assert((uint64_t)kBlock * kRoundsPerTile == kTilePhotons, "a tile is kRoundsPerTile rounds of a block");
constexpr double kCVacuum = 0.299792458;           // m/ns
constexpr float kLn2 = 0.69314718055994530942f;
constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kInv2Pi = 0.15915494309189533577f;


struct TilePartial {
    uint64_t n_steps;
    uint64_t n_hits;
    double sum_w, sum_wt, sum_wt2;
};

// The per-launch constants of run_tile (kernel-argument memory: run_tile
// reads them with scalar loads where it uses them).
struct TileCtx {
    const float4* table;  // medium table, two float4 per point (device memory; read once
                          // per photon, cached)
    const float* knots;   // spectrum knots (device memory)
    float xmax;           // n_points - 1
    float lmin, inv_dl;
    ppd::HG hg;
    SpecParams spec;
};

struct KParams {
    const DevReq* reqs;
    const uint64_t* tile_start;  // [n_req+1]
    uint32_t n_req;
    uint64_t n_tiles;
    uint32_t* tile_counter;
    TileCtx tile;
    pp_hit* hits;
    uint64_t hits_cap;
    unsigned long long* hit_cursor;
    TilePartial* partials;
};

// ---------------------------------------------------------------------------
// Propagation kernel
// ---------------------------------------------------------------------------

// Element i of data that no kernel writes during the launch, read through the
// constant address space: a uniform address then loads with scalar loads into
// SGPRs (through the scalar cache) instead of vector loads into VGPRs.  (The
// host pass of this file only parses device code, and has no address space 4.)
template <typename T>
__device__ __forceinline__ T load_constant(const T* p, uint64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T*)p)[i];
#else
    return p[i];
#endif
}

// The same for a pointer the compiler may hold in vector registers (the value
// is wave-uniform; readfirstlane moves it to scalar registers first).
template <typename T>
__device__ __forceinline__ const T* opaque_uniform_ptr(const T* p) {
    const uint64_t v = (uint64_t)p;
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    __asm__ volatile("" : "+s"(lo), "+s"(hi));
    return (const T*)(((uint64_t)hi << 32) | lo);
}

// Pointer p, opaque to the compiler at this point: a load through it is not
// hoisted or merged with another, so the loaded value lives only from here.
template <typename T>
__device__ __forceinline__ const T* opaque_ptr(const T* p) {
    __asm__ volatile("" : "+s"(p));
    return p;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Hit records leave a wave through an LDS buffer: a wave stages up to kHitBuf
// records and takes their output slots with one atomicAdd on the hit cursor
// when the next round's hits would not fit (or the tile ends).  One atomic per
// round with a hit, all on the same address, stalled the waves on its return:
// the launch ran 74 % longer than with no allocation at all (1,869 against
// 1,074 ms); 8, 16, 32 and 64 records per wave gave 1,487, 1,257, 1,160 and
// 1,135 ms.  64 records (16 KiB per block) also always take a whole round.
constexpr uint32_t kHitBuf = 64;
static_assert(kHitBuf >= 64, "a round's hits (up to 64) must fit an empty buffer");
static_assert((1ull << 32) % kTilePhotons == 0, "a tile never straddles photon id 2^32");

// 16 B of a record: plain, or (serving) as system-scope 32-bit stores that go
// straight to memory, where the host copies records while the kernel runs.
template <bool kSysStores>
__device__ __forceinline__ void store_quad(uint4* dst, const uint4& v) {
    if (kSysStores) {
        uint32_t* o = reinterpret_cast<uint32_t*>(dst);
        __hip_atomic_store(o + 0, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(o + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(o + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(o + 3, v.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
        *dst = v;
    }
}

// The wave's n staged records (n <= kHitBuf) to slots base .. base + n - 1
// of `hits`, base from one atomicAdd; slots at or beyond hits_cap are dropped
// (the cursor keeps counting, so overflow stays exact).  The lanes copy 16 B
// at a time.  The records' FP64 sums Σw, Σw·t, Σw·t² (record r on lane r, a
// fixed-order wave sum) and their count are added to the wave's row
// wsum[0..2], *whits, whether stored or not.  Called by the whole wave
// (uniform control flow).
template <bool kSysStores>
__device__ __forceinline__ void flush_hits(const uint4* buf, uint32_t n, pp_hit* hits,
                                           uint64_t hits_cap, unsigned long long* cursor,
                                           uint32_t lane, double* wsum, uint64_t* whits) {
    __asm__ volatile("" : "+v"(lane));
    double w = 0.0, wt = 0.0, wt2 = 0.0;
    if (lane < n) {
        const double t = (double)__uint_as_float(buf[4 * lane].w);
        w = (double)__uint_as_float(buf[4 * lane + 2].w);
        wt = w * t;
        wt2 = wt * t;
    }
    w = wave_sum_f64(w);
    wt = wave_sum_f64(wt);
    wt2 = wave_sum_f64(wt2);
    if (lane == 0) {
        wsum[0] += w;
        wsum[1] += wt;
        wsum[2] += wt2;
        *whits += n;
    }
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(cursor, (unsigned long long)n);
    base = __shfl(base, 0, 64);
    for (uint32_t c = lane; c < 4 * n; c += 64) {
        const unsigned long long slot = base + (c >> 2);
        if (slot < hits_cap)
            store_quad<kSysStores>(reinterpret_cast<uint4*>(hits + slot) + (c & 3), buf[c]);
    }
}


// Photons [first, first + count) of request R on one 256-lane block: emit,
// step, append hits (record field `request` = req_index), and reduce the
// tile's counts and FP64 sums in a fixed order.  The result is valid in
// thread 0 only.  Shared by the batch kernel and the serving kernel, so a
// request's tile partials are the same whichever kernel ran them.  kIso
// (= T.hg.iso, uniform per launch) takes the isotropic-scattering test out of
// the step loop; run_tile_any picks the instance.
template <bool kSysStores, bool kIso>
__device__ __forceinline__ TilePartial run_tile(const TileCtx* Tc, const DevReq& R,
                                                uint32_t req_index, uint64_t first,
                                                uint32_t count, pp_hit* hits, uint64_t hits_cap,
                                                unsigned long long* hit_cursor,
                                                double (*s_red)[3], uint64_t (*s_redi)[2],
                                                uint4 (*s_hitbuf)[kHitBuf * 4],
                                                float (*s_idir)[kBlock],
                                                uint2 (*s_keys)[kKeyRounds]) {
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    uint4* const hbuf = s_hitbuf[wave];
    // the request's Philox key schedule, one copy per wave
    const uint2* const rk = s_keys[wave];
    ppd::philox_keys_to_lds(s_keys[wave], lane, R.seed_lo, R.seed_hi);
    uint32_t fill = 0;  // records staged in hbuf (wave-uniform)
    // The wave's row of s_red / s_redi is addressed through a wave index taken
    // afresh at each use (an empty asm stops the compiler from computing the
    // row addresses once and holding them, spilled, through the step loop;
    // the uses are per flush and per tile).
    auto row_wave = [tid]() {
        uint32_t w = tid >> 6;
        __asm__ volatile("" : "+v"(w));
        return w;
    };
    // The hits' count and FP64 sums accumulate per wave in s_red / s_redi
    // (lane 0 owns its wave's row; the caller's barrier before the tile has
    // released the previous tile's rows), added at each flush.
    if (lane == 0) {
        const uint32_t rw = row_wave();
        s_red[rw][0] = 0.0;
        s_red[rw][1] = 0.0;
        s_red[rw][2] = 0.0;
        s_redi[rw][1] = 0ull;
    }
    const uint32_t rounds = (count + kBlock - 1) / kBlock;
    const uint32_t max_steps = R.max_steps;
    const float r2 = R.r2;
    const float cx = R.center[0], cy = R.center[1], cz = R.center[2];
    const ppd::HG hg = load_constant(&Tc->hg, 0);  // the step loop's constants

    // steps of this lane's photons: at most kRoundsPerTile * max_steps <= 2^29
    // (max_steps <= 2^24); the wave's sum is taken in 64 bits
    uint32_t acc_steps = 0;

    for (uint32_t k = 0; k < rounds; ++k) {
        const uint32_t local = k * kBlock + tid;
        const bool active = local < count;
        // a tile's photons share the high word of their id: `first` is a
        // multiple of kTilePhotons, which divides 2^32, and local < kTilePhotons
        const uint32_t id_lo = (uint32_t)first + local, id_hi = (uint32_t)(first >> 32);

        // ---- emission (src/photon_prop_cuda.jl:308-316) ----
        // Draws (ppd::draw_a, ppd::draw_p) per Philox block (x, y, z, w):
        //   block 0:      λ <- a(x), emission θ <- a(y), emission φ <- p(x, y);
        //                 step 0 <- {a(z), a(w), p(z, w)}
        //   block b >= 1: step 2b-1 <- {a(x), a(y), p(x, y)},
        //                 step 2b   <- {a(z), a(w), p(z, w)}
        // with {step length, scattering angle, azimuth} per step.  The final
        // step (max_steps - 1) needs only its length: when it is the first
        // step of its block and the emission is the Cherenkov cone (which
        // leaves a(y) of block 0 unused), its length is a(y) of block 0 and
        // its block is never computed.
        const ppd::U4 w0 = ppd::philox10(0u, id_lo, id_hi, R.stream, R.seed_lo, R.seed_hi, rk);
        const bool cone = R.profile == PP_EMIT_CHERENKOV_CONE;
        const float u_l = 2.0f - ppd::draw_a(w0.x);
        float lambda;
        // the launch constants of the emission, loaded field by field for this
        // round only (scalar loads; no registers held for them through the step
        // loop; a whole-struct copy here became vector loads)
        const TileCtx* const Tu = opaque_uniform_ptr(Tc);
        const uint32_t spec_kind = load_constant(&Tu->spec.kind, 0);
        if (spec_kind == PP_SPEC_MONO) {
            lambda = load_constant(&Tu->spec.mono, 0);
        } else if (spec_kind == PP_SPEC_KNOTS) {
            lambda = ppd::knot_interp(load_constant(&Tu->knots, 0), load_constant(&Tu->spec.n_knots, 0),
                                      load_constant(&Tu->spec.knot_step, 0), u_l);
        } else {
            // hardware reciprocal (1 ulp) rather than the correctly rounded
            // division sequence, here and for the cone angle below
            lambda = __builtin_amdgcn_rcpf(
                __builtin_fmaf(-u_l, load_constant(&Tu->spec.d, 0), load_constant(&Tu->spec.a, 0)));
        }
        const float3 med = ppd::table_lerp_slopes(
            load_constant(&Tu->table, 0), load_constant(&Tu->xmax, 0),
            (lambda - load_constant(&Tu->lmin, 0)) * load_constant(&Tu->inv_dl, 0));
        const float n_phase = med.x;
        const float sca_ln2 = med.y;  // the table holds λ_sca·ln2
        const float abs_len = med.z;

        float dx, dy, dz;
        {
            const float u_phi = ppd::draw_p(w0.x, w0.y);
            const float sp = __builtin_amdgcn_sinf(u_phi);
            const float cp = __builtin_amdgcn_cosf(u_phi);
            const float u_th = 2.0f - ppd::draw_a(w0.y);
            if (R.profile == PP_EMIT_REF_ISOTROPIC) {
                const float theta = __builtin_fmaf(u_th, 2.0f, -1.0f);
                const float st = __builtin_amdgcn_sinf(theta * kInv2Pi);
                const float ct = __builtin_amdgcn_cosf(theta * kInv2Pi);
                dx = cp * st;
                dy = __builtin_fmaf(sp, st, -0x1p-60f);  // the step loop's y bias
                dz = ct;
            } else {
                const float ct =
                    R.profile == PP_EMIT_ISOTROPIC
                        ? __builtin_fmaf(2.0f, u_th, -1.0f)
                        : fminf(__builtin_amdgcn_rcpf(n_phase * R.beta), 1.0f);
                const float st = __builtin_amdgcn_sqrtf(__builtin_fmaf(-ct, ct, 1.0f));
                // not renormalised here: step 0's rotation renormalises; y
                // carries the step loop's bias
                ppd::rotate_polar<false, false, true>(R.axis[0], R.axis[1], R.axis[2], cp, sp, st,
                                                      ct, dx, dy, dz);
            }
        }
        // (dy carries a bias of -2^-60, which the rotations need at the exact
        // south pole; it leaves every |y| above 2^-36 unchanged)
        // the emission direction waits in LDS for the hit record (registers
        // stay free for the step loop)
        s_idir[0][tid] = dx;
        s_idir[1][tid] = dy;
        s_idir[2][tid] = dz;
        // position relative to the target centre: the sphere test needs no
        // per-step subtraction; the absolute position is rebuilt for a hit
        float qx = R.pos[0] - cx, qy = R.pos[1] - cy, qz = R.pos[2] - cz;
        float dist = 0.0f;
        bool alive = active;
        bool hit = false;
        // steps of this photon: max_steps unless it hits at step k (k + 1)
        uint32_t nst = 0u;

        // One scattering-loop iteration (src/photon_prop_cuda.jl:320-376, D3).
        // LAST (wave-uniform): the photon's final allowed step.  A photon that
        // does not hit there ends with it, so its new direction would never be
        // used and the scatter is skipped; results are unchanged.
#define PP_STEP(K, AS, AC, PH, RN, LAST)                                                 \
    if (alive) {                                                                        \
        ++nst;                                                                          \
        float s = -__builtin_amdgcn_logf(2.0f - (AS)) * sca_ln2;                        \
        const float a = __builtin_fmaf(dz, qz, __builtin_fmaf(dy, qy, dx * qx));        \
        /* |q|² - r² in one chain (r² folded into the first fma) */                     \
        const float pmr = __builtin_fmaf(qz, qz, __builtin_fmaf(qy, qy, __builtin_fmaf(qx, qx, -r2))); \
        const float b = __builtin_fmaf(a, a, -pmr);                                     \
        bool h = false;                                                                 \
        if (b >= 0.0f) {                                                                \
            const float d = -a - __builtin_amdgcn_sqrtf(b);                             \
            if (d > 0.0f && d < s) {                                                    \
                h = true;                                                               \
                s = d;                                                                  \
            }                                                                           \
        }                                                                               \
        qx = __builtin_fmaf(s, dx, qx);                                                 \
        qy = __builtin_fmaf(s, dy, qy);                                                 \
        qz = __builtin_fmaf(s, dz, qz);                                                 \
        dist += s;                                                                      \
        __asm__ volatile("" : "+v"(qx), "+v"(qy), "+v"(qz), "+v"(dist));               \
        if (h) {                                                                        \
            hit = true;                                                                 \
            alive = false;                                                              \
        } else if (!(LAST)) {                                                           \
            const float ct = ppd::hg_costheta_a_t<kIso>(hg, (AC));                      \
            const float st = ppd::sin_from_cos(ct);                                     \
            const float sph = __builtin_amdgcn_sinf(PH);                                \
            const float cph = __builtin_amdgcn_cosf(PH);                                \
            ppd::rotate_polar<false, true, true>(dx, dy, dz, cph, sph, st, ct, dx, dy, dz);       \
            if (RN) ppd::renormalize(dx, dy, dz);                                       \
        }                                                                               \
    }

        // No per-step ballot: a wave whose lanes have all hit skips the
        // (divergent) step bodies and Philox blocks on an empty exec mask.
        // The direction is renormalised after every eighth rotation, counting
        // the emission as the first (after steps 6, 14, 22, ...): a unit
        // vector's length moves by ~2e-7 per rotation, so it stays within a
        // few 1e-6 of 1, where one Newton step renormalises exactly.  Two
        // Philox blocks per iteration, one per pair of steps.
        const uint32_t w_lone = w0.y;  // the lone final step's length word (cone)
        PP_STEP(0u, ppd::draw_a(w0.z), ppd::draw_a(w0.w), ppd::draw_p(w0.z, w0.w), false,
                max_steps == 1u)
        // Pairs of steps s0, s0 + 1, one Philox block each; the first of a pair
        // is never the final step.  A final step that opens a block (max_steps
        // even) comes after the loop: it needs only its length, which under
        // cone emission is the unused word of block 0 (no block computed).
        uint32_t s0 = 1;
        for (; s0 + 1 < max_steps; s0 += 2) {
            ppd::U4 Q;
            if (alive)
                Q = ppd::philox10((s0 + 1u) >> 1, id_lo, id_hi, R.stream, R.seed_lo, R.seed_hi,
                                  rk);
            PP_STEP(s0, ppd::draw_a(Q.x), ppd::draw_a(Q.y), ppd::draw_p(Q.x, Q.y), false, false)
            PP_STEP(s0 + 1u, ppd::draw_a(Q.z), ppd::draw_a(Q.w), ppd::draw_p(Q.z, Q.w),
                    (s0 & 7u) == 5u, s0 + 2 >= max_steps)
        }
        if (s0 < max_steps) {
            uint32_t wl = w_lone;
            if (!cone && alive)
                wl = ppd::philox10((s0 + 1u) >> 1, id_lo, id_hi, R.stream, R.seed_lo, R.seed_hi,
                                   rk).x;
            PP_STEP(s0, ppd::draw_a(wl), 1.0f, 1.0f, false, true)
        }
#undef PP_STEP
        acc_steps += nst;

        // ---- hit append: staged per wave, one slot allocation per flush ----
        const uint64_t mask = __ballot(hit);
        if (mask != 0) {
            const uint32_t h = (uint32_t)__popcll(mask);
            if (fill + h > kHitBuf) {
                const uint32_t rw = row_wave();
                flush_hits<kSysStores>(hbuf, fill, hits, hits_cap, hit_cursor, lane, s_red[rw],
                                       &s_redi[rw][1]);
                fill = 0;
            }
            if (hit) {
                const float px = qx + cx, py = qy + cy, pz = qz + cz;
                const float e = -dist / abs_len;
                const float wgt = __builtin_amdgcn_exp2f(e * kLog2e);
                const float t = R.t0 + dist * n_phase / (float)kCVacuum;
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                const uint4 q0 = make_uint4(__float_as_uint(px), __float_as_uint(py),
                                            __float_as_uint(pz), __float_as_uint(t));
                const uint4 q1 = make_uint4(__float_as_uint(dx), __float_as_uint(dy),
                                            __float_as_uint(dz), __float_as_uint(lambda));
                const uint4 q2 = make_uint4(__float_as_uint(s_idir[0][tid]),
                                            __float_as_uint(s_idir[1][tid]),
                                            __float_as_uint(s_idir[2][tid]), __float_as_uint(wgt));
                const uint4 q3 = make_uint4(__float_as_uint(dist), req_index, id_lo, id_hi);
                uint4* o = hbuf + (fill + rank) * 4;
                o[0] = q0;
                o[1] = q1;
                o[2] = q2;
                o[3] = q3;
            }
            fill += h;
        }
    }
    if (fill) {
        const uint32_t rw = row_wave();
        flush_hits<kSysStores>(hbuf, fill, hits, hits_cap, hit_cursor, lane, s_red[rw],
                               &s_redi[rw][1]);
    }

    // ---- tile partial: fixed-order reduction ----
    const unsigned long long ws = wave_sum_u64(acc_steps);
    if (lane == 0) s_redi[row_wave()][0] = ws;
    __syncthreads();
    TilePartial tp;
    tp.n_steps = 0;
    tp.n_hits = 0;
    tp.sum_w = 0.0;
    tp.sum_wt = 0.0;
    tp.sum_wt2 = 0.0;
    if (tid == 0) {
        for (int q = 0; q < kWaves; ++q) {
            tp.n_steps += s_redi[q][0];
            tp.n_hits += s_redi[q][1];
            tp.sum_w += s_red[q][0];
            tp.sum_wt += s_red[q][1];
            tp.sum_wt2 += s_red[q][2];
        }
    }
    return tp;
}

template <bool kSysStores>
__device__ __forceinline__ TilePartial run_tile_any(const TileCtx* Tc, const DevReq& R,
                                                    uint32_t req_index, uint64_t first,
                                                    uint32_t count, pp_hit* hits,
                                                    uint64_t hits_cap,
                                                    unsigned long long* hit_cursor,
                                                    double (*s_red)[3], uint64_t (*s_redi)[2],
                                                    uint4 (*s_hitbuf)[kHitBuf * 4],
                                                    float (*s_idir)[kBlock],
                                                    uint2 (*s_keys)[kKeyRounds]) {
    if (load_constant(&Tc->hg, 0).iso)
        return run_tile<kSysStores, true>(Tc, R, req_index, first, count, hits, hits_cap,
                                          hit_cursor, s_red, s_redi, s_hitbuf, s_idir, s_keys);
    return run_tile<kSysStores, false>(Tc, R, req_index, first, count, hits, hits_cap,
                                       hit_cursor, s_red, s_redi, s_hitbuf, s_idir, s_keys);
}


// P.tile where the kernel arguments lie (P is the kernel's only argument, at
// offset 0): taking &P instead would make the compiler copy P to private memory.
__device__ __forceinline__ const TileCtx* kernarg_tile() {
#if defined(__HIP_DEVICE_COMPILE__)
    return (const TileCtx*)((const char*)__builtin_amdgcn_kernarg_segment_ptr() +
                            offsetof(KParams, tile));
#else
    return nullptr;  // host pass: device code is only parsed
#endif
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) propagate_kernel(KParams P) {
    __shared__ uint32_t s_tile;
    __shared__ double s_red[kWaves][3];
    __shared__ uint64_t s_redi[kWaves][2];
    __shared__ uint4 s_hitbuf[kWaves][kHitBuf * 4];
    __shared__ float s_idir[3][kBlock];
    __shared__ uint2 s_keys[kWaves][kKeyRounds];
    for (;;) {
        __syncthreads();  // previous tile's s_tile and s_red consumed
        if (threadIdx.x == 0) s_tile = atomicAdd(P.tile_counter, 1u);
        __syncthreads();
        const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
        if ((uint64_t)tile >= P.n_tiles) break;

        // Request owning this tile: last r with tile_start[r] <= tile.
        uint32_t lo = 0, hi = P.n_req;  // invariant: tile_start[lo] <= tile < tile_start[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (load_constant(P.tile_start, mid) <= (uint64_t)tile) lo = mid; else hi = mid;
        }
        const uint32_t r = __builtin_amdgcn_readfirstlane(lo);
        // through the constant address space: the request is not written
        // during the launch, so it is read with scalar loads into SGPRs
        const DevReq R = load_constant(P.reqs, r);
        const uint64_t first = ((uint64_t)tile - load_constant(P.tile_start, r)) * kTilePhotons;
        const uint64_t left = R.n_photons - first;
        const uint32_t count = (uint32_t)(left < kTilePhotons ? left : kTilePhotons);
        const TilePartial tp = run_tile_any<false>(kernarg_tile(), R, r, first, count, P.hits, P.hits_cap,
                                                   P.hit_cursor, s_red, s_redi, s_hitbuf, s_idir, s_keys);
        if (threadIdx.x == 0) P.partials[tile] = tp;
    }
}

// Sum of a request's tile partials in tile order: lane l adds tiles l, l+64, ...
// then a fixed butterfly.  Valid in every lane of the calling wave.
template <bool kSysLoads>
__device__ __forceinline__ pp_summary reduce_partials(const TilePartial* parts, uint64_t n_tiles,
                                                      uint64_t n_sim) {
    // (an empty asm: the per-lane offsets are formed here, not hoisted into a
    // persistent caller's registers)
    uint32_t lane = threadIdx.x & 63u;
    __asm__ volatile("" : "+v"(lane));
    uint64_t st = 0, nh = 0;
    double sw = 0.0, swt = 0.0, swt2 = 0.0;
    for (uint64_t t = lane; t < n_tiles; t += 64) {
        TilePartial p;
        if (kSysLoads) {  // written by blocks on other XCDs with system-scope stores
            const uint64_t* q = reinterpret_cast<const uint64_t*>(parts + t);
            uint64_t w[5];
#pragma unroll
            for (int k = 0; k < 5; ++k)
                w[k] = __hip_atomic_load(q + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            p.n_steps = w[0];
            p.n_hits = w[1];
            p.sum_w = __longlong_as_double((long long)w[2]);
            p.sum_wt = __longlong_as_double((long long)w[3]);
            p.sum_wt2 = __longlong_as_double((long long)w[4]);
        } else {
            p = parts[t];
        }
        st += p.n_steps;
        nh += p.n_hits;
        sw += p.sum_w;
        swt += p.sum_wt;
        swt2 += p.sum_wt2;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        st += __shfl_xor(st, o, 64);
        nh += __shfl_xor(nh, o, 64);
        sw += __shfl_xor(sw, o, 64);
        swt += __shfl_xor(swt, o, 64);
        swt2 += __shfl_xor(swt2, o, 64);
    }
    pp_summary s;
    s.n_sim = n_sim;
    s.n_steps = st;
    s.n_hits = nh;
    s.n_stored = 0;
    s.sum_w = sw;
    s.sum_wt = swt;
    s.sum_wt2 = swt2;
    s.status = PP_OK;
    s.reserved = 0;
    return s;
}

// Per-request sums over its tiles in tile order (one wave per request).
__global__ void __launch_bounds__(64) reduce_kernel(const TilePartial* partials,
                                                    const uint64_t* tile_start,
                                                    const DevReq* reqs, uint32_t n_req,
                                                    pp_summary* out) {
    const uint32_t r = blockIdx.x;
    if (r >= n_req) return;
    const uint64_t b = tile_start[r], e = tile_start[r + 1];
    const pp_summary s = reduce_partials<false>(partials + b, e - b, reqs[r].n_photons);
    if (threadIdx.x == 0) out[r] = s;
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Hit compaction.  Staged records [0, min(cursor, cap)) are split into
// gridDim.x contiguous chunks, the same in both passes.  Per chunk, an LDS
// histogram counts records per request, so a block issues one global atomic
// per request it saw instead of one per record (the staged order follows the
// tiles, so a chunk mostly holds one or two requests).  Requests beyond
// kLdsHistMax fall back to per-record global atomics.  The count pass also
// writes each record's request to a 4-byte array, so the scatter's own
// histogram pass reads 4 bytes per record instead of a 64-byte record line.
constexpr uint32_t kLdsHistMax = 4096;

__device__ __forceinline__ uint32_t staged_request(const pp_hit* staged, uint64_t i) {
    return reinterpret_cast<const uint32_t*>(staged + i)[13];
}

__global__ void __launch_bounds__(256) count_kernel(const pp_hit* staged,
                                                    const unsigned long long* cursor,
                                                    uint64_t cap, uint32_t n_req, uint32_t* cnt,
                                                    uint32_t* rec_req) {
    __shared__ uint32_t hist[kLdsHistMax];
    const uint64_t n = *cursor < cap ? *cursor : cap;
    const bool lds = n_req <= kLdsHistMax;
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b = blockIdx.x * chunk;
    const uint64_t e = b + chunk < n ? b + chunk : n;
    if (lds)
        for (uint32_t r = threadIdx.x; r < n_req; r += blockDim.x) hist[r] = 0;
    __syncthreads();
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const uint32_t r = staged_request(staged, i);
        rec_req[i] = r;
        if (lds) atomicAdd(&hist[r], 1u);
        else atomicAdd(&cnt[r], 1u);
    }
    __syncthreads();
    if (lds)
        for (uint32_t r = threadIdx.x; r < n_req; r += blockDim.x)
            if (hist[r]) atomicAdd(&cnt[r], hist[r]);
}

// Exclusive scan of cnt into offsets[0..n], single block of 1024 threads;
// also records n_stored and the overflow status per request.
__device__ __forceinline__ uint64_t kept(uint32_t c, uint64_t per_cap) {
    return per_cap && (uint64_t)c > per_cap ? per_cap : (uint64_t)c;
}

__global__ void __launch_bounds__(1024) scan_kernel(const uint32_t* cnt, uint32_t n,
                                                    uint64_t per_cap, uint64_t* offsets,
                                                    pp_summary* sums) {
    __shared__ uint64_t s_part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t b = t * per;
    const uint32_t e = std::min(n, b + per);
    uint64_t local = 0;
    for (uint32_t i = b; i < e; ++i) local += kept(cnt[i], per_cap);
    s_part[t] = local;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint64_t v = t >= o ? s_part[t - o] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    uint64_t run = s_part[t] - local;
    for (uint32_t i = b; i < e; ++i) {
        offsets[i] = run;
        run += kept(cnt[i], per_cap);
        sums[i].n_stored = kept(cnt[i], per_cap);
        // overflow = records lost to the launch-wide capacity, not to per_cap
        sums[i].status = sums[i].n_hits > (uint64_t)cnt[i] ? PP_ERR_OVERFLOW : PP_OK;
    }
    if (t == 1023) offsets[n] = s_part[1023];
}

__global__ void __launch_bounds__(256) scatter_kernel(const pp_hit* staged,
                                                      const uint32_t* rec_req,
                                                      const unsigned long long* cursor,
                                                      uint64_t cap, uint32_t n_req,
                                                      uint64_t per_cap, const uint64_t* offsets,
                                                      uint32_t* fill, pp_hit* out) {
    __shared__ uint32_t hist[kLdsHistMax];
    __shared__ uint64_t base[kLdsHistMax];
    const uint64_t n = *cursor < cap ? *cursor : cap;
    const bool lds = n_req <= kLdsHistMax;
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b = blockIdx.x * chunk;
    const uint64_t e = b + chunk < n ? b + chunk : n;
    if (lds) {
        for (uint32_t r = threadIdx.x; r < n_req; r += blockDim.x) hist[r] = 0;
        __syncthreads();
        for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) atomicAdd(&hist[rec_req[i]], 1u);
        __syncthreads();
        for (uint32_t r = threadIdx.x; r < n_req; r += blockDim.x) {
            const uint32_t h = hist[r];
            base[r] = h ? atomicAdd(&fill[r], h) : 0;  // rank of this chunk's first record
            hist[r] = 0;
        }
        __syncthreads();
    }
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(staged + i);
        const uint4 a = src[0], bb = src[1], c = src[2], d = src[3];
        const uint32_t r = d.y;
        const uint64_t rank = lds ? base[r] + atomicAdd(&hist[r], 1u) : atomicAdd(&fill[r], 1u);
        if (per_cap && rank >= per_cap) continue;
        uint4* dst = reinterpret_cast<uint4*>(out + offsets[r] + rank);
        dst[0] = a;
        dst[1] = bb;
        dst[2] = c;
        dst[3] = d;
    }
}

// ---- self-test kernels (device helpers over arrays) ----
__global__ void selftest_rotate_kernel(const float* od, const float* nd, float* out, uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x, y, z;
    ppd::rotate_to_axis(od[3 * i], od[3 * i + 1], od[3 * i + 2], nd[3 * i], nd[3 * i + 1],
                        nd[3 * i + 2], x, y, z);
    out[3 * i] = x;
    out[3 * i + 1] = y;
    out[3 * i + 2] = z;
}

__global__ void selftest_philox_kernel(const uint32_t* ctr, uint32_t k0, uint32_t k1,
                                       uint32_t* out, uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ppd::U4 o = ppd::philox10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], k0, k1);
    out[4 * i] = o.x;
    out[4 * i + 1] = o.y;
    out[4 * i + 2] = o.z;
    out[4 * i + 3] = o.w;
}

// Word i of (seed, stream 0): block i/4 of counter (i/4, 0, 0, 0), word i%4.
__device__ __forceinline__ uint32_t hg_word(uint64_t i, uint32_t k0, uint32_t k1) {
    const ppd::U4 o = ppd::philox10((uint32_t)(i >> 2), (uint32_t)(i >> 34), 0u, 0u, k0, k1);
    const uint32_t j = (uint32_t)(i & 3u);
    return j == 0 ? o.x : j == 1 ? o.y : j == 2 ? o.z : o.w;
}

__global__ void selftest_hg_kernel(ppd::HG hg, uint32_t k0, uint32_t k1, float* out, uint64_t n) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = ppd::hg_costheta_a(hg, ppd::draw_a(hg_word(i, k0, k1)));
}

__global__ void __launch_bounds__(256) selftest_hg_moments_kernel(ppd::HG hg, uint32_t k0,
                                                                  uint32_t k1, uint64_t n,
                                                                  double* partial) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const double c = (double)ppd::hg_costheta_a(hg, ppd::draw_a(hg_word(i, k0, k1)));
        a += c;
        b += c * c;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s1[0];
        partial[2 * blockIdx.x + 1] = s2[0];
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

}  // namespace


namespace {


// The per-launch constants of run_tile, as the kernels' arguments carry them.
TileCtx make_tile_ctx(const float4* table, uint32_t n_points, const float* knots,
                      const SpecParams& spec, float lmin, float inv_dl, const ppd::HG& hg) {
    TileCtx T;
    T.table = table;
    T.knots = knots;
    T.xmax = (float)(n_points - 1u);
    T.lmin = lmin;
    T.inv_dl = inv_dl;
    T.hg = hg;
    T.spec = spec;
    return T;
}

int upload_medium(pp_ctx* c, const pp_medium_tables* m) {
    int rc = check_medium(m);
    if (rc) return rc;
    // point i: {n, λ_sca·ln2, λ_abs, 0}, then the FP32 differences to point
    // i+1 (zero at the last point), as the kernel's table_lerp_slopes reads it
    const uint32_t n = m->n_points;
    std::vector<float> sca(n);
    for (uint32_t i = 0; i < n; ++i) sca[i] = m->sca_len[i] * kLn2;
    std::vector<float4> t(2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t j = i + 1 < n ? i + 1 : i;
        t[2 * i] = make_float4(m->n_phase[i], sca[i], m->abs_len[i], 0.0f);
        t[2 * i + 1] = make_float4(m->n_phase[j] - m->n_phase[i], sca[j] - sca[i],
                                   m->abs_len[j] - m->abs_len[i], 0.0f);
    }
    if ((rc = ensure(c->table, t.size() * sizeof(float4)))) return rc;
    PP_HIP(hipMemcpy(c->table.p, t.data(), t.size() * sizeof(float4), hipMemcpyHostToDevice));
    c->n_points = m->n_points;
    c->lmin = m->lambda_min;
    c->lmax = m->lambda_max;
    c->inv_dl = (float)((double)(m->n_points - 1) / ((double)m->lambda_max - (double)m->lambda_min));
    c->g = m->g;
    c->n_min = *std::min_element(m->n_phase, m->n_phase + m->n_points);
    return PP_OK;
}


// A checked spectrum's knots to d_knots (PP_MAX_KNOTS floats) on `stream`;
// nothing for the other kinds.
int upload_knots(const pp_spectrum* s, float* d_knots, hipStream_t stream) {
    if (s->kind != PP_SPEC_KNOTS) return PP_OK;
    PP_HIP(hipMemcpyAsync(d_knots, s->knots, s->n_knots * sizeof(float), hipMemcpyHostToDevice,
                          stream));
    return PP_OK;
}




// Henyey-Greenstein constants for the device: 1-g², 1+g², 1/(2g).
ppd::HG make_hg(float g) {
    ppd::HG h;
    h.g = g;
    h.iso = g == 0.0f;
    const double gd = g, num = 1.0 - gd * gd;
    h.c = h.iso ? 0.0f : (float)((1.0 + gd * gd) / (2.0 * gd));
    h.k = h.iso ? 0.0f : (float)(num * num / (2.0 * gd));
    return h;
}

int grid_for(pp_ctx* c, uint64_t n_tiles, int* grid) {
    if (!c->occ_known) {
        int per_cu = 0;
        PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, propagate_kernel, kBlock, 0));
        c->blocks_per_cu = per_cu < 1 ? 1 : per_cu;
        c->occ_known = true;
    }
    const uint64_t resident = (uint64_t)c->blocks_per_cu * (uint64_t)c->n_cu;
    *grid = (int)std::min<uint64_t>(n_tiles, resident);
    return PP_OK;
}

// Pinned I/O buffer of a context: one request's upload at 0, its summary and
// hit cursor read back at kIoDownload.
constexpr size_t kIoBytes = 4096;
constexpr size_t kIoDownload = 2048;

int ensure_io(pp_ctx* c) {
    if (!c->io) PP_HIP(hipHostMalloc(&c->io, kIoBytes, 0));
    return PP_OK;
}

// Device work buffers of one launch.  The allocating entry points take them
// from the context's grow-only buffers (ctx_work), the allocation-free one
// from the fixed per-device scratch array (scratch_work).
struct WorkBufs {
    DevReq* reqs;           // [n_req]
    uint64_t* tile_start;   // [n_req + 1]
    TilePartial* partials;  // [n_tiles]
    uint32_t* counters;     // [16]: [0] tile counter, [2..3] hit cursor
    pp_summary* sums;       // [n_req]
    float* knots;           // [PP_MAX_KNOTS] (knot spectra)
    pp_hit* staging;        // [hits_cap] hit records as the kernel appends them
    uint32_t* cnt;          // [n_req] compaction: records per request
    uint32_t* fill;         // [n_req] compaction: records placed per request
    uint32_t* rec_req;      // [hits_cap] compaction: request of each staged record
};

// Host side of a launch: check the spectrum and the requests, and build the
// kernel's request records and tile offsets in c->h_reqs / c->h_tiles.
int plan_launch(pp_ctx* c, const pp_request* reqs, uint32_t n_req, const pp_spectrum* spec,
                SpecParams* sp, uint64_t* n_tiles_out) {
    if (n_req == 0) return fail(PP_ERR_INVALID_ARG, "no requests");
    if (!reqs) return fail(PP_ERR_INVALID_ARG, "requests is NULL");
    int rc = check_spectrum(spec, sp);
    if (rc) return rc;
    if ((rc = check_cone(reqs, n_req, c->n_min))) return rc;
    c->h_reqs.resize(n_req);
    c->h_tiles.resize((size_t)n_req + 1);
    c->h_tiles[0] = 0;
    for (uint32_t i = 0; i < n_req; ++i) {
        if ((rc = make_devreq(&reqs[i], i, &c->h_reqs[i]))) return rc;
        c->h_tiles[i + 1] = c->h_tiles[i] + (reqs[i].emitter.n_photons + kTilePhotons - 1) / kTilePhotons;
    }
    const uint64_t n_tiles = c->h_tiles[n_req];
    if (n_tiles > 0xFFFFFFFFull - (1ull << 24))
        return fail(PP_ERR_INVALID_ARG, "too many photons in one launch (%llu tiles)",
                    (unsigned long long)n_tiles);
    *n_tiles_out = n_tiles;
    return PP_OK;
}

// The context's work buffers, grown to fit (compaction: also its buffers).
int ctx_work(pp_ctx* c, uint32_t n_req, uint64_t n_tiles, uint64_t hits_cap, bool knots,
             bool compaction, WorkBufs* W) {
    int rc;
    if ((rc = ensure(c->reqs, n_req * sizeof(DevReq)))) return rc;
    if ((rc = ensure(c->tile_start, (n_req + 1) * sizeof(uint64_t)))) return rc;
    if ((rc = ensure(c->partials, std::max<uint64_t>(n_tiles, 1) * sizeof(TilePartial)))) return rc;
    if ((rc = ensure(c->counters, 64))) return rc;
    if ((rc = ensure(c->sums, n_req * sizeof(pp_summary)))) return rc;
    if (knots && (rc = ensure(c->knots, PP_MAX_KNOTS * sizeof(float)))) return rc;
    if ((rc = ensure(c->staging, std::max<uint64_t>(hits_cap, 1) * sizeof(pp_hit)))) return rc;
    if (compaction) {
        if ((rc = ensure(c->cnt, n_req * sizeof(uint32_t)))) return rc;
        if ((rc = ensure(c->fill, n_req * sizeof(uint32_t)))) return rc;
        if ((rc = ensure(c->rec_req, std::max<uint64_t>(hits_cap, 1) * sizeof(uint32_t)))) return rc;
    }
    W->reqs = (DevReq*)c->reqs.p;
    W->tile_start = (uint64_t*)c->tile_start.p;
    W->partials = (TilePartial*)c->partials.p;
    W->counters = (uint32_t*)c->counters.p;
    W->sums = (pp_summary*)c->sums.p;
    W->knots = (float*)c->knots.p;
    W->staging = (pp_hit*)c->staging.p;
    W->cnt = (uint32_t*)c->cnt.p;
    W->fill = (uint32_t*)c->fill.p;
    W->rec_req = (uint32_t*)c->rec_req.p;
    return PP_OK;
}

// Enqueue on c->stream: knots and requests up, zero the counters, run the
// persistent propagation kernel over all tiles of all requests (hit records to
// W.staging, capacity hits_cap), reduce per request into W.sums.  pinned_io:
// upload through the context's pinned buffer (callers that synchronise before
// returning, so that the buffer is free again at the next call).
int enqueue_propagation(pp_ctx* c, uint32_t n_req, uint64_t n_tiles, const SpecParams& sp,
                        const pp_spectrum* spec, const WorkBufs& W, uint64_t hits_cap,
                        bool pinned_io = false) {
    int rc;
    if ((rc = upload_knots(spec, W.knots, c->stream))) return rc;
    const size_t req_bytes = n_req * sizeof(DevReq), tile_bytes = (n_req + 1) * sizeof(uint64_t);
    if (pinned_io && req_bytes + tile_bytes <= kIoDownload) {
        if ((rc = ensure_io(c))) return rc;
        char* io = static_cast<char*>(c->io);
        memcpy(io, c->h_reqs.data(), req_bytes);
        memcpy(io + req_bytes, c->h_tiles.data(), tile_bytes);
        PP_HIP(hipMemcpyAsync(W.reqs, io, req_bytes, hipMemcpyHostToDevice, c->stream));
        PP_HIP(hipMemcpyAsync(W.tile_start, io + req_bytes, tile_bytes, hipMemcpyHostToDevice,
                              c->stream));
    } else {
        PP_HIP(hipMemcpyAsync(W.reqs, c->h_reqs.data(), req_bytes, hipMemcpyHostToDevice,
                              c->stream));
        PP_HIP(hipMemcpyAsync(W.tile_start, c->h_tiles.data(), tile_bytes, hipMemcpyHostToDevice,
                              c->stream));
    }
    // counters: [0] tile counter (u32), [2..3] hit cursor (u64)
    fill_u32_kernel<<<1, 64, 0, c->stream>>>(W.counters, 16, 0u);
    PP_HIP(hipGetLastError());

    int grid = 0;
    if ((rc = grid_for(c, n_tiles, &grid))) return rc;
    KParams P;
    memset(&P, 0, sizeof P);
    P.reqs = W.reqs;
    P.tile_start = W.tile_start;
    P.n_req = n_req;
    P.n_tiles = n_tiles;
    P.tile_counter = W.counters;
    P.tile = make_tile_ctx((const float4*)c->table.p, c->n_points, W.knots, sp, c->lmin,
                           c->inv_dl, make_hg(c->g));
    P.hits = W.staging;
    P.hits_cap = hits_cap;
    P.hit_cursor = (unsigned long long*)(W.counters + 2);
    P.partials = W.partials;
    PP_HIP(hipEventRecord(c->ev0, c->stream));
    if (grid > 0) {
        propagate_kernel<<<grid, kBlock, 0, c->stream>>>(P);
        PP_HIP(hipGetLastError());
    }
    PP_HIP(hipEventRecord(c->ev1, c->stream));
    c->timed = true;
    reduce_kernel<<<n_req, 64, 0, c->stream>>>(W.partials, W.tile_start, W.reqs, n_req, W.sums);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// Group the staged hits by request into d_out with offsets; fills n_stored and
// status of W.sums.
int enqueue_compaction(pp_ctx* c, uint32_t n_req, const WorkBufs& W, uint64_t hits_cap,
                       uint64_t per_cap, pp_hit* d_out, uint64_t* d_offsets) {
    const int zb = (int)std::min<uint64_t>(1024, (n_req + 255) / 256);
    fill_u32_kernel<<<zb, 256, 0, c->stream>>>(W.cnt, n_req, 0u);
    fill_u32_kernel<<<zb, 256, 0, c->stream>>>(W.fill, n_req, 0u);
    const unsigned long long* cursor = (const unsigned long long*)(W.counters + 2);
    const int gb = 2048;
    count_kernel<<<gb, 256, 0, c->stream>>>(W.staging, cursor, hits_cap, n_req, W.cnt, W.rec_req);
    scan_kernel<<<1, 1024, 0, c->stream>>>(W.cnt, n_req, per_cap, d_offsets, W.sums);
    scatter_kernel<<<gb, 256, 0, c->stream>>>(W.staging, W.rec_req, cursor, hits_cap, n_req, per_cap,
                                              d_offsets, W.fill, d_out);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

int validate_ctx(pp_ctx* c) {
    if (!c) return fail(PP_ERR_INVALID_ARG, "context is NULL");
    PP_HIP(hipSetDevice(c->device));
    return PP_OK;
}


// Propagation and compaction of a batch into d_out / d_offsets, summaries in
// the context's buffer (returned in *sums).
int run_batch(pp_ctx* c, const pp_request* reqs, uint32_t n_req, const pp_spectrum* spec,
              pp_hit* d_out, uint64_t hits_cap, uint64_t per_cap, uint64_t* d_offsets,
              pp_summary** sums) {
    int rc;
    if ((rc = check_records(hits_cap, "hits_cap"))) return rc;
    SpecParams sp;
    uint64_t n_tiles = 0;
    if ((rc = plan_launch(c, reqs, n_req, spec, &sp, &n_tiles))) return rc;
    WorkBufs W;
    if ((rc = ctx_work(c, n_req, n_tiles, hits_cap, sp.kind == PP_SPEC_KNOTS, true, &W))) return rc;
    if ((rc = enqueue_propagation(c, n_req, n_tiles, sp, spec, W, hits_cap))) return rc;
    *sums = W.sums;
    return enqueue_compaction(c, n_req, W, hits_cap, per_cap, d_out, d_offsets);
}


}  // namespace

// ---- work buffers of the allocation-free entry point ------------------------
// pp_propagate_batch_device_noalloc takes every work buffer from one
// file-scope device array, carved into fixed regions (offsets below, each
// 256-byte aligned); the summaries go straight to the caller's array.  Every
// GPU the library runs on holds its own copy until the process ends.  The
// array's name is a host-side stand-in: host code uses only the address that
// hipGetSymbolAddress returns for the current device (na_scratch_setup).
namespace {

constexpr uint64_t na_align(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t kNaReqs = 0;
constexpr uint64_t kNaTileStart = kNaReqs + na_align(sizeof(DevReq) * PP_NOALLOC_MAX_REQUESTS);
constexpr uint64_t kNaPartials = kNaTileStart + na_align(8ull * (PP_NOALLOC_MAX_REQUESTS + 1));
constexpr uint64_t kNaCounters = kNaPartials + na_align(sizeof(TilePartial) * (uint64_t)PP_NOALLOC_MAX_TILES);
constexpr uint64_t kNaKnots = kNaCounters + 256;
constexpr uint64_t kNaStaging = kNaKnots + na_align(sizeof(float) * PP_MAX_KNOTS);
constexpr uint64_t kNaCnt = kNaStaging + na_align(sizeof(pp_hit) * (uint64_t)PP_NOALLOC_MAX_HITS);
constexpr uint64_t kNaFill = kNaCnt + na_align(4ull * PP_NOALLOC_MAX_REQUESTS);
constexpr uint64_t kNaRecReq = kNaFill + na_align(4ull * PP_NOALLOC_MAX_REQUESTS);
constexpr uint64_t kNaBytes = kNaRecReq + na_align(4ull * PP_NOALLOC_MAX_HITS);
static_assert(kNaBytes == PP_NOALLOC_SCRATCH_BYTES, "scratch layout and the header's stated size differ");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};  // device address of `scratch`, per device

// Once per device, with that device current: the device address of its copy
// of `scratch`.  Devices beyond kNaMaxDevices get none (the entry point then
// refuses).
int na_scratch_setup(pp_ctx* c) {
    if (c->device >= kNaMaxDevices) return PP_OK;
    std::lock_guard<std::mutex> g(g_na_mu);
    if (!g_na_ptr[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PP_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", c->device,
                        hipGetErrorString(e));
        g_na_ptr[c->device] = p;
    }
    c->na_scratch = g_na_ptr[c->device];
    return PP_OK;
}

// The allocation-free entry point's work buffers in the device's scratch array.
WorkBufs scratch_work(void* base, pp_summary* d_summaries) {
    char* b = static_cast<char*>(base);
    WorkBufs W;
    W.reqs = reinterpret_cast<DevReq*>(b + kNaReqs);
    W.tile_start = reinterpret_cast<uint64_t*>(b + kNaTileStart);
    W.partials = reinterpret_cast<TilePartial*>(b + kNaPartials);
    W.counters = reinterpret_cast<uint32_t*>(b + kNaCounters);
    W.sums = d_summaries;
    W.knots = reinterpret_cast<float*>(b + kNaKnots);
    W.staging = reinterpret_cast<pp_hit*>(b + kNaStaging);
    W.cnt = reinterpret_cast<uint32_t*>(b + kNaCnt);
    W.fill = reinterpret_cast<uint32_t*>(b + kNaFill);
    W.rec_req = reinterpret_cast<uint32_t*>(b + kNaRecReq);
    return W;
}

}  // namespace

extern "C" {

const char* pp_last_error(void) { return last_error().c_str(); }

int pp_abi_version(void) { return PP_ABI_VERSION; }

int pp_create(pp_ctx** out, int device, const pp_medium_tables* medium) {
    if (!out) return fail(PP_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n_dev = 0;
    PP_HIP(hipGetDeviceCount(&n_dev));
    if (device < 0 || device >= n_dev)
        return fail(PP_ERR_INVALID_ARG, "device %d not in [0, %d)", device, n_dev);
    int rc = check_medium(medium);
    if (rc) return rc;
    PP_HIP(hipSetDevice(device));
    pp_ctx* c = new pp_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return fail(PP_ERR_DEVICE, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    }
    c->n_cu = prop.multiProcessorCount;
    snprintf(c->name, sizeof c->name, "%s (%s)", prop.name, prop.gcnArchName);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        pp_destroy(c);
        return fail(PP_ERR_DEVICE, "stream/event creation: %s", hipGetErrorString(e));
    }
    // host request records of up to PP_NOALLOC_MAX_REQUESTS requests without
    // growing later (the allocation-free entry point's bound)
    c->h_reqs.reserve(PP_NOALLOC_MAX_REQUESTS);
    c->h_tiles.reserve(PP_NOALLOC_MAX_REQUESTS + 1);
    if ((rc = upload_medium(c, medium)) || (rc = na_scratch_setup(c))) {
        std::string msg = last_error();
        pp_destroy(c);
        last_error() = msg;
        return rc;
    }
    *out = c;
    return PP_OK;
}

int pp_set_medium(pp_ctx* c, const pp_medium_tables* medium) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    PP_HIP(hipStreamSynchronize(c->stream));
    return upload_medium(c, medium);
}

int pp_destroy(pp_ctx* c) {
    if (!c) return PP_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->table, &c->knots, &c->reqs, &c->tile_start, &c->partials,
                      &c->counters, &c->staging, &c->cnt, &c->fill, &c->sums,
                      &c->offsets, &c->out_hits, &c->scratch, &c->rec_req};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int h = 0; h < 2; ++h) {
        if (c->pinned[h]) (void)hipHostFree(c->pinned[h]);
        if (c->copy_ev[h]) (void)hipEventDestroy(c->copy_ev[h]);
    }
    if (c->io) (void)hipHostFree(c->io);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PP_OK;
}

int pp_device_info(pp_ctx* c, char* name, size_t name_cap, int* n_cu) {
    if (!c) return fail(PP_ERR_INVALID_ARG, "context is NULL");
    if (name && name_cap) snprintf(name, name_cap, "%s", c->name);
    if (n_cu) *n_cu = c->n_cu;
    return PP_OK;
}

}  // extern "C"

namespace {

// Hit records to the caller's host memory (pageable as a rule): in chunks
// through two pinned staging halves, the DMA of one chunk overlapping the host
// copy of the one before.  hipMemcpy straight into pageable memory stages
// through the runtime's own small buffers, at a fraction of the link's rate.
constexpr uint64_t kHostStageBytes = 16ull << 20;
constexpr int kHostCopyThreads = 4;

// memcpy split over kHostCopyThreads threads: the first touch of fresh
// pageable memory (a page fault per 4 KiB) bounds one thread well below the
// DMA rate.  Below 1 MiB the thread starts would cost more than they save.
void parallel_memcpy(char* d, const char* s, uint64_t n) {
    if (n < (1u << 20)) {
        memcpy(d, s, n);
        return;
    }
    const uint64_t part = ((n + kHostCopyThreads - 1) / kHostCopyThreads + 4095) & ~4095ull;
    std::thread th[kHostCopyThreads - 1];
    for (int i = 1; i < kHostCopyThreads; ++i) {
        const uint64_t o = (uint64_t)i * part;
        if (o < n) th[i - 1] = std::thread([=] { memcpy(d + o, s + o, std::min(part, n - o)); });
    }
    memcpy(d, s, std::min(part, n));
    for (auto& t : th)
        if (t.joinable()) t.join();
}

int copy_records_to_host(pp_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (bytes <= kHostStageBytes / 4) {
        PP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return PP_OK;
    }
    for (int h = 0; h < 2; ++h) {
        if (!c->pinned[h]) PP_HIP(hipHostMalloc(&c->pinned[h], kHostStageBytes, 0));
        if (!c->copy_ev[h]) PP_HIP(hipEventCreateWithFlags(&c->copy_ev[h], hipEventDisableTiming));
    }
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    PP_HIP(hipMemcpyAsync(c->pinned[0], s, std::min(kHostStageBytes, bytes), hipMemcpyDeviceToHost,
                          c->stream));
    PP_HIP(hipEventRecord(c->copy_ev[0], c->stream));
    for (uint64_t off = 0, k = 0; off < bytes; off += kHostStageBytes, ++k) {
        const int h = (int)(k & 1);
        const uint64_t next = off + kHostStageBytes;
        if (next < bytes) {  // the other half's previous chunk was copied out last pass
            PP_HIP(hipMemcpyAsync(c->pinned[h ^ 1], s + next, std::min(kHostStageBytes, bytes - next),
                                  hipMemcpyDeviceToHost, c->stream));
            PP_HIP(hipEventRecord(c->copy_ev[h ^ 1], c->stream));
        }
        PP_HIP(hipEventSynchronize(c->copy_ev[h]));
        parallel_memcpy(d + off, static_cast<const char*>(c->pinned[h]),
                        std::min(kHostStageBytes, bytes - off));
    }
    return PP_OK;
}

}  // namespace

extern "C" {

int pp_propagate(pp_ctx* c, const pp_emitter* emitter, const pp_spectrum* spectrum,
                 const pp_sphere* target, uint64_t seed, uint32_t stream, uint32_t max_steps,
                 pp_hit* hits_out, uint64_t hits_cap, pp_summary* summary_out) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (!emitter || !target) return fail(PP_ERR_INVALID_ARG, "emitter/target is NULL");
    if (hits_cap && !hits_out) return fail(PP_ERR_INVALID_ARG, "hits_out is NULL with hits_cap > 0");
    if ((rc = check_records(hits_cap, "hits_cap"))) return rc;
    pp_request q;
    memset(&q, 0, sizeof q);
    q.emitter = *emitter;
    q.target = *target;
    q.seed = seed;
    q.stream = stream;
    q.max_steps = max_steps;
    SpecParams sp;
    uint64_t n_tiles = 0;
    if ((rc = plan_launch(c, &q, 1, spectrum, &sp, &n_tiles))) return rc;
    WorkBufs W;
    if ((rc = ctx_work(c, 1, n_tiles, hits_cap, sp.kind == PP_SPEC_KNOTS, false, &W))) return rc;
    if ((rc = enqueue_propagation(c, 1, n_tiles, sp, spectrum, W, hits_cap, true))) return rc;
    if ((rc = ensure_io(c))) return rc;
    char* down = static_cast<char*>(c->io) + kIoDownload;
    PP_HIP(hipMemcpyAsync(down, W.sums, sizeof(pp_summary), hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipMemcpyAsync(down + sizeof(pp_summary), W.counters + 2, sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    pp_summary s;
    unsigned long long cursor = 0;
    memcpy(&s, down, sizeof s);
    memcpy(&cursor, down + sizeof(pp_summary), sizeof cursor);
    const uint64_t stored = std::min<uint64_t>(cursor, hits_cap);
    if (stored && (rc = copy_records_to_host(c, hits_out, W.staging, stored * sizeof(pp_hit))))
        return rc;
    s.n_stored = stored;
    s.status = s.n_hits > stored ? PP_ERR_OVERFLOW : PP_OK;
    if (summary_out) *summary_out = s;
    if (s.status == PP_ERR_OVERFLOW)
        return fail(PP_ERR_OVERFLOW, "hit buffer overflow: %llu hits, capacity %llu",
                    (unsigned long long)s.n_hits, (unsigned long long)hits_cap);
    return PP_OK;
}

int pp_propagate_batch(pp_ctx* c, const pp_request* requests, uint32_t n_req,
                       const pp_spectrum* spectrum, pp_hit* hits_out, uint64_t hits_cap,
                       uint64_t per_request_cap, uint64_t* offsets_out,
                       pp_summary* summaries_out) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (hits_cap && !hits_out) return fail(PP_ERR_INVALID_ARG, "hits_out is NULL with hits_cap > 0");
    if (!offsets_out || !summaries_out)
        return fail(PP_ERR_INVALID_ARG, "offsets_out/summaries_out is NULL");
    if ((rc = check_records(hits_cap, "hits_cap"))) return rc;
    const uint64_t out_cap = out_records(hits_cap, per_request_cap, n_req);
    if ((rc = ensure(c->out_hits, std::max<uint64_t>(out_cap, 1) * sizeof(pp_hit)))) return rc;
    if ((rc = ensure(c->offsets, ((uint64_t)n_req + 1) * sizeof(uint64_t)))) return rc;
    pp_summary* d_sums = nullptr;
    if ((rc = run_batch(c, requests, n_req, spectrum, (pp_hit*)c->out_hits.p, hits_cap,
                        per_request_cap, (uint64_t*)c->offsets.p, &d_sums)))
        return rc;
    PP_HIP(hipMemcpyAsync(offsets_out, c->offsets.p, ((uint64_t)n_req + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipMemcpyAsync(summaries_out, d_sums, (uint64_t)n_req * sizeof(pp_summary),
                          hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    const uint64_t stored = offsets_out[n_req];
    if (stored && (rc = copy_records_to_host(c, hits_out, c->out_hits.p, stored * sizeof(pp_hit))))
        return rc;
    for (uint32_t i = 0; i < n_req; ++i)
        if (summaries_out[i].status == PP_ERR_OVERFLOW)
            return fail(PP_ERR_OVERFLOW, "hit buffer overflow (capacity %llu)",
                        (unsigned long long)hits_cap);
    return PP_OK;
}

int pp_propagate_batch_device(pp_ctx* c, const pp_request* requests, uint32_t n_req,
                              const pp_spectrum* spectrum, pp_hit* d_hits, uint64_t hits_cap,
                              uint64_t per_request_cap, uint64_t* d_offsets,
                              pp_summary* d_summaries) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if ((hits_cap && !d_hits) || !d_offsets || !d_summaries)
        return fail(PP_ERR_INVALID_ARG, "device output pointer is NULL");
    pp_summary* d_sums = nullptr;
    if ((rc = run_batch(c, requests, n_req, spectrum, d_hits, hits_cap, per_request_cap,
                        d_offsets, &d_sums)))
        return rc;
    PP_HIP(hipMemcpyAsync(d_summaries, d_sums, (uint64_t)n_req * sizeof(pp_summary),
                          hipMemcpyDeviceToDevice, c->stream));
    return PP_OK;
}

int pp_propagate_batch_device_noalloc(pp_ctx* c, const pp_request* requests, uint32_t n_req,
                                      const pp_spectrum* spectrum, pp_hit* d_hits,
                                      uint64_t hits_cap, uint64_t per_request_cap,
                                      uint64_t* d_offsets, pp_summary* d_summaries) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if ((hits_cap && !d_hits) || !d_offsets || !d_summaries)
        return fail(PP_ERR_INVALID_ARG, "device output pointer is NULL");
    if (!c->na_scratch)
        return fail(PP_ERR_DEVICE, "no scratch array on device %d", c->device);
    if (n_req > PP_NOALLOC_MAX_REQUESTS)
        return fail(PP_ERR_INVALID_ARG, "n_req %u exceeds PP_NOALLOC_MAX_REQUESTS (%u)", n_req,
                    (unsigned)PP_NOALLOC_MAX_REQUESTS);
    if (hits_cap > PP_NOALLOC_MAX_HITS)
        return fail(PP_ERR_INVALID_ARG, "hits_cap %llu exceeds PP_NOALLOC_MAX_HITS (%u)",
                    (unsigned long long)hits_cap, (unsigned)PP_NOALLOC_MAX_HITS);
    SpecParams sp;
    uint64_t n_tiles = 0;
    if ((rc = plan_launch(c, requests, n_req, spectrum, &sp, &n_tiles))) return rc;
    if (n_tiles > PP_NOALLOC_MAX_TILES)
        return fail(PP_ERR_INVALID_ARG, "%llu tiles of %d photons exceed PP_NOALLOC_MAX_TILES (%u)",
                    (unsigned long long)n_tiles, PP_TILE_PHOTONS, (unsigned)PP_NOALLOC_MAX_TILES);
    const WorkBufs W = scratch_work(c->na_scratch, d_summaries);
    if ((rc = enqueue_propagation(c, n_req, n_tiles, sp, spectrum, W, hits_cap))) return rc;
    return enqueue_compaction(c, n_req, W, hits_cap, per_request_cap, d_hits, d_offsets);
}

int pp_synchronize(pp_ctx* c) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    PP_HIP(hipStreamSynchronize(c->stream));
    return PP_OK;
}

int pp_last_kernel_ms(pp_ctx* c, float* ms) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (!ms) return fail(PP_ERR_INVALID_ARG, "ms is NULL");
    if (!c->timed) return fail(PP_ERR_INVALID_ARG, "no kernel launched yet");
    PP_HIP(hipEventSynchronize(c->ev1));
    PP_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
    return PP_OK;
}

int pp_selftest_rotate(pp_ctx* c, const float* old_dirs, const float* new_dirs, float* out,
                       uint64_t n) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (n == 0) return PP_OK;
    if (!old_dirs || !new_dirs || !out) return fail(PP_ERR_INVALID_ARG, "NULL array");
    const size_t bytes = n * 3 * sizeof(float);
    if ((rc = ensure(c->scratch, 3 * bytes))) return rc;
    float* d = (float*)c->scratch.p;
    PP_HIP(hipMemcpyAsync(d, old_dirs, bytes, hipMemcpyHostToDevice, c->stream));
    PP_HIP(hipMemcpyAsync(d + 3 * n, new_dirs, bytes, hipMemcpyHostToDevice, c->stream));
    selftest_rotate_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(d, d + 3 * n, d + 6 * n, n);
    PP_HIP(hipGetLastError());
    PP_HIP(hipMemcpyAsync(out, d + 6 * n, bytes, hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    return PP_OK;
}

int pp_selftest_philox(pp_ctx* c, const uint32_t* ctr, uint64_t seed, uint32_t* words_out,
                       uint64_t n) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (n == 0) return PP_OK;
    if (!ctr || !words_out) return fail(PP_ERR_INVALID_ARG, "NULL array");
    const size_t bytes = n * 4 * sizeof(uint32_t);
    if ((rc = ensure(c->scratch, 2 * bytes))) return rc;
    uint32_t* d = (uint32_t*)c->scratch.p;
    PP_HIP(hipMemcpyAsync(d, ctr, bytes, hipMemcpyHostToDevice, c->stream));
    selftest_philox_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
        d, (uint32_t)seed, (uint32_t)(seed >> 32), d + 4 * n, n);
    PP_HIP(hipGetLastError());
    PP_HIP(hipMemcpyAsync(words_out, d + 4 * n, bytes, hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    return PP_OK;
}

int pp_selftest_hg(pp_ctx* c, float g, uint64_t seed, float* out, uint64_t n) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (!(std::fabs(g) < 1.0f)) return fail(PP_ERR_INVALID_ARG, "g=%g not in (-1,1)", (double)g);
    if (n == 0) return PP_OK;
    if (!out) return fail(PP_ERR_INVALID_ARG, "NULL array");
    if ((rc = ensure(c->scratch, n * sizeof(float)))) return rc;
    selftest_hg_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
        make_hg(g), (uint32_t)seed, (uint32_t)(seed >> 32), (float*)c->scratch.p, n);
    PP_HIP(hipGetLastError());
    PP_HIP(hipMemcpyAsync(out, c->scratch.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    return PP_OK;
}

int pp_selftest_hg_moments(pp_ctx* c, float g, uint64_t seed, uint64_t n, double* sums_out) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (!(std::fabs(g) < 1.0f)) return fail(PP_ERR_INVALID_ARG, "g=%g not in (-1,1)", (double)g);
    if (!sums_out) return fail(PP_ERR_INVALID_ARG, "NULL array");
    const int blocks = 1024;
    if ((rc = ensure(c->scratch, 2 * blocks * sizeof(double)))) return rc;
    selftest_hg_moments_kernel<<<blocks, 256, 0, c->stream>>>(
        make_hg(g), (uint32_t)seed, (uint32_t)(seed >> 32), n, (double*)c->scratch.p);
    PP_HIP(hipGetLastError());
    std::vector<double> h(2 * blocks);
    PP_HIP(hipMemcpyAsync(h.data(), c->scratch.p, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    double a = 0.0, b = 0.0;
    for (int i = 0; i < blocks; ++i) {
        a += h[2 * i];
        b += h[2 * i + 1];
    }
    sums_out[0] = a;
    sums_out[1] = b;
    return PP_OK;
}

}  // extern "C"

// ===========================================================================
// Serving loop: continuous batching into a persistent kernel
// ===========================================================================
//
// Host side: pp_serve_core.hpp (admission by slots and hit-pool room,
// publishing of range entries, reaping of completions, kernel stop and
// relaunch), run here on a HIP backend.
// Device side (persistent kernel, a fixed share of the resident block slots):
// a block takes a ticket (the next global tile index) with one atomicAdd and
// waits until the tile is published.  The published count is mirrored from
// pinned memory into a device-memory gate word; only the block holding the
// lowest unpublished ticket (the head) reads the host's counters over PCIe, so
// the link carries one poller however many blocks wait.  Before it moves the
// gate the head also copies the newly published range entries into a
// device-memory ring, so a block finds its range entry by a 64-ary search in
// device memory (a search over pinned memory, some 70 scattered reads a tile,
// had held the serving kernel to 40 % of the batch kernel's throughput on the
// same work).  The request's slot descriptor is read over PCIe only when the
// tile belongs to another request than the block's previous tile (the entry
// carries the request's tag); the block keeps it in LDS.  It runs the tile with
// the same run_tile as the batch kernel, stores the tile partial, and the block
// that completes a request's last tile adds the partials in tile order and
// posts the summary to pinned memory.  The host stops the kernel (`stop`,
// raised into the gate by the head) only when every published tile is done, so
// the tickets held at exit are all unpublished and the next launch restarts the
// ticket counter at the published count.  The head also stops the kernel by
// itself after waiting wait_ticks with nothing published, reporting the tile
// to resume at.

namespace {

using pps::kSrvRing;
using pps::SrvControl;
using pps::SrvDone;
using pps::SrvEntry;
using pps::SrvSlot;

constexpr uint64_t kSrvExit = ~0ull;
constexpr unsigned long long kGateStop = 1ull << 63;  // gate word: published | stop
constexpr size_t kSrvGateOffset = 4096;  // ticket, gate and mirror counts on separate pages
// Room left on every SIMD for other kernels under PP_SERVER_LEAVE_ROOM: VGPRs
// per lane and wave slots (of 8) the persistent grid does not take.  Result
// copies never need compute units (an SDMA engine, or a hit pool in host
// memory), so this is only for the caller's other work on the device.  64
// requests of 1e8 photons through the server took 95.2 ms at 8 waves per SIMD
// and 97.9 ms at 7 (the batch kernel 94.1 ms).
constexpr int kSrvVgprReserve = 32;
constexpr int kSrvRoomWavesPerSimd = 7;
constexpr uint64_t kSrvStageRecords = 1u << 16;  // 4 MiB per half of the result staging

// The serving kernel's addresses and limits, in device memory (written once
// per server).  The kernel reads each field where it uses it (SA below), so
// that none of them occupies registers through the tile's step loop.
struct SrvAddr {
    SrvControl* ctl;
    const SrvEntry* entries;
    const SrvSlot* slots;
    SrvDone* done;
    unsigned long long* ticket;       // next ticket (global tile index)
    unsigned long long* gate;         // published tiles mirrored from ctl | kGateStop
    unsigned long long* dmeta;        // [0] entries mirrored, [1] first_live when mirrored
    SrvEntry* dentries;               // device copy of the range-entry ring
    uint32_t* tiles_done;             // [max_slots]
    unsigned long long* hit_cursor;   // [max_slots]
    TilePartial* partials;            // [max_slots][max_tiles]
    pp_hit* pool;                     // device memory, or pinned host memory
    uint64_t wait_ticks;              // s_memrealtime ticks the head waits before stopping
    uint64_t max_tiles;
    TileCtx tile;                     // the server's medium, spectrum and scattering
};

struct SrvParams {
    const SrvAddr* addr;
};


// A field of the serving addresses, read here with a scalar load.
#define SA(field) load_constant(&opaque_ptr(P.addr)->field, 0)

template <typename T>
__device__ __forceinline__ T sys_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Agent-scope atomic loads and stores (sc1): coherent across the XCDs' L2s.
template <typename T>
__device__ __forceinline__ T agent_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void agent_store(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A value that is the same in every lane, moved to scalar registers.
// kGateStop as an atomic's operand, formed in scalar registers where it is
// used (an empty asm): hoisted, the 64-bit constant was held in two VGPRs
// through the whole kernel and spilled
__device__ __forceinline__ unsigned long long stop_bit() {
    unsigned long long v = kGateStop;
    __asm__ volatile("" : "+s"(v));
    return v;
}

__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// A trivially copyable value that is the same in every lane (read from LDS),
// moved word by word into scalar registers.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    T r;
    const uint32_t* s = reinterpret_cast<const uint32_t*>(&v);
    uint32_t* d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for (uint32_t k = 0; k < sizeof(T) / 4; ++k) d[k] = __builtin_amdgcn_readfirstlane(s[k]);
    return r;
}

// Copy the host's range entries [mirrored, n_entries) into the device ring
// (one wave; each lane moves one 8-byte word, 16 entries a pass, four passes
// in flight).  Run by the head block before it moves the gate, so that every
// tile below the gate has its entry in device memory and the other blocks
// search there, not over PCIe.  The entries copied are those of unfinished
// jobs: the host does not reuse their ring positions meanwhile.
__device__ void mirror_entries(const SrvParams& P, uint32_t lane) {
    const uint64_t ne = uniform_u64(sys_load(&SA(ctl)->n_entries));
    const uint64_t fl = uniform_u64(sys_load(&SA(ctl)->first_live));
    const uint64_t m = uniform_u64(agent_load(&SA(dmeta)[0]));
    constexpr uint32_t kWords = sizeof(SrvEntry) / 8;
    constexpr uint32_t kPerPass = 64 / kWords;
    const uint64_t* src = reinterpret_cast<const uint64_t*>(SA(entries));
    uint64_t* dst = reinterpret_cast<uint64_t*>(SA(dentries));
    const uint32_t word = lane % kWords;
    for (uint64_t e0 = m; e0 < ne; e0 += 4 * kPerPass) {
        uint64_t v[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint64_t e = e0 + k * kPerPass + lane / kWords;
            v[k] = e < ne ? sys_load(src + (e % kSrvRing) * kWords + word) : 0;
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint64_t e = e0 + k * kPerPass + lane / kWords;
            if (e < ne) agent_store(dst + (e % kSrvRing) * kWords + word, v[k]);
        }
    }
    // the entries have landed before the count that admits them to searches
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) {
        agent_store(&SA(dmeta)[1], (unsigned long long)fl);
        agent_store(&SA(dmeta)[0], (unsigned long long)ne);
    }
    // the counts have landed before the gate moves
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

// At most 96 VGPRs (5 waves per SIMD's worth): the grid takes 4 or 5 waves per
// SIMD and (512 - reserve) / VGPRs must allow that; 113 would allow 3 or 4.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) serve_kernel(SrvParams P) {
    __shared__ unsigned long long s_tile;
    __shared__ SrvSlot s_desc;
    __shared__ uint64_t s_entry_t0, s_entry_local0;
    __shared__ uint32_t s_slot;
    __shared__ uint32_t s_desc_slot, s_desc_tag;  // request whose descriptor s_desc holds
    __shared__ uint32_t s_last;
    __shared__ double s_red[kWaves][3];
    __shared__ uint64_t s_redi[kWaves][2];
    __shared__ uint4 s_hitbuf[kWaves][kHitBuf * 4];
    __shared__ float s_idir[3][kBlock];
    __shared__ uint2 s_keys[kWaves][kKeyRounds];
    if (threadIdx.x == 0) s_desc_slot = ~0u;

    for (;;) {
        // the thread index taken afresh per tile (an empty asm): the per-lane
        // addresses of the scheduling code are recomputed per tile instead of
        // being held, spilled, through the whole kernel
        uint32_t tid = threadIdx.x;
        __asm__ volatile("" : "+v"(tid));
        __syncthreads();  // the previous tile's shared words consumed
        if (tid < 64) {   // wave 0: ticket, gate, range entry, slot descriptor
            const uint32_t lane = tid;
            unsigned long long tk = 0;
            if (lane == 0) tk = atomicAdd(SA(ticket), 1ull);
            // lane 0 took the ticket; it is the first active lane (the
            // whole wave is here), so a first-lane read broadcasts it
            const uint64_t t = uniform_u64(tk);
            uint64_t got = kSrvExit;
            uint64_t wait_start = 0;  // device clock when this block began to wait as head
            for (;;) {
                const uint64_t g = uniform_u64(agent_load(SA(gate)));
                if (t < (g & ~kGateStop)) {
                    got = t;
                    break;
                }
                if (g & kGateStop) break;
                if (t == g) {
                    // head of the queue: the one block that reads pinned memory.
                    // Relaxed is enough: every later read of host-written data
                    // is itself an uncached system-scope load, issued after
                    // the branch on this one.
                    const uint64_t hp = uniform_u64(sys_load(&SA(ctl)->published));
                    if (hp > t) {
                        mirror_entries(P, lane);
                        if (lane == 0) atomicMax(SA(gate), (unsigned long long)hp);
                        got = t;
                        break;
                    }
                    if (uniform_u64(sys_load(&SA(ctl)->stop))) {
                        if (lane == 0) atomicOr(SA(gate), stop_bit());
                        break;
                    }
                    // nothing published for wait_ticks: stop, and report the
                    // tile to resume at (every ticket below t was published
                    // and is done or being run; none above it was published)
                    const uint64_t now = __builtin_amdgcn_s_memrealtime();
                    if (wait_start == 0) {
                        wait_start = now;
                    } else if (now - wait_start > SA(wait_ticks)) {
                        if (lane == 0) {
                            __hip_atomic_store(&SA(ctl)->self_stop, (uint64_t)(t + 1),
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                            atomicOr(SA(gate), stop_bit());
                        }
                        break;
                    }
                    __builtin_amdgcn_s_sleep(16);
                } else {
                    // the farther behind the head, the longer between polls
                    const uint64_t behind = t - g;
                    const uint32_t reps = behind >= 448 ? 8u : 1u + (uint32_t)(behind >> 6);
                    for (uint32_t k = 0; k < reps; ++k) __builtin_amdgcn_s_sleep(127);
                }
            }
            if (lane == 0) {
                // how far the claims have got, for the host's publishing pace
                if (got != kSrvExit && (got & 63u) == 0)
                    __hip_atomic_store(&SA(ctl)->claimed, got, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                s_tile = got;
            }
            if (got != kSrvExit) {
                // Range entry holding tile t: 64-ary search of the device ring
                // over [first_live, mirrored).  The tile's own entry lies in
                // that range and stays in place while the tile is live; an
                // entry below it may meanwhile be overwritten by one 16384
                // later (a larger t0, so "not ok").  Such entries form a
                // prefix of the range, so the highest "ok" lane is still the
                // right one; if no lane is ok, the counts are read again.
                uint64_t lo = 0;
                for (bool found = false; !found;) {
                    lo = uniform_u64(agent_load(&SA(dmeta)[1]));
                    uint64_t hi = uniform_u64(agent_load(&SA(dmeta)[0]));
                    found = true;
                    while (hi - lo > 1) {
                        const uint64_t stride = (hi - lo + 63) / 64;
                        const uint64_t idx = lo + lane * stride;
                        const bool ok = idx < hi && agent_load(&SA(dentries)[idx % kSrvRing].t0) <= t;
                        const uint64_t bits = __ballot(ok);
                        if (bits == 0) {
                            found = false;
                            break;
                        }
                        const uint32_t l = 63u - (uint32_t)__clzll((long long)bits);
                        lo = lo + (uint64_t)l * stride;
                        hi = std::min(hi, lo + stride);
                    }
                }
                const SrvEntry* e = &SA(dentries)[lo % kSrvRing];
                const uint32_t slot = __builtin_amdgcn_readfirstlane(agent_load(&e->slot));
                const uint32_t tag = __builtin_amdgcn_readfirstlane(agent_load(&e->tag));
                if (lane == 0) {
                    s_entry_t0 = agent_load(&e->t0);
                    s_entry_local0 = agent_load(&e->local0);
                    s_slot = slot;
                }
                // the slot descriptor over PCIe only for another request than
                // the block's previous tile's (a slot's tag changes with the
                // request it holds)
                if (slot != __builtin_amdgcn_readfirstlane(s_desc_slot) ||
                    tag != __builtin_amdgcn_readfirstlane(s_desc_tag)) {
                    if (lane < sizeof(SrvSlot) / 4)
                        reinterpret_cast<uint32_t*>(&s_desc)[lane] =
                            sys_load(reinterpret_cast<const uint32_t*>(&SA(slots)[slot]) + lane);
                    if (lane == 0) {
                        s_desc_slot = slot;
                        s_desc_tag = tag;
                    }
                }
            }
        }
        __syncthreads();
        const unsigned long long t = s_tile;
        if (t == kSrvExit) break;
        const uint32_t slot = __builtin_amdgcn_readfirstlane(s_slot);
        const SrvSlot* sd = &s_desc;
        // the request's words from LDS into scalar registers, as the batch
        // kernel holds them
        const DevReq R = uniform_copy(sd->req);
        const uint64_t local = uniform_u64(t - s_entry_t0 + s_entry_local0);
        const uint64_t first = local * kTilePhotons;
        const uint64_t left = R.n_photons - first;
        const uint32_t count = (uint32_t)(left < kTilePhotons ? left : kTilePhotons);
        const uint64_t n_tiles = uniform_u64(sd->n_tiles);
        const uint64_t hit_cap = uniform_u64(sd->hit_cap);
        const uint64_t hit_base = uniform_u64(sd->hit_base);
        const uint64_t seq = uniform_u64(sd->seq);
        TilePartial* parts = SA(partials) + (uint64_t)slot * SA(max_tiles);
        const TilePartial tp = run_tile_any<true>(&opaque_ptr(P.addr)->tile, R, slot, first, count, SA(pool) + hit_base,
                                                  hit_cap, &SA(hit_cursor)[slot], s_red, s_redi,
                                                  s_hitbuf, s_idir, s_keys);
        // every wave waits for its own (system-scope) hit stores to complete
        // before the barrier; a workgroup-scope fence would not wait for them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            uint64_t* q = reinterpret_cast<uint64_t*>(parts + local);
            const uint64_t w[5] = {tp.n_steps, tp.n_hits, (uint64_t)__double_as_longlong(tp.sum_w),
                                   (uint64_t)__double_as_longlong(tp.sum_wt),
                                   (uint64_t)__double_as_longlong(tp.sum_wt2)};
#pragma unroll
            for (int k = 0; k < 5; ++k)
                __hip_atomic_store(q + k, w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // stores done before the count
            const uint32_t done = atomicAdd(&SA(tiles_done)[slot], 1u) + 1u;
            s_last = done == n_tiles;
        }
        __syncthreads();
        if (s_last && tid < 64) {
            const pp_summary s = reduce_partials<true>(parts, n_tiles, R.n_photons);
            if (tid == 0) {
                const unsigned long long cur = __hip_atomic_load(
                    &SA(hit_cursor)[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                pp_summary o = s;
                o.n_stored = cur < hit_cap ? cur : hit_cap;
                o.status = o.n_hits > o.n_stored ? PP_ERR_OVERFLOW : PP_OK;
                const uint64_t* w = reinterpret_cast<const uint64_t*>(&o);
                for (int k = 0; k < 8; ++k)
                    __hip_atomic_store(&SA(done)[slot].words[k], w[k], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                atomicExch(&SA(tiles_done)[slot], 0u);
                atomicExch(&SA(hit_cursor)[slot], 0ull);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // summary before its tag
                __hip_atomic_store(&SA(done)[slot].seq, seq, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

// Restart the ticket counter and the gate at the first unprocessed tile.
__global__ void serve_reset_kernel(unsigned long long* ticket, unsigned long long* gate,
                                   unsigned long long first) {
    if (threadIdx.x == 0) {
        atomicExch(ticket, first);
        atomicExch(gate, first);
    }
}

}  // namespace

namespace {

hsa_status_t find_cpu_agent(hsa_agent_t a, void* data) {
    hsa_device_type_t t;
    if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) == HSA_STATUS_SUCCESS &&
        t == HSA_DEVICE_TYPE_CPU) {
        *static_cast<hsa_agent_t*>(data) = a;
        return HSA_STATUS_INFO_BREAK;
    }
    return HSA_STATUS_SUCCESS;
}

// The serving loop's device side on HIP: the persistent kernel, its buffers,
// the pinned memory it shares with the host, and the result copies.  Result
// records never need compute units to leave the device: they go from a device
// pool through an SDMA engine, or the pool is pinned host memory (flag
// PP_SERVER_HOST_POOL, or no SDMA engine), which the kernel writes directly.
struct HipBackend {
    pp_ctx* ctx = nullptr;
    pp_server_config cfg{};
    SpecParams spec{};
    uint32_t n_points = 0;
    float lmin = 0.0f, inv_dl = 0.0f, g = 0.0f;
    int grid_ = 0;
    uint64_t wait_ticks = 0;
    hipStream_t kstream = nullptr;
    hipEvent_t kdone = nullptr;
    bool launched = false;
    // pinned, shared with the kernel
    SrvControl* ctl = nullptr;
    SrvEntry* entries = nullptr;
    SrvSlot* slots = nullptr;
    SrvDone* done = nullptr;
    // device
    DevBuf table, knots, dctl, dentries, tiles_done, hit_cursor, partials, pool, addr;
    pp_hit* host_pool = nullptr;  // the hit pool in pinned host memory
    // result copies on an SDMA engine (no compute units): two pinned staging
    // halves of kSrvStageRecords records, one completion signal each
    bool hsa_up = false, sdma = false;
    hsa_agent_t gpu_agent{}, cpu_agent{};
    hsa_amd_sdma_engine_id_t engine{};
    hsa_signal_t csig[2] = {};
    pp_hit* stage = nullptr;
    std::mutex copy_mu;

    ~HipBackend() {
        if (ctx) (void)hipSetDevice(ctx->device);
        if (kstream) (void)hipStreamSynchronize(kstream);
        if (ctl) (void)hipHostFree(ctl);
        if (entries) (void)hipHostFree(entries);
        if (slots) (void)hipHostFree(slots);
        if (done) (void)hipHostFree(done);
        if (host_pool) (void)hipHostFree(host_pool);
        DevBuf* bufs[] = {&table, &knots, &dctl, &dentries, &tiles_done, &hit_cursor, &partials, &pool,
                          &addr};
        for (DevBuf* b : bufs)
            if (b->p) (void)hipFree(b->p);
        if (kdone) (void)hipEventDestroy(kdone);
        if (kstream) (void)hipStreamDestroy(kstream);
        for (hsa_signal_t& sg : csig)
            if (sg.handle) (void)hsa_signal_destroy(sg);
        if (stage) (void)hipHostFree(stage);
        if (hsa_up) (void)hsa_shut_down();
    }

    pps::SrvShared shared() { return {ctl, entries, slots, done}; }

    int thread_init() {
        PP_HIP(hipSetDevice(ctx->device));
        return PP_OK;
    }

    uint64_t grid() const { return (uint64_t)grid_; }

    bool running() { return launched && hipEventQuery(kdone) == hipErrorNotReady; }

    void finish() {
        if (kstream) (void)hipStreamSynchronize(kstream);
    }

    int launch(uint64_t resume_tile) {
        SrvParams P;
        memset(&P, 0, sizeof P);
        P.addr = (const SrvAddr*)addr.p;
        // every tile below resume_tile is done and none above it was started
        serve_reset_kernel<<<1, 64, 0, kstream>>>(
            (unsigned long long*)dctl.p, (unsigned long long*)((char*)dctl.p + kSrvGateOffset),
            resume_tile);
        serve_kernel<<<grid_, kBlock, 0, kstream>>>(P);
        PP_HIP(hipGetLastError());
        PP_HIP(hipEventRecord(kdone, kstream));
        launched = true;
        return PP_OK;
    }

    // Start an SDMA copy of n records from the pool to staging half h.
    bool sdma_issue(int h, const pp_hit* src, uint64_t n) {
        hsa_signal_store_screlease(csig[h], 1);
        return hsa_amd_memory_async_copy_on_engine(stage + (uint64_t)h * kSrvStageRecords, cpu_agent,
                                                   src, gpu_agent, n * sizeof(pp_hit), 0, nullptr,
                                                   csig[h], engine, false) == HSA_STATUS_SUCCESS;
    }

    // Wait for staging half h; false on a copy error or after timeout_ms.
    bool sdma_wait(int h, int64_t timeout_ms) {
        const auto t_end = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms);
        for (;;) {
            const hsa_signal_value_t v = hsa_signal_wait_scacquire(
                csig[h], HSA_SIGNAL_CONDITION_LT, 1, 1000000, HSA_WAIT_STATE_ACTIVE);
            if (v < 0) return false;
            if (v == 0) return true;
            if (std::chrono::steady_clock::now() > t_end) return false;
        }
    }

    // Set up result copies on an SDMA engine: the HSA agents of the pool's GPU
    // and of host memory, an engine with a path between them, pinned staging
    // and one test copy.  Returns false (sdma stays false) if any step fails.
    bool setup_sdma() {
        if (hsa_init() != HSA_STATUS_SUCCESS) return false;
        hsa_up = true;
        hsa_amd_pointer_info_t pi;
        memset(&pi, 0, sizeof pi);
        pi.size = sizeof pi;
        if (hsa_amd_pointer_info(pool.p, &pi, nullptr, nullptr, nullptr) != HSA_STATUS_SUCCESS ||
            pi.type != HSA_EXT_POINTER_TYPE_HSA)
            return false;
        gpu_agent = pi.agentOwner;
        if (hsa_iterate_agents(find_cpu_agent, &cpu_agent) != HSA_STATUS_INFO_BREAK) return false;
        uint32_t mask = 0;
        if (hsa_amd_memory_copy_engine_status(cpu_agent, gpu_agent, &mask) != HSA_STATUS_SUCCESS ||
            mask == 0)
            return false;
        engine = (hsa_amd_sdma_engine_id_t)(mask & (~mask + 1));  // lowest available
        for (hsa_signal_t& sg : csig)
            if (hsa_signal_create(0, 0, nullptr, &sg) != HSA_STATUS_SUCCESS) return false;
        if (hipHostMalloc((void**)&stage, 2 * kSrvStageRecords * sizeof(pp_hit), 0) != hipSuccess) {
            stage = nullptr;
            return false;
        }
        // one record through the engine and back to the host
        if (!sdma_issue(0, (const pp_hit*)pool.p, 1) || !sdma_wait(0, 2000)) return false;
        sdma = true;
        return true;
    }

    int copy_hits(uint64_t base, uint64_t n, pp_hit* out) {
        if (host_pool) {  // the kernel wrote them to host memory; seen through the completion tag
            parallel_memcpy(reinterpret_cast<char*>(out), reinterpret_cast<const char*>(host_pool + base),
                            n * sizeof(pp_hit));
            return PP_OK;
        }
        // pool -> pinned staging on the SDMA engine, staging -> caller on the
        // host, the two halves alternating so that both overlap
        std::lock_guard<std::mutex> cg(copy_mu);
        const pp_hit* src = (const pp_hit*)pool.p + base;
        uint64_t off = 0;
        int h = 0;
        if (!sdma_issue(0, src, std::min(n, kSrvStageRecords)))
            return fail(PP_ERR_DEVICE, "SDMA copy of result records failed to start");
        while (off < n) {
            const uint64_t m = std::min(n - off, kSrvStageRecords);
            if (!sdma_wait(h, 60000)) return fail(PP_ERR_DEVICE, "SDMA copy of result records failed");
            const uint64_t next = off + m;
            if (next < n && !sdma_issue(h ^ 1, src + next, std::min(n - next, kSrvStageRecords)))
                return fail(PP_ERR_DEVICE, "SDMA copy of result records failed to start");
            parallel_memcpy(reinterpret_cast<char*>(out + off),
                            reinterpret_cast<const char*>(stage + (uint64_t)h * kSrvStageRecords),
                            m * sizeof(pp_hit));
            off = next;
            h ^= 1;
        }
        return PP_OK;
    }

    int setup(pp_ctx* c, const pp_spectrum* spectrum, const pp_server_config& config) {
        ctx = c;
        cfg = config;
        int rc;
        PP_HIP(hipStreamCreateWithFlags(&kstream, hipStreamNonBlocking));
        PP_HIP(hipEventCreateWithFlags(&kdone, hipEventDisableTiming));
        if ((rc = check_spectrum(spectrum, &spec))) return rc;
        if ((rc = ensure(knots, PP_MAX_KNOTS * sizeof(float)))) return rc;
        if ((rc = upload_knots(spectrum, (float*)knots.p, c->stream))) return rc;
        // the server keeps its own copy of the medium table (pp_set_medium on
        // the context does not change requests already being served)
        if ((rc = ensure(table, 2 * (size_t)c->n_points * sizeof(float4)))) return rc;
        PP_HIP(hipMemcpyAsync(table.p, c->table.p, 2 * (size_t)c->n_points * sizeof(float4),
                              hipMemcpyDeviceToDevice, c->stream));
        n_points = c->n_points;
        lmin = c->lmin;
        inv_dl = c->inv_dl;
        g = c->g;
        PP_HIP(hipStreamSynchronize(c->stream));
        const uint32_t S = cfg.max_slots;
        PP_HIP(hipHostMalloc((void**)&ctl, sizeof(SrvControl), hipHostMallocCoherent));
        PP_HIP(hipHostMalloc((void**)&entries, kSrvRing * sizeof(SrvEntry), hipHostMallocCoherent));
        PP_HIP(hipHostMalloc((void**)&slots, S * sizeof(SrvSlot), hipHostMallocCoherent));
        PP_HIP(hipHostMalloc((void**)&done, S * sizeof(SrvDone), hipHostMallocCoherent));
        memset(ctl, 0, sizeof(SrvControl));
        memset(entries, 0, kSrvRing * sizeof(SrvEntry));
        memset(slots, 0, S * sizeof(SrvSlot));
        memset(done, 0, S * sizeof(SrvDone));
        if ((rc = ensure(dctl, 3 * kSrvGateOffset))) return rc;
        if ((rc = ensure(dentries, kSrvRing * sizeof(SrvEntry)))) return rc;
        if ((rc = ensure(tiles_done, S * sizeof(uint32_t)))) return rc;
        if ((rc = ensure(hit_cursor, S * sizeof(unsigned long long)))) return rc;
        if ((rc = ensure(partials, (size_t)S * cfg.max_tiles_per_request * sizeof(TilePartial))))
            return rc;
        fill_u32_kernel<<<8, 256, 0, kstream>>>((uint32_t*)dctl.p, 3 * kSrvGateOffset / 4, 0u);
        fill_u32_kernel<<<64, 256, 0, kstream>>>((uint32_t*)tiles_done.p, S, 0u);
        fill_u32_kernel<<<64, 256, 0, kstream>>>((uint32_t*)hit_cursor.p, 2ull * S, 0u);
        PP_HIP(hipGetLastError());
        PP_HIP(hipStreamSynchronize(kstream));
        // the hit pool: device memory with result copies on an SDMA engine,
        // else (or when asked for) pinned host memory
        const uint64_t pool_bytes = std::max<uint64_t>(cfg.hit_pool_records, 1) * sizeof(pp_hit);
        if (!(cfg.flags & PP_SERVER_HOST_POOL)) {
            if ((rc = ensure(pool, pool_bytes))) return rc;
            if (!setup_sdma()) {
                (void)hipFree(pool.p);
                pool.p = nullptr;
                pool.cap = 0;
            }
        }
        if (!sdma) PP_HIP(hipHostMalloc((void**)&host_pool, pool_bytes, hipHostMallocCoherent));
        // the kernel's own stop after device_wait_us with nothing published
        int rate_khz = 0;
        PP_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device));
        const uint64_t wait_us =
            cfg.device_wait_us ? cfg.device_wait_us
                               : std::max<uint64_t>(4ull * cfg.idle_exit_us, 1000000ull);
        wait_ticks = std::max<uint64_t>(1, wait_us * (uint64_t)std::max(rate_khz, 1) / 1000);
        SrvAddr A;
        memset(&A, 0, sizeof A);
        A.ctl = ctl;
        A.entries = entries;
        A.slots = slots;
        A.done = done;
        A.ticket = (unsigned long long*)dctl.p;
        A.gate = (unsigned long long*)((char*)dctl.p + kSrvGateOffset);
        A.dmeta = (unsigned long long*)((char*)dctl.p + 2 * kSrvGateOffset);
        A.dentries = (SrvEntry*)dentries.p;
        A.tiles_done = (uint32_t*)tiles_done.p;
        A.hit_cursor = (unsigned long long*)hit_cursor.p;
        A.partials = (TilePartial*)partials.p;
        A.pool = host_pool ? host_pool : (pp_hit*)pool.p;
        A.wait_ticks = wait_ticks;
        A.max_tiles = cfg.max_tiles_per_request;
        A.tile = make_tile_ctx((const float4*)table.p, n_points, (const float*)knots.p, spec,
                               lmin, inv_dl, make_hg(g));
        if ((rc = ensure(addr, sizeof A))) return rc;
        PP_HIP(hipMemcpy(addr.p, &A, sizeof A, hipMemcpyHostToDevice));
        // The grid fills the device, or leaves room beside it on every SIMD for
        // the caller's other kernels (32 VGPRs and one of the 8 wave slots).
        int per_cu = 0;
        PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_kernel, kBlock, 0));
        hipFuncAttributes fa;
        PP_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(serve_kernel)));
        const int vgpr = std::max(8, (fa.numRegs + 7) & ~7);  // allocation granule of 8
        if (cfg.flags & PP_SERVER_LEAVE_ROOM)
            per_cu = std::min({per_cu, (512 - kSrvVgprReserve) / vgpr, kSrvRoomWavesPerSimd});
        grid_ = std::max(1, per_cu) * c->n_cu;
        return PP_OK;
    }
};

}  // namespace

struct pp_server {
    pps::ServerCore<HipBackend> core;
    float n_min = 0.0f;  // smallest n of the server's table snapshot (Cherenkov check)
    explicit pp_server(const pp_server_config& cfg) : core(cfg) {}
};

extern "C" {

int pp_server_create(pp_ctx* c, const pp_spectrum* spectrum, const pp_server_config* cfg,
                     pp_server** out) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (!cfg || !out) return fail(PP_ERR_INVALID_ARG, "cfg/out is NULL");
    *out = nullptr;
    if (cfg->max_slots < 1 || cfg->max_slots > 4096)
        return fail(PP_ERR_INVALID_ARG, "max_slots %u not in [1, 4096]", cfg->max_slots);
    if ((rc = check_records(cfg->hit_pool_records, "hit_pool_records"))) return rc;
    if (cfg->max_tiles_per_request < 1 || cfg->max_tiles_per_request > (1u << 20))
        return fail(PP_ERR_INVALID_ARG, "max_tiles_per_request %u not in [1, 2^20]",
                    cfg->max_tiles_per_request);
    if (cfg->flags & ~(PP_SERVER_HOST_POOL | PP_SERVER_LEAVE_ROOM))
        return fail(PP_ERR_INVALID_ARG, "unknown server flags 0x%x", cfg->flags);
    pp_server_config k = *cfg;
    if (!k.idle_exit_us) k.idle_exit_us = 20000;
    k.idle_exit_us = std::min<uint32_t>(k.idle_exit_us, 10000000u);
    pp_server* s = new pp_server(k);
    s->n_min = c->n_min;
    if ((rc = s->core.be.setup(c, spectrum, k)) || (rc = s->core.start())) {
        std::string msg = last_error();
        delete s;
        last_error() = msg;
        return rc;
    }
    *out = s;
    return PP_OK;
}

int pp_server_submit(pp_server* s, const pp_request* request, uint64_t hits_cap,
                     uint64_t* ticket_out) {
    if (!s || !request || !ticket_out) return fail(PP_ERR_INVALID_ARG, "NULL argument");
    DevReq dr;
    int rc = make_devreq(request, 0, &dr);
    if (rc) return rc;
    // against the server's own table snapshot, not the context's current medium
    if ((rc = check_cone(request, 1, s->n_min))) return rc;
    return s->core.submit(dr, hits_cap, ticket_out);
}

int pp_server_wait(pp_server* s, uint64_t ticket, uint32_t timeout_ms, int* done_out) {
    if (!s || !done_out) return fail(PP_ERR_INVALID_ARG, "NULL argument");
    return s->core.wait(ticket, timeout_ms, done_out);
}

int pp_server_result(pp_server* s, uint64_t ticket, pp_summary* summary, pp_hit* hits_out,
                     uint64_t hits_cap, uint64_t* n_hits_out, double* latency_us) {
    if (!s) return fail(PP_ERR_INVALID_ARG, "server is NULL");
    if (hits_cap && !hits_out) return fail(PP_ERR_INVALID_ARG, "hits_out is NULL with hits_cap > 0");
    return s->core.result(ticket, summary, hits_out, hits_cap, n_hits_out, latency_us);
}

int pp_server_stats_get(pp_server* s, pp_server_stats* out) {
    if (!s || !out) return fail(PP_ERR_INVALID_ARG, "NULL argument");
    s->core.stats(out);
    out->host_pool = s->core.be.host_pool ? 1 : 0;
    return PP_OK;
}

// Drain: stop admitting, let the kernel finish in-flight work and exit.
int pp_server_destroy(pp_server* s) {
    if (!s) return PP_OK;
    s->core.shutdown();
    delete s;
    return PP_OK;
}

}  // extern "C"
