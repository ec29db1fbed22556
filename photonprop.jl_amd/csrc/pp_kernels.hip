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
//     (block, photon_lo, photon_hi, stream).  Block 0 holds the emission draws
//     {λ, φ, θ, -}; step s uses words 3s..3s+2 of blocks 1, 2, ...  Steps run in
//     groups of four that consume exactly three Philox blocks, so the word a
//     step reads is fixed at compile time and lives in registers.
//   * Medium tables {n, λ_sca, λ_abs} and the spectrum knots sit in LDS.
//   * Hits: wavefront ballot + mbcnt rank, one global atomic per wave that has
//     any hit, 64-byte records written with 16-byte stores.
//   * Sums: per-lane FP64 accumulators, fixed-order wave butterfly and block
//     combine into one partial per tile; a second kernel adds a request's tile
//     partials in tile order.  The grouping depends only on the request, so a
//     request's sums are bitwise the same alone or inside any batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/photonprop.h"
#include "pp_device.hpp"

namespace {

constexpr int kBlock = 256;                       // lanes per block (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kRoundsPerTile = 32;            // photons per lane per tile
constexpr uint64_t kTilePhotons = (uint64_t)kBlock * kRoundsPerTile;  // 8192
constexpr double kCVacuum = 0.299792458;           // m/ns
constexpr float kLn2 = 0.69314718055994530942f;
constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kInv2Pi = 0.15915494309189533577f;

// Request as the kernel reads it (host-precomputed).
struct DevReq {
    float pos[3];
    float t0;
    float axis[3];
    float beta;
    float center[3];
    float r2;
    uint64_t n_photons;
    uint32_t profile;
    uint32_t max_steps;
    uint32_t seed_lo, seed_hi, stream, pad;
};

struct TilePartial {
    uint64_t n_steps;
    uint64_t n_hits;
    double sum_w, sum_wt, sum_wt2;
};

struct SpecParams {
    uint32_t kind;
    uint32_t n_knots;
    float a;          // INV_LAMBDA2: 1/lo
    float d;          // INV_LAMBDA2: 1/lo - 1/hi
    float mono;
    float knot_step;  // KNOTS: 1/(n_knots-1)
};

struct KParams {
    const DevReq* reqs;
    const uint64_t* tile_start;  // [n_req+1]
    uint32_t n_req;
    uint32_t n_points;
    uint64_t n_tiles;
    uint32_t* tile_counter;
    const float4* table;
    const float* knots;
    float lmin;
    float inv_dl;
    ppd::HG hg;
    SpecParams spec;
    pp_hit* hits;
    uint64_t hits_cap;
    unsigned long long* hit_cursor;
    TilePartial* partials;
};

// ---------------------------------------------------------------------------
// Propagation kernel
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void __launch_bounds__(kBlock) propagate_kernel(KParams P) {
    extern __shared__ float4 s_table[];  // n_points float4, then knots
    __shared__ uint32_t s_tile;
    __shared__ double s_red[kWaves][3];
    __shared__ uint32_t s_redi[kWaves][2];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    float* s_knots = reinterpret_cast<float*>(s_table + P.n_points);

    for (uint32_t i = tid; i < P.n_points; i += kBlock) s_table[i] = P.table[i];
    if (P.spec.kind == PP_SPEC_KNOTS)
        for (uint32_t i = tid; i < P.spec.n_knots; i += kBlock) s_knots[i] = P.knots[i];

    const ppd::HG hg = P.hg;
    const SpecParams spec = P.spec;

    for (;;) {
        __syncthreads();  // tables staged / previous tile's s_tile consumed
        if (tid == 0) s_tile = atomicAdd(P.tile_counter, 1u);
        __syncthreads();
        const uint32_t tile = __builtin_amdgcn_readfirstlane(s_tile);
        if ((uint64_t)tile >= P.n_tiles) break;

        // Request owning this tile: last r with tile_start[r] <= tile.
        uint32_t lo = 0, hi = P.n_req;  // invariant: tile_start[lo] <= tile < tile_start[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (P.tile_start[mid] <= (uint64_t)tile) lo = mid; else hi = mid;
        }
        const uint32_t r = __builtin_amdgcn_readfirstlane(lo);
        const DevReq R = P.reqs[r];
        const uint64_t first = ((uint64_t)tile - P.tile_start[r]) * kTilePhotons;
        const uint64_t left = R.n_photons - first;
        const uint32_t count = (uint32_t)(left < kTilePhotons ? left : kTilePhotons);
        const uint32_t rounds = (count + kBlock - 1) / kBlock;
        const uint32_t max_steps = R.max_steps;
        const float r2 = R.r2;
        const float cx = R.center[0], cy = R.center[1], cz = R.center[2];

        uint32_t acc_steps = 0, acc_hits = 0;
        double acc_w = 0.0, acc_wt = 0.0, acc_wt2 = 0.0;

        for (uint32_t k = 0; k < rounds; ++k) {
            const uint32_t local = k * kBlock + tid;
            const bool active = local < count;
            const uint64_t pid = first + local;
            const uint32_t id_lo = (uint32_t)pid, id_hi = (uint32_t)(pid >> 32);

            // ---- emission (src/photon_prop_cuda.jl:308-316) ----
            const ppd::U4 w0 = ppd::philox10(0u, id_lo, id_hi, R.stream, R.seed_lo, R.seed_hi);
            const float u_l = ppd::u01(w0.x);
            float lambda;
            if (spec.kind == PP_SPEC_MONO) {
                lambda = spec.mono;
            } else if (spec.kind == PP_SPEC_KNOTS) {
                lambda = ppd::knot_interp(s_knots, spec.n_knots, spec.knot_step, u_l);
            } else {
                lambda = 1.0f / __builtin_fmaf(-u_l, spec.d, spec.a);
            }
            const float4 med = ppd::table_lerp(s_table, P.n_points, (lambda - P.lmin) * P.inv_dl);
            const float n_phase = med.x;
            const float sca_ln2 = med.y * kLn2;
            const float abs_len = med.z;

            float dx, dy, dz;
            {
                const float u_phi = ppd::u01(w0.y);
                const float sp = __builtin_amdgcn_sinf(u_phi);
                const float cp = __builtin_amdgcn_cosf(u_phi);
                const float u_th = ppd::u01(w0.z);
                if (R.profile == PP_EMIT_REF_ISOTROPIC) {
                    const float theta = __builtin_fmaf(u_th, 2.0f, -1.0f);
                    const float st = __builtin_amdgcn_sinf(theta * kInv2Pi);
                    const float ct = __builtin_amdgcn_cosf(theta * kInv2Pi);
                    dx = cp * st;
                    dy = sp * st;
                    dz = ct;
                } else {
                    const float ct = R.profile == PP_EMIT_ISOTROPIC
                                         ? __builtin_fmaf(2.0f, u_th, -1.0f)
                                         : fminf(1.0f / (n_phase * R.beta), 1.0f);
                    const float st = sqrtf(__builtin_fmaf(-ct, ct, 1.0f));
                    ppd::rotate_to_axis(R.axis[0], R.axis[1], R.axis[2], cp * st, sp * st, ct, dx,
                                        dy, dz);
                }
            }
            const float idx = dx, idy = dy, idz = dz;
            float px = R.pos[0], py = R.pos[1], pz = R.pos[2];
            float dist = 0.0f;
            bool alive = active;
            bool hit = false;
            uint32_t nsteps = 0;

            // One scattering-loop iteration (src/photon_prop_cuda.jl:320-376, D3).
#define PP_STEP(WA, WB, WC)                                                              \
    if (alive) {                                                                         \
        const float u1 = ppd::u01(WA);                                                   \
        float s = -__builtin_amdgcn_logf(u1) * sca_ln2;                                  \
        const float rx = px - cx, ry = py - cy, rz = pz - cz;                            \
        const float a = __builtin_fmaf(dz, rz, __builtin_fmaf(dy, ry, dx * rx));         \
        const float pp = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));        \
        const float b = __builtin_fmaf(a, a, r2 - pp);                                   \
        ++nsteps;                                                                        \
        bool h = false;                                                                  \
        if (b >= 0.0f) {                                                                 \
            const float d = -a - sqrtf(b);                                               \
            if (d > 0.0f && d < s) {                                                     \
                h = true;                                                                \
                s = d;                                                                   \
            }                                                                            \
        }                                                                                \
        px = __builtin_fmaf(s, dx, px);                                                  \
        py = __builtin_fmaf(s, dy, py);                                                  \
        pz = __builtin_fmaf(s, dz, pz);                                                  \
        dist += s;                                                                       \
        if (h) {                                                                         \
            hit = true;                                                                  \
            alive = false;                                                               \
        } else {                                                                         \
            const float ct = ppd::hg_costheta(hg, ppd::u01(WB));                         \
            const float st = __builtin_amdgcn_sqrtf(__builtin_fmaf(-ct, ct, 1.0f));      \
            const float uph = ppd::u01(WC);                                              \
            const float sph = __builtin_amdgcn_sinf(uph);                                \
            const float cph = __builtin_amdgcn_cosf(uph);                                \
            ppd::rotate_to_axis(dx, dy, dz, cph * st, sph * st, ct, dx, dy, dz);        \
        }                                                                                \
    }

            for (uint32_t s0 = 0; s0 < max_steps; s0 += 4) {
                if (__ballot(alive) == 0) break;
                const uint32_t blk = 1u + 3u * (s0 >> 2);
                const ppd::U4 A = ppd::philox10(blk, id_lo, id_hi, R.stream, R.seed_lo, R.seed_hi);
                PP_STEP(A.x, A.y, A.z)
                if (s0 + 1 >= max_steps || __ballot(alive) == 0) break;
                const ppd::U4 B =
                    ppd::philox10(blk + 1u, id_lo, id_hi, R.stream, R.seed_lo, R.seed_hi);
                PP_STEP(A.w, B.x, B.y)
                if (s0 + 2 >= max_steps || __ballot(alive) == 0) break;
                const ppd::U4 C =
                    ppd::philox10(blk + 2u, id_lo, id_hi, R.stream, R.seed_lo, R.seed_hi);
                PP_STEP(B.z, B.w, C.x)
                if (s0 + 3 >= max_steps) break;
                PP_STEP(C.y, C.z, C.w)
            }
#undef PP_STEP
            acc_steps += nsteps;

            // ---- hit append: one atomic per wave with any hit ----
            const uint64_t mask = __ballot(hit);
            if (mask != 0) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(P.hit_cursor, (unsigned long long)__popcll(mask));
                base = __shfl(base, 0, 64);
                if (hit) {
                    const float e = -dist / abs_len;
                    const float wgt = __builtin_amdgcn_exp2f(e * kLog2e);
                    const float t = R.t0 + dist * n_phase / (float)kCVacuum;
                    acc_hits += 1;
                    acc_w += (double)wgt;
                    acc_wt += (double)wgt * (double)t;
                    acc_wt2 += (double)wgt * (double)t * (double)t;
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                    const unsigned long long slot = base + rank;
                    if (slot < P.hits_cap) {
                        float4* o = reinterpret_cast<float4*>(P.hits + slot);
                        o[0] = make_float4(px, py, pz, t);
                        o[1] = make_float4(dx, dy, dz, lambda);
                        o[2] = make_float4(idx, idy, idz, wgt);
                        uint4 tail;
                        tail.x = __float_as_uint(dist);
                        tail.y = r;
                        tail.z = id_lo;
                        tail.w = id_hi;
                        reinterpret_cast<uint4*>(o)[3] = tail;
                    }
                }
            }
        }

        // ---- tile partial: fixed-order reduction ----
        const uint32_t ws = wave_sum_u32(acc_steps);
        const uint32_t wh = wave_sum_u32(acc_hits);
        const double w1 = wave_sum_f64(acc_w);
        const double w2 = wave_sum_f64(acc_wt);
        const double w3 = wave_sum_f64(acc_wt2);
        if (lane == 0) {
            s_redi[wave][0] = ws;
            s_redi[wave][1] = wh;
            s_red[wave][0] = w1;
            s_red[wave][1] = w2;
            s_red[wave][2] = w3;
        }
        __syncthreads();
        if (tid == 0) {
            TilePartial tp;
            tp.n_steps = 0;
            tp.n_hits = 0;
            tp.sum_w = 0.0;
            tp.sum_wt = 0.0;
            tp.sum_wt2 = 0.0;
            for (int q = 0; q < kWaves; ++q) {
                tp.n_steps += s_redi[q][0];
                tp.n_hits += s_redi[q][1];
                tp.sum_w += s_red[q][0];
                tp.sum_wt += s_red[q][1];
                tp.sum_wt2 += s_red[q][2];
            }
            P.partials[tile] = tp;
        }
    }
}

// Per-request sums over its tiles in tile order (one wave per request).
__global__ void __launch_bounds__(64) reduce_kernel(const TilePartial* partials,
                                                    const uint64_t* tile_start,
                                                    const DevReq* reqs, uint32_t n_req,
                                                    pp_summary* out) {
    const uint32_t r = blockIdx.x;
    if (r >= n_req) return;
    const uint32_t lane = threadIdx.x;
    const uint64_t b = tile_start[r], e = tile_start[r + 1];
    uint64_t st = 0, nh = 0;
    double sw = 0.0, swt = 0.0, swt2 = 0.0;
    for (uint64_t t = b + lane; t < e; t += 64) {
        const TilePartial p = partials[t];
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
    if (lane == 0) {
        pp_summary s;
        s.n_sim = reqs[r].n_photons;
        s.n_steps = st;
        s.n_hits = nh;
        s.n_stored = 0;
        s.sum_w = sw;
        s.sum_wt = swt;
        s.sum_wt2 = swt2;
        s.status = PP_OK;
        s.reserved = 0;
        out[r] = s;
    }
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Stored hits per request.
__global__ void count_kernel(const pp_hit* staged, const unsigned long long* cursor,
                             uint64_t cap, uint32_t* cnt) {
    const uint64_t n = *cursor < cap ? *cursor : cap;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[staged[i].request], 1u);
}

// Exclusive scan of cnt into offsets[0..n], single block of 1024 threads;
// also records n_stored and the overflow status per request.
__global__ void __launch_bounds__(1024) scan_kernel(const uint32_t* cnt, uint32_t n,
                                                    uint64_t* offsets, pp_summary* sums) {
    __shared__ uint64_t s_part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t b = t * per;
    const uint32_t e = std::min(n, b + per);
    uint64_t local = 0;
    for (uint32_t i = b; i < e; ++i) local += cnt[i];
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
        run += cnt[i];
        sums[i].n_stored = cnt[i];
        sums[i].status = sums[i].n_hits > (uint64_t)cnt[i] ? PP_ERR_OVERFLOW : PP_OK;
    }
    if (t == 1023) offsets[n] = s_part[1023];
}

__global__ void scatter_kernel(const pp_hit* staged, const unsigned long long* cursor,
                               uint64_t cap, const uint64_t* offsets, uint32_t* fill,
                               pp_hit* out) {
    const uint64_t n = *cursor < cap ? *cursor : cap;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(staged + i);
        const uint4 a = src[0], b = src[1], c = src[2], d = src[3];
        const uint32_t r = d.y;
        const uint64_t slot = offsets[r] + atomicAdd(&fill[r], 1u);
        uint4* dst = reinterpret_cast<uint4*>(out + slot);
        dst[0] = a;
        dst[1] = b;
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
    out[i] = ppd::hg_costheta(hg, ppd::u01(hg_word(i, k0, k1)));
}

__global__ void __launch_bounds__(256) selftest_hg_moments_kernel(ppd::HG hg, uint32_t k0,
                                                                  uint32_t k1, uint64_t n,
                                                                  double* partial) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const double c = (double)ppd::hg_costheta(hg, ppd::u01(hg_word(i, k0, k1)));
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

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define PP_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(e_ == hipErrorOutOfMemory ? PP_ERR_OOM : PP_ERR_DEVICE,           \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,  \
                        __LINE__);                                                        \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

int ensure(DevBuf& b, size_t bytes) {
    if (bytes <= b.cap) return PP_OK;
    if (b.p) PP_HIP(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    PP_HIP(hipMalloc(&b.p, want));
    b.cap = want;
    return PP_OK;
}

}  // namespace

struct pp_ctx {
    int device = 0;
    int n_cu = 0;
    int blocks_per_cu = 1;
    char name[256] = {0};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    bool timed = false;
    // medium
    DevBuf table;
    uint32_t n_points = 0;
    float lmin = 0.0f, lmax = 0.0f, inv_dl = 0.0f, g = 0.0f, n_min = 0.0f;
    DevBuf knots;
    // work buffers
    DevBuf reqs, tile_start, partials, counters, staging, cnt, fill, sums, offsets, out_hits;
    DevBuf scratch;
    std::vector<DevReq> h_reqs;
    std::vector<uint64_t> h_tiles;
};

namespace {

int check_medium(const pp_medium_tables* m) {
    if (!m) return fail(PP_ERR_INVALID_ARG, "medium is NULL");
    if (m->n_points < 2 || m->n_points > PP_MAX_TABLE_POINTS)
        return fail(PP_ERR_INVALID_ARG, "medium n_points=%u outside [2, %d]", m->n_points,
                    PP_MAX_TABLE_POINTS);
    if (!(m->lambda_max > m->lambda_min) || !std::isfinite(m->lambda_min) ||
        !std::isfinite(m->lambda_max))
        return fail(PP_ERR_INVALID_ARG, "medium wavelength range [%g, %g] invalid",
                    (double)m->lambda_min, (double)m->lambda_max);
    if (!m->n_phase || !m->abs_len || !m->sca_len)
        return fail(PP_ERR_INVALID_ARG, "medium table pointer is NULL");
    if (!(std::fabs(m->g) < 1.0f)) return fail(PP_ERR_INVALID_ARG, "medium g=%g not in (-1,1)", (double)m->g);
    for (uint32_t i = 0; i < m->n_points; ++i) {
        if (!(m->abs_len[i] > 0.0f) || !(m->sca_len[i] > 0.0f) || !(m->n_phase[i] > 0.0f) ||
            !std::isfinite(m->abs_len[i]) || !std::isfinite(m->sca_len[i]) ||
            !std::isfinite(m->n_phase[i]))
            return fail(PP_ERR_INVALID_ARG, "medium table entry %u not positive and finite", i);
    }
    return PP_OK;
}

int upload_medium(pp_ctx* c, const pp_medium_tables* m) {
    int rc = check_medium(m);
    if (rc) return rc;
    std::vector<float4> t(m->n_points);
    for (uint32_t i = 0; i < m->n_points; ++i)
        t[i] = make_float4(m->n_phase[i], m->sca_len[i], m->abs_len[i], 0.0f);
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

// Validate a spectrum and fill the kernel parameters; knots go to the device.
int prepare_spectrum(pp_ctx* c, const pp_spectrum* s, SpecParams* out) {
    if (!s) return fail(PP_ERR_INVALID_ARG, "spectrum is NULL");
    memset(out, 0, sizeof *out);
    out->kind = s->kind;
    switch (s->kind) {
    case PP_SPEC_MONO:
        if (!(s->mono > 0.0f) || !std::isfinite(s->mono))
            return fail(PP_ERR_INVALID_ARG, "monochromatic wavelength %g invalid", (double)s->mono);
        out->mono = s->mono;
        return PP_OK;
    case PP_SPEC_INV_LAMBDA2:
        if (!(s->lo > 0.0f) || !(s->hi > s->lo) || !std::isfinite(s->hi))
            return fail(PP_ERR_INVALID_ARG, "spectrum range [%g, %g] invalid", (double)s->lo,
                        (double)s->hi);
        out->a = (float)(1.0 / (double)s->lo);
        out->d = (float)(1.0 / (double)s->lo - 1.0 / (double)s->hi);
        return PP_OK;
    case PP_SPEC_KNOTS: {
        if (s->n_knots < 2 || s->n_knots > PP_MAX_KNOTS || !s->knots)
            return fail(PP_ERR_INVALID_ARG, "spectrum knots: n_knots=%u (need 2..%d) or NULL",
                        s->n_knots, PP_MAX_KNOTS);
        for (uint32_t i = 0; i < s->n_knots; ++i)
            if (!(s->knots[i] > 0.0f) || !std::isfinite(s->knots[i]))
                return fail(PP_ERR_INVALID_ARG, "spectrum knot %u = %g invalid", i, (double)s->knots[i]);
        int rc = ensure(c->knots, PP_MAX_KNOTS * sizeof(float));
        if (rc) return rc;
        PP_HIP(hipMemcpyAsync(c->knots.p, s->knots, s->n_knots * sizeof(float),
                              hipMemcpyHostToDevice, c->stream));
        out->n_knots = s->n_knots;
        out->knot_step = 1.0f / (float)(s->n_knots - 1);
        return PP_OK;
    }
    default:
        return fail(PP_ERR_UNSUPPORTED, "Cannot initialize spectrum kind %u", s->kind);
    }
}

int make_devreq(const pp_request* q, uint32_t idx, DevReq* d) {
    const pp_emitter& e = q->emitter;
    if (e.profile > PP_EMIT_REF_ISOTROPIC)
        return fail(PP_ERR_UNSUPPORTED, "request %u: Cannot initialize emission profile %u", idx,
                    e.profile);
    for (int j = 0; j < 3; ++j)
        if (!std::isfinite(e.pos[j]) || !std::isfinite(e.axis[j]) ||
            !std::isfinite(q->target.center[j]))
            return fail(PP_ERR_INVALID_ARG, "request %u: non-finite position/axis/centre", idx);
    if (!std::isfinite(e.t0)) return fail(PP_ERR_INVALID_ARG, "request %u: non-finite t0", idx);
    if (!(q->target.radius > 0.0f) || !std::isfinite(q->target.radius))
        return fail(PP_ERR_INVALID_ARG, "request %u: target radius %g must be > 0", idx,
                    (double)q->target.radius);
    const double an = std::sqrt((double)e.axis[0] * e.axis[0] + (double)e.axis[1] * e.axis[1] +
                                (double)e.axis[2] * e.axis[2]);
    if (!(an > 0.0)) return fail(PP_ERR_INVALID_ARG, "request %u: zero emitter axis", idx);
    if (e.profile == PP_EMIT_CHERENKOV_CONE && !(e.beta > 0.0f && std::isfinite(e.beta)))
        return fail(PP_ERR_INVALID_ARG, "request %u: beta %g must be > 0", idx, (double)e.beta);
    if (q->max_steps == 0) return fail(PP_ERR_INVALID_ARG, "request %u: max_steps must be >= 1", idx);
    if (q->max_steps > (1u << 28))
        return fail(PP_ERR_INVALID_ARG, "request %u: max_steps %u too large", idx, q->max_steps);
    memset(d, 0, sizeof *d);
    for (int j = 0; j < 3; ++j) {
        d->pos[j] = e.pos[j];
        d->axis[j] = (float)(e.axis[j] / an);
        d->center[j] = q->target.center[j];
    }
    d->t0 = e.t0;
    d->beta = e.beta;
    d->r2 = q->target.radius * q->target.radius;
    d->n_photons = e.n_photons;
    d->profile = e.profile;
    d->max_steps = q->max_steps;
    d->seed_lo = (uint32_t)q->seed;
    d->seed_hi = (uint32_t)(q->seed >> 32);
    d->stream = q->stream;
    return PP_OK;
}

// Cherenkov cone needs n(λ)·β >= 1 everywhere on the table.
int check_cone(const pp_ctx* c, const pp_request* reqs, uint32_t n_req, float n_min) {
    for (uint32_t i = 0; i < n_req; ++i) {
        if (reqs[i].emitter.profile != PP_EMIT_CHERENKOV_CONE) continue;
        if (!((double)n_min * (double)reqs[i].emitter.beta >= 1.0 - 1e-6))
            return fail(PP_ERR_INVALID_ARG,
                        "request %u: no Cherenkov emission (n_min=%g, beta=%g, n*beta < 1)", i,
                        (double)n_min, (double)reqs[i].emitter.beta);
    }
    (void)c;
    return PP_OK;
}


size_t dyn_lds_bytes(const pp_ctx* c, const SpecParams& sp) {
    return (size_t)c->n_points * sizeof(float4) +
           (sp.kind == PP_SPEC_KNOTS ? (size_t)sp.n_knots * sizeof(float) : 0);
}

int grid_for(pp_ctx* c, size_t lds, uint64_t n_tiles, int* grid) {
    int per_cu = 0;
    PP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, propagate_kernel, kBlock, lds));
    if (per_cu < 1) per_cu = 1;
    c->blocks_per_cu = per_cu;
    const uint64_t resident = (uint64_t)per_cu * (uint64_t)c->n_cu;
    *grid = (int)std::min<uint64_t>(n_tiles, resident);
    return PP_OK;
}

// Enqueue on c->stream: zero the counters, run the persistent propagation
// kernel over all tiles of all requests, reduce per request into c->sums.
// Hit records go to d_staging (capacity hits_cap).
int enqueue_propagation(pp_ctx* c, const pp_request* reqs, uint32_t n_req,
                        const pp_spectrum* spec, pp_hit* d_staging, uint64_t hits_cap) {
    if (n_req == 0) return fail(PP_ERR_INVALID_ARG, "no requests");
    if (!reqs) return fail(PP_ERR_INVALID_ARG, "requests is NULL");
    SpecParams sp;
    int rc = prepare_spectrum(c, spec, &sp);
    if (rc) return rc;
    if ((rc = check_cone(c, reqs, n_req, c->n_min))) return rc;
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
    if ((rc = ensure(c->reqs, n_req * sizeof(DevReq)))) return rc;
    if ((rc = ensure(c->tile_start, (n_req + 1) * sizeof(uint64_t)))) return rc;
    if ((rc = ensure(c->partials, std::max<uint64_t>(n_tiles, 1) * sizeof(TilePartial)))) return rc;
    if ((rc = ensure(c->counters, 64))) return rc;
    if ((rc = ensure(c->sums, n_req * sizeof(pp_summary)))) return rc;
    PP_HIP(hipMemcpyAsync(c->reqs.p, c->h_reqs.data(), n_req * sizeof(DevReq),
                          hipMemcpyHostToDevice, c->stream));
    PP_HIP(hipMemcpyAsync(c->tile_start.p, c->h_tiles.data(), (n_req + 1) * sizeof(uint64_t),
                          hipMemcpyHostToDevice, c->stream));
    // counters: [0] tile counter (u32), [2..3] hit cursor (u64)
    fill_u32_kernel<<<1, 64, 0, c->stream>>>((uint32_t*)c->counters.p, 16, 0u);
    PP_HIP(hipGetLastError());

    const size_t lds = dyn_lds_bytes(c, sp);
    int grid = 0;
    if ((rc = grid_for(c, lds, n_tiles, &grid))) return rc;
    KParams P;
    memset(&P, 0, sizeof P);
    P.reqs = (const DevReq*)c->reqs.p;
    P.tile_start = (const uint64_t*)c->tile_start.p;
    P.n_req = n_req;
    P.n_points = c->n_points;
    P.n_tiles = n_tiles;
    P.tile_counter = (uint32_t*)c->counters.p;
    P.table = (const float4*)c->table.p;
    P.knots = (const float*)c->knots.p;
    P.lmin = c->lmin;
    P.inv_dl = c->inv_dl;
    P.hg.g = c->g;
    P.hg.iso = c->g == 0.0f;
    P.hg.num = std::fma(-c->g, c->g, 1.0f);
    P.hg.opg2 = std::fma(c->g, c->g, 1.0f);
    P.hg.inv2g = P.hg.iso ? 0.0f : 1.0f / (2.0f * c->g);
    P.spec = sp;
    P.hits = d_staging;
    P.hits_cap = hits_cap;
    P.hit_cursor = (unsigned long long*)((uint32_t*)c->counters.p + 2);
    P.partials = (TilePartial*)c->partials.p;
    PP_HIP(hipEventRecord(c->ev0, c->stream));
    if (grid > 0) {
        propagate_kernel<<<grid, kBlock, lds, c->stream>>>(P);
        PP_HIP(hipGetLastError());
    }
    PP_HIP(hipEventRecord(c->ev1, c->stream));
    c->timed = true;
    reduce_kernel<<<n_req, 64, 0, c->stream>>>((const TilePartial*)c->partials.p,
                                               (const uint64_t*)c->tile_start.p,
                                               (const DevReq*)c->reqs.p, n_req,
                                               (pp_summary*)c->sums.p);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// Group staged hits by request into d_out with offsets; fills n_stored/status.
int enqueue_compaction(pp_ctx* c, uint32_t n_req, const pp_hit* d_staging, uint64_t hits_cap,
                       pp_hit* d_out, uint64_t* d_offsets) {
    int rc;
    if ((rc = ensure(c->cnt, n_req * sizeof(uint32_t)))) return rc;
    if ((rc = ensure(c->fill, n_req * sizeof(uint32_t)))) return rc;
    const int zb = (int)std::min<uint64_t>(1024, (n_req + 255) / 256);
    fill_u32_kernel<<<zb, 256, 0, c->stream>>>((uint32_t*)c->cnt.p, n_req, 0u);
    fill_u32_kernel<<<zb, 256, 0, c->stream>>>((uint32_t*)c->fill.p, n_req, 0u);
    const unsigned long long* cursor = (const unsigned long long*)((uint32_t*)c->counters.p + 2);
    const int gb = 2048;
    count_kernel<<<gb, 256, 0, c->stream>>>(d_staging, cursor, hits_cap, (uint32_t*)c->cnt.p);
    scan_kernel<<<1, 1024, 0, c->stream>>>((const uint32_t*)c->cnt.p, n_req, d_offsets,
                                           (pp_summary*)c->sums.p);
    scatter_kernel<<<gb, 256, 0, c->stream>>>(d_staging, cursor, hits_cap, d_offsets,
                                              (uint32_t*)c->fill.p, d_out);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

int validate_ctx(pp_ctx* c) {
    if (!c) return fail(PP_ERR_INVALID_ARG, "context is NULL");
    PP_HIP(hipSetDevice(c->device));
    return PP_OK;
}

int run_batch(pp_ctx* c, const pp_request* reqs, uint32_t n_req, const pp_spectrum* spec,
              pp_hit* d_out, uint64_t hits_cap, uint64_t* d_offsets) {
    int rc;
    if ((rc = ensure(c->staging, std::max<uint64_t>(hits_cap, 1) * sizeof(pp_hit)))) return rc;
    if ((rc = enqueue_propagation(c, reqs, n_req, spec, (pp_hit*)c->staging.p, hits_cap))) return rc;
    return enqueue_compaction(c, n_req, (const pp_hit*)c->staging.p, hits_cap, d_out, d_offsets);
}

}  // namespace

extern "C" {

const char* pp_last_error(void) { return g_err.c_str(); }

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
    if ((rc = upload_medium(c, medium))) {
        std::string msg = g_err;
        pp_destroy(c);
        g_err = msg;
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
                      &c->offsets, &c->out_hits, &c->scratch};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
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

int pp_propagate(pp_ctx* c, const pp_emitter* emitter, const pp_spectrum* spectrum,
                 const pp_sphere* target, uint64_t seed, uint32_t stream, uint32_t max_steps,
                 pp_hit* hits_out, uint64_t hits_cap, pp_summary* summary_out) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (!emitter || !target) return fail(PP_ERR_INVALID_ARG, "emitter/target is NULL");
    if (hits_cap && !hits_out) return fail(PP_ERR_INVALID_ARG, "hits_out is NULL with hits_cap > 0");
    pp_request q;
    memset(&q, 0, sizeof q);
    q.emitter = *emitter;
    q.target = *target;
    q.seed = seed;
    q.stream = stream;
    q.max_steps = max_steps;
    if ((rc = ensure(c->staging, std::max<uint64_t>(hits_cap, 1) * sizeof(pp_hit)))) return rc;
    if ((rc = enqueue_propagation(c, &q, 1, spectrum, (pp_hit*)c->staging.p, hits_cap))) return rc;
    pp_summary s;
    unsigned long long cursor = 0;
    PP_HIP(hipMemcpyAsync(&s, c->sums.p, sizeof s, hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipMemcpyAsync(&cursor, (uint32_t*)c->counters.p + 2, sizeof cursor,
                          hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    const uint64_t stored = std::min<uint64_t>(cursor, hits_cap);
    if (stored)
        PP_HIP(hipMemcpy(hits_out, c->staging.p, stored * sizeof(pp_hit), hipMemcpyDeviceToHost));
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
                       uint64_t* offsets_out, pp_summary* summaries_out) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if (hits_cap && !hits_out) return fail(PP_ERR_INVALID_ARG, "hits_out is NULL with hits_cap > 0");
    if (!offsets_out || !summaries_out)
        return fail(PP_ERR_INVALID_ARG, "offsets_out/summaries_out is NULL");
    if ((rc = ensure(c->out_hits, std::max<uint64_t>(hits_cap, 1) * sizeof(pp_hit)))) return rc;
    if ((rc = ensure(c->offsets, ((uint64_t)n_req + 1) * sizeof(uint64_t)))) return rc;
    if ((rc = run_batch(c, requests, n_req, spectrum, (pp_hit*)c->out_hits.p, hits_cap,
                        (uint64_t*)c->offsets.p)))
        return rc;
    PP_HIP(hipMemcpyAsync(offsets_out, c->offsets.p, ((uint64_t)n_req + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipMemcpyAsync(summaries_out, c->sums.p, (uint64_t)n_req * sizeof(pp_summary),
                          hipMemcpyDeviceToHost, c->stream));
    PP_HIP(hipStreamSynchronize(c->stream));
    const uint64_t stored = offsets_out[n_req];
    if (stored)
        PP_HIP(hipMemcpy(hits_out, c->out_hits.p, stored * sizeof(pp_hit), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n_req; ++i)
        if (summaries_out[i].status == PP_ERR_OVERFLOW)
            return fail(PP_ERR_OVERFLOW, "hit buffer overflow (capacity %llu)",
                        (unsigned long long)hits_cap);
    return PP_OK;
}

int pp_propagate_batch_device(pp_ctx* c, const pp_request* requests, uint32_t n_req,
                              const pp_spectrum* spectrum, pp_hit* d_hits, uint64_t hits_cap,
                              uint64_t* d_offsets, pp_summary* d_summaries) {
    int rc = validate_ctx(c);
    if (rc) return rc;
    if ((hits_cap && !d_hits) || !d_offsets || !d_summaries)
        return fail(PP_ERR_INVALID_ARG, "device output pointer is NULL");
    if ((rc = run_batch(c, requests, n_req, spectrum, d_hits, hits_cap, d_offsets))) return rc;
    PP_HIP(hipMemcpyAsync(d_summaries, c->sums.p, (uint64_t)n_req * sizeof(pp_summary),
                          hipMemcpyDeviceToDevice, c->stream));
    return PP_OK;
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

static ppd::HG make_hg(float g) {
    ppd::HG h;
    h.g = g;
    h.iso = g == 0.0f;
    h.num = std::fma(-g, g, 1.0f);
    h.opg2 = std::fma(g, g, 1.0f);
    h.inv2g = h.iso ? 0.0f : 1.0f / (2.0f * g);
    return h;
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
