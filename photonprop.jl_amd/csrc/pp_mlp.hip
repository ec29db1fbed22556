// pp_mlp.hip — MLP surrogate of the propagation (training and inference) with
// hand-written BF16 MFMA kernels for gfx950.
//
// Role: PhotonProp.jl's Modelling module fits per-(distance, angle) arrival-time
// distributions and trains a Flux MLP on them (src/modelling.jl:52-111,
// :202-267).  Here an MLP maps request features to the log expected yield and a
// 3-component gamma mixture of the residual arrival time, trained directly on
// Monte-Carlo output of the propagation kernel.
//
// Layout: weights of layer l are Wt_l[n_out][n_in] (row-major, "out x in"),
// kept in FP32 (master, Adam state) and in two BF16 copies, Wt_l and W_l =
// Wt_lᵀ.  Every product is one kernel, gemm_nt: C[i][j] = Σ_k A[i][k]·B[j][k]
// with both operands contiguous along k, which lets each lane fetch its MFMA
// fragment (8 consecutive bf16 = 16 bytes) straight from global memory:
//     forward    Z_l  = A_l · Wt_lᵀ              (A = A_l,   B = Wt_l)
//     backward   dA_l = dZ_l · Wt_l              (A = dZ_l,  B = W_l)
//                dWt_l = dZ_lᵀ · A_l             (A = dZ_lᵀ, B = A_lᵀ, split over the batch)
// Inner dimensions are padded to multiples of 16 with zeros.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "pp_host.hpp"

using namespace pph;

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMaxLayers = 16;
constexpr uint32_t kMaxSplits = 64;  // split-K partials of a weight gradient at most
constexpr float kLogMuCap = 30.0f;   // Poisson log mean beyond which the loss grows quadratically
constexpr uint32_t kTimeScale = 10;  // residual times enter the mixture in units of 10 ns

inline uint32_t pad16(uint32_t x) { return (x + 15u) & ~15u; }

enum Epilogue : int {
    EPI_STORE = 0,      // C = acc (fp32)
    EPI_PARTIAL = 1,    // C[split] = acc (fp32; split-K partials, summed in order afterwards)
    EPI_BIAS_ACT = 2,   // C = z = acc + bias (fp32), Y = act(z) (bf16)
    EPI_ACT_ONLY = 3,   // Y = act(acc + bias) (bf16), C untouched
    EPI_BIAS = 4,       // C = acc + bias (fp32)
};

struct GemmArgs {
    const bf16* A;
    const bf16* B;
    float* C;
    const float* bias;
    bf16* Y;
    uint32_t lda, ldb, ldc, ldy;
    uint32_t M, N, K, k_split;
    uint64_t split_stride;  // EPI_PARTIAL: floats between the partials of consecutive splits
    int epi;
    uint32_t act;  // PP_MLP_SILU or PP_MLP_RELU (EPI_BIAS_ACT, EPI_ACT_ONLY)
};

// z·sigmoid(z); the 1-ulp hardware reciprocal instead of an IEEE division (the
// result is rounded to bf16 or feeds a bf16 product).
__device__ __forceinline__ float silu(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

__device__ __forceinline__ float hidden_act(float z, uint32_t act) {
    return act == PP_MLP_RELU ? fmaxf(z, 0.0f) : silu(z);
}

// 64x64 block tile, 4 waves of one 32x32 MFMA tile each.  The products here
// are small (a 4096-row minibatch against 256-wide layers: 256 blocks at this
// tile), so the kernel is bound by memory latency rather than MFMA rate: each
// wave keeps kGemmAhead k-steps of its A and B fragments in flight (a rolling
// window, refilled as each step's MFMA is issued).  The 128x128 tile it
// replaces loaded each k-step's fragments right before their MFMAs, on 64
// blocks, and took 16 µs per product.
constexpr int kGemmAhead = 8;

__global__ void __launch_bounds__(256) gemm_nt_kernel(GemmArgs g) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t i0 = blockIdx.y * 64u + (wave >> 1) * 32u;
    const uint32_t j0 = blockIdx.x * 64u + (wave & 1u) * 32u;
    const uint32_t kb = blockIdx.z * g.k_split;
    const uint32_t ke = min(g.K, kb + g.k_split);
    const uint32_t nst = ke > kb ? (ke - kb) / 16u : 0u;  // K is a multiple of 16
    const uint32_t r = lane & 31u, kh = (lane >> 5) * 8u;
    const bool rok = i0 + r < g.M, cok = j0 + r < g.N;
    const bf16* pa = g.A + (size_t)(rok ? i0 + r : 0u) * g.lda + kb + kh;
    const bf16* pb = g.B + (size_t)(cok ? j0 + r : 0u) * g.ldb + kb + kh;
    const bf16x8 zero = {};
    f32x16 acc;
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0f;
    bf16x8 fa[kGemmAhead], fb[kGemmAhead];
#pragma unroll
    for (int u = 0; u < kGemmAhead; ++u) {
        const bool in = (uint32_t)u < nst;
        fa[u] = in && rok ? *reinterpret_cast<const bf16x8*>(pa + u * 16) : zero;
        fb[u] = in && cok ? *reinterpret_cast<const bf16x8*>(pb + u * 16) : zero;
    }
    for (uint32_t s0 = 0; s0 < nst; s0 += kGemmAhead) {
#pragma unroll
        for (int u = 0; u < kGemmAhead; ++u) {
            const uint32_t st = s0 + u;
            if (st < nst) {  // uniform
                const bf16x8 a = fa[u], b = fb[u];
                const uint32_t sn = st + kGemmAhead;
                const bool in = sn < nst;
                fa[u] = in && rok ? *reinterpret_cast<const bf16x8*>(pa + sn * 16u) : zero;
                fb[u] = in && cok ? *reinterpret_cast<const bf16x8*>(pb + sn * 16u) : zero;
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
        }
    }
    // D layout of v_mfma_f32_32x32x16: lane holds column lane%32, rows
    // 8*(q/4) + 4*(lane/32) + q%4 for q = 0..15.
    const uint32_t j = j0 + r;
    if (j >= g.N) return;
    const float bj = (g.epi >= EPI_BIAS_ACT) ? g.bias[j] : 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t i = i0 + 8u * (q >> 2) + 4u * (lane >> 5) + (q & 3);
        if (i >= g.M) continue;
        const float v = acc[q];
        switch (g.epi) {
        case EPI_STORE: g.C[(size_t)i * g.ldc + j] = v; break;
        case EPI_PARTIAL: g.C[(size_t)blockIdx.z * g.split_stride + (size_t)i * g.ldc + j] = v; break;
        case EPI_BIAS_ACT: {
            const float z = v + bj;
            g.C[(size_t)i * g.ldc + j] = z;
            g.Y[(size_t)i * g.ldy + j] = (bf16)hidden_act(z, g.act);
            break;
        }
        case EPI_ACT_ONLY: g.Y[(size_t)i * g.ldy + j] = (bf16)hidden_act(v + bj, g.act); break;
        default: g.C[(size_t)i * g.ldc + j] = v + bj; break;
        }
    }
}

// Fused inference: every layer of the MLP for a tile of kFusedRows rows in
// one block, the activations resident in LDS as bf16 rows (input features,
// then each hidden layer's output written over its input).  Each layer is
// D = Wt_l · Aᵀ on v_mfma_f32_32x32x16_bf16: the weights are the A operand,
// read from global memory in fragment order (L2-resident: every block reads
// the same weights) kFusedAhead k-steps ahead of use, the activations the B
// operand, read from LDS one k-step ahead.  Wave w owns output features
// [32w, 32w + 32) for all 128 rows (four 32x32 accumulator tiles).  In D's
// layout a lane holds one row of the tile (lane % 32) and four consecutive
// output features per group of four accumulator registers, so the epilogue
// (bias, activation, bf16) writes 8 bytes per group back to LDS.  The output
// layer (<= 16 features) runs on 16x16x32 MFMAs, one 16-row tile per wave.
// Only the fp32 inputs and outputs touch HBM.  Limits: every padded width <=
// kFusedMaxWidth.  Two blocks share a CU (LDS 2 x 66 KiB, 4 waves per SIMD).
// Measured and dropped: 64x64 tiles per wave (each LDS fragment feeding two
// MFMAs, each weight fragment read by two waves) ran at half the speed, as
// the L2 requests doubled; the L1 does not merge the second wave's reads.
// 128x64 tiles per wave (4 waves per block, each weight fragment still read
// once per block, half the LDS reads) ran 6-11 % slower: 255 VGPRs leave 2
// waves per SIMD, too few to hide the weights' L2 latency.
constexpr uint32_t kFusedRows = 128;      // 4 sample tiles of 32
constexpr uint32_t kFusedMaxWidth = 256;  // 8 feature tiles: one per wave
constexpr uint32_t kFusedLd = kFusedMaxWidth + 8;  // LDS row stride (bf16): rows 16 B apart
constexpr uint32_t kFusedThreads = 512;
constexpr int kFusedAhead = 2;  // weight fragments in flight per wave (3 or 4 spill VGPRs)

struct FusedArgs {
    const float* x;
    float* out;
    uint32_t M, n_in, ld_x, n_out, ld_out, n_layers;
    uint32_t dims[kMaxLayers + 1];
    uint32_t pdim[kMaxLayers + 1];
    const bf16x8* wfrag[kMaxLayers];  // weights in fragment order (frag_geom)
    const float* bias[kMaxLayers];
};

// The fused kernel reads each layer's weights in MFMA fragment order, so that
// one wave's load of a fragment is 1 KiB of contiguous memory (8 full cache
// lines): wfrag[(tile·steps + s)·64 + lane] holds the 8 bf16
// Wt[tile·rows + lane % rows][s·kstep + 8·(lane / rows) + 0..7], zero past the
// padded matrix.  Hidden layers feed v_mfma_f32_32x32x16 (tiles of 32 output
// features, k-steps of 16), the output layer v_mfma_f32_16x16x32 (16 features,
// k-steps of 32).  Read straight from the row-major Wt, a fragment is 32 B of
// 32 different rows: 32 cache lines for every 1 KiB, which the L1 could not
// keep for the 16 waves of a CU, so the L2 delivered every line four times.
struct FragGeom {
    uint32_t rows, kstep, tiles, steps;
};
__host__ __device__ inline FragGeom frag_geom(uint32_t no_pad, uint32_t k_pad, bool output_layer) {
    FragGeom g;
    g.rows = output_layer ? 16u : 32u;
    g.kstep = output_layer ? 32u : 16u;
    g.tiles = (no_pad + g.rows - 1) / g.rows;
    g.steps = (k_pad + g.kstep - 1) / g.kstep;
    return g;
}

// One layer's k-loop for a width known at compile time (kSteps k-steps of 16),
// fully unrolled: the weight fragments are prefetched kFusedAhead steps ahead,
// and the next step's four activation fragments are read from LDS before this
// step's four MFMAs are issued, so the LDS latency runs under the MFMAs (left
// to itself the compiler read each fragment right before its MFMA and waited
// for it there).  The loads are unconditional (a lane past the layer's
// features reads row 0: its accumulator rows are output features the epilogue
// does not keep), so the code is straight-line.
// li·kFusedLd + col as an opaque value.  Added to it, the constant offsets of
// the other row tiles (t·32·kFusedLd) become immediate offsets of the LDS
// instructions; left visible, the compiler folds li·kFusedLd + t·32·kFusedLd
// back into (li | 32t)·kFusedLd, one product per row tile, and those stayed
// live across the layer loop and spilled to scratch.
__device__ __forceinline__ uint32_t lane_row_offset(uint32_t li, uint32_t col) {
    uint32_t o = li * kFusedLd + col;
    asm volatile("" : "+v"(o));
    return o;
}

template <int kSteps>
__device__ __forceinline__ void fused_mfma_steps(const bf16x8* wfr, const bf16* act, uint32_t li,
                                                 uint32_t kh, f32x16 (&acc)[4]) {
    bf16x8 wf[kFusedAhead];
#pragma unroll
    for (int u = 0; u < kFusedAhead && u < kSteps; ++u) wf[u] = wfr[u * 64];
    const bf16* arow = act + lane_row_offset(li, kh);
    bf16x8 xb[2][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
        xb[0][t] = *reinterpret_cast<const bf16x8*>(arow + t * 32u * kFusedLd);
#pragma unroll
    for (int st = 0; st < kSteps; ++st) {
        const bf16x8 w = wf[st % kFusedAhead];
        if (st + kFusedAhead < kSteps) wf[st % kFusedAhead] = wfr[(st + kFusedAhead) * 64];
        if (st + 1 < kSteps) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                xb[(st + 1) & 1][t] = *reinterpret_cast<const bf16x8*>(
                    arow + t * 32u * kFusedLd + (st + 1) * 16u);
        }
        // keep both prefetches ahead of this step's MFMAs
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, xb[st & 1][t], acc[t], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
static_assert(kFusedRows == 16 * (kFusedThreads / 64), "one 16-row output tile per wave");

// The output layer (at most 16 features, pdim 16) on v_mfma_f32_16x16x32_bf16,
// one 16-row tile per wave, so that all eight waves share it (as one 32-wide
// feature tile it ran on a single wave, as long as a whole hidden layer).
// D[feature][row] = Wt[feature][k] · act[row][k]; lane l holds A row l%16 and
// B column l%16 with k = 8·(l/16) + 0..7 of each 32-wide k-step, and D rows
// 4·(l/16) + 0..3 of column l%16.  K is a multiple of 16; a k-group past it
// is zero.
__device__ __forceinline__ void fused_output_layer(const FusedArgs& f, uint32_t l, bf16* act,
                                                   uint32_t wave, uint32_t lane, uint32_t row0) {
    const uint32_t K = f.pdim[l];
    const uint32_t c = lane & 15u, kg = (lane >> 4) * 8u;
    const uint32_t rtile = wave * 16u;
    const bf16x8* wfr = f.wfrag[l] + lane;  // one 16-feature tile
    const bf16* arow = act + (rtile + c) * kFusedLd + kg;
    const bf16x8 zero = {};
    f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0, s = 0; k < K; k += 32u, ++s) {
        // the weights are zero past K; the LDS row there is not
        const bf16x8 b = k + kg < K ? *reinterpret_cast<const bf16x8*>(arow + k) : zero;
        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfr[s * 64u], b, d, 0, 0, 0);
    }
    // The tile goes out through this wave's own LDS rows (only this wave read
    // them, and its reads have completed): D is row-scattered across lanes, so
    // stored directly each 4-byte store would write a few bytes of many lines.
    float* stage = reinterpret_cast<float*>(act + rtile * kFusedLd);  // [16 rows][16]
    *reinterpret_cast<f32x4*>(stage + c * 16u + (lane >> 4) * 4u) = d;
    // other lanes of this wave read what a lane wrote
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const uint32_t n_out = f.n_out, n_rows = min(16u, f.M - min(f.M, row0 + rtile));
    for (uint32_t e = lane; e < n_rows * n_out; e += 64u) {
        const uint32_t r = e / n_out, j = e - r * n_out;
        f.out[(size_t)(row0 + rtile + r) * f.ld_out + j] = stage[r * 16u + j] + f.bias[l][j];
    }
}

// One hidden layer of the fused kernel: wave w owns feature tile w for all
// 128 rows (acc[t] = row tile t); products, then (after every wave has read
// the layer's input) bias, activation and bf16 written over the input.
template <bool kRelu>
__device__ __forceinline__ void fused_hidden_layer(const FusedArgs& f, uint32_t l, bf16* act,
                                                   uint32_t wave, uint32_t lane) {
    const uint32_t li = lane & 31u, kh = (lane >> 5) * 8u;
    const uint32_t K = f.pdim[l], NO = f.pdim[l + 1], n_true = f.dims[l + 1];
    const bool mine = wave * 32u < NO;  // wave-uniform: this wave has a feature tile
    // this wave's feature tile; rows past NO are zero
    const bf16x8* wfr = f.wfrag[l] + (mine ? (size_t)wave * (K / 16u) * 64u : 0u) + lane;
    f32x16 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[t][q] = 0.0f;
    if (mine && K == kFusedMaxWidth) {
        fused_mfma_steps<kFusedMaxWidth / 16>(wfr, act, li, kh, acc);
    } else if (mine) {
        const bf16x8 zero = {};
        bf16x8 wf[kFusedAhead];
#pragma unroll
        for (int u = 0; u < kFusedAhead; ++u) wf[u] = (uint32_t)u * 16u < K ? wfr[u * 64] : zero;
        for (uint32_t k = 0; k < K; k += 16u * kFusedAhead) {
#pragma unroll
            for (int u = 0; u < kFusedAhead; ++u) {
                const uint32_t kk = k + 16u * u;
                if (kk >= K) break;
                const bf16x8 w = wf[u];
                const uint32_t kn = kk + 16u * kFusedAhead;
                if (kn < K) wf[u] = wfr[(kn / 16u) * 64u];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bf16x8 xb = *reinterpret_cast<const bf16x8*>(
                        act + lane_row_offset(li, kh) + t * 32u * kFusedLd + kk);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, xb, acc[t], 0, 0, 0);
                }
            }
        }
    }
    __syncthreads();  // every wave has read this layer's input
    if (mine) {
        const float* bias = f.bias[l];
        // this lane's 16 biases (D rows 8g + 4*(lane/32) + r), loaded once per
        // layer rather than once per row tile, and the activation computed for
        // every lane and then selected: the guarded per-element loads and
        // branches this replaces took a fifth of the kernel's time
        float bj[4][4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint32_t jg = wave * 32u + 8u * g + 4u * (lane >> 5);
#pragma unroll
            for (int r = 0; r < 4; ++r) bj[g][r] = jg + r < n_true ? bias[jg + r] : 0.0f;
        }
        // row li, first feature of this lane's D rows; row tiles and feature
        // groups are constant offsets from it
        bf16* const erow = act + lane_row_offset(li, wave * 32u + 4u * (lane >> 5));
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                // D rows 8g + 4*(lane/32) + 0..3 of this wave's feature tile
                const uint32_t jg = wave * 32u + 8u * g + 4u * (lane >> 5);
                typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                bf16x4 y;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float v = acc[t][4 * g + r] + bj[g][r];
                    // computed for every lane, then selected: no branch
                    const float sv = kRelu ? fmaxf(v, 0.0f) : silu(v);
                    y[r] = (bf16)(jg + r < n_true ? sv : 0.0f);
                }
                *reinterpret_cast<bf16x4*>(erow + t * 32u * kFusedLd + 8u * g) = y;
            }
        }
    }
    __syncthreads();  // this layer's output is the next layer's input
}

template <bool kRelu>
__global__ void __launch_bounds__(kFusedThreads) __attribute__((amdgpu_waves_per_eu(4)))
mlp_forward_fused_kernel(FusedArgs f) {
    extern __shared__ bf16 act[];  // [kFusedRows][kFusedLd]
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    const uint32_t row0 = blockIdx.x * kFusedRows;
    // input tile: fp32 -> bf16, zero beyond n_in and beyond M
    const uint32_t K0 = f.pdim[0];
    for (uint32_t e = tid; e < kFusedRows * K0; e += kFusedThreads) {
        const uint32_t i = e / K0, k = e - i * K0;
        float v = 0.0f;
        if (row0 + i < f.M && k < f.n_in) v = f.x[(size_t)(row0 + i) * f.ld_x + k];
        act[i * kFusedLd + k] = (bf16)v;
    }
    __syncthreads();
    for (uint32_t l = 0; l + 1 < f.n_layers; ++l) fused_hidden_layer<kRelu>(f, l, act, wave, lane);
    fused_output_layer(f, f.n_layers - 1, act, wave, lane, row0);
}

// out[c][r] = in[r][c] for an R x C bf16 matrix (ld_in >= C, ld_out >= R).
__global__ void __launch_bounds__(256) transpose_kernel(const bf16* in, uint32_t R, uint32_t C,
                                                        uint32_t ld_in, bf16* out,
                                                        uint32_t ld_out) {
    __shared__ bf16 tile[32][33];
    const uint32_t c0 = blockIdx.x * 32u, r0 = blockIdx.y * 32u;
    const uint32_t tx = threadIdx.x & 31u, ty = threadIdx.x >> 5;  // 32 x 8
    for (uint32_t y = ty; y < 32; y += 8) {
        const uint32_t rr = r0 + y, cc = c0 + tx;
        tile[y][tx] = (rr < R && cc < C) ? in[(size_t)rr * ld_in + cc] : (bf16)0.0f;
    }
    __syncthreads();
    for (uint32_t y = ty; y < 32; y += 8) {
        const uint32_t cc = c0 + y, rr = r0 + tx;
        if (cc < C && rr < R) out[(size_t)cc * ld_out + rr] = tile[tx][y];
    }
}

// fp32 [M][n_in] -> bf16 [M][ld] with zero padding (rows beyond M zeroed up to M_pad).
__global__ void pack_input_kernel(const float* x, uint32_t M, uint32_t M_pad, uint32_t n_in,
                                  bf16* out, uint32_t ld) {
    const uint64_t n = (uint64_t)M_pad * ld;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(t / ld), c = (uint32_t)(t % ld);
        out[t] = (i < M && c < n_in) ? (bf16)x[(size_t)i * n_in + c] : (bf16)0.0f;
    }
}

// dZ = dA ⊙ act'(z), bf16 out (rows >= M zeroed up to M_pad).  ReLU's
// derivative at z = 0 is taken as 0, as in Flux/Zygote.
__global__ void act_backward_kernel(const float* dA, const float* Z, uint32_t M, uint32_t M_pad,
                                    uint32_t N, uint32_t ld, uint32_t act, bf16* dZ) {
    const uint64_t n = (uint64_t)M_pad * N;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(t / N), j = (uint32_t)(t % N);
        float d = 0.0f;
        if (i < M) {
            const float z = Z[(size_t)i * ld + j];
            if (act == PP_MLP_RELU) {
                d = z > 0.0f ? dA[(size_t)i * ld + j] : 0.0f;
            } else {
                const float s = 1.0f / (1.0f + __expf(-z));
                d = dA[(size_t)i * ld + j] * s * (1.0f + z * (1.0f - s));
            }
        }
        dZ[(size_t)i * ld + j] = (bf16)d;
    }
}

// db[j] = Σ_i dZ[i][j] in two passes with a fixed summation order (the same
// gradient on every run).  Pass 1: block (x, y) sums rows [y·chunk, (y+1)·chunk)
// of columns [64x, 64x + 64) into part[y][j]; pass 2 adds the partials in y
// order.  (One block per 64 columns walking every row took 250 µs of a 4096-row
// step: four blocks on the whole device.)
__global__ void __launch_bounds__(256) bias_grad_partial_kernel(const bf16* dZ, uint32_t M,
                                                                uint32_t N, uint32_t ld,
                                                                uint32_t chunk, float* part) {
    __shared__ float red[4][64];
    const uint32_t c = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t j = blockIdx.x * 64u + c;
    const uint32_t i0 = blockIdx.y * chunk, i1 = min(M, i0 + chunk);
    float s = 0.0f;
    if (j < N)
        for (uint32_t i = i0 + w; i < i1; i += 4) s += (float)dZ[(size_t)i * ld + j];
    red[w][c] = s;
    __syncthreads();
    if (w == 0 && j < N)
        part[(size_t)blockIdx.y * N + j] = red[0][c] + red[1][c] + red[2][c] + red[3][c];
}

// One wave per column: lane y holds partial y (parts <= 64), then a butterfly
// of fixed shape.
static_assert(kMaxSplits <= 64, "one partial per lane");
__global__ void __launch_bounds__(256) bias_grad_sum_kernel(const float* part, uint32_t parts,
                                                            uint32_t N, float* db) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t j = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (j >= N) return;  // wave-uniform
    float s = lane < parts ? part[(size_t)lane * N + j] : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) db[j] += s;
}

__device__ float digamma(float x) {
    // Recurrence to x >= 6, then the asymptotic series of ψ.
    float r = 0.0f;
    while (x < 6.0f) {
        r -= 1.0f / x;
        x += 1.0f;
    }
    const float f = 1.0f / (x * x);
    return r + __logf(x) - 0.5f / x -
           f * (1.0f / 12 - f * (1.0f / 120 - f * (1.0f / 252 - f * (1.0f / 240 - f / 132))));
}

// Loss per row: Poisson NLL μ - k·log μ with μ = N·exp(y0), plus time_weight ×
// the mean negative log-likelihood of the row's residual times under the
// 3-component gamma mixture y[1..3] (logits), y[4..6] (log shape),
// y[7..9] (log rate).  Gradients (divided by M) go to dY as bf16; loss sums
// to loss[0] (Poisson) and loss[1] (time) by atomics.
__global__ void loss_kernel(const float* Y, uint32_t ldy, const float* logn, const float* kobs,
                            const float* times, uint32_t n_times, uint32_t M, uint32_t M_pad,
                            float time_weight, bf16* dY, uint32_t ld_dy, double* loss) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M_pad) return;
    float g[16];
    for (int c = 0; c < 16; ++c) g[c] = 0.0f;
    float lp = 0.0f, lt = 0.0f;
    if (i < M) {
        const float* y = Y + (size_t)i * ldy;
        const float inv_m = 1.0f / (float)M;
        // Poisson NLL mu - k·l of the log mean l = y0 + log N.  Above
        // l = kLogMuCap (mu = 1.1e13, far above any row's count) mu's growth
        // continues linearly, e^c·(1 + (l - c)), so a step that overshoots
        // keeps a bounded gradient (e^l would overflow Adam's second moment and
        // freeze the network there) and the loss stays convex.
        const float l = y[0] + logn[i];
        const float ex = fmaxf(l - kLogMuCap, 0.0f);
        const float ec = __expf(fminf(l, kLogMuCap));
        const float mu = ec * (1.0f + ex);  // the derivative of the mean term
        lp = ec * (1.0f + ex + 0.5f * ex * ex) - kobs[i] * l;
        g[0] = (mu - kobs[i]) * inv_m;
        // gamma mixture
        float logit[3], al[3], be[3], lal[3], lbe[3];
        float mx = fmaxf(y[1], fmaxf(y[2], y[3]));
        float zs = 0.0f;
        for (int c = 0; c < 3; ++c) {
            logit[c] = __expf(y[1 + c] - mx);
            zs += logit[c];
            lal[c] = fminf(fmaxf(y[4 + c], -4.0f), 6.0f);
            lbe[c] = fminf(fmaxf(y[7 + c], -8.0f), 6.0f);
            al[c] = __expf(lal[c]);
            be[c] = __expf(lbe[c]);
        }
        float pi[3], lg[3], dig[3];
        for (int c = 0; c < 3; ++c) {
            pi[c] = logit[c] / zs;
            lg[c] = lgammaf(al[c]);
            dig[c] = digamma(al[c]);
        }
        uint32_t nt = 0;
        float gl[3] = {0, 0, 0}, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
        for (uint32_t s = 0; s < n_times; ++s) {
            const float t = times[(size_t)i * n_times + s];
            if (!(t >= 0.0f)) continue;
            const float ts = t / (float)kTimeScale + 1e-3f;
            const float lt_ = __logf(ts);
            float lpc[3], m2 = -INFINITY;
            for (int c = 0; c < 3; ++c) {
                lpc[c] = __logf(pi[c]) + al[c] * lbe[c] + (al[c] - 1.0f) * lt_ - be[c] * ts - lg[c];
                m2 = fmaxf(m2, lpc[c]);
            }
            float den = 0.0f;
            for (int c = 0; c < 3; ++c) den += __expf(lpc[c] - m2);
            lt -= m2 + __logf(den);
            for (int c = 0; c < 3; ++c) {
                const float resp = __expf(lpc[c] - m2) / den;
                gl[c] -= resp - pi[c];
                ga[c] -= resp * al[c] * (lbe[c] + lt_ - dig[c]);
                gb[c] -= resp * (al[c] - be[c] * ts);
            }
            ++nt;
        }
        if (nt) {
            const float sc = time_weight * inv_m / (float)nt;
            lt /= (float)nt;
            for (int c = 0; c < 3; ++c) {
                g[1 + c] = gl[c] * sc;
                g[4 + c] = (y[4 + c] > -4.0f && y[4 + c] < 6.0f) ? ga[c] * sc : 0.0f;
                g[7 + c] = (y[7 + c] > -8.0f && y[7 + c] < 6.0f) ? gb[c] * sc : 0.0f;
            }
        }
    }
    for (uint32_t c = 0; c < ld_dy; ++c) dY[(size_t)i * ld_dy + c] = (bf16)(c < 16 ? g[c] : 0.0f);
    if (i < M) {
        atomicAdd(&loss[0], (double)lp);
        atomicAdd(&loss[1], (double)lt);
    }
}

// Adam (FP32 master weights).
__global__ void adam_kernel(float* w, const float* g, float* m, float* v, uint64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const float gi = g[t];
        const float mi = b1 * m[t] + (1.0f - b1) * gi;
        const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
        m[t] = mi;
        v[t] = vi;
        w[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// BF16 copies of one layer: Wt [n_out][ldk] and W = Wtᵀ [n_in_pad][ldn] (zero padded).
__global__ void sync_weights_kernel(const float* wt, uint32_t n_out, uint32_t n_in, bf16* wt16,
                                    uint32_t ldk, bf16* w16, uint32_t ldn, uint32_t n_in_pad,
                                    uint32_t n_out_pad) {
    const uint64_t n = (uint64_t)n_out_pad * ldk;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = (uint32_t)(t / ldk), k = (uint32_t)(t % ldk);
        const bf16 v = (o < n_out && k < n_in) ? (bf16)wt[(size_t)o * n_in + k] : (bf16)0.0f;
        if (o < n_out) wt16[(size_t)o * ldk + k] = v;
        if (k < n_in_pad) w16[(size_t)k * ldn + o] = v;
    }
}

// Layer weights in the fused kernel's fragment order (see frag_geom) from the
// padded row-major bf16 copy wt16 [no_pad][k_pad].
__global__ void pack_fragments_kernel(const bf16* wt16, uint32_t no_pad, uint32_t k_pad,
                                      FragGeom g, bf16x8* out) {
    const uint64_t n = (uint64_t)g.tiles * g.steps * 64u;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lane = (uint32_t)(t & 63u);
        const uint32_t s = (uint32_t)((t >> 6) % g.steps), tile = (uint32_t)((t >> 6) / g.steps);
        const uint32_t r = tile * g.rows + lane % g.rows;
        const uint32_t k0 = s * g.kstep + 8u * (lane / g.rows);
        bf16x8 v = {};
        if (r < no_pad) {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (k0 + i < k_pad) v[i] = wt16[(size_t)r * k_pad + k0 + i];
        }
        out[t] = v;
    }
}

// Gradient of layer weights in [n_out][ldk] padded form -> packed [n_out][n_in].
// The weight gradient from its split-K partials, added in split order (the
// same sum on every run, unlike atomics), unpadded into the packed vector.
__global__ void unpack_grad_splits_kernel(const float* gp, uint32_t splits, uint64_t stride,
                                          uint32_t n_out, uint32_t n_in, uint32_t ldk, float* g) {
    const uint64_t n = (uint64_t)n_out * n_in;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = (uint32_t)(t / n_in), k = (uint32_t)(t % n_in);
        const float* p = gp + (size_t)o * ldk + k;
        float acc = p[0];
        for (uint32_t sp = 1; sp < splits; ++sp) acc += p[sp * stride];
        g[t] = acc;
    }
}

__global__ void zero_f32_kernel(float* p, uint64_t n) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x)
        p[t] = 0.0f;
}

__global__ void add_f64_kernel(const double* src, uint32_t n, double* acc) {
    if (threadIdx.x < n) acc[threadIdx.x] += src[threadIdx.x];
}

__global__ void zero_f64_kernel(double* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0.0;
}

unsigned grid1d(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

}  // namespace

struct pp_mlp {
    pp_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t n_layers = 0;
    uint32_t dims[kMaxLayers + 1];      // true widths: dims[0] = n_in, dims[L] = n_out
    uint32_t pdim[kMaxLayers + 1];      // padded to 16
    uint64_t off_w[kMaxLayers], off_b[kMaxLayers], n_params = 0;
    DevBuf w, grad, adam_m, adam_v;     // fp32 master / grads / Adam state (packed)
    DevBuf gpad;                        // fp32 padded weight-gradient scratch
    DevBuf bpart;                       // fp32 bias-gradient partials [kMaxSplits][width]
    DevBuf wt16[kMaxLayers], w16[kMaxLayers];
    DevBuf wfrag[kMaxLayers];           // bf16 weights in fragment order (fused inference)
    // activations (grow with batch)
    uint32_t cap_rows = 0;
    DevBuf x16, act[kMaxLayers + 1], actT[kMaxLayers + 1], z[kMaxLayers + 1], dA, dZ, dZT, dY,
        yout, loss, loss_acc;
    uint64_t step = 0;
    uint32_t activation = PP_MLP_SILU;  // hidden activation
    bool frags_stale = true;            // wfrag lags the weights
};

namespace {

int mlp_reserve(pp_mlp* m, uint32_t rows) {
    if (rows <= m->cap_rows) return PP_OK;
    const uint32_t R = pad16(rows);
    int rc;
    const uint32_t L = m->n_layers;
    uint32_t wmax = 16;
    for (uint32_t l = 0; l <= L; ++l) wmax = std::max(wmax, m->pdim[l]);
    if ((rc = ensure(m->x16, (size_t)R * m->pdim[0] * sizeof(bf16)))) return rc;
    for (uint32_t l = 1; l < L; ++l) {
        if ((rc = ensure(m->act[l], (size_t)R * m->pdim[l] * sizeof(bf16)))) return rc;
        if ((rc = ensure(m->z[l], (size_t)R * m->pdim[l] * sizeof(float)))) return rc;
        // rows beyond a batch's M are read (times zero gradients) by the weight
        // gradient GEMM; start them finite
        PP_HIP(hipMemsetAsync(m->act[l].p, 0, m->act[l].cap, m->stream));
    }
    for (uint32_t l = 0; l < L; ++l)
        if ((rc = ensure(m->actT[l], (size_t)R * m->pdim[l] * sizeof(bf16)))) return rc;
    if ((rc = ensure(m->dA, (size_t)R * wmax * sizeof(float)))) return rc;
    if ((rc = ensure(m->dZ, (size_t)R * wmax * sizeof(bf16)))) return rc;
    if ((rc = ensure(m->dZT, (size_t)R * wmax * sizeof(bf16)))) return rc;
    if ((rc = ensure(m->dY, (size_t)R * m->pdim[L] * sizeof(bf16)))) return rc;
    if ((rc = ensure(m->yout, (size_t)R * m->pdim[L] * sizeof(float)))) return rc;
    if ((rc = ensure(m->loss, 2 * sizeof(double)))) return rc;
    if ((rc = ensure(m->loss_acc, 2 * sizeof(double)))) return rc;
    m->cap_rows = R;
    return PP_OK;
}

int gemm(pp_mlp* m, const bf16* A, uint32_t lda, const bf16* B, uint32_t ldb, uint32_t M,
         uint32_t N, uint32_t K, int epi, float* C, uint32_t ldc, const float* bias, bf16* Y,
         uint32_t ldy, uint32_t splits = 1, uint64_t split_stride = 0,
         uint32_t* splits_used = nullptr) {
    GemmArgs g;
    g.A = A;
    g.B = B;
    g.C = C;
    g.bias = bias;
    g.Y = Y;
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    g.ldy = ldy;
    g.M = M;
    g.N = N;
    g.K = K;
    const uint32_t per = ((K + splits - 1) / splits + 15) & ~15u;
    g.k_split = per;
    g.split_stride = split_stride;
    g.epi = epi;
    g.act = m->activation;
    dim3 grid((N + 63) / 64, (M + 63) / 64, (K + per - 1) / per);
    if (splits_used) *splits_used = grid.z;
    gemm_nt_kernel<<<grid, 256, 0, m->stream>>>(g);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

int sync_weights(pp_mlp* m) {
    for (uint32_t l = 0; l < m->n_layers; ++l) {
        const uint32_t ni = m->dims[l], no = m->dims[l + 1];
        const uint32_t ldk = m->pdim[l], ldn = m->pdim[l + 1];
        sync_weights_kernel<<<grid1d((uint64_t)ldn * ldk), 256, 0, m->stream>>>(
            (const float*)m->w.p + m->off_w[l], no, ni, (bf16*)m->wt16[l].p, ldk,
            (bf16*)m->w16[l].p, ldn, ldk, ldn);
    }
    PP_HIP(hipGetLastError());
    m->frags_stale = true;  // repacked by the next fused forward, not after every training step
    return PP_OK;
}

// The fused kernel's fragment-order weights from wt16, if the weights changed.
int pack_fragments(pp_mlp* m) {
    if (!m->frags_stale) return PP_OK;
    for (uint32_t l = 0; l < m->n_layers; ++l) {
        const uint32_t ldk = m->pdim[l], ldn = m->pdim[l + 1];
        const FragGeom g = frag_geom(ldn, ldk, l + 1 == m->n_layers);
        pack_fragments_kernel<<<grid1d((uint64_t)g.tiles * g.steps * 64u), 256, 0, m->stream>>>(
            (const bf16*)m->wt16[l].p, ldn, ldk, g, (bf16x8*)m->wfrag[l].p);
    }
    PP_HIP(hipGetLastError());
    m->frags_stale = false;
    return PP_OK;
}

// Forward over M rows already packed in x16.  With `train`, keep z and act.
int forward(pp_mlp* m, uint32_t M, bool train, float* out, uint32_t ld_out) {
    const uint32_t L = m->n_layers;
    const bf16* a = (const bf16*)m->x16.p;
    int rc;
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t K = m->pdim[l], N = m->dims[l + 1];
        const float* b = (const float*)m->w.p + m->off_b[l];
        if (l + 1 < L) {
            bf16* y = (bf16*)m->act[l + 1].p;
            if (train)
                rc = gemm(m, a, K, (const bf16*)m->wt16[l].p, K, M, N, K, EPI_BIAS_ACT,
                          (float*)m->z[l + 1].p, m->pdim[l + 1], b, y, m->pdim[l + 1]);
            else
                rc = gemm(m, a, K, (const bf16*)m->wt16[l].p, K, M, N, K, EPI_ACT_ONLY, nullptr, 0,
                          b, y, m->pdim[l + 1]);
            if (rc) return rc;
            a = y;
        } else {
            if ((rc = gemm(m, a, K, (const bf16*)m->wt16[l].p, K, M, N, K, EPI_BIAS, out, ld_out, b,
                           nullptr, 0)))
                return rc;
        }
    }
    return PP_OK;
}

bool fused_fits(const pp_mlp* m) {
    for (uint32_t l = 0; l <= m->n_layers; ++l)
        if (m->pdim[l] > kFusedMaxWidth) return false;
    return true;
}

// All layers in one kernel (mlp_forward_fused_kernel); fp32 x [M][n_in] in,
// fp32 y [M][n_out] out, both device pointers.  Times the kernel alone.
int forward_fused(pp_mlp* m, const float* d_x, uint32_t M, float* d_y) {
    int rc;
    if ((rc = pack_fragments(m))) return rc;
    FusedArgs fa;
    memset(&fa, 0, sizeof fa);
    const uint32_t L = m->n_layers;
    fa.x = d_x;
    fa.out = d_y;
    fa.M = M;
    fa.n_in = m->dims[0];
    fa.ld_x = m->dims[0];
    fa.n_out = m->dims[L];
    fa.ld_out = m->dims[L];
    fa.n_layers = L;
    for (uint32_t l = 0; l <= L; ++l) {
        fa.dims[l] = m->dims[l];
        fa.pdim[l] = m->pdim[l];
    }
    for (uint32_t l = 0; l < L; ++l) {
        fa.wfrag[l] = (const bf16x8*)m->wfrag[l].p;
        fa.bias[l] = (const float*)m->w.p + m->off_b[l];
    }
    const size_t lds = (size_t)kFusedRows * kFusedLd * sizeof(bf16);
    const bool relu = m->activation == PP_MLP_RELU;
    const void* fn = relu ? reinterpret_cast<const void*>(mlp_forward_fused_kernel<true>)
                          : reinterpret_cast<const void*>(mlp_forward_fused_kernel<false>);
    PP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const uint32_t grid = (M + kFusedRows - 1) / kFusedRows;
    PP_HIP(hipEventRecord(m->ev0, m->stream));
    if (relu)
        mlp_forward_fused_kernel<true><<<grid, kFusedThreads, lds, m->stream>>>(fa);
    else
        mlp_forward_fused_kernel<false><<<grid, kFusedThreads, lds, m->stream>>>(fa);
    PP_HIP(hipGetLastError());
    PP_HIP(hipEventRecord(m->ev1, m->stream));
    PP_HIP(hipStreamSynchronize(m->stream));
    return PP_OK;
}

int pack_input(pp_mlp* m, const float* d_x, uint32_t M) {
    const uint32_t R = pad16(M);
    pack_input_kernel<<<grid1d((uint64_t)R * m->pdim[0]), 256, 0, m->stream>>>(
        d_x, M, R, m->dims[0], (bf16*)m->x16.p, m->pdim[0]);
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// Loss and gradients into m->grad (packed, fp32) for M rows in x16.
int loss_grad(pp_mlp* m, uint32_t M, const float* d_logn, const float* d_k, const float* d_times,
              uint32_t n_times, float time_weight) {
    const uint32_t L = m->n_layers, R = pad16(M);
    int rc;
    const uint32_t ldo = m->pdim[L];
    if ((rc = forward(m, M, true, (float*)m->yout.p, ldo))) return rc;
    zero_f64_kernel<<<1, 64, 0, m->stream>>>((double*)m->loss.p, 2);
    loss_kernel<<<(R + 255) / 256, 256, 0, m->stream>>>(
        (const float*)m->yout.p, ldo, d_logn, d_k, d_times, n_times, M, R, time_weight,
        (bf16*)m->dY.p, ldo, (double*)m->loss.p);
    zero_f32_kernel<<<grid1d(m->n_params), 256, 0, m->stream>>>((float*)m->grad.p, m->n_params);
    PP_HIP(hipGetLastError());
    // transposed activations for the weight gradients: act_l^T [pdim_l][R]
    {
        const bf16* a0 = (const bf16*)m->x16.p;
        for (uint32_t l = 0; l < L; ++l) {
            const bf16* src = l == 0 ? a0 : (const bf16*)m->act[l].p;
            dim3 g((m->pdim[l] + 31) / 32, (R + 31) / 32);
            transpose_kernel<<<g, 256, 0, m->stream>>>(src, R, m->pdim[l], m->pdim[l],
                                                       (bf16*)m->actT[l].p, R);
        }
    }
    const bf16* dz = (const bf16*)m->dY.p;  // gradient w.r.t. the current layer's output
    uint32_t ldz = ldo;
    for (int l = (int)L - 1; l >= 0; --l) {
        const uint32_t K = m->pdim[l], N = m->dims[l + 1], Np = m->pdim[l + 1];
        // dZ^T [Np][R]
        {
            dim3 g((Np + 31) / 32, (R + 31) / 32);
            transpose_kernel<<<g, 256, 0, m->stream>>>(dz, R, Np, ldz, (bf16*)m->dZT.p, R);
        }
        // dWt_l [N][K] = dZ^T · act_l  (inner dimension = batch), as split-K
        // partials added in split order: the same gradient on every run
        float* gp = (float*)m->gpad.p;
        const uint64_t stride = (uint64_t)N * K;
        const uint32_t splits = std::max(1u, std::min(kMaxSplits, R / 256));
        uint32_t used = 0;
        if ((rc = gemm(m, (const bf16*)m->dZT.p, R, (const bf16*)m->actT[l].p, R, N, K, R,
                       EPI_PARTIAL, gp, K, nullptr, nullptr, 0, splits, stride, &used)))
            return rc;
        unpack_grad_splits_kernel<<<grid1d((uint64_t)N * m->dims[l]), 256, 0, m->stream>>>(
            gp, used, stride, N, m->dims[l], K, (float*)m->grad.p + m->off_w[l]);
        {
            const uint32_t parts = std::max(1u, std::min(kMaxSplits, M / 64u));
            const uint32_t chunk = (M + parts - 1) / parts;
            bias_grad_partial_kernel<<<dim3((N + 63) / 64, parts), 256, 0, m->stream>>>(
                dz, M, N, ldz, chunk, (float*)m->bpart.p);
            bias_grad_sum_kernel<<<(N + 3) / 4, 256, 0, m->stream>>>(
                (const float*)m->bpart.p, parts, N, (float*)m->grad.p + m->off_b[l]);
        }
        if (l == 0) break;
        // dA_l = dZ · Wt_l  (B operand = W_l = Wt_l^T, [K][Np])
        if ((rc = gemm(m, dz, ldz, (const bf16*)m->w16[l].p, Np, M, K, Np, EPI_STORE,
                       (float*)m->dA.p, K, nullptr, nullptr, 0)))
            return rc;
        act_backward_kernel<<<grid1d((uint64_t)R * K), 256, 0, m->stream>>>(
            (const float*)m->dA.p, (const float*)m->z[l].p, M, R, K, K, m->activation, (bf16*)m->dZ.p);
        dz = (const bf16*)m->dZ.p;
        ldz = K;
    }
    PP_HIP(hipGetLastError());
    return PP_OK;
}

// One Adam step (beta 0.9/0.999) from m->grad, then the bf16 weight copies;
// enqueued on the network's stream.
int adam_update(pp_mlp* m, float lr) {
    const float b1 = 0.9f, b2 = 0.999f;
    ++m->step;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    adam_kernel<<<grid1d(m->n_params), 256, 0, m->stream>>>(
        (float*)m->w.p, (const float*)m->grad.p, (float*)m->adam_m.p, (float*)m->adam_v.p,
        m->n_params, lr, b1, b2, 1e-8f, c1, c2);
    PP_HIP(hipGetLastError());
    return sync_weights(m);
}

}  // namespace

extern "C" {

int pp_mlp_create(pp_ctx* ctx, uint32_t n_in, uint32_t width, uint32_t depth, uint32_t n_out,
                  uint64_t seed, pp_mlp** out) {
    if (!ctx || !out) return fail(PP_ERR_INVALID_ARG, "ctx/out is NULL");
    *out = nullptr;
    if (n_in < 1 || n_in > 4096 || width < 16 || width > 4096 || width % 16 || depth < 1 ||
        depth > kMaxLayers - 1 || n_out < 1 || n_out > 16)
        return fail(PP_ERR_INVALID_ARG,
                    "unsupported MLP shape n_in=%u width=%u depth=%u n_out=%u (width multiple of "
                    "16, depth <= %d, n_out <= 16)", n_in, width, depth, n_out, kMaxLayers - 1);
    PP_HIP(hipSetDevice(ctx->device));
    pp_mlp* m = new pp_mlp();
    m->ctx = ctx;
    m->n_layers = depth + 1;
    m->dims[0] = n_in;
    for (uint32_t l = 1; l <= depth; ++l) m->dims[l] = width;
    m->dims[depth + 1] = n_out;
    for (uint32_t l = 0; l <= m->n_layers; ++l) m->pdim[l] = pad16(m->dims[l]);
    uint64_t off = 0;
    for (uint32_t l = 0; l < m->n_layers; ++l) {
        m->off_w[l] = off;
        off += (uint64_t)m->dims[l + 1] * m->dims[l];
        m->off_b[l] = off;
        off += m->dims[l + 1];
    }
    m->n_params = off;
    int rc = PP_OK;
    auto bail = [&](int code) {
        pp_mlp_destroy(m);
        return code;
    };
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
        return bail(fail(PP_ERR_DEVICE, "stream/event creation failed"));
    if ((rc = ensure(m->w, off * 4)) || (rc = ensure(m->grad, off * 4)) ||
        (rc = ensure(m->adam_m, off * 4)) || (rc = ensure(m->adam_v, off * 4)))
        return bail(rc);
    uint64_t gmax = 0;
    for (uint32_t l = 0; l < m->n_layers; ++l) {
        gmax = std::max<uint64_t>(gmax, (uint64_t)m->dims[l + 1] * m->pdim[l]);
        const FragGeom g = frag_geom(m->pdim[l + 1], m->pdim[l], l + 1 == m->n_layers);
        if ((rc = ensure(m->wt16[l], (size_t)m->pdim[l + 1] * m->pdim[l] * sizeof(bf16))) ||
            (rc = ensure(m->w16[l], (size_t)m->pdim[l] * m->pdim[l + 1] * sizeof(bf16))) ||
            (rc = ensure(m->wfrag[l], (size_t)g.tiles * g.steps * 64u * sizeof(bf16x8))))
            return bail(rc);
        if (hipMemset(m->wt16[l].p, 0, m->wt16[l].cap) != hipSuccess)
            return bail(fail(PP_ERR_DEVICE, "hipMemset failed"));
    }
    if ((rc = ensure(m->gpad, gmax * 4 * kMaxSplits))) return bail(rc);
    uint32_t nmax = 0;
    for (uint32_t l = 1; l <= m->n_layers; ++l) nmax = std::max(nmax, m->dims[l]);
    if ((rc = ensure(m->bpart, (size_t)nmax * 4 * kMaxSplits))) return bail(rc);
    // He-style uniform init, biases zero.
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (uint32_t l = 0; l < m->n_layers; ++l) {
        const float lim = std::sqrt(6.0f / (float)m->dims[l]) * (l + 1 == m->n_layers ? 0.1f : 1.0f);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (uint64_t i = 0; i < (uint64_t)m->dims[l + 1] * m->dims[l]; ++i) h[m->off_w[l] + i] = U(rng);
    }
    if (hipMemcpy(m->w.p, h.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->adam_m.p, 0, off * 4) != hipSuccess ||
        hipMemset(m->adam_v.p, 0, off * 4) != hipSuccess)
        return bail(fail(PP_ERR_DEVICE, "weight upload failed"));
    if ((rc = sync_weights(m))) return bail(rc);
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return bail(fail(PP_ERR_DEVICE, "stream sync failed"));
    *out = m;
    return PP_OK;
}

int pp_mlp_destroy(pp_mlp* m) {
    if (!m) return PP_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    DevBuf* bufs[] = {&m->w, &m->grad, &m->adam_m, &m->adam_v, &m->gpad, &m->bpart, &m->x16, &m->dA,
                      &m->dZ, &m->dZT, &m->dY, &m->yout, &m->loss, &m->loss_acc};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (int l = 0; l <= kMaxLayers; ++l) {
        if (l < kMaxLayers) {
            if (m->wt16[l].p) (void)hipFree(m->wt16[l].p);
            if (m->w16[l].p) (void)hipFree(m->w16[l].p);
            if (m->wfrag[l].p) (void)hipFree(m->wfrag[l].p);
        }
        if (m->act[l].p) (void)hipFree(m->act[l].p);
        if (m->actT[l].p) (void)hipFree(m->actT[l].p);
        if (m->z[l].p) (void)hipFree(m->z[l].p);
    }
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return PP_OK;
}

int pp_mlp_set_activation(pp_mlp* m, uint32_t activation) {
    if (!m) return fail(PP_ERR_INVALID_ARG, "mlp is NULL");
    if (activation != PP_MLP_SILU && activation != PP_MLP_RELU)
        return fail(PP_ERR_INVALID_ARG, "unknown activation %u", activation);
    m->activation = activation;
    return PP_OK;
}

int pp_mlp_n_params(pp_mlp* m, uint64_t* n) {
    if (!m || !n) return fail(PP_ERR_INVALID_ARG, "NULL argument");
    *n = m->n_params;
    return PP_OK;
}

int pp_mlp_get_params(pp_mlp* m, float* host, uint64_t n) {
    if (!m || !host || n != m->n_params)
        return fail(PP_ERR_INVALID_ARG, "params buffer must hold exactly %llu floats",
                    m ? (unsigned long long)m->n_params : 0ull);
    PP_HIP(hipSetDevice(m->ctx->device));
    PP_HIP(hipMemcpyAsync(host, m->w.p, n * 4, hipMemcpyDeviceToHost, m->stream));
    PP_HIP(hipStreamSynchronize(m->stream));
    return PP_OK;
}

int pp_mlp_set_params(pp_mlp* m, const float* host, uint64_t n) {
    if (!m || !host || n != m->n_params)
        return fail(PP_ERR_INVALID_ARG, "params buffer must hold exactly %llu floats",
                    m ? (unsigned long long)m->n_params : 0ull);
    PP_HIP(hipSetDevice(m->ctx->device));
    PP_HIP(hipMemcpyAsync(m->w.p, host, n * 4, hipMemcpyHostToDevice, m->stream));
    int rc = sync_weights(m);
    if (rc) return rc;
    PP_HIP(hipStreamSynchronize(m->stream));
    return PP_OK;
}

int pp_mlp_forward_device(pp_mlp* m, const float* d_x, uint64_t n_rows, float* d_y) {
    if (!m || (!d_x && n_rows) || (!d_y && n_rows)) return fail(PP_ERR_INVALID_ARG, "NULL argument");
    if (n_rows == 0) return PP_OK;
    if (n_rows > (1ull << 26)) return fail(PP_ERR_INVALID_ARG, "at most 2^26 rows per call");
    PP_HIP(hipSetDevice(m->ctx->device));
    const uint32_t M = (uint32_t)n_rows;
    if (fused_fits(m)) return forward_fused(m, d_x, M, d_y);
    int rc;
    if ((rc = mlp_reserve(m, M))) return rc;
    if ((rc = pack_input(m, d_x, M))) return rc;
    PP_HIP(hipEventRecord(m->ev0, m->stream));
    const uint32_t L = m->n_layers;
    // outputs straight into d_y [M][n_out]
    if ((rc = forward(m, M, false, (float*)m->yout.p, m->pdim[L]))) return rc;
    PP_HIP(hipEventRecord(m->ev1, m->stream));
    PP_HIP(hipMemcpy2DAsync(d_y, m->dims[L] * 4, m->yout.p, m->pdim[L] * 4, m->dims[L] * 4, M,
                            hipMemcpyDeviceToDevice, m->stream));
    PP_HIP(hipStreamSynchronize(m->stream));
    return PP_OK;
}

int pp_mlp_forward(pp_mlp* m, const float* x, uint64_t n_rows, float* y) {
    if (!m || (!x && n_rows) || (!y && n_rows)) return fail(PP_ERR_INVALID_ARG, "NULL argument");
    if (n_rows == 0) return PP_OK;
    PP_HIP(hipSetDevice(m->ctx->device));
    void *dx = nullptr, *dy = nullptr;
    PP_HIP(hipMalloc(&dx, n_rows * m->dims[0] * 4));
    if (hipMalloc(&dy, n_rows * m->dims[m->n_layers] * 4) != hipSuccess) {
        (void)hipFree(dx);
        return fail(PP_ERR_OOM, "hipMalloc failed");
    }
    int rc = PP_OK;
    if (hipMemcpy(dx, x, n_rows * m->dims[0] * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(PP_ERR_DEVICE, "copy failed");
    if (!rc) rc = pp_mlp_forward_device(m, (const float*)dx, n_rows, (float*)dy);
    if (!rc && hipMemcpy(y, dy, n_rows * m->dims[m->n_layers] * 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(PP_ERR_DEVICE, "copy failed");
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

int pp_mlp_loss_grad_device(pp_mlp* m, const float* d_x, const float* d_logn, const float* d_k,
                            const float* d_times, uint32_t n_times, uint32_t n_rows,
                            float time_weight, float* grads_host, double* loss_out) {
    if (!m || !d_x || !d_logn || !d_k || (n_times && !d_times) || n_rows == 0)
        return fail(PP_ERR_INVALID_ARG, "NULL argument or empty batch");
    PP_HIP(hipSetDevice(m->ctx->device));
    int rc;
    if ((rc = mlp_reserve(m, n_rows))) return rc;
    if ((rc = pack_input(m, d_x, n_rows))) return rc;
    if ((rc = loss_grad(m, n_rows, d_logn, d_k, d_times, n_times, time_weight))) return rc;
    if (grads_host)
        PP_HIP(hipMemcpyAsync(grads_host, m->grad.p, m->n_params * 4, hipMemcpyDeviceToHost,
                              m->stream));
    double l[2];
    PP_HIP(hipMemcpyAsync(l, m->loss.p, sizeof l, hipMemcpyDeviceToHost, m->stream));
    PP_HIP(hipStreamSynchronize(m->stream));
    if (loss_out) {
        loss_out[0] = l[0] / n_rows;
        loss_out[1] = l[1] / n_rows;
    }
    return PP_OK;
}

int pp_mlp_train_step_device(pp_mlp* m, const float* d_x, const float* d_logn, const float* d_k,
                             const float* d_times, uint32_t n_times, uint32_t n_rows, float lr,
                             float time_weight, double* loss_out) {
    int rc = pp_mlp_loss_grad_device(m, d_x, d_logn, d_k, d_times, n_times, n_rows, time_weight,
                                     nullptr, loss_out);
    if (rc) return rc;
    if ((rc = adam_update(m, lr))) return rc;
    PP_HIP(hipStreamSynchronize(m->stream));
    return PP_OK;
}

int pp_mlp_train_steps_device(pp_mlp* m, const float* d_x, const float* d_logn, const float* d_k,
                              const float* d_times, uint32_t n_times, uint32_t batch,
                              uint32_t n_steps, const float* lr, float time_weight,
                              double* loss_out) {
    if (!m || !d_x || !d_logn || !d_k || (n_times && !d_times) || !lr || batch == 0 ||
        n_steps == 0)
        return fail(PP_ERR_INVALID_ARG, "NULL argument, empty batch or no steps");
    PP_HIP(hipSetDevice(m->ctx->device));
    int rc;
    if ((rc = mlp_reserve(m, batch))) return rc;
    zero_f64_kernel<<<1, 64, 0, m->stream>>>((double*)m->loss_acc.p, 2);
    for (uint32_t i = 0; i < n_steps; ++i) {
        const size_t o = (size_t)i * batch;
        if ((rc = pack_input(m, d_x + o * m->dims[0], batch))) return rc;
        if ((rc = loss_grad(m, batch, d_logn + o, d_k + o, n_times ? d_times + o * n_times : nullptr,
                            n_times, time_weight)))
            return rc;
        add_f64_kernel<<<1, 64, 0, m->stream>>>((const double*)m->loss.p, 2,
                                                 (double*)m->loss_acc.p);
        if ((rc = adam_update(m, lr[i]))) return rc;
    }
    double l[2];
    PP_HIP(hipMemcpyAsync(l, m->loss_acc.p, sizeof l, hipMemcpyDeviceToHost, m->stream));
    PP_HIP(hipStreamSynchronize(m->stream));
    if (loss_out) {
        loss_out[0] = l[0] / ((double)batch * n_steps);
        loss_out[1] = l[1] / ((double)batch * n_steps);
    }
    return PP_OK;
}

int pp_mlp_last_forward_ms(pp_mlp* m, float* ms) {
    if (!m || !ms) return fail(PP_ERR_INVALID_ARG, "NULL argument");
    PP_HIP(hipEventElapsedTime(ms, m->ev0, m->ev1));
    return PP_OK;
}

}  // extern "C"
