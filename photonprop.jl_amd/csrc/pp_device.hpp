// pp_device.hpp — device helpers inlined by the gfx950 propagation kernel.
//
// Built with -ffp-contract=off: every fused multiply-add below is written out,
// so the operation order is fixed and matches the CPU oracle wherever the two
// use the same operations.  Per-step transcendentals use the hardware
// instructions (v_log_f32, v_sin_f32/v_cos_f32 in revolutions, v_rcp_f32,
// v_rsq_f32, v_sqrt_f32); per-photon and per-hit math uses IEEE division and
// square root.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppd {

// Philox4x32 constants (Salmon et al., SC'11).
constexpr uint32_t kM0 = 0xD2511F53u;
constexpr uint32_t kM1 = 0xCD9E8D57u;
constexpr uint32_t kW0 = 0x9E3779B9u;
constexpr uint32_t kW1 = 0xBB67AE85u;

struct U4 {
    uint32_t x, y, z, w;
};

// a ^ b ^ c in one gfx950 v_bitop3_b32 (truth table 0x96 over the operand
// bit patterns 0xF0, 0xCC, 0xAA); plain XORs would take two instructions.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Philox4x32-10.  The key (k0, k1) is wave-uniform, so its schedule lives in
// scalar registers; each round costs two 32x32->64 multiplies (v_mad_u64_u32)
// and two 3-input XORs per lane.  The propagation calls it with the counter
// words c0 (block index), c2 (high photon-id word) and c3 (stream) uniform
// and only c1 (low photon-id word) varying, so in the first rounds some words
// and their multiplies stay uniform:
//   round 0: only word 0 varies (c1 ^ uniform);
//   round 1: words 2, 3 vary; word 0 is uniform again;
//   round 2: words 0, 1, 2 vary; round 3 on: all four.
// Each XOR is grouped so that its uniform operands combine first (one scalar
// XOR, then at most one vector XOR); a v_bitop3_b32 would take only one
// scalar operand and need a v_mov for the other.  The grouping is only
// associativity, so any counter (uniform or not) gives the same words.
// With `rk` (the key's schedule in LDS, rk[r] = (k0 + r·W0, k1 + r·W1), see
// philox_keys_to_lds), the keys that only feed vector XORs (k0 of rounds 3-9,
// k1 of rounds 4-9) come from LDS reads instead of scalar adds: a round's key
// add and its use had been a scalar-to-vector dependency in every round, and
// the 13 adds per block took scalar issue slots the step loop is short of.
// The keys of the first rounds, which combine with uniform words on the scalar
// unit, stay scalar adds.  Without `rk` every key is a scalar add.
template <bool kRk>
__device__ __forceinline__ U4 philox10_t(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1, const uint2* rk) {
    // opaque key: each round's key is one scalar add from the last, instead of
    // a 20-register key schedule that the compiler would hoist out of the loops
    k0 = __builtin_amdgcn_readfirstlane(k0);
    k1 = __builtin_amdgcn_readfirstlane(k1);
    __asm__ volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kM0 * c0;
        const uint64_t p1 = (uint64_t)kM1 * c2;
        const uint32_t h1 = (uint32_t)(p1 >> 32), h0 = (uint32_t)(p0 >> 32);
        uint32_t n0, n2;
        if (r == 0) {
            n0 = c1 ^ (h1 ^ k0);
            n2 = (h0 ^ k1) ^ c3;
        } else if (r == 1) {
            n0 = (h1 ^ k0) ^ c1;
            n2 = h0 ^ (c3 ^ k1);
        } else if (r == 2) {
            n0 = h1 ^ (c1 ^ k0);
            n2 = c3 ^ (h0 ^ k1);
        } else if (r == 3) {
            n0 = xor3(h1, c1, kRk ? rk[3].x : k0);
            n2 = h0 ^ (c3 ^ k1);
        } else {
            const uint2 kk = kRk ? rk[r] : make_uint2(k0, k1);
            n0 = xor3(h1, c1, kk.x);
            n2 = xor3(h0, c3, kk.y);
        }
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        if (!kRk || r < 3) {
            k0 += kW0;
            k1 += kW1;
        }
    }
    return U4{c0, c1, c2, c3};
}

__device__ __forceinline__ U4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1) {
    return philox10_t<false>(c0, c1, c2, c3, k0, k1, nullptr);
}

__device__ __forceinline__ U4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, const uint2* rk) {
    return philox10_t<true>(c0, c1, c2, c3, k0, k1, rk);
}

// The schedule of key (k0, k1) for philox10's `rk`: lane r < 10 of a wave
// writes round key r.  A wave reads its own entries only, and LDS operations
// of one wave complete in order, so no barrier is needed before the reads.
__device__ __forceinline__ void philox_keys_to_lds(uint2* rk, uint32_t lane, uint32_t k0,
                                                   uint32_t k1) {
    __asm__ volatile("" : "+v"(lane));
    if (lane < 10u) rk[lane] = make_uint2(k0 + lane * kW0, k1 + lane * kW1);
}

// Uniform on (0, 1]: (top 24 bits + 1) * 2^-24, exact in FP32 (self-tests).
__device__ __forceinline__ float u01(uint32_t w) {
    return __builtin_fmaf((float)(w >> 8), 0x1p-24f, 0x1p-24f);
}

// The propagation's draws from one Philox block (words w0..w3):
//   a_k = 1 + (w_k >> 9)·2^-23 on [1, 2), exact in FP32: the word's top 23 bits
//         as the mantissa under the exponent of 1.0 (one v_alignbit_b32); the
//         matching uniform on (0, 1] is u_k = 2 - a_k, also exact;
//   p0  = ((w0 & 0xFF) << 8 | (w1 & 0xFF))·2^-16 and p1 likewise from w2, w3:
//         16-bit azimuths in revolutions on [0, 1), the unit the hardware
//         sin/cos take.  They are carried as 1 + p on [1, 2) (one v_perm_b32
//         for the 16 bits, one v_lshl_or_b32 to put them under the exponent of
//         1.0, exact), which is the same angle, so no integer-to-float
//         conversion is needed.
__device__ __forceinline__ float draw_a(uint32_t w) {
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, w, 9u));
}
__device__ __forceinline__ float draw_p(uint32_t hi, uint32_t lo) {
    const uint32_t p = __builtin_amdgcn_perm(hi, lo, 0x0C0C0400u);
    return __uint_as_float((p << 7) | 0x3F800000u);
}

// Henyey-Greenstein cosθ by inverse CDF (src/photon_prop_cuda.jl:31-38):
// cosθ = (1+g² - ((1-g²)/(1+g(2η-1)))²)/(2g), clamped; g == 0 is isotropic.
// The uniform comes as a = 2 - η on [1, 2), so 2η-1 = 3-2a, exact in FP32;
// the denominator 1 + g(2η-1) is then rounded once, as in the reference form.
// With r = 1/(1 + g(2η-1)) the same expression is C - K·r², where
// C = (1+g²)/(2g) and K = (1-g²)²/(2g) are rounded once on the host: one
// multiply per draw fewer than (1/(2g))·(1+g² - ((1-g²)·r)²), with the same
// error (about 1e-6 absolute for g in [0.3, 0.99]).
struct HG {
    float g, c, k;
    int iso;
};

// The same with the isotropic case fixed at compile time (kIso = h.iso), for
// loops that are instantiated once per case, and without the final clamp to
// [-1, 1]: C - K·r² can leave the interval by an ulp or two.  The step loop
// clamps 1 - cos² to [0, 1] instead, which costs nothing there (the clamp bit
// of the fma that forms it), and takes the unclamped cosine as the rotation's
// nz, an error the renormalisation absorbs.
template <bool kIso>
__device__ __forceinline__ float hg_costheta_a_t(const HG& h, float a) {
    const float x = __builtin_fmaf(-2.0f, a, 3.0f);
    if (kIso) return x;
    const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(h.g, x, 1.0f));
    // (K·r)·r: each instruction reads one scalar operand (the VALU limit)
    return __builtin_fmaf(-(h.k * r), r, h.c);
}

// sinθ from cosθ, 1 - cos² clamped to [0, 1].
__device__ __forceinline__ float sin_from_cos(float c) {
    return __builtin_amdgcn_sqrtf(__builtin_amdgcn_fmed3f(__builtin_fmaf(-c, c, 1.0f), 0.0f, 1.0f));
}

__device__ __forceinline__ float hg_costheta_a(const HG& h, float a) {
    const float c = h.iso ? hg_costheta_a_t<true>(h, a) : hg_costheta_a_t<false>(h, a);
    return __builtin_amdgcn_fmed3f(c, -1.0f, 1.0f);
}

// rotate_to_axis (src/photon_prop_cuda.jl:171-219): apply to (nx,ny,nz) the
// rotation taking e_z to (x,y,z), then renormalise.  The reference's form
// divides by sinθ = sqrt(1-z²) twice.  Written as Rodrigues' rotation about
// e_z × r = (-y, x, 0), with p = x·nx + y·ny and x² + y² = (1-z)(1+z), the
// same rotation reads
//     t = nz - p/(1+z),  x' = nx + x·t,  y' = ny + y·t,  z' = nz·z - p:
// one reciprocal, one multiply, one fma for t and one fma per component.
// 1/(1+z) is taken as 2/((1+z)² + x² + y²), equal to it for a unit vector and
// free of cancellation at both poles: near z = -1 it follows x and y (as
// (x²+y²)/(1-z) would), near z = 1 it is 1+z.  For |r|² = 1 + e the stand-in
// is off by e/2, like (x²+y²)/(1-z).  Near z = -1, p/(1+z) grows like
// 1/sqrt(1+z), but x·t and y·t stay bounded (x² + y² <= 2(1+z) + e), so the
// result keeps absolute errors of a few ulp.  (The factor 2 rides in the fma
// that forms t as the inline constant -2: the output modifier mul:2 is
// ignored in IEEE mode, which compute kernels run in.)
// y enters as y - 2^-60: that changes no y of magnitude above 2^-36 (the
// subtraction rounds back to y), and at the exact south pole (x = y = 0,
// z = -1), where the rotation is otherwise undetermined (0/0), it selects the
// rotation by π about the x axis, so the result is the reference's
// (nx, -ny, -nz) to rounding: x' = nx exactly, y' = -ny + 2^-60·(1 - nz),
// z' = -nz + 2^-60·ny.  (x² + y² >= 2^-120 and |p/(1+z)| <= 2^61 stay in
// range.)
// Unit inputs give |r|² = 1 + e with |e| of a few 1e-6 (rounding and the
// hardware sin/cos), so the renormalising 1/sqrt(|r|²) is one Newton step from
// 1, 1.5 - |r|²/2, whose error 3e²/8 is far below an FP32 ulp; it needs no
// transcendental.  kRenorm = false skips it: one rotation of a unit vector
// moves |r| by ~2e-7, so a caller may renormalise every few rotations (after
// eight, |e| stays within a few 1e-6 and the Newton step is still exact).
// kInBiased: the caller's y already carries the -2^-60.  kOutBiased: the
// result's y carries a bias too, ny - 2^-60 taking the place of ny (and in p,
// where it changes p by at most 2^-60), so y' = y·t + ny - 2^-60·(1 + t) when
// y entered biased; rotate_polar folds it into the product that forms ny, so a
// loop that feeds each result into the next rotation pays for the bias with
// no instruction.
template <bool kRenorm, bool kInBiased>
__device__ __forceinline__ void rotate_core(float x, float y_in, float z, float nx, float ny,
                                            float nz, float& ox, float& oy, float& oz) {
    const float y = kInBiased ? y_in : y_in - 0x1p-60f;
    const float s2 = __builtin_fmaf(y, y, x * x);
    const float zp = 1.0f + z;
    const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(zp, zp, s2));  // 1/(2(1+z))
    const float p = __builtin_fmaf(x, nx, y * ny);
    // -t = 2p/(1+z) - nz; the components take it negated (a source
    // modifier), which keeps their fmas in the three-address form, so the
    // results can land in the registers of x and y
    const float tn = __builtin_fmaf(p * r, 2.0f, -nz);
    const float rx = __builtin_fmaf(x, -tn, nx);
    const float ry = __builtin_fmaf(y, -tn, ny);
    const float rz = __builtin_fmaf(nz, z, -p);
    if (!kRenorm) {
        ox = rx;
        oy = ry;
        oz = rz;
        return;
    }
    const float n2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
    const float inv = __builtin_fmaf(-0.5f, n2, 1.5f);
    ox = rx * inv;
    oy = ry * inv;
    oz = rz * inv;
}

template <bool kRenorm = true, bool kInBiased = false, bool kOutBiased = false>
__device__ __forceinline__ void rotate_to_axis(float x, float y_in, float z, float nx, float ny,
                                               float nz, float& ox, float& oy, float& oz) {
    rotate_core<kRenorm, kInBiased>(x, y_in, z, nx, kOutBiased ? ny - 0x1p-60f : ny, nz, ox, oy,
                                    oz);
}

// The same for n = (st·cφ, st·sφ, ct), the form every caller has: the output
// bias is folded into the product that forms ny (one fmaak).
template <bool kRenorm = true, bool kInBiased = false, bool kOutBiased = false>
__device__ __forceinline__ void rotate_polar(float x, float y_in, float z, float cphi, float sphi,
                                             float st, float ct, float& ox, float& oy, float& oz) {
    const float nx = cphi * st;
    const float ny = kOutBiased ? __builtin_fmaf(sphi, st, -0x1p-60f) : sphi * st;
    rotate_core<kRenorm, kInBiased>(x, y_in, z, nx, ny, ct, ox, oy, oz);
}

// Renormalise a direction of squared length 1 + e, |e| of a few 1e-6: one
// Newton step of 1/sqrt from 1 (error 3e²/8, far below an FP32 ulp).
__device__ __forceinline__ void renormalize(float& x, float& y, float& z) {
    const float n2 = __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x));
    const float inv = __builtin_fmaf(-0.5f, n2, 1.5f);
    x *= inv;
    y *= inv;
    z *= inv;
}

// Linear interpolation in the medium table, clamped.  Point i holds two
// float4: {n, λ_sca·ln2, λ_abs, 0} at t[2i] and the differences to point i+1
// (formed in FP32 on the host, as hi - lo would be here; zero at the last
// point) at t[2i+1], so a lookup is one clamp (v_med3), one address and three
// fmas, lo + f·(hi - lo) in the oracle's operation order (or_table_), with no
// clamp of the index: xs = n-1 gives i = n-1 and f = 0.
__device__ __forceinline__ float3 table_lerp_slopes(const float4* t, float xmax, float xs) {
    xs = __builtin_amdgcn_fmed3f(xs, 0.0f, xmax);
    const uint32_t i = (uint32_t)xs;
    const float f = xs - (float)i;
    const float4 v = t[2u * i];
    const float4 d = t[2u * i + 1u];
    return make_float3(__builtin_fmaf(f, d.x, v.x), __builtin_fmaf(f, d.y, v.y),
                       __builtin_fmaf(f, d.z, v.z));
}

// Linear interpolation in a uniform-grid table, clamped (same operation order
// as the oracle's or_table_).  The medium table interleaves {n, λ_sca, λ_abs, 0}.
__device__ __forceinline__ float4 table_lerp(const float4* t, uint32_t n, float xs) {
    const float xmax = (float)(n - 1);
    xs = fminf(fmaxf(xs, 0.0f), xmax);
    uint32_t i = (uint32_t)xs;
    i = i > n - 2 ? n - 2 : i;
    const float f = xs - (float)i;
    const float4 lo = t[i];
    const float4 hi = t[i + 1];
    float4 r;
    r.x = __builtin_fmaf(f, hi.x - lo.x, lo.x);
    r.y = __builtin_fmaf(f, hi.y - lo.y, lo.y);
    r.z = __builtin_fmaf(f, hi.z - lo.z, lo.z);
    r.w = 0.0f;
    return r;
}

// fast_linear_interp(u, knots, 0, 1) on equidistant knots (src/utils.jl:37-60).
__device__ __forceinline__ float knot_interp(const float* knots, uint32_t n, float step_size,
                                             float u) {
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    const float along = u / step_size;
    const float fl = floorf(along);
    const int64_t k = (int64_t)fl;
    if (k >= (int64_t)n - 1) return knots[n - 1];
    const float along_step = along - fl;
    const float y_low = knots[k];
    const float slope = knots[k + 1] - y_low;
    return y_low + slope * along_step;
}

}  // namespace ppd
