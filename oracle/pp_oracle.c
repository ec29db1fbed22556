/*
 * pp_oracle.c — CPU oracle for the photon-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * the timed CPU comparison.  The product (photonprop.jl_amd/) never links,
 * loads or calls it.
 *
 * What it restates (PhotonProp.jl, file:line):
 *   or_photon_*                 propagation loop, src/photon_prop_cuda.jl:301-380
 *   or_hg_costheta_*            src/photon_prop_cuda.jl:31-38
 *   or_rotate_to_axis_*         src/photon_prop_cuda.jl:171-219
 *   or_update_position_*        src/photon_prop_cuda.jl:255-259
 *   or_ray_sphere_*             src/photon_prop_cuda.jl:327-351
 *   or_fast_linear_interp_*     src/utils.jl:10-34 and :37-60
 *   or_wavelength_*             src/photon_prop_cuda.jl:70-75 (+ closed-form 1/λ²)
 *   weight / time at the hit    src/photon_prop_cuda.jl:622, :634-638
 *   or_rodrigues_rotation       src/photon_prop_cuda.jl:161-169 (textbook form)
 *   or_philox4x32_10            Salmon, Moraes, Dror, Shaw, "Parallel random
 *                               numbers: as easy as 1, 2, 3", SC'11 (Philox4x32
 *                               with 10 rounds).  The reference's own device RNG
 *                               (CUDA.jl 3.10.1, not vendored) is replaced by
 *                               design; its stream is not reproduced.
 *
 * Deviations from the reference kernel, shared with the device kernel:
 *   D1 Cherenkov-cone emission about the emitter axis (the reference's
 *      :IsotropicEmission draws θ ~ U(-1,1) rad and ignores the axis; it is
 *      available as PP_EMIT_REF_ISOTROPIC).
 *   D2 g comes from the medium (reference hard-codes 0.99 at :229).
 *   D3 A photon whose line meets the sphere outside the current step moves and
 *      scatters (notebook CPU path, photon_prob.ipynb:3984-4005); the kernel
 *      at :345-356 neither moves nor scatters it.
 *   D4 A sphere behind the photon is no hit (same as the reference).
 *   D5 max_steps is a parameter (reference: 10, :319).
 *   D6 The hit record keeps arrival and initial direction.
 *   D7 t = t0 + L·n(λ)/c with the phase index, c = 0.299792458 m/ns.
 *   D8 Absorption weight exp(-L/λ_abs) at the hit, summed in FP64.
 *   D9 Exactly n_photons photons (reference over-simulates, :288, :624-631).
 *   D10 Hits beyond the capacity are counted, never silently dropped.
 *
 * How each function is pinned (tests/test_oracle_*.py): Philox against the
 * rocRAND header's known answers (tests/golden/philox_kat.json); HG against its
 * closed-form mean and variance; rotation against the reference's unit-test
 * KATs (src/tests/test_photon_prob.jl:26-41) and the Rodrigues formula;
 * interpolation against the KATs at src/tests/test_photon_prob.jl:75-90;
 * ray-sphere against brute-force geometry; whole propagation against closed
 * forms (no-scatter cone geometry, straight-line survival exp(-d/λ_abs)).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/photonprop.h"

#define PP_TWO_PI 6.283185307179586
#define PP_C_VACUUM 0.299792458 /* m/ns */

/* Philox4x32 round constants (Salmon et al. 2011, Table 2 / Random123). */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Uniform on (0, 1]: the top 24 bits plus one, times 2^-24 (exact in FP32).
 * Used by the device self-tests of single helpers. */
float or_uniform(uint32_t w) { return fmaf((float)(w >> 8), 0x1p-24f, 0x1p-24f); }

/* The propagation's draws from one Philox block (words w0..w3):
 *   u[k] = 1 - (w_k >> 9)·2^-23 on (0, 1], k = 0..3 (the top 23 bits; exact in
 *          FP32; the device forms a_k = 1 + (w_k >> 9)·2^-23 and u = 2 - a_k);
 *   p[0] = ((w0 & 0xFF) << 8 | (w1 & 0xFF))·2^-16, p[1] likewise from w2, w3:
 *          16-bit azimuths in revolutions on [0, 1) from the low bytes. */
void or_block_draws(const uint32_t w[4], float u[4], float p[2]) {
    for (int k = 0; k < 4; ++k) u[k] = 1.0f - (float)(w[k] >> 9) * 0x1p-23f;
    p[0] = (float)(((w[0] & 0xFFu) << 8) | (w[1] & 0xFFu)) * 0x1p-16f;
    p[1] = (float)(((w[2] & 0xFFu) << 8) | (w[3] & 0xFFu)) * 0x1p-16f;
}

/* Rodrigues rotation taking a to b, applied to operand, in FP64
 * (src/photon_prop_cuda.jl:161-169). */
void or_rodrigues_rotation(const double* a, const double* b, const double* v, double* out) {
    double ax[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                    a[0] * b[1] - a[1] * b[0]};
    double axn = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    double dotab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    if (axn == 0.0) { /* parallel: identity; antiparallel: undefined axis */
        out[0] = v[0];
        out[1] = v[1];
        out[2] = v[2];
        return;
    }
    for (int j = 0; j < 3; ++j) ax[j] /= axn;
    /* theta from atan2 so that rotations beyond 90 degrees are correct (the
     * reference's asin(|a×b|) is exact only for angles up to 90°). */
    double th = atan2(axn, dotab);
    double c = cos(th), s = sin(th);
    double cr[3] = {ax[1] * v[2] - ax[2] * v[1], ax[2] * v[0] - ax[0] * v[2],
                    ax[0] * v[1] - ax[1] * v[0]};
    double d = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
    for (int j = 0; j < 3; ++j) out[j] = v[j] * c + cr[j] * s + ax[j] * (1 - c) * d;
}

/* ---- FP32 instantiation -------------------------------------------------- */
#define REAL float
#define SUF _f32
#define R_FMA fmaf
#define R_SQRT sqrtf
#define R_LOG logf
#define R_EXP expf
#define R_SIN sinf
#define R_COS cosf
#define R_FLOOR floorf
#define R_FABS fabsf
#include "pp_oracle_impl.h"
#undef REAL
#undef SUF
#undef R_FMA
#undef R_SQRT
#undef R_LOG
#undef R_EXP
#undef R_SIN
#undef R_COS
#undef R_FLOOR
#undef R_FABS

/* ---- FP64 instantiation -------------------------------------------------- */
#define REAL double
#define SUF _f64
#define R_FMA fma
#define R_SQRT sqrt
#define R_LOG log
#define R_EXP exp
#define R_SIN sin
#define R_COS cos
#define R_FLOOR floor
#define R_FABS fabs
#include "pp_oracle_impl.h"
