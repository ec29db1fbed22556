// pp_types.hpp — host-side definitions that need no HIP: error reporting, the
// request record the kernels read, and the checks of the C-ABI inputs.  The
// library compiles them with hipcc; the CPU test harness of the serving
// scheduler (tests/native/) compiles them with g++ and the sanitizers.
#pragma once

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/photonprop.h"

namespace pph {

// Message of the last failure on this thread (pp_last_error).
inline std::string& last_error() {
    thread_local std::string msg;
    return msg;
}

// Record a formatted message and return `code`.
inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    last_error() = buf;
    return code;
}

constexpr uint64_t kTilePhotons = PP_TILE_PHOTONS;  // photons per work tile

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

// Spectrum as the kernel reads it.
struct SpecParams {
    uint32_t kind;
    uint32_t n_knots;
    float a;          // INV_LAMBDA2: 1/lo
    float d;          // INV_LAMBDA2: 1/lo - 1/hi
    float mono;
    float knot_step;  // KNOTS: 1/(n_knots-1)
};

inline int check_medium(const pp_medium_tables* m) {
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

// Validate a spectrum and fill the kernel parameters.
inline int check_spectrum(const pp_spectrum* s, SpecParams* out) {
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
        out->n_knots = s->n_knots;
        out->knot_step = 1.0f / (float)(s->n_knots - 1);
        return PP_OK;
    }
    default:
        return fail(PP_ERR_UNSUPPORTED, "Cannot initialize spectrum kind %u", s->kind);
    }
}

// Check one request and build its kernel record.
inline int make_devreq(const pp_request* q, uint32_t idx, DevReq* d) {
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
    if (q->max_steps > PP_MAX_STEPS)
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

// Cherenkov cone needs n(λ)·β >= 1 everywhere on the medium table (n_min its
// smallest n).
inline int check_cone(const pp_request* reqs, uint32_t n_req, float n_min) {
    for (uint32_t i = 0; i < n_req; ++i) {
        if (reqs[i].emitter.profile != PP_EMIT_CHERENKOV_CONE) continue;
        if (!((double)n_min * (double)reqs[i].emitter.beta >= 1.0 - 1e-6))
            return fail(PP_ERR_INVALID_ARG,
                        "request %u: no Cherenkov emission (n_min=%g, beta=%g, n*beta < 1)", i,
                        (double)n_min, (double)reqs[i].emitter.beta);
    }
    return PP_OK;
}

inline int check_records(uint64_t n, const char* what) {
    if (n > PP_MAX_HIT_RECORDS)
        return fail(PP_ERR_INVALID_ARG, "%s %llu exceeds PP_MAX_HIT_RECORDS (%llu)", what,
                    (unsigned long long)n, (unsigned long long)PP_MAX_HIT_RECORDS);
    return PP_OK;
}

// Records the compacted output of a batch can hold.
inline uint64_t out_records(uint64_t hits_cap, uint64_t per_cap, uint32_t n_req) {
    if (!per_cap) return hits_cap;
    const unsigned __int128 all = (unsigned __int128)per_cap * n_req;  // no wrap-around
    return all < hits_cap ? (uint64_t)all : hits_cap;
}

}  // namespace pph
