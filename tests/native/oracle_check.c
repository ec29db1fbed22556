/*
 * oracle_check.c — drives the CPU oracle (oracle/pp_oracle.c) through its
 * edge cases so that tests/test_native_host.py can run it under
 * AddressSanitizer + UndefinedBehaviorSanitizer.
 *
 * TEST CODE.  Reads Philox4x32-10 known answers from stdin, one per line as
 * ten decimal words (counter[4] key[2] out[4]), then checks, in FP32 and FP64:
 *   - propagation over empty, one-photon and ragged id ranges, with hit
 *     capacities 0, 3 and unlimited, 1 and 3 threads, every emission profile
 *     and spectrum kind, the largest medium table and knot count: counts and
 *     sums agree between the runs, stored hits are the first hits in photon-id
 *     order, the records of a sub-range equal that range's part of the whole,
 *     hit points lie on the sphere, weights in (0, 1];
 *   - interpolation clamped outside the table, rotation at the poles,
 *     ray-sphere from inside the sphere and along a tangent.
 * Prints PASS and exits 0 when all hold.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/photonprop.h"

void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]);
#define DECL(R, S)                                                                                \
    int64_t or_propagate##S(const pp_medium_tables*, const pp_spectrum*, const pp_emitter*,      \
                            const pp_sphere*, uint64_t, uint32_t, uint32_t, uint64_t, uint64_t,   \
                            pp_hit*, uint64_t, pp_summary*, int);                                \
    R or_fast_linear_interp_knots##S(R, const R*, uint32_t, R, R);                                \
    R or_fast_linear_interp_sorted##S(R, const R*, const R*, uint32_t);                           \
    void or_rotate_to_axis##S(const R*, const R*, R*);                                           \
    int or_ray_sphere##S(const R*, const R*, const R*, R, R, R*);                                 \
    R or_hg_costheta##S(R, R);
DECL(float, _f32)
DECL(double, _f64)

static int failures = 0;
#define CHECK(cond, ...)                                                        \
    do {                                                                        \
        if (!(cond)) {                                                          \
            ++failures;                                                         \
            fprintf(stderr, "CHECK failed line %d: %s: ", __LINE__, #cond);     \
            fprintf(stderr, __VA_ARGS__);                                       \
            fprintf(stderr, "\n");                                              \
        }                                                                       \
    } while (0)

static float tab_n[PP_MAX_TABLE_POINTS], tab_abs[PP_MAX_TABLE_POINTS], tab_sca[PP_MAX_TABLE_POINTS];
static float knots[PP_MAX_KNOTS];

static pp_medium_tables medium(uint32_t n_points) {
    for (uint32_t i = 0; i < n_points; ++i) {
        const float f = (float)i / (float)(n_points - 1);
        tab_n[i] = 1.36f - 0.02f * f;
        tab_abs[i] = 20.0f + 40.0f * f * (1.0f - f);
        tab_sca[i] = 30.0f + 20.0f * f;
    }
    pp_medium_tables m;
    memset(&m, 0, sizeof m);
    m.lambda_min = 300.0f;
    m.lambda_max = 600.0f;
    m.n_points = n_points;
    m.n_phase = tab_n;
    m.abs_len = tab_abs;
    m.sca_len = tab_sca;
    m.g = 0.95f;
    return m;
}

static pp_spectrum spectrum(uint32_t kind) {
    pp_spectrum s;
    memset(&s, 0, sizeof s);
    s.kind = kind;
    s.mono = 420.0f;
    s.lo = 300.0f;
    s.hi = 600.0f;
    for (int i = 0; i < PP_MAX_KNOTS; ++i) knots[i] = 300.0f + 300.0f * (float)i / (PP_MAX_KNOTS - 1);
    s.n_knots = PP_MAX_KNOTS;
    s.knots = knots;
    return s;
}

static int hits_equal(const pp_hit* a, const pp_hit* b) { return memcmp(a, b, sizeof *a) == 0; }

#define PROPAGATION_CHECKS(S, R)                                                                   \
    static void propagation##S(const pp_medium_tables* m, const pp_spectrum* sp, uint32_t profile) {\
        pp_emitter em;                                                                             \
        memset(&em, 0, sizeof em);                                                                 \
        em.axis[0] = 0.0f;                                                                         \
        em.axis[2] = 2.0f; /* normalised by the oracle */                                          \
        em.beta = 1.0f;                                                                            \
        em.t0 = 5.0f;                                                                              \
        em.n_photons = 6001;                                                                       \
        em.profile = profile;                                                                      \
        pp_sphere tg = {{1.5f, 0.0f, 1.5f}, 0.4f};                                                 \
        const uint64_t N = em.n_photons;                                                           \
        pp_hit* all = (pp_hit*)calloc(N, sizeof(pp_hit));                                          \
        pp_hit* part = (pp_hit*)calloc(N, sizeof(pp_hit));                                         \
        pp_summary s_all, s;                                                                       \
        const int64_t h = or_propagate##S(m, sp, &em, &tg, 11, 3, 10, 0, N, all, N, &s_all, 3);    \
        CHECK(h >= 0 && (uint64_t)h == s_all.n_hits && s_all.n_stored == s_all.n_hits, "all");    \
        CHECK(s_all.n_sim == N && s_all.n_steps >= N && s_all.n_steps <= 10 * N, "steps");        \
        CHECK(s_all.status == PP_OK, "status");                                                    \
        if (profile == PP_EMIT_CHERENKOV_CONE) CHECK(s_all.n_hits > 20, "cone hits the sphere");   \
        double sw = 0.0;                                                                           \
        for (uint64_t i = 0; i < s_all.n_stored; ++i) {                                            \
            const pp_hit* x = &all[i];                                                             \
            if (i) CHECK(x->photon_id > all[i - 1].photon_id, "photon-id order at %llu",           \
                         (unsigned long long)i);                                                   \
            const double dx = x->pos[0] - 1.5, dy = x->pos[1], dz = x->pos[2] - 1.5;               \
            CHECK(fabs(sqrt(dx * dx + dy * dy + dz * dz) - 0.4) < 1e-3, "hit on the sphere");      \
            CHECK(x->weight > 0.0f && x->weight <= 1.0f, "weight");                                \
            CHECK(x->time >= 5.0f && x->dist > 0.0f, "time and path length");                      \
            sw += x->weight;                                                                       \
        }                                                                                          \
        CHECK(fabs(sw - s_all.sum_w) <= 1e-4 * (1.0 + s_all.sum_w), "sum of weights");             \
        /* one thread, capacity 3: same counts, the first three records */                         \
        or_propagate##S(m, sp, &em, &tg, 11, 3, 10, 0, N, part, 3, &s, 1);                         \
        CHECK(s.n_hits == s_all.n_hits && s.n_steps == s_all.n_steps, "threads change counts");   \
        CHECK(s.n_stored == (s_all.n_hits < 3 ? s_all.n_hits : 3), "capacity 3");                  \
        CHECK(s.status == (s_all.n_hits > 3 ? PP_ERR_OVERFLOW : PP_OK), "overflow status");        \
        for (uint64_t i = 0; i < s.n_stored; ++i) CHECK(hits_equal(&part[i], &all[i]), "first hits");\
        /* capacity 0 with a NULL-free buffer of one record */                                     \
        or_propagate##S(m, sp, &em, &tg, 11, 3, 10, 0, N, part, 0, &s, 2);                         \
        CHECK(s.n_stored == 0 && s.n_hits == s_all.n_hits, "capacity 0");                          \
        /* ragged sub-ranges reproduce the whole */                                                \
        const uint64_t cuts[4] = {0, 1, 2777, N};                                                  \
        uint64_t k = 0;                                                                            \
        for (int c = 0; c < 3; ++c) {                                                              \
            or_propagate##S(m, sp, &em, &tg, 11, 3, 10, cuts[c], cuts[c + 1], part, N, &s, 3);     \
            CHECK(s.n_sim == cuts[c + 1] - cuts[c], "sub-range size");                             \
            for (uint64_t i = 0; i < s.n_stored; ++i, ++k)                                         \
                CHECK(k < s_all.n_stored && hits_equal(&part[i], &all[k]), "sub-range record");    \
        }                                                                                          \
        CHECK(k == s_all.n_stored, "sub-ranges cover the hits");                                   \
        /* empty range */                                                                          \
        CHECK(or_propagate##S(m, sp, &em, &tg, 11, 3, 10, 7, 7, part, N, &s, 3) == 0 &&            \
                  s.n_sim == 0 && s.n_steps == 0,                                                  \
              "empty range");                                                                      \
        /* one step only */                                                                        \
        or_propagate##S(m, sp, &em, &tg, 11, 3, 1, 0, N, part, N, &s, 3);                          \
        CHECK(s.n_steps == N, "one step per photon");                                              \
        free(all);                                                                                 \
        free(part);                                                                                \
    }
PROPAGATION_CHECKS(_f32, float)
PROPAGATION_CHECKS(_f64, double)

#define GEOMETRY_CHECKS(S, R)                                                                      \
    static void geometry##S(void) {                                                                \
        const R xs[2] = {1, 3}, ys[2] = {10, 30};                                                  \
        CHECK(or_fast_linear_interp_sorted##S(0, xs, ys, 2) == 10, "clamp below");                 \
        CHECK(or_fast_linear_interp_sorted##S(9, xs, ys, 2) == 30, "clamp above");                 \
        CHECK(or_fast_linear_interp_sorted##S(2, xs, ys, 2) == 20, "midpoint");                    \
        const R kn[2] = {10, 30};                                                                  \
        CHECK(or_fast_linear_interp_knots##S(-1, kn, 2, 0, 1) == 10, "knots clamp below");         \
        CHECK(or_fast_linear_interp_knots##S(2, kn, 2, 0, 1) == 30, "knots clamp above");          \
        const R up[3] = {0, 0, 1}, down[3] = {0, 0, -1}, v[3] = {0.6, 0, 0.8};                     \
        R out[3];                                                                                  \
        or_rotate_to_axis##S(up, v, out);                                                          \
        CHECK(out[0] == v[0] && out[1] == v[1] && out[2] == v[2], "identity at the north pole");   \
        or_rotate_to_axis##S(down, v, out);                                                        \
        CHECK(fabs((double)out[2] + 0.8) < 1e-6, "south pole flips z");                            \
        const R c[3] = {0, 0, 0}, inside[3] = {0.1, 0, 0}, dir[3] = {1, 0, 0};                     \
        R d = -1;                                                                                  \
        CHECK(or_ray_sphere##S(inside, dir, c, 1, 5, &d) == 0, "start inside: no hit");            \
        const R tang[3] = {-2, 1, 0}, behind[3] = {2, 0, 0};                                       \
        CHECK(or_ray_sphere##S(tang, dir, c, 1, 5, &d) == 1 && d == 2, "tangent ray");             \
        CHECK(or_ray_sphere##S(tang, dir, c, 1, 2, &d) == 0, "entry beyond the step");             \
        CHECK(or_ray_sphere##S(behind, dir, c, 1, 5, &d) == 0, "sphere behind");                   \
        const R cos1 = or_hg_costheta##S(0.95, 1), cos0 = or_hg_costheta##S(0.95, 0);              \
        CHECK(fabs((double)cos1 - 1) < 1e-5 && fabs((double)cos0 + 1) < 1e-5, "HG end points");    \
    }
GEOMETRY_CHECKS(_f32, float)
GEOMETRY_CHECKS(_f64, double)

int main(int argc, char** argv) {
    if (argc > 1 && strcmp(argv[1], "canary_overflow") == 0) {
        /* deliberate read one past a heap array: shows AddressSanitizer is live */
        int* v = (int*)calloc(8, sizeof(int));
        volatile int i = 8;
        const int r = v[i];
        free(v);
        return r;
    }
    unsigned long w[10];
    int n_kat = 0;
    while (scanf("%lu %lu %lu %lu %lu %lu %lu %lu %lu %lu", &w[0], &w[1], &w[2], &w[3], &w[4],
                 &w[5], &w[6], &w[7], &w[8], &w[9]) == 10) {
        uint32_t ctr[4] = {(uint32_t)w[0], (uint32_t)w[1], (uint32_t)w[2], (uint32_t)w[3]};
        uint32_t key[2] = {(uint32_t)w[4], (uint32_t)w[5]}, out[4];
        or_philox4x32_10(ctr, key, out);
        for (int j = 0; j < 4; ++j) CHECK(out[j] == (uint32_t)w[6 + j], "Philox KAT %d word %d", n_kat, j);
        ++n_kat;
    }
    CHECK(n_kat >= 3, "known answers read: %d", n_kat);
    const uint32_t points[2] = {2, PP_MAX_TABLE_POINTS};
    const uint32_t kinds[3] = {PP_SPEC_MONO, PP_SPEC_KNOTS, PP_SPEC_INV_LAMBDA2};
    for (int p = 0; p < 2; ++p) {
        pp_medium_tables m = medium(points[p]);
        for (int k = 0; k < 3; ++k) {
            pp_spectrum sp = spectrum(kinds[k]);
            for (uint32_t prof = 0; prof <= PP_EMIT_REF_ISOTROPIC; ++prof) {
                propagation_f32(&m, &sp, prof);
                if (prof == 0) propagation_f64(&m, &sp, prof);
            }
        }
    }
    geometry_f32();
    geometry_f64();
    if (failures) {
        printf("FAIL (%d checks)\n", failures);
        return 1;
    }
    printf("PASS (%d known answers)\n", n_kat);
    return 0;
}
