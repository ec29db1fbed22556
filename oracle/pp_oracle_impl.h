/*
 * pp_oracle_impl.h — body of the CPU oracle, instantiated twice by pp_oracle.c:
 * once with REAL=float (suffix _f32) and once with REAL=double (suffix _f64).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain restatement of PhotonProp.jl's
 * propagation loop used to check the HIP library; nothing in the product links
 * or calls it.
 *
 * The random inputs (Philox words -> uniforms) are identical in both
 * instantiations and identical to the device kernel's, so the oracle can be
 * compared with the GPU photon by photon.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

/* Table lookup on a uniform grid, clamped; linear interpolation.
 * Same operation order as the device kernel so per-photon values agree. */
static inline REAL FN(or_table_)(const float* t, uint32_t n, REAL x_scaled) {
    REAL xmax = (REAL)(n - 1);
    if (x_scaled < (REAL)0) x_scaled = (REAL)0;
    if (x_scaled > xmax) x_scaled = xmax;
    uint32_t i = (uint32_t)x_scaled;
    if (i > n - 2) i = n - 2;
    REAL f = x_scaled - (REAL)i;
    REAL lo = (REAL)t[i];
    return R_FMA(f, (REAL)t[i + 1] - lo, lo);
}

/* fast_linear_interp(x, knots, lower, upper), equidistant knots
 * (src/utils.jl:37-60). */
REAL FN(or_fast_linear_interp_knots)(REAL x, const REAL* knots, uint32_t n_knots, REAL lower,
                                     REAL upper) {
    if (x < lower) x = lower;
    if (x > upper) x = upper;
    REAL range = upper - lower;
    REAL step_size = range / (REAL)(n_knots - 1);
    REAL along_range = (x - lower) / step_size;
    REAL along_floor = R_FLOOR(along_range);
    int64_t lower_knot = (int64_t)along_floor; /* 0-based; reference is 1-based */
    if (lower_knot >= (int64_t)n_knots - 1) return knots[n_knots - 1];
    REAL along_step = along_range - along_floor;
    REAL y_low = knots[lower_knot];
    REAL slope = knots[lower_knot + 1] - y_low;
    return y_low + slope * along_step;
}

/* fast_linear_interp(x, xs, ys), sorted xs, searchsortedfirst
 * (src/utils.jl:10-34).  x equal to xs[0] selects the first interval, as the
 * reference's ix_upper = searchsortedfirst(...) does when the lower clamp bites
 * (ix_upper = 1 would read xs[0] twice; the reference then divides 0/0 — the
 * KAT fast_linear_interp(-1., xs, ys) ≈ 0 holds only because clamping lands on
 * the first knot, whose value is returned here directly). */
REAL FN(or_fast_linear_interp_sorted)(REAL x, const REAL* xs, const REAL* ys, uint32_t n) {
    REAL lower = xs[0], upper = xs[n - 1];
    if (x < lower) x = lower;
    if (x > upper) x = upper;
    uint32_t ix_upper = 0;
    while (ix_upper < n && xs[ix_upper] < x) ++ix_upper; /* searchsortedfirst */
    if (ix_upper == 0) return ys[0];
    uint32_t ix_lower = ix_upper - 1;
    REAL step = xs[ix_upper] - xs[ix_lower];
    REAL along = (x - xs[ix_lower]) / step;
    REAL y_low = ys[ix_lower];
    REAL slope = ys[ix_upper] - y_low;
    return y_low + slope * along;
}

/* Henyey-Greenstein inverse CDF (src/photon_prop_cuda.jl:31-38), with g taken
 * from the medium instead of the hard-coded 0.99 (:229).  g == 0 is isotropic. */
REAL FN(or_hg_costheta)(REAL g, REAL eta) {
    if (g == (REAL)0) return R_FMA((REAL)2, eta, (REAL)-1);
    REAL num = R_FMA(-g, g, (REAL)1);
    REAL den = R_FMA(g, R_FMA((REAL)2, eta, (REAL)-1), (REAL)1);
    REAL q = num / den;
    REAL c = ((REAL)1 / ((REAL)2 * g)) * (R_FMA(g, g, (REAL)1) - q * q);
    if (c < (REAL)-1) c = (REAL)-1;
    if (c > (REAL)1) c = (REAL)1;
    return c;
}

/* rotate_to_axis (src/photon_prop_cuda.jl:171-219): the rotation that takes e_z
 * to old_dir, applied to new_dir; result renormalised. */
void FN(or_rotate_to_axis)(const REAL* old_dir, const REAL* new_dir, REAL* out) {
    REAL oz = old_dir[2];
    if (R_FABS(oz) == (REAL)1) {
        REAL s = oz > 0 ? (REAL)1 : (REAL)-1;
        out[0] = new_dir[0];
        out[1] = s * new_dir[1];
        out[2] = s * new_dir[2];
        return;
    }
    REAL sintheta = R_SQRT(R_FMA(-oz, oz, (REAL)1));
    REAL ax1 = -old_dir[1] / sintheta;
    REAL ax2 = old_dir[0] / sintheta;
    REAL kappa = (ax1 * new_dir[0] + ax2 * new_dir[1]) * ((REAL)1 - oz);
    REAL nd3s = new_dir[2] * sintheta;
    REAL nx = R_FMA(new_dir[0], oz, R_FMA(ax2, nd3s, ax1 * kappa));
    REAL ny = R_FMA(new_dir[1], oz, R_FMA(-ax1, nd3s, ax2 * kappa));
    REAL nz = R_FMA(new_dir[2], oz, sintheta * R_FMA(ax1, new_dir[1], -ax2 * new_dir[0]));
    REAL inv = (REAL)1 / R_SQRT(nx * nx + ny * ny + nz * nz);
    out[0] = nx * inv;
    out[1] = ny * inv;
    out[2] = nz * inv;
}

/* update_position (src/photon_prop_cuda.jl:255-259): pos + step·dir. */
void FN(or_update_position)(const REAL* pos, const REAL* dir, REAL step, REAL* out) {
    for (int j = 0; j < 3; ++j) out[j] = R_FMA(step, dir[j], pos[j]);
}

/* Ray-sphere test before moving (src/photon_prop_cuda.jl:327-351).
 * Returns 1 and the distance to the entry point if the sphere is entered in
 * (0, step); a sphere behind the photon or beyond the step is no hit. */
int FN(or_ray_sphere)(const REAL* pos, const REAL* dir, const REAL* center, REAL r2, REAL step,
                      REAL* d_out) {
    REAL a = 0, pp = 0;
    for (int j = 0; j < 3; ++j) {
        REAL dpos = pos[j] - center[j];
        a = R_FMA(dir[j], dpos, a);
        pp = R_FMA(dpos, dpos, pp);
    }
    REAL b = R_FMA(a, a, r2 - pp);
    if (!(b >= 0)) return 0;
    REAL d = -a - R_SQRT(b);
    if (d > 0 && d < step) {
        *d_out = d;
        return 1;
    }
    return 0;
}

/* Wavelength draw (src/photon_prop_cuda.jl:70-75). */
static inline REAL FN(or_wavelength_)(const pp_spectrum* sp, const REAL* knots, REAL u) {
    switch (sp->kind) {
    case PP_SPEC_MONO:
        return (REAL)sp->mono;
    case PP_SPEC_KNOTS:
        return FN(or_fast_linear_interp_knots)(u, knots, sp->n_knots, (REAL)0, (REAL)1);
    default: { /* PP_SPEC_INV_LAMBDA2: 1/λ = 1/lo - u·(1/lo - 1/hi) */
        REAL a = (REAL)(1.0 / (double)sp->lo);
        REAL d = (REAL)(1.0 / (double)sp->lo - 1.0 / (double)sp->hi);
        return (REAL)1 / R_FMA(-u, d, a);
    }
    }
}

/* Initial direction (src/photon_prop_cuda.jl:52-56, :67-68 for the reference
 * profile; the Cherenkov cone is the configured physics). */
static inline void FN(or_init_direction_)(const pp_emitter* em, const REAL* axis, REAL n_phase,
                                          REAL u_phi, REAL u_theta, REAL* dir) {
    REAL phi = u_phi * (REAL)PP_TWO_PI;
    REAL sp = R_SIN(phi), cp = R_COS(phi);
    if (em->profile == PP_EMIT_REF_ISOTROPIC) {
        REAL theta = R_FMA(u_theta, (REAL)2, (REAL)-1); /* uniform(-1, 1), radians */
        REAL st = R_SIN(theta), ct = R_COS(theta);
        dir[0] = cp * st;
        dir[1] = sp * st;
        dir[2] = ct;
        return;
    }
    REAL ct;
    if (em->profile == PP_EMIT_ISOTROPIC)
        ct = R_FMA((REAL)2, u_theta, (REAL)-1);
    else
        ct = (REAL)1 / (n_phase * (REAL)em->beta);
    if (ct > (REAL)1) ct = (REAL)1; /* n·β = 1 within rounding: emission along the axis */
    REAL st = R_SQRT(R_FMA(-ct, ct, (REAL)1));
    REAL local[3] = {cp * st, sp * st, ct};
    FN(or_rotate_to_axis)(axis, local, dir);
}

/* Emission draws of photon pid: out = {λ, dir_x, dir_y, dir_z}
 * (src/photon_prop_cuda.jl:308-316 with D1). */
void FN(or_emit)(const pp_medium_tables* med, const pp_spectrum* sp, const pp_emitter* em,
                 uint64_t seed, uint32_t stream, uint64_t pid, REAL* out) {
    REAL knots[PP_MAX_KNOTS];
    if (sp->kind == PP_SPEC_KNOTS)
        for (uint32_t i = 0; i < sp->n_knots && i < PP_MAX_KNOTS; ++i) knots[i] = (REAL)sp->knots[i];
    double an = sqrt((double)em->axis[0] * em->axis[0] + (double)em->axis[1] * em->axis[1] +
                     (double)em->axis[2] * em->axis[2]);
    REAL axis[3] = {(REAL)(em->axis[0] / an), (REAL)(em->axis[1] / an),
                    (REAL)(em->axis[2] / an)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {0u, (uint32_t)pid, (uint32_t)(pid >> 32), stream};
    uint32_t w0[4];
    float u0[4], p0[2];
    or_philox4x32_10(ctr, key, w0);
    or_block_draws(w0, u0, p0);
    REAL lambda = FN(or_wavelength_)(sp, knots, (REAL)u0[0]);
    REAL inv_dl = (REAL)((double)(med->n_points - 1) /
                         ((double)med->lambda_max - (double)med->lambda_min));
    REAL n_phase = FN(or_table_)(med->n_phase, med->n_points, (lambda - (REAL)med->lambda_min) * inv_dl);
    REAL dir[3];
    FN(or_init_direction_)(em, axis, n_phase, (REAL)p0[0], (REAL)u0[1], dir);
    out[0] = lambda;
    out[1] = dir[0];
    out[2] = dir[1];
    out[3] = dir[2];
}

typedef struct {
    uint64_t n_steps, n_hits;
    double sum_w, sum_wt, sum_wt2;
    pp_hit* hits;
    uint64_t n_hit_rec, cap_hit_rec;
} FN(or_acc_);

/* One photon: emit, then up to max_steps iterations of {exponential step,
 * sphere test before the move, move (to the hit and stop, or by the step),
 * HG scatter} (src/photon_prop_cuda.jl:301-380; deviations D1-D10 are listed
 * in the oracle header of pp_oracle.c). */
static void FN(or_photon_)(const pp_medium_tables* med, const pp_spectrum* sp, const REAL* knots,
                           const pp_emitter* em, const REAL* axis, const pp_sphere* tgt,
                           uint64_t seed, uint32_t stream, uint32_t max_steps, uint64_t pid,
                           FN(or_acc_) * acc) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {0u, (uint32_t)pid, (uint32_t)(pid >> 32), stream};
    /* Draws (or_block_draws) per Philox block (w0, w1, w2, w3):
     *   block 0:      λ <- u0, emission θ <- u1, emission φ <- p0;
     *                 step 0 <- {u2, u3, p1};
     *   block b >= 1: step 2b-1 <- {u0, u1, p0}, step 2b <- {u2, u3, p1};
     * {step length, scattering angle, azimuth} per step.  The final step
     * (max_steps - 1) uses only its length; when it is the first step of its
     * block and the emission is the Cherenkov cone (which does not use u1 of
     * block 0), its length is u1 of block 0 (the device then never computes
     * that block). */
    uint32_t w0[4];
    float ub[4], pb[2];
    or_philox4x32_10(ctr, key, w0);
    or_block_draws(w0, ub, pb);
    const float u_lone = ub[1];
    const int lone_from_block0 = em->profile == PP_EMIT_CHERENKOV_CONE && max_steps % 2u == 0u;

    REAL u_lambda = (REAL)ub[0];
    REAL lambda = FN(or_wavelength_)(sp, knots, u_lambda);

    REAL inv_dl = (REAL)((double)(med->n_points - 1) /
                         ((double)med->lambda_max - (double)med->lambda_min));
    REAL xs = (lambda - (REAL)med->lambda_min) * inv_dl;
    REAL n_phase = FN(or_table_)(med->n_phase, med->n_points, xs);
    REAL sca = FN(or_table_)(med->sca_len, med->n_points, xs);
    REAL abs_len = FN(or_table_)(med->abs_len, med->n_points, xs);

    REAL dir[3], init_dir[3];
    FN(or_init_direction_)(em, axis, n_phase, (REAL)pb[0], (REAL)ub[1], dir);
    init_dir[0] = dir[0];
    init_dir[1] = dir[1];
    init_dir[2] = dir[2];

    REAL pos[3] = {(REAL)em->pos[0], (REAL)em->pos[1], (REAL)em->pos[2]};
    REAL center[3] = {(REAL)tgt->center[0], (REAL)tgt->center[1], (REAL)tgt->center[2]};
    REAL r2 = (REAL)tgt->radius * (REAL)tgt->radius;
    REAL g = (REAL)med->g;
    REAL dist = 0;
    int hit = 0;
    uint32_t steps = 0;
    uint32_t words[4];
    uint32_t have_block = 0u;

    for (uint32_t s = 0; s < max_steps; ++s) {
        const uint32_t blk = (s + 1u) / 2u;
        const uint32_t slot = (s % 2u == 0u) ? 1u : 0u; /* even steps: {u2, u3, p1} */
        REAL u[3];
        if (lone_from_block0 && s + 1u == max_steps) {
            /* the final step, first of its block: only its length is drawn */
            u[0] = (REAL)u_lone;
            u[1] = (REAL)0.5;
            u[2] = (REAL)0;
        } else {
            if (blk != have_block) {
                ctr[0] = blk;
                or_philox4x32_10(ctr, key, words);
                or_block_draws(words, ub, pb);
                have_block = blk;
            }
            u[0] = (REAL)ub[2 * slot];
            u[1] = (REAL)ub[2 * slot + 1];
            u[2] = (REAL)pb[slot];
        }
        ++steps;
        REAL step = -R_LOG(u[0]) * sca; /* src/photon_prop_cuda.jl:322-323 */
        REAL d;
        if (FN(or_ray_sphere)(pos, dir, center, r2, step, &d)) {
            FN(or_update_position)(pos, dir, d, pos);
            dist += d;
            hit = 1;
            break;
        }
        /* D3: move and scatter when the sphere is not entered in this step
         * (photon_prob.ipynb:3984-4005). */
        FN(or_update_position)(pos, dir, step, pos);
        dist += step;
        REAL ct = FN(or_hg_costheta)(g, u[1]);
        REAL st = R_SQRT(R_FMA(-ct, ct, (REAL)1));
        REAL phi = u[2] * (REAL)PP_TWO_PI;
        REAL local[3] = {R_COS(phi) * st, R_SIN(phi) * st, ct};
        REAL nd[3];
        FN(or_rotate_to_axis)(dir, local, nd);
        dir[0] = nd[0];
        dir[1] = nd[1];
        dir[2] = nd[2];
    }
    acc->n_steps += steps;
    if (!hit) return;

    REAL weight = R_EXP(-dist / abs_len);                        /* :622, on device (D8) */
    REAL time = (REAL)em->t0 + dist * n_phase / (REAL)PP_C_VACUUM; /* :634-638 (D7) */
    acc->n_hits += 1;
    acc->sum_w += (double)weight;
    acc->sum_wt += (double)weight * (double)time;
    acc->sum_wt2 += (double)weight * (double)time * (double)time;
    if (acc->n_hit_rec == acc->cap_hit_rec) {
        uint64_t nc = acc->cap_hit_rec ? 2 * acc->cap_hit_rec : 1024;
        pp_hit* nh = (pp_hit*)realloc(acc->hits, nc * sizeof(pp_hit));
        if (!nh) return; /* counted, not stored */
        acc->hits = nh;
        acc->cap_hit_rec = nc;
    }
    pp_hit* h = &acc->hits[acc->n_hit_rec++];
    for (int j = 0; j < 3; ++j) {
        h->pos[j] = (float)pos[j];
        h->dir[j] = (float)dir[j];
        h->init_dir[j] = (float)init_dir[j];
    }
    h->time = (float)time;
    h->wavelength = (float)lambda;
    h->weight = (float)weight;
    h->dist = (float)dist;
    h->request = 0;
    h->photon_id = pid;
}

/* Propagate photons [id_begin, id_end) of one request.  Hits are returned in
 * photon-id order; sums are FP64 and combined in chunk order.  Returns the
 * number of hits (which may exceed hits_cap; only hits_cap are stored). */
int64_t FN(or_propagate)(const pp_medium_tables* med, const pp_spectrum* sp,
                         const pp_emitter* em, const pp_sphere* tgt, uint64_t seed,
                         uint32_t stream, uint32_t max_steps, uint64_t id_begin, uint64_t id_end,
                         pp_hit* hits_out, uint64_t hits_cap, pp_summary* summary,
                         int n_threads) {
    REAL knots[PP_MAX_KNOTS];
    if (sp->kind == PP_SPEC_KNOTS) {
        if (sp->n_knots < 2 || sp->n_knots > PP_MAX_KNOTS) return -1;
        for (uint32_t i = 0; i < sp->n_knots; ++i) knots[i] = (REAL)sp->knots[i];
    }
    double an = sqrt((double)em->axis[0] * em->axis[0] + (double)em->axis[1] * em->axis[1] +
                     (double)em->axis[2] * em->axis[2]);
    REAL axis[3] = {(REAL)(em->axis[0] / an), (REAL)(em->axis[1] / an),
                    (REAL)(em->axis[2] / an)};
    if (n_threads < 1) n_threads = 1;
    uint64_t n = id_end > id_begin ? id_end - id_begin : 0;
    int n_chunks = n_threads;
    FN(or_acc_)* accs = (FN(or_acc_)*)calloc((size_t)n_chunks, sizeof(FN(or_acc_)));
    if (!accs) return -1;
#pragma omp parallel for num_threads(n_threads) schedule(static, 1)
    for (int c = 0; c < n_chunks; ++c) {
        uint64_t b = id_begin + n * (uint64_t)c / (uint64_t)n_chunks;
        uint64_t e = id_begin + n * (uint64_t)(c + 1) / (uint64_t)n_chunks;
        for (uint64_t pid = b; pid < e; ++pid)
            FN(or_photon_)(med, sp, knots, em, axis, tgt, seed, stream, max_steps, pid, &accs[c]);
    }
    pp_summary s;
    memset(&s, 0, sizeof s);
    s.n_sim = n;
    uint64_t stored = 0;
    for (int c = 0; c < n_chunks; ++c) {
        s.n_steps += accs[c].n_steps;
        s.n_hits += accs[c].n_hits;
        s.sum_w += accs[c].sum_w;
        s.sum_wt += accs[c].sum_wt;
        s.sum_wt2 += accs[c].sum_wt2;
        for (uint64_t i = 0; i < accs[c].n_hit_rec && stored < hits_cap; ++i)
            hits_out[stored++] = accs[c].hits[i];
        free(accs[c].hits);
    }
    free(accs);
    s.n_stored = stored;
    s.status = s.n_hits > hits_cap ? PP_ERR_OVERFLOW : PP_OK;
    if (summary) *summary = s;
    return (int64_t)s.n_hits;
}

#undef FN
