/*
 * photonprop.h — C ABI of the MI355X (gfx950) photon-propagation library.
 *
 * The library replaces the device side of PhotonProp.jl's propagation path:
 *   - the kernel cuda_propagate_photons!      (src/photon_prop_cuda.jl:262-386)
 *   - its host driver propagate_distance      (src/photon_prop_cuda.jl:588-644)
 *   - the buffer helpers initialize_photon_arrays / process_output
 *                                              (src/photon_prop_cuda.jl:390-422)
 * Everything crossing this boundary is plain C: fixed-layout structs, raw
 * pointers and sizes.  No C++ or torch types.
 *
 * Units: metres, nanometres, nanoseconds.
 *
 * Conventions
 *   - Every function returns PP_OK (0) or a negative pp_status code.  The
 *     message of the last failure on the calling thread is pp_last_error().
 *   - A pp_ctx owns all device memory it uses (medium tables, hit staging,
 *     per-tile partials).  The caller owns every host buffer it passes in.
 *   - Calls on one pp_ctx are synchronous unless the name ends in _device, and
 *     a pp_ctx must not be used from two host threads at once.  Use one
 *     pp_ctx per host thread; each has its own HIP stream.
 *   - Photon i of a request draws its random numbers from the Philox4x32-10
 *     counter stream keyed by (seed) with counter (block, i_lo, i_hi, stream),
 *     so a photon's history depends only on (seed, stream, i, inputs) and not
 *     on launch geometry, batch composition or device.
 */
#ifndef PHOTONPROP_H
#define PHOTONPROP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_ABI_VERSION 1

typedef enum pp_status {
    PP_OK = 0,
    PP_ERR_INVALID_ARG = -1, /* bad argument; see pp_last_error()                   */
    PP_ERR_OVERFLOW = -2,    /* hit buffer too small: counts and sums are exact,
                                stored records are truncated to the capacity       */
    PP_ERR_DEVICE = -3,      /* HIP runtime error                                   */
    PP_ERR_OOM = -4,         /* device allocation failed                            */
    PP_ERR_UNSUPPORTED = -5, /* unknown spectrum kind or emission profile; mirrors
                                the reference's ArgumentError("Cannot initialize")
                                (src/photon_prop_cuda.jl:67, :70)                   */
    PP_ERR_BUSY = -6,        /* serving queue full / capacity exceeded (back-pressure) */
    PP_ERR_NOT_FOUND = -7    /* unknown request ticket                              */
} pp_status;

/* Spectrum kinds (reference: src/spectrum.jl:19-103). */
typedef enum pp_spectrum_kind {
    PP_SPEC_MONO = 0,        /* Monochromatic (src/spectrum.jl:21-23)                 */
    PP_SPEC_KNOTS = 1,       /* CherenkovSpectrum: inverse-CDF knots equidistant in u
                                on [0,1] (src/spectrum.jl:90-103, utils.jl:37-60)   */
    PP_SPEC_INV_LAMBDA2 = 2  /* dN/dλ ∝ 1/λ² on [lo,hi], closed-form inverse CDF    */
} pp_spectrum_kind;

/* Emission profiles (reference: AngularEmissionProfile, src/emission.jl:13). */
typedef enum pp_emission_profile {
    PP_EMIT_CHERENKOV_CONE = 0, /* cosθ = 1/(n(λ)·β) about the emitter axis, φ uniform */
    PP_EMIT_ISOTROPIC = 1,      /* uniform on the sphere (cosθ uniform on [-1,1])       */
    PP_EMIT_REF_ISOTROPIC = 2   /* the reference's :IsotropicEmission as it behaves:
                                   θ ~ U(-1,1) radians, φ ~ U(0,2π), axis ignored
                                   (src/photon_prop_cuda.jl:52-56)                     */
} pp_emission_profile;

/* Medium as tables on a uniform wavelength grid of n_points points spanning
 * [lambda_min, lambda_max] nm.  Replaces WaterProperties + get_refractive_index /
 * get_scattering_length / get_absorption_length (src/medium.jl:23-251), which
 * the reference bakes into the kernel as a compile-time Val(medium)
 * (src/photon_prop_cuda.jl:274, :609).  Values are linearly interpolated and
 * clamped at the grid ends. */
typedef struct pp_medium_tables {
    float lambda_min;       /* nm */
    float lambda_max;       /* nm, > lambda_min */
    uint32_t n_points;      /* 2 .. PP_MAX_TABLE_POINTS */
    uint32_t reserved;
    const float* n_phase;   /* [n_points] phase refractive index            */
    const float* abs_len;   /* [n_points] absorption length, m (> 0)        */
    const float* sca_len;   /* [n_points] scattering length, m (> 0)        */
    float g;                /* Henyey-Greenstein asymmetry, |g| < 1          */
    float reserved2;
} pp_medium_tables;

#define PP_MAX_TABLE_POINTS 2048
#define PP_MAX_KNOTS 1024

/* Wavelength spectrum (replaces PhotonSource.spectrum, src/emission.jl:107). */
typedef struct pp_spectrum {
    uint32_t kind;          /* pp_spectrum_kind */
    uint32_t n_knots;       /* PP_SPEC_KNOTS: 2 .. PP_MAX_KNOTS */
    float lo, hi;           /* PP_SPEC_INV_LAMBDA2: range in nm */
    float mono;             /* PP_SPEC_MONO: wavelength in nm */
    float reserved;
    const float* knots;     /* PP_SPEC_KNOTS: λ at u = k/(n_knots-1) */
} pp_spectrum;

/* Point emitter (replaces PhotonSource, src/emission.jl:102-109). */
typedef struct pp_emitter {
    float pos[3];           /* m  */
    float t0;               /* ns */
    float axis[3];          /* unit vector (normalised by the library) */
    float beta;             /* particle speed / c for the Cherenkov cone */
    uint64_t n_photons;     /* photons to simulate, exactly */
    uint32_t profile;       /* pp_emission_profile */
    uint32_t reserved;
} pp_emitter;

/* Spherical target (replaces DetectionSphere, src/detection.jl:12-15). */
typedef struct pp_sphere {
    float center[3];        /* m */
    float radius;           /* m, > 0 */
} pp_sphere;

/* One detected photon, 64 bytes.  Replaces the reference's four SoA hit stacks
 * (position, initial direction, wavelength, distance; src/photon_prop_cuda.jl:369-372)
 * and adds what the reference computes on the host afterwards (weight, time). */
typedef struct pp_hit {
    float pos[3];           /* hit point on the sphere, m */
    float time;             /* t0 + L·n(λ)/c, ns (phase index, src/photon_prop_cuda.jl:634-638) */
    float dir[3];           /* direction at arrival */
    float wavelength;       /* nm */
    float init_dir[3];      /* direction at emission */
    float weight;           /* absorption weight exp(-L/λ_abs(λ)) */
    float dist;             /* path length L, m */
    uint32_t request;       /* request index within the launch (0 for pp_propagate) */
    uint64_t photon_id;     /* photon index within its request */
} pp_hit;

/* Per-request result summary.  Sums accumulate in FP64 in a fixed order that
 * depends only on the request itself. */
typedef struct pp_summary {
    uint64_t n_sim;         /* photons simulated (== n_photons) */
    uint64_t n_steps;       /* scattering-loop iterations executed */
    uint64_t n_hits;        /* photons that reached the target (exact) */
    uint64_t n_stored;      /* hit records written to the output */
    double sum_w;           /* Σ weight */
    double sum_wt;          /* Σ weight·time */
    double sum_wt2;         /* Σ weight·time² */
    int32_t status;         /* PP_OK or PP_ERR_OVERFLOW */
    uint32_t reserved;
} pp_summary;

/* One propagation request: emitter + target + random stream. */
typedef struct pp_request {
    pp_emitter emitter;
    pp_sphere target;
    uint64_t seed;          /* Philox key */
    uint32_t stream;        /* Philox counter word 3 */
    uint32_t max_steps;     /* scattering-loop cap, 1 .. PP_MAX_STEPS; reference uses 10
                               (src/photon_prop_cuda.jl:319) */
} pp_request;

/* Largest max_steps any call accepts (2^24 scattering-loop iterations per photon). */
#define PP_MAX_STEPS (1u << 24)

typedef struct pp_ctx pp_ctx;

/* ---- library / context ---------------------------------------------------- */

/* Thread-local message of the last failed call. */
const char* pp_last_error(void);

/* PP_ABI_VERSION of the loaded library. */
int pp_abi_version(void);

/* Create a context on HIP device `device` and upload the medium tables once.
 * Replaces the reference's compile-time Val(medium) and its module-load warm-up
 * _init_workaround (src/photon_prop_cuda.jl:646-654). */
int pp_create(pp_ctx** out, int device, const pp_medium_tables* medium);

/* Replace the resident medium tables. */
int pp_set_medium(pp_ctx* ctx, const pp_medium_tables* medium);

int pp_destroy(pp_ctx* ctx);

/* Name and compute-unit count of the context's device (name_cap bytes incl. NUL). */
int pp_device_info(pp_ctx* ctx, char* name, size_t name_cap, int* n_cu);

/* ---- plain propagation ---------------------------------------------------- */

/* Propagate emitter->n_photons photons towards one sphere and return their hits.
 * Replaces @cuda cuda_propagate_photons! + process_output + the host weight
 * pass of propagate_distance (src/photon_prop_cuda.jl:262-386, :588-644).
 * hits_out: host buffer of hits_cap records (may be NULL when hits_cap == 0).
 * Returns PP_ERR_OVERFLOW (summary still exact) if more photons hit than fit. */
int pp_propagate(pp_ctx* ctx, const pp_emitter* emitter, const pp_spectrum* spectrum,
                 const pp_sphere* target, uint64_t seed, uint32_t stream, uint32_t max_steps,
                 pp_hit* hits_out, uint64_t hits_cap, pp_summary* summary_out);

/* ---- batched propagation (one persistent launch for many requests) --------- */

/* Propagate n_req requests of uneven size in one launch.  The reference's only
 * batching is the unfinished split_source / propagate_sources
 * (src/photon_prop_cuda.jl:474-531).  Hits come back grouped by request:
 * request r owns hits_out[offsets_out[r] .. offsets_out[r+1]).
 * hits_cap bounds all hits of the launch (exceeding it is PP_ERR_OVERFLOW);
 * per_request_cap (0: none) keeps at most that many records of each request,
 * a subset independent of the photons' histories.  Within a request, records
 * are in the order the kernel's waves flushed them, which varies from run to
 * run; so does the subset a cap keeps (the summaries do not vary).
 * offsets_out: n_req+1 entries; summaries_out: n_req entries. */
int pp_propagate_batch(pp_ctx* ctx, const pp_request* requests, uint32_t n_req,
                       const pp_spectrum* spectrum, pp_hit* hits_out, uint64_t hits_cap,
                       uint64_t per_request_cap, uint64_t* offsets_out,
                       pp_summary* summaries_out);

/* Device-resident form of pp_propagate_batch: enqueue on the context's stream and
 * return.  d_hits, d_offsets (n_req+1) and d_summaries (n_req) are device
 * pointers; d_hits must hold min(hits_cap, n_req·per_request_cap) records (or
 * hits_cap when per_request_cap is 0).  Results are grouped by request as above
 * and are valid after pp_synchronize(ctx).  Requests are copied before the call
 * returns. */
int pp_propagate_batch_device(pp_ctx* ctx, const pp_request* requests, uint32_t n_req,
                              const pp_spectrum* spectrum, pp_hit* d_hits, uint64_t hits_cap,
                              uint64_t per_request_cap, uint64_t* d_offsets,
                              pp_summary* d_summaries);

/* Allocation-free form of pp_propagate_batch_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and bitwise the same summaries and per-request hit
 * sets; the call itself allocates nothing.  Its work buffers (requests, tile
 * partials, counters, hit staging, compaction arrays) live in one fixed
 * device array of PP_NOALLOC_SCRATCH_BYTES per GPU, set up by pp_create and
 * held until the process ends; the summaries go straight to d_summaries.
 * Limits, refused with PP_ERR_INVALID_ARG: n_req <= PP_NOALLOC_MAX_REQUESTS,
 * at most PP_NOALLOC_MAX_TILES tiles of PP_TILE_PHOTONS photons in all,
 * hits_cap <= PP_NOALLOC_MAX_HITS.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several contexts or threads) race;
 * serialise them. */
#define PP_NOALLOC_MAX_REQUESTS 4096u
#define PP_NOALLOC_MAX_TILES (1u << 20)     /* 8.6e9 photons per call */
#define PP_NOALLOC_MAX_HITS (1u << 18)      /* staged hit records per call */
#define PP_NOALLOC_SCRATCH_BYTES 60166656ull /* 57.4 MiB per GPU */
int pp_propagate_batch_device_noalloc(pp_ctx* ctx, const pp_request* requests, uint32_t n_req,
                                      const pp_spectrum* spectrum, pp_hit* d_hits,
                                      uint64_t hits_cap, uint64_t per_request_cap,
                                      uint64_t* d_offsets, pp_summary* d_summaries);

/* Wait for all work enqueued on the context's stream. */
int pp_synchronize(pp_ctx* ctx);

/* Duration in milliseconds of the last propagation kernel launched on this
 * context, measured with HIP events on the context's stream. */
int pp_last_kernel_ms(pp_ctx* ctx, float* ms);

/* ---- serving loop: continuous batching into a persistent kernel ------------ */
/* No counterpart in the reference (which launches one compile-time-specialised
 * kernel per call, src/photon_prop_cuda.jl:607).  Requests are queued on the
 * host, admitted when a device slot and room in the device hit pool (the HBM
 * budget) are free, and their photon tiles are published to a persistent
 * kernel that keeps running while work arrives.  Each request's summary and
 * hits are bitwise the same as pp_propagate gives for it alone.  All server
 * calls are thread-safe. */

#define PP_TILE_PHOTONS 8192

/* Largest hit-record capacity any call accepts (hits_cap, hit_pool_records):
 * 2^36 records of 64 B, beyond any device; larger values are rejected with
 * PP_ERR_INVALID_ARG rather than wrapped in a size computation. */
#define PP_MAX_HIT_RECORDS (1ull << 36)

/* pp_server_config.flags */
#define PP_SERVER_HOST_POOL 1u  /* hit pool in pinned host memory: the kernel writes records
                                   straight to the host and results are copied out by the
                                   host alone.  Used by itself when the device has no SDMA
                                   engine to copy results from device memory. */
#define PP_SERVER_LEAVE_ROOM 2u /* the persistent grid takes 7 of the 8 wave slots and leaves
                                   32 VGPRs on every SIMD, so the caller's other kernels run
                                   beside a busy server (about 3 % lower serving throughput);
                                   by default it takes the whole device */

typedef struct pp_server_config {
    uint32_t max_slots;             /* requests on the device at once, 1..4096 */
    uint32_t max_tiles_per_request; /* photons per request <= this * PP_TILE_PHOTONS */
    uint64_t hit_pool_records;      /* hit-record pool shared by admitted requests */
    uint32_t idle_exit_us;          /* the host stops the persistent kernel after this long
                                       without work (relaunched on demand); 0 -> 20000 */
    uint32_t device_wait_us;        /* the kernel stops by itself after waiting this long
                                       (device clock) with nothing published, should the
                                       host not stop it; 0 -> max(4 * idle_exit_us, 1 s) */
    uint32_t flags;                 /* PP_SERVER_* */
    uint32_t reserved;
} pp_server_config;

typedef struct pp_server_stats {
    uint64_t submitted, completed, in_flight, pending;
    uint64_t kernel_launches;       /* persistent-kernel launches so far */
    uint64_t pool_records_free;
    uint64_t tiles_published;
    uint64_t grid_blocks;           /* persistent-kernel blocks (one per resident slot) */
    uint64_t device_stops;          /* kernel exits after device_wait_us without work */
    uint64_t kernel_running;        /* 1 while the persistent kernel runs */
    uint64_t host_pool;             /* 1 if the hit pool is in host memory */
} pp_server_stats;

typedef struct pp_server pp_server;

/* Create a server on ctx's device for requests drawn from `spectrum`. */
int pp_server_create(pp_ctx* ctx, const pp_spectrum* spectrum, const pp_server_config* cfg,
                     pp_server** out);

/* Queue a request; never blocks.  hits_cap records of the device pool are
 * reserved for it while it is admitted (0: summary only).  Fails with
 * PP_ERR_INVALID_ARG if the request can never fit (too many photons or a
 * hits_cap above the pool). */
int pp_server_submit(pp_server* srv, const pp_request* request, uint64_t hits_cap,
                     uint64_t* ticket_out);

/* Wait up to timeout_ms (0: poll) for a ticket; *done_out = 1 when finished. */
int pp_server_wait(pp_server* srv, uint64_t ticket, uint32_t timeout_ms, int* done_out);

/* Results of a finished ticket: summary, hits (up to hits_cap records, the
 * count in *n_hits_out) and the submit-to-completion latency.  Frees the
 * ticket and its pool records. */
int pp_server_result(pp_server* srv, uint64_t ticket, pp_summary* summary, pp_hit* hits_out,
                     uint64_t hits_cap, uint64_t* n_hits_out, double* latency_us);

int pp_server_stats_get(pp_server* srv, pp_server_stats* out);

/* Drain: stop admitting, let the kernel finish in-flight work and exit. */
int pp_server_destroy(pp_server* srv);

/* ---- MLP surrogate (BF16 MFMA training and inference) ---------------------- */
/* Role of PhotonProp.jl's Modelling module (src/modelling.jl:52-111, :202-267):
 * an MLP n_in -> width^depth -> n_out with SiLU or ReLU hidden activations,
 * BF16 weights and activations, FP32 accumulation, FP32 master weights and Adam.
 * The training loss is the Poisson NLL of the observed (weighted) yield k with
 * expectation N·exp(y0) plus time_weight times the NLL of residual arrival
 * times (ns) under a 3-component gamma mixture parametrised by y[1..3]
 * (mixture logits), y[4..6] (log shapes), y[7..9] (log rates per 10 ns).
 * Parameters are packed layer by layer: Wt_l [n_out_l][n_in_l] then b_l. */

typedef struct pp_mlp pp_mlp;

int pp_mlp_create(pp_ctx* ctx, uint32_t n_in, uint32_t width, uint32_t depth, uint32_t n_out,
                  uint64_t seed, pp_mlp** out);
int pp_mlp_destroy(pp_mlp* mlp);
int pp_mlp_n_params(pp_mlp* mlp, uint64_t* n);
int pp_mlp_get_params(pp_mlp* mlp, float* params, uint64_t n);
int pp_mlp_set_params(pp_mlp* mlp, const float* params, uint64_t n);

/* Hidden-layer activation.  The reference's network is ReLU throughout
 * (src/modelling.jl:212-220, Dense(..., relu)); SiLU is smooth in the inputs. */
#define PP_MLP_SILU 0u
#define PP_MLP_RELU 1u
/* Select the hidden activation (PP_MLP_SILU after pp_mlp_create).  It applies
 * to every later forward, loss/gradient and training call; the parameters and
 * the Adam state are kept. */
int pp_mlp_set_activation(pp_mlp* mlp, uint32_t activation);

/* y[n_rows][n_out] = MLP(x[n_rows][n_in]); device pointers, FP32 in and out. */
int pp_mlp_forward_device(pp_mlp* mlp, const float* d_x, uint64_t n_rows, float* d_y);
/* Same with host buffers. */
int pp_mlp_forward(pp_mlp* mlp, const float* x, uint64_t n_rows, float* y);
/* Duration (ms) of the last pp_mlp_forward_device kernels (HIP events). */
int pp_mlp_last_forward_ms(pp_mlp* mlp, float* ms);

/* Loss (mean over rows: {Poisson, time}) and its gradient w.r.t. the packed
 * parameters (grads_host may be NULL).  Device inputs: x [n_rows][n_in],
 * log_n [n_rows] (log photons), k [n_rows] (observed yield), times
 * [n_rows][n_times] (residual times in ns; negative = absent). */
int pp_mlp_loss_grad_device(pp_mlp* mlp, const float* d_x, const float* d_log_n,
                            const float* d_k, const float* d_times, uint32_t n_times,
                            uint32_t n_rows, float time_weight, float* grads_host,
                            double* loss_out);
/* One Adam step (beta 0.9/0.999) on that loss. */
int pp_mlp_train_step_device(pp_mlp* mlp, const float* d_x, const float* d_log_n,
                             const float* d_k, const float* d_times, uint32_t n_times,
                             uint32_t n_rows, float lr, float time_weight, double* loss_out);
/* n_steps Adam steps over consecutive minibatches, the inner loop of the
 * reference's train_mlp (src/modelling.jl:253-261: gradient, then update, per
 * minibatch): step i uses rows [i·batch, (i+1)·batch) of the device arrays
 * (laid out as for pp_mlp_loss_grad_device) and learning rate lr[i] (host
 * array of n_steps).  The steps are enqueued back to back with no host
 * synchronisation between them; the result equals n_steps calls of
 * pp_mlp_train_step_device bit for bit.  loss_out[2] (may be NULL) receives
 * the mean over the steps of each step's mean loss. */
int pp_mlp_train_steps_device(pp_mlp* mlp, const float* d_x, const float* d_log_n,
                              const float* d_k, const float* d_times, uint32_t n_times,
                              uint32_t batch, uint32_t n_steps, const float* lr,
                              float time_weight, double* loss_out);

/* ---- device self-tests of the kernel's own helper functions ---------------- */
/* The reference checks its device helpers through tiny test kernels
 * (apply_rot, src/tests/test_photon_prob.jl:32-41).  These run the helpers the
 * propagation kernel inlines, on the device, over host arrays. */

/* out[3i..3i+2] = rotate_to_axis(old_dirs[3i..], new_dirs[3i..]) (src/photon_prop_cuda.jl:171-219). */
int pp_selftest_rotate(pp_ctx* ctx, const float* old_dirs, const float* new_dirs, float* out,
                       uint64_t n);

/* out[i] = Henyey-Greenstein cosθ for g from Philox word i of (seed, stream 0),
 * drawn as the propagation draws it: a = 1 + (w >> 9)·2^-23, η = 2 - a
 * (src/photon_prop_cuda.jl:31-38). */
int pp_selftest_hg(pp_ctx* ctx, float g, uint64_t seed, float* out, uint64_t n);

/* words_out[4i..4i+3] = Philox4x32-10(counter = ctr[4i..4i+3], key = (seed_lo, seed_hi)). */
int pp_selftest_philox(pp_ctx* ctx, const uint32_t* ctr, uint64_t seed, uint32_t* words_out,
                       uint64_t n);

/* Sums over n draws of HG cosθ computed on the device in FP64:
 * sums_out = {Σcosθ, Σcos²θ}.  For the mean-cosine check at ≥1e7 draws. */
int pp_selftest_hg_moments(pp_ctx* ctx, float g, uint64_t seed, uint64_t n, double* sums_out);

#ifdef __cplusplus
}
#endif

#endif /* PHOTONPROP_H */
