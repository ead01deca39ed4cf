/*
 * lgf_flow.h -- C-ABI of the MI355X (gfx950) Wilson-flow and clover-observable
 * path.  Plain pointers and sizes only; every call returns an lgf_status and,
 * on failure, leaves a message retrievable with lgf_ctx_last_error().
 *
 * The reference (LGF) has no FFI layer of its own; these entry points replace
 * the CPU work underneath its C++ host API:
 *
 *   lgf_field_upload / lgf_field_download
 *       the link storage of SU3_field after loadSU3FromFile / before
 *       saveSU3ToFile (LGF/Field/SU3_field.cpp:260-352, 89-172); the buffer is
 *       in the reference's saved-configuration order, links[((g*4+mu)*9+k)*2+ri].
 *   lgf_flow_steps
 *       GradientFlow::flow() (LGF/Gradient_Flow/GradientFlow.cpp:22-60) with
 *       Wilson::calculate_FlowGradient (LGF/Action/Gauge/Wilson.cpp:55-71)
 *       and HermTrLessExp (LGF/Math/SU3_mat.cpp:216-289).
 *   lgf_measure
 *       EnergyDensity::calculate (LGF/Observables/EnergyDensity.cpp:9-34),
 *       TopologicalCharge::calculate (LGF/Observables/TopologicalCharge.cpp:8-29)
 *       and the per-site sum of QdensityDistribution::calculate
 *       (LGF/Observables/QdensityDistribution.cpp:15-28).
 *   lgf_diagnostics
 *       SU3_field::Avg_Plaquette (LGF/Field/SU3_field.cpp:485-499) and the
 *       unitarity check of isSpecialUnitary (LGF/Math/SU3_mat.cpp:420-454).
 *   lgf_batch_*
 *       the reference's loop over configurations in
 *       FlowSavedGaugeConfigurations (LGF/LGF.cpp:169-195), done as one set of
 *       launches over all configurations of a batch.
 *   lgf_server_*
 *       no counterpart in the reference (a batch CLI tool); a continuous-
 *       batching request loop for flow jobs.
 *
 * Lifetimes: destroy batches before the fields they hold and fields before
 * their context; out-of-order destroys return LGF_E_STATE and change nothing.
 *
 * Threading: calls on one context are not reentrant; use one context per
 * host thread.  The server runs its own scheduler thread and is safe to
 * submit to from any thread.
 */
#ifndef LGF_FLOW_H
#define LGF_FLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lgf_ctx lgf_ctx;
typedef struct lgf_field lgf_field;
typedef struct lgf_batch lgf_batch;
typedef struct lgf_server lgf_server;

enum lgf_status {
    LGF_OK = 0,
    LGF_E_ARG = 1,       /* null pointer, bad enum, bad count */
    LGF_E_SHAPE = 2,     /* odd or non-positive extent, volume too large */
    LGF_E_IO = 3,        /* file missing, short or unwritable */
    LGF_E_NOMEM = 4,     /* device or host allocation failed */
    LGF_E_HIP = 5,       /* HIP runtime error */
    LGF_E_NAN = 6,       /* non-finite observable */
    LGF_E_STATE = 7,     /* object used in an invalid state */
    LGF_E_ADMISSION = 8, /* request larger than the server's memory budget */
    LGF_E_UNSUPPORTED = 9
};

enum lgf_precision {
    LGF_FP64 = 0, /* FP64 links and arithmetic: the parity path */
    LGF_FP32 = 1  /* FP32 links and arithmetic, FP64 reductions, FP64 reunitarisation per step */
};

enum lgf_measure_flags {
    LGF_MEASURE_DEFAULT = 0, /* clover arithmetic in the field's precision */
    LGF_MEASURE_FP64 = 1     /* clover arithmetic in FP64 also for FP32 fields */
};

/* observables of one clover sweep (lgf_measure_observables) */
enum lgf_observable {
    LGF_OBS_E = 0,            /* EnergyDensity: (1/V) sum_x sum_{mu<nu} ReTr(F F), 1x1 clover F */
    LGF_OBS_Q = 1,            /* TopologicalCharge: 1/(4 pi^2) sum_x qraw(x), improved F */
    LGF_OBS_QABS = 2,         /* 1/(4 pi^2) sum_x |qraw(x)|, a scale for judging Q differences */
    LGF_OBS_E_IMPROVED = 3,   /* ImprovedEnergyDensity: as E with the improved F */
    LGF_OBS_Q_UNIMPROVED = 4, /* UnImprovedTopologicalCharge: as Q with the 1x1 clover F */
    LGF_NUM_OBS = 5
};

/* ---- context ---------------------------------------------------------- */
int lgf_ctx_create(int device, lgf_ctx** out);
int lgf_ctx_destroy(lgf_ctx* ctx);
const char* lgf_ctx_last_error(const lgf_ctx* ctx);
/* the hipStream_t every call of this context is ordered on */
int lgf_ctx_stream(lgf_ctx* ctx, void** hip_stream);
int lgf_ctx_synchronize(lgf_ctx* ctx);
/* RK3 stage-kernel variant (register budget x workgroup shape) used by this
 * context for one precision; 0 = default.  Variants below
 * LGF_FP32_SHARED_PLAQUETTE_VARIANTS (mixed precision) and
 * LGF_FP64_SHARED_PLAQUETTE_VARIANTS (FP64) compute each plaquette once for
 * the four links of a site; the others sum each link's staples in the
 * reference's order.  Of the mixed-precision shared-plaquette variants,
 * those from LGF_FP32_COMPACT_VARIANTS on also share the backward links and
 * rebuild row 2 of the shared links from rows 0 and 1.  Variants of one
 * group compute bitwise identical links; the groups agree to rounding in the
 * path's precision.  For FP64 the autotuners keep the group of the
 * context's current variant, so tuning never changes an FP64 result. */
#define LGF_FP32_SHARED_PLAQUETTE_VARIANTS 8
#define LGF_FP32_COMPACT_VARIANTS 6
#define LGF_FP64_SHARED_PLAQUETTE_VARIANTS 3
int lgf_ctx_set_stage_variant(lgf_ctx* ctx, int precision, int variant);
int lgf_ctx_get_stage_variant(lgf_ctx* ctx, int precision, int* variant);
int lgf_num_stage_variants(int precision, int* n);
/* traversal order of the lattice tiles in batches created afterwards:
 * 0 = lexicographic, B in 2..64 = x0 blocked by B (better cache reuse of
 * the x0-neighbours), 100*B2 + B0 = x0 blocked by B0 and x2 by B2 (for
 * lattices whose x2-x3 planes are large); results do not depend on it */
int lgf_ctx_set_tile_order(lgf_ctx* ctx, int block);
int lgf_ctx_get_tile_order(lgf_ctx* ctx, int* block);
/* cache groups of batches created afterwards: the fields of a batch are
 * split, in order, into groups whose device footprint (lgf_field_footprint)
 * stays within `bytes` (a larger field forms a group of its own), and each
 * group takes all the steps of an lgf_batch_flow_steps call before the next
 * group starts, so that a group that fits the GPU's last-level cache
 * (256 MB) stays there across its stages.  0 = the whole batch is one group
 * (the default).  Results do not depend on it; lgf_autotune_batch chooses
 * it. */
int lgf_ctx_set_batch_group_bytes(lgf_ctx* ctx, size_t bytes);
int lgf_ctx_get_batch_group_bytes(lgf_ctx* ctx, size_t* bytes);
/* device scratch a measurement may hold for its line products (2- and
 * 3-link products of every link, 576 B per site for FP32 clover arithmetic,
 * 1152 B for FP64): the fields of a batch are measured in consecutive groups
 * that fit `bytes`, a single larger field on its own.  Default 1 GiB.  The
 * scratch is kept by the context until it is destroyed. */
int lgf_ctx_set_measure_scratch(lgf_ctx* ctx, size_t bytes);
/* tile orders the autotuners time (lgf_ctx_set_tile_order codes) */
#define LGF_AUTOTUNE_TILE_ORDERS {0, 2, 4, 8, 404, 408, 604, 1204, 6402, 6404}
/* time every pair of tile order (LGF_AUTOTUNE_TILE_ORDERS) and stage variant for
 * nsteps graph-replayed RK3 steps on a seeded lattice of `shape` (each pair three times, its
 * fastest run counting); time the four fastest pairs again with 3*nsteps
 * steps, then every variant at the order of the fastest of those, and make
 * that order and the fastest of those variants this context's tile order and
 * stage variant for `precision`; ms_per_step (optional) gets those last
 * times, one entry per variant (HUGE_VAL for FP64 variants outside the group
 * of the context's variant, which are not timed).  No counterpart in the reference
 * (its flow has no launch parameters to choose). */
int lgf_autotune(lgf_ctx* ctx, const int shape[4], int precision, int nsteps, int* best, double* ms_per_step);
/* the same search on `nshapes` seeded lattices (shapes[4*i .. 4*i+3])
 * flowed together as one batch: tunes the context for batches of that mix
 * of shapes, whose best order and variant can differ from any one
 * lattice's.  With more than one lattice the pairs are timed with the
 * batch as one group and in cache groups of 224 MiB
 * (lgf_ctx_set_batch_group_bytes), and the cap is then chosen among 0, 128,
 * 192, 224 and 256 MiB at the chosen pair; ms_per_step is at the cap of the
 * search that the chosen pair was timed with.  The lattices of all shapes are held at once (the sum of their
 * lgf_field_footprint) and released before it returns.  LGF_E_ARG for
 * nshapes <= 0, LGF_E_SHAPE for a bad shape; on failure the context's
 * tile order, variant and group cap are left as they were. */
int lgf_autotune_batch(lgf_ctx* ctx, const int* shapes, int nshapes, int precision, int nsteps, int* best,
                       double* ms_per_step);

/* ---- one lattice ------------------------------------------------------ */
/* device bytes held by a field of this shape and precision */
int lgf_field_footprint(const int shape[4], int precision, size_t* bytes);
int lgf_field_create(lgf_ctx* ctx, const int shape[4], int precision, lgf_field** out);
int lgf_field_destroy(lgf_field* f);
int lgf_field_shape(const lgf_field* f, int shape[4], int* precision);
/* links: volume*4*18 doubles in the reference's saved-configuration order */
int lgf_field_upload(lgf_field* f, const double* links);
int lgf_field_download(lgf_field* f, double* links);
/* flow time of the field: advanced by eps at every RK3 step it takes, as
 * GradientFlow::flow accumulates it (LGF/Gradient_Flow/GradientFlow.cpp:
 * 52-58); 0 at creation.  Set it as GradientFlow::SetFlowTime does
 * (GradientFlow.cpp:64-66) when a flow is resumed from a saved
 * configuration.  It labels measurements (lgf_batch_track_t0); the flow
 * itself does not depend on it. */
int lgf_field_flow_time(const lgf_field* f, double* t);
int lgf_field_set_flow_time(lgf_field* f, double t);
int lgf_field_set_cold(lgf_field* f);
/* seeded Haar-random SU(3) links (a "hot start"; same Philox4x32-10 counters
 * as the CPU oracle's, values equal to rounding) */
int lgf_field_set_random(lgf_field* f, uint64_t seed);
/* seeded configuration generator: nsweeps iterations of [one Cabibbo-Marinari
 * heatbath sweep + or_per_hb overrelaxation sweeps] at coupling beta, with
 * iteration indices first_sweep.. so that a chain can be continued; links
 * are reunitarised after every iteration.  Same algorithm and Philox
 * counters as the CPU oracle's generator (the reference's own generator,
 * LGF/SU3_Heatbath/SU3_Heatbath.cpp:16-224, is not seedable). */
int lgf_field_heatbath(lgf_field* f, double beta, uint64_t seed, int first_sweep, int nsweeps, int or_per_hb);
/* copy links of src into dst (same shape; precision converted) */
int lgf_field_copy(lgf_field* dst, const lgf_field* src);
/* read/write a configuration file of the reference's format at `path`
 * (headerless, 144 B per link, 64-bit offsets) */
int lgf_field_load_file(lgf_field* f, const char* path);
int lgf_field_save_file(lgf_field* f, const char* path);

/* nsteps RK3 steps of size eps */
int lgf_flow_steps(lgf_field* f, double eps, int nsteps);
/* E = (1/V) sum_x sum_{mu<nu} ReTr(F_munu F_munu) with the 1x1 clover F;
 * Q = 1/(4 pi^2) sum_x qraw(x) with the improved F;
 * Qabs = 1/(4 pi^2) sum_x |qraw(x)| (a scale for judging Q differences);
 * qraw (optional, volume doubles at the global index) =
 *   ReTr(F32 F01) + ReTr(F31 F20) + ReTr(F30 F12) per site. */
int lgf_measure(lgf_field* f, int flags, double* E, double* Q, double* Qabs, double* qraw);
/* all lgf_observable values of one sweep: obs[LGF_NUM_OBS]; covers also
 * ImprovedEnergyDensity::calculate (LGF/Observables/ImprovedEnergyDensity.cpp:9-32)
 * and UnImprovedTopologicalCharge::calculate
 * (LGF/Observables/UnImprovedTopologicalCharge.cpp:8-29), the observables of
 * MeasureFlowedGaugeConfigurations (LGF/LGF.cpp:228-255) */
int lgf_measure_observables(lgf_field* f, int flags, double* obs, double* qraw);
/* Allocation-free lgf_measure_observables (without q(x)).  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its scratch (descriptor, tile list, clover line products,
 * tile partials, sums) is a static device array of 256 MiB that each GPU the
 * library runs on holds until the process ends; it calls no hipMalloc or
 * hipFree.  Fields up to 465408 sites (FP32 clover arithmetic) or 232832
 * sites (FP64 arithmetic: LGF_FP64 fields or LGF_MEASURE_FP64) fit; larger
 * ones give LGF_E_ARG.  The observables are bitwise those of
 * lgf_measure_observables.  Calls on one GPU share the array and are
 * serialised within the process; each returns after its results are back
 * on the host. */
int lgf_measure_observables_noalloc(lgf_field* f, int flags, double* obs);
int lgf_diagnostics(lgf_field* f, double* avg_plaquette, double* unitarity_max, double* unitarity_mean);
/* flow-time scale t0 of a measured series: the t at which t^2 E(t) first
 * reaches `target` (0.3 defines t0, doc/thesis.tex:1214-1216; the reference
 * leaves it to its analysis scripts, DataAnalysis_and_Plotting/
 * t0_extrapolation.py), by linear interpolation of t^2 E between the two
 * bracketing measurements.  t ascending, n >= 2.  LGF_E_STATE if the series
 * never reaches the target.  Host-side; needs no context. */
int lgf_flow_scale_t0(const double* t, const double* E, int n, double target, double* t0);

/* ---- many lattices (any shapes) in one set of launches ----------------- */
int lgf_batch_create(lgf_ctx* ctx, lgf_field* const* fields, int nfields, lgf_batch** out);
int lgf_batch_destroy(lgf_batch* b);
/* eps: one value per field.  use_graph != 0 replays the RK3 steps from a
 * captured hipGraph (two steps per graph). */
int lgf_batch_flow_steps(lgf_batch* b, const double* eps, int nsteps, int use_graph);
int lgf_batch_measure(lgf_batch* b, int flags, double* E, double* Q, double* Qabs);
/* obs[LGF_NUM_OBS * field + lgf_observable] */
int lgf_batch_measure_observables(lgf_batch* b, int flags, double* obs);
/* deferred measurement: _start enqueues the clover sweep of the current
 * links (after all work enqueued on the context so far) and returns at once,
 * so flow steps can be enqueued behind it; _finish waits and returns the
 * observables of that flow time (as lgf_batch_measure_observables).  One
 * measurement per batch may be outstanding (else LGF_E_STATE).  No
 * counterpart in the reference, which measures between flow steps
 * (LGF/Gradient_Flow/GradientFlow.cpp:52-58). */
int lgf_batch_measure_start(lgf_batch* b, int flags);
int lgf_batch_measure_finish(lgf_batch* b, double* obs);
/* On-device flow-scale tracking, the t0 of lgf_flow_scale_t0 kept on the
 * GPU: after lgf_batch_track_t0, every measurement of the batch (synchronous
 * or deferred) also updates, in stream order and without synchronising,
 * each field's t0 = the first flow time at which t^2 E reaches `target`,
 * by linear interpolation of t^2 E between that measurement and the
 * previous one, t being the field's flow time (lgf_field_flow_time) when
 * the measurement was enqueued.  lgf_batch_t0 waits for the stream and
 * returns t0 per field, NaN for a field that has not reached the target.
 * Calling lgf_batch_track_t0 again restarts the tracking.  The reference
 * finds t0 in post-processing (DataAnalysis_and_Plotting/t0_extrapolation.py,
 * t^2 E(t0) = 0.3, doc/thesis.tex:1214-1216). */
int lgf_batch_track_t0(lgf_batch* b, double target);
int lgf_batch_t0(lgf_batch* b, double* t0);
int lgf_measure_start(lgf_field* f, int flags);
int lgf_measure_finish(lgf_field* f, double* obs);

/* ---- continuous-batching flow server ----------------------------------- */
typedef struct {
    int shape[4];
    int precision;          /* lgf_precision */
    double beta;            /* carried into the result, not used by the flow */
    double eps;
    int nsteps;
    int measure_every;      /* measure at t=0 and every k steps; 0 = only at t=0 and the end */
    const double* links;    /* initial links in file order (copied at submit), or NULL: */
    int start;              /*   links == NULL: 0 = cold start, 1 = seeded Haar-random links */
    uint64_t seed;          /*   seed of the Haar-random start */
    int measure_flags;      /* lgf_measure_flags */
} lgf_request;

typedef struct {
    uint64_t id;
    int status;             /* lgf_status of the job */
    int nmeas;
    double submit_s, start_s, finish_s; /* seconds on the server's steady clock */
    double t[1024];         /* flow times of the measurements (first nmeas entries) */
    double E[1024];
    double Q[1024];
} lgf_result;

/* budget_bytes: device memory the server may hold (0 = 80% of free): the
 * jobs' fields (lgf_field_footprint) plus the measurement scratch, kept
 * aside as max(1 GiB, the line products of the largest admitted lattice:
 * 576 B per site for FP32 clover arithmetic, 1152 B for FP64);
 * max_active: jobs flowing at once (0 = unlimited within the budget) */
int lgf_server_create(int device, size_t budget_bytes, int max_active, lgf_server** out);
int lgf_server_destroy(lgf_server* s);
/* LGF_E_ARG for a request that would record more than 1024 measurements
 * (the capacity of lgf_result): 1 at t = 0, one per multiple of
 * measure_every up to nsteps, and one at the end when nsteps is not such a
 * multiple (measure_every = 0: t = 0 and the end) */
int lgf_server_submit(lgf_server* s, const lgf_request* req, uint64_t* id);
/* wait (timeout_s < 0: forever) for the job; LGF_E_STATE on timeout */
int lgf_server_wait(lgf_server* s, uint64_t id, double timeout_s, lgf_result* out);
/* scheduling policy.  LGF_SERVE_ALL: every admitted job steps in every
 * round (processor sharing).  LGF_SERVE_SHORTEST_FIRST (default): the
 * admitted jobs step in the order of their remaining link-updates less
 * age_rate times the seconds since they were submitted (default age_rate 0:
 * shortest remaining work first), until they hold min_sites lattice sites
 * (default 2^20), so short jobs finish first while each launch still fills
 * the GPU, and a positive age_rate lets waiting long jobs in; the choice is
 * revisited whenever a job arrives or finishes.  Results do not depend on
 * the policy. */
enum lgf_serve_policy { LGF_SERVE_ALL = 0, LGF_SERVE_SHORTEST_FIRST = 1 };
int lgf_server_set_policy(lgf_server* s, int policy, int64_t min_sites, double age_rate);
/* lgf_autotune_batch on the server's own context, for the mix of lattice
 * shapes it is expected to serve.  Only while the server is idle (every
 * submitted job finished): LGF_E_STATE otherwise, also while another
 * call of it runs.  Jobs submitted while it runs wait for it.  Not to be
 * called concurrently with lgf_server_destroy. */
int lgf_server_autotune(lgf_server* s, const int* shapes, int nshapes, int precision, int nsteps, int* best);
int lgf_server_now(lgf_server* s, double* now_s);
/* Fault injection for tests of the scheduler's error path (no counterpart in
 * the reference): the next `count` batch creations of the scheduler fail as
 * if the device were out of memory.  The jobs that were to step in them then
 * finish with status LGF_E_NOMEM, keep the measurements recorded before the
 * failure, and free their memory; other jobs are not affected. */
int lgf_server_inject_batch_failures(lgf_server* s, int count);
/* counters: jobs done, steps executed, link-updates executed, busy seconds */
int lgf_server_stats(lgf_server* s, uint64_t* jobs_done, uint64_t* batch_steps, double* link_updates,
                     double* busy_s);
const char* lgf_server_last_error(const lgf_server* s);

#ifdef __cplusplus
}
#endif
#endif /* LGF_FLOW_H */
