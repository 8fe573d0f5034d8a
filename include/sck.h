/* sck.h -- C-ABI of libsck.so, the MI355X backend of the multi-colour Metropolis sweep.
 *
 * The reference has no native boundary: its kernels are Triton source generated per system
 * (SCK/generator.py:322-337 compile_system, SCK/bitwise.py:250-335) and launched by the Sampler operator class
 * (SCK/sampler.py:39-345). This header is the boundary that replaces that pair: the host side generates HIP
 * source per system, and libsck compiles it with hipRTC for gfx950, launches its class kernels, captures
 * sweeps in HIP graphs and runs the observable reductions. Plain pointers and sizes only; every device buffer
 * (spins, mask, k1, beta, tables, state) belongs to the caller and is read by address, so buffers updated in
 * place are seen by graphs captured earlier (SCK/sampler.py:134-150).
 *
 * Errors: every entry point returns 0 on success and a negative SCK_E* code otherwise; sck_last_error()
 * returns the calling thread's message for the last failure.
 */
#ifndef SCK_H
#define SCK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCK_ABI_VERSION 6

enum sck_status {
    SCK_OK = 0,
    SCK_E_INVALID = -1,   /* bad argument (ValueError on the Python side) */
    SCK_E_COMPILE = -2,   /* hipRTC compilation failed (log in sck_last_error) */
    SCK_E_HIP = -3,       /* a HIP runtime call failed */
    SCK_E_NOMEM = -4
};

/* Spin storage / kernel family of a module (SCK/sampler.py:40-69). */
enum sck_mode {
    SCK_MODE_INT8 = 0,          /* Ising, int8 +-1 per replica               */
    SCK_MODE_F32 = 1,           /* float32 components (Ising stored as float, XY, Heisenberg) */
    SCK_MODE_PACKED = 2,        /* Ising, 32 replicas per int32 word, per-replica Philox draws */
    SCK_MODE_BITWISE = 3,       /* bitwise multi-spin Ising, one uniform per (word, site) */
    SCK_MODE_BITWISE_INDEP = 4, /* bitwise multi-spin Ising, one 32-bit U per replica */
    SCK_MODE_F64 = 5            /* float64 components, updated in float32 arithmetic: the reference's storage
                                   dtype cast (SCK/generator.py:214 loads as float32, :310-319 stores the cast) */
};

/* Arguments every generated kernel receives, by value (one struct parameter).
 * Mirrors the runtime arguments of the reference kernel kernel(src_ptr, dst_ptr, mask_ptr, k1_ptr, beta_ptr,
 * state_ptr, x_only, row0, ...) (SCK/generator.py:191-194; SCK/bitwise.py:278-281). */
typedef struct sck_args {
    void *spins;              /* S [B][n][b][L3][L2][L1] (or words [W][1][b][L3][L2][L1]); updated in place */
    void *dst;                /* synchronous control: destination of the update. Whole sweeps of a module
                                 with a tiled launch: a workspace of the spins' size (the sweeps alternate
                                 between the two buffers; the result ends in spins). Else NULL          */
    const int8_t *mask;       /* [b][L3][L2][L1], 1 = spin present                               */
    const float *k1;          /* [b][L3][L2][L1] cubic K1 per site                               */
    const float *beta;        /* [Bp] inverse temperature per replica                            */
    const void *table;        /* acceptance table: Ising fp32 [Bp][T]; bitwise fp32/int64 [W][2][Z+1][Z+1] */
    const int64_t *state;     /* [2] = (base seed, sweep counter)                                */
    const float *width;       /* cone proposal: half-angle per replica [Bp], or the learned width table */
    const float *site_ku;     /* [b][L3][L2][L1] per-site uniaxial strength                      */
    const float *site_axis;   /* [b][L3][L2][L1][n] per-site uniaxial axis                        */
    double *obs;              /* observables: output E [B] then M [B][b][n]; partial sums in obs_scratch */
    double *obs_scratch;      /* [B / obs_blocks][chunks][1 + b*n] per-block partial sums          */
    int64_t sweep_offset;     /* sweep index = state[1] + sweep_offset                            */
    int32_t nrep;             /* replicas B (words W for packed/bitwise)                          */
    int32_t nsweep;           /* sweeps one launch of a resident kernel performs (set by the library) */
    const void *aux;          /* batched systems: the per-system descriptor table                  */
    const int32_t *tiles;     /* batched systems: the launch's work list (set from sck_launch)     */
    int32_t ntile;            /* its length                                                        */
    int32_t rep_base;         /* first replica of the launch's replica block (set from sck_launch) */
    double *track;            /* NULL, or running float64 totals E [B] then M [B][b][n], kept current by whole
                                 sweeps of the tiled kernel (which add the energy and magnetization changes of
                                 their accepted moves); the caller anchors them with sck_observe            */
    double *track_parts;      /* with track: [tiles][B][1 + b n] per-(replica, tile) partial sums of one sweep */
} sck_args;

/* One kernel launch of a plan. */
typedef struct sck_launch {
    int32_t kernel;           /* index into the module's kernel list */
    uint32_t grid[3];
    uint32_t block[3];
    uint32_t shmem;
    const int32_t *tiles;     /* NULL, or the work list handed to the kernel as args.tiles / args.ntile */
    int32_t ntile;
    int32_t rep_base;         /* replica-major class launches: first replica of the block it covers, else 0 */
} sck_launch;

/* Launch plans of a system, computed by the host-side generator. */
typedef struct sck_desc {
    int32_t abi_version;      /* SCK_ABI_VERSION */
    int32_t mode;             /* enum sck_mode */
    int32_t n, b, L1, L2, L3; /* model and box */
    int32_t batch;            /* B, replicas */
    int32_t words;            /* W = ceil(B/32) for packed/bitwise, else B */
    int32_t obs_chunks;       /* observable partial sums per replica; args.obs_scratch holds
                                 batch / max(1, obs_blocks) x obs_chunks x (1 + b n) doubles */
    const char *const *kernel_names; int32_t n_kernels;
    const sck_launch *sweep;  int32_t n_sweep;   /* one coloured sweep: class launches in sweep order */
    const sck_launch *ref;    int32_t n_ref;     /* one sequential sweep (reference chain)           */
    const sck_launch *sync;   int32_t n_sync;    /* one synchronous-control sweep                    */
    const sck_launch *obs;    int32_t n_obs;     /* energy and magnetization                          */
    const sck_launch *bump;                      /* state[1] += sweep_offset                          */
    const sck_launch *resident;                  /* NULL, or whole sweeps with each replica's lattice in LDS:
                                                    one launch performs args.nsweep sweeps                */
    int32_t rep_blocks;       /* >= 1: the sweep list is rep_blocks groups of n_sweep / rep_blocks class
                                 launches, one group per replica block. sck_sweep runs each block through
                                 all n sweeps before the next (replicas are independent chains, so the chain
                                 of every replica is unchanged) and a block's lattice stays cache-resident */
    int32_t obs_blocks;       /* 0 or 1: the observation list covers every replica at once. > 1 (equal to
                                 rep_blocks): it is obs_blocks equal groups, one per replica block, each
                                 writing its partial sums to the same first batch / obs_blocks rows of
                                 args.obs_scratch; a graph captured with an observation observes each block
                                 right after that block's sweeps */
    const sck_launch *tiled;                     /* NULL, or one whole sweep in one launch: each workgroup runs
                                                    every class of one (replica, tile) in LDS, reading
                                                    args.spins and writing args.dst. Whole sweeps use it when
                                                    args.dst is a workspace of the spins' size (unless
                                                    `resident` is set); otherwise the class launches run */
    const sck_launch *track_fold;                /* with `tiled`: adds one sweep's partial sums (args.track_parts)
                                                    to the running totals args.track                            */
} sck_desc;

/* ---------------------------------------------------------------- learned proposal width (no reference
 * counterpart: the reference proposes a uniform redraw, SCK/generator.py:274-289). */

/* One training sample, taken from a warm-up configuration: a site's spin, the local field of its neighbours
 * (exchange + DMI + field) and the site's constants. */
typedef struct sck_mlp_sample {
    float s[3];
    float v[3];
    float beta;
    float k1;                 /* cubic K1 of the site        */
    float ku;                 /* per-site uniaxial strength  */
    float axis[3];            /* per-site uniaxial axis      */
} sck_mlp_sample;

typedef struct sck_mlp_config {
    int32_t n;                /* spin components, 2 or 3                      */
    int32_t ndraw;            /* cone draws per sample                        */
    int32_t seed;             /* Philox key of the draws                      */
    float wmin, wmax;         /* half-angle range of the cone                 */
    float k2;                 /* cubic K2                                     */
    float du;                 /* uniform uniaxial strength                    */
    float axis[3];            /* its axis, also the feature axis e            */
    float xs[3];              /* input scaling of the three features          */
} sck_mlp_config;

typedef struct sck_mlp_adam_config {
    float lr, beta1, beta2, eps;
    int32_t step;             /* 1-based Adam step; lr = 0 computes the gradient only */
} sck_mlp_adam_config;

typedef struct sck_width_grid {
    int32_t n[3];             /* bins per feature        */
    float s[3];               /* bins per unit feature   */
} sck_width_grid;

typedef struct sck_handle sck_handle;
typedef struct sck_graph sck_graph;

/* Library version and the device architecture the module cache is keyed on ("gfx950"). */
int sck_version(void);
const char *sck_last_error(void);

/* Compile `hip_src` with hipRTC for gfx950 (or load it from the code-object cache `cache_dir`, NULL = no
 * cache), load the module on the current device and resolve the kernels named in desc.
 * Replaces compile_system (SCK/generator.py:322-337) plus Sampler.__init__'s launch geometry
 * (SCK/sampler.py:70-131). */
int sck_create(const sck_desc *desc, const char *hip_src, const char *cache_dir, sck_handle **out);

/* Compile only (no device needed): writes the code object into cache_dir. Used to pre-build kernels. */
int sck_precompile(const char *hip_src, const char *cache_dir);

int sck_destroy(sck_handle *h);

/* n coloured sweeps in place, then state[1] += n (SCK/sampler.py:234-240). stream: a hipStream_t or NULL.
 * class_lo/class_hi select launches [lo, hi) of one sweep (0, -1: all); used by per-class checks.
 * A module with a tiled launch runs whole sweeps through it when args->dst points to a workspace of the spins'
 * size: the sweeps alternate between the two buffers and the result ends in args->spins.
 * With args->track set, the sweeps must run tiled (whole sweeps, args->dst given), and each adds the energy and
 * magnetization changes of its accepted moves to args->track: float64 totals that stay equal to sck_observe's
 * up to the fp32 rounding of each move's dE (SCK_E_INVALID for any other path). */
int sck_sweep(sck_handle *h, const sck_args *args, int32_t n, int32_t class_lo, int32_t class_hi, void *stream);

/* sck_sweep of n whole sweeps with a HIP event pair on `stream` around every sweep-kernel launch (the class
 * kernels, or the resident kernel; not the counter bump). Synchronizes the stream, then writes the summed
 * elapsed milliseconds of those launches to *kernel_ms and their number to *launches. For measuring the
 * sweep kernel's own duration; no counterpart in the reference (its benchmark times whole sweeps,
 * SCK/sampler.py:234-240). */
int sck_sweep_timed(sck_handle *h, const sck_args *args, int32_t n, void *stream, float *kernel_ms,
                    int32_t *launches);

/* n sequential sweeps (one thread per replica or word visiting the sites in class order), then
 * state[1] += n (SCK/sampler.py:281-293). */
int sck_reference_sweep(sck_handle *h, const sck_args *args, int32_t n, void *stream);

/* n synchronous-control sweeps: every active site from the old state (SCK/sampler.py:271-279).
 * args->dst must hold a copy of args->spins' size; the result is copied back into args->spins. */
int sck_sync_sweep(sck_handle *h, const sck_args *args, int32_t n, int64_t spin_bytes, void *stream);

/* Float64 energy [B] and per-sublattice magnetization [B][b][n] into args->obs (SCK/sampler.py:326-345). */
int sck_observe(sck_handle *h, const sck_args *args, void *stream);

/* Bytes of the static device array behind sck_observe_noalloc (per GPU, held until the process ends). */
#define SCK_OBS_SCRATCH_BYTES (32u << 20)

/* sck_observe without a caller-provided workspace: the per-block partial sums go to a static device array of
 * SCK_OBS_SCRATCH_BYTES in the library (args->obs_scratch is ignored), so the call allocates nothing (no hipMalloc,
 * hipFree or allocating library routine). No counterpart in the reference; it is for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses that forbid allocation. Returns
 * SCK_E_INVALID when the partial sums (batch / obs_blocks x obs_chunks x (1 + b n) doubles) exceed that size. Every call on
 * one GPU shares its array, which keeps the last call's contents; calls that overlap on one GPU race. */
int sck_observe_noalloc(sck_handle *h, const sck_args *args, void *stream);

/* Flags of sck_graph_capture's `observe` argument. */
#define SCK_CAPTURE_OBSERVE 1   /* an observation after the sweeps */
#define SCK_CAPTURE_IN_DST 2    /* whole tiled sweeps leave an odd count's result in args->dst instead of copying it
                                   back into args->spins, and the observation reads it there: two graphs with the
                                   buffers exchanged, replayed in turn, measure after every sweep with no copies.
                                   No counterpart in the reference; other paths update args->spins in place as
                                   always. */

/* Capture k sweeps (launches and counter bump; with SCK_CAPTURE_OBSERVE an observation after them) in a HIP graph
 * (SCK/sampler.py:242-269). The graph reads the buffers of `args` by address. With replica blocks and a
 * per-block observation list, each block's observation follows that block's k sweeps (the same values as
 * observing after all of them: the blocks are independent chains). */
int sck_graph_capture(sck_handle *h, const sck_args *args, int32_t k, int32_t observe, sck_graph **out);
int sck_graph_replay(sck_graph *g, void *stream);
int sck_graph_destroy(sck_graph *g);

/* Elapsed milliseconds of `reps` replays of a graph (or of k sweeps when g is NULL), timed with HIP events
 * on `stream`. */
int sck_time_replays(sck_graph *g, int32_t reps, void *stream, float *ms);

/* Learned proposal: number of network parameters (3 -> 32 -> 32 -> 1), workspace bytes of one training
 * step, one training step (forward + backward over nsample samples of ndraw cone draws each, fixed-order
 * gradient sums, Adam; grad_out receives the gradient of L = -mean log(1e-4 + J_i), J_i a sample's expected
 * squared jump, and, last, the sum of log(1e-4 + J_i)), and the width table of the frozen network on a feature
 * grid.
 * No counterpart in the reference. */
int sck_mlp_nparam(void);
int64_t sck_mlp_workspace_bytes(int32_t nsample, int32_t ndraw);
int sck_mlp_train_step(float *params, float *adam_m, float *adam_v, const sck_mlp_sample *samples, int32_t nsample,
                       const sck_mlp_config *cfg, const sck_mlp_adam_config *adam, void *workspace, float *grad_out,
                       void *stream);
int sck_mlp_table(const float *params, const sck_mlp_config *cfg, const sck_width_grid *grid, float *table,
                  void *stream);

/* Training samples from the current spins: for rep_site[2i], rep_site[2i+1] (replica, site) the module's
 * sck_sample kernel writes out[i]. */
int sck_sample_sites(sck_handle *h, const sck_args *args, const int32_t *rep_site, int32_t nsample,
                     sck_mlp_sample *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif
