/* motion_hip — MI355X (gfx950) contact step behind the reference ContactEngine protocol.
 *
 * C ABI of libmotion_hip.so. Plain pointers and sizes only. Every device buffer belongs to the
 * caller: the body state arrays (me_bodies) and one workspace block whose size me_workspace_size
 * reports. The library itself calls no hipMalloc; it keeps a small pinned host block for counter
 * read-backs. Return codes follow the reference's only C ABI
 * (src/motion_engine/body_pairs_rt_native/api.cpp:130-163): 0 ok, 1 bad argument, 2 runtime
 * failure, 3 wrong thread, plus 4 capacity exceeded and 5 non-finite state.
 *
 * Reference interfaces replaced (antonbj3/3fold-motion-engine, src/motion_engine/...):
 *   me_create / me_destroy     the engine constructors + _build
 *                              (contact_engine_gpu.py:195-239, contact_engine_gpu_colored.py:798-889,
 *                               contact_engine_gpu_color_cache.py:21-35)
 *   me_step                    ContactEngine.step(dt, substeps) (contact_engine.py:52;
 *                              contact_engine_gpu.py:252-280, contact_engine_gpu_colored.py:1302-1375,
 *                              contact_engine_gpu_deferred_force.py:13-64)
 *   me_buffer                  the harness-visible solver arrays (ckeys, sbi, sbj, spA, spB, sn, spen,
 *                              jn, jt1, jt2, jp; tests/test_island_partition.py:70-71)
 *   me_read_counters           n_contacts, n_contacts_solved, n_pairs, n_colors, color_cache_hits/misses
 *   me_last_error              the exception text of the reference's ValueError / RuntimeError
 * get_state, set_kinematic and contact_forces (contact_engine.py:53-55) read or write the caller's
 * body tensors directly; contact_forces uses the ME_BUF_VSNAP snapshot.
 */
#ifndef MOTION_HIP_H
#define MOTION_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ME_OK 0
#define ME_EARG 1
#define ME_ERUNTIME 2
#define ME_ETHREAD 3
#define ME_ECAPACITY 4
#define ME_ENONFINITE 5

#define ME_SOLVER_JACOBI 0       /* RelaxedJacobiContactEngine (contact_engine_gpu.py:191) */
#define ME_SOLVER_PAIR_CHUNKS 1  /* GraphColoredContactEngine, pair_chunks=True (contact_engine_gpu_colored.py:793) */
#define ME_SOLVER_COLOR_CACHE 2  /* ColorCacheContactEngine (contact_engine_gpu_color_cache.py:21) */

#define ME_FLAG_TILED 1          /* solve islands in LDS-resident tiles (one launch per substep) */
#define ME_FLAG_PER_CONTACT 2    /* coloured solver without pair chunks: every contact its own chunk
                                    (GraphColoredContactEngine pair_chunks=False) */
#define ME_FLAG_GRAPH 4          /* me_step replays one hipGraph of the whole step (every substep, every
                                    sweep), captured on the first call for each (dt, substeps), at most 8
                                    kept; the graph solve capture of contact_engine_gpu_colored.py:1228-1299.
                                    Needs ME_FLAG_TILED, box geometry and a coloured solver. Steps taken
                                    while me_profile is on run eagerly. */
#define ME_FLAG_FAST_MATH 8      /* fast arithmetic in the tiled ColorCache sweeps: the same Gauss-Seidel
                                    order, colours and clamps, with the bodies' angular responses and the
                                    reciprocal effective masses computed once per substep, fused
                                    multiply-adds and the hardware reciprocal square root in the friction
                                    cone. Bit-identical from run to run, not to the reference kernels:
                                    after one step impulses within 1e-5 of the largest and positions
                                    within 1e-6 m of the exact solve. Needs ME_FLAG_TILED and the
                                    ColorCache solver; no adjoint recording. Not in the reference. */

typedef struct me_config {
    int32_t solver;        /* ME_SOLVER_* */
    int32_t n_bodies;      /* N */
    int32_t n_beads;       /* 0 = box geometry; 0 < P <= 64 = bead geometry (the reference's body model).
                            * Under the coloured solvers bead contacts are generated per body pair and
                            * reduced to at most 4 per pair (E-optimal); needs pair chunks. */
    int32_t vit, pit;      /* velocity / position iterations */
    float relax;           /* Jacobi relaxation (reference 0.2) */
    float t2d;             /* pair-chunk friction t2 factor (reference 1.0) */
    int32_t warm_start;    /* coloured solvers: carry impulses by feature key */
    float bead_radius;     /* bead geometry: R (reference 0.05) */
    float margin;          /* box geometry: contact margin (0 = penetrating points only) */
    int32_t max_pairs;     /* capacity: candidate body pairs (box) */
    int32_t max_contacts;  /* capacity: contacts */
    int32_t n_colors;      /* 0 = Jones-Plassmann colouring as is; K > 0: split to max(K, JP) classes */
    int32_t tile_bodies;   /* tiled solve: nominal bodies per tile */
    int32_t flags;         /* ME_FLAG_* */
} me_config;

/* Body state, device pointers, SoA float32 (the caller's tensors). */
typedef struct me_bodies {
    float *x;              /* [N,3] centre of mass */
    float *q;              /* [N,4] orientation, xyzw */
    float *v;              /* [N,3] linear velocity */
    float *w;              /* [N,3] angular velocity */
    float *inv_mass;       /* [N]   0 for kinematic bodies */
    float *inv_inertia;    /* [N,3,3] body-frame inverse inertia, row-major */
    float *half;           /* [N,3] box half extents (box geometry) */
    float *mu;             /* [N]   friction coefficient */
    int32_t *kinematic;    /* [N]   1 = kinematic */
    int32_t *scene;        /* [N]   scene id; bodies of different scenes never touch */
    float *rest;           /* [P,3] bead rest positions (bead geometry) */
} me_bodies;

typedef struct me_engine me_engine;

/* buffer ids for me_buffer */
enum {
    ME_BUF_KEY = 0,    /* int64 [C]   feature key per contact (sorted) */
    ME_BUF_BI,         /* int32 [C]   body A */
    ME_BUF_BJ,         /* int32 [C]   body B (-1 = ground) */
    ME_BUF_PA,         /* float [C,3] contact point on A */
    ME_BUF_PB,         /* float [C,3] contact point on B */
    ME_BUF_N,          /* float [C,3] normal, B -> A */
    ME_BUF_PEN,        /* float [C]   penetration */
    ME_BUF_JN,         /* float [C]   accumulated normal impulse */
    ME_BUF_JT1,        /* float [C] */
    ME_BUF_JT2,        /* float [C] */
    ME_BUF_JP,         /* float [C]   accumulated position impulse */
    ME_BUF_PBI,        /* int32 [Pp]  pair body A */
    ME_BUF_PBJ,        /* int32 [Pp]  pair body B */
    ME_BUF_PSTART,     /* int32 [Pp]  first contact of the pair */
    ME_BUF_PCOUNT,     /* int32 [Pp]  contacts of the pair */
    ME_BUF_COLOR,      /* int32 [Pp]  colour of the pair */
    ME_BUF_VSNAP,      /* float [N,3] velocity after gravity, before the contact solve */
    ME_BUF_PV,         /* float [N,3] pseudo-velocity of the last substep */
    ME_BUF_PO,         /* float [N,3] pseudo angular velocity */
    ME_BUF_INVIW,      /* float [N,3,3] world inverse inertia */
    ME_BUF_ISLAND,     /* int32 [N]   island root (tiled solve) */
    ME_BUF_DEBUG,      /* int64 [16]  phase timestamps of the largest big tile (wall clock) */
    ME_BUF_BIG_WORK,   /* int64 [4]   big-tile work since me_profile(e, 1): tiles, pairs, dynamic pair
                                      ends, contacts (for the algorithmic bytes of ME_K_TILE_BIG) */
    ME_BUF_MU,         /* float [C]   friction coefficient of the contact's pair */
    ME_BUF_XSNAP,      /* float [N,3] centres at the start of the last substep */
    ME_BUF_WSNAP,      /* float [N,3] angular velocity before the contact solve */
    ME_BUF_TILE_WORK,  /* int64 [12]  tile work since me_profile(e, 1), four counts (tiles, pairs, dynamic
                                      pair ends, contacts) for each of ME_K_TILE_BIG, ME_K_TILE_SMALL,
                                      ME_K_TILE_GIANT, in that order (the first four are ME_BUF_BIG_WORK) */
    ME_BUF_COUNT_
};

/* counters for me_read_counters */
enum {
    ME_CNT_CONTACTS = 0, /* contacts generated in the last substep (before a manifold reduction) */
    ME_CNT_PAIRS,        /* body pairs with contacts */
    ME_CNT_COLORS,       /* colours of the last solve */
    ME_CNT_CACHE_HITS,
    ME_CNT_CACHE_MISSES,
    ME_CNT_OVERFLOW,     /* overflow flags: bit 0 contacts, 1 pairs, 2 candidates, 3 colours, 4 tile,
                            5 recorded tile beyond one small tile, 6 a giant island's grid barrier
                            waited past its bound (the step's state is not valid), 7 an adjoint
                            contribution beyond the fixed-point range, 8 a one-pass scan waited past
                            its bound for an earlier block (the step's state is not valid) */
    ME_CNT_CANDIDATES,   /* broad-phase candidate pairs */
    ME_CNT_ISLANDS,      /* islands of the last tiled solve */
    ME_CNT_TILES,        /* tiles of the last tiled solve */
    ME_CNT_MAX_TILE_BODIES,
    ME_CNT_GRAPH_CAPTURES, /* whole-step graphs captured so far (ME_FLAG_GRAPH) */
    ME_CNT_SOLVED,       /* contacts solved in the last substep (after the bead manifold reduction) */
    ME_CNT_COUNT_
};

/* Workspace bytes for a configuration (no device work). */
int me_workspace_size(const me_config *cfg, size_t *bytes);

/* Create an engine over caller-owned body arrays and a caller-owned workspace of at least
 * me_workspace_size bytes (256-byte aligned). Zeroes the workspace on `stream`. */
int me_create(const me_config *cfg, const me_bodies *bodies, void *workspace, size_t bytes,
              void *stream, me_engine **out);
int me_destroy(me_engine *e);

/* Hand the engine to the calling thread: every entry point checks that it runs on the owning
 * thread (ME_ETHREAD otherwise, the convention of body_pairs_rt_native/api.cpp:130-163). For a
 * caller whose work moves between threads, e.g. a backward pass run by torch's autograd thread;
 * the previous owner must not use the engine afterwards. */
int me_adopt(me_engine *e);

/* Re-point the engine at (possibly re-allocated) body arrays; same N. */
int me_set_bodies(me_engine *e, const me_bodies *bodies);

/* Advance dt in `substeps` equal substeps (ContactEngine.step). Work is queued on `stream`
 * (a hipStream_t; NULL = the null stream). Returns after queueing; counters and the overflow
 * flag are read with me_read_counters. */
int me_step(me_engine *e, double dt, int substeps, void *stream);

/* Device pointer and element count of an internal buffer (a view into the workspace). */
int me_buffer(me_engine *e, int id, void **ptr, int64_t *count);

/* Copy ME_CNT_COUNT_ counters to host memory (synchronises `stream`). */
int me_read_counters(me_engine *e, int64_t *out, void *stream);

/* kernel ids for me_profile_read */
enum {
    ME_K_SWEEP_VEL = 0,  /* per-colour pair-chunk velocity sweep */
    ME_K_SWEEP_POS,      /* per-colour pair-chunk position sweep */
    ME_K_NARROW,         /* box narrow phase */
    ME_K_TILED,          /* the tiled solve's sweep region on the step's stream (k_t_prep, k_t_small and the
                            wait for the side stream's big and giant tiles): a region, not one kernel */
    ME_K_JAC_VEL,        /* Jacobi velocity pass */
    ME_K_TILE_BIG,       /* k_t_big: big tiles (islands beyond one small tile), on the side stream */
    ME_K_TILE_SMALL,     /* k_t_small: small tiles, on the step's stream */
    ME_K_TILE_GIANT,     /* k_t_giant: islands beyond the big tile's LDS, on the side stream */
    ME_K_ISLANDS,        /* k_t_islands_lds: island build of scenes up to 12 288 bodies */
    ME_K_COLOR,          /* k_t_color: JP colouring of the tiles */
    ME_K_COUNT_
};

/* Time launches of the kernels above with HIP events on the stream each is launched on (the first
 * 8192 launches per kernel after enabling). Not part of the reference interface: measurement only. */
int me_profile(me_engine *e, int enable);
int me_profile_read(me_engine *e, int kernel, double *total_ms, int64_t *launches);

/* Adjoint of the step (tiled ColorCache solver, box bodies, islands within one small tile).
 * No counterpart entry points in the reference's solver classes; they replace the role of the
 * reference's batched NCP adjoint (ncp/ncp_batch_gpu.py:1172-1187, BatchProxADMMGPU.grad_affine)
 * for the contact step itself. A caller differentiates a rollout substep by substep, last first:
 * restore the engine's state (bodies + workspace) from before the substep, record it with
 * me_step_taped, then me_adjoint_substep turns the adjoint of the state after the substep into the
 * adjoint of the state before it. */

/* Leading bytes of the workspace that carry the engine's state from one substep to the next and
 * what me_adjoint_substep reads of a recorded substep; the rest is scratch every substep rewrites
 * before reading. A copy of these bytes plus the body arrays is a complete checkpoint. */
int me_state_size(me_engine *e, size_t *bytes);

/* Copy an engine's state block (me_state_size bytes; dst and src each the workspace or a caller's
 * snapshot of it, 16-byte aligned) by the used items of its capacity-sized arrays: the contacts, pairs,
 * candidates, warm store and colour cache as far as the status words of src count them, the rest
 * whole. No counterpart in the reference: for the adjoint's per-substep snapshots, which copied the
 * whole block. Queued on `stream`. */
int me_state_copy(me_engine *e, void *dst, const void *src, void *stream);

/* NCP-level sensitivity, batched, float64 (replaces the reference's BatchProxADMMGPU.build_active +
 * grad_affine, ncp/ncp_batch_gpu.py:1160-1187 with kernels k_active :555-597 and k_solvesig :716-767,
 * and its CPU form ncp/ncp_ref.py: classify :251, _active_system :786, sensitivity :907).
 * n_env problems of n_c <= 64 contacts each, row-major per environment:
 *   G [3n_c, 3n_c] Delassus operator, b [3n_c] free velocity, mu [n_c], lam [3n_c] solved impulse
 *   (n, t1, t2 per contact), dirs [3n_c, n_dir] directions of b (NULL: the identity, n_dir = 3n_c),
 *   labels_in [n_c] (NULL: classify with tol_lam, tol_rel as ncp_ref.classify).
 * Outputs dlam_dmu [3n_c, n_c], dlam_db [3n_c, n_dir] = (dlam/db) dirs, labels [n_c] (bits 0-1:
 * 0 open, 1 stick, 2 slip; bits 4-7: branch-boundary reasons), status (bit 0 singular: a pivot below
 * rcond * max|A|; bit 1 a branch boundary; bit 2 non-finite input). Derivatives are NaN where status
 * is non-zero, as the reference returns them without a minimum-norm selection. No engine and no
 * thread ownership; everything on `stream`. workspace: me_ncp_workspace_size bytes (0 when the
 * per-environment system fits LDS). */
int me_ncp_workspace_size(int32_t n_env, int32_t n_c, int32_t n_dir, size_t *bytes);
int me_ncp_sensitivity(int32_t n_env, int32_t n_c, int32_t n_dir, const double *G, const double *b,
                       const double *mu, const double *lam, const double *dirs, const int32_t *labels_in,
                       double tol_lam, double tol_rel, double rcond, double *dlam_dmu, double *dlam_db,
                       int32_t *labels, int32_t *status, void *workspace, size_t bytes, void *stream);

/* me_ncp_sensitivity without a workspace argument, for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart
 * in the reference. The matrices of systems beyond the LDS limit go into one fixed device array of
 * ME_NCP_STATIC_BYTES per GPU, held until the process ends; a call whose me_ncp_workspace_size exceeds
 * it fails with ME_ECAPACITY before any launch. The call allocates nothing (no hipMalloc, hipFree or
 * allocating library routine). Every call on a GPU shares that GPU's array, so calls that overlap on
 * one GPU (other streams, other threads) race; serialise them. Same results as me_ncp_sensitivity. */
#define ME_NCP_STATIC_BYTES (16u << 20)
int me_ncp_sensitivity_noalloc(int32_t n_env, int32_t n_c, int32_t n_dir, const double *G, const double *b,
                               const double *mu, const double *lam, const double *dirs, const int32_t *labels_in,
                               double tol_lam, double tol_rel, double rcond, double *dlam_dmu, double *dlam_db,
                               int32_t *labels, int32_t *status, void *stream);

/* NCP-level sensitivity with the reference's SVD rank guard, batched, float64: ncp_ref.sensitivity
 * with selection='minimum_norm' and observable rows (ncp/ncp_ref.py:907-959) and contact_diagnostics
 * (:845-866), over rank_factor and guarded_response (ncp/contact_rank.py:13-30, 33-70). Same inputs
 * as me_ncp_sensitivity for n_env problems of n_c <= 256 contacts (dirs NULL: the identity with
 * n_dir = 3n_c, or no b columns with n_dir = 0), plus n_obs observable rows Q
 * [n_obs, 3n_c] per environment (q_shared: one Q for all). rcond is the rank cutoff relative to the
 * largest singular value; tol the compatibility and invariance tolerance (the reference's 1e-10).
 * Per environment (R = n_c + n_dir response columns, the mu columns first):
 *   X [3n_c, R] the guarded response: -A^+ B (A^+ the pseudo-inverse); NaN in every column when the
 *     branch is on a boundary, and, when A is rank-deficient, in every column unless minimum_norm is
 *     set, and in incompatible columns in either mode; NaN where the response is not finite;
 *   S [3n_c] singular values, descending; V [3n_c, 3n_c] right singular vectors (column k for S[k]),
 *   U [3n_c, 3n_c] left singular vectors (from the same method on A^T, by its descending order);
 *   rank: singular values above rcond * S[0]; the nullspace is V[:, rank:], the left nullspace
 *   U[:, rank:];
 *   compat, finite, resid [R]: column compatible (|L^T B_j| <= tol |B_j|, max-scaled), unguarded
 *   response finite, and the unscaled residual |L^T B_j|;
 *   Y, valid [n_obs, R]: Q X where the row is invariant (Q_k N = 0 at tol), the column compatible and
 *   the value finite, else NaN; invariant, leak [n_obs]: the invariance flag and |Q_k N|;
 *   labels [n_c] and status as me_ncp_sensitivity, with the singular bit set from the rank.
 * workspace: me_ncp_svd_workspace_size bytes. No engine and no thread ownership; everything on
 * `stream`. */
typedef struct {
    double *X, *S, *V, *U;
    int32_t *rank, *compat, *finite;
    double *resid, *Y;
    int32_t *valid, *invariant;
    double *leak;
    int32_t *labels, *status;
} me_ncp_svd_out;
int me_ncp_svd_workspace_size(int32_t n_env, int32_t n_c, int32_t n_dir, size_t *bytes);
int me_ncp_svd(int32_t n_env, int32_t n_c, int32_t n_dir, int32_t n_obs, const double *G, const double *b,
               const double *mu, const double *lam, const double *dirs, const int32_t *labels_in, const double *Q,
               int32_t q_shared, double tol_lam, double tol_rel, double rcond, double tol, int32_t minimum_norm,
               const me_ncp_svd_out *out, void *workspace, size_t bytes, void *stream);

/* Exact-cone NCP projected Gauss-Seidel, the reference's NCPSolverGPU (ncp/ncp_gpu.py:783-1069):
 * lambda_c <- proj_K(lambda_c - rho_c (u_c + Gamma(u_c))) over a deterministic colouring, de Saxce
 * correction on or off, optional TGS soft-step normal row (ncp_gpu.py:491-557), per-body
 * v = v_free + M^-1 J^T lambda from an int64 fixed-point total (scale 1e10). Reals are float
 * (real_bytes 4) or double (8). The solver keeps its colouring between solves (colour cache:
 * ME_PGS_COLOR_CACHE) and lives in one caller-owned workspace block. */
typedef struct me_pgs me_pgs;

#define ME_PGS_DESAXCE 1      /* Gamma(u) = (mu |u_t|, 0, 0) on; off = the convex relaxation */
#define ME_PGS_COLOR_CACHE 2  /* reuse the colouring while pairs, features and masses are unchanged */

/* A BlockScene (ncp_gpu.py:185-205) in device memory; row-major, reals of the solver's width. */
typedef struct me_pgs_scene {
    const void *Ja, *Jb;         /* [n_c,3,6] Jacobian blocks; Jb carries the minus of the pair */
    const void *Minv;            /* [n_b,6,6] inverse mass, world frame */
    const void *v_free;          /* [n_b,6] */
    const void *mu;              /* [n_c] */
    const void *bias;            /* [n_c,3] added to u */
    const void *sep;             /* [n_c] separation; read only in soft-step mode (may be NULL else) */
    const int32_t *body_pairs;   /* [n_c,2], -1 = static */
    const int32_t *feature;      /* [n_c] stable feature ids (the colouring's canonical key) */
} me_pgs_scene;

/* status slots of me_pgs_info */
enum { ME_PGS_NCOLORS = 0, ME_PGS_CACHE_HITS, ME_PGS_CACHE_MISSES, ME_PGS_ISLANDS, ME_PGS_TILES,
       ME_PGS_OVERFLOW, ME_PGS_INFO_COUNT_ };
/* buffers of me_pgs_buffer (device, solver width unless noted) */
enum { ME_PGS_LAM = 0,  /* [n_c,3] impulses (n, t1, t2) */
       ME_PGS_V,        /* [n_b,6] velocity after the iterations so far */
       ME_PGS_VBIAS,    /* [n_b,6] velocity after the biased pass (me_pgs_snapshot) */
       ME_PGS_RHO,      /* [n_c] 1 / lambda_max of the contact's Delassus block */
       ME_PGS_RES,      /* [n_c] residual per contact (me_pgs_residual) */
       ME_PGS_U,        /* [n_c,3] contact velocity (me_pgs_residual) */
       ME_PGS_COLOR,    /* int32 [n_c] colour of each contact */
       ME_PGS_BUF_COUNT_ };

/* Workspace bytes for n_bodies bodies and at most max_contacts contacts. */
int me_pgs_workspace_size(int32_t n_bodies, int32_t max_contacts, int32_t real_bytes, size_t *bytes);
/* NCPSolverGPU.__init__ (ncp_gpu.py:790-840): capacity fixed at construction. */
int me_pgs_create(int32_t n_bodies, int32_t max_contacts, int32_t real_bytes, int32_t flags, void *workspace,
                  size_t bytes, me_pgs **out);
int me_pgs_destroy(me_pgs *s);
/* The set-up block of NCPSolverGPU.solve (ncp_gpu.py:980-998) and _color_cached (:928-947): rho per
 * contact, lambda = warm ([n_c,3] device, NULL = zero) with its warm scatter, v = v_free + M^-1 J^T
 * lambda, and the colouring (reused on a cache hit). soft: 5 host doubles (bias_rate, mass_scale,
 * impulse_scale, inv_h, contact_speed) of SoftStep.params, NULL = the Baumgarte path of `bias`.
 * *cache_hit (may be NULL) receives 1 on a colour-cache hit. The scene arrays are read again by
 * me_pgs_iterate and me_pgs_residual: keep them unchanged until the last call for this scene.
 * Fails with ME_EARG when body_pairs holds an index outside [-1, n_bodies). Synchronises `stream`
 * (the index check, the cache test). */
int me_pgs_prepare(me_pgs *s, const me_pgs_scene *scene, int32_t n_c, const void *warm, const double *soft,
                   int32_t *cache_hit, void *stream);
/* `iters` full sweeps (every colour of every tile), soft-step use_bias flag as given
 * (the graph replays of _iter_launches, ncp_gpu.py:949-962, 1001-1036). One launch. */
int me_pgs_iterate(me_pgs *s, int32_t iters, int32_t use_bias, void *stream);
/* Copy the current velocity into ME_PGS_VBIAS (SolveResult.v_bias: the biased pass's velocity). */
int me_pgs_snapshot(me_pgs *s, void *stream);
/* Natural-map residual max_c ||lambda_c - proj_K(z_c)||_inf of the map the last pass iterated
 * (use_bias as given), per-contact residuals and contact velocities (k_resid, ncp_gpu.py:601-646).
 * Synchronises `stream`. */
int me_pgs_residual(me_pgs *s, int32_t use_bias, double *residual, void *stream);
int me_pgs_buffer(me_pgs *s, int32_t which, void **ptr, int64_t *count);
/* out[ME_PGS_INFO_COUNT_]: colours of the current colouring, cache hits and misses, islands and tiles
 * of the current colouring, overflow (1: more than 1024 colours). Synchronous. */
int me_pgs_info(me_pgs *s, int64_t *out);

/* Bytes of the caller-owned tape block (256-byte aligned device memory) for this engine. */
int me_tape_size(me_engine *e, size_t *bytes);

/* One substep of length dt (as one substep of me_step), recorded into the tape: the centres and
 * orientations at its start, the warm-start matches, and the body state and impulses at the start
 * of every sweep. A tile beyond one small tile sets overflow bit 32 (not recorded). */
int me_step_taped(me_engine *e, double dt, void *tape, size_t bytes, void *stream);

/* Reverse of the recorded substep; the engine must hold the state right after me_step_taped.
 * In: xbar [N,3], qbar [N,4], vbar [N,3], wbar [N,3] = adjoint of the state after the substep;
 *     jbar_end [capC,3] = adjoint of the substep's final (jn, jt1, jt2), the values the next substep
 *     warm-starts from (zero for the last substep).
 * Out: xbar, qbar, vbar, wbar = adjoint of the state before the substep (in place);
 *     jbar_prev [capC,3] = adjoint of the previous substep's final impulses, by its contact index;
 *     mubar [N] += adjoint of the per-body friction coefficients. */
int me_adjoint_substep(me_engine *e, double dt, void *tape, size_t bytes, float *xbar, float *qbar, float *vbar,
                       float *wbar, const float *jbar_end, float *jbar_prev, float *mubar, void *stream);

/* Text of the last error on the calling thread. */
const char *me_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
