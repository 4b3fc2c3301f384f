/* CPU restatement of the reference contact step: TEST INFRASTRUCTURE ONLY.
 *
 * This library is the parity checker for the HIP path. Only tests/, the smoke check and the
 * benchmark's CPU-baseline leg may load it; the product never links or calls it.
 *
 * Every routine follows the reference kernel it restates in operation order, float32 state, the
 * int64 fixed-point scale 1e10 and float64 quantisation. The build keeps the compiler's
 * contraction off (-ffp-contract=off); where the reference kernel is built with FMA contraction,
 * its fused multiply-adds are written out explicitly (RFMA; the rules are in oracle.c's header). Citations are to files of the reference
 * repository antonbj3/3fold-motion-engine (src/motion_engine/...).
 *
 * Parity pins (see tests/test_oracle_golden.py and tests/test_oracle_invariants.py):
 *   - body build, world points, bead contacts, relaxed Jacobi, integrate: the reference's
 *     archived trajectories reports/position_solve_matrix/inputs/cross_hardware_l4.npz
 *     (jacobi stack8 x120, lattice1000 x40), bit for bit, and per-step contact counts;
 *   - box narrow phase (an extension; the reference's bead model cannot express boxes):
 *     closed forms (resting box -> 4 corners at known depth; aligned stacked boxes -> the
 *     overlap rectangle), brute-force SAT on tiny inputs;
 *   - pair build, JP colouring: invariants (proper colouring, sortedness) + brute force;
 *   - bead manifold reduction, pair build, JP colouring (PCG hash), ColorCache and pair-chunk
 *     solvers: the same archive's pair_chunks and color_cache trajectories (stack8 x120,
 *     lattice1000 x40), bit for bit; invariants (proper colouring, sortedness) + brute force;
 *   - coloured chunk solvers on boxes: physics checks (resting force = Mg, stick/slide at atan mu).
 */
#ifndef MOTION_ORACLE_H
#define MOTION_ORACLE_H
#include <stdint.h>

/* State and arithmetic type: float (the restatement of the float32 reference kernels, the parity
 * oracle) or, built with -DOR_REAL_DOUBLE, double -- the same algorithm in float64, the function the
 * adjoint's central finite differences are taken of. Literal constants keep their float32 values in
 * both builds. */
#ifdef OR_REAL_DOUBLE
typedef double real;
#define RSQRT sqrt
#define RFABS fabs
#define RFLOOR floor
#define RFMA fma
#else
typedef float real;
#define RSQRT sqrtf
#define RFABS fabsf
#define RFLOOR floorf
#define RFMA fmaf
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_JACOBI = 0, OR_PAIR_CHUNKS = 1, OR_COLOR_CACHE = 2 };

typedef struct or_params {
    int solver;        /* OR_JACOBI / OR_PAIR_CHUNKS / OR_COLOR_CACHE */
    int vit, pit;      /* velocity / position iterations */
    float relax;      /* Jacobi relaxation (0.2) */
    float t2d;        /* coloured friction t2 factor (1.0) */
    int warm_start;    /* coloured solvers: warm start by feature key */
    float R;          /* bead radius (bead geometry) */
    float margin;     /* box contact margin (0 = penetrating points only) */
    int cap;           /* contact capacity */
    int n_colors;      /* 0 = JP colouring as is; K > 0 = split into max(K, JP colours) classes */
    int per_contact;   /* coloured solver without pair chunks: every contact its own chunk */
    int manifold;      /* bead geometry, coloured solvers: keep <= manifold points per body pair
                        * (E-optimal reduction); 0 = no reduction */
} or_params;

typedef struct or_world or_world;

or_world *or_create(int n_bodies, int n_beads, const or_params *prm);
void or_destroy(or_world *w);
/* SoA body state: x[3N], q[4N] (xyzw), v[3N], w[3N], invM[N], IbInv[9N] row-major, kin[N],
 * half[3N] (box half extents; unused in bead mode), mu[N], scene[N]. */
void or_set_bodies(or_world *w, const real *x, const real *q, const real *v, const real *om,
                   const real *invM, const real *IbInv, const int *kin, const real *half,
                   const real *mu, const int *scene);
void or_set_rest(or_world *w, const real *rest); /* bead geometry: rest[3P] */
void or_set_velocity(or_world *w, const real *v, const real *om);
void or_set_mass(or_world *w, const double *M);      /* fp64 masses for contact forces */
int or_step(or_world *w, double dt, int substeps);   /* 0 ok; 4 capacity; 2 colouring */
void or_get_state(const or_world *w, real *x, real *q, real *v, real *om);
void or_get_aux(const or_world *w, real *pv, real *po, real *invIw);
void or_contact_forces(const or_world *w, double *F);
/* counters: [n_contacts, n_pairs, n_colors, cache_hits, cache_misses, overflow, n_generated]:
 * n_contacts = contacts solved (after a manifold reduction), n_generated = before it */
void or_counters(const or_world *w, int64_t *out);
int or_get_contacts(const or_world *w, int64_t *key, int *bi, int *bj, real *pA, real *pB,
                    real *n, real *pen, real *jn, real *jt1, real *jt2, real *jp);
int or_get_pairs(const or_world *w, int *pbi, int *pbj, int *pstart, int *pcount, int *color);

/* single-stage entry points for unit tests */
void or_quat_to_matrix(const real *q, real *R);
void or_quat_rotate(const real *q, const real *x, real *out);
void or_integrate(int N, real *x, real *q, const real *v, const real *om, const real *pv,
                  const real *po, const real *invM, real dt);
int or_box_box(const real *xa, const real *qa, const real *ha, const real *xb,
               const real *qb, const real *hb, real margin, real *p, real *n, real *pen,
               int *code);
int or_box_ground(const real *x, const real *q, const real *h, real margin, real *p,
                  real *pen, int *code);
uint32_t or_pcg(uint32_t s);
int64_t or_jp_key(int rank);
double or_time_steps(or_world *w, double dt, int substeps, int steps);

#ifdef __cplusplus
}
#endif
#endif
