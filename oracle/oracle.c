/* CPU restatement of the reference contact step: TEST INFRASTRUCTURE ONLY (the parity
 * checker; see oracle.h for what pins each part). Build: oracle/Makefile, -ffp-contract=off.
 *
 * Vector conventions restate Warp's builtins as the reference kernels use them:
 *   dot(a,b)   = a.x*b.x + a.y*b.y + a.z*b.z                  (left to right)
 *   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   length(a)  = sqrt(dot(a,a)); normalize(a) = a / length(a) (0 if length is 0)
 *   M*v        = row i: (M[i][0]*v.x + M[i][1]*v.y) + M[i][2]*v.z
 *   M*N        = element ij: ((0 + M[i][0]*N[0][j]) + M[i][1]*N[1][j]) + M[i][2]*N[2][j]
 *   quat_rotate(q,x) = x*(2w*w-1) + qv*(2*(qv.x)) + cross(qv,x)*w*2
 *   quat_to_matrix(q) = columns quat_rotate(q, e0), quat_rotate(q, e1), quat_rotate(q, e2)
 *   quat product a*b (Hamilton, xyzw storage), quat normalize = q * (1/length(q))
 * Those are the uncontracted forms, used where the reference builds a kernel with fuse_fp=False: the
 * manifold selection (contact_engine_gpu_colored.py:651) and the chunk sweeps of its unfused engines
 * (contact_engine_gpu_unfused_solve.py:223-231, contact_engine_gpu_unfused_position.py). Every other
 * kernel the restatement follows (gravity, world inertia, world points and bead contacts, warm apply,
 * the relaxed Jacobi solve and apply, the integrator) is built with Warp's default contraction, so
 * a product that feeds an add is rounded once, as a fused multiply-add; the c-prefixed forms below
 * write those fusions out (RFMA) as the compiler makes them:
 *   a*b + c*d -> fma(a, b, c*d); s + a*b and a*b + s -> fma(a, b, s); s - a*b -> fma(-a, b, s);
 *   a*b - c*d -> fma(a, b, -(c*d)); a product by the literal 1 folds away before fusion; an add whose
 *   operands are not products (a fused result plus a value) stays an add.
 *   cdot(a,b) = fma(a.z,b.z, fma(a.x,b.x, a.y*b.y)); cmv row i = fma(M[i][2],v.z, fma(M[i][0],v.x, M[i][1]*v.y));
 *   cmm element ij = fma(M[i][2],N[2][j], fma(M[i][1],N[1][j], fma(M[i][0],N[0][j], 0))).
 * Pinned by the reference's archived trajectories (reports/position_solve_matrix/inputs/
 * cross_hardware_l4.npz): with these forms the Jacobi, pair-chunk and ColorCache restatements
 * reproduce every archived state, force and per-step hash bit for bit; with the uncontracted forms
 * in their place they drift by ulps from the first rotating step.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct { real x, y, z; } v3;
typedef struct { real x, y, z, w; } qt;
typedef struct { real m[3][3]; } m33;

static inline v3 V3(real x, real y, real z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline v3 vmul(v3 a, real s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 smul(real s, v3 a) { return V3(s * a.x, s * a.y, s * a.z); }
static inline v3 vdiv(v3 a, real s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline real vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real vlen(v3 a) { return RSQRT(vdot(a, a)); }
/* max(a, b) as a > b ? a : b (explicit, so signed zeros and NaNs behave the same everywhere) */
static inline real wmax(real a, real b) { return a > b ? a : b; }
static inline v3 vnormalize(v3 a) {
    real l = vlen(a);
    if (l > 0.0f) return vdiv(a, l);
    return V3(0.0f, 0.0f, 0.0f);
}
static inline v3 mv(const m33 *M, v3 b) {
    return V3(M->m[0][0] * b.x + M->m[0][1] * b.y + M->m[0][2] * b.z,
              M->m[1][0] * b.x + M->m[1][1] * b.y + M->m[1][2] * b.z,
              M->m[2][0] * b.x + M->m[2][1] * b.y + M->m[2][2] * b.z);
}
static inline m33 mm(const m33 *A, const m33 *B) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = 0.0f;
            for (int k = 0; k < 3; ++k) s += A->m[i][k] * B->m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
static inline m33 mtrans(const m33 *A) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = A->m[j][i];
    return r;
}
static inline v3 ld3(const real *p, int i) { return V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st3(real *p, int i, v3 a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
static inline qt ldq(const real *p, int i) {
    qt r = {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
    return r;
}
static inline m33 ldm(const real *p, int i) {
    m33 r;
    for (int k = 0; k < 9; ++k) r.m[k / 3][k % 3] = p[9 * i + k];
    return r;
}
static inline void stm(real *p, int i, const m33 *M) {
    for (int k = 0; k < 9; ++k) p[9 * i + k] = M->m[k / 3][k % 3];
}

/* Warp quat_rotate: c = 2*w*w - 1; d = 2*(qv.x); x*c + qv*d + cross(qv, x)*w*2 */
static inline v3 quat_rotate(qt q, v3 x) {
    real c = 2.0f * q.w * q.w - 1.0f;
    real d = 2.0f * (q.x * x.x + q.y * x.y + q.z * x.z);
    v3 qv = V3(q.x, q.y, q.z);
    return vadd(vadd(vmul(x, c), vmul(qv, d)), vmul(vmul(vcross(qv, x), q.w), 2.0f));
}
static inline m33 quat_to_matrix(qt q) {
    v3 c1 = quat_rotate(q, V3(1.0f, 0.0f, 0.0f));
    v3 c2 = quat_rotate(q, V3(0.0f, 1.0f, 0.0f));
    v3 c3 = quat_rotate(q, V3(0.0f, 0.0f, 1.0f));
    m33 R;
    R.m[0][0] = c1.x; R.m[1][0] = c1.y; R.m[2][0] = c1.z;
    R.m[0][1] = c2.x; R.m[1][1] = c2.y; R.m[2][1] = c2.z;
    R.m[0][2] = c3.x; R.m[1][2] = c3.y; R.m[2][2] = c3.z;
    return R;
}
static inline qt qmul(qt a, qt b) {
    qt r;
    r.x = a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z;
    r.y = a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x;
    r.z = a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
static inline qt qnormalize(qt q) {
    real l = RSQRT(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    qt r;
    if (l > 0.0f) {
        real inv = 1.0f / l;
        r.x = q.x * inv; r.y = q.y * inv; r.z = q.z * inv; r.w = q.w * inv;
    } else {
        r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 1.0f;
    }
    return r;
}

/* contracted forms (kernels the reference builds with FMA contraction; rules in the file header) */
static inline real cdot(v3 a, v3 b) { return RFMA(a.z, b.z, RFMA(a.x, b.x, a.y * b.y)); }
static inline v3 ccross(v3 a, v3 b) {
    return V3(RFMA(a.y, b.z, -(a.z * b.y)), RFMA(a.z, b.x, -(a.x * b.z)), RFMA(a.x, b.y, -(a.y * b.x)));
}
static inline v3 cmv(const m33 *M, v3 b) {
    return V3(RFMA(M->m[0][2], b.z, RFMA(M->m[0][0], b.x, M->m[0][1] * b.y)),
              RFMA(M->m[1][2], b.z, RFMA(M->m[1][0], b.x, M->m[1][1] * b.y)),
              RFMA(M->m[2][2], b.z, RFMA(M->m[2][0], b.x, M->m[2][1] * b.y)));
}
static inline m33 cmm(const m33 *A, const m33 *B) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            real s = RFMA(A->m[i][0], B->m[0][j], 0.0f);
            s = RFMA(A->m[i][1], B->m[1][j], s);
            r.m[i][j] = RFMA(A->m[i][2], B->m[2][j], s);
        }
    return r;
}
/* s + a*u, per component */
static inline v3 cmadd(real a, v3 u, v3 s) { return V3(RFMA(a, u.x, s.x), RFMA(a, u.y, s.y), RFMA(a, u.z, s.z)); }
static inline v3 cnormalize(v3 a) {
    real l = RSQRT(cdot(a, a));
    if (l > 0.0f) return vdiv(a, l);
    return V3(0.0f, 0.0f, 0.0f);
}
/* normalize(aa - dot(aa, n)*n) with aa = e_x, or e_y when |n.x| >= 0.9 */
static inline v3 cfixed_t1(v3 n) {
    v3 aa = V3(1.0f, 0.0f, 0.0f);
    if (RFABS(n.x) >= 0.9f) aa = V3(0.0f, 1.0f, 0.0f);
    return cnormalize(cmadd(-cdot(aa, n), n, aa));
}
/* quat_rotate: c = fma(2w, w, -1), d = 2*cdot(qv, x); x*c + qv*d + (cross(qv,x)*w)*2 */
static inline v3 cquat_rotate(qt q, v3 x) {
    real c = RFMA(2.0f * q.w, q.w, -1.0f);
    real d = 2.0f * RFMA(q.z, x.z, RFMA(q.x, x.x, q.y * x.y));
    v3 cr = ccross(V3(q.x, q.y, q.z), x);
    return V3(RFMA(cr.x * q.w, 2.0f, RFMA(x.x, c, q.x * d)), RFMA(cr.y * q.w, 2.0f, RFMA(x.y, c, q.y * d)),
              RFMA(cr.z * q.w, 2.0f, RFMA(x.z, c, q.z * d)));
}
/* quat_to_matrix: column k is quat_rotate(q, e_k) with the unit factors folded (d = 2*q_k, the
 * diagonal fma(q_k, d, c)) */
static inline m33 cquat_to_matrix(qt q) {
    real c = RFMA(2.0f * q.w, q.w, -1.0f);
    real qq[3] = {q.x, q.y, q.z};
    m33 R;
    for (int k = 0; k < 3; ++k) {
        real d = 2.0f * qq[k];
        v3 cr = vcross(V3(q.x, q.y, q.z), V3(k == 0, k == 1, k == 2));
        real cj[3] = {cr.x, cr.y, cr.z};
        for (int j = 0; j < 3; ++j) {
            real inner = j == k ? RFMA(qq[j], d, c) : RFMA(qq[j], d, 0.0f * c);
            R.m[j][k] = RFMA(cj[j] * q.w, 2.0f, inner);
        }
    }
    return R;
}
static inline qt cqmul(qt a, qt b) {
    qt r;
    r.x = RFMA(-b.y, a.z, RFMA(a.y, b.z, RFMA(a.w, b.x, b.w * a.x)));
    r.y = RFMA(-b.z, a.x, RFMA(a.z, b.x, RFMA(a.w, b.y, b.w * a.y)));
    r.z = RFMA(-b.x, a.y, RFMA(a.x, b.y, RFMA(a.w, b.z, b.w * a.z)));
    r.w = RFMA(-a.z, b.z, RFMA(-a.y, b.y, RFMA(a.w, b.w, -(a.x * b.x))));
    return r;
}
static inline qt cqnormalize(qt q) {
    real l = RSQRT(RFMA(q.w, q.w, RFMA(q.z, q.z, RFMA(q.x, q.x, q.y * q.y))));
    qt r;
    if (l > 0.0f) {
        real inv = 1.0f / l;
        r.x = q.x * inv; r.y = q.y * inv; r.z = q.z * inv; r.w = q.w * inv;
    } else {
        r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 1.0f;
    }
    return r;
}

/* ───────────────────────── world ───────────────────────── */
struct or_world {
    int N, P;
    or_params prm;
    real *x, *q, *v, *w, *invM, *IbInv, *invIw, *pv, *po, *half, *mu, *rest, *allp, *vsnap;
    real *Rw; /* 9N world rotation */
    real *aabb; /* 6N lo/hi */
    int *kin, *scene;
    double *M, *force;
    double force_dt;
    /* contacts (sorted by feature key) */
    int C, cap;
    int64_t *key;
    int *cbi, *cbj;
    real *cpA, *cpB, *cn, *cpen, *cmu, *jn, *jt1, *jt2, *jp;
    int Craw;  /* contacts generated, before a manifold reduction */
    /* bead manifold stage: float64 warm normal impulse, redistributed one, keep flags, pair order */
    double *jnd, *jnr;
    int *keep, *gord;
    /* pairs: the k-th contact of pair p is pidx[pstart[p] + k] */
    int Pp, pcap;
    int *pbi, *pbj, *pstart, *pcount, *col, *porder, *cstart, *ccount, *pidx;
    int ncol;
    /* warm store */
    int wn;
    int64_t *wkey;
    real *wjn, *wjt1, *wjt2;
    /* colour cache */
    int cached_P, cached_ncol;
    int *cache_a, *cache_b, *cache_col;
    real *cache_invM;
    int64_t hits, misses, overflow;
    /* Jacobi accumulators */
    int64_t *acc;
    /* scratch */
    int64_t *tmpk;
    int *tmpi;
};

static void *xc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

or_world *or_create(int N, int P, const or_params *prm) {
    or_world *w = (or_world *)xc(1, sizeof(or_world));
    w->N = N; w->P = P; w->prm = *prm;
    w->x = xc(3 * N, sizeof(real)); w->q = xc(4 * N, sizeof(real)); w->v = xc(3 * N, sizeof(real)); w->w = xc(3 * N, sizeof(real));
    w->invM = xc(N, sizeof(real)); w->IbInv = xc(9 * N, sizeof(real)); w->invIw = xc(9 * N, sizeof(real)); w->pv = xc(3 * N, sizeof(real));
    w->po = xc(3 * N, sizeof(real)); w->half = xc(3 * N, sizeof(real)); w->mu = xc(N, sizeof(real)); w->vsnap = xc(3 * N, sizeof(real));
    w->Rw = xc(9 * N, sizeof(real)); w->aabb = xc(6 * N, sizeof(real));
    w->kin = xc(N, 4); w->scene = xc(N, 4); w->M = xc(N, 8); w->force = xc(3 * N, 8);
    w->rest = xc(3 * (P > 0 ? P : 1), sizeof(real));
    w->allp = xc(3 * (size_t)N * (P > 0 ? P : 1), sizeof(real));
    w->cap = prm->cap;
    int cap = prm->cap;
    w->key = xc(cap, 8); w->cbi = xc(cap, 4); w->cbj = xc(cap, 4);
    w->cpA = xc(3 * cap, sizeof(real)); w->cpB = xc(3 * cap, sizeof(real)); w->cn = xc(3 * cap, sizeof(real));
    w->cpen = xc(cap, sizeof(real)); w->cmu = xc(cap, sizeof(real)); w->jn = xc(cap, sizeof(real)); w->jt1 = xc(cap, sizeof(real));
    w->jt2 = xc(cap, sizeof(real)); w->jp = xc(cap, sizeof(real));
    w->pcap = cap;
    w->pbi = xc(cap, 4); w->pbj = xc(cap, 4); w->pstart = xc(cap, 4); w->pcount = xc(cap, 4);
    w->col = xc(cap, 4); w->porder = xc(cap, 4); w->cstart = xc(cap + 1, 4); w->ccount = xc(cap + 1, 4);
    w->pidx = xc(cap, 4); w->keep = xc(cap, 4); w->gord = xc(cap, 4);
    w->jnd = xc(cap, 8); w->jnr = xc(cap, 8);
    w->wkey = xc(cap, 8); w->wjn = xc(cap, sizeof(real)); w->wjt1 = xc(cap, sizeof(real)); w->wjt2 = xc(cap, sizeof(real));
    w->cache_a = xc(cap, 4); w->cache_b = xc(cap, 4); w->cache_col = xc(cap, 4);
    w->cache_invM = xc(N, sizeof(real));
    w->cached_P = -1;
    w->acc = xc(6 * (size_t)N, 8);
    w->tmpk = xc(cap > N ? cap : N, 8);
    w->tmpi = xc((cap > N ? cap : N) + 1, 4);
    w->force_dt = 0.0;
    return w;
}

void or_destroy(or_world *w) {
    if (!w) return;
    void *ptrs[] = {w->x, w->q, w->v, w->w, w->invM, w->IbInv, w->invIw, w->pv, w->po, w->half,
                    w->mu, w->rest, w->allp, w->vsnap, w->Rw, w->aabb, w->kin, w->scene, w->M,
                    w->force, w->key, w->cbi, w->cbj, w->cpA, w->cpB, w->cn, w->cpen, w->cmu,
                    w->jn, w->jt1, w->jt2, w->jp, w->pbi, w->pbj, w->pstart, w->pcount, w->col,
                    w->porder, w->cstart, w->ccount, w->wkey, w->wjn, w->wjt1, w->wjt2,
                    w->cache_a, w->cache_b, w->cache_col, w->cache_invM, w->acc, w->tmpk, w->tmpi,
                    w->pidx, w->keep, w->gord, w->jnd, w->jnr};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(w);
}

void or_set_bodies(or_world *w, const real *x, const real *q, const real *v, const real *om,
                   const real *invM, const real *IbInv, const int *kin, const real *half,
                   const real *mu, const int *scene) {
    int N = w->N;
    memcpy(w->x, x, 3 * sizeof(real) * (size_t)N); memcpy(w->q, q, 4 * sizeof(real) * (size_t)N);
    memcpy(w->v, v, 3 * sizeof(real) * (size_t)N); memcpy(w->w, om, 3 * sizeof(real) * (size_t)N);
    memcpy(w->invM, invM, sizeof(real) * (size_t)N); memcpy(w->IbInv, IbInv, 9 * sizeof(real) * (size_t)N);
    memcpy(w->kin, kin, 4 * (size_t)N);
    if (half) memcpy(w->half, half, 3 * sizeof(real) * (size_t)N);
    if (mu) memcpy(w->mu, mu, sizeof(real) * (size_t)N);
    if (scene) memcpy(w->scene, scene, 4 * (size_t)N);
    for (int i = 0; i < N; ++i) w->M[i] = invM[i] > 0.0f ? 1.0 / (double)invM[i] : INFINITY;
}
void or_set_mass(or_world *w, const double *M) { memcpy(w->M, M, 8 * (size_t)w->N); }
void or_set_rest(or_world *w, const real *rest) { memcpy(w->rest, rest, 3 * sizeof(real) * (size_t)w->P); }
void or_set_velocity(or_world *w, const real *v, const real *om) {
    memcpy(w->v, v, 3 * sizeof(real) * (size_t)w->N); memcpy(w->w, om, 3 * sizeof(real) * (size_t)w->N);
}

/* ── A2 k_grav (contact_engine_gpu.py:41-44) ── */
static void grav(or_world *w, real dt) {
    for (int i = 0; i < w->N; ++i)
        if (w->invM[i] > 0.0f) {
            st3(w->v, i, cmadd(dt, V3(0.0f, 0.0f, -9.81f), ld3(w->v, i)));
        }
}

/* ── A3 k_invIw (contact_engine_gpu.py:47-53): invIw = (R*IbInv)*R^T, zero if kinematic ── */
static void inv_inertia_world(or_world *w) {
    for (int i = 0; i < w->N; ++i) {
        m33 R = cquat_to_matrix(ldq(w->q, i));
        stm(w->Rw, i, &R);
        if (w->kin[i] == 1) {
            m33 z; memset(&z, 0, sizeof z);
            stm(w->invIw, i, &z);
        } else {
            m33 Ib = ldm(w->IbInv, i);
            m33 t = cmm(&R, &Ib);
            m33 Rt = mtrans(&R);
            m33 r = cmm(&t, &Rt);
            stm(w->invIw, i, &r);
        }
    }
}

/* ── A11 k_integ (contact_engine_gpu.py:172-180) ── */
static void integrate_range(int N, real *x, real *q, const real *v, const real *om,
                            const real *pv, const real *po, const real *invM, real dt) {
    for (int i = 0; i < N; ++i) {
        if (invM[i] <= 0.0f) continue;
        v3 s = vadd(ld3(v, i), ld3(pv, i)), xi = ld3(x, i);
        st3(x, i, V3(RFMA(s.x, dt, xi.x), RFMA(s.y, dt, xi.y), RFMA(s.z, dt, xi.z)));
        v3 ww = vadd(ld3(om, i), ld3(po, i));
        qt h = {0.5f * ww.x, 0.5f * ww.y, 0.5f * ww.z, 0.5f * 0.0f};
        qt qi = ldq(q, i);
        qt m = cqmul(h, qi);
#ifdef OR_PLANT_UNFUSED_INTEG
        qt qn = {qi.x + m.x * dt, qi.y + m.y * dt, qi.z + m.z * dt, qi.w + m.w * dt};
#else
        qt qn = {RFMA(m.x, dt, qi.x), RFMA(m.y, dt, qi.y), RFMA(m.z, dt, qi.z), RFMA(m.w, dt, qi.w)};
#endif
        qt r = cqnormalize(qn);
        q[4 * i] = r.x; q[4 * i + 1] = r.y; q[4 * i + 2] = r.z; q[4 * i + 3] = r.w;
    }
}
void or_integrate(int N, real *x, real *q, const real *v, const real *om, const real *pv,
                  const real *po, const real *invM, real dt) {
    integrate_range(N, x, q, v, om, pv, po, invM, dt);
}
void or_quat_to_matrix(const real *q, real *R) {
    m33 M = quat_to_matrix(ldq(q, 0));
    stm(R, 0, &M);
}
void or_quat_rotate(const real *q, const real *x, real *out) {
    st3(out, 0, quat_rotate(ldq(q, 0), ld3(x, 0)));
}

/* ───────────────────────── contact generation ───────────────────────── */
static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : x > y;
}

static int push_contact(or_world *w, int64_t key, int bi, int bj, v3 pA, v3 pB, v3 n, real pen,
                        real mu) {
    if (w->C >= w->cap) { w->overflow++; return 0; }
    int c = w->C++;
    w->key[c] = key; w->cbi[c] = bi; w->cbj[c] = bj;
    st3(w->cpA, c, pA); st3(w->cpB, c, pB); st3(w->cn, c, n);
    w->cpen[c] = pen; w->cmu[c] = mu;
    return 1;
}

/* ── A4 k_worldp (contact_engine_gpu.py:56-59) + A5 k_gen_fid (contact_engine_gpu_colored.py:103-122)
 * The contact SET is what the reference's hash-grid query yields after its exact tests; the list is
 * emitted in feature-key order fa*(N*P+1) + fb + 1 (contact_engine_gpu_colored.py:903-917), which is
 * the order the reference's stable feature sort produces. */
typedef struct { int64_t cell; int idx; } cellrec;
static int cmp_cell(const void *a, const void *b) {
    const cellrec *x = a, *y = b;
    if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
    return x->idx - y->idx;
}
static int64_t cell_key(int64_t cx, int64_t cy, int64_t cz, int64_t s) {
    const int64_t B = 1 << 20;
    return (((s * B + (cx + B / 2)) * B + (cy + B / 2)) * B + (cz + B / 2));
}
static int find_cell(const cellrec *cr, int n, int64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cr[mid].cell < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void bead_contacts(or_world *w) {
    int N = w->N, P = w->P, NP = N * P;
    real R = w->prm.R;
    real mu = w->mu[0];
    for (int g = 0; g < NP; ++g) {
        int b = g / P;
        qt q = ldq(w->q, b);
        v3 p = vadd(ld3(w->x, b), cquat_rotate(q, ld3(w->rest, g % P)));
        st3(w->allp, g, p);
    }
    real h = 2.0f * R;
    cellrec *cr = (cellrec *)malloc(sizeof(cellrec) * (NP > 0 ? (size_t)NP : (size_t)1));
    for (int g = 0; g < NP; ++g) {
        v3 p = ld3(w->allp, g);
        cr[g].cell = cell_key((int64_t)RFLOOR(p.x / h), (int64_t)RFLOOR(p.y / h), (int64_t)RFLOOR(p.z / h), 0);
        cr[g].idx = g;
    }
    qsort(cr, NP, sizeof(cellrec), cmp_cell);
    int *cand = (int *)malloc(sizeof(int) * 4096);
    w->C = 0;
    int64_t stride = (int64_t)NP + 1;
    for (int g = 0; g < NP; ++g) {
        v3 xi = ld3(w->allp, g);
        int bi = g / P;
        real s = xi.z - R;
        if (s < 0.0f)
            push_contact(w, (int64_t)g * stride + 0, bi, -1, xi, V3(0.0f, 0.0f, 0.0f),
                         V3(0.0f, 0.0f, 1.0f), -s, mu);
        int64_t cx = (int64_t)RFLOOR(xi.x / h), cy = (int64_t)RFLOOR(xi.y / h), cz = (int64_t)RFLOOR(xi.z / h);
        int nc = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int64_t k = cell_key(cx + dx, cy + dy, cz + dz, 0);
                    for (int t = find_cell(cr, NP, k); t < NP && cr[t].cell == k; ++t) {
                        int j = cr[t].idx;
                        if (j > g && j / P != bi && nc < 4096) cand[nc++] = j;
                    }
                }
        /* ascending j = ascending feature key */
        for (int a = 1; a < nc; ++a) {
            int t = cand[a], b2 = a - 1;
            while (b2 >= 0 && cand[b2] > t) { cand[b2 + 1] = cand[b2]; --b2; }
            cand[b2 + 1] = t;
        }
        for (int a = 0; a < nc; ++a) {
            int j = cand[a];
            v3 xj = ld3(w->allp, j);
            v3 dv = vsub(xi, xj);
            real dist = RSQRT(cdot(dv, dv));
            if (dist < 2.0f * R && dist > 1e-9f)
                push_contact(w, (int64_t)g * stride + (int64_t)j + 1, bi, j / P, xi, xj,
                             vdiv(dv, dist), 2.0f * R - dist, mu);
        }
    }
    free(cand);
    free(cr);
}

/* ── box geometry (extension: per-body extents, solid-box inertia, <=4 points per pair) ──
 * Follows the reference's box precedent ncp/ncp_ref.py:1065-1211 (face-axis SAT, reference-face
 * clipping, deepest-first farthest-point reduction to 4), in float32, with the 9 edge-edge axes
 * used as separation tests only, and feature codes per clipped vertex for the warm start. */
static inline v3 box_vertex(v3 c, const m33 *R, v3 h, int k) {
    v3 l = V3((k & 1) ? h.x : -h.x, (k & 2) ? h.y : -h.y, (k & 4) ? h.z : -h.z);
    return vadd(c, mv(R, l));
}
static inline v3 mcol(const m33 *R, int i) { return V3(R->m[0][i], R->m[1][i], R->m[2][i]); }
static inline real hget(v3 h, int i) { return i == 0 ? h.x : (i == 1 ? h.y : h.z); }

/* face (axis a, sign s) polygon, ncp_ref._face_polygon order: (-,-), (-,+), (+,+), (+,-) */
static void face_poly(v3 c, const m33 *R, v3 h, int a, real s, v3 out[4]) {
    int o0 = a == 0 ? 1 : 0, o1 = a == 2 ? 1 : 2;
    static const real S0[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    static const real S1[4] = {-1.0f, 1.0f, 1.0f, -1.0f};
    for (int k = 0; k < 4; ++k) {
        real l[3] = {0, 0, 0};  /* a, o0, o1 are a permutation of 0, 1, 2 */
        l[a] = s * hget(h, a);
        l[o0] = S0[k] * hget(h, o0);
        l[o1] = S1[k] * hget(h, o1);
        out[k] = vadd(c, mv(R, V3(l[0], l[1], l[2])));
    }
}

#define CLIP_EPS 1e-6f
#define MAXPOLY 16
typedef struct { v3 p[MAXPOLY]; int code[MAXPOLY]; int edge[MAXPOLY]; int n; } poly_t;

static void clip_poly(const poly_t *in, v3 sn, real d, int plane, poly_t *out) {
    out->n = 0;
    int m = in->n;
    for (int i = 0; i < m; ++i) {
        v3 a = in->p[i], b = in->p[(i + 1) % m];
        real da = vdot(sn, a) - d;
        real db = vdot(sn, b) - d;
        int E = in->edge[i];
        if (da <= CLIP_EPS && out->n < MAXPOLY) {
            out->p[out->n] = a;
            out->code[out->n] = in->code[i];
            out->edge[out->n] = (db > CLIP_EPS && !(da < -CLIP_EPS)) ? 4 + plane : E;
            out->n++;
        }
        int enter = da > CLIP_EPS && db < -CLIP_EPS;
        int exitp = da < -CLIP_EPS && db > CLIP_EPS;
        if ((enter || exitp) && out->n < MAXPOLY) {
            real t = da / (da - db);
            out->p[out->n] = vadd(a, smul(t, vsub(b, a)));
            out->code[out->n] = E < 4 ? 4 + 4 * E + plane : 20 + 4 * (E - 4) + plane;
            out->edge[out->n] = exitp ? 4 + plane : E;
            out->n++;
        }
    }
}

/* deepest first (stable by gap), then greedily the candidate farthest (squared distance) from
 * those kept; returns kept indices ordered by ascending feature code.
 * OR_PLANT_REDUCE_UNSORTED / OR_PLANT_REDUCE_NEAREST exist so that a test can build the oracle with a
 * planted mistake (start from the first point instead of the deepest; take the nearest candidate
 * instead of the farthest) and show that its closed-form case rejects it. */
static int reduce4(int n, const v3 *p, const real *gap, const int *code, int *keep) {
    int ord[MAXPOLY];
    for (int i = 0; i < n; ++i) ord[i] = i;
#ifndef OR_PLANT_REDUCE_UNSORTED
    for (int a = 1; a < n; ++a) {
        int t = ord[a], b = a - 1;
        while (b >= 0 && gap[ord[b]] > gap[t]) { ord[b + 1] = ord[b]; --b; }
        ord[b + 1] = t;
    }
#endif
    int nk;
    if (n <= 4) {
        for (int i = 0; i < n; ++i) keep[i] = ord[i];
        nk = n;
    } else {
        int used[MAXPOLY] = {0};
        keep[0] = ord[0]; used[0] = 1; nk = 1;
        while (nk < 4) {
            int best = -1; real bd = -1.0f;
            for (int r = 1; r < n; ++r) {
                if (used[r]) continue;
                real md = 3.4e38f;
                for (int k = 0; k < nk; ++k) {
                    v3 e = vsub(p[ord[r]], p[keep[k]]);
                    real d2 = vdot(e, e);
                    if (d2 < md) md = d2;
                }
#ifdef OR_PLANT_REDUCE_NEAREST
                if (best < 0 || md < bd) { bd = md; best = r; }
#else
                if (md > bd) { bd = md; best = r; }
#endif
            }
            used[best] = 1; keep[nk++] = ord[best];
        }
    }
    for (int a = 1; a < nk; ++a) {
        int t = keep[a], b = a - 1;
        while (b >= 0 && code[keep[b]] > code[t]) { keep[b + 1] = keep[b]; --b; }
        keep[b + 1] = t;
    }
    return nk;
}

/* box vs ground plane z = 0, normal +z (ground -> box); up to 4 vertices with z < margin */
static int box_ground_impl(v3 c, const m33 *R, v3 h, real margin, v3 *pout, real *pen, int *code) {
    v3 p[8]; real gap[8]; int cd[8]; int n = 0;
    for (int k = 0; k < 8; ++k) {
        v3 vk = box_vertex(c, R, h, k);
        if (vk.z < margin) { p[n] = vk; gap[n] = vk.z; cd[n] = k; n++; }
    }
    int keep[MAXPOLY];
    int nk = reduce4(n, p, gap, cd, keep);
    for (int i = 0; i < nk; ++i) { pout[i] = p[keep[i]]; pen[i] = -gap[keep[i]]; code[i] = cd[keep[i]]; }
    return nk;
}

static inline real sat_overlap(v3 ax, const m33 *RA, v3 hA, const m33 *RB, v3 hB, v3 d) {
    real ra = RFABS(vdot(ax, mcol(RA, 0))) * hA.x + RFABS(vdot(ax, mcol(RA, 1))) * hA.y +
               RFABS(vdot(ax, mcol(RA, 2))) * hA.z;
    real rb = RFABS(vdot(ax, mcol(RB, 0))) * hB.x + RFABS(vdot(ax, mcol(RB, 1))) * hB.y +
               RFABS(vdot(ax, mcol(RB, 2))) * hB.z;
    return ra + rb - RFABS(vdot(ax, d));
}

/* box A vs box B; normal n points from B to A; code = feature code (13 bits) */
static int box_box_impl(v3 cA, const m33 *RA, v3 hA, v3 cB, const m33 *RB, v3 hB, real margin,
                        v3 *pout, v3 *nout, real *pen, int *code) {
    v3 d = vsub(cA, cB);
    real best = 3.4e38f; int bfrom = -1;
    for (int i = 0; i < 6; ++i) {
        v3 ax = i < 3 ? mcol(RA, i) : mcol(RB, i - 3);
        real o = sat_overlap(ax, RA, hA, RB, hB, d);
        if (o < best) { best = o; bfrom = i; }
    }
    if (bfrom < 0 || best < -margin) return 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            v3 L = vcross(mcol(RA, i), mcol(RB, j));
            real l2 = vdot(L, L);
            if (l2 < 1e-12f) continue;
            v3 ax = vdiv(L, RSQRT(l2));
            if (sat_overlap(ax, RA, hA, RB, hB, d) < -margin) return 0;
        }
    v3 n = bfrom < 3 ? mcol(RA, bfrom) : mcol(RB, bfrom - 3);
    if (vdot(n, d) < 0.0f) n = vneg(n);
    int refIsB = bfrom >= 3, ai = refIsB ? bfrom - 3 : bfrom;
    v3 refc, incc, refh, inch; const m33 *refR, *incR; v3 ref_n; real ref_sign;
    if (!refIsB) {
        refc = cA; refR = RA; refh = hA; incc = cB; incR = RB; inch = hB;
        ref_n = vneg(n);
        ref_sign = vdot(mcol(RA, ai), n) > 0.0f ? -1.0f : 1.0f;
    } else {
        refc = cB; refR = RB; refh = hB; incc = cA; incR = RA; inch = hA;
        ref_n = n;
        ref_sign = vdot(mcol(RB, ai), n) > 0.0f ? 1.0f : -1.0f;
    }
    int bi = 0; real bs = 1.0f, bv = 3.4e38f;
    for (int i = 0; i < 3; ++i)
        for (int si = 0; si < 2; ++si) {
            real s = si == 0 ? -1.0f : 1.0f;
            real val = vdot(vmul(mcol(incR, i), s), ref_n);
            if (val < bv) { bv = val; bi = i; bs = s; }
        }
    poly_t A, B;
    face_poly(incc, incR, inch, bi, bs, A.p);
    A.n = 4;
    for (int k = 0; k < 4; ++k) { A.code[k] = k; A.edge[k] = k; }
    v3 rp[4];
    face_poly(refc, refR, refh, ai, ref_sign, rp);
    poly_t *cur = &A, *nxt = &B;
    for (int i = 0; i < 4; ++i) {
        v3 e = vsub(rp[(i + 1) % 4], rp[i]);
        v3 sn = vcross(e, ref_n);
        real nl = vlen(sn);
        if (nl < 1e-12f) continue;
        sn = vdiv(sn, nl);
        if (vdot(sn, vsub(refc, rp[i])) > 0.0f) sn = vneg(sn);
        clip_poly(cur, sn, vdot(sn, rp[i]), i, nxt);
        poly_t *t = cur; cur = nxt; nxt = t;
        if (cur->n == 0) return 0;
    }
    /* drop coincident points, keep penetrating ones */
    v3 p[MAXPOLY]; real gap[MAXPOLY]; int cd[MAXPOLY]; int n2 = 0;
    for (int i = 0; i < cur->n; ++i) {
        int dup = 0;
        for (int k = 0; k < i; ++k) {
            v3 e = vsub(cur->p[i], cur->p[k]);
            if (vdot(e, e) <= 1e-12f) { dup = 1; break; }
        }
        if (dup) continue;
        real g = vdot(ref_n, vsub(cur->p[i], rp[0]));
        if (g < margin) {
            p[n2] = vsub(cur->p[i], vmul(ref_n, 0.5f * g));
            gap[n2] = g; cd[n2] = cur->code[i]; n2++;
        }
    }
    int keep[MAXPOLY];
    int nk = reduce4(n2, p, gap, cd, keep);
    int face = (refIsB << 11) | ((ai * 2 + (ref_sign > 0.0f)) << 8) | ((bi * 2 + (bs > 0.0f)) << 5);
    for (int i = 0; i < nk; ++i) {
        pout[i] = p[keep[i]]; nout[i] = n; pen[i] = -gap[keep[i]];
        code[i] = face | cd[keep[i]];
    }
    /* codes inside one face pair are distinct; the list is ordered by code */
    return nk;
}

int or_box_box(const real *xa, const real *qa, const real *ha, const real *xb, const real *qb,
               const real *hb, real margin, real *p, real *n, real *pen, int *code) {
    m33 RA = quat_to_matrix(ldq(qa, 0)), RB = quat_to_matrix(ldq(qb, 0));
    v3 pp[4], nn[4];
    int k = box_box_impl(ld3(xa, 0), &RA, ld3(ha, 0), ld3(xb, 0), &RB, ld3(hb, 0), margin, pp, nn, pen, code);
    for (int i = 0; i < k; ++i) { st3(p, i, pp[i]); st3(n, i, nn[i]); }
    return k;
}
int or_box_ground(const real *x, const real *q, const real *h, real margin, real *p, real *pen,
                  int *code) {
    m33 R = quat_to_matrix(ldq(q, 0));
    v3 pp[4];
    int k = box_ground_impl(ld3(x, 0), &R, ld3(h, 0), margin, pp, pen, code);
    for (int i = 0; i < k; ++i) st3(p, i, pp[i]);
    return k;
}

/* friction of a body pair: the geometric mean of the two bodies' coefficients (equal ones kept as
 * they are); the ground takes the body's own. OR_PLANT_MU_MEAN (the arithmetic mean) exists so that a
 * test can show that the slide distance v0^2 / (2 mu g) rejects it. */
static inline real pair_mu(const or_world *w, int a, int b) {
    if (b < 0) return w->mu[a];
    real ma = w->mu[a], mb = w->mu[b];
#ifdef OR_PLANT_MU_MEAN
    return 0.5f * (ma + mb);
#else
    return ma == mb ? ma : RSQRT(ma * mb);
#endif
}

/* body AABB half extents ext_k = sum_j |R_kj| h_j */
static void box_aabbs(or_world *w) {
    for (int i = 0; i < w->N; ++i) {
        m33 R = ldm(w->Rw, i);
        v3 h = ld3(w->half, i), c = ld3(w->x, i);
        real e[3];
        for (int k = 0; k < 3; ++k)
            e[k] = RFABS(R.m[k][0]) * h.x + RFABS(R.m[k][1]) * h.y + RFABS(R.m[k][2]) * h.z;
        w->aabb[6 * i + 0] = c.x - e[0]; w->aabb[6 * i + 1] = c.y - e[1]; w->aabb[6 * i + 2] = c.z - e[2];
        w->aabb[6 * i + 3] = c.x + e[0]; w->aabb[6 * i + 4] = c.y + e[1]; w->aabb[6 * i + 5] = c.z + e[2];
    }
}
/* broad-phase slack: candidate tests are conservative by this much beyond the contact margin */
#define BP_SLACK 1e-4f
static inline int aabb_overlap(const real *a, const real *b, real m) {
    return a[0] <= b[3] + m && b[0] <= a[3] + m && a[1] <= b[4] + m && b[1] <= a[4] + m &&
           a[2] <= b[5] + m && b[2] <= a[5] + m;
}

/* A6 body pairs (contact_body_pairs_gpu.py:7-66, 90-107): candidate pairs are AABB overlaps within
 * one scene plus ground candidates, keyed a*(N+1)+b+1 (ground b=-1), emitted in key order; the
 * narrow phase then writes each pair's contacts in feature-code order. */
static void box_contacts(or_world *w) {
    int N = w->N;
    real margin = w->prm.margin;
    box_aabbs(w);
    real hmax = 0.0f;
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) {
            real e = w->aabb[6 * i + 3 + k] - w->x[3 * i + k];
            if (e > hmax) hmax = e;
        }
    real cell = (2.0f * hmax + margin + BP_SLACK) * 1.001f;
    if (!(cell > 0.0f)) cell = 1.0f;
    cellrec *cr = (cellrec *)malloc(sizeof(cellrec) * (size_t)(N ? N : 1));
    for (int i = 0; i < N; ++i) {
        cr[i].cell = cell_key((int64_t)RFLOOR(w->x[3 * i] / cell), (int64_t)RFLOOR(w->x[3 * i + 1] / cell),
                              (int64_t)RFLOOR(w->x[3 * i + 2] / cell), w->scene[i]);
        cr[i].idx = i;
    }
    qsort(cr, N, sizeof(cellrec), cmp_cell);
    int *cand = (int *)malloc(sizeof(int) * 4096);
    w->C = 0;
    int64_t N1 = (int64_t)N + 1;
    for (int a = 0; a < N; ++a) {
        m33 RA = ldm(w->Rw, a);
        v3 cA = ld3(w->x, a), hA = ld3(w->half, a);
        v3 p[4], nn[4]; real pen[4]; int code[4];
        if (w->invM[a] > 0.0f && w->aabb[6 * a + 2] < margin + BP_SLACK) {
            int k = box_ground_impl(cA, &RA, hA, margin, p, pen, code);
            int64_t pk = (int64_t)a * N1 + 0;
            for (int i = 0; i < k; ++i)
                push_contact(w, pk * 8192 + code[i], a, -1, p[i], p[i], V3(0.0f, 0.0f, 1.0f), pen[i],
                             pair_mu(w, a, -1));
        }
        int64_t cx = (int64_t)RFLOOR(cA.x / cell), cy = (int64_t)RFLOOR(cA.y / cell), cz = (int64_t)RFLOOR(cA.z / cell);
        int nc = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int64_t k = cell_key(cx + dx, cy + dy, cz + dz, w->scene[a]);
                    for (int t = find_cell(cr, N, k); t < N && cr[t].cell == k; ++t) {
                        int b = cr[t].idx;
                        if (b > a && w->scene[b] == w->scene[a] &&
                            aabb_overlap(w->aabb + 6 * a, w->aabb + 6 * b, margin + BP_SLACK) && nc < 4096)
                            cand[nc++] = b;
                    }
                }
        for (int i = 1; i < nc; ++i) {
            int t = cand[i], j = i - 1;
            while (j >= 0 && cand[j] > t) { cand[j + 1] = cand[j]; --j; }
            cand[j + 1] = t;
        }
        for (int i = 0; i < nc; ++i) {
            int b = cand[i];
            if (w->invM[a] <= 0.0f && w->invM[b] <= 0.0f) continue;
            m33 RB = ldm(w->Rw, b);
            int k = box_box_impl(cA, &RA, hA, ld3(w->x, b), &RB, ld3(w->half, b), margin, p, nn, pen, code);
            int64_t pk = (int64_t)a * N1 + (int64_t)b + 1;
            for (int j = 0; j < k; ++j)
                push_contact(w, pk * 8192 + code[j], a, b, p[j], p[j], nn[j], pen[j], pair_mu(w, a, b));
        }
    }
    free(cand);
    free(cr);
}

/* ───────────────────────── A8-A10 relaxed Jacobi (contact_engine_gpu.py:32-38, 82-169) ───────────────────────── */
static inline void acc3(int64_t *acc, int N, int b, int off, v3 d, real sign, real scale) {
    double sc = (double)(sign * scale);
    acc[(off + 0) * (size_t)N + b] += (int64_t)round((double)d.x * sc);
    acc[(off + 1) * (size_t)N + b] += (int64_t)round((double)d.y * sc);
    acc[(off + 2) * (size_t)N + b] += (int64_t)round((double)d.z * sc);
}

static void jac_vel(or_world *w, int c, real scale) {
    int N = w->N;
    int bi = w->cbi[c], bj = w->cbj[c];
    v3 nrm = ld3(w->cn, c), rA = vsub(ld3(w->cpA, c), ld3(w->x, bi));
    real iMA = w->invM[bi];
    m33 IA = ldm(w->invIw, bi), IB;
    v3 va = vadd(ld3(w->v, bi), ccross(ld3(w->w, bi), rA));
    v3 vb = V3(0.0f, 0.0f, 0.0f), rB = V3(0.0f, 0.0f, 0.0f);
    real iMB = 0.0f;
    if (bj >= 0) {
        rB = vsub(ld3(w->cpB, c), ld3(w->x, bj));
        vb = vadd(ld3(w->v, bj), ccross(ld3(w->w, bj), rB));
        iMB = w->invM[bj];
        IB = ldm(w->invIw, bj);
    }
    v3 vrel = vsub(va, vb);
    real vn = cdot(vrel, nrm);
    v3 vt = cmadd(-vn, nrm, vrel);
    real mt = RSQRT(cdot(vt, vt));
    v3 t1 = vdiv(vt, mt);
    if (mt <= 5e-3f) {
        t1 = cfixed_t1(nrm);
    }
    v3 t2 = ccross(nrm, t1);
    v3 rnA = ccross(rA, nrm);
    real meff = iMA + cdot(rnA, cmv(&IA, rnA));
    if (bj >= 0) {
        v3 rnB = ccross(rB, nrm);
        meff = meff + iMB + cdot(rnB, cmv(&IB, rnB));
    }
    if (meff < 1e-12f) return;
    real dj = -vn / meff;
    real nw = wmax(0.0f, w->jn[c] + dj);
    dj = nw - w->jn[c];
    w->jn[c] = nw;
    v3 Jn = vmul(nrm, dj);
    acc3(w->acc, N, bi, 0, vmul(Jn, iMA), 1.0f, scale);
    acc3(w->acc, N, bi, 3, cmv(&IA, ccross(rA, Jn)), 1.0f, scale);
    if (bj >= 0) {
        acc3(w->acc, N, bj, 0, vmul(Jn, iMB), -1.0f, scale);
        acc3(w->acc, N, bj, 3, cmv(&IB, ccross(rB, Jn)), -1.0f, scale);
    }
    v3 rtA1 = ccross(rA, t1);
    real meft1 = iMA + cdot(rtA1, cmv(&IA, rtA1));
    v3 rtA2 = ccross(rA, t2);
    real meft2 = iMA + cdot(rtA2, cmv(&IA, rtA2));
    if (bj >= 0) {
        v3 rtB1 = ccross(rB, t1);
        meft1 = meft1 + iMB + cdot(rtB1, cmv(&IB, rtB1));
        v3 rtB2 = ccross(rB, t2);
        meft2 = meft2 + iMB + cdot(rtB2, cmv(&IB, rtB2));
    }
    real o1 = w->jt1[c], o2 = w->jt2[c];
    real a1 = o1 - cdot(vrel, t1) / meft1;
    real a2 = o2 - cdot(vrel, t2) / meft2;
    real mag = RSQRT(RFMA(a1, a1, a2 * a2));
    real lim = w->cmu[c] * w->jn[c];
    if (mag > lim && mag > 1e-12f) {
        a1 = a1 * lim / mag;
        a2 = a2 * lim / mag;
    }
    w->jt1[c] = a1; w->jt2[c] = a2;
    v3 Jt = cmadd(a1 - o1, t1, vmul(t2, 0.5f * (a2 - o2)));
    acc3(w->acc, N, bi, 0, vmul(Jt, iMA), 1.0f, scale);
    acc3(w->acc, N, bi, 3, cmv(&IA, ccross(rA, Jt)), 1.0f, scale);
    if (bj >= 0) {
        acc3(w->acc, N, bj, 0, vmul(Jt, iMB), -1.0f, scale);
        acc3(w->acc, N, bj, 3, cmv(&IB, ccross(rB, Jt)), -1.0f, scale);
    }
}

static void jac_pos(or_world *w, int c, real dt, real scale) {
    int N = w->N;
    int bi = w->cbi[c], bj = w->cbj[c];
    v3 nrm = ld3(w->cn, c), rA = vsub(ld3(w->cpA, c), ld3(w->x, bi));
    real iMA = w->invM[bi];
    m33 IA = ldm(w->invIw, bi), IB;
    v3 rel = vadd(ld3(w->pv, bi), ccross(ld3(w->po, bi), rA));
    v3 rB = V3(0.0f, 0.0f, 0.0f);
    real iMB = 0.0f;
    if (bj >= 0) {
        rB = vsub(ld3(w->cpB, c), ld3(w->x, bj));
        rel = vsub(rel, vadd(ld3(w->pv, bj), ccross(ld3(w->po, bj), rB)));
        iMB = w->invM[bj];
        IB = ldm(w->invIw, bj);
    }
    v3 rnA = ccross(rA, nrm);
    real meff = iMA + cdot(rnA, cmv(&IA, rnA));
    if (bj >= 0) {
        v3 rnB = ccross(rB, nrm);
        meff = meff + iMB + cdot(rnB, cmv(&IB, rnB));
    }
    if (meff < 1e-12f) return;
    real bias = 0.2f * wmax(w->cpen[c] - 1e-4f, 0.0f) / dt;
    real dj = (bias - cdot(rel, nrm)) / meff;
    real nw = wmax(0.0f, w->jp[c] + dj);
    dj = nw - w->jp[c];
    w->jp[c] = nw;
    v3 Jn = vmul(nrm, dj);
    acc3(w->acc, N, bi, 0, vmul(Jn, iMA), 1.0f, scale);
    acc3(w->acc, N, bi, 3, cmv(&IA, ccross(rA, Jn)), 1.0f, scale);
    if (bj >= 0) {
        acc3(w->acc, N, bj, 0, vmul(Jn, iMB), -1.0f, scale);
        acc3(w->acc, N, bj, 3, cmv(&IB, ccross(rB, Jn)), -1.0f, scale);
    }
}

/* k_apply (contact_engine_gpu.py:134-143): dv = f32(f64(acc) * f64(f32(1/scale))), v = fma(relax, dv, v) */
static void jac_apply(or_world *w, real *vv, real *ww, real relax, real scale) {
    int N = w->N;
    real inv = 1.0f / scale;
    for (int i = 0; i < N; ++i) {
        v3 dv = V3((real)((double)w->acc[0 * (size_t)N + i] * (double)inv),
                   (real)((double)w->acc[1 * (size_t)N + i] * (double)inv),
                   (real)((double)w->acc[2 * (size_t)N + i] * (double)inv));
        v3 dw = V3((real)((double)w->acc[3 * (size_t)N + i] * (double)inv),
                   (real)((double)w->acc[4 * (size_t)N + i] * (double)inv),
                   (real)((double)w->acc[5 * (size_t)N + i] * (double)inv));
        st3(vv, i, cmadd(relax, dv, ld3(vv, i)));
        st3(ww, i, cmadd(relax, dw, ld3(ww, i)));
        for (int k = 0; k < 6; ++k) w->acc[k * (size_t)N + i] = 0;
    }
}

/* ───────────────────────── pairs, warm start, colouring ───────────────────────── */
/* A16 pair build (contact_engine_gpu_colored.py:342-389; device form :1138-1157): a stable sort of the
 * contacts by the pair key a*(N+1) + b + 1, then one pair per run of equal keys. Box contacts are
 * generated pair-major, so the sort is the identity there; bead contacts are in feature-key order. */
static void build_pairs(or_world *w) {
    w->Pp = 0;
    if (w->prm.per_contact) { /* contact_engine_gpu_colored.py:1105-1136: the contact graph */
        for (int c = 0; c < w->C; ++c) {
            w->pstart[c] = c; w->pcount[c] = 1; w->pbi[c] = w->cbi[c]; w->pbj[c] = w->cbj[c]; w->pidx[c] = c;
        }
        w->Pp = w->C;
        return;
    }
    int C = w->C;
    if (w->P > 0) {
        cellrec *r = (cellrec *)malloc(sizeof(cellrec) * (C > 0 ? (size_t)C : (size_t)1));
        for (int c = 0; c < C; ++c) {
            r[c].cell = (int64_t)w->cbi[c] * (int64_t)(w->N + 1) + (int64_t)w->cbj[c] + 1;
            r[c].idx = c;
        }
        qsort(r, C, sizeof(cellrec), cmp_cell); /* (key, index): the stable order */
        for (int c = 0; c < C; ++c) w->pidx[c] = r[c].idx;
        free(r);
    } else {
        for (int c = 0; c < C; ++c) w->pidx[c] = c;
    }
    for (int t = 0; t < C; ++t) {
        int c = w->pidx[t];
        if (t == 0 || w->cbi[c] != w->cbi[w->pidx[t - 1]] || w->cbj[c] != w->cbj[w->pidx[t - 1]]) {
            int p = w->Pp++;
            w->pstart[p] = t; w->pcount[p] = 0; w->pbi[p] = w->cbi[c]; w->pbj[p] = w->cbj[c];
        }
        w->pcount[w->Pp - 1]++;
    }
}

/* A14 k_warm (contact_engine_gpu_colored.py:602-626): binary search in the previous kept keys */
static void warm_lookup(or_world *w) {
    for (int c = 0; c < w->C; ++c) {
        w->jn[c] = 0.0f; w->jt1[c] = 0.0f; w->jt2[c] = 0.0f; w->jp[c] = 0.0f;
        if (!w->prm.warm_start || w->wn == 0) continue;
        int64_t k = w->key[c];
        int lo = 0, hi = w->wn - 1, hit = -1;
        while (lo <= hi) {
            int mid = (lo + hi) / 2;
            int64_t km = w->wkey[mid];
            if (km == k) { hit = mid; lo = hi + 1; }
            else if (km < k) lo = mid + 1;
            else hi = mid - 1;
        }
        if (hit >= 0) { w->jn[c] = w->wjn[hit]; w->jt1[c] = w->wjt1[hit]; w->jt2[c] = w->wjt2[hit]; }
    }
}
/* A23 warm store (contact_engine_gpu_colored.py:1089-1093) */
static void warm_store(or_world *w) {
    w->wn = w->C;
    memcpy(w->wkey, w->key, 8 * (size_t)w->C);
    memcpy(w->wjn, w->jn, sizeof(real) * (size_t)w->C);
    memcpy(w->wjt1, w->jt1, sizeof(real) * (size_t)w->C);
    memcpy(w->wjt2, w->jt2, sizeof(real) * (size_t)w->C);
}

/* ── A15 manifold reduction of bead contacts, the device stage (contact_engine_gpu_colored.py:
 * 1050-1087 with the kernels k_fkey / k_warm / k_pkey_pair / k_runflag / k_select / k_compact at
 * :587-740): warm start of every generated contact by binary search of its feature key in the
 * previous step's kept keys (normal impulse promoted to float64), grouping by body pair (stable sort
 * of the pair key), and per pair with more than mx contacts the E-optimal choice of mx of them; the
 * normal impulse of every contact of the pair goes to the kept point nearest to it in the contact
 * plane, summed in float64. The kept contacts stay in feature-key order. */

/* planar coordinates about ref in the (u, w) basis: float32 difference, then float64
 * (contact_engine_gpu_colored.py:585-590) */
static inline void planar2(v3 p, v3 ref, const double u[3], const double wv[3], double r[2]) {
    real dx = p.x - ref.x, dy = p.y - ref.y, dz = p.z - ref.z;
    double ax = (double)dx, ay = (double)dy, az = (double)dz;
    r[0] = ax * u[0] + ay * u[1] + az * u[2];
    r[1] = ax * wv[0] + ay * wv[1] + az * wv[2];
}

/* k_select for one pair (contact_engine_gpu_colored.py:647-720); g[0..L) = the pair's contacts in
 * feature order. Returns 0, or 1 when the segment exceeds the kernel's fixed sizes. */
static int select_segment(or_world *w, const int *g, int L, int mx) {
    if (L <= mx) {
        for (int t = 0; t < L; ++t) { w->keep[g[t]] = 1; w->jnr[g[t]] = w->jnd[g[t]]; }
        return 0;
    }
    if (L > 1024 || mx > 8) {
        for (int t = 0; t < L; ++t) { w->keep[g[t]] = 1; w->jnr[g[t]] = w->jnd[g[t]]; }
        return 1;
    }
    int deep = 0;
    real dbest = w->cpen[g[0]];
    for (int t = 1; t < L; ++t)
#ifdef OR_PLANT_DEEP_LAST /* planted mistake for a test: last index wins ties */
        if (w->cpen[g[t]] >= dbest) { dbest = w->cpen[g[t]]; deep = t; }
#else
        if (w->cpen[g[t]] > dbest) { dbest = w->cpen[g[t]]; deep = t; }
#endif
    int gd = g[deep];
    v3 nr = ld3(w->cn, gd);
    double ax = 1.0, ay = 0.0, az = 0.0;
    if (RFABS(nr.x) >= 0.9f) { ax = 0.0; ay = 1.0; az = 0.0; }
    double n0 = (double)nr.x, n1 = (double)nr.y, n2 = (double)nr.z;
    double dn = ax * n0 + ay * n1 + az * n2;
    double u0 = ax - dn * n0, u1 = ay - dn * n1, u2 = az - dn * n2;
    double ln = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    double u[3] = {u0 / ln, u1 / ln, u2 / ln};
    double wv[3] = {n1 * u[2] - n2 * u[1], n2 * u[0] - n0 * u[2], n0 * u[1] - n1 * u[0]};
    v3 ref = ld3(w->x, w->cbi[gd]);
    double rd[2];
    planar2(ld3(w->cpA, gd), ref, u, wv, rd);
    double S00 = rd[0] * rd[0], S01 = rd[0] * rd[1], S10 = rd[1] * rd[0], S11 = rd[1] * rd[1];
    int pick[8];
    pick[0] = deep;
    for (int k = 1; k < mx; ++k) {
        double sbest = -1.0e300;
        int tbest = -1;
        double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
        for (int t = 0; t < L; ++t) {
            int taken = 0;
            for (int q = 0; q < k; ++q)
                if (pick[q] == t) taken = 1;
            if (taken) continue;
            double rt[2];
            planar2(ld3(w->cpA, g[t]), ref, u, wv, rt);
            double a00 = S00 + rt[0] * rt[0], a01 = S01 + rt[0] * rt[1];
            double a10 = S10 + rt[1] * rt[0], a11 = S11 + rt[1] * rt[1];
            double tr = a00 + a11;
            double det = a00 * a11 - a01 * a10;
            double disc = tr * tr / 4.0 - det;
            if (!(disc > 0.0)) disc = 0.0; /* max(., 0) */
            double lam = tr / 2.0 - sqrt(disc);
            double kk = (double)(k + 1);
            double sc = lam < kk ? lam : kk; /* min(k + 1, lambda_min) */
            if (sc > sbest) { sbest = sc; tbest = t; c00 = a00; c01 = a01; c10 = a10; c11 = a11; }
        }
        if (tbest < 0) return 1;
        pick[k] = tbest;
        S00 = c00; S01 = c01; S10 = c10; S11 = c11;
    }
    for (int k = 0; k < mx; ++k) { w->keep[g[pick[k]]] = 1; w->jnr[g[pick[k]]] = 0.0; }
    for (int t = 0; t < L; ++t) {
        double rt[2];
        planar2(ld3(w->cpA, g[t]), ref, u, wv, rt);
        double dbest2 = 1.0e300;
        int kbest = 0;
        for (int k = 0; k < mx; ++k) {
            double rk[2];
            planar2(ld3(w->cpA, g[pick[k]]), ref, u, wv, rk);
            double e0 = rt[0] - rk[0], e1 = rt[1] - rk[1];
            double d2 = e0 * e0 + e1 * e1;
            if (d2 < dbest2) { dbest2 = d2; kbest = k; }
        }
        int gk = g[pick[kbest]];
        w->jnr[gk] = w->jnr[gk] + w->jnd[g[t]];
    }
    return 0;
}

static int bead_manifold(or_world *w) {
    int C = w->C, mx = w->prm.manifold;
    for (int c = 0; c < C; ++c) {
        w->jnd[c] = 0.0; w->jt1[c] = 0.0f; w->jt2[c] = 0.0f; w->keep[c] = 0;
        if (!w->prm.warm_start || w->wn == 0) continue;
        int64_t k = w->key[c];
        int lo = 0, hi = w->wn - 1, hit = -1;
        while (lo <= hi) {
            int mid = (lo + hi) / 2;
            int64_t km = w->wkey[mid];
            if (km == k) { hit = mid; lo = hi + 1; }
            else if (km < k) lo = mid + 1;
            else hi = mid - 1;
        }
        if (hit >= 0) { w->jnd[c] = (double)w->wjn[hit]; w->jt1[c] = w->wjt1[hit]; w->jt2[c] = w->wjt2[hit]; }
    }
    cellrec *r = (cellrec *)malloc(sizeof(cellrec) * (C > 0 ? (size_t)C : (size_t)1));
    for (int c = 0; c < C; ++c) {
        r[c].cell = (int64_t)w->cbi[c] * (int64_t)(w->N + 1) + (int64_t)w->cbj[c] + 1;
        r[c].idx = c;
    }
    qsort(r, C, sizeof(cellrec), cmp_cell);
    for (int c = 0; c < C; ++c) w->gord[c] = r[c].idx;
    int bad = 0;
    for (int i = 0; i < C;) {
        int L = 1;
        while (i + L < C && r[i + L].cell == r[i].cell) ++L;
        bad += select_segment(w, w->gord + i, L, mx);
        i += L;
    }
    free(r);
    if (bad) return 5;
    int n = 0;
    for (int c = 0; c < C; ++c) {
        if (!w->keep[c]) continue;
        w->key[n] = w->key[c]; w->cbi[n] = w->cbi[c]; w->cbj[n] = w->cbj[c];
        st3(w->cpA, n, ld3(w->cpA, c)); st3(w->cpB, n, ld3(w->cpB, c)); st3(w->cn, n, ld3(w->cn, c));
        w->cpen[n] = w->cpen[c]; w->cmu[n] = w->cmu[c];
        w->jn[n] = (real)w->jnr[c]; w->jt1[n] = w->jt1[c]; w->jt2[n] = w->jt2[c]; w->jp[n] = 0.0f;
        ++n;
    }
    w->C = n;
    return 0;
}

/* PCG hash (Jarzynski & Olano 2020, "Hash Functions for GPU Rendering"), the generator behind the
 * reference's rand_init(seed, offset) / randi (contact_engine_gpu_colored.py:145). */
uint32_t or_pcg(uint32_t s) {
    uint32_t b = s * 747796405u + 2891336453u;
    uint32_t c = ((b >> ((b >> 28u) + 4u)) ^ b) * 277803737u;
    return (c >> 22u) ^ c;
}
/* k_prio (contact_engine_gpu_colored.py:137-148): |randi(rand_init(7919, c))| * 2^32 + c.
 * OR_JP_SEED exists so that a test can build the oracle with a planted wrong seed and show that
 * the reference's archived trajectories reject it. */
#ifndef OR_JP_SEED
#define OR_JP_SEED 7919
#endif
int64_t or_jp_key(int rank) {
    uint32_t st = or_pcg((uint32_t)OR_JP_SEED + or_pcg((uint32_t)rank));
    st = or_pcg(st);
    int64_t p = (int64_t)(int32_t)st;
    if (p < 0) p = -p;
    return p * (int64_t)4294967296LL + (int64_t)rank;
}

/* A17 JP colouring of the pair graph (k_claim / k_win, contact_engine_gpu_colored.py:151-185,
 * rounds at :1178-1208); then, for n_colors = K > 0, every JP class c is split into sub[c]
 * sub-colours (pair rank mod sub[c]) so that there are max(K, JP colours) classes: the extra
 * classes go one at a time to the class with the most pairs per sub-colour (lowest class first). */
static int color_pairs(or_world *w) {
    int P = w->Pp, N = w->N;
    int64_t *claim = w->tmpk;
    for (int p = 0; p < P; ++p) w->col[p] = -1;
    int cur = 0;
    for (;;) {
        for (int i = 0; i < N; ++i) claim[i] = INT64_MAX;
        for (int p = 0; p < P; ++p) {
            if (w->col[p] >= 0) continue;
            int64_t k = or_jp_key(p);
            int a = w->pbi[p], b = w->pbj[p];
            if (w->invM[a] > 0.0f && k < claim[a]) claim[a] = k;
            if (b >= 0 && w->invM[b] > 0.0f && k < claim[b]) claim[b] = k;
        }
        int left = 0;
        for (int p = 0; p < P; ++p) {
            if (w->col[p] >= 0) continue;
            int64_t k = or_jp_key(p);
            int a = w->pbi[p], b = w->pbj[p], ok = 1;
            if (w->invM[a] > 0.0f && claim[a] != k) ok = 0;
            if (b >= 0 && w->invM[b] > 0.0f && claim[b] != k) ok = 0;
            if (ok) w->col[p] = cur; else left++;
        }
        cur++;
        if (left == 0) break;
        if (cur > P + 1) return -1;
    }
    if (P == 0) cur = 0;
    int K = w->prm.n_colors;
    if (K > cur && cur > 0) {
        int *cnt = (int *)calloc((size_t)cur, sizeof(int)), *sub = (int *)calloc((size_t)cur, sizeof(int));
        int *base = (int *)calloc((size_t)cur + 1, sizeof(int));
        for (int p = 0; p < P; ++p) cnt[w->col[p]]++;
        for (int c = 0; c < cur; ++c) sub[c] = 1;
        for (int e = cur; e < K; ++e) {
            /* class with the largest cnt/sub (compare cnt_a*sub_b > cnt_b*sub_a), lowest first */
            int bc = 0;
            for (int c = 1; c < cur; ++c)
#ifdef OR_PLANT_SPLIT_BY_COUNT /* planted mistake for a test: the largest class, whatever its sub */
                if (cnt[c] > cnt[bc]) bc = c;
#else
                if ((int64_t)cnt[c] * sub[bc] > (int64_t)cnt[bc] * sub[c]) bc = c;
#endif
            sub[bc]++;
        }
        for (int c = 0; c < cur; ++c) base[c + 1] = base[c] + sub[c];
        for (int p = 0; p < P; ++p) {
            int c = w->col[p];
            w->col[p] = base[c] + p % sub[c];
        }
        cur = base[cur];
        free(cnt); free(sub); free(base);
    }
    return cur;
}

static void group_by_color(or_world *w) {
    int P = w->Pp, nc = w->ncol;
    memset(w->ccount, 0, sizeof(int) * (size_t)(nc + 1));
    for (int p = 0; p < P; ++p) w->ccount[w->col[p]]++;
    w->cstart[0] = 0;
    for (int c = 0; c < nc; ++c) w->cstart[c + 1] = w->cstart[c] + w->ccount[c];
    int *fill = w->tmpi;
    for (int c = 0; c < nc; ++c) fill[c] = w->cstart[c];
    for (int p = 0; p < P; ++p) w->porder[fill[w->col[p]]++] = p;
}

/* ColorCache (contact_engine_gpu_color_cache.py:7-18, 37-54): reuse when the pair list and invM are
 * identical to the cached ones */
static int color_cached(or_world *w) {
    int P = w->Pp;
    if (P == w->cached_P) {
        int mismatch = 0;
        for (int p = 0; p < P && !mismatch; ++p)
            if (w->pbi[p] != w->cache_a[p] || w->pbj[p] != w->cache_b[p]) mismatch = 1;
        for (int i = 0; i < w->N && !mismatch; ++i)
            if (w->invM[i] != w->cache_invM[i]) mismatch = 1;
        if (!mismatch) {
            w->hits++;
            memcpy(w->col, w->cache_col, sizeof(int) * (size_t)P);
            return w->cached_ncol;
        }
    }
    int nc = color_pairs(w);
    if (nc < 0) return nc;
    memcpy(w->cache_a, w->pbi, sizeof(int) * (size_t)P);
    memcpy(w->cache_b, w->pbj, sizeof(int) * (size_t)P);
    memcpy(w->cache_col, w->col, sizeof(int) * (size_t)P);
    memcpy(w->cache_invM, w->invM, sizeof(real) * (size_t)w->N);
    w->cached_P = P; w->cached_ncol = nc;
    w->misses++;
    return nc;
}

/* ───────────────────────── A18-A20 coloured chunk sweeps ───────────────────────── */
static inline v3 fixed_t1(v3 nrm) {
    v3 aa = V3(1.0f, 0.0f, 0.0f);
    if (RFABS(nrm.x) >= 0.9f) aa = V3(0.0f, 1.0f, 0.0f);
    return vnormalize(vsub(aa, vmul(nrm, vdot(aa, nrm))));
}

/* apply_warm (contact_engine_gpu_applied_warm.py:7-54) */
static void apply_warm_pair(or_world *w, int p) {
    int a = w->pbi[p], b = w->pbj[p];
    v3 va = ld3(w->v, a), wa = ld3(w->w, a);
    v3 vb = V3(0.0f, 0.0f, 0.0f), wb = V3(0.0f, 0.0f, 0.0f);
    if (b >= 0) { vb = ld3(w->v, b); wb = ld3(w->w, b); }
    m33 IA = ldm(w->invIw, a), IB;
    if (b >= 0) IB = ldm(w->invIw, b);
    for (int k = 0; k < w->pcount[p]; ++k) {
        int c = w->pidx[w->pstart[p] + k];
        v3 n = ld3(w->cn, c);
        v3 t1 = cfixed_t1(n);
        v3 t2 = ccross(n, t1);
        v3 imp = cmadd(w->jt2[c], t2, cmadd(w->jn[c], n, smul(w->jt1[c], t1)));
        if (w->invM[a] > 0.0f) {
            va = cmadd(w->invM[a], imp, va);
            wa = vadd(wa, cmv(&IA, ccross(vsub(ld3(w->cpA, c), ld3(w->x, a)), imp)));
        }
        if (b >= 0 && w->invM[b] > 0.0f) {
            vb = cmadd(-w->invM[b], imp, vb);
            wb = vsub(wb, cmv(&IB, ccross(vsub(ld3(w->cpB, c), ld3(w->x, b)), imp)));
        }
    }
    if (w->invM[a] > 0.0f) { st3(w->v, a, va); st3(w->w, a, wa); }
    if (b >= 0 && w->invM[b] > 0.0f) { st3(w->v, b, vb); st3(w->w, b, wb); }
}

/* chunk velocity sweep: fixed = 1 -> k_fixed_frame (contact_engine_gpu_fixed_frame.py:7-77, the
 * noncontracting form contact_engine_gpu_unfused_solve.py:117-220); fixed = 0 -> k_chunk_vel_r
 * (contact_engine_gpu_colored.py:413-490, slip-direction frame) */
static void chunk_vel(or_world *w, int p, int fixed, real t2d) {
    int bi = w->pbi[p], bj = w->pbj[p];
    real iMA = w->invM[bi];
    m33 IA = ldm(w->invIw, bi), IB = ldm(w->invIw, bi);
    v3 vA = ld3(w->v, bi), wA = ld3(w->w, bi);
    real iMB = 0.0f;
    v3 vB = V3(0.0f, 0.0f, 0.0f), wB = V3(0.0f, 0.0f, 0.0f);
    if (bj >= 0) { iMB = w->invM[bj]; IB = ldm(w->invIw, bj); vB = ld3(w->v, bj); wB = ld3(w->w, bj); }
    for (int k = 0; k < w->pcount[p]; ++k) {
        int c = w->pidx[w->pstart[p] + k];
        v3 nrm = ld3(w->cn, c), rA = vsub(ld3(w->cpA, c), ld3(w->x, bi));
        v3 va = vadd(vA, vcross(wA, rA));
        v3 vb = V3(0.0f, 0.0f, 0.0f), rB = V3(0.0f, 0.0f, 0.0f);
        if (bj >= 0) { rB = vsub(ld3(w->cpB, c), ld3(w->x, bj)); vb = vadd(vB, vcross(wB, rB)); }
        v3 vrel = vsub(va, vb);
        real vn = vdot(vrel, nrm);
        v3 t1;
        if (fixed) {
            t1 = fixed_t1(nrm);
        } else {
            v3 vt = vsub(vrel, vmul(nrm, vn));
            real mt = vlen(vt);
            t1 = vdiv(vt, mt);
            if (mt <= 5e-3f) t1 = fixed_t1(nrm);
        }
        v3 t2 = vcross(nrm, t1);
        v3 rnA = vcross(rA, nrm);
        real meff = iMA + vdot(rnA, mv(&IA, rnA));
        if (bj >= 0) {
            v3 rnB = vcross(rB, nrm);
            meff = meff + iMB + vdot(rnB, mv(&IB, rnB));
        }
        if (!(meff >= 1e-12f)) continue;
        real dj = -vn / meff;
        real nw = wmax(0.0f, w->jn[c] + dj);
        dj = nw - w->jn[c];
        w->jn[c] = nw;
        v3 Jn = vmul(nrm, dj);
        if (iMA > 0.0f) { vA = vadd(vA, vmul(Jn, iMA)); wA = vadd(wA, mv(&IA, vcross(rA, Jn))); }
        if (bj >= 0 && iMB > 0.0f) { vB = vsub(vB, vmul(Jn, iMB)); wB = vsub(wB, mv(&IB, vcross(rB, Jn))); }
        va = vadd(vA, vcross(wA, rA));
        vb = V3(0.0f, 0.0f, 0.0f);
        if (bj >= 0) vb = vadd(vB, vcross(wB, rB));
        vrel = vsub(va, vb);
        v3 rtA1 = vcross(rA, t1);
        real meft1 = iMA + vdot(rtA1, mv(&IA, rtA1));
        v3 rtA2 = vcross(rA, t2);
        real meft2 = iMA + vdot(rtA2, mv(&IA, rtA2));
        if (bj >= 0) {
            v3 rtB1 = vcross(rB, t1);
            meft1 = meft1 + iMB + vdot(rtB1, mv(&IB, rtB1));
            v3 rtB2 = vcross(rB, t2);
            meft2 = meft2 + iMB + vdot(rtB2, mv(&IB, rtB2));
        }
        if (meft1 >= 1e-12f && meft2 >= 1e-12f) {
            real o1 = w->jt1[c], o2 = w->jt2[c];
            real a1 = o1 - vdot(vrel, t1) / meft1;
            real a2 = o2 - vdot(vrel, t2) / meft2;
            real mag = RSQRT(a1 * a1 + a2 * a2);
            real lim = w->cmu[c] * w->jn[c];
            if (mag > lim && mag > 1e-12f) { a1 = a1 * lim / mag; a2 = a2 * lim / mag; }
            w->jt1[c] = a1; w->jt2[c] = a2;
            v3 Jt = vadd(vmul(t1, a1 - o1), vmul(t2, t2d * (a2 - o2)));
            if (iMA > 0.0f) { vA = vadd(vA, vmul(Jt, iMA)); wA = vadd(wA, mv(&IA, vcross(rA, Jt))); }
            if (bj >= 0 && iMB > 0.0f) { vB = vsub(vB, vmul(Jt, iMB)); wB = vsub(wB, mv(&IB, vcross(rB, Jt))); }
        }
    }
    if (iMA > 0.0f) { st3(w->v, bi, vA); st3(w->w, bi, wA); }
    if (bj >= 0 && iMB > 0.0f) { st3(w->v, bj, vB); st3(w->w, bj, wB); }
}

/* k_chunk_pos_r (contact_engine_gpu_colored.py:511-554; unfused_position.py:7-65) */
static void chunk_pos(or_world *w, int p, real dt) {
    int bi = w->pbi[p], bj = w->pbj[p];
    real iMA = w->invM[bi];
    m33 IA = ldm(w->invIw, bi), IB = ldm(w->invIw, bi);
    v3 pvA = ld3(w->pv, bi), poA = ld3(w->po, bi);
    real iMB = 0.0f;
    v3 pvB = V3(0.0f, 0.0f, 0.0f), poB = V3(0.0f, 0.0f, 0.0f);
    if (bj >= 0) { iMB = w->invM[bj]; IB = ldm(w->invIw, bj); pvB = ld3(w->pv, bj); poB = ld3(w->po, bj); }
    for (int k = 0; k < w->pcount[p]; ++k) {
        int c = w->pidx[w->pstart[p] + k];
        v3 nrm = ld3(w->cn, c), rA = vsub(ld3(w->cpA, c), ld3(w->x, bi));
        v3 rel = vadd(pvA, vcross(poA, rA));
        v3 rB = V3(0.0f, 0.0f, 0.0f);
        if (bj >= 0) { rB = vsub(ld3(w->cpB, c), ld3(w->x, bj)); rel = vsub(rel, vadd(pvB, vcross(poB, rB))); }
        v3 rnA = vcross(rA, nrm);
        real meff = iMA + vdot(rnA, mv(&IA, rnA));
        if (bj >= 0) {
            v3 rnB = vcross(rB, nrm);
            meff = meff + iMB + vdot(rnB, mv(&IB, rnB));
        }
        if (!(meff >= 1e-12f)) continue;
        real bias = 0.2f * wmax(w->cpen[c] - 1e-4f, 0.0f) / dt;
        real dj = (bias - vdot(rel, nrm)) / meff;
        real nw = wmax(0.0f, w->jp[c] + dj);
        dj = nw - w->jp[c];
        w->jp[c] = nw;
        v3 Jn = vmul(nrm, dj);
        if (iMA > 0.0f) { pvA = vadd(pvA, vmul(Jn, iMA)); poA = vadd(poA, mv(&IA, vcross(rA, Jn))); }
        if (bj >= 0 && iMB > 0.0f) { pvB = vsub(pvB, vmul(Jn, iMB)); poB = vsub(poB, mv(&IB, vcross(rB, Jn))); }
    }
    if (iMA > 0.0f) { st3(w->pv, bi, pvA); st3(w->po, bi, poA); }
    if (bj >= 0 && iMB > 0.0f) { st3(w->pv, bj, pvB); st3(w->po, bj, poB); }
}

/* ───────────────────────── step drivers ───────────────────────── */
static void generate(or_world *w) {
    if (w->P > 0) bead_contacts(w); else box_contacts(w);
}

static void forces_from(or_world *w, const real *before, double sdt) {
    for (int i = 0; i < w->N; ++i)
        for (int k = 0; k < 3; ++k) {
            if (w->kin[i] != 0) { w->force[3 * i + k] = 0.0; continue; }
            real dv = w->v[3 * i + k] - before[3 * i + k];
            w->force[3 * i + k] = w->M[i] * (double)dv / sdt;
        }
}

/* A12 Jacobi step (contact_engine_gpu.py:252-280) */
static int step_jacobi(or_world *w, double sdt64) {
    real sdt = (real)sdt64;
    const real S = 1.0e10f;
    grav(w, sdt);
    inv_inertia_world(w);
    generate(w);
    w->Craw = w->C;
    memcpy(w->vsnap, w->v, 3 * sizeof(real) * (size_t)w->N);
    int C = w->C;
    if (C > 0) {
        memset(w->jn, 0, sizeof(real) * (size_t)C); memset(w->jt1, 0, sizeof(real) * (size_t)C);
        memset(w->jt2, 0, sizeof(real) * (size_t)C); memset(w->jp, 0, sizeof(real) * (size_t)C);
        memset(w->pv, 0, 3 * sizeof(real) * (size_t)w->N); memset(w->po, 0, 3 * sizeof(real) * (size_t)w->N);
        memset(w->acc, 0, 48 * (size_t)w->N);
        for (int it = 0; it < w->prm.vit; ++it) {
            for (int c = 0; c < C; ++c) jac_vel(w, c, S);
            jac_apply(w, w->v, w->w, w->prm.relax, S);
        }
        for (int it = 0; it < w->prm.pit; ++it) {
            for (int c = 0; c < C; ++c) jac_pos(w, c, sdt, S);
            jac_apply(w, w->pv, w->po, w->prm.relax, S);
        }
    }
    integrate_range(w->N, w->x, w->q, w->v, w->w, w->pv, w->po, w->invM, sdt);
    forces_from(w, w->vsnap, sdt64);
    w->force_dt = sdt64;
    return 0;
}

/* coloured pair-chunk step (contact_engine_gpu_colored.py:1302-1375, pair_chunks path) and the
 * ColorCache step (contact_engine_gpu_deferred_force.py:13-64 with the colour cache) */
static int step_colored(or_world *w, double sdt64, int colorcache) {
    real sdt = (real)sdt64;
    grav(w, sdt);
    inv_inertia_world(w);
    generate(w);
    memcpy(w->vsnap, w->v, 3 * sizeof(real) * (size_t)w->N);
    int C = w->C;
    w->Craw = C;
    memset(w->pv, 0, 3 * sizeof(real) * (size_t)w->N); memset(w->po, 0, 3 * sizeof(real) * (size_t)w->N);
    if (C > 0) {
        if (w->P > 0 && w->prm.manifold > 0) {
            int rc = bead_manifold(w);
            if (rc) return rc;
        } else {
            warm_lookup(w);
        }
        build_pairs(w);
        int nc = colorcache ? color_cached(w) : color_pairs(w);
        if (nc < 0) return 2;
        w->ncol = nc;
        group_by_color(w);
        /* the pairs of one colour share no dynamic body: the OpenMP build (timing only) runs a
           colour's pairs in parallel, with the same result as the serial loop */
        if (colorcache)
            for (int ci = 0; ci < nc; ++ci) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
                for (int t = w->cstart[ci]; t < w->cstart[ci + 1]; ++t) apply_warm_pair(w, w->porder[t]);
            }
        for (int it = 0; it < w->prm.vit; ++it)
            for (int ci = 0; ci < nc; ++ci) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
                for (int t = w->cstart[ci]; t < w->cstart[ci + 1]; ++t)
                    chunk_vel(w, w->porder[t], colorcache, colorcache ? 1.0f : w->prm.t2d);
            }
        for (int it = 0; it < w->prm.pit; ++it)
            for (int ci = 0; ci < nc; ++ci) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
                for (int t = w->cstart[ci]; t < w->cstart[ci + 1]; ++t) chunk_pos(w, w->porder[t], sdt);
            }
        if (w->prm.warm_start) warm_store(w);
    } else {
        w->Pp = 0; w->ncol = 0; w->wn = 0;
    }
    integrate_range(w->N, w->x, w->q, w->v, w->w, w->pv, w->po, w->invM, sdt);
    forces_from(w, w->vsnap, sdt64);
    w->force_dt = sdt64;
    return 0;
}

int or_step(or_world *w, double dt, int substeps) {
    double sdt = dt / substeps;
    for (int s = 0; s < substeps; ++s) {
        int rc = w->prm.solver == OR_JACOBI ? step_jacobi(w, sdt)
                                            : step_colored(w, sdt, w->prm.solver == OR_COLOR_CACHE);
        if (rc) return rc;
        if (w->overflow) return 4;
    }
    return 0;
}

double or_time_steps(or_world *w, double dt, int substeps, int steps) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int s = 0; s < steps; ++s)
        if (or_step(w, dt, substeps)) return -1.0;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

void or_get_state(const or_world *w, real *x, real *q, real *v, real *om) {
    int N = w->N;
    if (x) memcpy(x, w->x, 3 * sizeof(real) * (size_t)N);
    if (q) memcpy(q, w->q, 4 * sizeof(real) * (size_t)N);
    if (v) memcpy(v, w->v, 3 * sizeof(real) * (size_t)N);
    if (om) memcpy(om, w->w, 3 * sizeof(real) * (size_t)N);
}
void or_get_aux(const or_world *w, real *pv, real *po, real *invIw) {
    int N = w->N;
    if (pv) memcpy(pv, w->pv, 3 * sizeof(real) * (size_t)N);
    if (po) memcpy(po, w->po, 3 * sizeof(real) * (size_t)N);
    if (invIw) memcpy(invIw, w->invIw, 9 * sizeof(real) * (size_t)N);
}
void or_contact_forces(const or_world *w, double *F) { memcpy(F, w->force, 24 * (size_t)w->N); }
void or_counters(const or_world *w, int64_t *out) {
    out[0] = w->C; out[1] = w->Pp; out[2] = w->ncol; out[3] = w->hits; out[4] = w->misses;
    out[5] = w->overflow; out[6] = w->Craw;
}
int or_get_contacts(const or_world *w, int64_t *key, int *bi, int *bj, real *pA, real *pB, real *n,
                    real *pen, real *jn, real *jt1, real *jt2, real *jp) {
    int C = w->C;
    if (key) memcpy(key, w->key, 8 * (size_t)C);
    if (bi) memcpy(bi, w->cbi, 4 * (size_t)C);
    if (bj) memcpy(bj, w->cbj, 4 * (size_t)C);
    if (pA) memcpy(pA, w->cpA, 3 * sizeof(real) * (size_t)C);
    if (pB) memcpy(pB, w->cpB, 3 * sizeof(real) * (size_t)C);
    if (n) memcpy(n, w->cn, 3 * sizeof(real) * (size_t)C);
    if (pen) memcpy(pen, w->cpen, sizeof(real) * (size_t)C);
    if (jn) memcpy(jn, w->jn, sizeof(real) * (size_t)C);
    if (jt1) memcpy(jt1, w->jt1, sizeof(real) * (size_t)C);
    if (jt2) memcpy(jt2, w->jt2, sizeof(real) * (size_t)C);
    if (jp) memcpy(jp, w->jp, sizeof(real) * (size_t)C);
    return C;
}
int or_get_pairs(const or_world *w, int *pbi, int *pbj, int *pstart, int *pcount, int *color) {
    int P = w->Pp;
    if (pbi) memcpy(pbi, w->pbi, 4 * (size_t)P);
    if (pbj) memcpy(pbj, w->pbj, 4 * (size_t)P);
    if (pstart) memcpy(pstart, w->pstart, 4 * (size_t)P);
    if (pcount) memcpy(pcount, w->pcount, 4 * (size_t)P);
    if (color) memcpy(color, w->col, 4 * (size_t)P);
    return P;
}
