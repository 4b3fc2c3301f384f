/* CPU oracle of the exact-cone NCP projected Gauss-Seidel (test infrastructure only: the tests and the
 * CPU baseline call it; the product never does). A restatement of the reference's NCPSolverGPU
 * (src/motion_engine/ncp/ncp_gpu.py) in its own global structure: one colour at a time over ALL
 * contacts, sweep then apply (k_sweep :491-566, k_apply_c :585-599), global Jones-Plassmann rounds
 * (_color :873-926), the warm scatter to every endpoint (k_scatter :459-474). The device build cuts
 * the same work into islands and tiles; the tests hold the two to bit identity.
 *
 * Compiled twice (Makefile): float (real = float, symbols *_f32) and double (*_f64). Contraction off.
 * acos and cos are the series below (the device file uses the same series), not libm's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef PGS_REAL_DOUBLE
typedef double real;
#define SFX(n) n##_f64
#define NA 26
#define NBC 12
#else
typedef float real;
#define SFX(n) n##_f32
#define NA 11
#define NBC 9
#endif

#define SCALE 1.0e10

/* square root rounded to `real` (C's sqrt returns double: rounding it to float is the correctly
 * rounded float square root, and keeps every use in float arithmetic) */
static real rsqrt_(real x) { return (real)sqrt((double)x); }
#define NCOLMAX 1024

/* asin(y) = y sum a_n y^(2n), a_n = a_(n-1) (2n-1)^2 / (2n (2n+1)); cos(x) = sum b_n x^(2n),
 * b_n = -b_(n-1) / ((2n-1) 2n): the Taylor series, coefficients in float64 then rounded to real */
static double ca[26], cb[12];
static int coef_ready = 0;
static void coef_init(void) {
    if (coef_ready) return;
    ca[0] = 1.0;
    for (int n = 1; n < 26; ++n) {
        double t = (double)(2 * n - 1);
        ca[n] = ca[n - 1] * (t * t) / ((double)(2 * n) * (double)(2 * n + 1));
    }
    cb[0] = 1.0;
    for (int n = 1; n < 12; ++n) cb[n] = -cb[n - 1] / ((double)(2 * n - 1) * (double)(2 * n));
    coef_ready = 1;
}
static real asin_small(real y) {
    real y2 = y * y, s = (real)ca[NA - 1];
    for (int n = NA - 2; n >= 0; --n) s = (real)ca[n] + y2 * s;
    return y * s;
}
static real acos_series(real x) {
    const real half = (real)0.5, one = (real)1.0, two = (real)2.0;
    if (x <= half && x >= -half) return (real)1.5707963267948966 - asin_small(x);
    if (x > half) return two * asin_small(rsqrt_((one - x) * half));
    return (real)3.141592653589793 - two * asin_small(rsqrt_((one + x) * half));
}
static real cos_series(real x) {
    real x2 = x * x, s = (real)cb[NBC - 1];
    for (int n = NBC - 2; n >= 0; --n) s = (real)cb[n] + x2 * s;
    return s;
}

/* f_eigmax (ncp_gpu.py:392-407) */
static real eigmax(const real A[3][3]) {
    const real zero = 0, one = 1, two = 2, three = 3, six = 6, tiny = (real)1.0e-14;
    real p1 = (A[0][1] * A[0][1] + A[0][2] * A[0][2]) + A[1][2] * A[1][2];
    real q = ((A[0][0] + A[1][1]) + A[2][2]) / three;
    real m12 = A[1][1] > A[2][2] ? A[1][1] : A[2][2];
    real out = A[0][0] > m12 ? A[0][0] : m12;
    if (p1 > zero) {
        real d0 = A[0][0] - q, d1 = A[1][1] - q, d2 = A[2][2] - q;
        real p2 = ((d0 * d0 + d1 * d1) + d2 * d2) + two * p1;
        real p = rsqrt_(p2 / six);
        if (p > tiny) {
            real ip = one / p, B[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) B[i][j] = ip * (A[i][j] - q * (i == j ? one : zero));
            real c0 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
            real c1 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
            real c2 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
            real r = ((B[0][0] * c0 + B[0][1] * c1) + B[0][2] * c2) / two;
            if (!(r > -one)) r = -one;
            if (!(r < one)) r = one;
            real phi = acos_series(r) / three;
            out = q + two * p * cos_series(phi);
        }
    }
    return out;
}

/* f_proj (ncp_gpu.py:376-389) */
static void proj(real z[3], real mu) {
    real zn = z[0], zt = rsqrt_(z[1] * z[1] + z[2] * z[2]);
    if (zt <= mu * zn) return;
#ifdef PGS_PLANT_POLAR_SIGN /* planted mistake (tests only): the polar-cone test with the sign lost */
    if (mu * zt <= zn) {
#else
    if (mu * zt <= -zn) {
#endif
        z[0] = z[1] = z[2] = 0;
        return;
    }
    real s = (zn + mu * zt) / ((real)1 + mu * mu);
    real f = mu * s / zt;
    z[0] = s;
    z[1] = f * z[1];
    z[2] = f * z[2];
}

typedef struct {
    int NB, C;
    const real *Ja, *Jb, *Minv, *vfree, *mu, *bias, *sep;
    const int *pab, *feat;
    double soft[5];
    int soft_on, desaxce;
} Scene;

static void Jv_add(const real *J, const real *v, real u[3]) {
    real r[3];
    for (int i = 0; i < 3; ++i) {
        real s = J[6 * i] * v[0];
        for (int j = 1; j < 6; ++j) s = s + J[6 * i + j] * v[j];
        r[i] = s;
    }
    for (int i = 0; i < 3; ++i) u[i] = u[i] + r[i];
}
static void MJtz(const real *M, const real *J, const real z[3], real d[6]) {
    real g[6];
    for (int i = 0; i < 6; ++i) g[i] = (J[i] * z[0] + J[6 + i] * z[1]) + J[12 + i] * z[2];
    for (int i = 0; i < 6; ++i) {
        real s = M[6 * i] * g[0];
        for (int j = 1; j < 6; ++j) s = s + M[6 * i + j] * g[j];
        d[i] = s;
    }
}
static int64_t qz(double x) { return (int64_t)round(x * SCALE); }

static void contact_u(const Scene *S, const real *v, int c, real u[3]) {
    int a = S->pab[2 * c], b = S->pab[2 * c + 1];
    for (int k = 0; k < 3; ++k) u[k] = S->bias[3 * c + k];
    if (a >= 0) Jv_add(S->Ja + 18 * (size_t)c, v + 6 * (size_t)a, u);
    if (b >= 0) Jv_add(S->Jb + 18 * (size_t)c, v + 6 * (size_t)b, u);
}

/* z = proj_K(map argument) of contact c (the two branches of k_sweep) */
static void contact_z(const Scene *S, int use_bias, real rh, int c, const real u[3], const real lo[3], real z[3]) {
    real m = S->mu[c];
    if (S->soft_on) {
        real sp0 = (real)S->soft[0], sp1 = (real)S->soft[1], sp2 = (real)S->soft[2], sp3 = (real)S->soft[3],
             sp4 = (real)S->soft[4];
        real s = S->sep[c], vb = 0, ms = 1, isc = 0;
        if (s > (real)0) {
            vb = s * sp3;
        } else if (use_bias) {
            real t = sp1 * sp0 * s, lo4 = -sp4;
            vb = t > lo4 ? t : lo4;
            ms = sp1;
            isc = sp2;
        }
        real rn = u[0];
        if (S->desaxce) rn = rn + m * rsqrt_(u[1] * u[1] + u[2] * u[2]);
        z[0] = ((real)1 - isc) * lo[0] - rh * (ms * rn + vb);
        z[1] = lo[1] - rh * u[1];
        z[2] = lo[2] - rh * u[2];
    } else {
        real s0 = u[0];
        if (S->desaxce) {
            real ut = rsqrt_(u[1] * u[1] + u[2] * u[2]);
            s0 = u[0] + m * ut;
        }
        z[0] = lo[0] - rh * s0;
        z[1] = lo[1] - rh * u[1];
        z[2] = lo[2] - rh * u[2];
    }
    proj(z, m);
}

typedef struct {
    int64_t key;
    int idx;
} KI;
static int ki_cmp(const void *x, const void *y) {
    const KI *a = (const KI *)x, *b = (const KI *)y;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->idx < b->idx ? -1 : (a->idx > b->idx ? 1 : 0);
}

static uint32_t pcg(uint32_t s) {
    uint32_t b = s * 747796405u + 2891336453u;
    uint32_t c = ((b >> ((b >> 28u) + 4u)) ^ b) * 277803737u;
    return (c >> 22u) ^ c;
}

/* Global colouring (_color): canonical stable sort, rank-hashed priorities, claim/win rounds over the
 * dynamic bodies. col [C]; returns the number of colours (rounds), -1 past NCOLMAX. */
static int color_all(const Scene *S, const int *dyn, int *col) {
    int C = S->C, NB = S->NB;
    KI *kv = (KI *)malloc(sizeof(KI) * (size_t)C);
    int64_t *prio = (int64_t *)malloc(sizeof(int64_t) * (size_t)C);
    int64_t *claim = (int64_t *)malloc(sizeof(int64_t) * (size_t)(NB > 0 ? NB : 1));
    for (int i = 0; i < C; ++i) {
        uint64_t pair = (uint64_t)((int64_t)S->pab[2 * i] + 1) * (uint64_t)(NB + 1) +
                        (uint64_t)((int64_t)S->pab[2 * i + 1] + 1);
        kv[i].key = (int64_t)(pair * 1048576ull + (uint64_t)(int64_t)S->feat[i]);
        kv[i].idx = i;
    }
    qsort(kv, (size_t)C, sizeof(KI), ki_cmp);
    for (int r = 0; r < C; ++r) {
        uint32_t st = pcg(pcg(7919u + pcg((uint32_t)r)));
        int64_t p = (int64_t)(int32_t)st;
        if (p < 0) p = -p;
        prio[kv[r].idx] = (int64_t)((uint64_t)p * 4294967296ull + (uint64_t)(int64_t)r);
    }
    for (int i = 0; i < C; ++i) col[i] = -1;
    int cur = 0;
    while (1) {
        for (int i = 0; i < NB; ++i) claim[i] = INT64_MAX;
        for (int c = 0; c < C; ++c) {
            if (col[c] >= 0) continue;
            int a = S->pab[2 * c], b = S->pab[2 * c + 1];
            if (a >= 0 && dyn[a] && prio[c] < claim[a]) claim[a] = prio[c];
            if (b >= 0 && dyn[b] && prio[c] < claim[b]) claim[b] = prio[c];
        }
        int left = 0;
        for (int c = 0; c < C; ++c) {
            if (col[c] >= 0) continue;
            int a = S->pab[2 * c], b = S->pab[2 * c + 1];
            int ok = 1;
            if (a >= 0 && dyn[a] && claim[a] != prio[c]) ok = 0;
            if (b >= 0 && dyn[b] && claim[b] != prio[c]) ok = 0;
            if (ok) col[c] = cur;
            else ++left;
        }
        ++cur;
        if (left == 0) break;
        if (cur > NCOLMAX) {
            cur = -1;
            break;
        }
    }
    free(kv);
    free(prio);
    free(claim);
    return cur;
}

static void apply_body(const Scene *S, const int64_t *acc, real *v, int i) {
    for (int k = 0; k < 6; ++k)
        v[6 * (size_t)i + k] = S->vfree[6 * (size_t)i + k] + (real)((double)acc[6 * (size_t)i + k] / SCALE);
}

static void sweep_pass(const Scene *S, int use_bias, const real *rho, const int *order, const int *cstart,
                       const int *ccount, int ncol, real *lam, real *v, int64_t *acc) {
    for (int ci = 0; ci < ncol; ++ci) {
        for (int t = 0; t < ccount[ci]; ++t) {
            int c = order[cstart[ci] + t];
            int a = S->pab[2 * c], b = S->pab[2 * c + 1];
            real u[3], lo[3], z[3];
            contact_u(S, v, c, u);
            for (int k = 0; k < 3; ++k) lo[k] = lam[3 * c + k];
            contact_z(S, use_bias, rho[c], c, u, lo, z);
            for (int k = 0; k < 3; ++k) lam[3 * c + k] = z[k];
            int ab[2] = {a, b};
            for (int s = 0; s < 2; ++s) {
                int i = ab[s];
                if (i < 0) continue;
                const real *J = (s ? S->Jb : S->Ja) + 18 * (size_t)c;
                real dn[6], dd[6];
                MJtz(S->Minv + 36 * (size_t)i, J, z, dn);
                MJtz(S->Minv + 36 * (size_t)i, J, lo, dd);
#ifdef PGS_PLANT_DELTA /* planted mistake (tests only): the rounded delta, not the replaced total */
                for (int k = 0; k < 6; ++k) acc[6 * (size_t)i + k] += qz((double)(dn[k] - dd[k]));
#else
                for (int k = 0; k < 6; ++k) acc[6 * (size_t)i + k] += qz((double)dn[k]) - qz((double)dd[k]);
#endif
            }
        }
        /* k_apply_c: the bodies of this colour */
        for (int t = 0; t < ccount[ci]; ++t) {
            int c = order[cstart[ci] + t];
            if (S->pab[2 * c] >= 0) apply_body(S, acc, v, S->pab[2 * c]);
            if (S->pab[2 * c + 1] >= 0) apply_body(S, acc, v, S->pab[2 * c + 1]);
        }
    }
}

/* One NCPSolverGPU.solve from a fresh solver: set-up, `iters` biased iterations, then (relax_iters > 0)
 * the velocity snapshot and `relax_iters` unbiased ones, then the residual of the last pass.
 * soft: 5 doubles or NULL. Outputs may be NULL except lam and v. Returns the colour count (-1 past
 * 1024 colours). */
int SFX(pgs_solve)(int NB, int C, const real *Ja, const real *Jb, const real *Minv, const real *vfree,
                   const real *mu, const real *bias, const real *sep, const int *pab, const int *feat,
                   const real *warm, const double *soft, int desaxce, int iters, int relax_iters, real *lam,
                   real *v, real *vbias, real *rho_out, int *col_out, real *res_out, real *u_out,
                   double *residual) {
    coef_init();
    Scene S = {NB, C, Ja, Jb, Minv, vfree, mu, bias, sep, pab, feat, {0, 0, 0, 0, 0}, soft ? 1 : 0, desaxce};
    if (soft)
        for (int k = 0; k < 5; ++k) S.soft[k] = soft[k];
    size_t nb = (size_t)(NB > 0 ? NB : 1), nc = (size_t)(C > 0 ? C : 1);
    int *dyn = (int *)malloc(sizeof(int) * nb);
    real *rho = (real *)malloc(sizeof(real) * nc);
    int64_t *acc = (int64_t *)calloc(6 * nb, sizeof(int64_t));
    int *col = (int *)malloc(sizeof(int) * nc);
    int *order = (int *)malloc(sizeof(int) * nc);
    int *cstart = (int *)calloc(NCOLMAX + 1, sizeof(int));
    int *ccount = (int *)calloc(NCOLMAX + 1, sizeof(int));
    for (int i = 0; i < NB; ++i) {
        real m = 0;
        for (int k = 0; k < 36; ++k) {
            real x = fabs(Minv[36 * (size_t)i + k]);
            m = x > m ? x : m;
        }
        dyn[i] = m > (real)0 ? 1 : 0;
    }
    /* k_rho */
    for (int c = 0; c < C; ++c) {
        real W[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int s = 0; s < 2; ++s) {
            int i = pab[2 * c + s];
            if (i < 0) continue;
            const real *J = (s ? Jb : Ja) + 18 * (size_t)c, *M = Minv + 36 * (size_t)i;
            real JM[3][6];
            for (int r = 0; r < 3; ++r)
                for (int j = 0; j < 6; ++j) {
                    real t = 0;
                    for (int k = 0; k < 6; ++k) t = t + J[6 * r + k] * M[6 * k + j];
                    JM[r][j] = t;
                }
            for (int r = 0; r < 3; ++r)
                for (int j = 0; j < 3; ++j) {
                    real t = 0;
                    for (int k = 0; k < 6; ++k) t = t + JM[r][k] * J[6 * j + k];
                    W[r][j] = W[r][j] + t;
                }
        }
        real lm = eigmax(W);
        rho[c] = lm > (real)1.0e-14 ? (real)1 / lm : (real)0;
    }
    /* lambda, warm scatter, k_apply */
    for (int c = 0; c < C; ++c)
        for (int k = 0; k < 3; ++k) lam[3 * c + k] = warm ? warm[3 * c + k] : (real)0;
    if (warm)
        for (int c = 0; c < C; ++c)
            for (int s = 0; s < 2; ++s) {
                int i = pab[2 * c + s];
                if (i < 0) continue;
                real d[6];
                MJtz(Minv + 36 * (size_t)i, (s ? Jb : Ja) + 18 * (size_t)c, lam + 3 * c, d);
                for (int k = 0; k < 6; ++k) acc[6 * (size_t)i + k] += qz((double)d[k]);
            }
    for (int i = 0; i < NB; ++i) apply_body(&S, acc, v, i);
    int ncol = C > 0 ? color_all(&S, dyn, col) : 0;
    if (ncol >= 0 && C > 0) {
        /* colour-grouped order: a stable counting sort of the colour ids */
        for (int c = 0; c < C; ++c) ccount[col[c]]++;
        for (int k = 1; k < ncol; ++k) cstart[k] = cstart[k - 1] + ccount[k - 1];
        int *fill = (int *)calloc(NCOLMAX + 1, sizeof(int));
        for (int c = 0; c < C; ++c) order[cstart[col[c]] + fill[col[c]]++] = c;
        free(fill);
        for (int it = 0; it < iters; ++it) sweep_pass(&S, 1, rho, order, cstart, ccount, ncol, lam, v, acc);
        if (vbias) memcpy(vbias, v, sizeof(real) * 6 * (size_t)NB);
        for (int it = 0; it < relax_iters; ++it) sweep_pass(&S, 0, rho, order, cstart, ccount, ncol, lam, v, acc);
    } else if (vbias) {
        memcpy(vbias, v, sizeof(real) * 6 * (size_t)NB);
    }
    /* k_resid of the last pass */
    double rmax = 0.0;
    int ub = relax_iters > 0 ? 0 : 1;
    for (int c = 0; c < C; ++c) {
        real u[3], lo[3], z[3];
        contact_u(&S, v, c, u);
        for (int k = 0; k < 3; ++k) lo[k] = lam[3 * c + k];
        contact_z(&S, ub, rho[c], c, u, lo, z);
        real d0 = fabs(lo[0] - z[0]), d1 = fabs(lo[1] - z[1]), d2 = fabs(lo[2] - z[2]);
        real m12 = d1 > d2 ? d1 : d2;
        real r = d0 > m12 ? d0 : m12;
        if (res_out) res_out[c] = r;
        if (u_out)
            for (int k = 0; k < 3; ++k) u_out[3 * c + k] = u[k];
        if ((double)r > rmax || r != r) rmax = (double)r;
    }
    if (residual) *residual = rmax;
    if (rho_out) memcpy(rho_out, rho, sizeof(real) * (size_t)C);
    if (col_out) memcpy(col_out, col, sizeof(int) * (size_t)C);
    free(dyn);
    free(rho);
    free(acc);
    free(col);
    free(order);
    free(cstart);
    free(ccount);
    return ncol;
}

/* the pieces on their own, for the tests that pin them */
real SFX(pgs_eigmax)(const real *A9) {
    coef_init();
    real A[3][3];
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = A9[i];
    return eigmax(A);
}
real SFX(pgs_acos)(real x) {
    coef_init();
    return acos_series(x);
}
real SFX(pgs_cos)(real x) {
    coef_init();
    return cos_series(x);
}
void SFX(pgs_proj)(real *z, real mu) { proj(z, mu); }
