// Exact-cone NCP projected Gauss-Seidel: the reference's NCPSolverGPU (ncp/ncp_gpu.py:383-1069).
//
// Per contact c, with 3x6 Jacobian blocks Ja, Jb on bodies a, b (-1 = static) and the 6x6 inverse
// mass of each body, the fixed point
//     lambda_c <- proj_K(lambda_c - rho_c (u_c + Gamma(u_c))),   rho_c = 1 / lambda_max(W_cc),
// u_c = bias_c + Ja v_a + Jb v_b, Gamma(u) = (mu |u_t|, 0, 0) (de Saxce; off = the convex relaxation),
// or, in soft-step mode, the TGS soft normal row (ncp_gpu.py:491-557). Every body carries
// v = v_free + M^-1 J^T lambda rebuilt from an int64 fixed-point TOTAL (scale 1e10): a contact update
// replaces its own rounded contribution (f_swap, ncp_gpu.py:419-432), so v is a pure function of the
// current lambda and two runs agree bit for bit.
//
// MI355X layout. The reference launches one sweep kernel and one apply kernel per colour per
// iteration. Here the contacts are cut into islands (connected through dynamic bodies); islands are
// packed, in island order, into tiles of about PGS_TILE contacts; one workgroup colours its tile
// (Jones-Plassmann rounds with the reference's priorities, which are local to an island, so every
// contact gets the colour the global rounds give it) and then runs every iteration of the tile with
// one barrier per colour. A colour's contacts share no dynamic body, so the apply of a colour
// (k_apply_c, ncp_gpu.py:585-599) is fused into the contact update: each thread rebuilds v of the
// dynamic bodies it has just updated. Static bodies (all-zero inverse mass) gain exactly zero from
// every contact and keep v = v_free + 0.
//
// Floating point follows the reference's Warp kernels with contraction off: matrix-vector products
// sum left to right from the first product, matrix-matrix products from zero, acos/cos are this
// file's own series (pgs_acos, pgs_cos; the reference calls CUDA's), and the float64 rounding of the
// fixed point is round-half-away-from-zero.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kernels.hpp"

namespace me {

#define PGS_THREADS 256
#define PGS_TILE 256       // contacts of a tile before the next island opens a new one
#define PGS_NCOLMAX 1024   // colours (the reference's _NCOLMAX, ncp_gpu.py:103)
#define PGS_SCALE 1.0e10   // fixed-point scale (_S, ncp_gpu.py:98)

// int32 status slots of one solver
enum { PS_MISMATCH = 0, PS_NCOL, PS_OVF, PS_NISL, PS_NTILE, PS_BAD, PS_COUNT = 16 };

template <class T>
struct PgsScene {
    const T *Ja, *Jb, *Minv, *vfree, *mu, *bias, *sep;
    const int *pab, *feat;
};

template <class T>
struct PgsSoft {  // (bias_rate, mass_scale, impulse_scale, inv_h, contact_speed) in T, and the flags
    T p0, p1, p2, p3, p4;
    int on, use_bias;
};

// ── series for the closed-form eigenvalue (acos, cos) ──
// asin(y) = y * sum_n a_n y^(2n), a_0 = 1, a_n = a_(n-1) (2n-1)^2 / (2n (2n+1)), |y| <= 1/2;
// cos(x) = sum_n b_n x^(2n), b_0 = 1, b_n = -b_(n-1) / ((2n-1) 2n), 0 <= x <= pi/3.
// Coefficients in float64 (the recurrences evaluated once), rounded to T; Horner in T.
template <class T> struct PgsSeries;
template <> struct PgsSeries<float> { static constexpr int NA = 11, NB = 9; };
template <> struct PgsSeries<double> { static constexpr int NA = 26, NB = 12; };

struct PgsCoef {
    double a[26], b[12];
};
constexpr PgsCoef pgs_make_coef() {
    PgsCoef c{};
    c.a[0] = 1.0;
    for (int n = 1; n < 26; ++n) {
        double t = (double)(2 * n - 1);
        c.a[n] = c.a[n - 1] * (t * t) / ((double)(2 * n) * (double)(2 * n + 1));
    }
    c.b[0] = 1.0;
    for (int n = 1; n < 12; ++n) c.b[n] = -c.b[n - 1] / ((double)(2 * n - 1) * (double)(2 * n));
    return c;
}
__device__ constexpr PgsCoef kPgsCoef = pgs_make_coef();

template <class T>
__device__ __forceinline__ T pgs_asin_small(T y) {
    constexpr int N = PgsSeries<T>::NA;
    T y2 = y * y;
    T s = (T)kPgsCoef.a[N - 1];
#pragma unroll
    for (int n = N - 2; n >= 0; --n) s = (T)kPgsCoef.a[n] + y2 * s;
    return y * s;
}
template <class T>
__device__ __forceinline__ T pgs_acos(T x) {  // x in [-1, 1]
    const T half = (T)0.5, one = (T)1.0, two = (T)2.0;
    if (x <= half && x >= -half) return (T)1.5707963267948966 - pgs_asin_small<T>(x);
    if (x > half) return two * pgs_asin_small<T>(sqrt((one - x) * half));
    return (T)3.141592653589793 - two * pgs_asin_small<T>(sqrt((one + x) * half));
}
template <class T>
__device__ __forceinline__ T pgs_cos(T x) {  // x in [0, pi/3]
    constexpr int N = PgsSeries<T>::NB;
    T x2 = x * x;
    T s = (T)kPgsCoef.b[N - 1];
#pragma unroll
    for (int n = N - 2; n >= 0; --n) s = (T)kPgsCoef.b[n] + x2 * s;
    return s;
}

// largest eigenvalue of a symmetric 3x3 (f_eigmax, ncp_gpu.py:392-407)
template <class T>
__device__ __forceinline__ T pgs_eigmax(const T A[3][3]) {
    const T zero = (T)0.0, one = (T)1.0, two = (T)2.0, three = (T)3.0, six = (T)6.0, tiny = (T)1.0e-14;
    T p1 = (A[0][1] * A[0][1] + A[0][2] * A[0][2]) + A[1][2] * A[1][2];
    T q = ((A[0][0] + A[1][1]) + A[2][2]) / three;
    T m12 = A[1][1] > A[2][2] ? A[1][1] : A[2][2];
    T out = A[0][0] > m12 ? A[0][0] : m12;
    if (p1 > zero) {
        T d0 = A[0][0] - q, d1 = A[1][1] - q, d2 = A[2][2] - q;
        T p2 = ((d0 * d0 + d1 * d1) + d2 * d2) + two * p1;
        T p = sqrt(p2 / six);
        if (p > tiny) {
            T ip = one / p, B[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) B[i][j] = ip * (A[i][j] - q * (i == j ? one : zero));
            // determinant as the triple product row0 . (row1 x row2)
            T c0 = B[1][1] * B[2][2] - B[1][2] * B[2][1];
            T c1 = B[1][2] * B[2][0] - B[1][0] * B[2][2];
            T c2 = B[1][0] * B[2][1] - B[1][1] * B[2][0];
            T r = ((B[0][0] * c0 + B[0][1] * c1) + B[0][2] * c2) / two;
            r = r > -one ? r : -one;
            r = r < one ? r : one;
            T phi = pgs_acos<T>(r) / three;
            out = q + two * p * pgs_cos<T>(phi);
        }
    }
    return out;
}

// exact projection onto K_mu (f_proj, ncp_gpu.py:376-389)
template <class T>
__device__ __forceinline__ void pgs_proj(T &z0, T &z1, T &z2, T mu) {
    T zn = z0;
    T zt = sqrt(z1 * z1 + z2 * z2);
    if (zt <= mu * zn) return;
    if (mu * zt <= -zn) {
        z0 = (T)0.0; z1 = (T)0.0; z2 = (T)0.0;
        return;
    }
    T s = (zn + mu * zt) / ((T)1.0 + mu * mu);
    T f = mu * s / zt;
    z0 = s; z1 = f * z1; z2 = f * z2;
}

// u += J v (3x6 block, row-major)
template <class T>
__device__ __forceinline__ void pgs_Jv(const T *J, const T *v, T &u0, T &u1, T &u2) {
    T r[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T s = J[6 * i] * v[0];
#pragma unroll
        for (int j = 1; j < 6; ++j) s = s + J[6 * i + j] * v[j];
        r[i] = s;
    }
    u0 = u0 + r[0]; u1 = u1 + r[1]; u2 = u2 + r[2];
}

// d = Minv (J^T z)
template <class T>
__device__ __forceinline__ void pgs_MJtz(const T *M, const T *J, T z0, T z1, T z2, T d[6]) {
    T g[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) g[i] = (J[i] * z0 + J[6 + i] * z1) + J[12 + i] * z2;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        T s = M[6 * i] * g[0];
#pragma unroll
        for (int j = 1; j < 6; ++j) s = s + M[6 * i + j] * g[j];
        d[i] = s;
    }
}

__device__ __forceinline__ int64_t pgs_q(double x) { return (int64_t)round(x * PGS_SCALE); }

// the fixed-point map's argument z for contact c (Baumgarte or soft-step row), from u and lambda
template <class T>
__device__ __forceinline__ void pgs_z(const PgsSoft<T> &sp, int desaxce, T m, T sep, T rh, T u0, T u1, T u2, T lo0,
                                     T lo1, T lo2, T &z0, T &z1, T &z2) {
    if (sp.on) {
        T s = sep, vb = (T)0.0, ms = (T)1.0, isc = (T)0.0;
        if (s > (T)0.0) {
            vb = s * sp.p3;
        } else if (sp.use_bias) {
            T t = sp.p1 * sp.p0 * s, lo = -sp.p4;
            vb = t > lo ? t : lo;
            ms = sp.p1;
            isc = sp.p2;
        }
        T rn = u0;
        if (desaxce) rn = rn + m * sqrt(u1 * u1 + u2 * u2);
        z0 = ((T)1.0 - isc) * lo0 - rh * (ms * rn + vb);
        z1 = lo1 - rh * u1;
        z2 = lo2 - rh * u2;
    } else {
        T s0 = u0;
        if (desaxce) {
            T ut = sqrt(u1 * u1 + u2 * u2);
            s0 = u0 + m * ut;
        }
        z0 = lo0 - rh * s0;
        z1 = lo1 - rh * u1;
        z2 = lo2 - rh * u2;
    }
    pgs_proj<T>(z0, z1, z2, m);
}

template <class T>
__device__ __forceinline__ void pgs_u(const PgsScene<T> &S, const T *v, int c, int a, int b, T &u0, T &u1, T &u2) {
    u0 = S.bias[3 * c]; u1 = S.bias[3 * c + 1]; u2 = S.bias[3 * c + 2];
    if (a >= 0) pgs_Jv<T>(S.Ja + 18 * (size_t)c, v + 6 * (size_t)a, u0, u1, u2);
    if (b >= 0) pgs_Jv<T>(S.Jb + 18 * (size_t)c, v + 6 * (size_t)b, u0, u1, u2);
}

// ── set-up ──

// body indices in [-1, NB): every later kernel indexes bodies by them
__global__ void k_pgs_validate(int C, int NB, const int *pab, int *ps) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (c < C) {
        int a = pab[2 * c], b = pab[2 * c + 1];
        bad = a < -1 || a >= NB || b < -1 || b >= NB;
    }
    if (__ballot(bad) && __lane_id() == 0) atomicOr(&ps[PS_BAD], 1);
}

// dyn flag per body (any nonzero inverse-mass entry; the reference's upload, ncp_gpu.py:868)
template <class T>
__global__ void k_pgs_dyn(int NB, const T *Minv, int *dyn) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NB) return;
    T m = (T)0.0;
    for (int k = 0; k < 36; ++k) {
        T x = fabs(Minv[36 * (size_t)i + k]);
        m = x > m ? x : m;
    }
    dyn[i] = m > (T)0.0 ? 1 : 0;
}

// rho_c = 1 / lambda_max(W_cc), W = Ja Minv_a Ja^T + Jb Minv_b Jb^T (k_rho, ncp_gpu.py:435-451)
template <class T>
__device__ __forceinline__ void pgs_W_add(const T *J, const T *M, T W[3][3]) {
    T JM[3][6];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 6; ++j) {
            T s = (T)0.0;
            for (int k = 0; k < 6; ++k) s = s + J[6 * i + k] * M[6 * k + j];
            JM[i][j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T s = (T)0.0;
            for (int k = 0; k < 6; ++k) s = s + JM[i][k] * J[6 * j + k];
            W[i][j] = W[i][j] + s;
        }
}
template <class T>
__global__ void k_pgs_rho(int C, PgsScene<T> S, T *rho) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    T W[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) W[i][j] = (T)0.0;
    int a = S.pab[2 * c], b = S.pab[2 * c + 1];
    if (a >= 0) pgs_W_add<T>(S.Ja + 18 * (size_t)c, S.Minv + 36 * (size_t)a, W);
    if (b >= 0) pgs_W_add<T>(S.Jb + 18 * (size_t)c, S.Minv + 36 * (size_t)b, W);
    T lm = pgs_eigmax<T>(W);
    rho[c] = lm > (T)1.0e-14 ? (T)1.0 / lm : (T)0.0;
}

// lambda <- warm (or zero); the accumulator cleared
template <class T>
__global__ void k_pgs_lam_init(int C, int NB, const T *warm, T *lam, int64_t *acc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * C) lam[i] = warm ? warm[i] : (T)0.0;
    if (i < 6 * NB) acc[i] = 0;
}

// warm start: acc += round(S Minv J^T lambda) for every contact (k_scatter, ncp_gpu.py:459-474)
template <class T>
__global__ void k_pgs_scatter(int C, PgsScene<T> S, const int *dyn, const T *lam, int64_t *acc) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    T l0 = lam[3 * c], l1 = lam[3 * c + 1], l2 = lam[3 * c + 2];
    int ab[2] = {S.pab[2 * c], S.pab[2 * c + 1]};
    for (int s = 0; s < 2; ++s) {
        int i = ab[s];
        if (i < 0 || !dyn[i]) continue;  // a static body gains exactly zero
        T d[6];
        pgs_MJtz<T>(S.Minv + 36 * (size_t)i, (s ? S.Jb : S.Ja) + 18 * (size_t)c, l0, l1, l2, d);
        for (int k = 0; k < 6; ++k)
            atomicAdd(reinterpret_cast<unsigned long long *>(acc + 6 * (size_t)i + k),
                      (unsigned long long)pgs_q((double)d[k]));
    }
}

// v = v_free + acc / S for every body (k_apply, ncp_gpu.py:572-583); v_bias starts equal
template <class T>
__global__ void k_pgs_apply(int NB, const T *vfree, const int64_t *acc, T *v, T *vb) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * NB) return;
    T x = vfree[i] + (T)((double)acc[i] / PGS_SCALE);
    v[i] = x;
    vb[i] = x;
}

template <class T>
__global__ void k_pgs_copy(int n, const T *src, T *dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// colour-cache test: pairs, feature ids, the translational diagonal and the per-row dynamic flags of
// the inverse mass (k_cache_cmp, ncp_gpu.py:744-770)
template <class T>
__global__ void k_pgs_cache_cmp(int C, int NB, PgsScene<T> S, const int *cpab, const int *cfeat, const T *cMinv,
                                int *ps) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool diff = false;
    if (i < C) diff = S.pab[2 * i] != cpab[2 * i] || S.pab[2 * i + 1] != cpab[2 * i + 1] || S.feat[i] != cfeat[i];
    if (i < NB) {
        const T *m = S.Minv + 36 * (size_t)i, *o = cMinv + 36 * (size_t)i;
        for (int r = 0; r < 3; ++r)
            if (m[7 * r] != o[7 * r]) diff = true;
        for (int r = 0; r < 6; ++r) {
            T dm = (T)0.0, dq = (T)0.0;
            for (int k = 0; k < 6; ++k) {
                dm = dm + fabs(m[6 * r + k]);
                dq = dq + fabs(o[6 * r + k]);
            }
            if ((dm > (T)0.0) != (dq > (T)0.0)) diff = true;
        }
    }
    if (__ballot(diff) && __lane_id() == 0) atomicMax(&ps[PS_MISMATCH], 1);
}

template <class T>
__global__ void k_pgs_cache_store(int C, int NB, PgsScene<T> S, int *cpab, int *cfeat, T *cMinv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < C) { cpab[2 * i] = S.pab[2 * i]; cpab[2 * i + 1] = S.pab[2 * i + 1]; cfeat[i] = S.feat[i]; }
    if (i < 36 * NB) cMinv[i] = S.Minv[i];
}

// ── deterministic colouring ──

// canonical key (body_a + 1, body_b + 1, feature) (k_sortkey, ncp_gpu.py:672-683), wrapped as int64
// arithmetic wraps; its bucket floor(key / (2^20 (NB+1))) clamped to [0, NB], a monotone function of
// the key, so the stable-sort rank is the rank of the bucket plus the rank inside it
__global__ void k_pgs_keys(int C, int NB, const int *pab, const int *feat, int64_t *key, int *bucket, int *bcnt) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    uint64_t pair = (uint64_t)((int64_t)pab[2 * c] + 1) * (uint64_t)(NB + 1) + (uint64_t)((int64_t)pab[2 * c + 1] + 1);
    int64_t k = (int64_t)(pair * 1048576ull + (uint64_t)(int64_t)feat[c]);
    int64_t D = 1048576LL * (int64_t)(NB + 1);
    int64_t q = k / D;
    if (k % D != 0 && k < 0) q -= 1;
    int bk = q < 0 ? 0 : (q > NB ? NB : (int)q);
    key[c] = k;
    bucket[c] = bk;
    atomicAdd(&bcnt[bk], 1);
}

__global__ void k_pgs_bucket_fill(int C, const int *bucket, const int *boff, int *bfill, int *items) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    int bk = bucket[c];
    items[boff[bk] + atomicAdd(&bfill[bk], 1)] = c;
}

// rank in the stable order of the keys, and the Jones-Plassmann priority of that rank
// (k_prio, ncp_gpu.py:685-697: |randi(rand_init(7919, r))| << 32 | r, int64 wrap included)
__global__ void k_pgs_rank(int C, const int64_t *key, const int *bucket, const int *boff, const int *bcnt,
                           const int *items, int64_t *prio) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    int bk = bucket[c], lo = boff[bk], n = bcnt[bk];
    int64_t kc = key[c];
    int r = lo;
    for (int j = 0; j < n; ++j) {
        int o = items[lo + j];
        int64_t ko = key[o];
        r += (ko < kc || (ko == kc && o < c)) ? 1 : 0;
    }
    uint32_t st = pcg(pcg((uint32_t)7919 + pcg((uint32_t)r)));
    int64_t p = (int64_t)(int32_t)st;
    if (p < 0) p = -p;
    prio[c] = (int64_t)((uint64_t)p * 4294967296ull + (uint64_t)(int64_t)r);
}

// union-find over contacts whose two bodies are dynamic (static endpoints never join)
__device__ __forceinline__ int pgs_find(int *par, int x) {
    volatile int *vp = par;
    while (true) {
        int p = vp[x];
        if (p == x) return x;
        int g = vp[p];
        if (g == p) return p;
        vp[x] = g;  // x is not a root: only ever shortened to an ancestor
        x = g;
    }
}
__global__ void k_pgs_uf_init(int NB, int *par) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NB) par[i] = i;
}
__global__ void k_pgs_union(int C, const int *pab, const int *dyn, int *par) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    int a = pab[2 * c], b = pab[2 * c + 1];
    if (a < 0 || b < 0 || !dyn[a] || !dyn[b]) return;
    while (true) {
        int ra = pgs_find(par, a), rb = pgs_find(par, b);
        if (ra == rb) return;
        int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
        if (atomicCAS(&par[hi], hi, lo) == hi) return;
    }
}
// island of each contact: the root of a dynamic endpoint, else its own slot NB + c
__global__ void k_pgs_island(int C, int NB, const int *pab, const int *dyn, int *par, int *isl, int *icnt) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    int a = pab[2 * c], b = pab[2 * c + 1];
    int I;
    if (a >= 0 && dyn[a]) I = pgs_find(par, a);
    else if (b >= 0 && dyn[b]) I = pgs_find(par, b);
    else I = NB + c;
    isl[c] = I;
    atomicAdd(&icnt[I], 1);
}
// tile of an island = floor(its first contact offset / PGS_TILE); a tile spans the islands that start
// in its window, so each tile is one contiguous range of the island-ordered contacts
__global__ void k_pgs_tiles(int NI, const int *icnt, const int *ioff, int *tlo, int *thi, int *ps) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = I < NI && icnt[I] > 0;
    wave_atomic_count(&ps[PS_NISL], live);
    if (!live) return;
    int t = ioff[I] / PGS_TILE;
    atomicMin(&tlo[t], ioff[I]);
    atomicMax(&thi[t], ioff[I] + icnt[I]);
}
__global__ void k_pgs_island_fill(int C, const int *isl, const int *ioff, int *ifill, int *iorder) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    int I = isl[c];
    iorder[ioff[I] + atomicAdd(&ifill[I], 1)] = c;
}

__device__ __forceinline__ long long pgs_claim_load(long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup per tile: Jones-Plassmann rounds (k_claim / k_win, ncp_gpu.py:699-739) over the
// tile's contacts, then the tile's contacts counting-sorted by colour into order[lo, hi) with the
// colour segment offsets in seg[t * (PGS_NCOLMAX + 1) + k].
__global__ void __launch_bounds__(PGS_THREADS) k_pgs_color(const int *pab, const int *dyn, const int64_t *prio,
                                                          const int *tlo, const int *thi, const int *iorder,
                                                          long long *claim, int *col, int *order, int *seg,
                                                          int *tncol, int *ps) {
    __shared__ int left;
    __shared__ int cnt[PGS_NCOLMAX + 1];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int lo = tlo[t], hi = thi[t];
    if (lo >= hi) return;
    for (int j = lo + tid; j < hi; j += PGS_THREADS) col[iorder[j]] = -1;
    __syncthreads();
    int cur = 0;
    while (true) {
        if (tid == 0) left = 0;
        for (int j = lo + tid; j < hi; j += PGS_THREADS) {
            int c = iorder[j];
            if (col[c] >= 0) continue;
            int a = pab[2 * c], b = pab[2 * c + 1];
            if (a >= 0 && dyn[a]) atomicExch(reinterpret_cast<unsigned long long *>(&claim[a]), 0x7fffffffffffffffull);
            if (b >= 0 && dyn[b]) atomicExch(reinterpret_cast<unsigned long long *>(&claim[b]), 0x7fffffffffffffffull);
        }
        __syncthreads();
        for (int j = lo + tid; j < hi; j += PGS_THREADS) {
            int c = iorder[j];
            if (col[c] >= 0) continue;
            int a = pab[2 * c], b = pab[2 * c + 1];
            long long k = prio[c];
            if (a >= 0 && dyn[a]) atomicMin(&claim[a], k);
            if (b >= 0 && dyn[b]) atomicMin(&claim[b], k);
        }
        __syncthreads();
        for (int j = lo + tid; j < hi; j += PGS_THREADS) {
            int c = iorder[j];
            if (col[c] >= 0) continue;
            int a = pab[2 * c], b = pab[2 * c + 1];
            long long k = prio[c];
            bool ok = true;
            if (a >= 0 && dyn[a] && pgs_claim_load(&claim[a]) != k) ok = false;
            if (b >= 0 && dyn[b] && pgs_claim_load(&claim[b]) != k) ok = false;
            if (ok) col[c] = cur;
            else atomicAdd(&left, 1);
        }
        __syncthreads();
        ++cur;
        int l = left;
        __syncthreads();
        if (l == 0) break;
        if (cur >= PGS_NCOLMAX) {  // the reference raises past _NCOLMAX colours
            if (tid == 0) atomicOr(&ps[PS_OVF], 1);
            for (int j = lo + tid; j < hi; j += PGS_THREADS)
                if (col[iorder[j]] < 0) col[iorder[j]] = PGS_NCOLMAX - 1;
            __syncthreads();
            break;
        }
    }
    const int nc = cur;
    if (tid == 0) {
        atomicMax(&ps[PS_NCOL], nc);
        atomicAdd(&ps[PS_NTILE], 1);
        tncol[t] = nc;
    }
    for (int k = tid; k <= nc; k += PGS_THREADS) cnt[k] = 0;
    __syncthreads();
    for (int j = lo + tid; j < hi; j += PGS_THREADS) atomicAdd(&cnt[col[iorder[j]]], 1);
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        int *sg = seg + (size_t)t * (PGS_NCOLMAX + 1);
        for (int k = 0; k < nc; ++k) {
            int n = cnt[k];
            cnt[k] = run;
            sg[k] = lo + run;
            run += n;
        }
        sg[nc] = lo + run;
    }
    __syncthreads();
    for (int j = lo + tid; j < hi; j += PGS_THREADS) {
        int c = iorder[j];
        order[lo + atomicAdd(&cnt[col[c]], 1)] = c;
    }
}

// ── the sweeps ──

// dynamic bodies of each tile: their slot in the tile (LDS index) and the tile's body list
__global__ void k_pgs_body_slot(int NB, const int *dyn, int *par, const int *icnt, const int *ioff, int *tbcnt,
                                int *bslot, int *btile) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NB) return;
    int sl = -1, t = -1;
    if (dyn[i]) {
        int r = pgs_find(par, i);
        if (icnt[r] > 0) {
            t = ioff[r] / PGS_TILE;
            sl = atomicAdd(&tbcnt[t], 1);
        }
    }
    bslot[i] = sl;
    btile[i] = t;
}
__global__ void k_pgs_body_fill(int NB, const int *bslot, const int *btile, const int *tbstart, int *tbody) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NB || bslot[i] < 0) return;
    tbody[tbstart[btile[i]] + bslot[i]] = i;
}

#define PGS_LANE 40         // reals of a lane record: J block (18), 3 inverse-mass rows (18), 3 of v_free, pad

// Per-solve contact records in tile-position order j (the order the sweeps visit): the ints
// (contact, bodies, their LDS slots, dynamic flags), the scalars (bias, mu, rho, separation) and, per
// lane r of the contact's quad (side r >> 1, rows 3 (r & 1) .. +2), the side's J block, the three
// inverse-mass rows and v_free entries that lane updates. A colour step then reads its contacts at
// addresses known from j alone, in one round of loads.
template <class T>
__global__ void k_pgs_pack(int C, PgsScene<T> S, const int *dyn, const int *bslot, const int *order, const T *rho,
                           int *pint, T *pscal, T *plane) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= C) return;
    int c = order[j];
    int a = S.pab[2 * c], b = S.pab[2 * c + 1];
    bool da = a >= 0 && dyn[a], db = b >= 0 && dyn[b];
    int *I = pint + 8 * (size_t)j;
    I[0] = c; I[1] = a; I[2] = b;
    I[3] = da ? bslot[a] : -1;
    I[4] = db ? bslot[b] : -1;
    I[5] = (da ? 1 : 0) | (db ? 2 : 0);
    I[6] = 0; I[7] = 0;
    T *P = pscal + 8 * (size_t)j;
    P[0] = S.bias[3 * c]; P[1] = S.bias[3 * c + 1]; P[2] = S.bias[3 * c + 2];
    P[3] = S.mu[c]; P[4] = rho[c]; P[5] = S.sep ? S.sep[c] : (T)0.0; P[6] = (T)0.0; P[7] = (T)0.0;
    for (int r = 0; r < 4; ++r) {
        int side = r >> 1, h = r & 1, i = side ? b : a;
        T *L = plane + (4 * (size_t)j + r) * PGS_LANE;
        const T *J = (side ? S.Jb : S.Ja) + 18 * (size_t)c;
        for (int k = 0; k < 18; ++k) L[k] = i >= 0 ? J[k] : (T)0.0;
        bool d = side ? db : da;
        for (int k = 0; k < 18; ++k) L[18 + k] = d ? S.Minv[36 * (size_t)i + 18 * h + k] : (T)0.0;
        for (int k = 0; k < 3; ++k) L[36 + k] = d ? S.vfree[6 * (size_t)i + 3 * h + k] : (T)0.0;
        L[39] = (T)0.0;
    }
}

// the record of one contact as one lane of its quad needs it
template <class T>
struct PgsRec {
    int a, b, i, sl, dy;
    T sc[6];   // bias (3), mu, rho, separation
    T L[39];   // J block of the lane's side, its three inverse-mass rows, their v_free entries
    T lo[3];   // lambda at the start of the update
};

template <class T>
__device__ __forceinline__ void pgs_load_rec(PgsRec<T> &R, const int *pint, const T *pscal, const T *plane,
                                             const T *lamo, int j, int r) {
    const int s = r >> 1;
    const int *I = pint + 8 * (size_t)j;
    R.a = I[1];
    R.b = I[2];
    R.i = s ? R.b : R.a;
    R.sl = s ? I[4] : I[3];
    R.dy = (I[5] >> s) & 1;
    const T *Sc = pscal + 8 * (size_t)j;
#pragma unroll
    for (int k = 0; k < 6; ++k) R.sc[k] = Sc[k];
    const T *L = plane + (4 * (size_t)j + r) * PGS_LANE;
#pragma unroll
    for (int k = 0; k < 39; ++k) R.L[k] = L[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) R.lo[k] = lamo[3 * (size_t)j + k];
}

// One contact by the four lanes of a quad (lane r: side r >> 1 = body a / body b, half r & 1 = rows
// 0-2 / 3-5 of that body's six components). Each side forms its J v; the quad exchanges the two
// partial products and every lane forms u and the new lambda the same way; each lane then replaces
// its three components of its body's total and rebuilds them. Every output is computed by the same
// operations, in the same order, as when one thread updates the whole contact. The new lambda is
// returned in z and stored at tile position j by lane 0.
template <class T>
__device__ __forceinline__ void pgs_contact_quad(const PgsSoft<T> &sp, int desaxce, const PgsRec<T> &R, T *lamo,
                                                 T *gv, int64_t *gacc, T *lv, int64_t *lacc, bool lds, int j, int r,
                                                 T z[3]) {
    const int h = r & 1;
    const bool has = R.i >= 0;
    T *vi = nullptr;
    int64_t *ai = nullptr;
    if (has) {
        if (lds && R.dy) {
            vi = lv + 6 * R.sl;
            ai = lacc + 6 * R.sl;
        } else {
            vi = gv + 6 * (size_t)R.i;
            ai = gacc + 6 * (size_t)R.i;
        }
    }
    const T *J = R.L;
    T P[3] = {(T)0.0, (T)0.0, (T)0.0};
    if (has) {
        T x[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) x[k] = vi[k];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            T t = J[6 * q] * x[0];
#pragma unroll
            for (int k = 1; k < 6; ++k) t = t + J[6 * q + k] * x[k];
            P[q] = t;
        }
    }
    const int base = __lane_id() & ~3;
    T pa[3], pb[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        pa[q] = __shfl(P[q], base);
        pb[q] = __shfl(P[q], base + 2);
    }
    T u0 = R.sc[0], u1 = R.sc[1], u2 = R.sc[2];
    if (R.a >= 0) { u0 = u0 + pa[0]; u1 = u1 + pa[1]; u2 = u2 + pa[2]; }
    if (R.b >= 0) { u0 = u0 + pb[0]; u1 = u1 + pb[1]; u2 = u2 + pb[2]; }
    const T lo0 = R.lo[0], lo1 = R.lo[1], lo2 = R.lo[2];
    T z0, z1, z2;
    pgs_z<T>(sp, desaxce, R.sc[3], R.sc[5], R.sc[4], u0, u1, u2, lo0, lo1, lo2, z0, z1, z2);
    z[0] = z0; z[1] = z1; z[2] = z2;
    if (r == 0) { lamo[3 * (size_t)j] = z0; lamo[3 * (size_t)j + 1] = z1; lamo[3 * (size_t)j + 2] = z2; }
    if (!R.dy) return;
    T g[6], go[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        g[k] = (J[k] * z0 + J[6 + k] * z1) + J[12 + k] * z2;
        go[k] = (J[k] * lo0 + J[6 + k] * lo1) + J[12 + k] * lo2;
    }
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
        const T *M = R.L + 18 + 6 * kk;
        T dn = M[0] * g[0], dd = M[0] * go[0];
#pragma unroll
        for (int q = 1; q < 6; ++q) {
            dn = dn + M[q] * g[q];
            dd = dd + M[q] * go[q];
        }
        const int row = 3 * h + kk;
        int64_t t = ai[row] + (pgs_q((double)dn) - pgs_q((double)dd));
        ai[row] = t;
        vi[row] = R.L[36 + kk] + (T)((double)t / PGS_SCALE);
    }
}

#define PGS_LDS_BODIES 600    // dynamic bodies of a tile kept in LDS (velocity and int64 total)
#define PGS_SWEEP_THREADS 512  // a quad per contact: 128 contacts of a colour at a time

// One workgroup per tile, every iteration, one barrier per colour; a quad per contact. The tile's
// dynamic bodies (velocity, int64 total) sit in LDS for the launch when they fit, else in global
// memory; static bodies are read from global memory (never written). lambda is gathered into tile
// positions for the launch (lamo) and scattered back at its end.
template <class T>
__global__ void __launch_bounds__(PGS_SWEEP_THREADS) k_pgs_sweeps(int iters, PgsSoft<T> sp, int desaxce,
                                                                 const int *tlo, const int *thi, const int *tncol,
                                                                 const int *seg, const int *order, const int *tbstart,
                                                                 const int *tbcnt, const int *tbody, const int *pint,
                                                                 const T *pscal, const T *plane, T *lam, T *lamo, T *v,
                                                                 int64_t *acc) {
    extern __shared__ __align__(16) unsigned char pgs_smem[];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int lo = tlo[t], hi = thi[t];
    if (lo >= hi) return;
    const int nc = tncol[t];
    const int *sg = seg + (size_t)t * (PGS_NCOLMAX + 1);
    const int nb = tbcnt[t], b0 = tbstart[t];
    const bool lds = nb <= PGS_LDS_BODIES;
    int64_t *lacc = reinterpret_cast<int64_t *>(pgs_smem);
    T *lv = reinterpret_cast<T *>(pgs_smem + sizeof(int64_t) * 6 * PGS_LDS_BODIES);
    if (lds)
        for (int k = tid; k < 6 * nb; k += PGS_SWEEP_THREADS) {
            size_t g = 6 * (size_t)tbody[b0 + k / 6] + k % 6;
            lacc[k] = acc[g];
            lv[k] = v[g];
        }
    for (int k = lo * 3 + tid; k < hi * 3; k += PGS_SWEEP_THREADS) lamo[k] = lam[3 * (size_t)order[k / 3] + k % 3];
    __syncthreads();
    const int quad = tid >> 2, r = tid & 3;
    for (int it = 0; it < iters; ++it)
        for (int k = 0; k < nc; ++k) {
            const int e = sg[k + 1];
            for (int j = sg[k] + quad; j < e; j += PGS_SWEEP_THREADS / 4) {
                PgsRec<T> R;
                pgs_load_rec<T>(R, pint, pscal, plane, lamo, j, r);
                T z[3];
                pgs_contact_quad<T>(sp, desaxce, R, lamo, v, acc, lv, lacc, lds, j, r, z);
            }
            __syncthreads();
        }
    if (lds)
        for (int k = tid; k < 6 * nb; k += PGS_SWEEP_THREADS) {
            size_t g = 6 * (size_t)tbody[b0 + k / 6] + k % 6;
            acc[g] = lacc[k];
            v[g] = lv[k];
        }
    for (int k = lo * 3 + tid; k < hi * 3; k += PGS_SWEEP_THREADS) lam[3 * (size_t)order[k / 3] + k % 3] = lamo[k];
}

// natural-map residual ||lambda - proj_K(z)||_inf per contact and the contact velocity u
// (k_resid, ncp_gpu.py:601-646); the maximum as the bit pattern of a non-negative double
template <class T>
__global__ void k_pgs_residual(int C, PgsScene<T> S, PgsSoft<T> sp, int desaxce, const T *rho, const T *lam,
                               const T *v, T *res, T *uout, unsigned long long *rmax) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    double r = 0.0;
    if (c < C) {
        int a = S.pab[2 * c], b = S.pab[2 * c + 1];
        T u0, u1, u2;
        pgs_u<T>(S, v, c, a, b, u0, u1, u2);
        uout[3 * c] = u0; uout[3 * c + 1] = u1; uout[3 * c + 2] = u2;
        T l0 = lam[3 * c], l1 = lam[3 * c + 1], l2 = lam[3 * c + 2];
        T z0, z1, z2;
        pgs_z<T>(sp, desaxce, S.mu[c], sp.on ? S.sep[c] : (T)0.0, rho[c], u0, u1, u2, l0, l1, l2, z0, z1, z2);
        T d0 = fabs(l0 - z0), d1 = fabs(l1 - z1), d2 = fabs(l2 - z2);
        T m12 = d1 > d2 ? d1 : d2;
        T rr = d0 > m12 ? d0 : m12;
        res[c] = rr;
        r = (double)rr;
    }
    unsigned long long bits = (unsigned long long)__double_as_longlong(r);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(bits, o);
        bits = w > bits ? w : bits;
    }
    if (__lane_id() == 0) atomicMax(rmax, bits);
}

}  // namespace me
