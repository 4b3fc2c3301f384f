// Batched fixed-branch NCP sensitivity (float64): for every environment e, given the Delassus
// operator G, the free velocity b, friction mu and a solved impulse lam of n contacts, classify each
// contact open / stick / slip, assemble the fixed-branch residual Jacobian A and the parameter maps
// B_mu, B_b (ncp/ncp_ref.py:786-842), and solve A X = -[B_mu | B_b D] for dlam/dmu and dlam/db along
// the caller's directions D (ncp_ref.py:907-959; on the reference's GPU this is build_active +
// grad_affine, ncp/ncp_batch_gpu.py:555-597, 1160-1187).
//
// One wave per environment. The augmented matrix [A | B_mu | B_b D] (3n rows) is factorised in
// place by LU with partial pivoting (first maximal pivot on ties) and back-substituted per
// right-hand-side column; it lives in LDS when it fits, else in a caller-sized global scratch.
// A pivot below rcond * max|A| marks the system singular (the reference's SVD rank test at the same
// rcond); singular systems and branch-boundary cases return NaN derivatives, as the reference does
// without a minimum-norm selection.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace me {

#define NCP_THREADS 64
#define NCP_MAXC 64  // contacts per environment

enum { NCP_OPEN = 0, NCP_STICK = 1, NCP_SLIP = 2 };
// per-contact boundary reasons (bits above the label) and per-environment status bits
enum { NCP_B_OPEN_ZERO_UN = 16, NCP_B_SLIP_ZERO_SPEED = 32, NCP_B_STICK_ON_CONE = 64, NCP_B_NEAR_ZERO_N = 128 };
enum { NCP_ST_SINGULAR = 1, NCP_ST_BOUNDARY = 2, NCP_ST_NONFINITE = 4 };

__device__ __forceinline__ double ncp_wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// one wave: M is n3 x W (row-major, leading dimension W), LU with partial pivoting on the first n3
// columns applied to all W; returns false if a pivot is below thresh
__device__ bool ncp_lu(double *M, int n3, int W, double thresh) {
    int t = threadIdx.x;
    __shared__ int piv_row;
    for (int k = 0; k < n3; ++k) {
        // first row of maximal |M[r][k]|, r >= k
        double best = -1.0;
        int br = n3;
        for (int r = k + t; r < n3; r += NCP_THREADS) {
            double a = fabs(M[(size_t)r * W + k]);
            if (a > best) { best = a; br = r; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            double ob = __shfl_xor(best, o);
            int orow = __shfl_xor(br, o);
            if (ob > best || (ob == best && orow < br)) { best = ob; br = orow; }
        }
        if (t == 0) piv_row = br;
        __syncthreads();
        int p = piv_row;
        if (!(best > thresh)) return false;  // uniform: every lane holds the reduced best
        if (p != k)
            for (int j = t; j < W; j += NCP_THREADS) {
                double a = M[(size_t)k * W + j];
                M[(size_t)k * W + j] = M[(size_t)p * W + j];
                M[(size_t)p * W + j] = a;
            }
        __syncthreads();
        double d = M[(size_t)k * W + k];
        int rows = n3 - k - 1, cols = W - k - 1;
        for (int idx = t; idx < rows * cols; idx += NCP_THREADS) {
            int r = k + 1 + idx / cols, j = k + 1 + idx % cols;
            double f = M[(size_t)r * W + k] / d;
            M[(size_t)r * W + j] = M[(size_t)r * W + j] - f * M[(size_t)k * W + j];
        }
        __syncthreads();
    }
    return true;
}

// maximum of v over the NT threads of the workgroup (every thread gets it); red: NT/64 doubles of LDS
template <int NT>
__device__ __forceinline__ double ncp_block_max(double v, double *red) {
    v = ncp_wave_max(v);
    if (NT == 64) return v;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int w = 1; w < NT / 64; ++w) m = fmax(m, red[w]);
    __syncthreads();
    return m;
}

// Set-up of one environment by its NT threads: u = G lam + b, finiteness of every input, labels
// (classify, ncp_ref.py:251-266) with their boundary reasons (ncp_ref.py:838-842), and the row-major
// augmented matrix M = [A | B_mu | B_b D] (n3 rows, leading dimension W = n3 + n + nd; D = identity
// when dirs is null) of the fixed-branch residual (ncp_ref.py:786-836). u (n3), lab (n), st and red
// (NT/64) are LDS; st collects NCP_ST_NONFINITE and NCP_ST_BOUNDARY.
template <int NT>
__device__ void ncp_setup(int n, int nd, const double *Ge, const double *be, const double *me, const double *le,
                          const double *De, const int *lab_in, double tol_lam, double tol_rel, double *M, int W,
                          double *u, int *lab, int *st, double *red) {
    const int t = threadIdx.x, n3 = 3 * n;
    if (t == 0) *st = 0;
    // u = G lam + b; finiteness of every input
    bool bad = false;
    for (int r = t; r < n3; r += NT) {
        double s = 0.0;
        for (int j = 0; j < n3; ++j) {
            double g = Ge[(size_t)r * n3 + j];
            bad = bad || !isfinite(g);
            s = s + g * le[j];
        }
        u[r] = s + be[r];
        bad = bad || !isfinite(be[r]) || !isfinite(le[r]);
    }
    for (int c = t; c < n; c += NT) bad = bad || !isfinite(me[c]);
    if (De)
        for (int idx = t; idx < n3 * nd; idx += NT) bad = bad || !isfinite(De[idx]);
    double sc = 0.0;
    for (int r = t; r < n3; r += NT) sc = fmax(sc, fabs(le[r]));
    sc = fmax(ncp_block_max<NT>(sc, red), 1.0e-30);
    if (bad) atomicOr(st, NCP_ST_NONFINITE);
    __syncthreads();
    // labels (classify, ncp_ref.py:251-266) and boundary reasons (ncp_ref.py:838-842)
    for (int c = t; c < n; c += NT) {
        double ln = le[3 * c], lt = sqrt(le[3 * c + 1] * le[3 * c + 1] + le[3 * c + 2] * le[3 * c + 2]);
        int L;
        if (lab_in) {
            L = lab_in[c] & 3;
        } else if (ln <= tol_lam + tol_rel * sc) {
            L = NCP_OPEN;
        } else if (lt >= me[c] * ln * (1.0 - tol_rel) - tol_lam) {
            L = NCP_SLIP;
        } else {
            L = NCP_STICK;
        }
        int flags = 0;
        if (L == NCP_OPEN && fabs(u[3 * c]) < 1e-9) flags |= NCP_B_OPEN_ZERO_UN;
        if (L == NCP_SLIP && sqrt(u[3 * c + 1] * u[3 * c + 1] + u[3 * c + 2] * u[3 * c + 2]) < 1e-14)
            flags |= NCP_B_SLIP_ZERO_SPEED;
        double scl = fmax(fabs(ln), 1e-12);
        if (L == NCP_STICK && fabs(lt - me[c] * ln) < 1e-7 * me[c] * scl) flags |= NCP_B_STICK_ON_CONE;
        if (L != NCP_OPEN && ln < 1e-9) flags |= NCP_B_NEAR_ZERO_N;
        lab[c] = L | flags;
        if (flags) atomicOr(st, NCP_ST_BOUNDARY);
    }
    __syncthreads();
    // assemble [A | B_mu | B_b D] row by row
    for (int r = t; r < n3; r += NT) {
        double *row = M + (size_t)r * W;
        for (int j = 0; j < W; ++j) row[j] = 0.0;
        int c = r / 3, k = r % 3, i = 3 * c, L = lab[c] & 3;
        if (L == NCP_OPEN) {
            row[r] = 1.0;
            continue;
        }
        if (L == NCP_STICK || k == 0) {  // stick rows and the slip normal row: u = G lam + b
            for (int j = 0; j < n3; ++j) row[j] = Ge[(size_t)r * n3 + j];
            if (De) for (int j = 0; j < nd; ++j) row[n3 + n + j] = De[(size_t)r * nd + j];
            else row[n3 + n + r] = 1.0;
            continue;
        }
        // slip tangent row q: lam_t + mu lam_n u_t/|u_t| = 0, linearised about (lam, u)
        double s1 = u[i + 1], s2 = u[i + 2];
        double ns = sqrt(s1 * s1 + s2 * s2);
        if (ns < 1e-14) ns = 1e-14;
        double h1 = s1 / ns, h2 = s2 / ns;
        int q = k - 1;
        double hq = q == 0 ? h1 : h2;
        double p0 = ((q == 0 ? 1.0 : 0.0) - hq * h1) / ns, p1 = ((q == 1 ? 1.0 : 0.0) - hq * h2) / ns;
        double ml = me[c] * le[i];
        for (int j = 0; j < n3; ++j)
            row[j] = ml * (p0 * Ge[(size_t)(i + 1) * n3 + j] + p1 * Ge[(size_t)(i + 2) * n3 + j]);
        row[i] = row[i] + me[c] * hq;
        row[r] = row[r] + 1.0;
        row[n3 + c] = le[i] * hq;
        if (De) {
            for (int j = 0; j < nd; ++j)
                row[n3 + n + j] = ml * (p0 * De[(size_t)(i + 1) * nd + j] + p1 * De[(size_t)(i + 2) * nd + j]);
        } else {
            row[n3 + n + i + 1] = ml * p0;
            row[n3 + n + i + 2] = ml * p1;
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(NCP_THREADS)
k_ncp_sensitivity(int E, int n, int nd, const double *G, const double *b, const double *mu, const double *lam,
                  const double *dirs, const int *lab_in, double tol_lam, double tol_rel, double rcond,
                  double *dmu, double *dlb, int *lab_out, int *status, double *gscratch, int use_lds) {
    extern __shared__ double sm[];
    int e = blockIdx.x, t = threadIdx.x;
    if (e >= E) return;
    const int n3 = 3 * n, W = n3 + n + nd;
    __shared__ double u[3 * NCP_MAXC];
    __shared__ int lab[NCP_MAXC];
    __shared__ int st;
    __shared__ double red[1];
    double *M = use_lds ? sm : gscratch + (size_t)e * n3 * W;
    ncp_setup<NCP_THREADS>(n, nd, G + (size_t)e * n3 * n3, b + (size_t)e * n3, mu + (size_t)e * n,
                           lam + (size_t)e * n3, dirs ? dirs + (size_t)e * n3 * nd : nullptr,
                           lab_in ? lab_in + (size_t)e * n : nullptr, tol_lam, tol_rel, M, W, u, lab, &st, red);
    // singularity threshold from max|A|
    double amax = 0.0;
    for (int idx = t; idx < n3 * n3; idx += NCP_THREADS) amax = fmax(amax, fabs(M[(size_t)(idx / n3) * W + idx % n3]));
    amax = ncp_wave_max(amax);
    bool ok = (st & NCP_ST_NONFINITE) == 0 && ncp_lu(M, n3, W, rcond * amax);
    if (!ok && t == 0 && !(st & NCP_ST_NONFINITE)) st |= NCP_ST_SINGULAR;
    __syncthreads();
    bool nanout = st != 0;
    // back substitution per right-hand-side column; X = -A^-1 [B_mu | B_b D]
    for (int j = n3 + t; j < W; j += NCP_THREADS) {
        if (!nanout) {
            for (int r = n3 - 1; r >= 0; --r) {
                double s = M[(size_t)r * W + j];
                for (int q = r + 1; q < n3; ++q) s = s - M[(size_t)r * W + q] * M[(size_t)q * W + j];
                M[(size_t)r * W + j] = s / M[(size_t)r * W + r];
            }
        }
        for (int r = 0; r < n3; ++r) {
            double x = nanout ? __builtin_nan("") : -M[(size_t)r * W + j];
            if (j < n3 + n) dmu[((size_t)e * n3 + r) * n + (j - n3)] = x;
            else dlb[((size_t)e * n3 + r) * nd + (j - n3 - n)] = x;
        }
    }
    for (int c = t; c < n; c += NCP_THREADS) lab_out[(size_t)e * n + c] = lab[c];
    if (t == 0) status[e] = st;
}

}  // namespace me
