// Fixed-branch NCP sensitivity with the reference's SVD rank guard (float64): the numerical rank of
// the active-set Jacobian A at rcond (contact_rank.py:13-30), its right and left nullspaces, the
// response X = -A^+ [B_mu | B_b D] guarded as contact_rank.py:33-70 does (NaN for ambiguous columns
// unless the minimum-norm linear response is selected, NaN for incompatible columns in either mode),
// and the optional observable rows Q: Q X only for rows with Q N = 0 and compatible columns. The
// per-environment set-up (classify, active system) is ncp_setup, as in k_ncp_sensitivity.
//
// One workgroup of NT threads per environment. The SVD is the one-sided Jacobi (Hestenes) method:
// plane rotations of column pairs of A, taken in round-robin order so that a round's pairs are
// disjoint and its rotations run in parallel (one wave per pair), until every pair of columns is
// orthogonal to 1e-15 relative; the rotations accumulated in V give A V = W with orthogonal columns,
// singular values |W_i| and left vectors W_i / |W_i|. The same method on A^T gives the full left
// basis, whose trailing columns (by descending singular value) span the left nullspace. All
// matrices of an environment live in a caller-sized global scratch.
#pragma once
#include "ncp.hpp"

namespace me {

#define NCPS_MAXC 256     // contacts per environment (3n <= 768)
#define NCPS_SWEEPS 40    // Jacobi sweeps at most (quadratic convergence: about 6-10 are taken)

// outputs of k_ncp_svd, per environment e (n3 = 3n, R = n + m response columns, K observable rows)
struct NcpSvdOut {
    double *X;        // [E][n3][R] guarded response: dlam/dmu (first n columns), dlam/db D (next m)
    double *S;        // [E][n3] singular values of A, descending
    double *V;        // [E][n3][n3] right singular vectors, column k for S[k] (row-major)
    double *U;        // [E][n3][n3] left singular vectors of the A^T run, by its descending order
    int *rank;        // [E] numerical rank at rcond * S[0]
    int *compat;      // [E][R] column compatible (|L^T B_j| <= tol |B_j| on max-scaled columns)
    int *finite;      // [E][R] unguarded response of the column finite
    double *resid;    // [E][R] compatibility residual |L^T B_j| (unscaled)
    double *Y;        // [E][K][R] observable responses Q X (NaN where not valid)
    int *valid;       // [E][K][R]
    int *invariant;   // [E][K] Q_k N = 0 at tol
    double *leak;     // [E][K] |Q_k N| (unscaled)
    int *labels;      // [E][n] label | boundary bits (as k_ncp_sensitivity)
    int *status;      // [E] NCP_ST_* bits (NCP_ST_SINGULAR from the SVD rank)
};

// scratch doubles per environment
__host__ __device__ inline size_t ncp_svd_scratch(int n, int m) {
    size_t n3 = 3 * (size_t)n, W = n3 + n + m;
    return n3 * W + 4 * n3 * n3 + 2 * n3 + n3 * (n + (size_t)m);
}

template <int NT>
__device__ __forceinline__ double ncp_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One-sided Jacobi on the n3 columns of X (column-major, column j at X + j * n3), rotations
// accumulated into Vc (column-major, identity on entry). Every thread of the workgroup calls it.
template <int NT>
__device__ void ncp_jacobi(double *X, double *Vc, int n3, int *flag) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6, NW = NT / 64;
    const int np = n3 + (n3 & 1), half = np / 2, rounds = np - 1;
    for (int sweep = 0; sweep < NCPS_SWEEPS; ++sweep) {
        if (t == 0) *flag = 0;
        __syncthreads();
        for (int rd = 0; rd < rounds; ++rd) {
            // round-robin (circle) pairing: position np-1 is fixed, the others rotate by rd
            for (int k = wv; k < half; k += NW) {
                int i, j;
                if (k == 0) { i = rd; j = np - 1; }
                else { i = (rd + k) % rounds; j = (rd - k + rounds) % rounds; }
                if (i > j) { int s = i; i = j; j = s; }
                if (j >= n3) continue;  // the padding column of an odd n3
                double *xi = X + (size_t)i * n3, *xj = X + (size_t)j * n3;
                double a = 0.0, b = 0.0, g = 0.0;
                for (int r = lane; r < n3; r += 64) {
                    double p = xi[r], q = xj[r];
                    a += p * p; b += q * q; g += p * q;
                }
                a = ncp_wave_sum<NT>(a); b = ncp_wave_sum<NT>(b); g = ncp_wave_sum<NT>(g);
                if (!(fabs(g) > 1e-15 * sqrt(a * b)) || g == 0.0) continue;
                double zeta = (b - a) / (2.0 * g);
                double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
                for (int r = lane; r < n3; r += 64) {
                    double p = xi[r], q = xj[r];
                    xi[r] = c * p - s * q;
                    xj[r] = s * p + c * q;
                }
                double *vi = Vc + (size_t)i * n3, *vj = Vc + (size_t)j * n3;
                for (int r = lane; r < n3; r += 64) {
                    double p = vi[r], q = vj[r];
                    vi[r] = c * p - s * q;
                    vj[r] = s * p + c * q;
                }
                if (lane == 0) *flag = 1;
            }
            __syncthreads();
        }
        if (*flag == 0) break;
        __syncthreads();
    }
    __syncthreads();
}

// column norms of X into nrm and the descending order (ties: lower index first) into ord
template <int NT>
__device__ void ncp_order(const double *X, int n3, double *nrm, int *ord) {
    const int t = threadIdx.x;
    for (int j = t; j < n3; j += NT) {
        double s = 0.0;
        for (int r = 0; r < n3; ++r) s += X[(size_t)j * n3 + r] * X[(size_t)j * n3 + r];
        nrm[j] = sqrt(s);
    }
    __syncthreads();
    for (int j = t; j < n3; j += NT) {
        int rk = 0;
        for (int k = 0; k < n3; ++k) rk += nrm[k] > nrm[j] || (nrm[k] == nrm[j] && k < j);
        ord[rk] = j;
    }
    __syncthreads();
}

template <int NT>
__global__ void __launch_bounds__(NT)
k_ncp_svd(int E, int n, int m, int K, const double *G, const double *b, const double *mu, const double *lam,
          const double *dirs, const int *lab_in, const double *Q, int q_shared, double tol_lam, double tol_rel,
          double rcond, double tol, int minnorm, NcpSvdOut O, double *gscratch) {
    const int e = blockIdx.x, t = threadIdx.x;
    if (e >= E) return;
    const int n3 = 3 * n, R = n + m, W = n3 + R;
    __shared__ double u[3 * NCPS_MAXC];
    __shared__ int lab[NCPS_MAXC];
    __shared__ int ordA[3 * NCPS_MAXC], ordT[3 * NCPS_MAXC];
    __shared__ double nrmA[3 * NCPS_MAXC], nrmT[3 * NCPS_MAXC];
    __shared__ int st, flag, rank_s;
    __shared__ double red[NT / 64];
    double *M = gscratch + (size_t)e * ncp_svd_scratch(n, m);
    double *XA = M + (size_t)n3 * W, *VA = XA + (size_t)n3 * n3, *XT = VA + (size_t)n3 * n3,
           *VT = XT + (size_t)n3 * n3, *P = VT + (size_t)n3 * n3 + 2 * (size_t)n3;
    ncp_setup<NT>(n, m, G + (size_t)e * n3 * n3, b + (size_t)e * n3, mu + (size_t)e * n, lam + (size_t)e * n3,
                  dirs ? dirs + (size_t)e * n3 * m : nullptr, lab_in ? lab_in + (size_t)e * n : nullptr, tol_lam,
                  tol_rel, M, W, u, lab, &st, red);
    const bool nonfinite = (st & NCP_ST_NONFINITE) != 0;
    // A (column-major) for the run on A, A^T (column-major: column i = row i of A) for the run on A^T
    for (int idx = t; idx < n3 * n3; idx += NT) {
        int r = idx / n3, j = idx % n3;
        double a = nonfinite ? 0.0 : M[(size_t)r * W + j];
        XA[(size_t)j * n3 + r] = a;
        XT[(size_t)r * n3 + j] = a;
        VA[idx] = (r == j) ? 1.0 : 0.0;
        VT[idx] = (r == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    ncp_jacobi<NT>(XA, VA, n3, &flag);
    ncp_jacobi<NT>(XT, VT, n3, &flag);
    ncp_order<NT>(XA, n3, nrmA, ordA);
    ncp_order<NT>(XT, n3, nrmT, ordT);
    if (t == 0) {
        double cut = n3 > 0 ? rcond * nrmA[ordA[0]] : 0.0;
        int rk = 0;
        for (int k = 0; k < n3; ++k) rk += nrmA[ordA[k]] > cut;
        rank_s = rk;
        if (rk < n3 && !nonfinite) st |= NCP_ST_SINGULAR;
    }
    __syncthreads();
    const int rank = rank_s;
    const bool singular = rank < n3, boundary = (st & NCP_ST_BOUNDARY) != 0;
    double *Se = O.S + (size_t)e * n3, *Ve = O.V + (size_t)e * n3 * n3, *Ue = O.U + (size_t)e * n3 * n3;
    for (int idx = t; idx < n3 * n3; idx += NT) {
        int r = idx / n3, k = idx % n3;
        Ve[idx] = VA[(size_t)ordA[k] * n3 + r];
        Ue[idx] = VT[(size_t)ordT[k] * n3 + r];
    }
    for (int k = t; k < n3; k += NT) Se[k] = nrmA[ordA[k]];
    // response columns: c_k = U_k^T B / S_k for k < rank (U_k = W_k / S_k of the run on A), the
    // unguarded response P = -V_r c, and the compatibility residual through the left nullspace
    for (int j = t; j < R; j += NT) {
        const double *Bj = M + n3 + j;  // column j of [B_mu | B_b D], stride W
        double bs = 0.0;
        for (int r = 0; r < n3; ++r) bs = fmax(bs, fabs(Bj[(size_t)r * W]));
        const double sc = bs > 0.0 ? bs : 1.0;
        double nb2 = 0.0, mis2 = 0.0;
        for (int r = 0; r < n3; ++r) nb2 += (Bj[(size_t)r * W] / sc) * (Bj[(size_t)r * W] / sc);
        for (int k = rank; k < n3; ++k) {
            const double *Lk = VT + (size_t)ordT[k] * n3;
            double d = 0.0;
            for (int r = 0; r < n3; ++r) d += Lk[r] * (Bj[(size_t)r * W] / sc);
            mis2 += d * d;
        }
        const double relmis = sqrt(mis2);
        const bool comp = relmis <= tol * sqrt(nb2);
        double *Pj = P + (size_t)j;  // P is n3 x R row-major
        for (int r = 0; r < n3; ++r) Pj[(size_t)r * R] = 0.0;
        for (int k = 0; k < rank; ++k) {
            const int ck = ordA[k];
            const double sk = nrmA[ck];
            const double *Wk = XA + (size_t)ck * n3, *Vk = VA + (size_t)ck * n3;
            double d = 0.0;
            for (int r = 0; r < n3; ++r) d += (Wk[r] / sk) * Bj[(size_t)r * W];
            const double ckv = d / sk;
            for (int r = 0; r < n3; ++r) Pj[(size_t)r * R] -= Vk[r] * ckv;
        }
        bool fin = true;
        for (int r = 0; r < n3; ++r) fin = fin && isfinite(Pj[(size_t)r * R]);
        const bool keep = fin && !boundary && !nonfinite && (!singular || (minnorm && comp));
        double *Xe = O.X + (size_t)e * n3 * R;
        for (int r = 0; r < n3; ++r) Xe[(size_t)r * R + j] = keep ? Pj[(size_t)r * R] : __builtin_nan("");
        O.compat[(size_t)e * R + j] = comp;
        O.finite[(size_t)e * R + j] = fin && !nonfinite;
        O.resid[(size_t)e * R + j] = relmis * bs;
    }
    __syncthreads();
    // observable rows: leakage |Q_k N| on max-scaled rows, Y = Q P where valid
    if (K > 0) {
        const double *Qe = Q + (q_shared ? 0 : (size_t)e * K * n3);
        for (int k = t; k < K; k += NT) {
            const double *Qk = Qe + (size_t)k * n3;
            double qs = 0.0;
            for (int r = 0; r < n3; ++r) qs = fmax(qs, fabs(Qk[r]));
            const double sc = qs > 0.0 ? qs : 1.0;
            double nq2 = 0.0, lk2 = 0.0;
            for (int r = 0; r < n3; ++r) nq2 += (Qk[r] / sc) * (Qk[r] / sc);
            for (int z = rank; z < n3; ++z) {
                const double *Nz = VA + (size_t)ordA[z] * n3;
                double d = 0.0;
                for (int r = 0; r < n3; ++r) d += (Qk[r] / sc) * Nz[r];
                lk2 += d * d;
            }
            const double rl = sqrt(lk2);
            const bool inv = rl <= tol * sqrt(nq2);
            O.invariant[(size_t)e * K + k] = inv;
            O.leak[(size_t)e * K + k] = rl * qs;
            for (int j = 0; j < R; ++j) {
                double y = 0.0;
                for (int r = 0; r < n3; ++r) y += Qk[r] * P[(size_t)r * R + j];
                const bool ok = inv && O.compat[(size_t)e * R + j] && isfinite(y) && !boundary && !nonfinite;
                O.Y[((size_t)e * K + k) * R + j] = ok ? y : __builtin_nan("");
                O.valid[((size_t)e * K + k) * R + j] = ok;
            }
        }
    }
    for (int c = t; c < n; c += NT) O.labels[(size_t)e * n + c] = lab[c];
    if (t == 0) {
        O.rank[e] = rank;
        O.status[e] = st;
    }
}

}  // namespace me
