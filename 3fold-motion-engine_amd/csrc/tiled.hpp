// Island-tiled coloured solve: one launch sweeps every colour of every island with the island's
// bodies resident in LDS and a workgroup barrier between colours.
//
// Islands are the connected components of the pair graph over dynamic bodies (static/kinematic
// bodies do not merge islands, as the reference's partition_pairs,
// contact_engine_gpu_island_solve.py:7-33). Pairs of different islands never share a dynamic body,
// so the JP rounds of one island, and its colour sweeps, do not depend on any other island:
// solving island by island in colour order gives the global colour-sweep result bit for bit.
//
// Per substep (no host round trip):
//   union-find (root = minimum body index of the island, so labels are deterministic)
//   -> bodies and pairs grouped by island (island order = root order)
//   -> tiles: tile t holds the islands whose first body falls in [t*TB, (t+1)*TB)
//   -> k_t_color: JP rounds per tile in LDS (skipped on a ColorCache hit)
//   -> k_t_plan: split of the JP classes into n_colors classes
//   -> k_t_solve: bodies to LDS, pairs counting-sorted by colour in LDS, warm apply + vit velocity
//      sweeps + pit position sweeps over the colours, write-back and integration.
// Tiles larger than the LDS capacity run the same sweeps on global memory.
#pragma once
#include <climits>

#include "kernels.hpp"

namespace me {

#define T_MAXB 512   // bodies of a tile resident in LDS
#define T_MAXP 1024  // pairs of a tile sorted by colour in LDS

struct TiledBuffers {
    int *parent, *root, *proot;
    int *cnt_b, *cnt_p, *fill_b, *fill_p, *boff, *poff;
    int *bsorted, *bslot, *psorted;
    int *tb0, *tb1, *tp0, *tp1;
    int *ccount_jp, *csub, *cbase;
};

struct TiledArgs {
    int N, capC, capP, solver, vit, pit, warm, n_colors, tile_bodies;
    float t2d, dt;
    Bodies b;
    float *invIw, *pv, *po;
    Contacts Cc;
    Pairs Pp;
    int *cache_a, *cache_b, *cache_col;
    float *cache_invM;
    int64_t *wkey;
    float *wjn, *wjt1, *wjt2;
    int *istat, *bsum;
    int nbsum;
    unsigned long long *claim;
};

template <class BumpT>
inline void tiled_layout(TiledBuffers &t, BumpT &B, int N, int capP, int capC) {
    (void)capC;
    size_t n = (size_t)N, p = (size_t)capP;
    t.parent = B.template take<int>(n);
    t.root = B.template take<int>(n);
    t.proot = B.template take<int>(p);
    t.cnt_b = B.template take<int>(n);
    t.cnt_p = B.template take<int>(n);
    t.fill_b = B.template take<int>(n);
    t.fill_p = B.template take<int>(n);
    t.boff = B.template take<int>(n);
    t.poff = B.template take<int>(n);
    t.bsorted = B.template take<int>(n);
    t.bslot = B.template take<int>(n);
    t.psorted = B.template take<int>(p);
    t.tb0 = B.template take<int>(n);
    t.tb1 = B.template take<int>(n);
    t.tp0 = B.template take<int>(n);
    t.tp1 = B.template take<int>(n);
    t.ccount_jp = B.template take<int>(ME_NCOLMAX + 1);
    t.csub = B.template take<int>(ME_NCOLMAX + 1);
    t.cbase = B.template take<int>(ME_NCOLMAX + 2);
}

// IS_MISMATCH after k_t_mode: 0 = ColorCache hit, 1 = colour this substep, 2 = no contacts
__global__ void k_t_init(int N, TiledBuffers T, int *istat) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        T.parent[i] = i;
        T.cnt_b[i] = 0; T.cnt_p[i] = 0; T.fill_b[i] = 0; T.fill_p[i] = 0;
        T.tb0[i] = INT_MAX; T.tb1[i] = -1; T.tp0[i] = INT_MAX; T.tp1[i] = -1;
    }
    if (i <= ME_NCOLMAX) T.ccount_jp[i] = 0;
    if (i == 0) {
        istat[IS_NCOLJP] = 0; istat[IS_ISL] = 0; istat[IS_TILES] = 0; istat[IS_MAXTB] = 0;
        istat[IS_MISMATCH] = 0;
    }
}

__global__ void k_t_mode(int *istat, int cc_mode) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (istat[IS_C] <= 0 || istat[IS_P] <= 0) { istat[IS_MISMATCH] = 2; istat[IS_NCOL] = 0; return; }
    if (!cc_mode) { istat[IS_MISMATCH] = 1; return; }
    if (istat[IS_MISMATCH] == 0) {
        istat[IS_HITS] += 1;
        istat[IS_NCOL] = istat[IS_CACHED_NCOL];
    } else {
        istat[IS_MISMATCH] = 1;
        istat[IS_MISSES] += 1;
    }
}

__global__ void k_t_cc_restore(int capP, const int *istat, const int *ccol, int *col) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (istat[IS_MISMATCH] != 0) return;
    if (i < istat[IS_P] && i < capP) col[i] = ccol[i];
}

__global__ void k_t_cc_store(int N, int capP, int *istat, const int *pbi, const int *pbj, const int *col,
                             const float *invM, int *ca, int *cb, int *ccol, float *cinvM) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (istat[IS_MISMATCH] != 1) return;
    int P = istat[IS_P];
    if (i < P && i < capP) { ca[i] = pbi[i]; cb[i] = pbj[i]; ccol[i] = col[i]; }
    if (i < N) cinvM[i] = invM[i];
    if (i == 0) { istat[IS_CACHED_P] = P; istat[IS_CACHED_NCOL] = istat[IS_NCOL]; }
}

// parent chains only ever point to smaller indices, so the walk ends within N steps
__device__ __forceinline__ int uf_find(const int *parent, int x) {
    const volatile int *par = parent;
    int p = par[x];
    while (p != x) { x = p; p = par[x]; }
    return x;
}

// lock-free union: link the larger root under the smaller one (compare-and-swap on the root)
__global__ void k_t_union(int capP, int *istat, const int *pbi, const int *pbj, const float *invM, int *parent) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= istat[IS_P] || p >= capP) return;
    int a = pbi[p], b = pbj[p];
    if (b < 0 || !(invM[a] > 0.0f) || !(invM[b] > 0.0f)) return;
    for (int it = 0; it < (1 << 20); ++it) {
        int ra = uf_find(parent, a), rb = uf_find(parent, b);
        if (ra == rb) return;
        int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
        if (atomicCAS(&parent[hi], hi, lo) == hi) return;
    }
    atomicOr(&istat[IS_OVF], OVF_TILE);
}

__global__ void k_t_flatten(int N, const float *invM, TiledBuffers T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (invM[i] > 0.0f) {
        int r = uf_find(T.parent, i);
        T.root[i] = r;
        atomicAdd(&T.cnt_b[r], 1);
    } else {
        T.root[i] = -1;
    }
}

__global__ void k_t_pair_root(int capP, const int *istat, const int *pbi, const int *pbj, const float *invM,
                              TiledBuffers T) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= istat[IS_P] || p >= capP) return;
    int a = pbi[p], b = pbj[p];
    int r = invM[a] > 0.0f ? T.root[a] : ((b >= 0 && invM[b] > 0.0f) ? T.root[b] : -1);
    T.proot[p] = r;
    if (r >= 0) atomicAdd(&T.cnt_p[r], 1);
}

__global__ void k_t_scatter_b(int N, TiledBuffers T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int r = T.root[i];
    if (r < 0) return;
    int pos = T.boff[r] + atomicAdd(&T.fill_b[r], 1);
    T.bsorted[pos] = i;
    T.bslot[i] = pos;
}

__global__ void k_t_scatter_p(int capP, const int *istat, TiledBuffers T) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= istat[IS_P] || p >= capP) return;
    int r = T.proot[p];
    if (r < 0) return;
    T.psorted[T.poff[r] + atomicAdd(&T.fill_p[r], 1)] = p;
}

__global__ void k_t_tiles(int N, int TB, int *istat, TiledBuffers T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || T.cnt_b[i] == 0) return;
    int t = T.boff[i] / TB;
    atomicMin(&T.tb0[t], T.boff[i]);
    atomicMax(&T.tb1[t], T.boff[i] + T.cnt_b[i]);
    atomicMin(&T.tp0[t], T.poff[i]);
    atomicMax(&T.tp1[t], T.poff[i] + T.cnt_p[i]);
    atomicAdd(&istat[IS_ISL], 1);
}

// JP rounds of one tile (k_claim / k_win of contact_engine_gpu_colored.py:151-185, restricted to the
// tile's islands). LDS mode: claims by the body's slot in the tile, colours in LDS. Oversized
// tiles: claims in the global claim array by body id (tiles never share a dynamic body).
template <int WG>
__global__ void __launch_bounds__(WG) k_t_color(TiledArgs A, TiledBuffers T) {
    if (A.istat[IS_MISMATCH] != 1) return;
    int t = blockIdx.x;
    int b0 = T.tb0[t];
    if (b0 == INT_MAX) return;
    int nb = T.tb1[t] - b0, p0 = T.tp0[t], np = T.tp1[t] - T.tp0[t];
    if (np <= 0) return;
    __shared__ unsigned long long sclaim[T_MAXB];
    __shared__ int colL[T_MAXP];
    __shared__ int left;
    const bool lds = nb <= T_MAXB && np <= T_MAXP;
    unsigned long long *claim = lds ? sclaim : A.claim;
    const float *invM = A.b.invM;
    int *gcol = A.Pp.col;
    auto slot = [&](int body) { return lds ? T.bslot[body] - b0 : body; };
    auto getc = [&](int k, int p) { return lds ? colL[k] : gcol[p]; };
    for (int k = threadIdx.x; k < np; k += WG) {
        if (lds) colL[k] = -1;
        else gcol[T.psorted[p0 + k]] = -1;
    }
    int nround = 0;
    for (int cur = 0; cur <= np; ++cur) {
        for (int j = threadIdx.x; j < nb; j += WG) claim[slot(T.bsorted[b0 + j])] = ~0ull;
        if (threadIdx.x == 0) left = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < np; k += WG) {
            int p = T.psorted[p0 + k];
            if (getc(k, p) >= 0) continue;
            unsigned long long key = (unsigned long long)jp_key(p);
            int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
            if (invM[a] > 0.0f) atomicMin(&claim[slot(a)], key);
            if (b >= 0 && invM[b] > 0.0f) atomicMin(&claim[slot(b)], key);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < np; k += WG) {
            int p = T.psorted[p0 + k];
            if (getc(k, p) >= 0) continue;
            unsigned long long key = (unsigned long long)jp_key(p);
            int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
            bool ok = true;
            if (invM[a] > 0.0f && claim[slot(a)] != key) ok = false;
            if (b >= 0 && invM[b] > 0.0f && claim[slot(b)] != key) ok = false;
            if (ok) {
                if (lds) colL[k] = cur;
                else gcol[p] = cur;
            } else {
                atomicAdd(&left, 1);
            }
        }
        __syncthreads();
        int l = left;
        __syncthreads();
        if (l == 0) { nround = cur + 1; break; }
    }
    bool ovf = false;
    for (int k = threadIdx.x; k < np; k += WG) {
        int p = T.psorted[p0 + k];
        int c = getc(k, p);
        if (lds) gcol[p] = c;
        if (c >= 0 && c < ME_NCOLMAX) atomicAdd(&T.ccount_jp[c], 1);
        else ovf = true;
    }
    if (ovf) atomicOr(&A.istat[IS_OVF], OVF_COLORS);
    if (threadIdx.x == 0) atomicMax(&A.istat[IS_NCOLJP], nround);
}

// split of the JP classes (same rule as k_color_split_plan)
__global__ void k_t_plan(int *istat, int K, TiledBuffers T) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (istat[IS_MISMATCH] != 1) return;
    int nc = istat[IS_NCOLJP];
    if (nc > ME_NCOLMAX) { nc = ME_NCOLMAX; atomicOr(&istat[IS_OVF], OVF_COLORS); }
    for (int c = 0; c < nc; ++c) T.csub[c] = 1;
    int tot = nc;
    if (K > nc && nc > 0) {
        for (int e = nc; e < K && e < ME_NCOLMAX; ++e) {
            int bc = 0;
            for (int c = 1; c < nc; ++c)
                if ((long long)T.ccount_jp[c] * T.csub[bc] > (long long)T.ccount_jp[bc] * T.csub[c]) bc = c;
            T.csub[bc]++;
        }
        tot = 0;
        for (int c = 0; c < nc; ++c) tot += T.csub[c];
    }
    T.cbase[0] = 0;
    for (int c = 0; c < nc; ++c) T.cbase[c + 1] = T.cbase[c] + T.csub[c];
    istat[IS_NCOL] = tot;
}

// the sweeps of one tile; SOLVER: ME_SOLVER_PAIR_CHUNKS (slip frame, no warm apply) or
// ME_SOLVER_COLOR_CACHE (fixed frame, warm apply)
template <int WG, int SOLVER>
__global__ void __launch_bounds__(WG) k_t_solve(TiledArgs A, TiledBuffers T) {
    int t = blockIdx.x;
    int b0 = T.tb0[t];
    if (b0 == INT_MAX) return;
    int nb = T.tb1[t] - b0, p0 = T.tp0[t], np = T.tp1[t] - T.tp0[t];
    if (np < 0) np = 0;
    const int mode = A.istat[IS_MISMATCH];
    int ncol = A.istat[IS_NCOL];
    if (ncol > ME_NCOLMAX) ncol = ME_NCOLMAX;
    __shared__ float sv[3 * T_MAXB], sw[3 * T_MAXB], spv[3 * T_MAXB], spo[3 * T_MAXB], sx[3 * T_MAXB];
    __shared__ float sinvM[T_MAXB], sI[9 * T_MAXB];
    __shared__ int plist[T_MAXP], pcol[T_MAXP];
    __shared__ int cst[ME_NCOLMAX + 1], cfl[ME_NCOLMAX + 1];
    __shared__ int maxc;
    const bool lds = nb <= T_MAXB && np <= T_MAXP;
    if (threadIdx.x == 0) {
        atomicAdd(&A.istat[IS_TILES], 1);
        atomicMax(&A.istat[IS_MAXTB], nb);
        maxc = 0;
    }
    BodyView G;
    G.v = A.b.v; G.w = A.b.w; G.pv = A.pv; G.po = A.po; G.x = A.b.x; G.invM = A.b.invM; G.invIw = A.invIw;
    BodyView L;
    L.v = sv; L.w = sw; L.pv = spv; L.po = spo; L.x = sx; L.invM = sinvM; L.invIw = sI;
    ContactView Cv;
    Cv.pA = A.Cc.pA; Cv.pB = A.Cc.pB; Cv.n = A.Cc.n; Cv.pen = A.Cc.pen; Cv.mu = A.Cc.mu;
    Cv.jn = A.Cc.jn; Cv.jt1 = A.Cc.jt1; Cv.jt2 = A.Cc.jt2; Cv.jp = A.Cc.jp;
    if (lds) {
        for (int j = threadIdx.x; j < nb; j += WG) {
            int i = T.bsorted[b0 + j];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                sv[3 * j + k] = A.b.v[3 * i + k];
                sw[3 * j + k] = A.b.w[3 * i + k];
                sx[3 * j + k] = A.b.x[3 * i + k];
                spv[3 * j + k] = 0.0f;
                spo[3 * j + k] = 0.0f;
            }
            sinvM[j] = A.b.invM[i];
#pragma unroll
            for (int k = 0; k < 9; ++k) sI[9 * j + k] = A.invIw[9 * i + k];
        }
    }
    // final colours (split on a fresh colouring) and, in LDS mode, a counting sort by colour
    for (int c = threadIdx.x; c <= ncol; c += WG) { cst[c] = 0; cfl[c] = 0; }
    __syncthreads();
    for (int k = threadIdx.x; k < np; k += WG) {
        int p = T.psorted[p0 + k];
        int c = A.Pp.col[p];
        if (mode == 1 && c >= 0 && c < ME_NCOLMAX) {
            c = T.cbase[c] + p % T.csub[c];
            A.Pp.col[p] = c;
        }
        if (c < 0 || c >= ncol) c = ncol > 0 ? ncol - 1 : 0;  // only after a colour overflow (flagged)
        if (lds) {
            pcol[k] = c;
            atomicAdd(&cst[c], 1);
        }
        atomicMax(&maxc, c + 1);
    }
    __syncthreads();
    if (lds && threadIdx.x == 0) {
        int run = 0;
        for (int c = 0; c < maxc; ++c) {
            int n = cst[c];
            cst[c] = run;
            run += n;
        }
        cst[maxc] = run;
    }
    __syncthreads();
    if (lds)
        for (int k = threadIdx.x; k < np; k += WG) {
            int c = pcol[k];
            plist[cst[c] + atomicAdd(&cfl[c], 1)] = T.psorted[p0 + k];
        }
    __syncthreads();
    const int nc = maxc;
    const int *pbi = A.Pp.pbi, *pbj = A.Pp.pbj, *pst = A.Pp.pstart, *pcn = A.Pp.pcount;
    const float *invM = A.b.invM;
    // one colour step of one pass: 0 = warm apply, 1 = velocity, 2 = position
    auto solve_pair = [&](int p, int pass) {
        int a = pbi[p], b = pbj[p];
        bool dA = lds && invM[a] > 0.0f, dB = lds && b >= 0 && invM[b] > 0.0f;
        const BodyView &VA = dA ? L : G;
        const BodyView &VB = dB ? L : G;
        int ia = dA ? T.bslot[a] - b0 : a;
        int ib = b < 0 ? -1 : (dB ? T.bslot[b] - b0 : b);
        if (pass == 0) apply_warm_pair(VA, VB, Cv, ia, ib, pst[p], pcn[p]);
        else if (pass == 1) chunk_vel<SOLVER == ME_SOLVER_COLOR_CACHE>(VA, VB, Cv, ia, ib, pst[p], pcn[p],
                                                                      SOLVER == ME_SOLVER_COLOR_CACHE ? 1.0f : A.t2d);
        else chunk_pos(VA, VB, Cv, ia, ib, pst[p], pcn[p], A.dt);
    };
    auto sweep = [&](int pass) {
        for (int c = 0; c < nc; ++c) {
            if (lds) {
                for (int k = cst[c] + threadIdx.x; k < cst[c + 1]; k += WG) solve_pair(plist[k], pass);
            } else {
                for (int k = threadIdx.x; k < np; k += WG) {
                    int p = T.psorted[p0 + k];
                    int pc = A.Pp.col[p];
                    if (pc < 0 || pc >= ncol) pc = ncol > 0 ? ncol - 1 : 0;
                    if (pc == c) solve_pair(p, pass);
                }
            }
            __syncthreads();
        }
    };
    if (np > 0) {
        if (SOLVER == ME_SOLVER_COLOR_CACHE) sweep(0);
        for (int it = 0; it < A.vit; ++it) sweep(1);
        for (int it = 0; it < A.pit; ++it) sweep(2);
    }
    __syncthreads();
    // write back and integrate (k_integ, contact_engine_gpu.py:172-180)
    for (int j = threadIdx.x; j < nb; j += WG) {
        int i = T.bsorted[b0 + j];
        f3 v, w, pv, po;
        if (lds) {
            v = ld3(sv, j); w = ld3(sw, j); pv = ld3(spv, j); po = ld3(spo, j);
            st3(A.b.v, i, v); st3(A.b.w, i, w); st3(A.pv, i, pv); st3(A.po, i, po);
        } else {
            v = ld3(A.b.v, i); w = ld3(A.b.w, i); pv = ld3(A.pv, i); po = ld3(A.po, i);
        }
        if (A.b.invM[i] > 0.0f) integrate_body(A.b.x, A.b.q, v, w, pv, po, i, A.dt);
    }
}

inline int tiled_scan(const TiledArgs &A, const int *in, int *out, int n, int *total, hipStream_t s) {
    int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb < 1) nb = 1;
    if (nb > A.nbsum) return ME_ERUNTIME;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(ME_THREADS), 0, s, in, nullptr, n, n, A.bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, A.bsum, nb, total);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(ME_THREADS), 0, s, in, out, nullptr, n, n, A.bsum);
    return ME_OK;
}

inline int tblk(long long n) { return (int)((n + ME_THREADS - 1) / ME_THREADS); }

template <int WG>
inline void launch_tiles(const TiledArgs &A, TiledBuffers &T, int ntile, hipStream_t s) {
    hipLaunchKernelGGL(k_t_color<WG>, dim3(ntile), dim3(WG), 0, s, A, T);
    hipLaunchKernelGGL(k_t_plan, dim3(1), dim3(1), 0, s, A.istat, A.n_colors, T);
    if (A.solver == ME_SOLVER_COLOR_CACHE)
        hipLaunchKernelGGL((k_t_solve<WG, ME_SOLVER_COLOR_CACHE>), dim3(ntile), dim3(WG), 0, s, A, T);
    else
        hipLaunchKernelGGL((k_t_solve<WG, ME_SOLVER_PAIR_CHUNKS>), dim3(ntile), dim3(WG), 0, s, A, T);
}

// the whole coloured solve + integration of one substep
inline int tiled_solve(TiledBuffers &T, const TiledArgs &A, hipStream_t s) {
    int N = A.N, capP = A.capP, capC = A.capC;
    int TB = A.tile_bodies > 0 ? A.tile_bodies : 64;
    if (TB > T_MAXB) TB = T_MAXB;
    int ntile = (N + TB - 1) / TB;
    bool cc = A.solver == ME_SOLVER_COLOR_CACHE;
    int g = capP > N ? capP : N;
    hipLaunchKernelGGL(k_zero_pvpo, dim3(tblk(N)), dim3(ME_THREADS), 0, s, N, A.pv, A.po);
    hipLaunchKernelGGL(k_warm_lookup, dim3(tblk(capC)), dim3(ME_THREADS), 0, s, capC, A.istat, A.warm, A.Cc.key,
                       A.wkey, A.wjn, A.wjt1, A.wjt2, A.Cc.jn, A.Cc.jt1, A.Cc.jt2, A.Cc.jp);
    hipLaunchKernelGGL(k_t_init, dim3(tblk(N > ME_NCOLMAX + 1 ? N : ME_NCOLMAX + 1)), dim3(ME_THREADS), 0, s, N, T,
                       A.istat);
    if (cc)
        hipLaunchKernelGGL(k_cc_compare, dim3(tblk(g)), dim3(ME_THREADS), 0, s, N, A.istat, A.Pp.pbi, A.Pp.pbj,
                           A.cache_a, A.cache_b, A.b.invM, A.cache_invM);
    hipLaunchKernelGGL(k_t_mode, dim3(1), dim3(1), 0, s, A.istat, cc ? 1 : 0);
    if (cc) hipLaunchKernelGGL(k_t_cc_restore, dim3(tblk(capP)), dim3(ME_THREADS), 0, s, capP, A.istat, A.cache_col,
                               A.Pp.col);
    hipLaunchKernelGGL(k_t_union, dim3(tblk(capP)), dim3(ME_THREADS), 0, s, capP, A.istat, A.Pp.pbi, A.Pp.pbj, A.b.invM,
                       T.parent);
    hipLaunchKernelGGL(k_t_flatten, dim3(tblk(N)), dim3(ME_THREADS), 0, s, N, A.b.invM, T);
    hipLaunchKernelGGL(k_t_pair_root, dim3(tblk(capP)), dim3(ME_THREADS), 0, s, capP, A.istat, A.Pp.pbi, A.Pp.pbj,
                       A.b.invM, T);
    int rc = tiled_scan(A, T.cnt_b, T.boff, N, A.istat + IS_NBODYSORT, s);
    if (rc) return rc;
    rc = tiled_scan(A, T.cnt_p, T.poff, N, A.istat + IS_NPAIRSORT, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_t_scatter_b, dim3(tblk(N)), dim3(ME_THREADS), 0, s, N, T);
    hipLaunchKernelGGL(k_t_scatter_p, dim3(tblk(capP)), dim3(ME_THREADS), 0, s, capP, A.istat, T);
    hipLaunchKernelGGL(k_t_tiles, dim3(tblk(N)), dim3(ME_THREADS), 0, s, N, TB, A.istat, T);
    if (TB <= 64) launch_tiles<64>(A, T, ntile, s);
    else launch_tiles<256>(A, T, ntile, s);
    if (A.warm)
        hipLaunchKernelGGL(k_warm_store, dim3(tblk(capC)), dim3(ME_THREADS), 0, s, capC, A.istat, A.Cc.key, A.Cc.jn,
                           A.Cc.jt1, A.Cc.jt2, A.wkey, A.wjn, A.wjt1, A.wjt2);
    if (cc)
        hipLaunchKernelGGL(k_t_cc_store, dim3(tblk(g)), dim3(ME_THREADS), 0, s, N, capP, A.istat, A.Pp.pbi, A.Pp.pbj,
                           A.Pp.col, A.b.invM, A.cache_a, A.cache_b, A.cache_col, A.cache_invM);
    return hipGetLastError() == hipSuccess ? ME_OK : ME_ERUNTIME;
}

}  // namespace me

struct me_engine;
me::TiledArgs tiled_args_of(me_engine *e, float sdt);
