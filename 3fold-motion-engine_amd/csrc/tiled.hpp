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

#define T_COLB 8192   // bodies of a tile whose JP claims sit in LDS
#define T_COLP 16384  // pairs of a tile whose colours sit in LDS
#define T_COL_DYN (8 * T_COLB + 4 * T_COLP)
#define T_PREPF 40               // floats of a contact's fast-arithmetic constants (10 float4)
#define T_REC_HDR4 10            // float4 of a big-tile pair record's header
#define T_REC4 (T_REC_HDR4 + 20)   // float4 of a big-tile pair record: header + 4 x 5 (k_t_prep)

struct TiledBuffers {
    int *parent, *root, *proot;
    int *cnt_b, *cnt_p, *fill_b, *fill_p, *boff, *poff;
    int *bsorted, *bslot, *psorted;
    int *tb0, *tb1, *tp0, *tp1;
    int *ccount_jp, *csub, *cbase;
    int *ccls;                   // JP class of every final colour [ME_NCOLMAX], and at [ME_NCOLMAX] the colour
                                 // count it was written for (k_t_plan); kept with the colour cache
    float *invMs, *invIws;       // body constants in island order (index = position in bsorted)
    int *rec;                    // big tiles: 8 ints per colour-sorted pair (p, a, b, la, lb, st, cnt, flags)
    float *prep;                 // 20 floats per contact: n, rA, rB, t1, t2, meff, meft1, meft2, bias, mu
    float *prepf;                // fast arithmetic: T_PREPF floats per contact (k_t_prep<true, true>)
    float *sprep;                // big tiles: T_REC4 float4 fields per colour-sorted pair, field-major in blocks
                                 // of 8 pairs (read only in the sweeps: body header, then 4 x 5 fields of
                                 // prep; tiled_big.hpp)
    float *simp;                 // big tiles: the impulses (jn, jt1, jt2, jp) of a colour-sorted pair's
                                 // contacts, one float4 per contact slot, slot-major in blocks of 8 pairs
    int *srec;                   // big tiles: 4 ints per colour-sorted pair (p, a, b, first contact)
    unsigned *pwordg;            // big tiles: the 32-bit word of each colour-sorted pair (k_t_bigord)
    int *pclass;                 // big tiles: the colour class of each colour-sorted pair (k_t_bigord)
    long long *dbg;              // phase timestamps (wall clock) of the largest big tile of >= 1000 bodies
    long long *bigw;             // tile work since the last reset, 4 counts (tiles, pairs, dynamic pair ends,
                                 // contacts) for each of k_t_big [0..3], k_t_small [4..7], k_t_giant [8..11]
    int *biglist;                // tiles beyond one small tile (IS_NBIG of them), in k_t_biglist's order
    int *bigq;                   // the same tiles, most pairs first (ties: lower tile index): the work queue
    int *smalllist;              // tiles within one small tile (IS_NSMALL of them), in k_t_biglist's order
    int *smallq;                 // the same, most pairs first (ties: lower tile index): k_t_small's queue
    int *giantlist;              // tiles beyond the big tile's LDS (IS_NGIANT of them): k_t_giant's
    int *gcollist;               // tiles beyond the LDS colouring limits (IS_NGCOL): k_t_color_giant's
    int *gcst;                   // giant tile: colour-class starts (ME_NCOLMAX + 1), class count at [ME_NCOLMAX + 1]
    int *gcnt;                   // giant tile: pairs per class with two dynamic bodies, then with one (ME_NCOLMAX + 1 each)
    unsigned *gbar;              // giant tiles: grid-barrier counts of the sweep [0] and colouring [1]
                                 // launches, and the colouring's round counters [2..4]
};

// Recording of one substep for its adjoint (all device pointers into one caller buffer; V, P =
// velocity and position iterations): the centres and orientations at the substep start, the warm-start
// match of every contact, and the body state and impulses at the start of every sweep. The remaining
// arrays are scratch of the reverse pass.
struct AdjTape {
    float *xpre, *qpre;  // [3N], [4N]
    int *widx;           // [capC] index of the warm-start match in the previous substep's contacts, -1 none
    float *sst;          // [(V+P) * 6N] sweep-start (v, w) then (pv, po), by body
    float *simp;         // [(V+P) * 4 * capC] sweep-start (jn, jt1, jt2, jp), by contact
    float *cbar;         // [20 * capC] adjoint of the per-contact constants (k_t_prep layout)
    float *jbi;          // [3 * capC] adjoint of the warm-started impulses
    float *gbar;         // [7 * capC] adjoint of contact point, normal, penetration
    float *pvb, *pob;    // [3N] adjoint of the pseudo-velocities
    long long *acc;      // [ADJ_ACC * N] fixed-point sums of the body-level adjoint contributions
};

struct TiledArgs {
    int N, capC, capP, solver, vit, pit, warm, n_colors, tile_bodies;
    int warm_set;  // the contacts already carry their warm-start impulses (bead manifold stage)
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
    unsigned long long *scan_words;  // the one-pass scans' ticket counters and block words (k_scan1)
    unsigned long long *claim;
    void *side, *ev_fork, *ev_join;  // side stream and fork/join events of the sweep kernels
    // timing: next event of kernel `id` (ME_K_*) for the engine `prof`, null when not timed
    void *(*prof_ev)(void *prof, int id);
    void *prof;
    int taped;                       // record the substep into tape (k_t_small<.., true>)
    int giant_grid;                  // workgroups of the giant-island grid kernels, all resident together
    int big_grid;                    // workgroups of the big-tile kernels (one per CU the LDS allows)
    int small_grid;                  // workgroups of the persistent small-tile kernel (all resident ones)
    int fast;                        // fast arithmetic (ME_FLAG_FAST_MATH): the *_f contact updates
    int giant_possible, gcol_possible;  // the largest scene could hold an island beyond the big tile's LDS
                                        // (beyond the LDS colouring limits): islands never cross scenes
    long long giant_wait_ticks;      // bound of a grid barrier's wait, in wall-clock counter ticks
    AdjTape tape;
};

// tile arrays the adjoint of a recorded substep reads (bounds, body and pair order, constants)
template <class BumpT>
inline void tiled_layout_state(TiledBuffers &t, BumpT &B, int N, int capP, int capC) {
    size_t n = (size_t)N, p = (size_t)capP;
    t.bsorted = B.template take<int>(n);
    t.bslot = B.template take<int>(n);
    t.psorted = B.template takec<int>(p, IS_P, 1);
    t.tb0 = B.template take<int>(n);
    t.tb1 = B.template take<int>(n);
    t.tp0 = B.template take<int>(n);
    t.tp1 = B.template take<int>(n);
    t.prep = B.template takec<float>(20 * (size_t)capC, IS_C, 20);
    t.ccls = B.template take<int>(ME_NCOLMAX + 1);
}

// The JP class of final colour c (colours base[j] .. base[j+1]-1 split JP class j). Pairs of one JP
// class share no dynamic body, so a colour step can take the whole class: the same result as its
// sub-colours one after another, in fewer steps. The identity when the table was written for another
// colour count.
__device__ __forceinline__ int tile_class(const TiledBuffers &T, int c, int ncol) {
    return (T.ccls[ME_NCOLMAX] == ncol && c >= 0 && c < ncol) ? T.ccls[c] : c;
}

// tile arrays rebuilt within every substep
template <class BumpT>
inline void tiled_layout_scratch(TiledBuffers &t, BumpT &B, int N, int capP, int capC, bool fast) {
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
    t.ccount_jp = B.template take<int>(ME_NCOLMAX + 1);
    t.csub = B.template take<int>(ME_NCOLMAX + 1);
    t.cbase = B.template take<int>(ME_NCOLMAX + 2);
    t.invMs = B.template take<float>(n);
    t.invIws = B.template take<float>(12 * n);
    t.rec = B.template take<int>(8 * p);
    t.prepf = B.template take<float>(fast ? T_PREPF * (size_t)capC : 1);
    const size_t p8 = (p + 7) & ~size_t(7);  // whole blocks of 8 pairs (tiled_big.hpp: T_RB)
    t.sprep = B.template take<float>(4 * T_REC4 * p8);
    t.simp = B.template take<float>(16 * p8);
    t.srec = B.template take<int>(4 * p);
    t.pwordg = B.template take<unsigned>(p);
    t.pclass = B.template take<int>(p);
    t.dbg = B.template take<long long>(16);
    t.bigw = B.template take<long long>(12);
    t.biglist = B.template take<int>(n);
    t.bigq = B.template take<int>(n);
    t.smalllist = B.template take<int>(n);
    t.smallq = B.template take<int>(n);
    t.giantlist = B.template take<int>(n);
    t.gcollist = B.template take<int>(n);
    t.gcst = B.template take<int>(ME_NCOLMAX + 2);
    t.gcnt = B.template take<int>(2 * (ME_NCOLMAX + 1));
    t.gbar = B.template take<unsigned>(8);
}

// IS_MISMATCH after k_t_mode: 0 = ColorCache hit, 1 = colour this substep, 2 = no contacts
__device__ __forceinline__ void t_init_at(int i, int N, const TiledBuffers &T, int *istat, int lds_islands) {
    if (i < N) {
        if (!lds_islands) {
            T.parent[i] = i;
            T.cnt_b[i] = 0; T.cnt_p[i] = 0; T.fill_b[i] = 0; T.fill_p[i] = 0;
        }
        T.tb0[i] = INT_MAX; T.tb1[i] = -1; T.tp0[i] = INT_MAX; T.tp1[i] = -1;
    }
    if (i <= ME_NCOLMAX) T.ccount_jp[i] = 0;
    if (i == 0) {
        istat[IS_NCOLJP] = 0; istat[IS_ISL] = 0; istat[IS_TILES] = 0; istat[IS_MAXTB] = 0;
        istat[IS_MISMATCH] = 0;
        istat[IS_NBIG] = 0;
        istat[IS_BIGQ] = 0;
        istat[IS_COLQ] = 0;
        istat[IS_NSMALL] = 0;
        istat[IS_SMALLQ] = 0;
        istat[IS_NGIANT] = 0;
        istat[IS_NGCOL] = 0;
        T.dbg[7] = 0;
    }
}

// The first launch of the tiled set-up: the adjoint's warm-start matches (recording), the contacts'
// warm-start impulses (k_warm_lookup), the tile arrays' initial values and, under ColorCache, the
// comparison of the pair list with the colour cache (into IS_CCDIFF, clear since the last substep's
// k_t_biglist). One thread per contact, pair and body.
__global__ void k_t_setup(TiledArgs A, TiledBuffers T, int lds_islands, int cc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (A.taped) tape_widx_at(i, A.capC, A.istat, A.warm, A.Cc.key, A.wkey, A.tape.widx);
    if (!A.warm_set)
        warm_lookup_at(i, A.capC, A.istat, A.warm, A.Cc.key, A.wkey, A.wjn, A.wjt1, A.wjt2, A.Cc.jn, A.Cc.jt1,
                       A.Cc.jt2, A.Cc.jp);
    t_init_at(i, A.N, T, A.istat, lds_islands);
    if (cc) cc_compare_at(i, A.N, A.istat, IS_CCDIFF, A.Pp.pbi, A.Pp.pbj, A.cache_a, A.cache_b, A.b.invM, A.cache_invM);
}

// The substep's colouring mode, IS_MISMATCH: 2 without contacts, 1 to colour (a cache miss, or no
// cache), 0 to take the cached colours, which the launch copies back into the pairs. Every thread
// derives the mode from the same words; the first writes it and the cache counters.
__global__ void k_t_mode(int capP, int *istat, int cc_mode, const int *ccol, int *col) {
    const int P = istat[IS_P];
    const int mode = (istat[IS_C] <= 0 || P <= 0) ? 2 : (!cc_mode || istat[IS_CCDIFF] != 0) ? 1 : 0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        istat[IS_MISMATCH] = mode;
        if (mode == 2) istat[IS_NCOL] = 0;
        else if (cc_mode && mode == 0) { istat[IS_HITS] += 1; istat[IS_NCOL] = istat[IS_CACHED_NCOL]; }
        else if (cc_mode) istat[IS_MISSES] += 1;
    }
    if (mode == 0 && i < P && i < capP) col[i] = ccol[i];
}

// end of a tiled substep, one launch: the warm store (A23, contact_engine_gpu_colored.py:1089-1093)
// and, after a colour-cache miss, the cache's copy of the pair list, colours and invM
// (contact_engine_gpu_color_cache.py:48-53)
__global__ void k_t_finish(int N, int capP, int capC, int warm, int cc, int *istat, const int64_t *key,
                           const float *jn, const float *jt1, const float *jt2, int64_t *wkey, float *wjn, float *wjt1,
                           float *wjt2, const int *pbi, const int *pbj, const int *col, const float *invM, int *ca,
                           int *cb, int *ccol, float *cinvM) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (warm) {
        int C = istat[IS_C];
        if (C > capC) C = capC;
        if (i < C) { wkey[i] = key[i]; wjn[i] = jn[i]; wjt1[i] = jt1[i]; wjt2[i] = jt2[i]; }
    }
    if (cc && istat[IS_MISMATCH] == 1) {
        int P = istat[IS_P];
        if (i < P && i < capP) { ca[i] = pbi[i]; cb[i] = pbj[i]; ccol[i] = col[i]; }
        if (i < N) cinvM[i] = invM[i];
    }
    if (i == 0) {  // no thread of this launch reads the words written here
        int C = istat[IS_C];
        if (warm) istat[IS_WN] = C > capC ? capC : C;
        if (cc && istat[IS_MISMATCH] == 1) { istat[IS_CACHED_P] = istat[IS_P]; istat[IS_CACHED_NCOL] = istat[IS_NCOL]; }
    }
}

// parent chains only ever point to smaller indices, so the walk ends within N steps
__device__ __forceinline__ int uf_find(const int *parent, int x) {
    const volatile int *par = parent;
    int p = par[x];
    while (p != x) { x = p; p = par[x]; }
    return x;
}

// find with path halving: a node's parent is replaced by its grandparent, an ancestor, so every
// chain still ends at the same root whatever other threads read or write meanwhile
__device__ __forceinline__ int uf_find_halve(int *parent, int x) {
    volatile int *par = parent;
    int p = par[x];
    while (p != x) {
        int g = par[p];
        if (g != p) par[x] = g;
        x = p;
        p = par[x];
    }
    return x;
}

// Atomic add of 1 per lane to cnt[key] with the wave's most common case folded: the lanes that share
// the first active lane's key add their count once (one atomic instead of up to 64 on one address,
// the case of an island that holds most of the scene); the others add their own. Returns this
// lane's old-value slot (base + rank among the lanes of its key).
__device__ __forceinline__ int wave_add1(int *cnt, int key, bool active) {
    const unsigned long long act = __ballot(active);
    if (!active) return 0;
    const int lane = __lane_id();
    const int first = __ffsll((long long)act) - 1;
    const int k0 = __shfl(key, first);
    const unsigned long long same = __ballot(active && key == k0);
    if (key == k0) {
        int base = 0;
        if (lane == first) base = atomicAdd(&cnt[k0], __popcll(same));
        base = __shfl(base, first);
        return base + __popcll(same & ((1ull << lane) - 1ull));
    }
    return atomicAdd(&cnt[key], 1);
}

// wave_add1 over a workgroup of NT threads: every lane whose key is that of the workgroup's first
// active lane (in a large scene mostly the one island that holds most of it) gets its slot from one
// atomic per workgroup, the others from wave_add1. The same-address atomics of an island of 90 000
// bodies drop from one per wave to one per workgroup. Every thread of the workgroup calls it.
template <int NT>
__device__ __forceinline__ int block_add1(int *cnt, int key, bool active) {
    constexpr int NW = NT / 64;
    __shared__ int s_key[NW], s_has[NW], s_cnt[NW], s_base;
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    const unsigned long long act = __ballot(active);
    const int first = act ? __ffsll((long long)act) - 1 : 0;
    const int kw = __shfl(key, first);
    if (lane == 0) { s_key[wv] = kw; s_has[wv] = act != 0ull; }
    __syncthreads();
    int k0 = 0, any = 0;
    for (int w = 0; w < NW; ++w)
        if (s_has[w]) { k0 = s_key[w]; any = 1; break; }
    const bool mine = any && active && key == k0;
    const unsigned long long same = __ballot(mine);
    if (lane == 0) s_cnt[wv] = __popcll(same);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < NW; ++w) tot += s_cnt[w];
        s_base = tot ? atomicAdd(&cnt[k0], tot) : 0;
    }
    __syncthreads();
    int rank = s_base;
    for (int w = 0; w < wv; ++w) rank += s_cnt[w];
    rank += __popcll(same & ((1ull << lane) - 1ull));
    const int other = wave_add1(cnt, key, active && !mine);
    return mine ? rank : other;
}
#define T_AGG 1024  // workgroup of the island-build kernels that count or scatter by root

// lock-free union: link the larger root under the smaller one (compare-and-swap on the root)
__global__ void k_t_union(int capP, int *istat, const int *pbi, const int *pbj, const float *invM, int *parent) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= istat[IS_P] || p >= capP) return;
    int a = pbi[p], b = pbj[p];
    if (b < 0 || !(invM[a] > 0.0f) || !(invM[b] > 0.0f)) return;
    for (int it = 0; it < (1 << 20); ++it) {
        int ra = uf_find_halve(parent, a), rb = uf_find_halve(parent, b);
        if (ra == rb) return;
        int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
        if (atomicCAS(&parent[hi], hi, lo) == hi) return;
    }
    atomicOr(&istat[IS_OVF], OVF_TILE);
}

__global__ void __launch_bounds__(T_AGG) k_t_flatten(int N, const float *invM, TiledBuffers T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool dyn = i < N && invM[i] > 0.0f;
    int r = -1;
    if (dyn) {
        r = uf_find_halve(T.parent, i);
        T.root[i] = r;
    } else if (i < N) {
        T.root[i] = -1;
    }
    block_add1<T_AGG>(T.cnt_b, r, dyn);
}

__global__ void __launch_bounds__(T_AGG) k_t_pair_root(int capP, const int *istat, const int *pbi, const int *pbj, const float *invM,
                              TiledBuffers T) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int r = -1;
    if (p < istat[IS_P] && p < capP) {
        int a = pbi[p], b = pbj[p];
        r = invM[a] > 0.0f ? T.root[a] : ((b >= 0 && invM[b] > 0.0f) ? T.root[b] : -1);
        T.proot[p] = r;
    }
    block_add1<T_AGG>(T.cnt_p, r, r >= 0);
}

__global__ void __launch_bounds__(T_AGG) k_t_scatter_b(int N, TiledBuffers T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int r = i < N ? T.root[i] : -1;
    int slot = block_add1<T_AGG>(T.fill_b, r, r >= 0);
    if (r < 0) return;
    int pos = T.boff[r] + slot;
    T.bsorted[pos] = i;
    T.bslot[i] = pos;
}

__global__ void __launch_bounds__(T_AGG) k_t_scatter_p(int capP, const int *istat, TiledBuffers T) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    int r = (p < istat[IS_P] && p < capP) ? T.proot[p] : -1;
    int slot = block_add1<T_AGG>(T.fill_p, r, r >= 0);
    if (r >= 0) T.psorted[T.poff[r] + slot] = p;
}

// Tile bounds from one island: bodies [s, e) and pairs [ps, pe), the islands laid out back to back in
// root order over nb bodies; tile of an island = s / TB. The next island starts at e, so an island
// closes its tile when e / TB lies beyond it, and then opens the tile e / TB; the first island opens
// tile 0. Every tile that holds an island is written exactly once, with plain stores (the tile's
// first start, last end, first pair start, last pair end); the others keep their initial values.
__device__ __forceinline__ void island_tile_bounds(const TiledBuffers &T, int TB, int nb, int s, int e, int ps, int pe) {
    int t = s / TB;
    if (s == 0) { T.tb0[0] = 0; T.tp0[0] = ps; }
    if (e >= nb || e / TB > t) {
        T.tb1[t] = e; T.tp1[t] = pe;
        if (e < nb) { T.tb0[e / TB] = e; T.tp0[e / TB] = pe; }
    }
}

__global__ void k_t_tiles(int N, int TB, int *istat, TiledBuffers T) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool root = i < N && T.cnt_b[i] != 0;
    wave_atomic_count(&istat[IS_ISL], root);
    if (!root) return;
    island_tile_bounds(T, TB, istat[IS_NBODYSORT], T.boff[i], T.boff[i] + T.cnt_b[i], T.poff[i],
                       T.poff[i] + T.cnt_p[i]);
}

// ── the island build of a scene of at most T_ISL_MAXN bodies in one workgroup ──
// The same result as k_t_union .. k_t_tiles (islands of the dynamic-dynamic pair graph, their bodies
// and pairs laid out island-major by root, tiles of about TB bodies), with the union-find forest,
// the per-root counts and their offsets in LDS and no kernel boundary between the phases. Within an
// island the order of bodies and pairs is the order of LDS atomics, which no result depends on (a
// tile's pairs are coloured by key and a colour's pairs touch disjoint bodies).
#define T_ISL_THREADS 1024
#ifndef T_ISL_MAXN
#define T_ISL_MAXN 12288  // three int arrays of N in LDS
#endif
#define T_ISL_DYN (3 * T_ISL_MAXN * 4)

__device__ __forceinline__ int lds_find_halve(volatile int *par, int x) {
    int p = par[x];
    while (p != x) {
        int g = par[p];
        if (g != p) par[x] = g;
        x = p;
        p = par[x];
    }
    return x;
}

// exclusive scan of a[0..n) in place by T_ISL_THREADS threads (a contiguous chunk each); returns the total
__device__ int lds_exclusive_scan(int *a, int n, int *part) {
    int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int chunk = (n + T_ISL_THREADS - 1) / T_ISL_THREADS;
    int lo = t * chunk, hi = lo + chunk < n ? lo + chunk : n;
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += a[i];
    int incl = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) part[wv] = incl;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int w = 0; w < T_ISL_THREADS / 64; ++w) { int v = part[w]; part[w] = run; run += v; }
        part[T_ISL_THREADS / 64] = run;
    }
    __syncthreads();
    int base = part[wv] + incl - sum;
    for (int i = lo; i < hi; ++i) { int v = a[i]; a[i] = base; base += v; }
    int total = part[T_ISL_THREADS / 64];
    __syncthreads();
    return total;
}

#define T_ISL_PK 16  // pairs per thread held in registers (the first 16 384 pairs; the rest re-read)
#define T_ISL_BK ((T_ISL_MAXN + T_ISL_THREADS - 1) / T_ISL_THREADS)

// root of pair p's island (-1: no dynamic end) from its endpoints a, b (b < 0: ground) and their
// dynamic flags, once the forest points every body at its root
__device__ __forceinline__ int isl_pair_root(const int *par, int a, int b, bool da, bool db) {
    return da ? par[a] : (db ? par[b] : -1);
}

__global__ void __launch_bounds__(T_ISL_THREADS) k_t_islands_lds(int N, int capP, int TB, int *istat, const int *pbi,
                                                                 const int *pbj, const float *invM, TiledBuffers T) {
    extern __shared__ int sh[];
    __shared__ int part[T_ISL_THREADS / 64 + 1];
    int *par = sh, *cb = sh + N, *cp = sh + 2 * N;
    volatile int *vpar = par;
    const int t = threadIdx.x;
    int P = istat[IS_P];
    if (P > capP) P = capP;
    // this thread's pairs p = t + k * T_ISL_THREADS and bodies i = t + k * T_ISL_THREADS: endpoints and
    // dynamic flags loaded up front (all loads of a kind in flight together, not one per iteration)
    int pa[T_ISL_PK], pb[T_ISL_PK];
    unsigned dA = 0, dB = 0, dBody = 0;
#pragma unroll
    for (int k = 0; k < T_ISL_PK; ++k) {
        int p = t + k * T_ISL_THREADS;
        pa[k] = p < P ? pbi[p] : 0;
        pb[k] = p < P ? pbj[p] : -1;
    }
#pragma unroll
    for (int k = 0; k < T_ISL_PK; ++k) {
        int p = t + k * T_ISL_THREADS;
        if (p < P && invM[pa[k]] > 0.0f) dA |= 1u << k;
        if (p < P && pb[k] >= 0 && invM[pb[k]] > 0.0f) dB |= 1u << k;
    }
#pragma unroll
    for (int k = 0; k < T_ISL_BK; ++k) {
        int i = t + k * T_ISL_THREADS;
        if (i < N && invM[i] > 0.0f) dBody |= 1u << k;
    }
    for (int i = t; i < N; i += T_ISL_THREADS) { par[i] = i; cb[i] = 0; cp[i] = 0; }
    __syncthreads();
    // union: link the larger root under the smaller (compare-and-swap on the root)
    bool stuck = false;
    auto unite = [&](int a, int b) {
        int it = 0;
        for (; it < (1 << 20); ++it) {
            int ra = lds_find_halve(vpar, a), rb = lds_find_halve(vpar, b);
            if (ra == rb) break;
            int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
            if (atomicCAS(&par[hi], hi, lo) == hi) break;
        }
        stuck = stuck || it == (1 << 20);
    };
#pragma unroll
    for (int k = 0; k < T_ISL_PK; ++k)
        if (((dA & dB) >> k) & 1u) unite(pa[k], pb[k]);
    for (int p = t + T_ISL_PK * T_ISL_THREADS; p < P; p += T_ISL_THREADS) {
        int a = pbi[p], b = pbj[p];
        if (b >= 0 && invM[a] > 0.0f && invM[b] > 0.0f) unite(a, b);
    }
    if (stuck) atomicOr(&istat[IS_OVF], OVF_TILE);
    __syncthreads();
    // roots and bodies per root (wave_add1 folds the lanes that share a root into one LDS atomic)
    int rootk[T_ISL_BK];
#pragma unroll
    for (int k = 0; k < T_ISL_BK; ++k) {
        int i = t + k * T_ISL_THREADS;
        bool dyn = (dBody >> k) & 1u;
        int r = dyn ? lds_find_halve(vpar, i) : -1;
        rootk[k] = r;
        if (i < N) T.root[i] = r;
        wave_add1(cb, r, dyn);
    }
    __syncthreads();
    // the forest now points every body at its root; only after every find has ended (a path-halving
    // write of a concurrent find could replace a root written earlier by an older ancestor)
#pragma unroll
    for (int k = 0; k < T_ISL_BK; ++k)
        if ((dBody >> k) & 1u) par[t + k * T_ISL_THREADS] = rootk[k];
    __syncthreads();
    // pairs per root (a pair belongs to the island of its dynamic end)
#pragma unroll
    for (int k = 0; k < T_ISL_PK; ++k) {
        int r = isl_pair_root(par, pa[k], pb[k], (dA >> k) & 1u, (dB >> k) & 1u);
        wave_add1(cp, r, r >= 0);
    }
    for (int p0 = T_ISL_PK * T_ISL_THREADS; p0 < P; p0 += T_ISL_THREADS) {
        int p = p0 + t, r = -1;
        if (p < P) {
            int a = pbi[p], b = pbj[p];
            r = isl_pair_root(par, a, b, invM[a] > 0.0f, b >= 0 && invM[b] > 0.0f);
        }
        wave_add1(cp, r, r >= 0);
    }
    __syncthreads();
    int nbs = lds_exclusive_scan(cb, N, part);
    int nps = lds_exclusive_scan(cp, N, part);
    // tiles: an island goes to tile boff / TB; tile bounds over its islands
    int nroot = 0;
    for (int r = t; r < N; r += T_ISL_THREADS) {
        int b1 = r + 1 < N ? cb[r + 1] : nbs;
        if (b1 == cb[r]) continue;
        ++nroot;
        int p1 = r + 1 < N ? cp[r + 1] : nps;
        island_tile_bounds(T, TB, nbs, cb[r], b1, cp[r], p1);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nroot += __shfl_xor(nroot, o);
    if (t == 0) part[0] = 0;
    __syncthreads();  // every thread has read the offsets before they become cursors
    if ((t & 63) == 0 && nroot) atomicAdd(&part[0], nroot);
    // scatter: bodies, then pairs, island-major (the offsets are the cursors)
    int posk[T_ISL_BK];
#pragma unroll
    for (int k = 0; k < T_ISL_BK; ++k) posk[k] = wave_add1(cb, rootk[k], (dBody >> k) & 1u);
#pragma unroll
    for (int k = 0; k < T_ISL_BK; ++k) {
        if (!((dBody >> k) & 1u)) continue;
        int i = t + k * T_ISL_THREADS, pos = posk[k];
        T.bsorted[pos] = i;
        T.bslot[i] = pos;
    }
#pragma unroll
    for (int k = 0; k < T_ISL_PK; ++k) {
        int r = isl_pair_root(par, pa[k], pb[k], (dA >> k) & 1u, (dB >> k) & 1u);
        int pos = wave_add1(cp, r, r >= 0);
        if (r >= 0) T.psorted[pos] = t + k * T_ISL_THREADS;
    }
    for (int p0 = T_ISL_PK * T_ISL_THREADS; p0 < P; p0 += T_ISL_THREADS) {
        int p = p0 + t, r = -1;
        if (p < P) {
            int a = pbi[p], b = pbj[p];
            r = isl_pair_root(par, a, b, invM[a] > 0.0f, b >= 0 && invM[b] > 0.0f);
        }
        int pos = wave_add1(cp, r, r >= 0);
        if (r >= 0) T.psorted[pos] = p;
    }
    __syncthreads();
    if (t == 0) { istat[IS_NBODYSORT] = nbs; istat[IS_NPAIRSORT] = nps; istat[IS_ISL] += part[0]; }
}

// JP rounds of one tile (k_claim / k_win of contact_engine_gpu_colored.py:151-185, restricted to the
// tile's islands). LDS mode: claims by the body's slot in the tile, colours in LDS. Oversized
// tiles: claims in the global claim array by body id (tiles never share a dynamic body).
// LDSM: claims and colours in LDS (claim and colL point into it, so their accesses compile to ds_*
// instructions); otherwise the global claim array and the pair colours.
template <int WG, bool LDSM>
__device__ __forceinline__ void t_color_tile(const TiledArgs &A, const TiledBuffers &T, int b0, int nb, int p0, int np,
                                             unsigned long long *claim, int *colL, int &left, int *hist) {
    constexpr bool lds = LDSM;
    const float *invM = A.b.invM;
    int *gcol = A.Pp.col;
    auto slot = [&](int body) { return lds ? T.bslot[body] - b0 : body; };
    auto getc = [&](int k, int p) { return lds ? colL[k] : gcol[p]; };
    for (int k = threadIdx.x; k < np; k += WG) {
        if (lds) colL[k] = -1;
        else gcol[T.psorted[p0 + k]] = -1;
    }
    int nround = 0;
    if (lds && np <= T_COLP) {
        // the LDS rounds with each thread's pairs (keys, body slots; -1 for a static side) in registers,
        // read once instead of every round; the claims by tile slot (slot j is the tile's body j)
        constexpr int PER = T_COLP / WG;
        unsigned long long ck[PER];
        int sa[PER], sb[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int k = threadIdx.x + j * WG;
            ck[j] = 0ull;
            sa[j] = sb[j] = -1;
            if (k < np) {
                const int p = T.psorted[p0 + k];
                const int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
                ck[j] = (unsigned long long)jp_key(p);
                if (invM[a] > 0.0f) sa[j] = T.bslot[a] - b0;
                if (b >= 0 && invM[b] > 0.0f) sb[j] = T.bslot[b] - b0;
            }
        }
        for (int cur = 0; cur <= np; ++cur) {
            for (int j = threadIdx.x; j < nb; j += WG) claim[j] = ~0ull;
            if (threadIdx.x == 0) left = 0;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int k = threadIdx.x + j * WG;
                if (k >= np || colL[k] >= 0) continue;
                if (sa[j] >= 0) atomicMin(&claim[sa[j]], ck[j]);
                if (sb[j] >= 0) atomicMin(&claim[sb[j]], ck[j]);
            }
            __syncthreads();
            int mine = 0;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int k = threadIdx.x + j * WG;
                if (k >= np || colL[k] >= 0) continue;
                const bool ok = (sa[j] < 0 || claim[sa[j]] == ck[j]) && (sb[j] < 0 || claim[sb[j]] == ck[j]);
                if (ok) colL[k] = cur;
                else ++mine;
            }
            if (mine) atomicAdd(&left, mine);
            __syncthreads();
            int l = left;
            __syncthreads();
            if (l == 0) { nround = cur + 1; break; }
        }
    }
    for (int cur = 0; !(lds && np <= T_COLP) && cur <= np; ++cur) {
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
    // class sizes: a histogram in LDS, then one global add per class and tile
    for (int c = threadIdx.x; c < nround && c < ME_NCOLMAX; c += WG) hist[c] = 0;
    __syncthreads();
    bool ovf = false;
    for (int k = threadIdx.x; k < np; k += WG) {
        int p = T.psorted[p0 + k];
        int c = getc(k, p);
        if (lds) gcol[p] = c;
        if (c >= 0 && c < ME_NCOLMAX) atomicAdd(&hist[c], 1);
        else ovf = true;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nround && c < ME_NCOLMAX; c += WG)
        if (hist[c]) atomicAdd(&T.ccount_jp[c], hist[c]);
    if (ovf) atomicOr(&A.istat[IS_OVF], OVF_COLORS);
    if (threadIdx.x == 0) atomicMax(&A.istat[IS_NCOLJP], nround);
}

template <int WG>
// the big tiles (k_t_biglist), taken in turn by the launch's workgroups; small tiles go to
// k_t_color_small, many workgroups per CU instead of one 128 KB-LDS workgroup per tile
__global__ void __launch_bounds__(WG) k_t_color(TiledArgs A, TiledBuffers T, int giant) {
    if (A.istat[IS_MISMATCH] != 1) return;
    extern __shared__ unsigned long long cdsm[];
    __shared__ int left, hist[ME_NCOLMAX];
    const int nbig = A.istat[IS_NBIG];
    __shared__ int qi;
    for (;;) {
        if (threadIdx.x == 0) qi = atomicAdd(&A.istat[IS_COLQ], 1);
        __syncthreads();
        const int i = qi;
        __syncthreads();
        if (i >= nbig) break;
        int t = T.bigq[i];
        int b0 = T.tb0[t];
        int nb = T.tb1[t] - b0, p0 = T.tp0[t], np = T.tp1[t] - T.tp0[t];
        if (np > 0) {
            if (nb <= T_COLB && np <= T_COLP)
                t_color_tile<WG, true>(A, T, b0, nb, p0, np, cdsm, reinterpret_cast<int *>(cdsm + T_COLB), left, hist);
            else if (!giant)  // else k_t_color_giant's grid colours it
                t_color_tile<WG, false>(A, T, b0, nb, p0, np, A.claim, nullptr, left, hist);
        }
        __syncthreads();
    }
}

// split of the JP classes (same rule as k_color_split_plan)
// One wave: the greedy split rule of k_color_split_plan (each of the K - nc extra colours goes to the
// class with the most pairs per sub-colour, the lowest class index on ties) with the class counts in
// LDS and an argmax over the wave per extra colour.
__global__ void __launch_bounds__(64) k_t_plan(int *istat, int K, TiledBuffers T) {
    __shared__ int cnt[ME_NCOLMAX], sub[ME_NCOLMAX], base[ME_NCOLMAX + 1];
    if (istat[IS_MISMATCH] != 1) return;
    const int lane = threadIdx.x;
    int nc = istat[IS_NCOLJP];
    if (nc > ME_NCOLMAX) {
        nc = ME_NCOLMAX;
        if (lane == 0) atomicOr(&istat[IS_OVF], OVF_COLORS);
    }
    for (int c = lane; c < nc; c += 64) { cnt[c] = T.ccount_jp[c]; sub[c] = 1; }
    __syncthreads();
    for (int e = nc; K > nc && nc > 0 && e < K && e < ME_NCOLMAX; ++e) {
        int bc = -1;
        for (int c = lane; c < nc; c += 64)  // this lane's best, lowest index first
            if (bc < 0 || (long long)cnt[c] * sub[bc] > (long long)cnt[bc] * sub[c]) bc = c;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const int oc = __shfl_xor(bc, o, 64);
            if (oc >= 0 && (bc < 0 || (long long)cnt[oc] * sub[bc] > (long long)cnt[bc] * sub[oc] ||
                            ((long long)cnt[oc] * sub[bc] == (long long)cnt[bc] * sub[oc] && oc < bc)))
                bc = oc;
        }
        if (lane == 0) sub[bc]++;
        __syncthreads();
    }
    if (lane == 0) {
        base[0] = 0;
        for (int c = 0; c < nc; ++c) base[c + 1] = base[c] + sub[c];
    }
    __syncthreads();
    const int tot = nc > 0 ? base[nc] : 0;
    for (int c = lane; c < nc; c += 64) {
        T.csub[c] = sub[c];
        for (int s = base[c]; s < base[c + 1] && s < ME_NCOLMAX; ++s) T.ccls[s] = c;
    }
    for (int c = lane; c <= nc; c += 64) T.cbase[c] = base[c];
    if (lane == 0) {
        T.ccls[ME_NCOLMAX] = tot;
        istat[IS_NCOL] = tot;
    }
}

// final colour of every pair: on a fresh colouring, JP class c becomes base[c] + rank mod sub[c]
__global__ void k_t_apply_split(int capP, const int *istat, int *col, TiledBuffers T) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (istat[IS_MISMATCH] != 1 || p >= istat[IS_P] || p >= capP) return;
    int c = col[p];
    if (c >= 0 && c < ME_NCOLMAX) col[p] = T.cbase[c] + p % T.csub[c];
}

#define T_SMALL_P 256   // pairs of a small tile: one per thread, constants in registers
#define T_SMALL_B 512   // bodies of a small tile (velocities in LDS)
#ifndef T_BIG_B
#define T_BIG_B 5000    // bodies of a big tile whose v, w (then pv, po) sit in LDS (120 KB)
#endif

// per-contact constants of one substep's solve (identical to recomputing them every sweep: their
// inputs -- normal, contact points, centres, world inverse inertia -- do not change in the solve)
// A contact's constants for the substep and its impulses. Under fast arithmetic meff, meft1 and meft2
// hold the reciprocals of the effective masses (zero where the exact solve skips the update) and
// kA, kB, k1A, k2A, k1B, k2B the bodies' angular responses to unit impulses along n, t1, t2.
struct CCon {
    f3 n, rA, rB, t1, t2;
    float meff, meft1, meft2, bias, mu;
    float jn, jt1, jt2, jp;
    f3 kA, kB, k1A, k2A, k1B, k2B;
};

// Per-contact constants of the substep under fast arithmetic (ColorCache's fixed friction frame), 10
// float4 per contact in prepf, grouped so that the position pass reads the first five and the warm
// apply the first and the last seven:
//   (n, 1/meff) (rA, bias) (rB, mu) (kA, 1/meft1) (kB, 1/meft2) (t1, k1A.x) (t2, k1A.y)
//   (k1A.z, k2A) (k1B, k2B.x) (k2B.y, k2B.z, 0, 0)
// with kA = IA (rA x n), k1A = IA (rA x t1), k2A = IA (rA x t2) and the same for B (zero without B).
// The reciprocals are zero where the exact sweep skips the update (meff < 1e-12; meft1 or meft2 <
// 1e-12), and the effective masses are the exact sweep's own expressions.
__device__ __forceinline__ void prepf_write(float4 *o, f3 n, f3 rA, f3 rB, f3 t1, f3 t2, float meff, float m1,
                                            float m2, float bias, float mu, f3 kA, f3 kB, f3 k1A, f3 k2A, f3 k1B,
                                            f3 k2B) {
    const float rm = meff >= 1e-12f ? 1.0f / meff : 0.0f;
    const bool tok = m1 >= 1e-12f && m2 >= 1e-12f;
    const float r1 = tok ? 1.0f / m1 : 0.0f, r2 = tok ? 1.0f / m2 : 0.0f;
    o[0] = make_float4(n.x, n.y, n.z, rm);
    o[1] = make_float4(rA.x, rA.y, rA.z, bias);
    o[2] = make_float4(rB.x, rB.y, rB.z, mu);
    o[3] = make_float4(kA.x, kA.y, kA.z, r1);
    o[4] = make_float4(kB.x, kB.y, kB.z, r2);
    o[5] = make_float4(t1.x, t1.y, t1.z, k1A.x);
    o[6] = make_float4(t2.x, t2.y, t2.z, k1A.y);
    o[7] = make_float4(k1A.z, k2A.x, k2A.y, k2A.z);
    o[8] = make_float4(k1B.x, k1B.y, k1B.z, k2B.x);
    o[9] = make_float4(k2B.y, k2B.z, 0.0f, 0.0f);
}

// fields f0 .. f9 of prepf_write into a contact (impulses untouched)
__device__ __forceinline__ void prepf_unpack(CCon &C, float4 f0, float4 f1, float4 f2, float4 f3, float4 f4,
                                             float4 f5, float4 f6, float4 f7, float4 f8, float4 f9) {
    C.n = mk(f0.x, f0.y, f0.z); C.meff = f0.w;
    C.rA = mk(f1.x, f1.y, f1.z); C.bias = f1.w;
    C.rB = mk(f2.x, f2.y, f2.z); C.mu = f2.w;
    C.kA = mk(f3.x, f3.y, f3.z); C.meft1 = f3.w;
    C.kB = mk(f4.x, f4.y, f4.z); C.meft2 = f4.w;
    C.t1 = mk(f5.x, f5.y, f5.z); C.t2 = mk(f6.x, f6.y, f6.z);
    C.k1A = mk(f5.w, f6.w, f7.x); C.k2A = mk(f7.y, f7.z, f7.w);
    C.k1B = mk(f8.x, f8.y, f8.z); C.k2B = mk(f8.w, f9.x, f9.y);
}

// contact c of prepf, the fields pass PASS reads (0 warm: f0, f3..f9; 1 velocity: all; 2 position: f0..f4)
template <int PASS>
__device__ __forceinline__ void ld_prepf(const float4 *o, CCon &C) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 f0 = o[0], f3 = o[3], f4 = o[4];
    const float4 f1 = PASS != 0 ? o[1] : z, f2 = PASS != 0 ? o[2] : z;
    float4 f5 = z, f6 = z, f7 = z, f8 = z, f9 = z;
    if (PASS != 2) { f5 = o[5]; f6 = o[6]; f7 = o[7]; f8 = o[8]; f9 = o[9]; }
    prepf_unpack(C, f0, f1, f2, f3, f4, f5, f6, f7, f8, f9);
}

struct PairCache {
    int cnt, la, lb;
    bool hasB, dynA, dynB;
    float iMA, iMB;
    m3 IA, IB;
    f3 vAs, wAs, vBs, wBs, pvBs, poBs, pvAs, poAs;  // static bodies: constant state
    CCon c[4];
};

template <bool FIXED, bool FAST>
__device__ __forceinline__ void pc_load(PairCache &P, const TiledArgs &A, const TiledBuffers &T, int p, int b0) {
    int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
    const float *invM = A.b.invM;
    P.iMA = invM[a];
    P.dynA = P.iMA > 0.0f;
    P.hasB = b >= 0;
    P.iMB = P.hasB ? invM[b] : 0.0f;
    P.dynB = P.hasB && P.iMB > 0.0f;
    P.la = P.dynA ? T.bslot[a] - b0 : a;
    P.lb = P.dynB ? T.bslot[b] - b0 : b;
    P.IA = ldm(A.invIw, a);
    P.IB = P.hasB ? ldm(A.invIw, b) : P.IA;
    f3 z = mk(0.0f, 0.0f, 0.0f);
    f3 xA = ld3(A.b.x, a), xB = P.hasB ? ld3(A.b.x, b) : z;
    P.vAs = P.dynA ? z : ld3(A.b.v, a); P.wAs = P.dynA ? z : ld3(A.b.w, a);
    P.pvAs = P.dynA ? z : ld3(A.pv, a); P.poAs = P.dynA ? z : ld3(A.po, a);
    P.vBs = (P.hasB && !P.dynB) ? ld3(A.b.v, b) : z; P.wBs = (P.hasB && !P.dynB) ? ld3(A.b.w, b) : z;
    P.pvBs = (P.hasB && !P.dynB) ? ld3(A.pv, b) : z; P.poBs = (P.hasB && !P.dynB) ? ld3(A.po, b) : z;
    int st = A.Pp.pstart[p];
    P.cnt = A.Pp.pcount[p];
    if (FAST) {  // the substep's constants from k_t_prep (the inertias are in the responses)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= P.cnt) break;
            int c = st + k;
            CCon &C = P.c[k];
            ld_prepf<1>(reinterpret_cast<const float4 *>(T.prepf) + 10 * (size_t)c, C);
            C.jn = A.Cc.jn[c]; C.jt1 = A.Cc.jt1[c]; C.jt2 = A.Cc.jt2[c]; C.jp = A.Cc.jp[c];
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        int c = st + k;
        CCon &C = P.c[k];
        C.n = ld3(A.Cc.n, c);
        C.rA = ld3(A.Cc.pA, c) - xA;
        C.rB = P.hasB ? ld3(A.Cc.pB, c) - xB : z;
        C.mu = A.Cc.mu[c];
        C.jn = A.Cc.jn[c]; C.jt1 = A.Cc.jt1[c]; C.jt2 = A.Cc.jt2[c]; C.jp = A.Cc.jp[c];
        f3 rnA = cross(C.rA, C.n);
        float meff = P.iMA + dot(rnA, mv(P.IA, rnA));
        if (P.hasB) {
            f3 rnB = cross(C.rB, C.n);
            meff = meff + P.iMB + dot(rnB, mv(P.IB, rnB));
        }
        C.meff = meff;
        C.bias = 0.2f * wmax(A.Cc.pen[c] - 1e-4f, 0.0f) / A.dt;
        C.t1 = fixed_t1(C.n);
        C.t2 = cross(C.n, C.t1);
        if (FIXED) {
            f3 rtA1 = cross(C.rA, C.t1);
            float m1 = P.iMA + dot(rtA1, mv(P.IA, rtA1));
            f3 rtA2 = cross(C.rA, C.t2);
            float m2 = P.iMA + dot(rtA2, mv(P.IA, rtA2));
            if (P.hasB) {
                f3 rtB1 = cross(C.rB, C.t1);
                m1 = m1 + P.iMB + dot(rtB1, mv(P.IB, rtB1));
                f3 rtB2 = cross(C.rB, C.t2);
                m2 = m2 + P.iMB + dot(rtB2, mv(P.IB, rtB2));
            }
            C.meft1 = m1;
            C.meft2 = m2;
        }
    }
}

__device__ __forceinline__ void pc_store(const PairCache &P, const TiledArgs &A, int p) {
    int st = A.Pp.pstart[p];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        int c = st + k;
        A.Cc.jn[c] = P.c[k].jn; A.Cc.jt1[c] = P.c[k].jt1; A.Cc.jt2[c] = P.c[k].jt2; A.Cc.jp[c] = P.c[k].jp;
    }
}

// one contact of apply_warm (contact_engine_gpu_applied_warm.py:7-54)
__device__ __forceinline__ void cc_warm(const CCon &C, f3 &va, f3 &wa, f3 &vb, f3 &wb, bool hasB, float iMA,
                                        float iMB, const m3 &IA, const m3 &IB) {
    // the warm apply's own frame (contracted), not the sweep's C.t1, C.t2
    f3 t1 = cfixed_t1(C.n);
    f3 t2 = ccross(C.n, t1);
    f3 imp = cmadd(C.jt2, t2, cmadd(C.jn, C.n, C.jt1 * t1));
    if (iMA > 0.0f) {
        va = cmadd(iMA, imp, va);
        wa = wa + cmv(IA, ccross(C.rA, imp));
    }
    if (hasB && iMB > 0.0f) {
        vb = cmadd(-iMB, imp, vb);
        wb = wb - cmv(IB, ccross(C.rB, imp));
    }
}

// one contact of the chunk velocity sweep, expression for expression as chunk_vel
template <bool FIXED>
__device__ __forceinline__ void cc_vel(CCon &C, f3 &vA, f3 &wA, f3 &vB, f3 &wB, bool hasB, float iMA, float iMB,
                                       const m3 &IA, const m3 &IB, float t2d) {
    const f3 nrm = C.n, rA = C.rA, rB = C.rB;
    f3 va = vA + cross(wA, rA);
    f3 vb = mk(0.0f, 0.0f, 0.0f);
    if (hasB) vb = vB + cross(wB, rB);
    f3 vrel = va - vb;
    float vn = dot(vrel, nrm);
    f3 t1, t2;
    float meft1, meft2;
    if (FIXED) {
        t1 = C.t1; t2 = C.t2;
    } else {
        f3 vt = vrel - nrm * vn;
        float mt = length(vt);
        t1 = vt / mt;
        if (mt <= 5e-3f) t1 = C.t1;
        t2 = cross(nrm, t1);
    }
    if (!(C.meff >= 1e-12f)) return;
    float jo = C.jn;
    float dj = -vn / C.meff;
    float nw = wmax(0.0f, jo + dj);
    dj = nw - jo;
    C.jn = nw;
    f3 Jn = nrm * dj;
    if (iMA > 0.0f) { vA = vA + Jn * iMA; wA = wA + mv(IA, cross(rA, Jn)); }
    if (hasB && iMB > 0.0f) { vB = vB - Jn * iMB; wB = wB - mv(IB, cross(rB, Jn)); }
    va = vA + cross(wA, rA);
    vb = mk(0.0f, 0.0f, 0.0f);
    if (hasB) vb = vB + cross(wB, rB);
    vrel = va - vb;
    if (FIXED) {
        meft1 = C.meft1; meft2 = C.meft2;
    } else {
        f3 rtA1 = cross(rA, t1);
        meft1 = iMA + dot(rtA1, mv(IA, rtA1));
        f3 rtA2 = cross(rA, t2);
        meft2 = iMA + dot(rtA2, mv(IA, rtA2));
        if (hasB) {
            f3 rtB1 = cross(rB, t1);
            meft1 = meft1 + iMB + dot(rtB1, mv(IB, rtB1));
            f3 rtB2 = cross(rB, t2);
            meft2 = meft2 + iMB + dot(rtB2, mv(IB, rtB2));
        }
    }
    if (meft1 >= 1e-12f && meft2 >= 1e-12f) {
        float o1 = C.jt1, o2 = C.jt2;
        float a1 = o1 - dot(vrel, t1) / meft1;
        float a2 = o2 - dot(vrel, t2) / meft2;
        float mag = sqrtf(a1 * a1 + a2 * a2);
        float lim = C.mu * nw;
        if (mag > lim && mag > 1e-12f) { a1 = a1 * lim / mag; a2 = a2 * lim / mag; }
        C.jt1 = a1;
        C.jt2 = a2;
        f3 Jt = t1 * (a1 - o1) + t2 * (t2d * (a2 - o2));
        if (iMA > 0.0f) { vA = vA + Jt * iMA; wA = wA + mv(IA, cross(rA, Jt)); }
        if (hasB && iMB > 0.0f) { vB = vB - Jt * iMB; wB = wB - mv(IB, cross(rB, Jt)); }
    }
}

// one contact of the chunk position sweep, as chunk_pos
__device__ __forceinline__ void cc_pos(CCon &C, f3 &pvA, f3 &poA, f3 &pvB, f3 &poB, bool hasB, float iMA, float iMB,
                                       const m3 &IA, const m3 &IB) {
    const f3 nrm = C.n, rA = C.rA, rB = C.rB;
    f3 rel = pvA + cross(poA, rA);
    if (hasB) rel = rel - (pvB + cross(poB, rB));
    if (!(C.meff >= 1e-12f)) return;
    float dj = (C.bias - dot(rel, nrm)) / C.meff;
    float jo = C.jp;
    float nw = wmax(0.0f, jo + dj);
    dj = nw - jo;
    C.jp = nw;
    f3 Jn = nrm * dj;
    if (iMA > 0.0f) { pvA = pvA + Jn * iMA; poA = poA + mv(IA, cross(rA, Jn)); }
    if (hasB && iMB > 0.0f) { pvB = pvB - Jn * iMB; poB = poB - mv(IB, cross(rB, Jn)); }
}

// ── fast arithmetic (ME_FLAG_FAST_MATH; ColorCache's fixed friction frame) ──
// The contact updates of cc_warm, cc_vel<true> and cc_pos in a shorter instruction chain: the bodies'
// angular responses to unit impulses (kA = IA (rA x n), ...) and the reciprocals of the effective
// masses come from the substep's constants (prepf_write), products are fused into FMAs, and the
// friction cone is scaled through the hardware's reciprocal square root. Equal to the exact updates
// in real arithmetic; the rounding differs, and the order of the updates (the Gauss-Seidel order) does
// not, so the result is the same from run to run.

// v + w x r
__device__ __forceinline__ f3 f_vel(f3 v, f3 w, f3 r) {
    return mk(fmaf(w.y, r.z, fmaf(-w.z, r.y, v.x)), fmaf(w.z, r.x, fmaf(-w.x, r.z, v.y)),
              fmaf(w.x, r.y, fmaf(-w.y, r.x, v.z)));
}
__device__ __forceinline__ float f_dot(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }

// warm apply: the impulse jn n + jt1 t1 + jt2 t2 in the sweep's frame
__device__ __forceinline__ void cc_warm_f(const CCon &C, f3 &vA, f3 &wA, f3 &vB, f3 &wB, bool hasB, float iMA,
                                          float iMB) {
    if (iMA > 0.0f) {
        vA = cmadd(C.jn * iMA, C.n, cmadd(C.jt1 * iMA, C.t1, cmadd(C.jt2 * iMA, C.t2, vA)));
        wA = cmadd(C.jn, C.kA, cmadd(C.jt1, C.k1A, cmadd(C.jt2, C.k2A, wA)));
    }
    if (hasB && iMB > 0.0f) {
        vB = cmadd(-C.jn * iMB, C.n, cmadd(-C.jt1 * iMB, C.t1, cmadd(-C.jt2 * iMB, C.t2, vB)));
        wB = cmadd(-C.jn, C.kB, cmadd(-C.jt1, C.k1B, cmadd(-C.jt2, C.k2B, wB)));
    }
}

// velocity pass: normal impulse, then the friction pair clamped to the cone mu * jn
__device__ __forceinline__ void cc_vel_f(CCon &C, f3 &vA, f3 &wA, f3 &vB, f3 &wB, bool hasB, float iMA,
                                         float iMB) {
    if (!(C.meff > 0.0f)) return;
    f3 vrel = f_vel(vA, wA, C.rA);
    if (hasB) vrel = vrel - f_vel(vB, wB, C.rB);
    const float jo = C.jn;
    const float nw = wmax(0.0f, fmaf(-f_dot(vrel, C.n), C.meff, jo));
    float dj = nw - jo;
    C.jn = nw;
    if (iMA > 0.0f) { vA = cmadd(dj * iMA, C.n, vA); wA = cmadd(dj, C.kA, wA); }
    if (hasB && iMB > 0.0f) { vB = cmadd(-dj * iMB, C.n, vB); wB = cmadd(-dj, C.kB, wB); }
    if (!(C.meft1 > 0.0f)) return;
    vrel = f_vel(vA, wA, C.rA);
    if (hasB) vrel = vrel - f_vel(vB, wB, C.rB);
    const float o1 = C.jt1, o2 = C.jt2;
    float a1 = fmaf(-f_dot(vrel, C.t1), C.meft1, o1);
    float a2 = fmaf(-f_dot(vrel, C.t2), C.meft2, o2);
    const float s2 = fmaf(a1, a1, a2 * a2), lim = C.mu * nw;
    if (s2 > lim * lim && s2 > 1e-24f) {
        const float f = lim * __builtin_amdgcn_rsqf(s2);
        a1 = a1 * f;
        a2 = a2 * f;
    }
    C.jt1 = a1;
    C.jt2 = a2;
    const float d1 = a1 - o1, d2 = a2 - o2;
    if (iMA > 0.0f) {
        vA = cmadd(d1 * iMA, C.t1, cmadd(d2 * iMA, C.t2, vA));
        wA = cmadd(d1, C.k1A, cmadd(d2, C.k2A, wA));
    }
    if (hasB && iMB > 0.0f) {
        vB = cmadd(-d1 * iMB, C.t1, cmadd(-d2 * iMB, C.t2, vB));
        wB = cmadd(-d1, C.k1B, cmadd(-d2, C.k2B, wB));
    }
}

// position pass
__device__ __forceinline__ void cc_pos_f(CCon &C, f3 &pvA, f3 &poA, f3 &pvB, f3 &poB, bool hasB, float iMA,
                                         float iMB) {
    if (!(C.meff > 0.0f)) return;
    f3 rel = f_vel(pvA, poA, C.rA);
    if (hasB) rel = rel - f_vel(pvB, poB, C.rB);
    const float jo = C.jp;
    const float nw = wmax(0.0f, fmaf(C.bias - f_dot(rel, C.n), C.meff, jo));
    const float dj = nw - jo;
    C.jp = nw;
    if (iMA > 0.0f) { pvA = cmadd(dj * iMA, C.n, pvA); poA = cmadd(dj, C.kA, poA); }
    if (hasB && iMB > 0.0f) { pvB = cmadd(-dj * iMB, C.n, pvB); poB = cmadd(-dj, C.kB, poB); }
}

// The same fast updates with the angular responses formed from the bodies' inverse inertias at each
// use (for the big and giant tiles, whose records stream from memory every colour step: four more
// float4 per contact would not stay in the L2). C.meff, meft1, meft2 hold the reciprocals.
__device__ __forceinline__ void cc_warm_c(const CCon &C, f3 &vA, f3 &wA, f3 &vB, f3 &wB, bool hasB, float iMA,
                                          float iMB, const m3 &IA, const m3 &IB) {
    // the warm apply's own frame, as cc_warm (the records carry the sweep's frame to the velocity pass only)
    const f3 t1 = cfixed_t1(C.n), t2 = ccross(C.n, t1);
    const f3 imp = cmadd(C.jn, C.n, cmadd(C.jt1, t1, C.jt2 * t2));
    if (iMA > 0.0f) { vA = cmadd(iMA, imp, vA); wA = wA + cmv(IA, ccross(C.rA, imp)); }
    if (hasB && iMB > 0.0f) { vB = cmadd(-iMB, imp, vB); wB = wB - cmv(IB, ccross(C.rB, imp)); }
}

__device__ __forceinline__ void cc_vel_c(CCon &C, f3 &vA, f3 &wA, f3 &vB, f3 &wB, bool hasB, float iMA,
                                         float iMB, const m3 &IA, const m3 &IB) {
    if (!(C.meff > 0.0f)) return;
    f3 vrel = f_vel(vA, wA, C.rA);
    if (hasB) vrel = vrel - f_vel(vB, wB, C.rB);
    const float jo = C.jn;
    const float nw = wmax(0.0f, fmaf(-f_dot(vrel, C.n), C.meff, jo));
    const float dj = nw - jo;
    C.jn = nw;
    if (iMA > 0.0f) { vA = cmadd(dj * iMA, C.n, vA); wA = cmadd(dj, cmv(IA, ccross(C.rA, C.n)), wA); }
    if (hasB && iMB > 0.0f) { vB = cmadd(-dj * iMB, C.n, vB); wB = cmadd(-dj, cmv(IB, ccross(C.rB, C.n)), wB); }
    if (!(C.meft1 > 0.0f)) return;
    vrel = f_vel(vA, wA, C.rA);
    if (hasB) vrel = vrel - f_vel(vB, wB, C.rB);
    const float o1 = C.jt1, o2 = C.jt2;
    float a1 = fmaf(-f_dot(vrel, C.t1), C.meft1, o1);
    float a2 = fmaf(-f_dot(vrel, C.t2), C.meft2, o2);
    const float s2 = fmaf(a1, a1, a2 * a2), lim = C.mu * nw;
    if (s2 > lim * lim && s2 > 1e-24f) {
        const float f = lim * __builtin_amdgcn_rsqf(s2);
        a1 = a1 * f;
        a2 = a2 * f;
    }
    C.jt1 = a1;
    C.jt2 = a2;
    const f3 Jt = cmadd(a1 - o1, C.t1, (a2 - o2) * C.t2);
    if (iMA > 0.0f) { vA = cmadd(iMA, Jt, vA); wA = wA + cmv(IA, ccross(C.rA, Jt)); }
    if (hasB && iMB > 0.0f) { vB = cmadd(-iMB, Jt, vB); wB = wB - cmv(IB, ccross(C.rB, Jt)); }
}

__device__ __forceinline__ void cc_pos_c(CCon &C, f3 &pvA, f3 &poA, f3 &pvB, f3 &poB, bool hasB, float iMA,
                                         float iMB, const m3 &IA, const m3 &IB) {
    if (!(C.meff > 0.0f)) return;
    f3 rel = f_vel(pvA, poA, C.rA);
    if (hasB) rel = rel - f_vel(pvB, poB, C.rB);
    const float jo = C.jp;
    const float nw = wmax(0.0f, fmaf(C.bias - f_dot(rel, C.n), C.meff, jo));
    const float dj = nw - jo;
    C.jp = nw;
    if (iMA > 0.0f) { pvA = cmadd(dj * iMA, C.n, pvA); poA = cmadd(dj, cmv(IA, ccross(C.rA, C.n)), poA); }
    if (hasB && iMB > 0.0f) { pvB = cmadd(-dj * iMB, C.n, pvB); poB = cmadd(-dj, cmv(IB, ccross(C.rB, C.n)), poB); }
}

// the cached pair: bodies from LDS (dynamic) or the cached static state
template <bool FAST>
__device__ __forceinline__ void pc_warm(const PairCache &P, float *sv, float *sw) {
    f3 va = P.dynA ? ld3(sv, P.la) : P.vAs, wa = P.dynA ? ld3(sw, P.la) : P.wAs;
    f3 vb = mk(0.0f, 0.0f, 0.0f), wb = mk(0.0f, 0.0f, 0.0f);
    if (P.hasB) { vb = P.dynB ? ld3(sv, P.lb) : P.vBs; wb = P.dynB ? ld3(sw, P.lb) : P.wBs; }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        if (FAST) cc_warm_f(P.c[k], va, wa, vb, wb, P.hasB, P.iMA, P.iMB);
        else cc_warm(P.c[k], va, wa, vb, wb, P.hasB, P.iMA, P.iMB, P.IA, P.IB);
    }
    if (P.iMA > 0.0f) { st3(sv, P.la, va); st3(sw, P.la, wa); }
    if (P.hasB && P.iMB > 0.0f) { st3(sv, P.lb, vb); st3(sw, P.lb, wb); }
}

template <bool FIXED, bool FAST>
__device__ __forceinline__ void pc_vel(PairCache &P, float *sv, float *sw, float t2d) {
    f3 vA = P.dynA ? ld3(sv, P.la) : P.vAs, wA = P.dynA ? ld3(sw, P.la) : P.wAs;
    f3 vB = mk(0.0f, 0.0f, 0.0f), wB = mk(0.0f, 0.0f, 0.0f);
    if (P.hasB) { vB = P.dynB ? ld3(sv, P.lb) : P.vBs; wB = P.dynB ? ld3(sw, P.lb) : P.wBs; }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        if (FAST) cc_vel_f(P.c[k], vA, wA, vB, wB, P.hasB, P.iMA, P.iMB);
        else cc_vel<FIXED>(P.c[k], vA, wA, vB, wB, P.hasB, P.iMA, P.iMB, P.IA, P.IB, t2d);
    }
    if (P.iMA > 0.0f) { st3(sv, P.la, vA); st3(sw, P.la, wA); }
    if (P.hasB && P.iMB > 0.0f) { st3(sv, P.lb, vB); st3(sw, P.lb, wB); }
}

template <bool FAST>
__device__ __forceinline__ void pc_pos(PairCache &P, float *spv, float *spo) {
    f3 pvA = P.dynA ? ld3(spv, P.la) : P.pvAs, poA = P.dynA ? ld3(spo, P.la) : P.poAs;
    f3 pvB = mk(0.0f, 0.0f, 0.0f), poB = mk(0.0f, 0.0f, 0.0f);
    if (P.hasB) { pvB = P.dynB ? ld3(spv, P.lb) : P.pvBs; poB = P.dynB ? ld3(spo, P.lb) : P.poBs; }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= P.cnt) break;
        if (FAST) cc_pos_f(P.c[k], pvA, poA, pvB, poB, P.hasB, P.iMA, P.iMB);
        else cc_pos(P.c[k], pvA, poA, pvB, poB, P.hasB, P.iMA, P.iMB, P.IA, P.IB);
    }
    if (P.iMA > 0.0f) { st3(spv, P.la, pvA); st3(spo, P.la, poA); }
    if (P.hasB && P.iMB > 0.0f) { st3(spv, P.lb, pvB); st3(spo, P.lb, poB); }
}

// workgroup barrier that orders LDS only: s_waitcnt lgkmcnt(0) + s_barrier, global stores stay in
// flight (for sweeps whose cross-thread state is all in LDS)
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ void tile_bounds(const TiledBuffers &T, int t, int &b0, int &nb, int &p0, int &np) {
    b0 = T.tb0[t];
    nb = T.tb1[t] - b0;
    p0 = T.tp0[t];
    np = T.tp1[t] - T.tp0[t];
    if (np < 0) np = 0;
}

__device__ __forceinline__ bool is_small_tile(int nb, int np) { return np <= T_SMALL_P && nb <= T_SMALL_B; }

// write back the LDS body state and integrate (k_integ, contact_engine_gpu.py:172-180)
template <int WG>
__device__ __forceinline__ void tile_writeback(const TiledArgs &A, const TiledBuffers &T, int b0, int nb, bool lds,
                                               const float *sv, const float *sw, const float *spv, const float *spo) {
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

// small tiles: thread t owns the t-th pair of the tile in colour order, its constants and impulses
// in registers; body velocities in LDS; one workgroup barrier per colour step
// sweep-start record of a small tile: its bodies' state (LDS) and this thread's pair's impulses
__device__ __forceinline__ void tape_sweep(const TiledArgs &A, const TiledBuffers &T, int b0, int nb, int sweep,
                                           const float *s0, const float *s1, const PairCache &P, int p) {
    float *st = A.tape.sst + (size_t)sweep * 6 * A.N;
    for (int j = threadIdx.x; j < nb; j += T_SMALL_P) {
        int i = T.bsorted[b0 + j];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            st[6 * (size_t)i + k] = s0[3 * j + k];
            st[6 * (size_t)i + 3 + k] = s1[3 * j + k];
        }
    }
    if (p >= 0) {
        int c0 = A.Pp.pstart[p];
        float *si = A.tape.simp + (size_t)sweep * 4 * A.capC;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= P.cnt) break;
            float *o = si + 4 * (size_t)(c0 + k);
            o[0] = P.c[k].jn; o[1] = P.c[k].jt1; o[2] = P.c[k].jt2; o[3] = P.c[k].jp;
        }
    }
}

// Small tiles by a persistent grid: each workgroup takes the next tile of the small-tile queue
// (k_t_bigsort: most pairs first) until the queue is empty, so a batch of uneven scenes finishes close
// to max(largest tile, total work / workgroups) whatever the number of tiles.
template <bool FIXED, bool TAPE, bool FAST = false>
__global__ void __launch_bounds__(T_SMALL_P) k_t_small(TiledArgs A, TiledBuffers T) {
    __shared__ float sv[3 * T_SMALL_B], sw[3 * T_SMALL_B], spv[3 * T_SMALL_B], spo[3 * T_SMALL_B];
    __shared__ int cnt[ME_NCOLMAX + 1], fill[ME_NCOLMAX + 1], plist[T_SMALL_P], plcol[T_SMALL_P];
    __shared__ int maxc, wdyn, wcon, qi;
    const int tid = threadIdx.x;
    const int nq = A.istat[IS_NSMALL];
    for (;;) {
        if (tid == 0) qi = atomicAdd(&A.istat[IS_SMALLQ], 1);
        __syncthreads();
        const int iq = qi;
        __syncthreads();  // every thread has read qi before thread 0 takes the next entry
        if (iq >= nq) break;
        const int t = T.smallq[iq];
        int b0, nb, p0, np;
        tile_bounds(T, t, b0, nb, p0, np);
        if (tid == 0) {
            atomicAdd(&A.istat[IS_TILES], 1);
            atomicMax(&A.istat[IS_MAXTB], nb);
            maxc = 0;
            wdyn = 0;
            wcon = 0;
        }
        int ncol = A.istat[IS_NCOL];
        if (ncol > ME_NCOLMAX) ncol = ME_NCOLMAX;
        for (int j = tid; j < nb; j += T_SMALL_P) {
            int i = T.bsorted[b0 + j];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                sv[3 * j + k] = A.b.v[3 * i + k];
                sw[3 * j + k] = A.b.w[3 * i + k];
                spv[3 * j + k] = 0.0f;
                spo[3 * j + k] = 0.0f;
            }
        }
        for (int c = tid; c <= ncol; c += T_SMALL_P) { cnt[c] = 0; fill[c] = 0; }
        __syncthreads();
        int myc = -1;
        if (tid < np) {
            myc = A.Pp.col[T.psorted[p0 + tid]];
            if (myc < 0 || myc >= ncol) myc = ncol > 0 ? ncol - 1 : 0;  // only after a flagged colour overflow
            myc = tile_class(T, myc, ncol);
            atomicAdd(&cnt[myc], 1);
            atomicMax(&maxc, myc + 1);
        }
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            for (int c = 0; c < maxc; ++c) { int n = cnt[c]; cnt[c] = run; run += n; }
        }
        __syncthreads();
        if (tid < np) {
            int pos = cnt[myc] + atomicAdd(&fill[myc], 1);
            plist[pos] = T.psorted[p0 + tid];
            plcol[pos] = myc;
        }
        __syncthreads();
        const int nc = maxc;
        PairCache P;
        int p = -1, c_own = -1;
        if (tid < np) {
            p = plist[tid];
            c_own = plcol[tid];
            pc_load<FIXED, FAST>(P, A, T, p, b0);
        }
        if (np > 0) {
            if (FIXED) {
                for (int c = 0; c < nc; ++c) {
                    if (c_own == c) pc_warm<FAST>(P, sv, sw);
                    lds_barrier();
                }
            }
            const float t2d = FIXED ? 1.0f : A.t2d;
            for (int it = 0; it < A.vit; ++it) {
                if (TAPE) tape_sweep(A, T, b0, nb, it, sv, sw, P, p);
                for (int c = 0; c < nc; ++c) {
                    if (c_own == c) pc_vel<FIXED, FAST>(P, sv, sw, t2d);
                    lds_barrier();
                }
            }
            for (int it = 0; it < A.pit; ++it) {
                if (TAPE) tape_sweep(A, T, b0, nb, A.vit + it, spv, spo, P, p);
                for (int c = 0; c < nc; ++c) {
                    if (c_own == c) pc_pos<FAST>(P, spv, spo);
                    lds_barrier();
                }
            }
        } else if (TAPE) {
            for (int it = 0; it < A.vit; ++it) tape_sweep(A, T, b0, nb, it, sv, sw, P, -1);
            for (int it = 0; it < A.pit; ++it) tape_sweep(A, T, b0, nb, A.vit + it, spv, spo, P, -1);
        }
        if (p >= 0) {
            pc_store(P, A, p);
            atomicAdd(&wdyn, (int)P.dynA + (int)P.dynB);
            atomicAdd(&wcon, P.cnt);
        }
        __syncthreads();
        if (tid == 0) {  // work of this tile for the algorithmic-byte count of the launch
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[4]), 1ull);
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[5]), (unsigned long long)np);
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[6]), (unsigned long long)wdyn);
            atomicAdd(reinterpret_cast<unsigned long long *>(&T.bigw[7]), (unsigned long long)wcon);
        }

        tile_writeback<T_SMALL_P>(A, T, b0, nb, true, sv, sw, spv, spo);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// per-contact constants of the substep (k_t_prep): the same expressions pc_load evaluates; under fast
// arithmetic (FAST, with the fixed frame) the prepf record instead
template <bool FIXED, bool FAST>
__global__ void k_t_prep(TiledArgs A, TiledBuffers T) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int C = A.istat[IS_C];
    if (C > A.capC) C = A.capC;
    if (c >= C) return;
    int a = A.Cc.bi[c], b = A.Cc.bj[c];
    const float *invM = A.b.invM;
    float iMA = invM[a], iMB = b >= 0 ? invM[b] : 0.0f;
    m3 IA = ldm(A.invIw, a), IB = b >= 0 ? ldm(A.invIw, b) : IA;
    f3 n = ld3(A.Cc.n, c), rA = ld3(A.Cc.pA, c) - ld3(A.b.x, a);
    f3 rB = b >= 0 ? ld3(A.Cc.pB, c) - ld3(A.b.x, b) : mk(0.0f, 0.0f, 0.0f);
    if (FAST) {
        const f3 z = mk(0.0f, 0.0f, 0.0f);
        f3 rnA = cross(rA, n), kA = mv(IA, rnA), kB = z;
        float meff = iMA + dot(rnA, kA);
        f3 t1 = fixed_t1(n), t2 = cross(n, t1);
        f3 r1A = cross(rA, t1), r2A = cross(rA, t2);
        f3 k1A = mv(IA, r1A), k2A = mv(IA, r2A), k1B = z, k2B = z;
        float m1 = iMA + dot(r1A, k1A), m2 = iMA + dot(r2A, k2A);
        if (b >= 0) {
            f3 rnB = cross(rB, n), r1B = cross(rB, t1), r2B = cross(rB, t2);
            kB = mv(IB, rnB); k1B = mv(IB, r1B); k2B = mv(IB, r2B);
            meff = meff + iMB + dot(rnB, kB);
            m1 = m1 + iMB + dot(r1B, k1B);
            m2 = m2 + iMB + dot(r2B, k2B);
        }
        const float bias = 0.2f * wmax(A.Cc.pen[c] - 1e-4f, 0.0f) / A.dt;
        float4 *of = reinterpret_cast<float4 *>(T.prepf) + 10 * (size_t)c;
        prepf_write(of, n, rA, rB, t1, t2, meff, m1, m2, bias, A.Cc.mu[c], kA, kB, k1A, k2A, k1B, k2B);
        // the big and giant tiles' records: prep's layout with the reciprocals of prepf
        const float4 r0 = of[0], r3 = of[3], r4 = of[4];
        float *o = T.prep + 20 * (size_t)c;
        o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = rA.x; o[4] = rA.y; o[5] = rA.z; o[6] = rB.x; o[7] = rB.y;
        o[8] = rB.z; o[9] = t1.x; o[10] = t1.y; o[11] = t1.z; o[12] = t2.x; o[13] = t2.y; o[14] = t2.z;
        o[15] = r0.w; o[16] = r3.w; o[17] = r4.w; o[18] = bias; o[19] = A.Cc.mu[c];
        return;
    }
    f3 rnA = cross(rA, n);
    float meff = iMA + dot(rnA, mv(IA, rnA));
    if (b >= 0) {
        f3 rnB = cross(rB, n);
        meff = meff + iMB + dot(rnB, mv(IB, rnB));
    }
    f3 t1 = fixed_t1(n), t2 = cross(n, t1);
    float m1 = 0.0f, m2 = 0.0f;
    if (FIXED) {
        f3 rtA1 = cross(rA, t1);
        m1 = iMA + dot(rtA1, mv(IA, rtA1));
        f3 rtA2 = cross(rA, t2);
        m2 = iMA + dot(rtA2, mv(IA, rtA2));
        if (b >= 0) {
            f3 rtB1 = cross(rB, t1);
            m1 = m1 + iMB + dot(rtB1, mv(IB, rtB1));
            f3 rtB2 = cross(rB, t2);
            m2 = m2 + iMB + dot(rtB2, mv(IB, rtB2));
        }
    }
    float *o = T.prep + 20 * (size_t)c;
    o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = rA.x; o[4] = rA.y; o[5] = rA.z; o[6] = rB.x; o[7] = rB.y; o[8] = rB.z;
    o[9] = t1.x; o[10] = t1.y; o[11] = t1.z; o[12] = t2.x; o[13] = t2.y; o[14] = t2.z;
    o[15] = meff; o[16] = m1; o[17] = m2;
    o[18] = 0.2f * wmax(A.Cc.pen[c] - 1e-4f, 0.0f) / A.dt;
    o[19] = A.Cc.mu[c];
}

#include "tiled_big.hpp"

inline int tiled_scan(const TiledArgs &A, int site, const int *in, int *out, int n, int *total, hipStream_t s) {
    int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb < 1) nb = 1;
    if (nb > A.nbsum) return ME_ERUNTIME;
    hipLaunchKernelGGL(k_scan1<1>, dim3(nb), dim3(ME_THREADS), 0, s, in, out, nullptr, nullptr, n, n, total, nullptr,
                       A.scan_words + (size_t)site * (1 + 2 * (size_t)A.nbsum), A.istat, A.giant_wait_ticks, 0, 0);
    return hipGetLastError() == hipSuccess ? ME_OK : ME_ERUNTIME;
}

inline int tblk(long long n) { return (int)((n + ME_THREADS - 1) / ME_THREADS); }

// JP rounds of one small tile (the rounds of t_color_tile): one pair per thread, its key and body
// slots kept in registers across the rounds, claims in static LDS
__global__ void __launch_bounds__(T_SMALL_P) k_t_color_small(TiledArgs A, TiledBuffers T) {
    if (A.istat[IS_MISMATCH] != 1) return;
    int t = blockIdx.x;
    int b0 = T.tb0[t];
    if (b0 == INT_MAX) return;
    int nb = T.tb1[t] - b0, p0 = T.tp0[t], np = T.tp1[t] - T.tp0[t];
    if (np <= 0 || !is_small_tile(nb, np)) return;
    __shared__ unsigned long long sclaim[T_SMALL_B];
    __shared__ int left;
    const int tid = threadIdx.x;
    const float *invM = A.b.invM;
    int p = -1, la = -1, lb = -1, col = -1;
    unsigned long long key = 0;
    if (tid < np) {
        p = T.psorted[p0 + tid];
        key = (unsigned long long)jp_key(p);
        int a = A.Pp.pbi[p], b = A.Pp.pbj[p];
        if (invM[a] > 0.0f) la = T.bslot[a] - b0;
        if (b >= 0 && invM[b] > 0.0f) lb = T.bslot[b] - b0;
    }
    int nround = 0;
    for (int cur = 0; cur <= np; ++cur) {
        for (int j = tid; j < nb; j += T_SMALL_P) sclaim[j] = ~0ull;
        if (tid == 0) left = 0;
        __syncthreads();
        if (p >= 0 && col < 0) {
            if (la >= 0) atomicMin(&sclaim[la], key);
            if (lb >= 0) atomicMin(&sclaim[lb], key);
        }
        __syncthreads();
        if (p >= 0 && col < 0) {
            bool ok = (la < 0 || sclaim[la] == key) && (lb < 0 || sclaim[lb] == key);
            if (ok) col = cur;
            else atomicAdd(&left, 1);
        }
        __syncthreads();
        int l = left;
        __syncthreads();
        if (l == 0) { nround = cur + 1; break; }
    }
    // class sizes: a histogram in LDS, then one global add per class and tile
    __shared__ int hist[ME_NCOLMAX];
    for (int c = tid; c < nround && c < ME_NCOLMAX; c += T_SMALL_P) hist[c] = 0;
    __syncthreads();
    if (p >= 0) {
        A.Pp.col[p] = col;
        if (col >= 0 && col < ME_NCOLMAX) atomicAdd(&hist[col], 1);
        else atomicOr(&A.istat[IS_OVF], OVF_COLORS);
    }
    __syncthreads();
    for (int c = tid; c < nround && c < ME_NCOLMAX; c += T_SMALL_P)
        if (hist[c]) atomicAdd(&T.ccount_jp[c], hist[c]);
    if (tid == 0) atomicMax(&A.istat[IS_NCOLJP], nround);
}

// lists of the tiles that exceed one small tile (the big-tile kernel's work), and of those beyond the
// big tile's LDS and the LDS colouring (the grid kernels' work, usually none); and the big tiles' body
// constants in island order (position j <- body bsorted[j]: coalesced writes, gathered reads), 48-byte
// records of invIw for three 16-byte loads
__global__ void k_t_biglist(int ntile, int *istat, const float *invM, const float *invIw, TiledBuffers T) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 5) T.gbar[t] = 0u;  // the giant-tile grid barriers and round counters of this substep start at zero
    for (int j = t; j < 2 * (ME_NCOLMAX + 1); j += gridDim.x * blockDim.x) T.gcnt[j] = 0;  // its class totals
    if (t == 0) istat[IS_CCDIFF] = 0;  // read for the last time by k_t_mode: clear for the next substep
    if (t < istat[IS_NBODYSORT]) {
        int i = T.bsorted[t];
        T.invMs[t] = invM[i];
        float I[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) I[q] = invIw[9 * i + q];
        float4 *d = reinterpret_cast<float4 *>(T.invIws + 12 * (size_t)t);
        d[0] = make_float4(I[0], I[1], I[2], I[3]);
        d[1] = make_float4(I[4], I[5], I[6], I[7]);
        d[2] = make_float4(I[8], 0.0f, 0.0f, 0.0f);
    }
    if (t >= ntile || T.tb0[t] == INT_MAX) return;
    int b0, nb, p0, np;
    tile_bounds(T, t, b0, nb, p0, np);
    if (is_small_tile(nb, np)) {
        T.smalllist[atomicAdd(&istat[IS_NSMALL], 1)] = t;
        return;
    }
    T.biglist[atomicAdd(&istat[IS_NBIG], 1)] = t;
    if (!(nb <= T_BIG_B && np <= T_BIG_P)) T.giantlist[atomicAdd(&istat[IS_NGIANT], 1)] = t;
    if (np > 0 && !(nb <= T_COLB && np <= T_COLP)) T.gcollist[atomicAdd(&istat[IS_NGCOL], 1)] = t;
}

// the big- and small-tile queues: each list's tiles ranked by pair count, largest first (ties by tile
// index, so the order does not depend on k_t_biglist's atomics). Workgroups that take tiles from the
// front in turn finish the launch close to max(largest tile, total work / workgroups) -- the
// longest-processing-time rule -- where a fixed stride could stack two large tiles on one workgroup.
// One thread per entry: the big list's first, then the small list's.
// grid (tiles / 256, 2): y = 0 ranks the big list, y = 1 the small list; a list of up to T_SORT_LDS
// entries is staged in LDS first (its weights and tile ids), so each thread's pass over it reads LDS
#define T_SORT_LDS 4096
__global__ void __launch_bounds__(256) k_t_bigsort(int ntile, const int *istat, TiledBuffers T) {
    __shared__ int wl[T_SORT_LDS], ul[T_SORT_LDS];
    const bool big = blockIdx.y == 0;
    int n = big ? istat[IS_NBIG] : istat[IS_NSMALL];
    if (n > ntile) n = ntile;
    if ((int)(blockIdx.x * blockDim.x) >= n) return;
    const int *list = big ? T.biglist : T.smalllist;
    const bool lds = n <= T_SORT_LDS;
    if (lds) {
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const int u = list[j];
            ul[j] = u;
            wl[j] = T.tp1[u] - T.tp0[u];
        }
        __syncthreads();
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = list[i];
    const int w = T.tp1[t] - T.tp0[t];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
        const int u = lds ? ul[j] : list[j];
        const int wu = lds ? wl[j] : T.tp1[u] - T.tp0[u];
        rank += (wu > w || (wu == w && u < t)) ? 1 : 0;
    }
    (big ? T.bigq : T.smallq)[rank] = t;
}

// dynamic-LDS limits of the tile kernels beyond the 64 KB default; set on the current device before
// the first launch (at engine creation, so that no attribute call falls inside a stream capture)
inline int tiled_set_attributes() {
    if (hipFuncSetAttribute((const void *)k_t_big<ME_SOLVER_PAIR_CHUNKS>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)T_BIG_DYN) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_t_big<ME_SOLVER_COLOR_CACHE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)T_BIG_DYN) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_t_big<ME_SOLVER_COLOR_CACHE, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)T_BIG_DYN) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_t_color<1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)T_COL_DYN) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_t_islands_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)T_ISL_DYN) != hipSuccess)
        return ME_ERUNTIME;
    return ME_OK;
}

// timing event of kernel `id` on stream s (no-op when the engine is not timing); non-zero on failure
inline int prof_rec(const TiledArgs &A, int id, hipStream_t s) {
    if (!A.prof_ev) return 0;
    hipEvent_t ev = (hipEvent_t)A.prof_ev(A.prof, id);
    return ev && hipEventRecord(ev, s) != hipSuccess;
}

// small tiles and then giant islands on stream s, big tiles on the side stream s2 (disjoint islands),
// joined on s; FAST: fast arithmetic (ColorCache only, never recorded)
template <int SOLVER, bool FAST = false>
inline int launch_solve(const TiledArgs &A, TiledBuffers &T, int ntile, hipStream_t s) {
    static_assert(!FAST || SOLVER == ME_SOLVER_COLOR_CACHE, "fast arithmetic takes the fixed friction frame");
    hipStream_t s2 = (hipStream_t)A.side;
    hipEvent_t fork = (hipEvent_t)A.ev_fork, join = (hipEvent_t)A.ev_join;
    const size_t dyn = T_BIG_DYN;
    if (FAST && A.taped) return ME_ERUNTIME;
    hipLaunchKernelGGL((k_t_prep<SOLVER == ME_SOLVER_COLOR_CACHE, FAST>), dim3(tblk(A.capC)), dim3(ME_THREADS), 0, s,
                       A, T);
    // the big tiles' colour order and records, spread over the GPU before either tile kernel starts (on
    // the side stream they would wait for the small tiles' persistent workgroups to leave the CUs)
    hipLaunchKernelGGL(k_t_bigord, dim3(ntile < A.big_grid ? ntile : A.big_grid), dim3(1024), 0, s, A, T);
    hipLaunchKernelGGL(k_t_bigfill, dim3(ntile < A.big_grid ? ntile : A.big_grid, T_FILL_CHUNKS), dim3(T_FILL_WG), 0,
                       s, A, T);
    if (hipEventRecord(fork, s) != hipSuccess || hipStreamWaitEvent(s2, fork, 0) != hipSuccess) return ME_ERUNTIME;
    // the big tiles (the critical path when there are any) are submitted first, so that their
    // workgroups find CUs with free LDS before the small tiles fill them
    if (prof_rec(A, ME_K_TILE_BIG, s2)) return ME_ERUNTIME;
    hipLaunchKernelGGL((k_t_big<SOLVER, FAST>), dim3(ntile < A.big_grid ? ntile : A.big_grid), dim3(T_BIGT_WG), dyn, s2,
                       A, T);
    if (prof_rec(A, ME_K_TILE_BIG, s2) || prof_rec(A, ME_K_TILE_SMALL, s)) return ME_ERUNTIME;
    if (A.taped)
        hipLaunchKernelGGL((k_t_small<SOLVER == ME_SOLVER_COLOR_CACHE, true>), dim3(ntile < A.small_grid ? ntile : A.small_grid),
                           dim3(T_SMALL_P), 0, s, A, T);
    else
        hipLaunchKernelGGL((k_t_small<SOLVER == ME_SOLVER_COLOR_CACHE, false, FAST>),
                           dim3(ntile < A.small_grid ? ntile : A.small_grid), dim3(T_SMALL_P), 0, s, A, T);
    if (prof_rec(A, ME_K_TILE_SMALL, s)) return ME_ERUNTIME;
    // a tile beyond the big tile's LDS (more than T_BIG_B bodies or T_BIG_P pairs) is possible: the grid
    // sweep, which k_t_big leaves such tiles to, after the small tiles on their stream, so that it runs
    // beside k_t_big and its workgroups find the CUs that the small tiles' persistent workgroups leave.
    // Measured on a 100 000-box pile, per substep: a giant island's sweep after k_t_big on the side
    // stream, 6.47 ms for the tile sweeps; here, 5.90 ms; on a third stream beside both, 5.56 ms, but
    // that stream's fork and join cost the 10 000-box pile (no giant island) about 10 us per substep,
    // which here the empty launch hides behind k_t_big.
    if (A.giant_possible && (A.N > T_BIG_B || A.capP > T_BIG_P)) {
        if (prof_rec(A, ME_K_TILE_GIANT, s)) return ME_ERUNTIME;
        hipLaunchKernelGGL((k_t_giant<SOLVER, FAST>), dim3(A.giant_grid), dim3(T_BIG_WG), 0, s, A, T);
        if (prof_rec(A, ME_K_TILE_GIANT, s)) return ME_ERUNTIME;
    }
    if (hipEventRecord(join, s2) != hipSuccess || hipStreamWaitEvent(s, join, 0) != hipSuccess) return ME_ERUNTIME;
    return ME_OK;
}

// the coloured solve + integration of one substep, in three parts so that the sweep kernels can be
// timed on their own: prepare (warm start, cache, islands, tiles, colouring), sweeps, finish
inline int tiled_ntile(const TiledArgs &A) {
    int TB = A.tile_bodies > 0 ? A.tile_bodies : 128;
    if (TB > T_SMALL_B) TB = T_SMALL_B;
    return (A.N + TB - 1) / TB;
}

inline int tiled_prepare(TiledBuffers &T, const TiledArgs &A, hipStream_t s) {
    int N = A.N, capP = A.capP, capC = A.capC;
    int TB = A.tile_bodies > 0 ? A.tile_bodies : 128;
    if (TB > T_SMALL_B) TB = T_SMALL_B;
    int ntile = tiled_ntile(A);
    bool cc = A.solver == ME_SOLVER_COLOR_CACHE;
    int g = capP > N ? capP : N;
    // pv, po were zeroed by k_body_pre; the warm store still holds the previous substep's contacts
    bool lds_isl = N <= T_ISL_MAXN;
    int gs = g > capC ? g : capC;
    if (gs < ME_NCOLMAX + 1) gs = ME_NCOLMAX + 1;
    hipLaunchKernelGGL(k_t_setup, dim3(tblk(gs)), dim3(ME_THREADS), 0, s, A, T, lds_isl ? 1 : 0, cc ? 1 : 0);
    hipLaunchKernelGGL(k_t_mode, dim3(tblk(capP)), dim3(ME_THREADS), 0, s, capP, A.istat, cc ? 1 : 0, A.cache_col,
                       A.Pp.col);
    if (lds_isl) {
        if (prof_rec(A, ME_K_ISLANDS, s)) return ME_ERUNTIME;
        hipLaunchKernelGGL(k_t_islands_lds, dim3(1), dim3(T_ISL_THREADS), 3 * N * sizeof(int), s, N, capP, TB, A.istat,
                           A.Pp.pbi, A.Pp.pbj, A.b.invM, T);
        if (prof_rec(A, ME_K_ISLANDS, s)) return ME_ERUNTIME;
    } else {
        hipLaunchKernelGGL(k_t_union, dim3(tblk(capP)), dim3(ME_THREADS), 0, s, capP, A.istat, A.Pp.pbi, A.Pp.pbj,
                           A.b.invM, T.parent);
        hipLaunchKernelGGL(k_t_flatten, dim3((N + T_AGG - 1) / T_AGG), dim3(T_AGG), 0, s, N, A.b.invM, T);
        hipLaunchKernelGGL(k_t_pair_root, dim3((capP + T_AGG - 1) / T_AGG), dim3(T_AGG), 0, s, capP, A.istat, A.Pp.pbi, A.Pp.pbj,
                           A.b.invM, T);
        int rc = tiled_scan(A, SC_TILED_B, T.cnt_b, T.boff, N, A.istat + IS_NBODYSORT, s);
        if (rc) return rc;
        rc = tiled_scan(A, SC_TILED_P, T.cnt_p, T.poff, N, A.istat + IS_NPAIRSORT, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_t_scatter_b, dim3((N + T_AGG - 1) / T_AGG), dim3(T_AGG), 0, s, N, T);
        hipLaunchKernelGGL(k_t_scatter_p, dim3((capP + T_AGG - 1) / T_AGG), dim3(T_AGG), 0, s, capP, A.istat, T);
        hipLaunchKernelGGL(k_t_tiles, dim3(tblk(N)), dim3(ME_THREADS), 0, s, N, TB, A.istat, T);
    }
    hipLaunchKernelGGL(k_t_biglist, dim3(tblk(ntile > N ? ntile : N)), dim3(ME_THREADS), 0, s, ntile, A.istat, A.b.invM,
                       A.invIw, T);
    hipLaunchKernelGGL(k_t_bigsort, dim3(tblk(ntile), 2), dim3(ME_THREADS), 0, s, ntile, A.istat, T);
    hipLaunchKernelGGL(k_t_color_small, dim3(ntile), dim3(T_SMALL_P), 0, s, A, T);
    // a tile beyond the LDS colouring limits (possible only in a scene of that size): coloured by a grid
    // of workgroups
    const int giant = A.gcol_possible && N > T_COLB;
    if (prof_rec(A, ME_K_COLOR, s)) return ME_ERUNTIME;
    hipLaunchKernelGGL(k_t_color<1024>, dim3(ntile < A.big_grid ? ntile : A.big_grid), dim3(1024), T_COL_DYN, s, A, T,
                       giant);
    if (prof_rec(A, ME_K_COLOR, s)) return ME_ERUNTIME;
    if (giant) hipLaunchKernelGGL(k_t_color_giant, dim3(A.giant_grid), dim3(T_BIG_WG), 0, s, A, T);
    hipLaunchKernelGGL(k_t_plan, dim3(1), dim3(64), 0, s, A.istat, A.n_colors, T);
    hipLaunchKernelGGL(k_t_apply_split, dim3(tblk(capP)), dim3(ME_THREADS), 0, s, capP, A.istat, A.Pp.col, T);
    return hipGetLastError() == hipSuccess ? ME_OK : ME_ERUNTIME;
}

inline int tiled_sweeps(TiledBuffers &T, const TiledArgs &A, hipStream_t s) {
    int ntile = tiled_ntile(A);
    if (A.solver == ME_SOLVER_COLOR_CACHE)
        return A.fast ? launch_solve<ME_SOLVER_COLOR_CACHE, true>(A, T, ntile, s)
                      : launch_solve<ME_SOLVER_COLOR_CACHE>(A, T, ntile, s);
    return launch_solve<ME_SOLVER_PAIR_CHUNKS>(A, T, ntile, s);
}

inline int tiled_finish(TiledBuffers &T, const TiledArgs &A, hipStream_t s) {
    int N = A.N, capP = A.capP, capC = A.capC;
    int g = capP > N ? capP : N;
    bool cc = A.solver == ME_SOLVER_COLOR_CACHE;
    if (A.warm || cc) {
        int gg = g > capC ? g : capC;
        hipLaunchKernelGGL(k_t_finish, dim3(tblk(gg)), dim3(ME_THREADS), 0, s, N, capP, capC, A.warm, cc ? 1 : 0,
                           A.istat, A.Cc.key, A.Cc.jn, A.Cc.jt1, A.Cc.jt2, A.wkey, A.wjn, A.wjt1, A.wjt2, A.Pp.pbi,
                           A.Pp.pbj, A.Pp.col, A.b.invM, A.cache_a, A.cache_b, A.cache_col, A.cache_invM);
    }
    return hipGetLastError() == hipSuccess ? ME_OK : ME_ERUNTIME;
}

}  // namespace me

struct me_engine;
me::TiledArgs tiled_args_of(me_engine *e, float sdt);
