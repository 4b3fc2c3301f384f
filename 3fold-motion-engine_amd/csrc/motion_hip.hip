// libmotion_hip.so: the C ABI of include/motion_hip.h and the host orchestration of one contact
// step on one MI355X. All device memory is the caller's (body arrays + one workspace block).
#include "../../include/motion_hip.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kernels.hpp"
#include "tiled.hpp"
#include "beads.hpp"
#include "ncp.hpp"
#include "ncp_svd.hpp"
#include "pgs.hpp"
#include "adjoint_kernels.hpp"
#include "giant_grid.hpp"

using namespace me;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return fail(ME_ERUNTIME, std::string(#expr) + ": " + hipGetErrorString(_e));         \
    } while (0)

#define LAUNCH_CHECK()                                                                           \
    do {                                                                                         \
        hipError_t _e = hipGetLastError();                                                       \
        if (_e != hipSuccess) return fail(ME_ERUNTIME, std::string("kernel launch: ") + hipGetErrorString(_e)); \
    } while (0)

inline int nblk(long long n) { return (int)((n + ME_THREADS - 1) / ME_THREADS); }
inline int pow2_at_least(long long n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// bump allocator over the workspace (sizes in bytes, 256-byte aligned)
struct Bump {
    char *base;
    size_t off = 0;
    StateSegs *segs = nullptr;  // when set, every take is recorded as a state segment
    template <class T>
    T *take(size_t count) { return takec<T>(count, -1, 1); }
    // count/per items of per elements each, of which status word `slot` counts the used items (-1: all)
    template <class T>
    T *takec(size_t count, int slot, size_t per) {
        off = (off + 255) & ~size_t(255);
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        if (segs && segs->n < ME_STATE_SEGS) {
            StateSeg &g = segs->s[segs->n++];
            g.off = off;
            g.bytes = sizeof(T) * (count ? count : 1);
            g.unit = sizeof(T) * per;
            g.slot = slot;
        } else if (segs) {
            segs->overflow = 1;
        }
        off += sizeof(T) * (count ? count : 1);
        return p;
    }
};

}  // namespace

struct me_engine {
    me_config cfg;
    me_bodies b;
    int N, P, NP, T;
    int cap_cand, capC, capP;
    std::thread::id owner;
    // per body
    float *invIw, *Rw, *aabb, *pv, *po, *vsnap, *xsnap, *wsnap, *allp;
    // Jacobi: per-body entry counts and offsets, each contact's entry at its two bodies, the entries
    int *jcnt, *joff, *jposA, *jposB;
    int64_t *jd;
    unsigned long long *claim;
    // grid
    int *cellc, *bucket, *bcount, *bstart, *bcursor, *items;
    float4 *spos;  // bead grid: position and id of each item in bucket order
    int4 *scell;   // body grid: cell + scene of each item in bucket order
    float4 *sbox;  // body grid: AABB lo + invM, hi + id of each item in bucket order
    // candidates and narrow-phase slots
    int *cand_n, *cand_off, *cand_tmp, *cand_a, *cand_b;
    Slots S;
    int *slot_off, *pair_flag, *pair_off;
    // contacts, pairs
    Contacts Cc;
    Pairs Pp;
    int *cache_a, *cache_b, *cache_col;
    float *cache_invM;
    int *ccount, *cstart, *cfill, *csub, *cbase;
    // warm store
    int64_t *wkey;
    float *wjn, *wjt1, *wjt2;
    // bead geometry under a coloured solver: body-pair broad phase, per-pair bead tests, manifold stage
    bool bead_pairs = false;
    int bp_words = 0;      // 64-slot ballot words per candidate pair
    float bhalf[3] = {0.0f, 0.0f, 0.0f};  // half extents of the bead spheres' box in the body frame
    float *bp_half;        // [N,3] filled with bhalf every substep
    uint64_t *bp_mask;     // [cap_cand * bp_words]
    int *bp_kn, *bp_koff;  // kept contacts per candidate pair, their offsets
    BeadRaw raw;           // generated contacts, pair-major, before the reduction
    // tiled solve
    TiledBuffers tb;
    // scan scratch + status
    int *bsum;
    int nbsum;
    unsigned long long *scan_words;  // the one-pass scans' per-site ticket counter and block words
    int *istat;
    int *hstat;  // pinned host mirror
    size_t state_bytes = 0;  // leading workspace bytes that carry state between substeps
    StateSegs segs{};        // the arrays of those bytes and the status words counting their used items
    hipStream_t side = nullptr;  // second stream for the big-tile sweeps (default priority)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // whole-step graphs (ME_FLAG_GRAPH): captured and launched on gs, keyed by (dt, substeps)
    struct Graph {
        double dt;
        int substeps;
        hipGraphExec_t exec;
    };
    hipStream_t gs = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // recording of the current substep for its adjoint (me_step_taped / me_adjoint_substep)
    bool taping = false;
    AdjTape tape{};
    std::vector<Graph> graphs;
    int64_t n_captures = 0;
    // HIP-event timing of selected kernels (measurement only)
    bool prof_on = false;
    std::vector<hipEvent_t> prof_ev[ME_K_COUNT_];
    int prof_used[ME_K_COUNT_] = {0};
    // giant-island grid kernels: workgroups (all resident together) and the barrier's wait bound
    int giant_grid = 0;
    long long giant_wait_ticks = 0;
    int big_grid = 0;  // workgroups of k_t_big / k_t_color: one per CU their LDS allows
    int small_grid = 0;  // workgroups of the persistent k_t_small: as many as can be resident
    int small_grid_fast = 0;  // the same for its fast-arithmetic variant
    // whether the largest scene could hold an island beyond the big tile's LDS (beyond the LDS colouring
    // limits): islands never cross scenes, so below these sizes the giant-island kernels are not launched
    int giant_possible = 1, gcol_possible = 1;
};

namespace {

const int kProfMax = 8192;

// next timing event of kernel `id` (two per launch, the first kProfMax launches); null when off
hipEvent_t prof_next(me_engine *e, int id) {
    if (!e->prof_on || e->prof_used[id] >= 2 * kProfMax) return nullptr;
    std::vector<hipEvent_t> &v = e->prof_ev[id];
    if ((int)v.size() <= e->prof_used[id]) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) return nullptr;
        v.push_back(ev);
    }
    return v[e->prof_used[id]++];
}

// event pair around one launch of kernel `id` on stream s
void prof_mark(me_engine *e, int id, hipStream_t s) {
    hipEvent_t ev = prof_next(e, id);
    if (ev) (void)hipEventRecord(ev, s);
}

// the tiled solve's hook for its own kernels (TiledArgs::prof_ev)
void *prof_event(void *eng, int id) { return (void *)prof_next((me_engine *)eng, id); }

void layout(me_engine *e, char *base, size_t *bytes) {
    const me_config &c = e->cfg;
    int N = c.n_bodies, P = c.n_beads > 0 ? c.n_beads : 0;
    long long NP = (long long)N * (P > 0 ? P : 1);
    e->N = N; e->P = P; e->NP = (int)NP;
    // bead geometry: the Jacobi solver and the per-contact coloured solver take contacts from a grid
    // of bead points (feature-major, unreduced); the pair-chunk solvers take them per body pair from the
    // body broad phase (pair-major, then reduced)
    bool bp = P > 0 && c.solver != ME_SOLVER_JACOBI && !(c.flags & ME_FLAG_PER_CONTACT);
    bool point_grid = P > 0 && !bp;
    e->bead_pairs = bp;
    e->bp_words = bp ? (P * P + 63) / 64 : 0;
    long long items = point_grid ? NP : N;
    e->T = pow2_at_least(2 * items > 64 ? 2 * items : 64);
    e->cap_cand = c.max_pairs > 0 ? c.max_pairs : (point_grid ? 1 : 16 * N + 1024);
    e->capC = c.max_contacts > 0 ? c.max_contacts : (P > 0 ? 400000 : 4 * e->cap_cand);
    e->capP = (P > 0 || (c.flags & ME_FLAG_PER_CONTACT)) ? e->capC : e->cap_cand;
    long long scan_n = items > e->cap_cand ? items : e->cap_cand;
    if (e->T > scan_n) scan_n = e->T;
    if (e->capC > scan_n) scan_n = e->capC;
    size_t C = (size_t)e->capC, Pc = (size_t)e->capP;
    size_t ns = P > 0 ? 1 : 4 * (size_t)e->cap_cand;
    Bump B{base};
    e->segs = StateSegs{};
    B.segs = &e->segs;
    // ---- state: what a substep leaves for the next one (status words, warm store, colour cache)
    // and what the adjoint of a recorded substep reads (contacts, pairs, candidates, tiles,
    // per-contact constants, world inertia, pseudo-velocities); a snapshot of these leading bytes
    // plus the body arrays restores the engine exactly ----
    e->istat = B.take<int>(IS_COUNT);
    e->invIw = B.take<float>(9 * (size_t)N);
    e->pv = B.take<float>(3 * (size_t)N);
    e->po = B.take<float>(3 * (size_t)N);
    e->cand_a = B.takec<int>((size_t)e->cap_cand, IS_CAND, 1);
    e->cand_b = B.takec<int>((size_t)e->cap_cand, IS_CAND, 1);
    e->S.n = B.takec<int>((size_t)e->cap_cand, IS_CAND, 1);
    e->slot_off = B.takec<int>((size_t)e->cap_cand, IS_CAND, 1);
    e->Cc.key = B.takec<int64_t>(C, IS_C, 1);
    e->Cc.bi = B.takec<int>(C, IS_C, 1);
    e->Cc.bj = B.takec<int>(C, IS_C, 1);
    e->Cc.pA = B.takec<float>(3 * C, IS_C, 3);
    e->Cc.pB = B.takec<float>(3 * C, IS_C, 3);
    e->Cc.n = B.takec<float>(3 * C, IS_C, 3);
    e->Cc.pen = B.takec<float>(C, IS_C, 1);
    e->Cc.mu = B.takec<float>(C, IS_C, 1);
    e->Cc.jn = B.takec<float>(C, IS_C, 1);
    e->Cc.jt1 = B.takec<float>(C, IS_C, 1);
    e->Cc.jt2 = B.takec<float>(C, IS_C, 1);
    e->Cc.jp = B.takec<float>(C, IS_C, 1);
    e->Pp.pbi = B.takec<int>(Pc, IS_P, 1);
    e->Pp.pbj = B.takec<int>(Pc, IS_P, 1);
    e->Pp.pstart = B.takec<int>(Pc, IS_P, 1);
    e->Pp.pcount = B.takec<int>(Pc, IS_P, 1);
    e->Pp.col = B.takec<int>(Pc, IS_P, 1);
    e->Pp.porder = B.takec<int>(Pc, IS_P, 1);
    e->cache_a = B.takec<int>(Pc, IS_CACHED_P, 1);
    e->cache_b = B.takec<int>(Pc, IS_CACHED_P, 1);
    e->cache_col = B.takec<int>(Pc, IS_CACHED_P, 1);
    e->cache_invM = B.take<float>((size_t)N);
    e->ccount = B.take<int>(ME_NCOLMAX + 1);
    e->cstart = B.take<int>(ME_NCOLMAX + 1);
    e->cfill = B.take<int>(ME_NCOLMAX + 1);
    e->csub = B.take<int>(ME_NCOLMAX + 1);
    e->cbase = B.take<int>(ME_NCOLMAX + 2);
    e->wkey = B.takec<int64_t>(C, IS_WN, 1);
    e->wjn = B.takec<float>(C, IS_WN, 1);
    e->wjt1 = B.takec<float>(C, IS_WN, 1);
    e->wjt2 = B.takec<float>(C, IS_WN, 1);
    tiled_layout_state(e->tb, B, N, (int)Pc, (int)C);
    e->state_bytes = (B.off + 255) & ~size_t(255);
    B.segs = nullptr;
    // ---- scratch: rewritten by every substep before it is read ----
    e->Rw = B.take<float>(9 * (size_t)N);
    e->aabb = B.take<float>(6 * (size_t)N);
    e->vsnap = B.take<float>(3 * (size_t)N);
    e->xsnap = B.take<float>(3 * (size_t)N);
    e->wsnap = B.take<float>(3 * (size_t)N);
    e->allp = B.take<float>(3 * (size_t)(P > 0 ? NP : 1));
    {
        bool jac = c.solver == ME_SOLVER_JACOBI;
        size_t jc = jac ? C : 1;
        e->jcnt = B.take<int>(jac ? (size_t)N : 1);
        e->joff = B.take<int>(jac ? (size_t)N : 1);
        e->jposA = B.take<int>(jc);
        e->jposB = B.take<int>(jc);
        e->jd = B.take<int64_t>(jac ? 12 * C : 1);
    }
    e->claim = B.take<unsigned long long>((size_t)N);
    e->cellc = B.take<int>(4 * (size_t)items);
    e->bucket = B.take<int>((size_t)items);
    e->bcount = B.take<int>((size_t)e->T);
    e->bstart = B.take<int>((size_t)e->T);
    e->bcursor = B.take<int>((size_t)e->T);
    e->items = B.take<int>((size_t)items);
    e->spos = B.take<float4>(point_grid ? (size_t)items : 1);
    e->scell = B.take<int4>(point_grid ? 1 : (size_t)N);
    e->sbox = B.take<float4>(point_grid ? 1 : 2 * (size_t)N);
    e->cand_n = B.take<int>((size_t)items);
    e->cand_off = B.take<int>((size_t)items);
    e->cand_tmp = B.take<int>(point_grid ? 1 : (size_t)N * ME_MAXCAND);
    e->S.key = B.take<int64_t>(ns);
    e->S.p = B.take<float>(3 * ns);
    e->S.nrm = B.take<float>(3 * ns);
    e->S.pen = B.take<float>(ns);
    e->S.mu = B.take<float>(ns);
    e->pair_flag = B.take<int>((size_t)e->cap_cand);
    e->pair_off = B.take<int>((size_t)e->cap_cand);
    {
        size_t R = bp ? C : 1, cc = bp ? (size_t)e->cap_cand : 1;
        e->bp_half = B.take<float>(bp ? 3 * (size_t)N : 1);
        e->bp_mask = B.take<uint64_t>(cc * (size_t)(bp ? e->bp_words : 1));
        e->bp_kn = B.take<int>(cc);
        e->bp_koff = B.take<int>(cc);
        e->raw.key = B.take<int64_t>(R);
        e->raw.pA = B.take<float>(3 * R);
        e->raw.pB = B.take<float>(3 * R);
        e->raw.n = B.take<float>(3 * R);
        e->raw.pen = B.take<float>(R);
        e->raw.jnd = B.take<double>(R);
        e->raw.jnr = B.take<double>(R);
        e->raw.jt1 = B.take<float>(R);
        e->raw.jt2 = B.take<float>(R);
        e->raw.keep = B.take<int>(R);
    }
    tiled_layout_scratch(e->tb, B, N, (int)Pc, (int)C, (c.flags & ME_FLAG_FAST_MATH) != 0);
    e->nbsum = (int)((scan_n + SCAN_BLOCK - 1) / SCAN_BLOCK) + 1;
    e->bsum = B.take<int>((size_t)e->nbsum);
    // the one-pass scans' ticket counters and block words: zeroed with the workspace at creation and
    // never cleared again (each launch's words carry its own launch number)
    e->scan_words = B.take<unsigned long long>((size_t)SCAN_SITES * (1 + 2 * (size_t)e->nbsum));
    *bytes = (B.off + 255) & ~size_t(255);
}

int validate(const me_config *c) {
    if (!c) return fail(ME_EARG, "me_config is NULL");
    if (c->n_bodies <= 0) return fail(ME_EARG, "n_bodies must be positive");
    if (c->solver < 0 || c->solver > 2) return fail(ME_EARG, "solver must be 0 (jacobi), 1 (pair chunks) or 2 (color cache)");
    if (c->vit < 0 || c->pit < 0) return fail(ME_EARG, "iteration counts must be >= 0");
    if (c->n_beads < 0) return fail(ME_EARG, "n_beads must be >= 0");
    if (c->solver == ME_SOLVER_COLOR_CACHE && (c->flags & ME_FLAG_PER_CONTACT))
        return fail(ME_EARG, "ColorCache solves body-pair chunks; per-contact colouring is the pair-chunk solver's "
                             "pair_chunks=False mode");
    if (c->n_beads > 64) return fail(ME_EARG, "n_beads must be <= 64");
    if (c->n_colors < 0 || c->n_colors > ME_NCOLMAX) return fail(ME_EARG, "n_colors must be in [0, 512]");
    if ((long long)c->n_bodies * (c->n_beads > 0 ? c->n_beads : 1) > (1LL << 30))
        return fail(ME_EARG, "too many bodies/beads");
    if (!(c->margin >= 0.0f)) return fail(ME_EARG, "margin must be >= 0");
    if ((c->flags & ME_FLAG_GRAPH) && (!(c->flags & ME_FLAG_TILED) || c->solver == ME_SOLVER_JACOBI))
        return fail(ME_EARG, "graph capture needs the tiled solve and a coloured solver");
    if ((c->flags & ME_FLAG_FAST_MATH) && (!(c->flags & ME_FLAG_TILED) || c->solver != ME_SOLVER_COLOR_CACHE))
        return fail(ME_EARG, "fast arithmetic needs the tiled ColorCache solver");
    return ME_OK;
}

// call sites of the one-pass scan, each with its own ticket counter and block words (k_scan1)
enum { SC_GRID = 0, SC_CAND, SC_BP_SLOT, SC_BP_KEEP, SC_BP_PAIR, SC_BEAD, SC_NARROW, SC_JAC, SC_COLOR };
static_assert(SC_COLOR < SC_TILED_B, "scan sites");

unsigned long long *scan_site(me_engine *e, int site) { return e->scan_words + (size_t)site * (1 + 2 * (size_t)e->nbsum); }

// exclusive scan in -> out over n elements (n_dev device count or n_host), total -> *total_dev; one
// launch (k_scan1). scan2 also scans the flags in[i] > 0 into out2 (total2) and, with capC > 0, leaves
// the clamped contact and pair counts of a compaction in the status words.
int scan2(me_engine *e, int site, const int *in, int *out, int *out2, const int *n_dev, int n_host, int cap,
          int *total_dev, int *total2, int capC, int capP, hipStream_t s) {
    int nb = (cap + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb < 1) nb = 1;
    if (nb > e->nbsum) return fail(ME_ERUNTIME, "scan scratch too small");
    if (out2)
        hipLaunchKernelGGL(k_scan1<2>, dim3(nb), dim3(ME_THREADS), 0, s, in, out, out2, n_dev, n_host, cap, total_dev,
                           total2, scan_site(e, site), e->istat, e->giant_wait_ticks, capC, capP);
    else
        hipLaunchKernelGGL(k_scan1<1>, dim3(nb), dim3(ME_THREADS), 0, s, in, out, nullptr, n_dev, n_host, cap,
                           total_dev, nullptr, scan_site(e, site), e->istat, e->giant_wait_ticks, 0, 0);
    LAUNCH_CHECK();
    return ME_OK;
}

int scan(me_engine *e, int site, const int *in, int *out, const int *n_dev, int n_host, int cap, int *total_dev,
         hipStream_t s) {
    return scan2(e, site, in, out, nullptr, n_dev, n_host, cap, total_dev, nullptr, 0, 0, s);
}

Bodies bodies_of(me_engine *e) {
    Bodies b;
    b.x = e->b.x; b.q = e->b.q; b.v = e->b.v; b.w = e->b.w;
    b.invM = e->b.inv_mass; b.IbInv = e->b.inv_inertia; b.half = e->bead_pairs ? e->bp_half : e->b.half;
    b.mu = e->b.mu;
    b.kin = e->b.kinematic; b.scene = e->b.scene; b.rest = e->b.rest;
    return b;
}

// giant_possible / gcol_possible from the bodies per scene (read once from the device): an island
// of a scene of n bodies has at most n bodies and n(n-1)/2 + n pairs. Scene ids outside [0, N) keep
// both set.
int scene_bounds(me_engine *e) {
    const int N = e->N;
    long long mx = N;
    if (e->b.scene) {
        std::vector<int> sc((size_t)N), cnt((size_t)N, 0);
        HIP_TRY(hipMemcpy(sc.data(), e->b.scene, sizeof(int) * (size_t)N, hipMemcpyDeviceToHost));
        bool in_range = true;
        for (int i = 0; i < N && in_range; ++i) {
            if (sc[i] < 0 || sc[i] >= N) in_range = false;
            else cnt[sc[i]]++;
        }
        if (in_range) {
            mx = 0;
            for (int c : cnt) mx = c > mx ? c : mx;
        }
    }
    const long long pairs = mx * (mx - 1) / 2 + mx;
    e->giant_possible = (mx > T_BIG_B || pairs > T_BIG_P) ? 1 : 0;
    e->gcol_possible = (mx > T_COLB || pairs > T_COLP) ? 1 : 0;
    return ME_OK;
}

int read_stat(me_engine *e, hipStream_t s) {
    HIP_TRY(hipMemcpyAsync(e->hstat, e->istat, sizeof(int) * IS_COUNT, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ME_OK;
}

// grid over n items with positions pos; cell from IS_CELL (cell_fixed <= 0) or fixed
int build_grid(me_engine *e, int n, const float *pos, const int *scene, int div, float cell_fixed, float4 *spos,
               bool body_grid, hipStream_t s) {
    hipLaunchKernelGGL(k_grid_count, dim3(nblk(n)), dim3(ME_THREADS), 0, s, n, pos, scene, div, e->istat, cell_fixed,
                       e->cfg.margin, e->T, e->cellc, e->bucket, e->bcount);
    int rc = scan(e, SC_GRID, e->bcount, e->bstart, nullptr, e->T, e->T, nullptr, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_grid_scatter, dim3(nblk(n)), dim3(ME_THREADS), 0, s, n, e->bucket, e->bstart, e->bcursor,
                       e->items, pos, spos, e->cellc, e->aabb, e->b.inv_mass, body_grid ? e->scell : nullptr,
                       body_grid ? e->sbox : nullptr);
    LAUNCH_CHECK();
    return ME_OK;
}

// body pairs (A6): grid over body centres, AABB overlaps in pair-key order, ground pair first;
// all_pairs keeps pairs of static bodies (bead geometry)
int body_pairs(me_engine *e, int all_pairs, hipStream_t s) {
    Bodies b = bodies_of(e);
    int N = e->N;
    int rc = build_grid(e, N, e->b.x, e->b.scene, 1, -1.0f, nullptr, true, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_box_pairs, dim3((N + ME_BP_PER_WG - 1) / ME_BP_PER_WG), dim3(32 * ME_BP_PER_WG), 0, s, N, b,
                       e->aabb, e->cellc, e->bstart, e->bcount, e->scell, e->sbox,
                       e->T, e->cfg.margin, all_pairs, e->cand_n, e->cand_tmp, e->istat);
    rc = scan(e, SC_CAND, e->cand_n, e->cand_off, nullptr, N, N, e->istat + IS_CAND, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_box_cand_emit, dim3(nblk(N)), dim3(ME_THREADS), 0, s, N, e->cand_n, e->cand_off,
                       e->cand_tmp, e->cap_cand, e->cand_a, e->cand_b, e->istat);
    LAUNCH_CHECK();
    return ME_OK;
}

#define ME_BEAD_KEEP 4  // manifold_max_points of the reference's coloured engines

// bead geometry under a coloured solver: A4 world points, A6 body pairs over the boxes of the bead
// spheres, A7 per-pair bead tests (pair-major), A14 warm start and A15 manifold reduction per pair,
// compaction into contiguous pairs of at most ME_BEAD_KEEP contacts. Leaves the contacts with their
// warm-start impulses set.
int generate_bead_pairs(me_engine *e, hipStream_t s) {
    Bodies b = bodies_of(e);
    int N = e->N, P = e->P, NP = e->NP, cc = e->cap_cand, W = e->bp_words;
    float R = e->cfg.bead_radius;
    hipLaunchKernelGGL(k_bead_points, dim3(nblk(NP)), dim3(ME_THREADS), 0, s, N, P, b, e->allp);
    int rc = body_pairs(e, 1, s);
    if (rc) return rc;
    int wg = nblk(64LL * cc);
    hipLaunchKernelGGL(k_bp_count, dim3(wg), dim3(ME_THREADS), 0, s, cc, e->istat, e->cand_a, e->cand_b, e->allp, P,
                       R, W, e->bp_mask, e->S.n);
    rc = scan(e, SC_BP_SLOT, e->S.n, e->slot_off, e->istat + IS_CAND, 0, cc, e->istat + IS_CRAW, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bp_emit, dim3(wg), dim3(ME_THREADS), 0, s, cc, e->istat, e->cand_a, e->cand_b, e->allp, N, P,
                       R, W, e->bp_mask, e->slot_off, e->capC, e->raw, e->istat);
    int sel_blocks = nblk((long long)cc * BP_G);
    if (sel_blocks > 2048) sel_blocks = 2048;
    hipLaunchKernelGGL(k_bp_select, dim3(sel_blocks), dim3(ME_THREADS), 0, s, cc, e->istat, e->cand_a, e->b.x, e->S.n,
                       e->slot_off, e->capC, ME_BEAD_KEEP, e->cfg.warm_start, e->wkey, e->wjn, e->wjt1, e->wjt2,
                       e->raw, e->bp_kn, e->pair_flag);
    rc = scan(e, SC_BP_KEEP, e->bp_kn, e->bp_koff, e->istat + IS_CAND, 0, cc, e->istat + IS_C, s);
    if (rc) return rc;
    rc = scan(e, SC_BP_PAIR, e->pair_flag, e->pair_off, e->istat + IS_CAND, 0, cc, e->istat + IS_P, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bp_compact, dim3(nblk(cc)), dim3(ME_THREADS), 0, s, cc, e->istat, e->cand_a, e->cand_b,
                       e->b.mu, e->S.n, e->slot_off, e->bp_kn, e->bp_koff, e->pair_flag, e->pair_off, e->capC,
                       e->capP, e->raw, e->Cc, e->Pp, e->istat);
    hipLaunchKernelGGL(k_clamp_counts, dim3(1), dim3(1), 0, s, e->istat, e->capC, e->capP, 0);
    LAUNCH_CHECK();
    return ME_OK;
}

// contact generation: box geometry (A6 body pairs + narrow phase) or bead geometry (A4 + A5, or
// A4 + A6 + A7 + A15 under a coloured solver)
int generate(me_engine *e, hipStream_t s) {
    Bodies b = bodies_of(e);
    int N = e->N;
    if (e->bead_pairs) return generate_bead_pairs(e, s);
    if (e->P > 0) {
        int NP = e->NP, P = e->P;
        float R = e->cfg.bead_radius;
        hipLaunchKernelGGL(k_bead_points, dim3(nblk(NP)), dim3(ME_THREADS), 0, s, N, P, b, e->allp);
        int rc = build_grid(e, NP, e->allp, nullptr, 1, 2.0f * R, e->spos, false, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bead_contacts<false>, dim3(nblk(NP)), dim3(ME_THREADS), 0, s, NP, P, e->allp, e->cellc,
                           e->bstart, e->bcount, e->spos, e->T, R, e->b.mu, e->cand_n, e->cand_off, e->capC, e->Cc,
                           e->istat);
        rc = scan(e, SC_BEAD, e->cand_n, e->cand_off, nullptr, NP, NP, e->istat + IS_C, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bead_contacts<true>, dim3(nblk(NP)), dim3(ME_THREADS), 0, s, NP, P, e->allp, e->cellc,
                           e->bstart, e->bcount, e->spos, e->T, R, e->b.mu, e->cand_n, e->cand_off, e->capC, e->Cc,
                           e->istat);
        hipLaunchKernelGGL(k_clamp_counts, dim3(1), dim3(1), 0, s, e->istat, e->capC, e->capP, 1);
        // per-contact colouring (GraphColoredContactEngine pair_chunks=False): every raw bead contact its
        // own chunk, in the feature-key order the emit pass wrote
        if (e->cfg.solver != ME_SOLVER_JACOBI)
            hipLaunchKernelGGL(k_pairs_per_contact, dim3(nblk(e->capC)), dim3(ME_THREADS), 0, s, e->capC, e->istat,
                               e->Cc.bi, e->Cc.bj, e->Pp);
        LAUNCH_CHECK();
        return ME_OK;
    }
    int rc = body_pairs(e, 0, s);
    if (rc) return rc;
    int cc = e->cap_cand;
    prof_mark(e, ME_K_NARROW, s);
    if (N <= ME_NARROW_LDS_MAXN)
        hipLaunchKernelGGL(k_narrow<true>, dim3((cc + ME_NARROW_WG - 1) / ME_NARROW_WG), dim3(ME_NARROW_WG), 0, s, cc,
                           e->istat, b, e->Rw, e->cand_a, e->cand_b, e->cfg.margin, N, e->S);
    else
        hipLaunchKernelGGL(k_narrow<false>, dim3(nblk(cc)), dim3(ME_THREADS), 0, s, cc, e->istat, b, e->Rw, e->cand_a,
                           e->cand_b, e->cfg.margin, N, e->S);
    prof_mark(e, ME_K_NARROW, s);
    // contact slots and non-empty pairs in one scan, which leaves the clamped counts (k_clamp_counts)
    rc = scan2(e, SC_NARROW, e->S.n, e->slot_off, e->pair_off, e->istat + IS_CAND, 0, cc, e->istat + IS_C,
               e->istat + IS_P, e->capC, e->capP, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_compact, dim3(nblk(cc)), dim3(ME_THREADS), 0, s, cc, e->capC, e->capP, e->istat, e->cand_a,
                       e->cand_b, e->S, e->slot_off, e->pair_off, e->Cc, e->Pp);
    if (e->cfg.flags & ME_FLAG_PER_CONTACT)
        hipLaunchKernelGGL(k_pairs_per_contact, dim3(nblk(e->capC)), dim3(ME_THREADS), 0, s, e->capC, e->istat,
                           e->Cc.bi, e->Cc.bj, e->Pp);
    LAUNCH_CHECK();
    return ME_OK;
}

// A12 relaxed Jacobi substep body (contact_engine_gpu.py:266-275)
int solve_jacobi(me_engine *e, float sdt, hipStream_t s) {
    int rc = read_stat(e, s);
    if (rc) return rc;
    int C = e->hstat[IS_C];
    if (C <= 0) return ME_OK;
    int N = e->N, capC = e->capC;
    int gridN = nblk(N > C ? N : C);
    hipLaunchKernelGGL(k_jac_reset, dim3(gridN), dim3(ME_THREADS), 0, s, N, capC, e->istat, e->pv, e->po, e->jcnt,
                       e->Cc.jn, e->Cc.jt1, e->Cc.jt2, e->Cc.jp);
    // the body-major entry list of this substep's contacts
    hipLaunchKernelGGL(k_jac_slots, dim3(nblk(C)), dim3(ME_THREADS), 0, s, capC, e->istat, e->Cc.bi, e->Cc.bj, e->jcnt,
                       e->jposA, e->jposB);
    rc = scan(e, SC_JAC, e->jcnt, e->joff, nullptr, N, N, e->istat + IS_JTOT, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_jac_place, dim3(nblk(C)), dim3(ME_THREADS), 0, s, capC, e->istat, e->Cc.bi, e->Cc.bj, e->joff,
                       e->jposA, e->jposB);
    JacArgs A;
    A.bi = e->Cc.bi; A.bj = e->Cc.bj; A.pA = e->Cc.pA; A.pB = e->Cc.pB; A.n = e->Cc.n; A.pen = e->Cc.pen;
    A.mu = e->Cc.mu; A.jn = e->Cc.jn; A.jt1 = e->Cc.jt1; A.jt2 = e->Cc.jt2; A.jp = e->Cc.jp;
    A.x = e->b.x; A.invM = e->b.inv_mass; A.invIw = e->invIw; A.v = e->b.v; A.w = e->b.w; A.pv = e->pv; A.po = e->po;
    A.posA = e->jposA; A.posB = e->jposB; A.d = e->jd; A.N = N;
    const float S = 1.0e10f;
    const int *tot = e->istat + IS_JTOT;
    for (int it = 0; it < e->cfg.vit; ++it) {
        prof_mark(e, ME_K_JAC_VEL, s);
        hipLaunchKernelGGL(k_jac_vel, dim3(nblk(C)), dim3(ME_THREADS), 0, s, e->istat, A, S);
        prof_mark(e, ME_K_JAC_VEL, s);
        hipLaunchKernelGGL(k_jac_gather, dim3(nblk(6 * N)), dim3(ME_THREADS), 0, s, N, e->joff, tot, e->jd, e->b.v,
                           e->b.w, e->cfg.relax, S);
    }
    for (int it = 0; it < e->cfg.pit; ++it) {
        hipLaunchKernelGGL(k_jac_pos, dim3(nblk(C)), dim3(ME_THREADS), 0, s, e->istat, A, sdt, S);
        hipLaunchKernelGGL(k_jac_gather, dim3(nblk(6 * N)), dim3(ME_THREADS), 0, s, N, e->joff, tot, e->jd, e->pv,
                           e->po, e->cfg.relax, S);
    }
    LAUNCH_CHECK();
    return ME_OK;
}

SweepArgs sweep_args(me_engine *e) {
    SweepArgs S;
    S.B.v = e->b.v; S.B.w = e->b.w; S.B.pv = e->pv; S.B.po = e->po; S.B.x = e->b.x; S.B.invM = e->b.inv_mass;
    S.B.invIw = e->invIw;
    S.Cv.pA = e->Cc.pA; S.Cv.pB = e->Cc.pB; S.Cv.n = e->Cc.n; S.Cv.pen = e->Cc.pen; S.Cv.mu = e->Cc.mu;
    S.Cv.jn = e->Cc.jn; S.Cv.jt1 = e->Cc.jt1; S.Cv.jt2 = e->Cc.jt2; S.Cv.jp = e->Cc.jp;
    S.pbi = e->Pp.pbi; S.pbj = e->Pp.pbj; S.pstart = e->Pp.pstart; S.pcount = e->Pp.pcount; S.porder = e->Pp.porder;
    S.cstart = e->cstart; S.ccount = e->ccount;
    return S;
}

// JP colouring rounds with one read-back per round (contact_engine_gpu_colored.py:1178-1208)
int color_jp(me_engine *e, int P, hipStream_t s) {
    int capP = e->capP;
    hipLaunchKernelGGL(k_jp_init, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, e->Pp.col);
    int cur = 0;
    for (;;) {
        hipLaunchKernelGGL(k_fill_i64, dim3(nblk(e->N)), dim3(ME_THREADS), 0, s, (int64_t *)e->claim, e->N, INT64_MAX);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, e->istat + IS_LEFT, 1, 0);
        hipLaunchKernelGGL(k_jp_claim, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, e->Pp.col, e->Pp.pbi,
                           e->Pp.pbj, e->b.inv_mass, e->claim);
        hipLaunchKernelGGL(k_jp_win, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, cur, e->Pp.col, e->Pp.pbi,
                           e->Pp.pbj, e->b.inv_mass, e->claim);
        LAUNCH_CHECK();
        cur++;
        int rc = read_stat(e, s);
        if (rc) return rc;
        if (e->hstat[IS_LEFT] == 0) break;
        if (cur > P + 1) return fail(ME_ERUNTIME, "colouring did not terminate");
    }
    if (cur > ME_NCOLMAX) return fail(ME_ECAPACITY, "colour count exceeds the per-colour capacity 512");
    // class sizes, optional split into n_colors classes, final sizes
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(ME_THREADS), 0, s, e->istat + IS_NCOL, 1, cur);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ME_NCOLMAX + 1)), dim3(ME_THREADS), 0, s, e->ccount, ME_NCOLMAX + 1, 0);
    hipLaunchKernelGGL(k_color_count, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, e->Pp.col, e->ccount);
    if (e->cfg.n_colors > cur) {
        hipLaunchKernelGGL(k_color_split_plan, dim3(1), dim3(1), 0, s, e->istat, e->cfg.n_colors, e->ccount, e->csub,
                           e->cbase, ME_NCOLMAX);
        hipLaunchKernelGGL(k_color_split_apply, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, e->Pp.col,
                           e->csub, e->cbase);
    }
    LAUNCH_CHECK();
    return ME_OK;
}

// A17-A21 coloured pair-chunk solve: warm lookup, (cached) colouring, grouping, sweeps, warm store
int solve_colored(me_engine *e, float sdt, hipStream_t s) {
    bool cc_mode = e->cfg.solver == ME_SOLVER_COLOR_CACHE;
    int capC = e->capC, capP = e->capP, N = e->N;
    int rc = read_stat(e, s);  // pv, po were zeroed by k_body_pre
    if (rc) return rc;
    int C = e->hstat[IS_C], P = e->hstat[IS_P];
    if (C <= 0) {  // no contacts: warm store emptied (contact_engine_gpu_deferred_force.py:227-235)
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, e->istat + IS_WN, 1, 0);
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, e->istat + IS_NCOL, 1, 0);
        LAUNCH_CHECK();
        return ME_OK;
    }
    if (!e->bead_pairs)  // bead contacts carry their warm start from the manifold stage
        hipLaunchKernelGGL(k_warm_lookup, dim3(nblk(C)), dim3(ME_THREADS), 0, s, capC, e->istat, e->cfg.warm_start,
                           e->Cc.key, e->wkey, e->wjn, e->wjt1, e->wjt2, e->Cc.jn, e->Cc.jt1, e->Cc.jt2, e->Cc.jp);
    bool hit = false;
    if (cc_mode) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, e->istat + IS_MISMATCH, 1, 0);
        int g = P > N ? P : N;
        hipLaunchKernelGGL(k_cc_compare, dim3(nblk(g)), dim3(ME_THREADS), 0, s, N, e->istat, e->Pp.pbi, e->Pp.pbj,
                           e->cache_a, e->cache_b, e->b.inv_mass, e->cache_invM);
        LAUNCH_CHECK();
        rc = read_stat(e, s);
        if (rc) return rc;
        hit = e->hstat[IS_MISMATCH] == 0;
    }
    if (hit) {
        hipLaunchKernelGGL(k_cc_restore, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, e->cache_col,
                           e->Pp.col);
    } else {
        rc = color_jp(e, P, s);
        if (rc) return rc;
    }
    if (cc_mode) {
        int g = P > N ? P : N;
        hipLaunchKernelGGL(k_cc_store, dim3(nblk(g)), dim3(ME_THREADS), 0, s, N, capP, e->istat, e->Pp.pbi, e->Pp.pbj,
                           e->Pp.col, e->b.inv_mass, e->cache_a, e->cache_b, e->cache_col, e->cache_invM, hit ? 1 : 0);
    }
    rc = read_stat(e, s);
    if (rc) return rc;
    int ncol = e->hstat[IS_NCOL];
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ME_NCOLMAX + 1)), dim3(ME_THREADS), 0, s, e->ccount, ME_NCOLMAX + 1, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ME_NCOLMAX + 1)), dim3(ME_THREADS), 0, s, e->cfill, ME_NCOLMAX + 1, 0);
    hipLaunchKernelGGL(k_color_count, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, e->Pp.col, e->ccount);
    rc = scan(e, SC_COLOR, e->ccount, e->cstart, nullptr, ncol, ME_NCOLMAX + 1, nullptr, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_color_scatter, dim3(nblk(P)), dim3(ME_THREADS), 0, s, capP, e->istat, e->Pp.col, e->cstart,
                       e->cfill, e->Pp.porder);
    SweepArgs S = sweep_args(e);
    dim3 g(nblk(P)), t(ME_THREADS);
    if (cc_mode)
        for (int ci = 0; ci < ncol; ++ci) hipLaunchKernelGGL(k_sweep_warm, g, t, 0, s, S, ci);
    for (int it = 0; it < e->cfg.vit; ++it)
        for (int ci = 0; ci < ncol; ++ci) {
            prof_mark(e, ME_K_SWEEP_VEL, s);
            if (cc_mode) hipLaunchKernelGGL(k_sweep_vel<true>, g, t, 0, s, S, ci, 1.0f);
            else hipLaunchKernelGGL(k_sweep_vel<false>, g, t, 0, s, S, ci, e->cfg.t2d);
            prof_mark(e, ME_K_SWEEP_VEL, s);
        }
    for (int it = 0; it < e->cfg.pit; ++it)
        for (int ci = 0; ci < ncol; ++ci) {
            prof_mark(e, ME_K_SWEEP_POS, s);
            hipLaunchKernelGGL(k_sweep_pos, g, t, 0, s, S, ci, sdt);
            prof_mark(e, ME_K_SWEEP_POS, s);
        }
    if (e->cfg.warm_start)
        hipLaunchKernelGGL(k_warm_store, dim3(nblk(C)), dim3(ME_THREADS), 0, s, capC, e->istat, e->Cc.key, e->Cc.jn,
                           e->Cc.jt1, e->Cc.jt2, e->wkey, e->wjn, e->wjt1, e->wjt2);
    LAUNCH_CHECK();
    return ME_OK;
}

int substep(me_engine *e, double sdt64, hipStream_t s) {
    float sdt = (float)sdt64;
    int N = e->N;
    Bodies b = bodies_of(e);
    if (e->bead_pairs)
        hipLaunchKernelGGL(k_fill_half, dim3(nblk(N)), dim3(ME_THREADS), 0, s, N, e->bhalf[0], e->bhalf[1], e->bhalf[2],
                           e->bp_half);
    bool colored = e->cfg.solver != ME_SOLVER_JACOBI;
    hipLaunchKernelGGL(k_body_pre, dim3(nblk(N)), dim3(ME_THREADS), 0, s, N, b, e->invIw, e->Rw, e->aabb, e->vsnap,
                       e->xsnap, e->wsnap, e->istat, sdt, (e->P == 0 || e->bead_pairs) ? 1 : 0, e->T, e->bcount,
                       e->bcursor, colored ? e->pv : nullptr, colored ? e->po : nullptr);
    LAUNCH_CHECK();
    int rc = generate(e, s);
    if (rc) return rc;
    if (e->cfg.solver == ME_SOLVER_JACOBI) {
        rc = solve_jacobi(e, sdt, s);
    } else if (e->cfg.flags & ME_FLAG_TILED) {
        TiledArgs ta = tiled_args_of(e, sdt);
        rc = tiled_prepare(e->tb, ta, s);
        if (rc) return fail(rc, "tiled solve: launch failed");
        prof_mark(e, ME_K_TILED, s);
        rc = tiled_sweeps(e->tb, ta, s);
        prof_mark(e, ME_K_TILED, s);
        if (rc) return fail(rc, "tiled solve: launch failed");
        rc = tiled_finish(e->tb, ta, s);
    } else {
        rc = solve_colored(e, sdt, s);
    }
    if (rc) return rc;
    if (!(e->cfg.flags & ME_FLAG_TILED) || e->cfg.solver == ME_SOLVER_JACOBI)
        hipLaunchKernelGGL(k_integrate, dim3(nblk(N)), dim3(ME_THREADS), 0, s, N, b, e->pv, e->po, sdt);
    LAUNCH_CHECK();
    return ME_OK;
}

const size_t kGraphMax = 8;

void drop_graphs(me_engine *e) {
    for (auto &g : e->graphs) (void)hipGraphExecDestroy(g.exec);
    e->graphs.clear();
}

// everything an engine holds besides the caller's memory
void release(me_engine *e) {
    drop_graphs(e);
    for (int k = 0; k < ME_K_COUNT_; ++k)
        for (hipEvent_t ev : e->prof_ev[k]) (void)hipEventDestroy(ev);
    if (e->hstat) (void)hipHostFree(e->hstat);
    for (hipEvent_t ev : {e->ev_fork, e->ev_join, e->ev_in, e->ev_out})
        if (ev) (void)hipEventDestroy(ev);
    if (e->side) (void)hipStreamDestroy(e->side);
    if (e->gs) (void)hipStreamDestroy(e->gs);
    delete e;
}

// one step as a graph launch: capture every substep on gs the first time a (dt, substeps) pair is
// seen, then replay; gs is ordered after the caller's stream s and s after the replay
int step_graph(me_engine *e, double dt, int substeps, hipStream_t s) {
    hipGraphExec_t exec = nullptr;
    for (auto &g : e->graphs)
        if (g.dt == dt && g.substeps == substeps) exec = g.exec;
    if (!exec) {
        if (e->graphs.size() >= kGraphMax) {
            (void)hipGraphExecDestroy(e->graphs.front().exec);
            e->graphs.erase(e->graphs.begin());
        }
        HIP_TRY(hipStreamBeginCapture(e->gs, hipStreamCaptureModeThreadLocal));
        double sdt = dt / substeps;
        int rc = ME_OK;
        for (int k = 0; k < substeps && rc == ME_OK; ++k) rc = substep(e, sdt, e->gs);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->gs, &g);
        if (rc != ME_OK || ce != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return rc != ME_OK ? rc : fail(ME_ERUNTIME, "step capture failed");
        }
        hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return fail(ME_ERUNTIME, "graph instantiation failed");
        e->graphs.push_back({dt, substeps, exec});
        e->n_captures++;
    }
    HIP_TRY(hipEventRecord(e->ev_in, s));
    HIP_TRY(hipStreamWaitEvent(e->gs, e->ev_in, 0));
    HIP_TRY(hipGraphLaunch(exec, e->gs));
    HIP_TRY(hipEventRecord(e->ev_out, e->gs));
    HIP_TRY(hipStreamWaitEvent(s, e->ev_out, 0));
    return ME_OK;
}

// layout of a tape block (see AdjTape), 256-byte aligned arrays
size_t tape_layout(const me_engine *e, char *base, AdjTape *t) {
    size_t N = (size_t)e->N, C = (size_t)e->capC, S = (size_t)(e->cfg.vit + e->cfg.pit);
    Bump B{base};
    AdjTape x{};
    x.xpre = B.take<float>(3 * N);
    x.qpre = B.take<float>(4 * N);
    x.widx = B.take<int>(C);
    x.sst = B.take<float>(S * 6 * N);
    x.simp = B.take<float>(S * 4 * C);
    x.cbar = B.take<float>(20 * C);
    x.jbi = B.take<float>(3 * C);
    x.gbar = B.take<float>(7 * C);
    x.pvb = B.take<float>(3 * N);
    x.pob = B.take<float>(3 * N);
    x.acc = B.take<long long>(ADJ_ACC * N);
    if (t) *t = x;
    return B.off;
}

int tape_check(me_engine *e, void *tape, size_t bytes) {
    if (!e) return fail(ME_EARG, "engine is NULL");
    if (std::this_thread::get_id() != e->owner) return fail(ME_ETHREAD, "engine used from a thread that does not own it");
    const me_config &c = e->cfg;
    if (c.solver != ME_SOLVER_COLOR_CACHE || !(c.flags & ME_FLAG_TILED) || c.n_beads > 0)
        return fail(ME_EARG, "the adjoint covers the tiled ColorCache solver on box bodies");
    if (c.flags & ME_FLAG_FAST_MATH)
        return fail(ME_EARG, "the adjoint differentiates the exact arithmetic: an engine without ME_FLAG_FAST_MATH");
    if (!tape || ((uintptr_t)tape & 255) != 0) return fail(ME_EARG, "tape must be a 256-byte aligned device block");
    size_t need = tape_layout(e, nullptr, nullptr);
    if (bytes < need) return fail(ME_EARG, "tape too small: need " + std::to_string(need) + " bytes");
    tape_layout(e, (char *)tape, &e->tape);
    return ME_OK;
}

}  // namespace

// tiled-solve argument bundle (declared in tiled.hpp)
TiledArgs tiled_args_of(me_engine *e, float sdt) {
    TiledArgs a;
    a.N = e->N; a.capC = e->capC; a.capP = e->capP;
    a.solver = e->cfg.solver; a.vit = e->cfg.vit; a.pit = e->cfg.pit; a.t2d = e->cfg.t2d;
    a.warm = e->cfg.warm_start; a.n_colors = e->cfg.n_colors; a.tile_bodies = e->cfg.tile_bodies;
    a.warm_set = e->bead_pairs ? 1 : 0;
    a.dt = sdt;
    a.b = bodies_of(e);
    a.invIw = e->invIw; a.pv = e->pv; a.po = e->po;
    a.Cc = e->Cc; a.Pp = e->Pp;
    a.cache_a = e->cache_a; a.cache_b = e->cache_b; a.cache_col = e->cache_col; a.cache_invM = e->cache_invM;
    a.wkey = e->wkey; a.wjn = e->wjn; a.wjt1 = e->wjt1; a.wjt2 = e->wjt2;
    a.istat = e->istat;
    a.bsum = e->bsum; a.nbsum = e->nbsum;
    a.scan_words = e->scan_words;
    a.claim = e->claim;
    a.side = e->side; a.ev_fork = e->ev_fork; a.ev_join = e->ev_join;
    a.prof_ev = e->prof_on ? prof_event : nullptr;
    a.prof = e;
    a.giant_grid = e->giant_grid;
    a.big_grid = e->big_grid;
    a.fast = (e->cfg.flags & ME_FLAG_FAST_MATH) != 0;
    a.giant_possible = e->giant_possible;
    a.gcol_possible = e->gcol_possible;
    a.small_grid = a.fast ? e->small_grid_fast : e->small_grid;
    a.giant_wait_ticks = e->giant_wait_ticks;
    a.taped = e->taping ? 1 : 0;
    a.tape = e->tape;
    return a;
}

// bead geometry under a coloured solver: half extents of the box around the bead spheres (body
// frame), the extent the body-pair broad phase bounds each body by
int bead_half_extents(me_engine *e, hipStream_t s) {
    if (!e->bead_pairs) return ME_OK;
    std::vector<float> rest(3 * (size_t)e->P);
    HIP_TRY(hipMemcpyAsync(rest.data(), e->b.rest, sizeof(float) * rest.size(), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float R = e->cfg.bead_radius;
    for (int k = 0; k < 3; ++k) {
        float m = 0.0f;
        for (int i = 0; i < e->P; ++i) m = fmaxf(m, fabsf(rest[3 * i + k]));
        if (!(m == m) || !(R >= 0.0f)) return fail(ME_ENONFINITE, "bead rest positions or radius not finite");
        e->bhalf[k] = m + R;
    }
    return ME_OK;
}

// the giant-island kernels' grid: the preferred 64 workgroups lowered to what can be resident at
// once on this device (their workgroups wait for each other at grid barriers), and the barrier's
// wait bound, 100 ms of the constant-rate wall clock
int giant_setup(me_engine *e) {
    int dev = 0, ncu = 0, khz = 0, occ[3] = {0, 0, 0};
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) {
        (void)hipGetLastError();
        khz = 100000;  // gfx9 constant-rate counter: 100 MHz
    }
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[0], (const void *)k_t_giant<ME_SOLVER_PAIR_CHUNKS>,
                                                         T_BIG_WG, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[1], (const void *)k_t_giant<ME_SOLVER_COLOR_CACHE>,
                                                         T_BIG_WG, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[2], (const void *)k_t_color_giant, T_BIG_WG, 0));
    int sweep = occ[0] < occ[1] ? occ[0] : occ[1];
    {
        int of = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&of, (const void *)k_t_giant<ME_SOLVER_COLOR_CACHE, true>,
                                                             T_BIG_WG, 0));
        if (of < sweep) sweep = of;
    }
    e->giant_grid = giant_grid_size(T_GIANT_GRID, sweep, occ[2], ncu);
    // the big-tile kernels take tiles from a queue: as many workgroups as can be resident, at most
    // T_BIG_GRID_MAX (more would only find the queue empty)
    int ob[3] = {0, 0, 0};
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob[0], (const void *)k_t_big<ME_SOLVER_PAIR_CHUNKS>,
                                                         T_BIGT_WG, T_BIG_DYN));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob[1], (const void *)k_t_big<ME_SOLVER_COLOR_CACHE>,
                                                         T_BIGT_WG, T_BIG_DYN));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&ob[2], (const void *)k_t_color<1024>, 1024, T_COL_DYN));
    int obig = ob[0] < ob[1] ? ob[0] : ob[1];
    {
        int of = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&of, (const void *)k_t_big<ME_SOLVER_COLOR_CACHE, true>,
                                                             T_BIGT_WG, T_BIG_DYN));
        if (of < obig) obig = of;
    }
    if (ob[2] < obig) obig = ob[2];
    e->big_grid = giant_grid_size(T_BIG_GRID_MAX, obig, obig, ncu);
    if (e->big_grid < 1) e->big_grid = 1;  // a workgroup that fits nowhere fails at launch with an error
    // the small-tile kernel takes tiles from a queue too: every workgroup that can be resident
    int os[4] = {0, 0, 0, 0};
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&os[0], (const void *)k_t_small<false, false>, T_SMALL_P, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&os[1], (const void *)k_t_small<true, false>, T_SMALL_P, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&os[2], (const void *)k_t_small<false, true>, T_SMALL_P, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&os[3], (const void *)k_t_small<true, true>, T_SMALL_P, 0));
    int osm = os[0];
    for (int k = 1; k < 4; ++k) osm = os[k] < osm ? os[k] : osm;
    e->small_grid = (osm > 0 ? osm : 1) * (ncu > 0 ? ncu : 1);
    int osf = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&osf, (const void *)k_t_small<true, false, true>, T_SMALL_P, 0));
    e->small_grid_fast = (osf > 0 ? osf : 1) * (ncu > 0 ? ncu : 1);
    e->giant_wait_ticks = 100LL * khz;
    bool may_launch = (e->cfg.flags & ME_FLAG_TILED) && e->cfg.solver != ME_SOLVER_JACOBI &&
                      (e->N > T_BIG_B || e->capP > T_BIG_P || e->N > T_COLB);
    if (may_launch && e->giant_grid < 1)
        return fail(ME_ERUNTIME, "the giant-island grid kernels cannot hold one workgroup resident on this device "
                                 "(occupancy " + std::to_string(sweep) + "/" + std::to_string(occ[2]) + " per CU, " +
                                 std::to_string(ncu) + " CUs)");
    return ME_OK;
}

// ── exact-cone PGS solver (pgs.hpp) ──
struct me_pgs {
    int NB = 0, MC = 0, rb = 4, flags = 0;
    int C = 0;
    int cached_C = -1;
    int64_t hits = 0, misses = 0;
    me_pgs_scene sc{};
    double soft[5] = {0, 0, 0, 0, 0};
    int soft_on = 0;
    void *rho = nullptr, *lam = nullptr, *v = nullptr, *vbias = nullptr, *res = nullptr, *uout = nullptr,
         *cMinv = nullptr, *lamo = nullptr, *pscal = nullptr, *plane = nullptr;
    int *pint = nullptr;
    int64_t *acc = nullptr, *key = nullptr, *prio = nullptr;
    long long *claim = nullptr;
    unsigned long long *rmax = nullptr;
    int *dyn = nullptr, *bucket = nullptr, *bcnt = nullptr, *boff = nullptr, *bfill = nullptr, *items = nullptr,
        *par = nullptr, *isl = nullptr, *icnt = nullptr, *ioff = nullptr, *ifill = nullptr, *iorder = nullptr,
        *col = nullptr, *order = nullptr, *tlo = nullptr, *thi = nullptr, *tncol = nullptr, *seg = nullptr,
        *cpab = nullptr, *cfeat = nullptr, *ps = nullptr, *bsum = nullptr, *tbcnt = nullptr, *tbstart = nullptr,
        *tbody = nullptr, *bslot = nullptr, *btile = nullptr;
    int nbsum = 0;
    int *hps = nullptr;  // pinned host mirror of ps
    unsigned long long *hrmax = nullptr;
};

namespace {

int pgs_ntile(int C) { return C > 0 ? (C + PGS_TILE - 1) / PGS_TILE : 0; }

size_t pgs_layout(me_pgs *s, char *base) {
    const size_t NB = (size_t)s->NB, MC = (size_t)s->MC, R = (size_t)s->rb;
    const size_t NI = NB + MC, NT = (size_t)pgs_ntile(s->MC) + 1;
    Bump B{base};
    s->rho = B.take<char>(R * MC);
    s->lam = B.take<char>(R * 3 * MC);
    s->v = B.take<char>(R * 6 * NB);
    s->vbias = B.take<char>(R * 6 * NB);
    s->res = B.take<char>(R * MC);
    s->uout = B.take<char>(R * 3 * MC);
    s->cMinv = B.take<char>(R * 36 * NB);
    s->lamo = B.take<char>(R * 3 * MC);
    s->pscal = B.take<char>(R * 8 * MC);
    s->plane = B.take<char>(R * 4 * PGS_LANE * MC);
    s->pint = B.take<int>(8 * MC);
    s->acc = B.take<int64_t>(6 * NB);
    s->key = B.take<int64_t>(MC);
    s->prio = B.take<int64_t>(MC);
    s->claim = B.take<long long>(NB);
    s->rmax = B.take<unsigned long long>(1);
    s->dyn = B.take<int>(NB);
    s->bucket = B.take<int>(MC);
    s->bcnt = B.take<int>(NB + 1);
    s->boff = B.take<int>(NB + 1);
    s->bfill = B.take<int>(NB + 1);
    s->items = B.take<int>(MC);
    s->par = B.take<int>(NB);
    s->isl = B.take<int>(MC);
    s->icnt = B.take<int>(NI);
    s->ioff = B.take<int>(NI);
    s->ifill = B.take<int>(NI);
    s->iorder = B.take<int>(MC);
    s->col = B.take<int>(MC);
    s->order = B.take<int>(MC);
    s->tlo = B.take<int>(NT);
    s->thi = B.take<int>(NT);
    s->tncol = B.take<int>(NT);
    s->seg = B.take<int>(NT * (PGS_NCOLMAX + 1));
    s->tbcnt = B.take<int>(NT);
    s->tbstart = B.take<int>(NT);
    s->tbody = B.take<int>(NB);
    s->bslot = B.take<int>(NB);
    s->btile = B.take<int>(NB);
    s->cpab = B.take<int>(2 * MC);
    s->cfeat = B.take<int>(MC);
    s->ps = B.take<int>(PS_COUNT);
    size_t scan_n = NI > NB + 1 ? NI : NB + 1;
    s->nbsum = (int)((scan_n + SCAN_BLOCK - 1) / SCAN_BLOCK);
    if (s->nbsum < 1) s->nbsum = 1;
    s->bsum = B.take<int>((size_t)s->nbsum);
    return B.off;
}

int pgs_scan(me_pgs *p, const int *in, int *out, int n, hipStream_t s) {
    int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb < 1) nb = 1;
    if (nb > p->nbsum) return fail(ME_ERUNTIME, "scan scratch too small");
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(ME_THREADS), 0, s, in, nullptr, n, n, p->bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, p->bsum, nb, nullptr);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(ME_THREADS), 0, s, in, out, nullptr, n, n, p->bsum);
    LAUNCH_CHECK();
    return ME_OK;
}

void pgs_fill(int *p, long long n, int val, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(ME_THREADS), 0, s, p, (int)n, val);
}

template <class T>
PgsScene<T> pgs_scene_of(const me_pgs *p) {
    PgsScene<T> S;
    S.Ja = (const T *)p->sc.Ja; S.Jb = (const T *)p->sc.Jb; S.Minv = (const T *)p->sc.Minv;
    S.vfree = (const T *)p->sc.v_free; S.mu = (const T *)p->sc.mu; S.bias = (const T *)p->sc.bias;
    S.sep = (const T *)p->sc.sep; S.pab = p->sc.body_pairs; S.feat = p->sc.feature;
    return S;
}

template <class T>
PgsSoft<T> pgs_soft_of(const me_pgs *p, int use_bias) {
    PgsSoft<T> sp;
    sp.p0 = (T)p->soft[0]; sp.p1 = (T)p->soft[1]; sp.p2 = (T)p->soft[2]; sp.p3 = (T)p->soft[3];
    sp.p4 = (T)p->soft[4];
    sp.on = p->soft_on;
    sp.use_bias = use_bias ? 1 : 0;
    return sp;
}

// deterministic colouring of the current scene (NCPSolverGPU._color, ncp_gpu.py:873-926), as tiles
int pgs_color(me_pgs *p, hipStream_t s) {
    const int C = p->C, NB = p->NB, NI = NB + C, nt = pgs_ntile(C);
    const int *pab = p->sc.body_pairs;
    pgs_fill(p->ps + PS_NCOL, 4, 0, s);  // PS_NCOL, PS_OVF, PS_NISL, PS_NTILE
    pgs_fill(p->bcnt, NB + 1, 0, s);
    pgs_fill(p->bfill, NB + 1, 0, s);
    pgs_fill(p->icnt, NI, 0, s);
    pgs_fill(p->ifill, NI, 0, s);
    pgs_fill(p->tlo, nt, INT_MAX, s);
    pgs_fill(p->thi, nt, INT_MIN, s);
    pgs_fill(p->tbcnt, nt, 0, s);
    hipLaunchKernelGGL(k_pgs_keys, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, NB, pab, p->sc.feature, p->key, p->bucket,
                       p->bcnt);
    int rc = pgs_scan(p, p->bcnt, p->boff, NB + 1, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pgs_bucket_fill, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, p->bucket, p->boff, p->bfill,
                       p->items);
    hipLaunchKernelGGL(k_pgs_rank, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, p->key, p->bucket, p->boff, p->bcnt,
                       p->items, p->prio);
    hipLaunchKernelGGL(k_pgs_uf_init, dim3(nblk(NB)), dim3(ME_THREADS), 0, s, NB, p->par);
    hipLaunchKernelGGL(k_pgs_union, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, pab, p->dyn, p->par);
    hipLaunchKernelGGL(k_pgs_island, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, NB, pab, p->dyn, p->par, p->isl,
                       p->icnt);
    rc = pgs_scan(p, p->icnt, p->ioff, NI, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pgs_tiles, dim3(nblk(NI)), dim3(ME_THREADS), 0, s, NI, p->icnt, p->ioff, p->tlo, p->thi,
                       p->ps);
    hipLaunchKernelGGL(k_pgs_island_fill, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, p->isl, p->ioff, p->ifill,
                       p->iorder);
    hipLaunchKernelGGL(k_pgs_body_slot, dim3(nblk(NB)), dim3(ME_THREADS), 0, s, NB, p->dyn, p->par, p->icnt, p->ioff,
                       p->tbcnt, p->bslot, p->btile);
    rc = pgs_scan(p, p->tbcnt, p->tbstart, nt, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pgs_body_fill, dim3(nblk(NB)), dim3(ME_THREADS), 0, s, NB, p->bslot, p->btile, p->tbstart,
                       p->tbody);
    hipLaunchKernelGGL(k_pgs_color, dim3(nt), dim3(PGS_THREADS), 0, s, pab, p->dyn, p->prio, p->tlo, p->thi,
                       p->iorder, p->claim, p->col, p->order, p->seg, p->tncol, p->ps);
    LAUNCH_CHECK();
    return ME_OK;
}

template <class T>
int pgs_prepare_t(me_pgs *p, const void *warm, int *hit_out, hipStream_t s) {
    const int C = p->C, NB = p->NB;
    PgsScene<T> S = pgs_scene_of<T>(p);
    T *lam = (T *)p->lam, *v = (T *)p->v, *vb = (T *)p->vbias;
    if (C > 0) {
        pgs_fill(p->ps + PS_BAD, 1, 0, s);
        hipLaunchKernelGGL(k_pgs_validate, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, NB, S.pab, p->ps);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(p->hps, p->ps, sizeof(int) * PS_COUNT, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (p->hps[PS_BAD]) return fail(ME_EARG, "body_pairs holds an index outside [-1, n_bodies)");
    }
    hipLaunchKernelGGL(k_pgs_dyn<T>, dim3(nblk(NB)), dim3(ME_THREADS), 0, s, NB, S.Minv, p->dyn);
    if (C > 0) hipLaunchKernelGGL(k_pgs_rho<T>, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, S, (T *)p->rho);
    long long n0 = 3LL * C > 6LL * NB ? 3LL * C : 6LL * NB;
    hipLaunchKernelGGL(k_pgs_lam_init<T>, dim3(nblk(n0)), dim3(ME_THREADS), 0, s, C, NB, (const T *)warm, lam, p->acc);
    if (warm && C > 0)
        hipLaunchKernelGGL(k_pgs_scatter<T>, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, S, p->dyn, lam, p->acc);
    hipLaunchKernelGGL(k_pgs_apply<T>, dim3(nblk(6LL * NB)), dim3(ME_THREADS), 0, s, NB, S.vfree, p->acc, v, vb);
    LAUNCH_CHECK();
    if (C == 0) {
        if (hit_out) *hit_out = 0;
        return ME_OK;
    }
    bool hit = false;
    if ((p->flags & ME_PGS_COLOR_CACHE) && C == p->cached_C) {
        pgs_fill(p->ps + PS_MISMATCH, 1, 0, s);
        int n = C > NB ? C : NB;
        hipLaunchKernelGGL(k_pgs_cache_cmp<T>, dim3(nblk(n)), dim3(ME_THREADS), 0, s, C, NB, S, p->cpab, p->cfeat,
                           (const T *)p->cMinv, p->ps);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(p->hps, p->ps, sizeof(int) * PS_COUNT, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        hit = p->hps[PS_MISMATCH] == 0;
    }
    if (hit) {
        p->hits += 1;
    } else {
        int rc = pgs_color(p, s);
        if (rc) return rc;
        long long n = 36LL * NB > C ? 36LL * NB : C;
        hipLaunchKernelGGL(k_pgs_cache_store<T>, dim3(nblk(n)), dim3(ME_THREADS), 0, s, C, NB, S, p->cpab, p->cfeat,
                           (T *)p->cMinv);
        LAUNCH_CHECK();
        p->cached_C = C;
        p->misses += 1;
    }
    hipLaunchKernelGGL(k_pgs_pack<T>, dim3(nblk(C)), dim3(ME_THREADS), 0, s, C, S, p->dyn, p->bslot, p->order,
                       (const T *)p->rho, p->pint, (T *)p->pscal, (T *)p->plane);
    LAUNCH_CHECK();
    if (hit_out) *hit_out = hit ? 1 : 0;
    return ME_OK;
}

template <class T>
int pgs_iterate_t(me_pgs *p, int iters, int use_bias, hipStream_t s) {
    if (p->C == 0 || iters <= 0) return ME_OK;
    const size_t lds = (size_t)6 * PGS_LDS_BODIES * (sizeof(int64_t) + sizeof(T));
    hipLaunchKernelGGL(k_pgs_sweeps<T>, dim3(pgs_ntile(p->C)), dim3(PGS_SWEEP_THREADS), lds, s, iters,
                       pgs_soft_of<T>(p, use_bias), (p->flags & ME_PGS_DESAXCE) ? 1 : 0, p->tlo, p->thi, p->tncol,
                       p->seg, p->order, p->tbstart, p->tbcnt, p->tbody, p->pint, (const T *)p->pscal,
                       (const T *)p->plane, (T *)p->lam, (T *)p->lamo, (T *)p->v, p->acc);
    LAUNCH_CHECK();
    return ME_OK;
}

template <class T>
int pgs_residual_t(me_pgs *p, int use_bias, double *out, hipStream_t s) {
    HIP_TRY(hipMemsetAsync(p->rmax, 0, sizeof(unsigned long long), s));
    if (p->C > 0) {
        PgsScene<T> S = pgs_scene_of<T>(p);
        hipLaunchKernelGGL(k_pgs_residual<T>, dim3(nblk(p->C)), dim3(ME_THREADS), 0, s, p->C, S,
                           pgs_soft_of<T>(p, use_bias), (p->flags & ME_PGS_DESAXCE) ? 1 : 0, (const T *)p->rho,
                           (const T *)p->lam, (const T *)p->v, (T *)p->res, (T *)p->uout, p->rmax);
        LAUNCH_CHECK();
    }
    HIP_TRY(hipMemcpyAsync(p->hrmax, p->rmax, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    double r;
    memcpy(&r, p->hrmax, sizeof(double));
    *out = r;
    return ME_OK;
}

}  // namespace

extern "C" {

int me_workspace_size(const me_config *cfg, size_t *bytes) {
    int rc = validate(cfg);
    if (rc) return rc;
    if (!bytes) return fail(ME_EARG, "bytes is NULL");
    me_engine tmp{};
    tmp.cfg = *cfg;
    layout(&tmp, nullptr, bytes);
    return ME_OK;
}

int me_create(const me_config *cfg, const me_bodies *bodies, void *workspace, size_t bytes, void *stream,
              me_engine **out) {
    int rc = validate(cfg);
    if (rc) return rc;
    if (!bodies || !out || !workspace) return fail(ME_EARG, "NULL argument");
    if (((uintptr_t)workspace & 255) != 0) return fail(ME_EARG, "workspace must be 256-byte aligned");
    if (!bodies->x || !bodies->q || !bodies->v || !bodies->w || !bodies->inv_mass || !bodies->inv_inertia ||
        !bodies->mu || !bodies->kinematic || !bodies->scene)
        return fail(ME_EARG, "body array pointer is NULL");
    if (cfg->n_beads > 0 && !bodies->rest) return fail(ME_EARG, "bead geometry needs rest positions");
    if (cfg->n_beads == 0 && !bodies->half) return fail(ME_EARG, "box geometry needs half extents");
    size_t need = 0;
    me_engine *e = new me_engine{};
    e->cfg = *cfg;
    layout(e, nullptr, &need);
    if (bytes < need) {
        delete e;
        return fail(ME_EARG, "workspace too small: need " + std::to_string(need) + " bytes");
    }
    layout(e, (char *)workspace, &need);
    e->b = *bodies;
    e->owner = std::this_thread::get_id();
    if (hipHostMalloc((void **)&e->hstat, sizeof(int) * IS_COUNT, hipHostMallocDefault) != hipSuccess) {
        e->hstat = nullptr;
        release(e);
        return fail(ME_ERUNTIME, "hipHostMalloc failed");
    }
    bool graph = (cfg->flags & ME_FLAG_GRAPH) != 0;
    if (hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess ||
        (graph && (hipStreamCreateWithFlags(&e->gs, hipStreamNonBlocking) != hipSuccess ||
                   hipEventCreateWithFlags(&e->ev_in, hipEventDisableTiming) != hipSuccess ||
                   hipEventCreateWithFlags(&e->ev_out, hipEventDisableTiming) != hipSuccess))) {
        release(e);
        return fail(ME_ERUNTIME, "stream/event creation failed");
    }
    if (tiled_set_attributes() != ME_OK) {
        release(e);
        return fail(ME_ERUNTIME, "setting the tile kernels' LDS limits failed");
    }
    rc = giant_setup(e);
    if (rc) {
        release(e);
        return rc;
    }
    hipStream_t s = (hipStream_t)stream;
    // zero the workspace with a fill kernel (no memset nodes, so a caller may capture later steps)
    size_t words = need / 4;
    for (size_t off = 0; off < words; off += (size_t)1 << 30) {
        int n = (int)((words - off) < ((size_t)1 << 30) ? (words - off) : ((size_t)1 << 30));
        hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(ME_THREADS), 0, s, (int *)workspace + off, n, 0);
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, s, e->istat + IS_CACHED_P, 1, -1);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        release(e);
        return fail(ME_ERUNTIME, "workspace initialisation failed");
    }
    rc = scene_bounds(e);
    if (rc == ME_OK) rc = bead_half_extents(e, s);
    if (rc) {
        release(e);
        return rc;
    }
    *out = e;
    return ME_OK;
}

int me_destroy(me_engine *e) {
    if (!e) return fail(ME_EARG, "engine is NULL");
    release(e);
    return ME_OK;
}

int me_adopt(me_engine *e) {
    if (!e) return fail(ME_EARG, "engine is NULL");
    e->owner = std::this_thread::get_id();
    return ME_OK;
}

int me_set_bodies(me_engine *e, const me_bodies *bodies) {
    if (!e || !bodies) return fail(ME_EARG, "NULL argument");
    if (std::this_thread::get_id() != e->owner) return fail(ME_ETHREAD, "engine used from a thread that does not own it");
    if (e->P > 0 && !bodies->rest) return fail(ME_EARG, "bead geometry needs rest positions");
    e->b = *bodies;
    drop_graphs(e);  // captured kernels hold the old body pointers
    int rc = scene_bounds(e);
    if (rc) return rc;
    return bead_half_extents(e, nullptr);
}

int me_step(me_engine *e, double dt, int substeps, void *stream) {
    if (!e) return fail(ME_EARG, "engine is NULL");
    if (std::this_thread::get_id() != e->owner) return fail(ME_ETHREAD, "engine used from a thread that does not own it");
    if (substeps < 1) return fail(ME_EARG, "substeps must be >= 1");
    if (!(dt > 0.0) || dt != dt) return fail(ME_EARG, "dt must be positive and finite");
    hipStream_t s = (hipStream_t)stream;
    if ((e->cfg.flags & ME_FLAG_GRAPH) && !e->prof_on) return step_graph(e, dt, substeps, s);
    double sdt = dt / substeps;
    for (int k = 0; k < substeps; ++k) {
        int rc = substep(e, sdt, s);
        if (rc) return rc;
    }
    return ME_OK;
}

int me_buffer(me_engine *e, int id, void **ptr, int64_t *count) {
    if (!e || !ptr || !count) return fail(ME_EARG, "NULL argument");
    int64_t C = e->capC, P = e->capP, N = e->N;
    switch (id) {
        case ME_BUF_KEY: *ptr = e->Cc.key; *count = C; break;
        case ME_BUF_BI: *ptr = e->Cc.bi; *count = C; break;
        case ME_BUF_BJ: *ptr = e->Cc.bj; *count = C; break;
        case ME_BUF_PA: *ptr = e->Cc.pA; *count = 3 * C; break;
        case ME_BUF_PB: *ptr = e->Cc.pB; *count = 3 * C; break;
        case ME_BUF_N: *ptr = e->Cc.n; *count = 3 * C; break;
        case ME_BUF_PEN: *ptr = e->Cc.pen; *count = C; break;
        case ME_BUF_JN: *ptr = e->Cc.jn; *count = C; break;
        case ME_BUF_JT1: *ptr = e->Cc.jt1; *count = C; break;
        case ME_BUF_JT2: *ptr = e->Cc.jt2; *count = C; break;
        case ME_BUF_JP: *ptr = e->Cc.jp; *count = C; break;
        case ME_BUF_PBI: *ptr = e->Pp.pbi; *count = P; break;
        case ME_BUF_PBJ: *ptr = e->Pp.pbj; *count = P; break;
        case ME_BUF_PSTART: *ptr = e->Pp.pstart; *count = P; break;
        case ME_BUF_PCOUNT: *ptr = e->Pp.pcount; *count = P; break;
        case ME_BUF_COLOR: *ptr = e->Pp.col; *count = P; break;
        case ME_BUF_VSNAP: *ptr = e->vsnap; *count = 3 * N; break;
        case ME_BUF_PV: *ptr = e->pv; *count = 3 * N; break;
        case ME_BUF_PO: *ptr = e->po; *count = 3 * N; break;
        case ME_BUF_INVIW: *ptr = e->invIw; *count = 9 * N; break;
        case ME_BUF_ISLAND: *ptr = e->tb.root; *count = N; break;
        case ME_BUF_DEBUG: *ptr = e->tb.dbg; *count = 16; break;
        case ME_BUF_BIG_WORK: *ptr = e->tb.bigw; *count = 4; break;
        case ME_BUF_TILE_WORK: *ptr = e->tb.bigw; *count = 12; break;
        case ME_BUF_MU: *ptr = e->Cc.mu; *count = C; break;
        case ME_BUF_XSNAP: *ptr = e->xsnap; *count = 3 * N; break;
        case ME_BUF_WSNAP: *ptr = e->wsnap; *count = 3 * N; break;
        default: return fail(ME_EARG, "unknown buffer id");
    }
    return ME_OK;
}

int me_read_counters(me_engine *e, int64_t *out, void *stream) {
    if (!e || !out) return fail(ME_EARG, "NULL argument");
    int rc = read_stat(e, (hipStream_t)stream);
    if (rc) return rc;
    const int *h = e->hstat;
    out[ME_CNT_CONTACTS] = h[IS_CRAW];
    out[ME_CNT_SOLVED] = h[IS_C];
    out[ME_CNT_PAIRS] = h[IS_P];
    out[ME_CNT_COLORS] = h[IS_NCOL];
    out[ME_CNT_CACHE_HITS] = h[IS_HITS];
    out[ME_CNT_CACHE_MISSES] = h[IS_MISSES];
    out[ME_CNT_OVERFLOW] = h[IS_OVF];
    out[ME_CNT_CANDIDATES] = h[IS_CAND];
    out[ME_CNT_ISLANDS] = h[IS_ISL];
    out[ME_CNT_TILES] = h[IS_TILES];
    out[ME_CNT_MAX_TILE_BODIES] = h[IS_MAXTB];
    out[ME_CNT_GRAPH_CAPTURES] = e->n_captures;
    return ME_OK;
}

int me_profile(me_engine *e, int enable) {
    if (!e) return fail(ME_EARG, "engine is NULL");
    e->prof_on = enable != 0;
    if (e->prof_on) {  // a new measurement window; disabling keeps the recorded events for reading
        for (int k = 0; k < ME_K_COUNT_; ++k) e->prof_used[k] = 0;
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemset(e->tb.bigw, 0, 12 * sizeof(long long)));
        HIP_TRY(hipDeviceSynchronize());
    }
    return ME_OK;
}

int me_profile_read(me_engine *e, int kernel, double *total_ms, int64_t *launches) {
    if (!e || !total_ms || !launches) return fail(ME_EARG, "NULL argument");
    if (kernel < 0 || kernel >= ME_K_COUNT_) return fail(ME_EARG, "unknown kernel id");
    double tot = 0.0;
    int n = e->prof_used[kernel] / 2;
    for (int k = 0; k < n; ++k) {
        hipEvent_t a = e->prof_ev[kernel][2 * k], b = e->prof_ev[kernel][2 * k + 1];
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        tot += ms;
    }
    *total_ms = tot;
    *launches = n;
    return ME_OK;
}

// NCP sensitivity: LDS holds the augmented matrix up to this size, else the caller's scratch does
static const size_t kNcpLds = 64 * 1024;

static size_t ncp_matrix_bytes(int n_c, int n_dir) {
    size_t n3 = 3 * (size_t)n_c;
    return n3 * (n3 + (size_t)n_c + (size_t)n_dir) * sizeof(double);
}

int me_ncp_workspace_size(int32_t n_env, int32_t n_c, int32_t n_dir, size_t *bytes) {
    if (!bytes) return fail(ME_EARG, "bytes is NULL");
    if (n_env < 1 || n_c < 1 || n_c > NCP_MAXC || n_dir < 0)
        return fail(ME_EARG, "need n_env >= 1, 1 <= n_c <= 64, n_dir >= 0");
    size_t m = ncp_matrix_bytes(n_c, n_dir);
    *bytes = m <= kNcpLds ? 0 : m * (size_t)n_env;
    return ME_OK;
}

static int ncp_check(int32_t n_env, int32_t n_c, int32_t n_dir, const double *G, const double *b, const double *mu,
                     const double *lam, const double *dirs, double tol_lam, double tol_rel, double rcond,
                     double *dlam_dmu, double *dlam_db, int32_t *labels, int32_t *status) {
    if (n_env < 1 || n_c < 1 || n_c > NCP_MAXC || n_dir < 0)
        return fail(ME_EARG, "need n_env >= 1, 1 <= n_c <= 64, n_dir >= 0");
    if (!dirs && n_dir != 3 * n_c) return fail(ME_EARG, "without directions n_dir must be 3 * n_c (the identity)");
    if (!G || !b || !mu || !lam || !dlam_dmu || (n_dir > 0 && !dlam_db) || !labels || !status)
        return fail(ME_EARG, "NULL argument");
    if (!(tol_lam >= 0.0) || !(tol_rel >= 0.0) || !(rcond > 0.0 && rcond < 1.0))
        return fail(ME_EARG, "need tol_lam >= 0, tol_rel >= 0, 0 < rcond < 1");
    return ME_OK;
}

static int ncp_launch(int32_t n_env, int32_t n_c, int32_t n_dir, const double *G, const double *b, const double *mu,
                      const double *lam, const double *dirs, const int32_t *labels_in, double tol_lam, double tol_rel,
                      double rcond, double *dlam_dmu, double *dlam_db, int32_t *labels, int32_t *status,
                      double *scratch_rows, void *stream) {
    size_t m = ncp_matrix_bytes(n_c, n_dir);
    int lds = m <= kNcpLds ? 1 : 0;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_ncp_sensitivity, dim3(n_env), dim3(NCP_THREADS), lds ? m : 0, s, n_env, n_c, n_dir, G, b, mu,
                       lam, dirs, labels_in, tol_lam, tol_rel, rcond, dlam_dmu, dlam_db, labels, status, scratch_rows,
                       lds);
    LAUNCH_CHECK();
    return ME_OK;
}

int me_ncp_sensitivity(int32_t n_env, int32_t n_c, int32_t n_dir, const double *G, const double *b,
                       const double *mu, const double *lam, const double *dirs, const int32_t *labels_in,
                       double tol_lam, double tol_rel, double rcond, double *dlam_dmu, double *dlam_db,
                       int32_t *labels, int32_t *status, void *workspace, size_t bytes, void *stream) {
    int rc = ncp_check(n_env, n_c, n_dir, G, b, mu, lam, dirs, tol_lam, tol_rel, rcond, dlam_dmu, dlam_db, labels,
                       status);
    if (rc) return rc;
    size_t need = 0;
    (void)me_ncp_workspace_size(n_env, n_c, n_dir, &need);
    if (need > 0 && (!workspace || bytes < need))
        return fail(ME_EARG, "workspace too small: need " + std::to_string(need) + " bytes");
    return ncp_launch(n_env, n_c, n_dir, G, b, mu, lam, dirs, labels_in, tol_lam, tol_rel, rcond, dlam_dmu, dlam_db,
                      labels, status, (double *)workspace, stream);
}

// The no-allocation variant's scratch: the augmented matrices of systems beyond the LDS limit, one
// fixed array per GPU for the life of the process. Host code never passes the array's name to a
// kernel (in host code it is only a stand-in symbol): its device address is looked up once per
// device with hipGetSymbolAddress, on the entry point's first call there, and kept per device.
static __device__ __align__(256) unsigned char scratch[ME_NCP_STATIC_BYTES];

namespace {
const int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(ME_ERUNTIME, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !ptr)
            return fail(ME_ERUNTIME, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return ME_OK;
}
}  // namespace

int me_ncp_sensitivity_noalloc(int32_t n_env, int32_t n_c, int32_t n_dir, const double *G, const double *b,
                               const double *mu, const double *lam, const double *dirs, const int32_t *labels_in,
                               double tol_lam, double tol_rel, double rcond, double *dlam_dmu, double *dlam_db,
                               int32_t *labels, int32_t *status, void *stream) {
    int rc = ncp_check(n_env, n_c, n_dir, G, b, mu, lam, dirs, tol_lam, tol_rel, rcond, dlam_dmu, dlam_db, labels,
                       status);
    if (rc) return rc;
    size_t need = 0;
    (void)me_ncp_workspace_size(n_env, n_c, n_dir, &need);
    if (need > (size_t)ME_NCP_STATIC_BYTES)
        return fail(ME_ECAPACITY, "me_ncp_sensitivity_noalloc: " + std::to_string(need) + " bytes of matrices exceed "
                                  "the fixed scratch of " + std::to_string((size_t)ME_NCP_STATIC_BYTES) +
                                  " bytes; use fewer environments per call or me_ncp_sensitivity");
    void *ptr = nullptr;
    if (need > 0) {
        rc = scratch_address(&ptr);
        if (rc) return rc;
    }
    return ncp_launch(n_env, n_c, n_dir, G, b, mu, lam, dirs, labels_in, tol_lam, tol_rel, rcond, dlam_dmu, dlam_db,
                      labels, status, (double *)ptr, stream);
}

int me_ncp_svd_workspace_size(int32_t n_env, int32_t n_c, int32_t n_dir, size_t *bytes) {
    if (!bytes) return fail(ME_EARG, "bytes is NULL");
    if (n_env < 1 || n_c < 1 || n_c > NCPS_MAXC || n_dir < 0)
        return fail(ME_EARG, "need n_env >= 1, 1 <= n_c <= 256, n_dir >= 0");
    *bytes = ncp_svd_scratch(n_c, n_dir) * sizeof(double) * (size_t)n_env;
    return ME_OK;
}

int me_ncp_svd(int32_t n_env, int32_t n_c, int32_t n_dir, int32_t n_obs, const double *G, const double *b,
               const double *mu, const double *lam, const double *dirs, const int32_t *labels_in, const double *Q,
               int32_t q_shared, double tol_lam, double tol_rel, double rcond, double tol, int32_t minimum_norm,
               const me_ncp_svd_out *out, void *workspace, size_t bytes, void *stream) {
    if (n_env < 1 || n_c < 1 || n_c > NCPS_MAXC || n_dir < 0 || n_obs < 0)
        return fail(ME_EARG, "need n_env >= 1, 1 <= n_c <= 256, n_dir >= 0, n_obs >= 0");
    if (!dirs && n_dir != 3 * n_c && n_dir != 0)
        return fail(ME_EARG, "without directions n_dir must be 3 * n_c (the identity) or 0 (no b columns)");
    if (!G || !b || !mu || !lam || !out || (n_obs > 0 && !Q)) return fail(ME_EARG, "NULL argument");
    const me_ncp_svd_out &o = *out;
    if (!o.X || !o.S || !o.V || !o.U || !o.rank || !o.compat || !o.finite || !o.resid || !o.labels || !o.status ||
        (n_obs > 0 && (!o.Y || !o.valid || !o.invariant || !o.leak)))
        return fail(ME_EARG, "NULL output");
    if (!(tol_lam >= 0.0) || !(tol_rel >= 0.0) || !(rcond > 0.0 && rcond < 1.0) || !(tol > 0.0 && tol < 1.0))
        return fail(ME_EARG, "need tol_lam >= 0, tol_rel >= 0, 0 < rcond < 1, 0 < tol < 1");
    size_t need = 0;
    (void)me_ncp_svd_workspace_size(n_env, n_c, n_dir, &need);
    if (!workspace || bytes < need) return fail(ME_EARG, "workspace too small: need " + std::to_string(need) + " bytes");
    NcpSvdOut k{o.X, o.S, o.V, o.U, o.rank, o.compat, o.finite, o.resid, o.Y, o.valid, o.invariant, o.leak,
                o.labels, o.status};
    // no b columns: a non-null direction pointer of zero columns (never read) keeps the identity out
    const double *D = (!dirs && n_dir == 0) ? b : dirs;
    hipStream_t s = (hipStream_t)stream;
    const int n3 = 3 * n_c;
    double *ws = (double *)workspace;
    if (n3 <= 24)
        hipLaunchKernelGGL(k_ncp_svd<64>, dim3(n_env), dim3(64), 0, s, n_env, n_c, n_dir, n_obs, G, b, mu, lam, D,
                           labels_in, Q, q_shared, tol_lam, tol_rel, rcond, tol, minimum_norm, k, ws);
    else if (n3 <= 192)
        hipLaunchKernelGGL(k_ncp_svd<256>, dim3(n_env), dim3(256), 0, s, n_env, n_c, n_dir, n_obs, G, b, mu, lam, D,
                           labels_in, Q, q_shared, tol_lam, tol_rel, rcond, tol, minimum_norm, k, ws);
    else
        hipLaunchKernelGGL(k_ncp_svd<1024>, dim3(n_env), dim3(1024), 0, s, n_env, n_c, n_dir, n_obs, G, b, mu, lam, D,
                           labels_in, Q, q_shared, tol_lam, tol_rel, rcond, tol, minimum_norm, k, ws);
    LAUNCH_CHECK();
    return ME_OK;
}

int me_state_size(me_engine *e, size_t *bytes) {
    if (!e || !bytes) return fail(ME_EARG, "NULL argument");
    *bytes = e->state_bytes;
    return ME_OK;
}

int me_state_copy(me_engine *e, void *dst, const void *src, void *stream) {
    if (!e || !dst || !src) return fail(ME_EARG, "NULL argument");
    if (std::this_thread::get_id() != e->owner) return fail(ME_ETHREAD, "engine used from a thread that does not own it");
    if (((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) return fail(ME_EARG, "dst and src must be 16-byte aligned");
    if (e->segs.overflow || e->segs.n < 1) return fail(ME_ERUNTIME, "state segment table incomplete");
    hipLaunchKernelGGL(k_state_copy, dim3(64, e->segs.n), dim3(ME_THREADS), 0, (hipStream_t)stream, (char *)dst,
                       (const char *)src, e->segs);
    return hipGetLastError() == hipSuccess ? ME_OK : fail(ME_ERUNTIME, "state copy launch failed");
}

int me_tape_size(me_engine *e, size_t *bytes) {
    if (!e || !bytes) return fail(ME_EARG, "NULL argument");
    *bytes = tape_layout(e, nullptr, nullptr);
    return ME_OK;
}

int me_step_taped(me_engine *e, double dt, void *tape, size_t bytes, void *stream) {
    int rc = tape_check(e, tape, bytes);
    if (rc) return rc;
    if (!(dt > 0.0) || dt != dt) return fail(ME_EARG, "dt must be positive and finite");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_tape_xq, dim3(nblk(e->N)), dim3(ME_THREADS), 0, s, e->N, e->b.x, e->b.q, e->tape.xpre,
                       e->tape.qpre);
    e->taping = true;
    rc = substep(e, dt, s);
    e->taping = false;
    return rc;
}

int me_adjoint_substep(me_engine *e, double dt, void *tape, size_t bytes, float *xbar, float *qbar, float *vbar,
                       float *wbar, const float *jbar_end, float *jbar_prev, float *mubar, void *stream) {
    int rc = tape_check(e, tape, bytes);
    if (rc) return rc;
    if (!xbar || !qbar || !vbar || !wbar || !jbar_end || !jbar_prev || !mubar) return fail(ME_EARG, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    int N = e->N, capC = e->capC;
    const AdjTape &t = e->tape;
    // the reverse pass's accumulators, zeroed in one launch
    ZeroSpans z{};
    float *zp[7] = {t.cbar, t.jbi, t.gbar, t.pvb, t.pob, jbar_prev, (float *)t.acc};
    size_t zn[7] = {20 * (size_t)capC, 3 * (size_t)capC, 7 * (size_t)capC, 3 * (size_t)N, 3 * (size_t)N,
                    3 * (size_t)capC, 2 * ADJ_ACC * (size_t)N};
    for (int k = 0; k < 7; ++k) { z.p[k] = zp[k]; z.n[k] = zn[k]; }
    z.k = 7;
    hipLaunchKernelGGL(k_zero_spans, dim3(2048), dim3(ME_THREADS), 0, s, z);
    e->taping = true;
    TiledArgs A = tiled_args_of(e, (float)dt);
    e->taping = false;
    int ntile = tiled_ntile(A);
    hipLaunchKernelGGL(k_adj_integ, dim3(nblk(N)), dim3(ME_THREADS), 0, s, A, xbar, qbar, vbar, wbar);
    hipLaunchKernelGGL(k_adj_tiles, dim3(ntile), dim3(T_SMALL_P), 0, s, A, e->tb, vbar, wbar, jbar_end);
    hipLaunchKernelGGL(k_adj_contacts, dim3(nblk(capC)), dim3(ME_THREADS), 0, s, A, jbar_prev);
    hipLaunchKernelGGL(k_adj_narrow, dim3(nblk(2LL * e->cap_cand)), dim3(ME_THREADS), 0, s, A, e->cap_cand, e->cand_a,
                       e->cand_b, e->S.n, e->slot_off, e->cfg.margin);
    hipLaunchKernelGGL(k_adj_body, dim3(nblk(N)), dim3(ME_THREADS), 0, s, A, xbar, qbar, mubar);
    LAUNCH_CHECK();
    return ME_OK;
}

// ── exact-cone PGS ──

int me_pgs_workspace_size(int32_t n_bodies, int32_t max_contacts, int32_t real_bytes, size_t *bytes) {
    if (!bytes) return fail(ME_EARG, "bytes is NULL");
    if (n_bodies < 1 || max_contacts < 1) return fail(ME_EARG, "need n_bodies >= 1 and max_contacts >= 1");
    if (real_bytes != 4 && real_bytes != 8) return fail(ME_EARG, "real_bytes must be 4 (float32) or 8 (float64)");
    if ((long long)n_bodies + max_contacts > (1LL << 30)) return fail(ME_EARG, "n_bodies + max_contacts too large");
    me_pgs tmp;
    tmp.NB = n_bodies; tmp.MC = max_contacts; tmp.rb = real_bytes;
    *bytes = pgs_layout(&tmp, nullptr);
    return ME_OK;
}

int me_pgs_create(int32_t n_bodies, int32_t max_contacts, int32_t real_bytes, int32_t flags, void *workspace,
                  size_t bytes, me_pgs **out) {
    if (!out) return fail(ME_EARG, "out is NULL");
    *out = nullptr;
    size_t need = 0;
    int rc = me_pgs_workspace_size(n_bodies, max_contacts, real_bytes, &need);
    if (rc) return rc;
    if (!workspace || bytes < need) return fail(ME_EARG, "workspace too small: need " + std::to_string(need) + " bytes");
    if ((reinterpret_cast<uintptr_t>(workspace) & 255) != 0) return fail(ME_EARG, "workspace must be 256-byte aligned");
    me_pgs *p = new me_pgs();
    p->NB = n_bodies; p->MC = max_contacts; p->rb = real_bytes;
    p->flags = flags & (ME_PGS_DESAXCE | ME_PGS_COLOR_CACHE);
    pgs_layout(p, (char *)workspace);
    if (hipHostMalloc((void **)&p->hps, sizeof(int) * PS_COUNT + sizeof(unsigned long long)) != hipSuccess) {
        delete p;
        return fail(ME_ERUNTIME, "hipHostMalloc failed");
    }
    p->hrmax = reinterpret_cast<unsigned long long *>(p->hps + PS_COUNT);
    memset(p->hps, 0, sizeof(int) * PS_COUNT);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(ME_THREADS), 0, nullptr, p->ps, PS_COUNT, 0);
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipHostFree(p->hps);
        delete p;
        return fail(ME_ERUNTIME, std::string("me_pgs_create: ") + hipGetErrorString(e));
    }
    *out = p;
    return ME_OK;
}

int me_pgs_destroy(me_pgs *p) {
    if (!p) return ME_OK;
    if (p->hps) (void)hipHostFree(p->hps);
    delete p;
    return ME_OK;
}

int me_pgs_prepare(me_pgs *p, const me_pgs_scene *sc, int32_t n_c, const void *warm, const double *soft,
                   int32_t *cache_hit, void *stream) {
    if (!p || !sc) return fail(ME_EARG, "NULL argument");
    if (n_c < 0 || n_c > p->MC)
        return fail(ME_EARG, std::to_string(n_c) + " contacts exceeds capacity " + std::to_string(p->MC));
    if (!sc->Minv || !sc->v_free) return fail(ME_EARG, "scene Minv / v_free is NULL");
    if (n_c > 0 && (!sc->Ja || !sc->Jb || !sc->mu || !sc->bias || !sc->body_pairs || !sc->feature))
        return fail(ME_EARG, "scene contact array is NULL");
    if (soft && n_c > 0 && !sc->sep) return fail(ME_EARG, "soft-step mode needs the separations (sep)");
    p->sc = *sc;
    p->C = n_c;
    p->soft_on = soft ? 1 : 0;
    for (int k = 0; k < 5; ++k) p->soft[k] = soft ? soft[k] : 0.0;
    int hit = 0;
    hipStream_t s = (hipStream_t)stream;
    int rc = p->rb == 8 ? pgs_prepare_t<double>(p, warm, &hit, s) : pgs_prepare_t<float>(p, warm, &hit, s);
    if (cache_hit) *cache_hit = hit;
    return rc;
}

int me_pgs_iterate(me_pgs *p, int32_t iters, int32_t use_bias, void *stream) {
    if (!p) return fail(ME_EARG, "NULL solver");
    if (iters < 0) return fail(ME_EARG, "iters must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    return p->rb == 8 ? pgs_iterate_t<double>(p, iters, use_bias, s) : pgs_iterate_t<float>(p, iters, use_bias, s);
}

int me_pgs_snapshot(me_pgs *p, void *stream) {
    if (!p) return fail(ME_EARG, "NULL solver");
    int n = 6 * p->NB * (p->rb / 4);
    hipLaunchKernelGGL(k_pgs_copy<float>, dim3(nblk(n)), dim3(ME_THREADS), 0, (hipStream_t)stream, n,
                       (const float *)p->v, (float *)p->vbias);
    LAUNCH_CHECK();
    return ME_OK;
}

int me_pgs_residual(me_pgs *p, int32_t use_bias, double *residual, void *stream) {
    if (!p || !residual) return fail(ME_EARG, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    return p->rb == 8 ? pgs_residual_t<double>(p, use_bias, residual, s)
                      : pgs_residual_t<float>(p, use_bias, residual, s);
}

int me_pgs_buffer(me_pgs *p, int32_t which, void **ptr, int64_t *count) {
    if (!p || !ptr || !count) return fail(ME_EARG, "NULL argument");
    const int64_t C = p->C, NB = p->NB;
    switch (which) {
        case ME_PGS_LAM: *ptr = p->lam; *count = 3 * C; break;
        case ME_PGS_V: *ptr = p->v; *count = 6 * NB; break;
        case ME_PGS_VBIAS: *ptr = p->vbias; *count = 6 * NB; break;
        case ME_PGS_RHO: *ptr = p->rho; *count = C; break;
        case ME_PGS_RES: *ptr = p->res; *count = C; break;
        case ME_PGS_U: *ptr = p->uout; *count = 3 * C; break;
        case ME_PGS_COLOR: *ptr = p->col; *count = C; break;
        default: return fail(ME_EARG, "unknown buffer id " + std::to_string(which));
    }
    return ME_OK;
}

int me_pgs_info(me_pgs *p, int64_t *out) {
    if (!p || !out) return fail(ME_EARG, "NULL argument");
    HIP_TRY(hipMemcpy(p->hps, p->ps, sizeof(int) * PS_COUNT, hipMemcpyDeviceToHost));
    out[ME_PGS_NCOLORS] = p->hps[PS_NCOL];
    out[ME_PGS_CACHE_HITS] = p->hits;
    out[ME_PGS_CACHE_MISSES] = p->misses;
    out[ME_PGS_ISLANDS] = p->hps[PS_NISL];
    out[ME_PGS_TILES] = p->hps[PS_NTILE];
    out[ME_PGS_OVERFLOW] = p->hps[PS_OVF];
    return ME_OK;
}

const char *me_last_error(void) { return g_err.c_str(); }

}  // extern "C"
