// nbody_hip.hip -- host orchestration and C ABI of libnbody_hip.
//
// Replaces OpenClSimulation's host side (open_cl_simulation.cpp:15-616): the
// constructor's upload and octree build, step()'s phase sequence, and
// particles().  All per-step phases run on one HIP stream; the host reads the
// device counters twice per force evaluation (after the tree count, to size
// the node arrays, and after the walk's counting phase, to size the lists).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nbody_hip.h"
#include "common.h"
#include "expansion.h"
#include "launch.h"

using namespace nbh;

namespace {

thread_local std::string g_create_error = "no error";

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes && p) return hipSuccess;
        size_t nb = std::max(need, bytes + bytes / 2);
        nb = std::max<size_t>(nb, 256);
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&p, nb);
        if (e != hipSuccess) { p = nullptr; return e; }
        bytes = nb;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return (T*)p; }
};

struct StateBufs {
    DevBuf x, y, z, vx, vy, vz, mass, charge, ax, ay, az, pot, id;
    hipError_t alloc(size_t n) {
        DevBuf* d[] = {&x, &y, &z, &vx, &vy, &vz, &mass, &charge};
        for (DevBuf* b : d) { hipError_t e = b->ensure(n * 8); if (e) return e; }
        DevBuf* f[] = {&ax, &ay, &az, &pot, &id};
        for (DevBuf* b : f) { hipError_t e = b->ensure(n * 4); if (e) return e; }
        return hipSuccess;
    }
    State view() const {
        return State{x.as<double>(), y.as<double>(), z.as<double>(), vx.as<double>(), vy.as<double>(),
                     vz.as<double>(), mass.as<double>(), charge.as<double>(), ax.as<float>(), ay.as<float>(),
                     az.as<float>(), pot.as<float>(), id.as<uint32_t>()};
    }
    void release() {
        DevBuf* all[] = {&x, &y, &z, &vx, &vy, &vz, &mass, &charge, &ax, &ay, &az, &pot, &id};
        for (DevBuf* b : all) b->release();
    }
};

// Launch shapes of a captured step, taken from the last uncaptured step with
// headroom: the grids and the level range are fixed in the graph, the
// buffers sized so that their pointers stay valid across replays, and the
// kernels read the actual counts on the device.  A step that outgrows them
// flags an overflow, is rolled back and run uncaptured.
struct GraphShape {
    uint32_t node_cap = 0;       // node arrays
    uint32_t node_items = 0;     // grid sizing: nodes, leaves
    uint32_t leaf_items = 0;
    uint32_t depth_cap = 0;      // levels 0..depth_cap are processed
    uint32_t level_items[MAX_LEVELS + 1] = {};
    uint32_t walk_grid[MAX_LEVELS + 1] = {};
    uint32_t far_cap = 0;
    unsigned long long m2l_cap = 0, p2p_cap = 0;
    uint32_t sl_cap = 0, row_cap = 0;
};

}  // namespace

struct nbh_handle {
    nbh_config cfg{};
    std::string err = "no error";
    int device = 0;
    hipStream_t stream = nullptr;
    size_t n = 0;
    int nsys = 1;
    int bits = 21;
    uint32_t max_sys_n = 0;
    int theta_num = 1, theta_den = 2;
    double time = 0.0;
    bool acc_valid = false;
    // the evaluation being built computes potentials (FmmParams::pot); the
    // potentials in the state are those of its positions (steps evaluate
    // accelerations only, as the reference's step does)
    bool eval_pot = true;
    bool pot_valid = false;
    bool timing = false;
    int max_lds = 65536;  // LDS bytes per workgroup on this device
    uint32_t walk_grid = 1024;  // group-kernel workgroups resident at once
    std::vector<uint32_t> sys_off_h;

    StateBufs st[2];
    int cur = 0;
    DevBuf sys_off, keys, keys_alt, perm, perm_alt, sort_tmp, tree_tmp, scan_tmp;
    DevBuf box, inv, unit, bbox_part;
    DevBuf geo, link, nsys_of, node_begin, leaf_ids, M, L, Lx, pf, pleaf, dtab, xtab, far_src, far_tgt;
    DevBuf row_node, row_first, row_next, row_sl_start, row_sl_cnt, sl, row_m_start, row_m_cnt, row_mx_cnt,
        row_p_start, row_p_cnt, m2l_pool, p2p_pool;
    unsigned long long m2l_pool_cap = 0, p2p_pool_cap = 0;
    DevBuf counters, diag, tmp_small;
    Counters* hc = nullptr;  // pinned host copy
    DevBuf Mw, Lw, pw, wbuf, gbuf;   // adjoint
    // same-level M2L by translation offset (matrix cores): split moments
    // (forward and adjoint), dense tensor tables, candidate offsets
    DevBuf mb, mbw, d2tab, cand, hank, opar;
    DevBuf sl_offs, sl_map;  // same-level table: offsets that occur (rows), offset -> row map
    DevBuf acc_save;         // nbh_diagnostics: the accelerations kept across its evaluation
    int sl_rows = 0;         // rows of the same-level table per level, its zero row included
    // trajectory adjoint: checkpoints of the full state (every adj_k steps),
    // the final state, the recomputed states of one segment, the cotangents
    std::vector<DevBuf> adj_ckpt, adj_seg;
    std::vector<double> adj_ckpt_t, adj_seg_t;
    DevBuf adj_final, adj_xb, adj_vb, adj_mb, adj_w;
    double adj_final_t = 0.0;
    int adj_nsteps = 0, adj_k = 0;
    bool adj_ready = false;
    int ncand = 0, offK = 0;
    bool gravity = true;             // charge == mass for every particle
    FmmParams lastF{};               // parameters of the last FMM force evaluation
    uint32_t node_cap = 0, row_cap = 0, sl_cap = 0;
    uint32_t num_nodes = 0, num_leaves = 0, depth = 0;
    nbh_stats stats{};
    hipEvent_t ev[8] = {};           // phase timing on the main stream
    hipEvent_t evb[4] = {};          // phase timing on the side stream
    // The side stream runs the moments and tables while the walk runs, then
    // the near field while M2L and L2L run; these order the two streams.
    hipStream_t stream2 = nullptr;
    // the side streams as created (nbh_set_timing mode 2 points stream2..4 at
    // the main stream, which serializes the evaluation)
    hipStream_t own2 = nullptr, own3 = nullptr, own4 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_side = nullptr, ev_walk = nullptr, ev_near = nullptr;
    // A third stream runs each level's same-level M2L as soon as the walk
    // depth that emits it is done (ev_depth), while the walk goes deeper.
    // (two such streams, taking alternate levels, so that a big level
    // does not wait behind the one before it)
    hipStream_t stream3 = nullptr, stream4 = nullptr;
    hipEvent_t ev_depth[MAX_LEVELS + 1] = {};
    hipEvent_t ev_same = nullptr, ev_same4 = nullptr;  // same-level M2L done
    hipEvent_t evc0 = nullptr;     // timing: first same-level launch
    // timing: each level's same-level M2L launch, bracketed on its own stream
    hipEvent_t ev_lvl[2 * (MAX_LEVELS + 1)] = {};
    bool lvl_timed[MAX_LEVELS + 1] = {};
    // captured steps (cfg.graph): one executable graph per state buffer
    // parity (the tree build alternates the two state buffers), the shapes
    // they were captured with, and the state before the running step
    GraphShape gshape;
    bool gshape_valid = false;
    hipGraphExec_t gexec[2] = {};
    DevBuf gsnap;

    int fail(int code, const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

#define HIPCHK(h, expr)                                                                                     \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            return (h)->fail(e_ == hipErrorOutOfMemory ? NBH_E_NOMEM : NBH_E_HIP, "%s failed: %s (%s:%d)", \
                             #expr, hipGetErrorString(e_), __FILE__, __LINE__);                             \
    } while (0)

#define CHK(expr)                  \
    do {                           \
        int r_ = (expr);           \
        if (r_ != NBH_OK) return r_; \
    } while (0)

namespace {

int ncp_of(int p) { return (ncoef(p) + 3) & ~3; }

bool theta_rational(double theta, int* num, int* den) {
    for (int d = 1; d <= 64; ++d) {
        double nf = theta * d;
        int nn = (int)std::lround(nf);
        if (nn >= 1 && nn <= 64 && std::fabs(nf - nn) <= 1e-9 * std::max(1.0, nf)) {
            *num = nn;
            *den = d;
            return true;
        }
    }
    return false;
}

__global__ void zero_u32_kernel(uint32_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// A captured step processes levels 0..depth_cap: a deeper tree is flagged
// like one that outgrew its node arrays, so that every later kernel of the
// step stands down (those levels have no rows) and the host reruns it.
__global__ void graph_guard_kernel(Counters* C, uint32_t depth_cap) {
    if (threadIdx.x == 0 && C->depth > depth_cap) atomicOr(&C->overflow, (uint32_t)OVF_NODES);
}

void zero_u32(void* p, size_t n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)ceil_div((unsigned)std::min<size_t>(n, 0xffffffffu), 256)),
                       dim3(256), 0, st, (uint32_t*)p, n);
}

int validate_config(const nbh_config* c, std::string& why) {
    char b[256];
    if (!c) { why = "config is NULL"; return NBH_E_INVALID; }
    if (c->struct_size != sizeof(nbh_config)) {
        snprintf(b, sizeof b, "config struct_size %u != %zu (ABI mismatch)", c->struct_size, sizeof(nbh_config));
        why = b;
        return NBH_E_INVALID;
    }
    if (c->solver > NBH_SOLVER_DIRECT) { why = "unknown solver"; return NBH_E_INVALID; }
    if (!(c->dt > 0) || !std::isfinite(c->dt)) { why = "dt must be finite and > 0"; return NBH_E_INVALID; }
    if (!(c->eps >= 0) || !std::isfinite(c->eps)) { why = "eps must be finite and >= 0"; return NBH_E_INVALID; }
    if (!std::isfinite(c->G)) { why = "G must be finite"; return NBH_E_INVALID; }
    if (c->integrator > NBH_INT_KICK_DRIFT) { why = "unknown integrator"; return NBH_E_INVALID; }
    if (c->force_law > NBH_LAW_COMPAT_NAIVE) { why = "unknown force law"; return NBH_E_INVALID; }
    if (c->force_law == NBH_LAW_COMPAT_NAIVE && c->solver != NBH_SOLVER_DIRECT) {
        why = "the compat_naive force law needs the direct solver";
        return NBH_E_INVALID;
    }
    if (c->solver == NBH_SOLVER_FMM) {
        int num, den;
        if (!(c->theta > 0) || !theta_rational(c->theta, &num, &den)) {
            snprintf(b, sizeof b, "theta=%g must be > 0 and equal num/den with num, den <= 64", c->theta);
            why = b;
            return NBH_E_INVALID;
        }
        if (c->leaf_capacity < 1 || c->leaf_capacity > 2048) { why = "leaf_capacity must be in [1, 2048]"; return NBH_E_INVALID; }
        if (c->order < 1 || c->order > 6) { why = "order must be in [1, 6]"; return NBH_E_INVALID; }
        if (c->m2l_order != 0 && c->m2l_order != c->order && c->m2l_order != c->order + 1) {
            why = "m2l_order must be 0 (= order + 1), order or order + 1";
            return NBH_E_INVALID;
        }
        if (c->m2l_terms > NBH_M2L_TERMS_LOW) { why = "unknown m2l_terms"; return NBH_E_INVALID; }
        if (c->p2p_chunk != 0 && c->p2p_chunk != 64 && c->p2p_chunk != 128) {
            why = "p2p_chunk must be 0 (= 64), 64 or 128";
            return NBH_E_INVALID;
        }
        if (c->max_depth > 21) { why = "max_depth must be <= 21"; return NBH_E_INVALID; }
        if (c->m2l > NBH_M2L_MFMA_F32) { why = "unknown m2l kernel"; return NBH_E_INVALID; }
        if ((c->m2l == NBH_M2L_BF16X3 || c->m2l == NBH_M2L_MFMA_F32) && c->order < 2) {
            why = "the matrix-core M2L needs order >= 2";
            return NBH_E_INVALID;
        }
        if (c->p2p > NBH_P2P_LANE) { why = "unknown p2p kernel"; return NBH_E_INVALID; }
    }
    if (c->fixed_root && !(c->root_size > 0)) { why = "fixed_root needs root_size > 0"; return NBH_E_INVALID; }
    return NBH_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
static int alloc_nodes(nbh_handle* h, uint32_t need) {
    if (need <= h->node_cap) return NBH_OK;
    uint32_t cap = std::max<uint32_t>(need + need / 4, 1024);
    const int ncp = ncp_of((int)h->cfg.order);
    HIPCHK(h, h->geo.ensure((size_t)cap * 16));
    HIPCHK(h, h->link.ensure((size_t)cap * 16));
    HIPCHK(h, h->nsys_of.ensure((size_t)cap * 4));
    HIPCHK(h, h->node_begin.ensure((size_t)cap * 4));
    HIPCHK(h, h->M.ensure((size_t)cap * ncp * 4));
    HIPCHK(h, h->L.ensure((size_t)cap * ncp * 4));
    HIPCHK(h, h->Lx.ensure((size_t)cap * ncp * 4));
    HIPCHK(h, h->row_first.ensure((size_t)cap * 4));
    HIPCHK(h, h->scan_tmp.ensure(scan_tmp_bytes(cap + 1) + 256));
    h->node_cap = cap;
    return NBH_OK;
}

static int read_counters(nbh_handle* h) {
    HIPCHK(h, hipMemcpyAsync(h->hc, h->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NBH_OK;
}

static NodeView node_view(nbh_handle* h) {
    return NodeView{h->geo.as<int4>(), h->link.as<int4>(), h->nsys_of.as<int>()};
}

static void record(nbh_handle* h, int k) {
    if (h->timing) (void)hipEventRecord(h->ev[k], h->stream);
}
static void record_side(nbh_handle* h, int k) {
    if (h->timing) (void)hipEventRecord(h->evb[k], h->stream2);
}

// bbox -> keys -> sort -> gather: the state ends in Morton order
static int sort_particles(nbh_handle* h) {
    const uint32_t n = (uint32_t)h->n;
    State a = h->st[h->cur].view(), b = h->st[1 - h->cur].view();
    Counters* dC = h->counters.as<Counters>();
    launch_bbox(a.x, a.y, a.z, h->sys_off.as<uint32_t>(), h->nsys, h->max_sys_n, h->bits, (int)h->cfg.fixed_root,
                h->cfg.root_lo, h->cfg.root_size, h->bbox_part.as<double>(), h->box.as<double4>(),
                h->inv.as<double>(), h->unit.as<float>(), &dC->nonfinite, h->stream);
    launch_keys(a.x, a.y, a.z, n, h->sys_off.as<uint32_t>(), h->nsys, h->bits, h->box.as<double4>(),
                h->inv.as<double>(), h->keys.as<uint64_t>(), h->perm.as<uint32_t>(),
                h->cfg.fixed_root ? &dC->nonfinite : nullptr, h->stream);
    int keybits = 3 * h->bits;
    if (h->nsys > 1) {
        int sb = 0;
        while ((1 << sb) < h->nsys) ++sb;
        keybits += sb;
    }
    radix_sort_pairs(h->keys.as<uint64_t>(), h->perm.as<uint32_t>(), h->keys_alt.as<uint64_t>(),
                     h->perm_alt.as<uint32_t>(), n, keybits, h->sort_tmp.p, h->stream);
    launch_gather_state(a, b, h->perm.as<uint32_t>(), n, h->stream);
    h->cur = 1 - h->cur;
    return NBH_OK;
}

// gs: a captured step (no read-back; the node arrays were sized beforehand
// and the emission flags a tree that outgrows them)
static int build_tree(nbh_handle* h, const GraphShape* gs = nullptr) {
    const uint32_t n = (uint32_t)h->n;
    Counters* dC = h->counters.as<Counters>();
    zero_u32(dC, sizeof(Counters) / 4, h->stream);
    CHK(sort_particles(h));
    const int D = (int)h->cfg.max_depth;
    launch_tree_build(h->keys.as<uint64_t>(), n, h->sys_off.as<uint32_t>(), h->nsys, h->bits,
                      (int)h->cfg.leaf_capacity, D, node_view(h), h->node_begin.as<uint32_t>(),
                      h->leaf_ids.as<uint32_t>(), h->node_cap, dC, h->tree_tmp.p, h->stream);
    if (gs) {
        launch_tree_emit(h->keys.as<uint64_t>(), n, h->sys_off.as<uint32_t>(), h->nsys, h->bits,
                         (int)h->cfg.leaf_capacity, D, node_view(h), h->node_begin.as<uint32_t>(),
                         h->leaf_ids.as<uint32_t>(), h->node_cap, dC, h->tree_tmp.p, h->stream);
        launch_tree_link(h->keys.as<uint64_t>(), n, h->sys_off.as<uint32_t>(), h->nsys, h->bits,
                         (int)h->cfg.leaf_capacity, D, dC, h->node_cap, h->node_begin.as<uint32_t>(), node_view(h),
                         h->stream);
        hipLaunchKernelGGL(graph_guard_kernel, dim3(1), dim3(64), 0, h->stream, dC, gs->depth_cap);
        return NBH_OK;
    }
    CHK(read_counters(h));
    if (h->hc->nonfinite) return h->fail(NBH_E_NONFINITE, "non-finite particle position");
    if (h->hc->outside)
        return h->fail(NBH_E_OUTSIDE, "a particle lies outside the fixed root box [%g, %g) x [%g, %g) x [%g, %g)",
                       h->cfg.root_lo[0], h->cfg.root_lo[0] + h->cfg.root_size, h->cfg.root_lo[1],
                       h->cfg.root_lo[1] + h->cfg.root_size, h->cfg.root_lo[2], h->cfg.root_lo[2] + h->cfg.root_size);
    h->num_nodes = h->hc->num_nodes;
    h->num_leaves = h->hc->num_leaves;
    h->depth = h->hc->depth;
    CHK(alloc_nodes(h, h->num_nodes + 1));
    launch_tree_emit(h->keys.as<uint64_t>(), n, h->sys_off.as<uint32_t>(), h->nsys, h->bits,
                     (int)h->cfg.leaf_capacity, D, node_view(h), h->node_begin.as<uint32_t>(),
                     h->leaf_ids.as<uint32_t>(), h->node_cap, dC, h->tree_tmp.p, h->stream);
    launch_tree_link(h->keys.as<uint64_t>(), n, h->sys_off.as<uint32_t>(), h->nsys, h->bits,
                     (int)h->cfg.leaf_capacity, D, dC, h->num_nodes, h->node_begin.as<uint32_t>(), node_view(h),
                     h->stream);
    return NBH_OK;
}

static FmmParams fmm_params(nbh_handle* h) {
    FmmParams F{};
    F.nodes = node_view(h);
    F.C = h->counters.as<Counters>();
    F.box = h->box.as<double4>();
    F.unit = h->unit.as<float>();
    F.bits = h->bits;
    F.eps2 = (float)(h->cfg.eps * h->cfg.eps);
    F.G = (float)h->cfg.G;
    F.M = h->M.as<float>();
    F.L = h->L.as<float>();
    F.ncp = ncp_of((int)h->cfg.order);
    F.leaf_ids = h->leaf_ids.as<uint32_t>();
    F.pf = h->pf.as<float4>();
    F.lists.first = h->row_first.as<uint32_t>();
    F.lists.next = h->row_next.as<uint32_t>();
    F.lists.m_start = h->row_m_start.as<uint32_t>();
    F.lists.m_cnt = h->row_m_cnt.as<uint32_t>();
    F.lists.mx_cnt = h->row_mx_cnt.as<uint32_t>();
    F.lists.p_start = h->row_p_start.as<uint32_t>();
    F.lists.p_cnt = h->row_p_cnt.as<uint32_t>();
    F.lists.m_pool = h->m2l_pool.as<int>();
    F.lists.p_pool = h->p2p_pool.as<int>();
    F.p2p_inter = &h->counters.as<Counters>()->p2p_inter;
    F.pleaf = h->pleaf.as<uint32_t>();
    F.dtab = nullptr;
    F.Lx = nullptr;
    F.eps2_64 = h->cfg.eps * h->cfg.eps;
    F.pot = h->eval_pot ? 1 : 0;
    return F;
}

static int grow_rows(nbh_handle* h, uint32_t cap) {
    h->row_cap = cap;
    DevBuf* rows[] = {&h->row_node, &h->row_next, &h->row_sl_start, &h->row_sl_cnt,
                      &h->row_m_start, &h->row_m_cnt, &h->row_mx_cnt, &h->row_p_start, &h->row_p_cnt};
    for (DevBuf* b : rows) HIPCHK(h, b->ensure((size_t)cap * 4));
    return NBH_OK;
}

static M2LPlan m2l_plan(const nbh_handle* h);
static FmmParams fmm_params(nbh_handle* h);

// Same-level M2L launched level by level during the walk: every same-level
// entry of a level-l target is emitted at depth l (carried rows of a leaf
// only meet finer candidates), so level l's kernel can start after depth l.
struct EarlyM2L {
    bool on = false;
    int p = 0, q = 0;
    FmmParams T{};  // the table fields
};

static void copy_tables(FmmParams& F, const FmmParams& T) {
    F.dtab = T.dtab;
    F.dK = T.dK;
    F.dlevel0 = T.dlevel0;
    F.dnlev = T.dnlev;
    F.dcp = T.dcp;
    F.dsz = T.dsz;
    F.doffs = T.doffs;
    F.dmap = T.dmap;
    F.xtab = T.xtab;
    F.xE = T.xE;
    F.xlevel0 = T.xlevel0;
    F.xnlev = T.xnlev;
}

static M2LPlan graph_plan(const nbh_handle* h, const GraphShape& gs);

// gs: a captured step: one attempt with the grids and buffers of gs, no
// read-back (an overflow is flagged on the device and found after the step)
// side: enqueues the side stream's work of the evaluation (moments and
// tables, ending with ev_side), called once, after the first walk depths are
// queued: the walk needs only the tree, so the main stream starts it while the
// host is still enqueueing the side stream's launches
static int run_walk(nbh_handle* h, const EarlyM2L* em = nullptr, const GraphShape* gs = nullptr,
                    const std::function<int()>* side = nullptr) {
    Counters* dC = h->counters.as<Counters>();
    const uint32_t nn = h->num_nodes;
    const uint32_t depth = gs ? gs->depth_cap : h->depth;
    if (!gs && h->row_cap < 2 * nn + 1024) CHK(grow_rows(h, 2 * nn + 1024));
    // group-kernel grid per depth (one workgroup per group at a time): at
    // most what stays resident, so that the per-workgroup pool chunks waste
    // little; every workgroup may leave one unused chunk tail per depth
    uint32_t grid[MAX_LEVELS + 1] = {};
    unsigned long long blocks = 0;
    for (uint32_t s = 1; s <= depth; ++s) {
        const uint32_t groups = gs ? 0 : h->hc->level_count[s - 1];
        grid[s] = gs ? gs->walk_grid[s] : std::max<uint32_t>(std::min<uint32_t>(groups + 16, h->walk_grid), 64);
        blocks += std::min<unsigned long long>(grid[s], nn);
    }
    const unsigned long long waste_m = blocks * POOL_CHUNK_M, waste_p = blocks * POOL_CHUNK_P;
    const unsigned long long waste_s = blocks * POOL_CHUNK_S;
    // pools: sized from the last evaluation (or a per-node guess) and grown
    // on overflow (a captured step's were sized beforehand)
    if (!gs && h->sl_cap < 64ull * nn + waste_s) {
        h->sl_cap = (uint32_t)std::min<unsigned long long>(64ull * nn + waste_s, 0xffff0000ull);
        HIPCHK(h, h->sl.ensure((size_t)h->sl_cap * 4));
    }
    const unsigned long long mguess = 256ull * nn + waste_m, pguess = 64ull * nn + waste_p;
    if (!gs && h->m2l_pool_cap < mguess) {
        h->m2l_pool_cap = std::min<unsigned long long>(mguess, 0xffff0000ull);
        HIPCHK(h, h->m2l_pool.ensure((size_t)h->m2l_pool_cap * 4));
    }
    if (!gs && h->p2p_pool_cap < pguess) {
        h->p2p_pool_cap = std::min<unsigned long long>(pguess, 0xffff0000ull);
        HIPCHK(h, h->p2p_pool.ensure((size_t)h->p2p_pool_cap * 4));
    }
    WalkParams W{};
    W.nodes = node_view(h);
    W.C = dC;
    W.bits = h->bits;
    W.num4 = 4u * h->theta_num * h->theta_num;
    W.den3 = 3u * h->theta_den * h->theta_den;
    bool side_done = false;
    // the first walk depths are queued before the side stream's launches
    const uint32_t side_after = std::min<uint32_t>(depth, 2);
    auto run_side = [&]() -> int {
        if (side_done) return NBH_OK;
        side_done = true;
        return side ? (*side)() : NBH_OK;
    };
    for (int attempt = 0; attempt < 12; ++attempt) {
        W.row_cap = h->row_cap;
        W.row_node = h->row_node.as<int>();
        W.row_first = h->row_first.as<uint32_t>();
        W.row_next = h->row_next.as<uint32_t>();
        W.row_sl_start = h->row_sl_start.as<uint32_t>();
        W.row_sl_cnt = h->row_sl_cnt.as<uint32_t>();
        W.row_m_start = h->row_m_start.as<uint32_t>();
        W.row_m_cnt = h->row_m_cnt.as<uint32_t>();
        W.row_mx_cnt = h->row_mx_cnt.as<uint32_t>();
        W.row_p_start = h->row_p_start.as<uint32_t>();
        W.row_p_cnt = h->row_p_cnt.as<uint32_t>();
        W.sl = h->sl.as<int>();
        W.sl_cap = h->sl_cap;
        W.m2l_pool = h->m2l_pool.as<int>();
        W.p2p_pool = h->p2p_pool.as<int>();
        W.m2l_pool_cap = h->m2l_pool_cap;
        W.p2p_pool_cap = h->p2p_pool_cap;
        const bool early = em && em->on;
        FmmParams Fe{};
        M2LPlan plan{};
        bool started = false;
        if (early) {
            Fe = fmm_params(h);  // this attempt's rows and pools
            copy_tables(Fe, em->T);
            plan = gs ? graph_plan(h, *gs) : m2l_plan(h);
        }
        // the same-level streams wait for the moments and tables (ev_side,
        // recorded by the side stream's work)
        auto fork_same = [&]() -> int {
            if (!early) return NBH_OK;
            HIPCHK(h, hipStreamWaitEvent(h->stream3, h->ev_side, 0));
            HIPCHK(h, hipStreamWaitEvent(h->stream4, h->ev_side, 0));
            return NBH_OK;
        };
        // a captured step has no host latency to hide: its side stream's work
        // is captured first (captured after the first walk depths, the graph
        // ran 5.6 -> 6.1 ms per step at N=1M)
        if (gs && !side_done) CHK(run_side());
        if (side_done) CHK(fork_same());
        // reset the walk part of the counters
        zero_u32(&dC->carried[0], (offsetof(Counters, nonfinite) - offsetof(Counters, carried)) / 4, h->stream);
        launch_walk_root(W, h->stream);
        if (!side_done && side_after == 0) {
            CHK(run_side());
            CHK(fork_same());
        }
        for (uint32_t s = 1; s <= depth; ++s) {
            launch_walk_group(W, (int)s, grid[s], h->stream);
            if (!side_done && s == side_after) {
                CHK(run_side());
                CHK(fork_same());
            }
            if (early && (int)s >= Fe.dlevel0 && (int)s < Fe.dlevel0 + Fe.dnlev && (gs || h->hc->level_count[s])) {
                hipStream_t sl = (s & 1) ? h->stream4 : h->stream3;
                HIPCHK(h, hipEventRecord(h->ev_depth[s], h->stream));
                HIPCHK(h, hipStreamWaitEvent(sl, h->ev_depth[s], 0));
                if (!started && h->timing) (void)hipEventRecord(h->evc0, sl);
                started = true;
                if (h->timing) (void)hipEventRecord(h->ev_lvl[2 * s], sl);
                launch_m2l_parts(em->p, em->q, Fe, plan, M2L_LEVELS, (int)s, sl);
                if (h->timing) {
                    (void)hipEventRecord(h->ev_lvl[2 * s + 1], sl);
                    h->lvl_timed[s] = true;
                }
            }
        }
        if (gs) return NBH_OK;
        CHK(read_counters(h));
        const uint32_t ovf = h->hc->overflow;
        const uint32_t walk_bits = OVF_SL | OVF_ROWS | OVF_M2L | OVF_P2P;
        if (ovf & walk_bits)
            for (auto& t : h->lvl_timed) t = false;  // the levels are launched again
        if ((ovf & walk_bits) && early) {
            // this attempt's same-level kernels may still be reading the rows
            // and pools about to be re-allocated: drain their streams first
            HIPCHK(h, hipStreamSynchronize(h->stream3));
            HIPCHK(h, hipStreamSynchronize(h->stream4));
        }
        if (ovf & OVF_SL) {
            h->sl_cap = (uint32_t)std::min<unsigned long long>(
                std::max<unsigned long long>(2ull * h->sl_cap, h->hc->sl_used + h->hc->sl_used / 4 + waste_s),
                0xffff0000ull);
            HIPCHK(h, h->sl.ensure((size_t)h->sl_cap * 4));
        }
        if (ovf & OVF_ROWS) {
            uint64_t need = 0;
            for (uint32_t s = 0; s <= h->depth + 1; ++s) need += h->hc->level_count[s] + h->hc->carried[s];
            CHK(grow_rows(h, (uint32_t)std::max<uint64_t>((uint64_t)h->row_cap * 2, need + need / 4)));
        }
        if (ovf & OVF_M2L) {
            h->m2l_pool_cap = std::min<unsigned long long>(
                std::max<unsigned long long>(h->m2l_pool_cap * 2, h->hc->m2l_pool_used + waste_m), 0xffff0000ull);
            HIPCHK(h, h->m2l_pool.ensure((size_t)h->m2l_pool_cap * 4));
        }
        if (ovf & OVF_P2P) {
            h->p2p_pool_cap = std::min<unsigned long long>(
                std::max<unsigned long long>(h->p2p_pool_cap * 2, h->hc->p2p_pool_used + waste_p), 0xffff0000ull);
            HIPCHK(h, h->p2p_pool.ensure((size_t)h->p2p_pool_cap * 4));
        }
        if (ovf & walk_bits) {
            uint32_t keep = ovf & ~walk_bits;
            HIPCHK(h, hipMemcpyAsync(&dC->overflow, &keep, 4, hipMemcpyHostToDevice, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            continue;
        }
        break;
    }
    if (h->hc->overflow) return h->fail(NBH_E_INTERNAL, "walk buffers overflow (flags %u)", h->hc->overflow);
    return NBH_OK;
}

// M2L truncation code (expansion.h): q = m2l_order (0: p + 1), with M2L_ALL
// for every term of degree <= q where that differs from the low-degree set
static int m2l_q(const nbh_handle* h) {
    const int p = (int)h->cfg.order;
    const int q = h->cfg.m2l_order ? (int)h->cfg.m2l_order : p + 1;
    return (h->cfg.m2l_terms == NBH_M2L_TERMS_ALL && p < q && q < 2 * p) ? (q | M2L_ALL) : q;
}

// P2P tile: sources staged per wavefront round
static int p2p_chunk(const nbh_handle* h) { return h->cfg.p2p_chunk == 128 ? 128 : 64; }

// the same-level pairs go to the matrix cores (one system; the window and
// the offset encoding bound the opening ratio)
static bool use_offset_m2l(const nbh_handle* h) {
    return (h->cfg.m2l == NBH_M2L_BF16X3 || h->cfg.m2l == NBH_M2L_MFMA_F32) && h->nsys == 1 && h->ncand > 0;
}

static M2LPlan m2l_plan(const nbh_handle* h) {
    M2LPlan pl{};
    pl.mode = h->cfg.m2l == NBH_M2L_PLAIN ? 0 : 1;
    pl.mfma = use_offset_m2l(h) ? (h->cfg.m2l == NBH_M2L_BF16X3 ? 1 : 2) : 0;
    if (pl.mfma) {
        pl.off.mb = h->mb.p;
        pl.off.zero_node = h->num_nodes;
        pl.off.tab = h->d2tab.as<uint32_t>();
        pl.off.K = h->offK;
        pl.off.E = h->offK + 1;
        pl.off.W = 4 + 2 * h->offK;
        pl.off.cand = h->cand.as<uint32_t>();
        pl.off.ncand = h->ncand;
        pl.off.flags = &h->counters.as<Counters>()->overflow;
        pl.off.hank = h->hank.as<uint16_t>();
        pl.off.par = h->opar.as<uint8_t>();
    }
    pl.num_nodes = h->num_nodes;
    pl.max_lds = (size_t)h->max_lds;
    pl.nsys = h->nsys;
    std::copy(h->hc->level_off, h->hc->level_off + MAX_LEVELS + 2, pl.level_off);
    std::copy(h->hc->level_count, h->hc->level_count + MAX_LEVELS + 1, pl.level_count);
    return pl;
}

// One force evaluation; kick_dt != 0 applies v += a * kick_dt afterwards.
// for_vjp: only the tree, lists and mass expansions are needed (the adjoint
// kernels follow on the same stream): no near field, no closing read-back.
static void graph_invalidate(nbh_handle* h);
static void graph_shape_update(nbh_handle* h);

// gs: a captured step (graph mode): the launches of the evaluation with the
// counts read on the device, the grids, levels and buffers of gs, and no
// read-back; the host reads the counters after the graph.
static int force_fmm(nbh_handle* h, double kick_dt, bool for_vjp = false, const GraphShape* gs = nullptr,
                     bool pot = true) {
    const int p = (int)h->cfg.order;
    h->eval_pot = pot;
    h->pot_valid = pot;
    // an evaluation outside a graph may grow (move) the buffers a graph holds
    if (!gs) graph_invalidate(h);
    record(h, 0);
    CHK(build_tree(h, gs));
    record(h, 1);
    FmmParams F = fmm_params(h);
    State S = h->st[h->cur].view();
    const uint32_t depth = gs ? gs->depth_cap : h->depth;  // levels 0..depth
    // side stream, overlapping the walk: moments (P2M, M2M) and the tables.
    // The host part (table parameters, buffers) comes first; the launches are
    // in `side`, which run_walk enqueues once the walk's first depths are
    // queued (the walk needs only the tree)
    HIPCHK(h, hipEventRecord(h->ev_fork, h->stream));
    const int q = m2l_q(h);
    const bool tables = h->cfg.m2l != NBH_M2L_PLAIN && depth >= 2;
    FmmParams T{};  // the table part of the parameters
    int K = 0, nlev = 0, E = 0, xn = 0;
    const bool off = tables && use_offset_m2l(h), bf3 = h->cfg.m2l == NBH_M2L_BF16X3;
    if (tables) {
        // derivative tensors of the same-level offsets [0, K]^3 of levels
        // 2..depth (cells of levels 0 and 1 are never well separated), per
        // system, since the tensors depend on each system's root size
        const int Kp = (int)std::floor(std::sqrt(3.0) * h->theta_den / h->theta_num);
        K = 2 * Kp + 1;
        const int dsz = h->sl_rows, dcp = ncp_of(m2l_deg(q));
        nlev = (int)depth - 1;
        HIPCHK(h, h->dtab.ensure((size_t)dsz * nlev * dcp * 4 * h->nsys));
        T = F;
        T.dtab = h->dtab.as<float>();
        T.dK = K;
        T.dlevel0 = 2;
        T.dnlev = nlev;
        T.dcp = dcp;
        T.dsz = dsz;
        T.doffs = h->sl_offs.as<uint32_t>();
        T.dmap = h->sl_map.as<uint16_t>();
        // pairs one level apart, by finer level 1..depth: offsets |d_i| <=
        // 4 Kp + 1 half finer cells (the coarser side's parent pair was
        // split), tabled at d_i = 1, 3, .., 4 Kp + 1
        E = 2 * Kp + 1;
        xn = (int)depth;
        HIPCHK(h, h->xtab.ensure((size_t)E * E * E * xn * dcp * 4 * h->nsys));
        T.xtab = h->xtab.as<float>();
        T.xE = E;
        T.xlevel0 = 1;
        T.xnlev = xn;
        if (off) {
            // the matrix-core same-level M2L: signed split moments and the
            // dense tensors (orders <= 2p) of the offsets [0, K]^3, levels 2..depth
            HIPCHK(h, h->mb.ensure(off_moment_bytes(p, bf3, h->num_nodes)));
            const size_t cube = (size_t)(K + 1) * (K + 1) * (K + 1);  // every offset of [0, K]^3
            HIPCHK(h, h->d2tab.ensure(cube * nlev * off_table_words(p) * 4));
        }
    }
    const std::function<int()> side = [&]() -> int {
        HIPCHK(h, hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
        record_side(h, 0);
        if (gs) {
            launch_p2m(p, F, S, DEV_COUNT, h->stream2, gs->leaf_items);
            for (int l = (int)depth - 1; l >= 0; --l)
                launch_m2m_level(p, F, l, DEV_COUNT, 0, h->stream2, gs->level_items[l]);
        } else {
            launch_p2m(p, F, S, h->num_leaves, h->stream2);
            for (int l = (int)depth - 1; l >= 0; --l)
                launch_m2m_level(p, F, l, h->hc->level_count[l], h->hc->level_off[l], h->stream2);
        }
        record_side(h, 1);
        if (tables) {
            launch_dtable(m2l_deg(q), T, T.dtab, 2, nlev, K + 1, 0, 1, 0, (int)h->nsys, T.doffs, T.dsz, h->stream2);
            launch_dtable(m2l_deg(q), T, T.xtab, 1, xn, E, 1, 2, 1, (int)h->nsys, nullptr, 0, h->stream2);
            if (off) {
                launch_off_moments(p, bf3, F.M, F.ncp, h->num_nodes, h->geo.as<int4>(), h->box.as<double4>(),
                                   h->mb.p, h->stream2);
                launch_off_dtable(p, bf3, h->box.as<double4>(), h->cfg.eps * h->cfg.eps, 2, nlev, K + 1,
                                  h->d2tab.as<uint32_t>(), h->stream2);
            }
        }
        HIPCHK(h, hipEventRecord(h->ev_side, h->stream2));
        return NBH_OK;
    };
    EarlyM2L em;
    em.on = tables;
    em.p = p;
    em.q = q;
    em.T = T;
    CHK(run_walk(h, &em, gs, &side));
    record(h, 2);
    F = fmm_params(h);
    if (tables) {
        copy_tables(F, T);
        F.Lx = h->Lx.as<float>();
        // deferred pairs: at most every other-level entry (a captured step:
        // sized from the previous step, a shortfall flagged)
        const unsigned long long nx = gs ? gs->far_cap : std::max<unsigned long long>(h->hc->m2lx_total, 1);
        HIPCHK(h, h->far_src.ensure((size_t)nx * 4));
        HIPCHK(h, h->far_tgt.ensure((size_t)nx * 4));
        F.far_src = h->far_src.as<int>();
        F.far_tgt = h->far_tgt.as<int>();
        F.far_cap = (uint32_t)std::min<unsigned long long>(nx, 0xffffffffull);
        F.far_cnt = &h->counters.as<Counters>()->far_blocks;
    }
    // side stream, overlapping M2L and L2L: the near field (raw sums)
    if (!for_vjp) {
        HIPCHK(h, hipEventRecord(h->ev_walk, h->stream));
        HIPCHK(h, hipStreamWaitEvent(h->stream2, h->ev_walk, 0));
        record_side(h, 2);
        if (gs) launch_p2p_staged(p, p2p_chunk(h), F, S, DEV_COUNT, h->stream2, gs->leaf_items);
        else if (h->cfg.p2p == NBH_P2P_LANE) launch_p2p_staged(p, p2p_chunk(h), F, S, h->num_leaves, h->stream2);
        else launch_p2p(p, F, S, h->num_leaves, h->stream2);
        record_side(h, 3);
        HIPCHK(h, hipEventRecord(h->ev_near, h->stream2));
    }
    HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_side, 0));
    if (em.on) {
        // the same-level kernels run beside the walk; the other-level pairs
        // (into Lx) beside them; L2L waits for both
        launch_m2l_parts(p, q, F, gs ? graph_plan(h, *gs) : m2l_plan(h), M2L_REST, -1, h->stream);
        HIPCHK(h, hipEventRecord(h->ev_same, h->stream3));
        HIPCHK(h, hipEventRecord(h->ev_same4, h->stream4));
        HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_same, 0));
        HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_same4, 0));
    } else {
        launch_m2l(p, q, F, gs ? graph_plan(h, *gs) : m2l_plan(h), h->stream);
    }
    record(h, 3);
    for (uint32_t l = 1; l <= depth; ++l) {
        if (gs) launch_l2l_level(p, F, (int)l, DEV_COUNT, 0, h->stream, gs->level_items[l]);
        else launch_l2l_level(p, F, (int)l, h->hc->level_count[l], h->hc->level_off[l], h->stream);
    }
    record(h, 4);
    h->lastF = F;
    if (for_vjp) {
        HIPCHK(h, hipGetLastError());
        return NBH_OK;
    }
    // join: far field of every particle added to the near field, forces,
    // closing half kick
    HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_near, 0));
    record(h, 5);
    launch_l2p(p, F, S, (uint32_t)h->n, kick_dt, h->stream);
    record(h, 6);
    HIPCHK(h, hipGetLastError());
    if (gs) return NBH_OK;
    // stats (read at the end of the evaluation; one more sync)
    CHK(read_counters(h));
    if (h->hc->overflow) return h->fail(NBH_E_INTERNAL, "overflow flags %u", h->hc->overflow);
    h->stats.num_nodes = h->num_nodes;
    h->stats.num_leaves = h->num_leaves;
    h->stats.depth = h->depth;
    h->stats.m2l_pairs = h->hc->m2l_total;
    h->stats.p2p_pairs = h->hc->p2p_total;
    h->stats.p2p_interactions = h->hc->p2p_inter;
    h->stats.split_pairs = h->hc->split_total;
    h->stats.m2l_same_pairs = h->hc->m2l_total - h->hc->m2lx_total;
    graph_shape_update(h);
    h->stats.ms_m2l_same = 0.0;
    h->stats.m2l_same_launches = 0;
    for (uint32_t l = 0; l <= (uint32_t)MAX_LEVELS; ++l) {
        if (!h->lvl_timed[l]) continue;
        float m = 0;
        if (hipEventElapsedTime(&m, h->ev_lvl[2 * l], h->ev_lvl[2 * l + 1]) == hipSuccess) {
            h->stats.ms_m2l_same += m;
            h->stats.m2l_same_launches += 1;
        }
        h->lvl_timed[l] = false;
    }
    if (h->timing) {
        // phases on the side stream overlap the others: moments run beside
        // the walk, P2P beside M2L and L2L; the total is the main stream's
        float ms[6], mb[3];
        for (int k = 0; k < 6; ++k) (void)hipEventElapsedTime(&ms[k], h->ev[k], h->ev[k + 1]);
        for (int k = 0; k < 3; ++k) (void)hipEventElapsedTime(&mb[k], h->evb[k], h->evb[k + 1]);
        h->stats.ms_tree = ms[0];
        h->stats.ms_moments = mb[0];
        h->stats.ms_walk = ms[1];
        h->stats.ms_m2l = ms[2];
        if (tables) {  // from the first same-level launch, beside the walk
            float m = 0;
            if (hipEventElapsedTime(&m, h->evc0, h->ev[3]) == hipSuccess) h->stats.ms_m2l = m;
        }
        h->stats.ms_l2l = ms[3] + ms[5];  // L2L and the far-field join
        h->stats.ms_p2p = mb[2];
        float tot = 0;
        (void)hipEventElapsedTime(&tot, h->ev[0], h->ev[6]);
        h->stats.ms_total = tot;
    }
    return NBH_OK;
}

static int force_direct(nbh_handle* h, double kick_dt) {
    State S = h->st[h->cur].view();
    Counters* dC = h->counters.as<Counters>();
    zero_u32(dC, sizeof(Counters) / 4, h->stream);
    launch_bbox(S.x, S.y, S.z, h->sys_off.as<uint32_t>(), h->nsys, h->max_sys_n, h->bits, (int)h->cfg.fixed_root,
                h->cfg.root_lo, h->cfg.root_size, h->bbox_part.as<double>(), h->box.as<double4>(),
                h->inv.as<double>(), h->unit.as<float>(), &dC->nonfinite, h->stream);
    launch_direct_f32(S, h->max_sys_n, h->sys_off.as<uint32_t>(), h->nsys, h->box.as<double4>(), (float)h->cfg.G,
                      (float)(h->cfg.eps * h->cfg.eps), (int)h->cfg.force_law, kick_dt, h->stream);
    HIPCHK(h, hipGetLastError());
    CHK(read_counters(h));
    if (h->hc->nonfinite) return h->fail(NBH_E_NONFINITE, "non-finite particle position");
    return NBH_OK;
}

// ---- captured steps ---------------------------------------------------------
static M2LPlan graph_plan(const nbh_handle* h, const GraphShape& gs) {
    M2LPlan pl{};
    pl.mode = 1;
    pl.mfma = 0;
    pl.num_nodes = gs.node_items;
    pl.max_lds = (size_t)h->max_lds;
    pl.nsys = h->nsys;
    pl.dev = true;
    for (uint32_t l = 0; l <= gs.depth_cap; ++l) pl.level_count[l] = gs.level_items[l];
    return pl;
}

static void graph_invalidate(nbh_handle* h) {
    for (auto& g : h->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// the launch shapes of a step like the one just evaluated, with headroom
static void graph_shape_update(nbh_handle* h) {
    const Counters& c = *h->hc;
    GraphShape g;
    const uint32_t nn = c.num_nodes, nl = c.num_leaves;
    g.node_cap = nn + nn / 4 + 1024;
    g.node_items = nn + nn / 8 + 64;
    g.leaf_items = nl + nl / 8 + 64;
    g.depth_cap = std::min<uint32_t>(h->cfg.max_depth, c.depth + 2);
    unsigned long long blocks = 0;
    for (uint32_t l = 0; l <= g.depth_cap; ++l) {
        const uint32_t lc = l <= c.depth ? c.level_count[l] : 0;
        g.level_items[l] = lc + lc / 4 + 64;
        if (l >= 1) {
            const uint32_t groups = g.level_items[l - 1] + (l - 1 <= c.depth ? c.carried[l - 1] : 0);
            g.walk_grid[l] = std::max<uint32_t>(std::min<uint32_t>(groups + 16, h->walk_grid), 64);
            blocks += g.walk_grid[l];
        }
    }
    // every workgroup may leave one unused chunk tail per depth
    const unsigned long long cap32 = 0xffff0000ull;
    g.m2l_cap = std::min(std::max<unsigned long long>(h->m2l_pool_cap,
                                                      c.m2l_pool_used + c.m2l_pool_used / 4 + blocks * POOL_CHUNK_M),
                         cap32);
    g.p2p_cap = std::min(std::max<unsigned long long>(h->p2p_pool_cap,
                                                      c.p2p_pool_used + c.p2p_pool_used / 4 + blocks * POOL_CHUNK_P),
                         cap32);
    g.sl_cap = (uint32_t)std::min(std::max<unsigned long long>(h->sl_cap, c.sl_used + c.sl_used / 4 +
                                                                              blocks * POOL_CHUNK_S),
                                  cap32);
    g.row_cap = std::max<uint32_t>(h->row_cap, 2 * g.node_cap + 1024);
    g.far_cap = (uint32_t)std::min<unsigned long long>(c.m2lx_total + c.m2lx_total / 4 + 65536, 0xffffffffull);
    h->gshape = g;
    h->gshape_valid = true;
}

// captured steps apply to the KDK FMM step with the tabled VALU M2L and the
// per-lane P2P, timing off
static bool graph_ok(const nbh_handle* h) {
    return h->cfg.graph && h->cfg.solver == NBH_SOLVER_FMM && h->cfg.integrator == NBH_INT_KDK && h->acc_valid &&
           !h->timing && h->cfg.m2l == NBH_M2L_GROUPED && h->cfg.p2p == NBH_P2P_LANE && h->gshape_valid &&
           h->gshape.depth_cap >= 2 && !use_offset_m2l(h);
}

// the buffers of the captured shape, allocated before the capture so that no
// allocation happens inside it and the graph's pointers stay valid
static int graph_prepare(nbh_handle* h) {
    const GraphShape& g = h->gshape;
    CHK(alloc_nodes(h, g.node_cap));
    if (h->row_cap < g.row_cap) CHK(grow_rows(h, g.row_cap));
    if (h->sl_cap < g.sl_cap) {
        h->sl_cap = g.sl_cap;
        HIPCHK(h, h->sl.ensure((size_t)h->sl_cap * 4));
    }
    if (h->m2l_pool_cap < g.m2l_cap) {
        h->m2l_pool_cap = g.m2l_cap;
        HIPCHK(h, h->m2l_pool.ensure((size_t)h->m2l_pool_cap * 4));
    }
    if (h->p2p_pool_cap < g.p2p_cap) {
        h->p2p_pool_cap = g.p2p_cap;
        HIPCHK(h, h->p2p_pool.ensure((size_t)h->p2p_pool_cap * 4));
    }
    // the tables (as force_fmm sizes them) and the deferred pairs
    const int q = m2l_q(h);
    const int Kp = (int)std::floor(std::sqrt(3.0) * h->theta_den / h->theta_num);
    const int E = 2 * Kp + 1;
    const size_t dsz = (size_t)h->sl_rows, dcp = (size_t)ncp_of(m2l_deg(q));
    HIPCHK(h, h->dtab.ensure(dsz * (g.depth_cap - 1) * dcp * 4 * h->nsys));
    HIPCHK(h, h->xtab.ensure((size_t)E * E * E * g.depth_cap * dcp * 4 * h->nsys));
    HIPCHK(h, h->far_src.ensure((size_t)g.far_cap * 4));
    HIPCHK(h, h->far_tgt.ensure((size_t)g.far_cap * 4));
    HIPCHK(h, h->gsnap.ensure(h->n * 48));
    return NBH_OK;
}

// Positions and velocities of the current state buffer to (save) or from the
// graph's snapshot: a step changes only those in that buffer (the kick and
// drift; the tree build gathers into the other buffer, where the forces and
// the closing kick go), so they are all a rollback needs.
static int graph_snapshot(nbh_handle* h, bool save) {
    const size_t n = h->n;
    StateBufs& sb = h->st[h->cur];
    DevBuf* d[] = {&sb.x, &sb.y, &sb.z, &sb.vx, &sb.vy, &sb.vz};
    char* o = (char*)h->gsnap.p;
    for (DevBuf* b : d) {
        HIPCHK(h, save ? hipMemcpyAsync(o, b->p, n * 8, hipMemcpyDeviceToDevice, h->stream)
                       : hipMemcpyAsync(b->p, o, n * 8, hipMemcpyDeviceToDevice, h->stream));
        o += n * 8;
    }
    return NBH_OK;
}

// Records one KDK step (state snapshot, opening kick and drift, the force
// evaluation with its closing kick) for the current state buffer into
// gexec[h->cur].  Nothing runs; the host state is left as it was.
static int graph_capture(nbh_handle* h) {
    CHK(graph_prepare(h));
    const int cur0 = h->cur;
    const double dt = h->cfg.dt;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = graph_snapshot(h, true);
    if (rc == NBH_OK) {
        launch_integrate(h->st[h->cur].view(), (uint32_t)h->n, 0, 0.5 * dt, dt, h->stream);
        rc = force_fmm(h, 0.5 * dt, false, &h->gshape, false);
    }
    h->cur = cur0;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc != NBH_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != NBH_OK ? rc : h->fail(NBH_E_HIP, "step capture failed: %s", hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&h->gexec[cur0], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        h->gexec[cur0] = nullptr;
        return h->fail(NBH_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
    }
    ++h->stats.graph_captures;
    return NBH_OK;
}

// One KDK step from the captured graph of this state buffer (captured first
// if needed): 1 when done, 0 when the step must run uncaptured (not
// applicable, or it outgrew the captured shapes: then it was rolled back),
// < 0 on error.
static int graph_step(nbh_handle* h) {
    if (!graph_ok(h)) return 0;
    if (!h->gexec[h->cur]) CHK(graph_capture(h));
    const int cur0 = h->cur;
    HIPCHK(h, hipGraphLaunch(h->gexec[cur0], h->stream));
    CHK(read_counters(h));
    const Counters& c = *h->hc;
    if (c.nonfinite) return h->fail(NBH_E_NONFINITE, "non-finite particle position");
    if (c.outside) return h->fail(NBH_E_OUTSIDE, "a particle lies outside the fixed root box");
    if ((c.overflow & (OVF_WALK | OVF_FAR)) || c.depth > h->gshape.depth_cap) {
        // outgrew the captured shapes: back to the state before the step,
        // which the caller runs uncaptured (and the next step captures anew)
        h->cur = cur0;
        CHK(graph_snapshot(h, false));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        graph_invalidate(h);
        h->gshape_valid = false;
        ++h->stats.graph_fallbacks;
        return 0;
    }
    if (c.overflow) return h->fail(NBH_E_INTERNAL, "overflow flags %u", c.overflow);
    h->cur = 1 - cur0;
    h->num_nodes = c.num_nodes;
    h->num_leaves = c.num_leaves;
    h->depth = c.depth;
    h->stats.num_nodes = c.num_nodes;
    h->stats.num_leaves = c.num_leaves;
    h->stats.depth = c.depth;
    h->stats.m2l_pairs = c.m2l_total;
    h->stats.p2p_pairs = c.p2p_total;
    h->stats.p2p_interactions = c.p2p_inter;
    h->stats.split_pairs = c.split_total;
    h->stats.m2l_same_pairs = c.m2l_total - c.m2lx_total;
    ++h->stats.graph_replays;
    h->pot_valid = false;  // a captured step evaluates accelerations only
    return 1;
}

// pot: the potentials too (a step needs the accelerations only; the direct
// sum computes both always)
static int force_eval(nbh_handle* h, double kick_dt, bool pot = true) {
    if (h->cfg.solver == NBH_SOLVER_FMM) return force_fmm(h, kick_dt, false, nullptr, pot);
    h->pot_valid = true;
    return force_direct(h, kick_dt);
}

static int do_step(nbh_handle* h) {
    const double dt = h->cfg.dt;
    const uint32_t n = (uint32_t)h->n;
    switch (h->cfg.integrator) {
        case NBH_INT_KDK: {
            if (!h->acc_valid) { CHK(force_eval(h, 0.0, false)); h->acc_valid = true; }
            if (h->cfg.graph) {
                const int r = graph_step(h);
                if (r < 0) return r;
                if (r == 1) break;
            }
            launch_integrate(h->st[h->cur].view(), n, 0, 0.5 * dt, dt, h->stream);
            CHK(force_eval(h, 0.5 * dt, false));
            break;
        }
        case NBH_INT_EULER:
            CHK(force_eval(h, 0.0, false));
            launch_integrate(h->st[h->cur].view(), n, 1, dt, dt, h->stream);
            h->acc_valid = false;
            h->pot_valid = false;  // the positions moved after the evaluation
            break;
        default:
            CHK(force_eval(h, 0.0, false));
            launch_integrate(h->st[h->cur].view(), n, 2, dt, dt, h->stream);
            h->acc_valid = false;
            h->pot_valid = false;
            break;
    }
    h->time += dt;
    return NBH_OK;
}

// ---------------------------------------------------------------------------
extern "C" {

int nbh_abi_version(void) { return NBH_ABI_VERSION; }
size_t nbh_config_size(void) { return sizeof(nbh_config); }

int nbh_config_default(nbh_config* c) {
    if (!c) return NBH_E_INVALID;
    std::memset(c, 0, sizeof *c);
    c->struct_size = sizeof(nbh_config);
    c->solver = NBH_SOLVER_FMM;
    c->G = 1.0;              // FORCE_CONSTANT = -1 (field.cl:7-9): attractive, unit strength
    c->eps = 0.01;           // PARTICLE_RADIUS (field.cl:3-5)
    c->theta = 0.5;          // NODE_APPROX_RATIO (interaction.cl:3-5)
    c->dt = 1e-3;
    c->leaf_capacity = 8;    // open_cl_simulation.cpp:41-47
    c->order = 4;
    c->max_depth = 21;
    c->m2l = NBH_M2L_GROUPED;
    c->integrator = NBH_INT_KDK;
    c->force_law = NBH_LAW_SOFTENED;
    c->device = 0;
    c->p2p = NBH_P2P_LANE;
    c->m2l_terms = NBH_M2L_TERMS_ALL;  // q = p + 1 with every degree-(p + 1) term
    return NBH_OK;
}

int nbh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* nbh_last_error(const nbh_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

static int create_impl(const nbh_config* cfg, size_t nsys, const uint32_t* sys_off, const double* pos,
                       const double* vel, const double* mass, const double* charge, nbh_handle** out) {
    if (!out) { g_create_error = "out is NULL"; return NBH_E_INVALID; }
    *out = nullptr;
    std::string why;
    int rc = validate_config(cfg, why);
    if (rc) { g_create_error = why; return rc; }
    if (nsys < 1 || !sys_off) { g_create_error = "need at least one system"; return NBH_E_INVALID; }
    const size_t n = sys_off[nsys];
    if (sys_off[0] != 0 || n < 1 || n > 0x7fffffffu) { g_create_error = "bad particle count"; return NBH_E_INVALID; }
    for (size_t s = 0; s < nsys; ++s)
        if (sys_off[s + 1] <= sys_off[s]) { g_create_error = "every system needs >= 1 particle"; return NBH_E_INVALID; }
    if (!pos || !vel || !mass) { g_create_error = "pos, vel and mass are required"; return NBH_E_INVALID; }
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(pos[3 * i + k]) || !std::isfinite(vel[3 * i + k])) {
                g_create_error = "non-finite position or velocity";
                return NBH_E_NONFINITE;
            }
        if (!(mass[i] > 0) || !std::isfinite(mass[i])) { g_create_error = "masses must be finite and > 0"; return NBH_E_INVALID; }
        if (charge && !std::isfinite(charge[i])) { g_create_error = "non-finite charge"; return NBH_E_NONFINITE; }
    }
    if (cfg->fixed_root) {
        // the reference's octree spans [0, bounds) (open_cl_simulation.cpp:20,43-47); the
        // closed upper face is accepted (main.cpp:118-122 may produce it)
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                const double t = pos[3 * i + k] - cfg->root_lo[k];
                if (!(t >= 0.0 && t <= cfg->root_size)) {
                    char b[160];
                    snprintf(b, sizeof b, "particle %zu lies outside the fixed root box (axis %d: %g not in [%g, %g])",
                             i, k, pos[3 * i + k], cfg->root_lo[k], cfg->root_lo[k] + cfg->root_size);
                    g_create_error = b;
                    return NBH_E_OUTSIDE;
                }
            }
    }
    bool gravity = true;
    if (charge)
        for (size_t i = 0; i < n; ++i) gravity = gravity && charge[i] == mass[i];
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { g_create_error = "no HIP device available"; return NBH_E_HIP; }
    if (cfg->device < 0 || cfg->device >= ndev) { g_create_error = "device ordinal out of range"; return NBH_E_INVALID; }

    nbh_handle* h = new nbh_handle;
    h->cfg = *cfg;
    h->gravity = gravity;
    h->n = n;
    h->nsys = (int)nsys;
    h->device = cfg->device;
    h->sys_off_h.assign(sys_off, sys_off + nsys + 1);
    for (size_t s = 0; s < nsys; ++s) h->max_sys_n = std::max(h->max_sys_n, sys_off[s + 1] - sys_off[s]);
    if (nsys > 1) {
        int sb = 0;
        while ((size_t(1) << sb) < nsys) ++sb;
        h->bits = std::min(21, (64 - sb) / 3);
    }
    if (h->cfg.max_depth > (uint32_t)h->bits) h->cfg.max_depth = (uint32_t)h->bits;
    if (cfg->solver == NBH_SOLVER_FMM) theta_rational(cfg->theta, &h->theta_num, &h->theta_den);

    auto bail = [&](int code) {
        g_create_error = h->err;
        nbh_destroy(h);
        return code;
    };
#define CCHK(expr)                                                                                          \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            h->fail(e_ == hipErrorOutOfMemory ? NBH_E_NOMEM : NBH_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
            return bail(e_ == hipErrorOutOfMemory ? NBH_E_NOMEM : NBH_E_HIP);                               \
        }                                                                                                   \
    } while (0)
    CCHK(hipSetDevice(h->device));
    CCHK(hipDeviceGetAttribute(&h->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    {
        int ncu = 0;
        CCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device));
        h->walk_grid = walk_group_resident_blocks(ncu);
    }
    CCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    CCHK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    CCHK(hipStreamCreateWithFlags(&h->stream3, hipStreamNonBlocking));
    CCHK(hipStreamCreateWithFlags(&h->stream4, hipStreamNonBlocking));
    h->own2 = h->stream2;
    h->own3 = h->stream3;
    h->own4 = h->stream4;
    CCHK(hipEventCreateWithFlags(&h->ev_same4, hipEventDisableTiming));
    CCHK(hipEventCreate(&h->evc0));
    CCHK(hipEventCreateWithFlags(&h->ev_same, hipEventDisableTiming));
    for (auto& e : h->ev_depth) CCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : h->ev_lvl) CCHK(hipEventCreate(&e));
    for (auto& e : h->ev) CCHK(hipEventCreate(&e));
    for (auto& e : h->evb) CCHK(hipEventCreate(&e));
    for (hipEvent_t* e : {&h->ev_fork, &h->ev_side, &h->ev_walk, &h->ev_near})
        CCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    CCHK(h->st[0].alloc(n));
    CCHK(h->st[1].alloc(n));
    CCHK(h->sys_off.ensure((nsys + 1) * 4));
    CCHK(h->keys.ensure(n * 8));
    CCHK(h->keys_alt.ensure(n * 8));
    CCHK(h->perm.ensure(n * 4));
    CCHK(h->perm_alt.ensure(n * 4));
    CCHK(h->sort_tmp.ensure(radix_tmp_bytes((uint32_t)n)));
    CCHK(h->tree_tmp.ensure(tree_tmp_bytes((uint32_t)n, (int)h->cfg.max_depth)));
    CCHK(h->box.ensure(nsys * sizeof(double4)));
    CCHK(h->inv.ensure(nsys * 8));
    CCHK(h->unit.ensure(nsys * 4));
    CCHK(h->bbox_part.ensure(nsys * (size_t)ceil_div(h->max_sys_n, 4096) * 6 * 8 + 64));
    CCHK(h->leaf_ids.ensure(n * 4));
    CCHK(h->pf.ensure(n * 16));
    CCHK(h->pleaf.ensure(n * 4));
    CCHK(h->counters.ensure(sizeof(Counters)));
    CCHK(h->diag.ensure(1024 * 8 * 8));
    CCHK(h->tmp_small.ensure(64));
    CCHK(hipHostMalloc((void**)&h->hc, sizeof(Counters), hipHostMallocDefault));
    std::memset(h->hc, 0, sizeof(Counters));
    if (cfg->solver == NBH_SOLVER_FMM) {
        int r = alloc_nodes(h, (uint32_t)std::min<size_t>(n / 2 + 64, 0x7fffffff));
        if (r) return bail(r);
        // candidate same-level offsets o of the matrix-core M2L, with the mask of
        // the target classes c (position in the parent) that take them: the
        // criterion accepts o and rejects the parent offset floor((c + o) / 2)
        const int num = h->theta_num, den = h->theta_den;
        const int K = 2 * (int)std::floor(std::sqrt(3.0) * den / num) + 1;
        {
            // the same-level tables of the vector M2L hold the offsets (|o_x|,
            // |o_y|, |o_z|) that occur (same criterion, every target class),
            // in lexicographic order, then a zero row that the other entries
            // of the (K+1)^3 map point to
            auto acc = [&](int x, int y, int z) {
                return (long long)num * num * (x * x + y * y + z * z) > 3ll * den * den;
            };
            auto half = [](int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); };
            const int K1 = K + 1;
            std::vector<uint8_t> occurs((size_t)K1 * K1 * K1, 0);
            for (int ox = -K; ox <= K; ++ox)
                for (int oy = -K; oy <= K; ++oy)
                    for (int oz = -K; oz <= K; ++oz) {
                        if (!acc(ox, oy, oz)) continue;
                        for (int c = 0; c < 8; ++c)
                            if (!acc(half(((c >> 2) & 1) + ox), half(((c >> 1) & 1) + oy), half((c & 1) + oz))) {
                                occurs[((size_t)std::abs(ox) * K1 + std::abs(oy)) * K1 + std::abs(oz)] = 1;
                                break;
                            }
                    }
            std::vector<uint32_t> offs;
            std::vector<uint16_t> map((size_t)K1 * K1 * K1);
            for (int x = 0; x < K1; ++x)
                for (int y = 0; y < K1; ++y)
                    for (int z = 0; z < K1; ++z)
                        if (occurs[((size_t)x * K1 + y) * K1 + z]) offs.push_back((uint32_t)x | (uint32_t)y << 8 | (uint32_t)z << 16);
            if (K1 > 256 || offs.size() >= 0xffff) {
                h->fail(NBH_E_INVALID, "theta too small for the same-level tables");
                return bail(NBH_E_INVALID);
            }
            const uint16_t zero_row = (uint16_t)offs.size();
            for (size_t i = 0, r = 0; i < map.size(); ++i) map[i] = occurs[i] ? (uint16_t)r++ : zero_row;
            offs.push_back(~0u);
            CCHK(h->sl_offs.ensure(offs.size() * 4));
            CCHK(hipMemcpy(h->sl_offs.p, offs.data(), offs.size() * 4, hipMemcpyHostToDevice));
            CCHK(h->sl_map.ensure(map.size() * 2));
            CCHK(hipMemcpy(h->sl_map.p, map.data(), map.size() * 2, hipMemcpyHostToDevice));
            h->sl_rows = (int)offs.size();
        }
        int dev_lds = 0;
        CCHK(hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
        std::vector<uint32_t> cand;
        if (K <= 15 && h->nsys == 1 && off_supported((int)cfg->order)) {
            auto accepts = [&](int x, int y, int z) {
                return (long long)num * num * (x * x + y * y + z * z) > 3ll * den * den;
            };
            auto fdiv2 = [](int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); };
            for (int ox = -K; ox <= K; ++ox)
                for (int oy = -K; oy <= K; ++oy)
                    for (int oz = -K; oz <= K; ++oz) {
                        if (!accepts(ox, oy, oz)) continue;
                        uint32_t mask = 0;
                        for (int c = 0; c < 8; ++c) {
                            const int px = fdiv2(((c >> 2) & 1) + ox), py = fdiv2(((c >> 1) & 1) + oy),
                                      pz = fdiv2((c & 1) + oz);
                            if (!accepts(px, py, pz)) mask |= 1u << c;
                        }
                        if (mask)
                            cand.push_back((uint32_t)(ox + 16) | (uint32_t)(oy + 16) << 5 | (uint32_t)(oz + 16) << 10 |
                                           mask << 16);
                    }
            if (off_lds_bytes((int)cfg->order, cfg->m2l == NBH_M2L_BF16X3, 4 + 2 * K, (int)cand.size()) >
                (size_t)dev_lds)
                cand.clear();  // the window does not fit: the same-level pairs stay on the vector units
        }
        if (!cand.empty()) {
            CCHK(h->cand.ensure(cand.size() * 4));
            CCHK(hipMemcpy(h->cand.p, cand.data(), cand.size() * 4, hipMemcpyHostToDevice));
            std::vector<uint16_t> hk;
            std::vector<uint8_t> pr;
            off_tables((int)cfg->order, cfg->m2l == NBH_M2L_BF16X3, hk, pr);
            CCHK(h->hank.ensure(hk.size() * 2));
            CCHK(hipMemcpy(h->hank.p, hk.data(), hk.size() * 2, hipMemcpyHostToDevice));
            CCHK(h->opar.ensure(pr.size()));
            CCHK(hipMemcpy(h->opar.p, pr.data(), pr.size(), hipMemcpyHostToDevice));
            h->ncand = (int)cand.size();
            h->offK = K;
        }
    }
    h->stats.m2l_same_order = cfg->solver != NBH_SOLVER_FMM ? 0
                              : use_offset_m2l(h) ? 2 * cfg->order
                                                  : (cfg->m2l_order ? cfg->m2l_order : cfg->order + 1);
    CCHK(hipMemcpy(h->sys_off.p, sys_off, (nsys + 1) * 4, hipMemcpyHostToDevice));
    {
        std::vector<double> buf(n);
        State S = h->st[0].view();
        double* dst[6] = {S.x, S.y, S.z, S.vx, S.vy, S.vz};
        for (int k = 0; k < 6; ++k) {
            const double* srcp = k < 3 ? pos : vel;
            for (size_t i = 0; i < n; ++i) buf[i] = srcp[3 * i + (k % 3)];
            CCHK(hipMemcpy(dst[k], buf.data(), n * 8, hipMemcpyHostToDevice));
        }
        CCHK(hipMemcpy(S.mass, mass, n * 8, hipMemcpyHostToDevice));
        if (charge) CCHK(hipMemcpy(S.charge, charge, n * 8, hipMemcpyHostToDevice));
        else CCHK(hipMemcpy(S.charge, mass, n * 8, hipMemcpyHostToDevice));
        std::vector<uint32_t> ids(n);
        for (size_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
        CCHK(hipMemcpy(S.id, ids.data(), n * 4, hipMemcpyHostToDevice));
        CCHK(hipMemset(S.ax, 0, n * 4));
        CCHK(hipMemset(S.ay, 0, n * 4));
        CCHK(hipMemset(S.az, 0, n * 4));
        CCHK(hipMemset(S.pot, 0, n * 4));
    }
#undef CCHK
    *out = h;
    return NBH_OK;
}

int nbh_create(const nbh_config* cfg, size_t n, const double* pos, const double* vel, const double* mass,
               const double* charge, nbh_handle** out) {
    if (n < 1 || n > 0x7fffffffu) {
        g_create_error = "n must be in [1, 2^31)";
        if (out) *out = nullptr;
        return NBH_E_INVALID;
    }
    uint32_t off[2] = {0, (uint32_t)n};
    return create_impl(cfg, 1, off, pos, vel, mass, charge, out);
}

int nbh_create_ensemble(const nbh_config* cfg, size_t nsys, const uint32_t* sys_off, const double* pos,
                        const double* vel, const double* mass, const double* charge, nbh_handle** out) {
    if (nsys > (1u << 20)) {
        g_create_error = "too many systems";
        if (out) *out = nullptr;
        return NBH_E_INVALID;
    }
    return create_impl(cfg, nsys, sys_off, pos, vel, mass, charge, out);
}

void nbh_destroy(nbh_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    graph_invalidate(h);
    h->gsnap.release();
    h->st[0].release();
    h->st[1].release();
    DevBuf* all[] = {&h->sys_off, &h->keys, &h->keys_alt, &h->perm, &h->perm_alt, &h->sort_tmp, &h->tree_tmp,
                     &h->scan_tmp, &h->box, &h->inv, &h->unit, &h->bbox_part, &h->geo, &h->link, &h->nsys_of,
                     &h->node_begin, &h->leaf_ids, &h->M, &h->L, &h->Lx, &h->pf, &h->pleaf, &h->dtab, &h->xtab, &h->far_src, &h->far_tgt, &h->row_node, &h->row_first,
                     &h->row_next, &h->row_sl_start, &h->row_sl_cnt, &h->row_m_start, &h->row_m_cnt,
                     &h->row_mx_cnt, &h->row_p_start, &h->row_p_cnt, &h->m2l_pool, &h->p2p_pool, &h->sl, &h->counters, &h->diag,
                     &h->tmp_small, &h->Mw, &h->Lw, &h->pw, &h->wbuf, &h->gbuf, &h->mb, &h->mbw, &h->d2tab,
                     &h->cand, &h->hank, &h->opar, &h->sl_offs, &h->sl_map, &h->acc_save};
    for (DevBuf* b : all) b->release();
    for (DevBuf& b : h->adj_ckpt) b.release();
    for (DevBuf& b : h->adj_seg) b.release();
    DevBuf* adj[] = {&h->adj_final, &h->adj_xb, &h->adj_vb, &h->adj_mb, &h->adj_w};
    for (DevBuf* b : adj) b->release();
    if (h->hc) (void)hipHostFree(h->hc);
    for (auto& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : h->evb)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {h->ev_fork, h->ev_side, h->ev_walk, h->ev_near, h->ev_same, h->ev_same4, h->evc0})
        if (e) (void)hipEventDestroy(e);
    for (auto& e : h->ev_depth)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : h->ev_lvl)
        if (e) (void)hipEventDestroy(e);
    if (h->own4) (void)hipStreamDestroy(h->own4);
    if (h->own3) (void)hipStreamDestroy(h->own3);
    if (h->own2) (void)hipStreamDestroy(h->own2);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int nbh_step(nbh_handle* h, int nsteps, double* t_out) {
    if (!h) return NBH_E_INVALID;
    if (nsteps < 0) return h->fail(NBH_E_INVALID, "nsteps must be >= 0");
    HIPCHK(h, hipSetDevice(h->device));
    for (int s = 0; s < nsteps; ++s) CHK(do_step(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (t_out) *t_out = h->time;
    return NBH_OK;
}

int nbh_get_state(nbh_handle* h, double* pos, double* vel, double* mass, double* charge, double* t) {
    if (!h) return NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t n = h->n;
    State S = h->st[h->cur].view();
    std::vector<uint32_t> ids(n);
    std::vector<double> buf(8 * n);
    HIPCHK(h, hipMemcpyAsync(ids.data(), S.id, n * 4, hipMemcpyDeviceToHost, h->stream));
    const double* src[8] = {S.x, S.y, S.z, S.vx, S.vy, S.vz, S.mass, S.charge};
    double* dsts[8] = {pos, pos, pos, vel, vel, vel, mass, charge};
    for (int k = 0; k < 8; ++k)
        if (dsts[k]) HIPCHK(h, hipMemcpyAsync(&buf[k * n], src[k], n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < 8; ++k) {
        double* dst = dsts[k];
        if (!dst) continue;
        const double* b = &buf[k * n];
        if (k < 6)
            for (size_t i = 0; i < n; ++i) dst[3 * (size_t)ids[i] + (k % 3)] = b[i];
        else
            for (size_t i = 0; i < n; ++i) dst[ids[i]] = b[i];
    }
    if (t) *t = h->time;
    return NBH_OK;
}

int nbh_set_state(nbh_handle* h, const double* pos, const double* vel, double t) {
    if (!h || !pos || !vel) return h ? h->fail(NBH_E_INVALID, "pos and vel are required") : NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t n = h->n;
    for (size_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(pos[i]) || !std::isfinite(vel[i])) return h->fail(NBH_E_NONFINITE, "non-finite state");
    State S = h->st[h->cur].view();
    // masses/charges keep their current (sorted) slots: bring them to input order first
    std::vector<uint32_t> ids(n);
    std::vector<double> m(n), q(n), mi(n), qi(n), buf(n);
    HIPCHK(h, hipMemcpyAsync(ids.data(), S.id, n * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(m.data(), S.mass, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(q.data(), S.charge, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; ++i) { mi[ids[i]] = m[i]; qi[ids[i]] = q[i]; }
    double* dst[6] = {S.x, S.y, S.z, S.vx, S.vy, S.vz};
    buf.resize(6 * n);
    for (int k = 0; k < 6; ++k) {
        const double* srcp = k < 3 ? pos : vel;
        double* b = &buf[k * n];
        for (size_t i = 0; i < n; ++i) b[i] = srcp[3 * i + (k % 3)];
        HIPCHK(h, hipMemcpyAsync(dst[k], b, n * 8, hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(h, hipMemcpyAsync(S.mass, mi.data(), n * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(S.charge, qi.data(), n * 8, hipMemcpyHostToDevice, h->stream));
    for (size_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
    HIPCHK(h, hipMemcpyAsync(S.id, ids.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->time = t;
    h->acc_valid = false;
    h->pot_valid = false;
    return NBH_OK;
}

int nbh_compute_forces(nbh_handle* h, float* acc, float* pot) {
    if (!h) return NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    CHK(force_eval(h, 0.0));
    h->acc_valid = true;
    const size_t n = h->n;
    State S = h->st[h->cur].view();
    std::vector<uint32_t> ids(n);
    std::vector<float> a(n);
    HIPCHK(h, hipMemcpyAsync(ids.data(), S.id, n * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (acc) {
        float* src[3] = {S.ax, S.ay, S.az};
        for (int k = 0; k < 3; ++k) {
            HIPCHK(h, hipMemcpyAsync(a.data(), src[k], n * 4, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            for (size_t i = 0; i < n; ++i) acc[3 * (size_t)ids[i] + k] = a[i];
        }
    }
    if (pot) {
        HIPCHK(h, hipMemcpyAsync(a.data(), S.pot, n * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (size_t i = 0; i < n; ++i) pot[ids[i]] = a[i];
    }
    return NBH_OK;
}

int nbh_autotune(nbh_handle* h, const uint32_t* caps, size_t ncaps, const uint32_t* chunks, size_t nchunks,
                 int evals, nbh_tune_result* out, double* ms_table) {
    if (!h) return NBH_E_INVALID;
    if (h->cfg.solver != NBH_SOLVER_FMM) return h->fail(NBH_E_INVALID, "nbh_autotune: the FMM solver only");
    if (evals < 1 || evals > 1000) return h->fail(NBH_E_INVALID, "nbh_autotune: evals must be in [1, 1000]");
    HIPCHK(h, hipSetDevice(h->device));
    const uint32_t cap0 = h->cfg.leaf_capacity, chunk0 = (uint32_t)p2p_chunk(h);
    const std::vector<uint32_t> cs = (caps && ncaps) ? std::vector<uint32_t>(caps, caps + ncaps)
                                                     : std::vector<uint32_t>{cap0};
    const std::vector<uint32_t> ks = (chunks && nchunks) ? std::vector<uint32_t>(chunks, chunks + nchunks)
                                                         : std::vector<uint32_t>{chunk0};
    for (uint32_t c : cs)
        if (c < 1 || c > 2048) return h->fail(NBH_E_INVALID, "nbh_autotune: leaf capacity %u not in [1, 2048]", c);
    for (uint32_t k : ks)
        if (k != 64 && k != 128) return h->fail(NBH_E_INVALID, "nbh_autotune: p2p chunk %u is not 64 or 128", k);
    double best = 1e300;
    uint32_t best_c = cap0, best_k = chunk0;
    for (size_t i = 0; i < cs.size(); ++i)
        for (size_t j = 0; j < ks.size(); ++j) {
            h->cfg.leaf_capacity = cs[i];
            h->cfg.p2p_chunk = ks[j];
            // one untimed evaluation (buffers grow to this tree), then the
            // timed ones; each evaluation ends with a host synchronisation
            int rc = force_eval(h, 0.0, false);
            const auto t0 = std::chrono::steady_clock::now();
            for (int e = 0; e < evals && rc == NBH_OK; ++e) rc = force_eval(h, 0.0, false);
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / evals;
            if (rc != NBH_OK) {
                h->cfg.leaf_capacity = cap0;
                h->cfg.p2p_chunk = chunk0;
                h->acc_valid = false;
                h->pot_valid = false;
                return rc;
            }
            if (ms_table) ms_table[i * ks.size() + j] = ms;
            if (ms < best) {
                best = ms;
                best_c = cs[i];
                best_k = ks[j];
            }
        }
    h->cfg.leaf_capacity = best_c;
    h->cfg.p2p_chunk = best_k;
    // the accelerations of the chosen configuration at this state
    CHK(force_eval(h, 0.0));
    h->acc_valid = true;
    if (out) {
        out->leaf_capacity = best_c;
        out->p2p_chunk = best_k;
        out->ms_per_eval = best;
        out->tried = (uint32_t)(cs.size() * ks.size());
    }
    return NBH_OK;
}

int nbh_direct_fp64(nbh_handle* h, size_t nt, const int64_t* targets, double* acc, double* pot) {
    if (!h) return NBH_E_INVALID;
    if (!targets || !acc || !pot) return h->fail(NBH_E_INVALID, "targets, acc and pot are required");
    HIPCHK(h, hipSetDevice(h->device));
    const size_t n = h->n;
    if (!nt) return NBH_OK;
    State S = h->st[h->cur].view();
    std::vector<uint32_t> ids(n), slot(n), tg(nt);
    HIPCHK(h, hipMemcpyAsync(ids.data(), S.id, n * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; ++i) slot[ids[i]] = (uint32_t)i;
    for (size_t k = 0; k < nt; ++k) {
        if (targets[k] < 0 || (size_t)targets[k] >= n) return h->fail(NBH_E_INVALID, "target index out of range");
        tg[k] = slot[targets[k]];
    }
    // the box is needed only by direct_f32; fp64 uses raw positions
    DevBuf dt, da, dp;
    HIPCHK(h, dt.ensure(nt * 4));
    HIPCHK(h, da.ensure(nt * 24));
    HIPCHK(h, dp.ensure(nt * 8));
    HIPCHK(h, hipMemcpyAsync(dt.p, tg.data(), nt * 4, hipMemcpyHostToDevice, h->stream));
    launch_direct_fp64(S, h->sys_off.as<uint32_t>(), h->nsys, dt.as<uint32_t>(), (uint32_t)nt, h->cfg.G,
                       h->cfg.eps * h->cfg.eps, da.as<double>(), dp.as<double>(), h->stream);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(acc, da.p, nt * 24, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(pot, dp.p, nt * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dt.release(); da.release(); dp.release();
    return NBH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Fixed device scratch of nbh_direct_fp64_noalloc: the input-order targets,
// their sorted slots and the FP64 results of up to NBH_NOALLOC_MAX_TARGETS
// targets, 44 bytes per target (2.75 MiB), held by every GPU the library runs
// on until the process ends.  Host code uses only the per-device address from
// hipGetSymbolAddress, never the array's name.
namespace {
constexpr size_t kNoallocTg = 0;                                                 // int64[max]
constexpr size_t kNoallocSlot = kNoallocTg + 8 * (size_t)NBH_NOALLOC_MAX_TARGETS;  // uint32[max]
constexpr size_t kNoallocAcc = kNoallocSlot + 4 * (size_t)NBH_NOALLOC_MAX_TARGETS; // double[3 max]
constexpr size_t kNoallocPot = kNoallocAcc + 24 * (size_t)NBH_NOALLOC_MAX_TARGETS; // double[max]
constexpr size_t kNoallocBytes = kNoallocPot + 8 * (size_t)NBH_NOALLOC_MAX_TARGETS;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) char scratch[kNoallocBytes];

// The device address of this GPU's scratch (the current device must be
// h->device), looked up once per device.
static int noalloc_scratch(nbh_handle* h, char** out) {
    if (h->device < 0 || h->device >= kMaxDevices) return h->fail(NBH_E_INVALID, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
        void* p = nullptr;
        HIPCHK(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return h->fail(NBH_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[h->device] = (char*)p;
    }
    *out = g_scratch_ptr[h->device];
    return NBH_OK;
}

extern "C" {

int nbh_direct_fp64_noalloc(nbh_handle* h, size_t nt, const int64_t* targets, double* acc, double* pot) {
    if (!h) return NBH_E_INVALID;
    if (!targets || !acc || !pot) return h->fail(NBH_E_INVALID, "targets, acc and pot are required");
    if (nt > (size_t)NBH_NOALLOC_MAX_TARGETS)
        return h->fail(NBH_E_INVALID, "nbh_direct_fp64_noalloc: %zu targets exceed its fixed scratch of %d targets",
                       nt, NBH_NOALLOC_MAX_TARGETS);
    HIPCHK(h, hipSetDevice(h->device));
    if (!nt) return NBH_OK;
    const size_t n = h->n;
    for (size_t k = 0; k < nt; ++k)
        if (targets[k] < 0 || (size_t)targets[k] >= n) return h->fail(NBH_E_INVALID, "target index out of range");
    char* base = nullptr;
    CHK(noalloc_scratch(h, &base));
    int64_t* dtg = (int64_t*)(base + kNoallocTg);
    uint32_t* dslot = (uint32_t*)(base + kNoallocSlot);
    double* dacc = (double*)(base + kNoallocAcc);
    double* dpot = (double*)(base + kNoallocPot);
    State S = h->st[h->cur].view();
    HIPCHK(h, hipMemcpyAsync(dtg, targets, nt * 8, hipMemcpyHostToDevice, h->stream));
    launch_find_slots(S.id, (uint32_t)n, dtg, (uint32_t)nt, dslot, h->stream);
    launch_direct_fp64(S, h->sys_off.as<uint32_t>(), h->nsys, dslot, (uint32_t)nt, h->cfg.G,
                       h->cfg.eps * h->cfg.eps, dacc, dpot, h->stream);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(acc, dacc, nt * 24, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(pot, dpot, nt * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NBH_OK;
}

int nbh_diagnostics(nbh_handle* h, double* out) {
    if (!h || !out) return h ? h->fail(NBH_E_INVALID, "out is NULL") : NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    if (!h->pot_valid) {
        // steps evaluate accelerations only: one evaluation with potentials
        // at the current positions, the accelerations the integrator holds
        // kept as they were (the evaluation rewrites them, in another
        // summation order)
        // (the positions are already in the order the evaluation sorts
        // them into, so the saved accelerations go back in the same order,
        // into whichever state buffer is current afterwards)
        const size_t n = h->n;
        HIPCHK(h, h->acc_save.ensure(n * 12));
        char* o = (char*)h->acc_save.p;
        {
            const StateBufs& sb = h->st[h->cur];
            const DevBuf* a[3] = {&sb.ax, &sb.ay, &sb.az};
            for (int k = 0; k < 3; ++k)
                HIPCHK(h, hipMemcpyAsync(o + k * n * 4, a[k]->p, n * 4, hipMemcpyDeviceToDevice, h->stream));
        }
        CHK(force_eval(h, 0.0, true));
        {
            const StateBufs& sb = h->st[h->cur];
            const DevBuf* a[3] = {&sb.ax, &sb.ay, &sb.az};
            for (int k = 0; k < 3; ++k)
                HIPCHK(h, hipMemcpyAsync(a[k]->p, o + k * n * 4, n * 4, hipMemcpyDeviceToDevice, h->stream));
        }
    }
    const uint32_t nb = diagnostics_blocks((uint32_t)h->n);
    launch_diagnostics(h->st[h->cur].view(), (uint32_t)h->n, h->diag.as<double>(), nb, h->stream);
    std::vector<double> part((size_t)nb * 8);
    HIPCHK(h, hipMemcpyAsync(part.data(), h->diag.p, part.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < 8; ++k) {
        double s = 0;
        for (uint32_t b = 0; b < nb; ++b) s += part[(size_t)b * 8 + k];
        out[k] = s;
    }
    return NBH_OK;
}

int nbh_get_stats(nbh_handle* h, nbh_stats* s) {
    if (!h || !s) return NBH_E_INVALID;
    *s = h->stats;
    return NBH_OK;
}

int nbh_set_timing(nbh_handle* h, int enable) {
    if (!h) return NBH_E_INVALID;
    if (enable < 0 || enable > 2) return h->fail(NBH_E_INVALID, "timing mode must be 0, 1 or 2");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());
    h->timing = enable != 0;
    const bool serial = enable == 2;
    h->stream2 = serial ? h->stream : h->own2;
    h->stream3 = serial ? h->stream : h->own3;
    h->stream4 = serial ? h->stream : h->own4;
    return NBH_OK;
}

}  // extern "C"

// VJP of the accelerations at the current state with the cotangent w_dev
// (device, input order, 3n): g = (da/dx)^T w, (da/dm)^T w into h->gbuf in the
// current (sorted) order, as four arrays of n doubles (gx, gy, gz, gm).
static int vjp_device(nbh_handle* h, const double* w_dev) {
    const size_t n = h->n;
    // tree, lists and mass expansions at the current positions
    CHK(force_fmm(h, 0.0, true));
    h->acc_valid = false;
    h->pot_valid = false;
    const int p = (int)h->cfg.order;
    const size_t ncp = (size_t)ncp_of(p);
    HIPCHK(h, h->Mw.ensure((size_t)h->num_nodes * ncp * 4));
    HIPCHK(h, h->Lw.ensure((size_t)h->num_nodes * ncp * 4));
    HIPCHK(h, h->pw.ensure(n * 16));
    HIPCHK(h, h->gbuf.ensure(n * 32));
    FmmParams F = h->lastF;
    FmmParams Fw = F;
    Fw.M = h->Mw.as<float>();
    Fw.L = h->Lw.as<float>();
    State S = h->st[h->cur].view();
    float4* pw = h->pw.as<float4>();
    double* gdev = h->gbuf.as<double>();
    launch_vjp_gather(w_dev, S.id, pw, (uint32_t)n, h->stream);
    // the near field on the side stream, beside the transposed far-field chain
    HIPCHK(h, hipEventRecord(h->ev_fork, h->stream));
    HIPCHK(h, hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
    launch_vjp_near(p, F, pw, (uint32_t)n, h->num_leaves, gdev, h->stream2);
    HIPCHK(h, hipEventRecord(h->ev_near, h->stream2));
    M2LPlan plan = m2l_plan(h);
    if (plan.mfma) {
        HIPCHK(h, h->mbw.ensure(off_moment_bytes(p, plan.mfma == 1, h->num_nodes)));
        plan.off.mb = h->mbw.p;  // the chain splits the dipole moments into it
    }
    launch_vjp_chain(p, m2l_q(h), plan, F, Fw, pw, h->num_leaves, h->hc->level_count, h->hc->level_off,
                     (int)h->depth, h->stream);
    HIPCHK(h, hipStreamWaitEvent(h->stream, h->ev_near, 0));
    launch_vjp_far(p, F, Fw, pw, (uint32_t)n, gdev, h->stream);
    HIPCHK(h, hipGetLastError());
    return NBH_OK;
}

static int vjp_checks(nbh_handle* h, const char* what) {
    if (h->cfg.solver != NBH_SOLVER_FMM) return h->fail(NBH_E_INVALID, "%s needs the FMM solver", what);
    if (!h->gravity) return h->fail(NBH_E_INVALID, "%s needs charge == mass (gravity)", what);
    return NBH_OK;
}

// ---- trajectory adjoint ---------------------------------------------------
// A state snapshot: x, y, z, vx, vy, vz, mass, charge (FP64), ax, ay, az, pot
// (FP32) and id, in the current order (84 bytes per particle).
static size_t snapshot_bytes(size_t n) { return n * (8 * 8 + 4 * 4 + 4); }

static int snapshot_save(nbh_handle* h, DevBuf& dst) {
    const size_t n = h->n;
    HIPCHK(h, dst.ensure(snapshot_bytes(n)));
    const StateBufs& sb = h->st[h->cur];
    const DevBuf* d8[] = {&sb.x, &sb.y, &sb.z, &sb.vx, &sb.vy, &sb.vz, &sb.mass, &sb.charge};
    const DevBuf* d4[] = {&sb.ax, &sb.ay, &sb.az, &sb.pot, &sb.id};
    char* o = (char*)dst.p;
    for (const DevBuf* b : d8) { HIPCHK(h, hipMemcpyAsync(o, b->p, n * 8, hipMemcpyDeviceToDevice, h->stream)); o += n * 8; }
    for (const DevBuf* b : d4) { HIPCHK(h, hipMemcpyAsync(o, b->p, n * 4, hipMemcpyDeviceToDevice, h->stream)); o += n * 4; }
    return NBH_OK;
}

static int snapshot_restore(nbh_handle* h, const DevBuf& src, double t) {
    const size_t n = h->n;
    StateBufs& sb = h->st[h->cur];
    DevBuf* d8[] = {&sb.x, &sb.y, &sb.z, &sb.vx, &sb.vy, &sb.vz, &sb.mass, &sb.charge};
    DevBuf* d4[] = {&sb.ax, &sb.ay, &sb.az, &sb.pot, &sb.id};
    const char* o = (const char*)src.p;
    for (DevBuf* b : d8) { HIPCHK(h, hipMemcpyAsync(b->p, o, n * 8, hipMemcpyDeviceToDevice, h->stream)); o += n * 8; }
    for (DevBuf* b : d4) { HIPCHK(h, hipMemcpyAsync(b->p, o, n * 4, hipMemcpyDeviceToDevice, h->stream)); o += n * 4; }
    h->time = t;
    h->acc_valid = true;  // the snapshot carries the accelerations of its positions
    h->pot_valid = false;  // (and the potentials only if that evaluation computed them)
    return NBH_OK;
}

extern "C" {

int nbh_force_vjp(nbh_handle* h, const double* w, double* gx, double* gm) {
    if (!h) return NBH_E_INVALID;
    if (!w || !gx || !gm) return h->fail(NBH_E_INVALID, "w, gx and gm are required");
    CHK(vjp_checks(h, "nbh_force_vjp"));
    HIPCHK(h, hipSetDevice(h->device));
    const size_t n = h->n;
    HIPCHK(h, h->wbuf.ensure(n * 24));
    HIPCHK(h, hipMemcpyAsync(h->wbuf.p, w, n * 24, hipMemcpyHostToDevice, h->stream));
    CHK(vjp_device(h, h->wbuf.as<double>()));
    State S = h->st[h->cur].view();
    std::vector<double> g(4 * n);
    std::vector<uint32_t> ids(n);
    HIPCHK(h, hipMemcpyAsync(g.data(), h->gbuf.p, 4 * n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(ids.data(), S.id, n * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < n; ++i) {
        const size_t j = ids[i];
        gx[3 * j] = g[i];
        gx[3 * j + 1] = g[n + i];
        gx[3 * j + 2] = g[2 * n + i];
        gm[j] = g[3 * n + i];
    }
    return NBH_OK;
}

int nbh_adjoint_forward(nbh_handle* h, int nsteps, int ckpt_every) {
    if (!h) return NBH_E_INVALID;
    CHK(vjp_checks(h, "nbh_adjoint_forward"));
    if (h->cfg.integrator != NBH_INT_KDK) return h->fail(NBH_E_INVALID, "nbh_adjoint_forward needs the KDK integrator");
    if (nsteps < 1 || ckpt_every < 1) return h->fail(NBH_E_INVALID, "nsteps and ckpt_every must be >= 1");
    HIPCHK(h, hipSetDevice(h->device));
    const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
    if ((int)h->adj_ckpt.size() < nck) h->adj_ckpt.resize(nck);
    h->adj_ckpt_t.assign(nck, 0.0);
    h->adj_ready = false;
    if (!h->acc_valid) { CHK(force_eval(h, 0.0, false)); h->acc_valid = true; }
    for (int s = 0; s < nsteps; ++s) {
        if (s % ckpt_every == 0) {
            CHK(snapshot_save(h, h->adj_ckpt[s / ckpt_every]));
            h->adj_ckpt_t[s / ckpt_every] = h->time;
        }
        CHK(do_step(h));
    }
    CHK(snapshot_save(h, h->adj_final));
    h->adj_final_t = h->time;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->adj_nsteps = nsteps;
    h->adj_k = ckpt_every;
    h->adj_ready = true;
    return NBH_OK;
}

int nbh_adjoint_backward(nbh_handle* h, const double* xbar, const double* vbar, double* gx, double* gv, double* gm) {
    if (!h) return NBH_E_INVALID;
    if (!xbar || !vbar || !gx || !gv || !gm) return h->fail(NBH_E_INVALID, "xbar, vbar, gx, gv and gm are required");
    if (!h->adj_ready) return h->fail(NBH_E_INVALID, "nbh_adjoint_backward needs a preceding nbh_adjoint_forward");
    HIPCHK(h, hipSetDevice(h->device));
    const size_t n = h->n;
    const double dt = h->cfg.dt;
    HIPCHK(h, h->adj_xb.ensure(n * 24));
    HIPCHK(h, h->adj_vb.ensure(n * 24));
    HIPCHK(h, h->adj_mb.ensure(n * 8));
    HIPCHK(h, h->adj_w.ensure(n * 24));
    double* xb = h->adj_xb.as<double>();
    double* vb = h->adj_vb.as<double>();
    double* mbar = h->adj_mb.as<double>();
    double* w = h->adj_w.as<double>();
    HIPCHK(h, hipMemcpyAsync(xb, xbar, n * 24, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(vb, vbar, n * 24, hipMemcpyHostToDevice, h->stream));
    launch_fill_f64(mbar, n, 0.0, h->stream);
    // the last state: second half kick of the last step, cotangent dt/2 vbar
    CHK(snapshot_restore(h, h->adj_final, h->adj_final_t));
    launch_adj_kick(xb, vb, w, 3 * n, dt, 0.5 * dt, false, h->stream);
    CHK(vjp_device(h, w));
    launch_adj_accum(h->gbuf.as<double>(), h->st[h->cur].id.as<uint32_t>(), (uint32_t)n, xb, mbar, h->stream);
    const int nseg = (int)h->adj_ckpt_t.size();
    if ((int)h->adj_seg.size() < h->adj_k) h->adj_seg.resize(h->adj_k);
    h->adj_seg_t.assign(h->adj_k, 0.0);
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int start = sg * h->adj_k, len = std::min(h->adj_k, h->adj_nsteps - start);
        // recompute the segment from its checkpoint, keeping each state
        CHK(snapshot_restore(h, h->adj_ckpt[sg], h->adj_ckpt_t[sg]));
        for (int j = 0; j < len; ++j) {
            CHK(snapshot_save(h, h->adj_seg[j]));
            h->adj_seg_t[j] = h->time;
            if (j + 1 < len) CHK(do_step(h));
        }
        // reverse it: vbar_h = vbar + dt xbar; xbar += VJP(x_j, c vbar_h); vbar = vbar_h
        for (int j = len - 1; j >= 0; --j) {
            const double c = (start + j == 0 ? 0.5 : 1.0) * dt;
            CHK(snapshot_restore(h, h->adj_seg[j], h->adj_seg_t[j]));
            launch_adj_kick(xb, vb, w, 3 * n, dt, c, true, h->stream);
            CHK(vjp_device(h, w));
            launch_adj_accum(h->gbuf.as<double>(), h->st[h->cur].id.as<uint32_t>(), (uint32_t)n, xb, mbar,
                             h->stream);
        }
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(gx, xb, n * 24, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(gv, vb, n * 24, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(gm, mbar, n * 8, hipMemcpyDeviceToHost, h->stream));
    // the handle ends at the trajectory's last state
    CHK(snapshot_restore(h, h->adj_final, h->adj_final_t));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NBH_OK;
}

int nbh_debug_sizes(nbh_handle* h, int64_t* sizes) {
    if (!h || !sizes) return NBH_E_INVALID;
    sizes[0] = h->num_nodes;
    sizes[1] = h->num_leaves;
    sizes[2] = h->depth;
    sizes[3] = (int64_t)h->stats.m2l_pairs;
    sizes[4] = (int64_t)h->stats.p2p_pairs;
    sizes[5] = h->nsys;
    return NBH_OK;
}

int nbh_debug_keys(nbh_handle* h, uint64_t* keys, uint32_t* ids, double* box) {
    if (!h) return NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    if (keys) HIPCHK(h, hipMemcpyAsync(keys, h->keys.p, h->n * 8, hipMemcpyDeviceToHost, h->stream));
    if (ids) HIPCHK(h, hipMemcpyAsync(ids, h->st[h->cur].id.p, h->n * 4, hipMemcpyDeviceToHost, h->stream));
    if (box) HIPCHK(h, hipMemcpyAsync(box, h->box.p, h->nsys * 32, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NBH_OK;
}

int nbh_debug_tree(nbh_handle* h, int32_t* level, int32_t* sys, int32_t* ix, int32_t* iy, int32_t* iz,
                   int32_t* begin, int32_t* count, int32_t* first_child, int32_t* nchild, int32_t* parent) {
    if (!h) return NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t nn = h->num_nodes;
    std::vector<int4> g(nn), l(nn);
    std::vector<int> s(nn);
    HIPCHK(h, hipMemcpyAsync(g.data(), h->geo.p, nn * 16, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(l.data(), h->link.p, nn * 16, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(s.data(), h->nsys_of.p, nn * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < nn; ++i) {
        if (level) level[i] = g[i].w & 0xff;
        if (nchild) nchild[i] = g[i].w >> 8;
        if (ix) ix[i] = g[i].x;
        if (iy) iy[i] = g[i].y;
        if (iz) iz[i] = g[i].z;
        if (sys) sys[i] = s[i];
        if (begin) begin[i] = l[i].x;
        if (count) count[i] = l[i].y;
        if (first_child) first_child[i] = l[i].z;
        if (parent) parent[i] = l[i].w;
    }
    return NBH_OK;
}

int nbh_debug_lists(nbh_handle* h, int64_t* m2l_off, int32_t* m2l_src, int64_t* p2p_off, int32_t* p2p_src) {
    if (!h) return NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    // the device keeps each node's lists as a chain of row segments; here they
    // are concatenated in chain (depth) order into CSR
    const size_t nn = h->num_nodes;
    size_t rows = 0;
    for (uint32_t s = 0; s <= h->depth + 1 && s <= (uint32_t)MAX_LEVELS; ++s)
        rows += h->hc->level_count[s] + h->hc->carried[s];
    rows = std::min<size_t>(rows, h->row_cap);
    std::vector<uint32_t> first(nn), next(rows), ms(rows), mc(rows), ps(rows), pc(rows);
    std::vector<int32_t> mpool(h->hc->m2l_pool_used), ppool(h->hc->p2p_pool_used);
    struct { void* dst; const void* src; size_t bytes; } cp[] = {
        {first.data(), h->row_first.p, nn * 4},     {next.data(), h->row_next.p, rows * 4},
        {ms.data(), h->row_m_start.p, rows * 4},    {mc.data(), h->row_m_cnt.p, rows * 4},
        {ps.data(), h->row_p_start.p, rows * 4},    {pc.data(), h->row_p_cnt.p, rows * 4},
        {mpool.data(), h->m2l_pool.p, mpool.size() * 4}, {ppool.data(), h->p2p_pool.p, ppool.size() * 4}};
    for (auto& c : cp)
        if (c.bytes) HIPCHK(h, hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int64_t om = 0, op = 0;
    for (size_t a = 0; a < nn; ++a) {
        if (m2l_off) m2l_off[a] = om;
        if (p2p_off) p2p_off[a] = op;
        for (uint32_t r = first[a]; r != NO_ROW; r = next[r]) {
            if (r >= rows) return h->fail(NBH_E_INTERNAL, "row chain of node %zu leaves the row table", a);
            if ((size_t)ms[r] + mc[r] > mpool.size() || (size_t)ps[r] + pc[r] > ppool.size() ||
                (uint64_t)om + mc[r] > h->stats.m2l_pairs || (uint64_t)op + pc[r] > h->stats.p2p_pairs)
                return h->fail(NBH_E_INTERNAL, "row segment of node %zu out of range", a);
            for (uint32_t k = 0; k < mc[r]; ++k, ++om)
                if (m2l_src) m2l_src[om] = mpool[ms[r] + k];
            for (uint32_t k = 0; k < pc[r]; ++k, ++op)
                if (p2p_src) p2p_src[op] = ppool[ps[r] + k];
        }
        // rows keep same-level M2L sources ahead of the others; report each
        // list in ascending source order
        if (m2l_src && m2l_off) std::sort(m2l_src + m2l_off[a], m2l_src + om);
    }
    if (m2l_off) m2l_off[nn] = om;
    if (p2p_off) p2p_off[nn] = op;
    return NBH_OK;
}

int nbh_debug_expansions(nbh_handle* h, float* M, float* L) {
    if (!h) return NBH_E_INVALID;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t nn = h->num_nodes;
    const int nc = ncoef((int)h->cfg.order), ncp = ncp_of((int)h->cfg.order);
    std::vector<float> buf(nn * ncp);
    float* outs[2] = {M, L};
    DevBuf* srcs[2] = {&h->M, &h->L};
    for (int k = 0; k < 2; ++k) {
        if (!outs[k]) continue;
        HIPCHK(h, hipMemcpyAsync(buf.data(), srcs[k]->p, buf.size() * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (size_t i = 0; i < nn; ++i)
            for (int c = 0; c < nc; ++c) outs[k][i * nc + c] = buf[i * ncp + c];
    }
    return NBH_OK;
}

}  // extern "C"
