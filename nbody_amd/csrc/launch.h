// launch.h -- host-side launchers of the device kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "common.h"

namespace nbh {

// ---- scan / sort (sort.hip)
template <typename TI, typename TO>
void exclusive_scan(const TI* in, TO* out, const uint32_t* n_dev, uint32_t n_max, TO* total, void* tmp,
                    hipStream_t st);
size_t scan_tmp_bytes(uint32_t n_max);
size_t radix_tmp_bytes(uint32_t n);
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                      int nbits, void* tmp, hipStream_t st);

// ---- tree (tree.hip)
void launch_bbox(const double* x, const double* y, const double* z, const uint32_t* sys_off, int nsys,
                 uint32_t max_sys_n, int bits, int fixed, const double* flo, double fsize, double* part,
                 double4* box, double* inv, float* unit, uint32_t* nonfinite, hipStream_t st);
void launch_keys(const double* x, const double* y, const double* z, uint32_t n, const uint32_t* sys_off,
                 int nsys, int bits, const double4* box, const double* inv, uint64_t* keys, uint32_t* perm,
                 uint32_t* fixed_flags, hipStream_t st);
size_t tree_tmp_bytes(uint32_t n, int max_depth);
void launch_tree_build(const uint64_t* keys, uint32_t n, const uint32_t* sys_off, int nsys, int bits, int cap,
                       int max_depth, NodeView nodes, uint32_t* node_begin, uint32_t* leaf_ids, uint32_t node_cap,
                       Counters* C, void* tmp, hipStream_t st);
void launch_tree_emit(const uint64_t* keys, uint32_t n, const uint32_t* sys_off, int nsys, int bits, int cap,
                      int max_depth, NodeView nodes, uint32_t* node_begin, uint32_t* leaf_ids, uint32_t node_cap,
                      Counters* C, void* tmp, hipStream_t st);
void launch_tree_link(const uint64_t* keys, uint32_t n, const uint32_t* sys_off, int nsys, int bits, int cap,
                      int max_depth, const Counters* C, uint32_t num_nodes, const uint32_t* node_begin,
                      NodeView nodes, hipStream_t st);

// ---- walk (walk.hip)
// Per-wave chunk sizes of the walk's segment pools (entries).  A wave reserves
// a chunk with one atomic and carves row segments from it; the host sizes the
// pools for one unused tail per wave per depth.
constexpr uint32_t POOL_CHUNK_M = 2048, POOL_CHUNK_P = 512, POOL_CHUNK_S = 512;
constexpr uint32_t WALK_WAVES = 4;

// A node's interaction lists are the segments of its rows, one row per walk
// depth that reached it: row_first[node] is its level row, row_next[row] the
// row one depth deeper (a carried leaf), NO_ROW ends the chain.
constexpr uint32_t NO_ROW = 0xffffffffu;

struct WalkParams {
    NodeView nodes;
    Counters* C;
    int bits;
    uint32_t num4;  // 4 num^2
    uint32_t den3;  // 3 den^2
    uint32_t row_cap;
    int* row_node;
    uint32_t* row_first;  // [nodes]
    uint32_t* row_next;
    uint32_t* row_sl_start;
    uint32_t* row_sl_cnt;
    uint32_t* row_m_start;
    uint32_t* row_m_cnt;   // M2L entries of the row: same level first, then
    uint32_t* row_mx_cnt;  // the last row_mx_cnt at other levels
    uint32_t* row_p_start;
    uint32_t* row_p_cnt;
    int* sl;
    uint32_t sl_cap;
    int* m2l_pool;
    int* p2p_pool;
    unsigned long long m2l_pool_cap, p2p_pool_cap;
};
void launch_walk_root(const WalkParams& W, hipStream_t st);
void launch_walk_group(const WalkParams& W, int s, uint32_t grid, hipStream_t st);
// workgroups of the group kernel resident at once on a device with num_cu CUs
uint32_t walk_group_resident_blocks(int num_cu);

// Read-only view of the lists for the force kernels.
struct ListView {
    const uint32_t* first;
    const uint32_t* next;
    const uint32_t* m_start;
    const uint32_t* m_cnt;
    const uint32_t* mx_cnt;  // trailing M2L entries at another level than the target
    const uint32_t* p_start;
    const uint32_t* p_cnt;
    const int* m_pool;
    const int* p_pool;
};

// ---- particle state in sorted order (FP64 positions/velocities)
struct State {
    double *x, *y, *z, *vx, *vy, *vz, *mass, *charge;
    float *ax, *ay, *az, *pot;
    uint32_t* id;
};

// ---- force pipeline (fmm.hip)
struct FmmParams {
    NodeView nodes;
    const Counters* C;
    const double4* box;
    const float* unit;     // per system: half of the finest cell, float
    int bits;
    float eps2;
    float G;
    float* M;              // [nodes][ncp]
    float* L;              // [nodes][ncp]
    // with the tables: the other-level and deferred pairs' part of L, kept
    // apart so that it can be computed beside the same-level kernels; L2L
    // adds it in (nullptr: those kernels add to L)
    float* Lx;
    int ncp;               // padded coefficient stride
    const uint32_t* leaf_ids;
    float4* pf;            // per particle: offset from its leaf centre, charge
    ListView lists;
    unsigned long long* p2p_inter;  // statistics: particle pairs of the P2P kernels
    uint32_t* pleaf;       // per particle: its leaf node
    float* dtab;           // per system and level, derivative tensors of the same-level offsets (or null);
                           // system s, level l at row ((s dnlev + l - dlevel0) dsz + row)
    int dK, dlevel0, dnlev;
    int dcp;               // tensor row stride (floats): ncoef(Q) padded to 4
    int dsz;               // tensors per level: the offsets that occur, and a zero row last
    const uint32_t* doffs; // [dsz] offset (|o_x|, |o_y|, |o_z|) of each row, 8 bits each (the zero row: ~0)
    const uint16_t* dmap;  // [(K+1)^3] row of the offset (|o_x| (K+1) + |o_y|) (K+1) + |o_z|
    // pairs one level apart: per system and finer level lf, tensors of the
    // centre offsets d (units of half the finer cell, odd) with |d_i| = 1, 3,
    // .., 2 xE - 1
    float* xtab;
    int xE, xlevel0, xnlev;
    // pairs further apart in level: deferred as (target, source) pairs, the
    // pairs of one target contiguous
    int* far_tgt;
    int* far_src;
    uint32_t far_cap;
    uint32_t* far_cnt;     // [targets, pairs]
    double eps2_64;
    int pot;               // 1: potentials too; 0: accelerations only (the tuned P2P and M2L kernels leave
                           // out the potential's terms, l2p_kernel does not write it)
};

void launch_gather_state(const State& src, const State& dst, const uint32_t* perm, uint32_t n, hipStream_t st);
// count (num_leaves) == DEV_COUNT: read on the device, the grid sized for
// grid_items; the per-level launchers then read the level's count and offset
void launch_p2m(int p, const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st,
                uint32_t grid_items = 0);
void launch_m2m_level(int p, const FmmParams& F, int level, uint32_t count, uint32_t off, hipStream_t st,
                      uint32_t grid_items = 0);
void launch_l2l_level(int p, const FmmParams& F, int level, uint32_t count, uint32_t off, hipStream_t st,
                      uint32_t grid_items = 0);
// the P2P kernels leave the raw near-field sums of every particle in (ax, ay,
// az, pot); launch_l2p adds the far field of the leaf's local expansion,
// applies G q/m and the sign of the potential, and the kick v += a kick_dt
void launch_p2p(int p, const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st);
// chunk: sources staged per wavefront round, 64 or 128 (the tile shape)
void launch_p2p_staged(int p, int chunk, const FmmParams& F, const State& S, uint32_t num_leaves, hipStream_t st,
                       uint32_t grid_items = 0);
void launch_l2p(int p, const FmmParams& F, const State& S, uint32_t n, double kick_dt, hipStream_t st);
// M2L of order p with the truncation code q (expansion.h: m2l_deg(q) = p or
// p + 1; q | M2L_ALL keeps every term of degree <= p + 1, q = p + 1 only
// those with |n| = 1 or |k| = 1 beyond degree p).
// ---- same-level M2L by translation offset on the matrix cores (m2l_offset.hip)
constexpr int OFF_J = 4;      // translation offsets per batch (along the contraction)
constexpr int OFF_WAVES = 4;  // wavefronts per workgroup (one 4x4x4 block of targets)
struct OffsetParams {
    const void* mb;         // signed split moments per node, zero row at zero_node
    uint32_t zero_node;
    const uint32_t* tab;    // dense tensors, per level 2.. : [E^3][off_table_words(p)]
    int E, K, W;            // E = K + 1 (octant table edge), K = max |o_i|, W = 4 + 2K (window edge)
    const uint32_t* cand;   // candidate offsets: (ox+16) | (oy+16) << 5 | (oz+16) << 10 | class mask << 16
    int ncand;
    uint32_t* flags;        // Counters::overflow (OVF_GEOM)
    const uint16_t* hank;   // off_tables: Hankel byte offsets of the A operand
    const uint8_t* par;     // off_tables: octant parity bits of the tensor words
};
// the host tables of the kernel (uploaded once per handle)
void off_tables(int p, bool bf3, std::vector<uint16_t>& hank, std::vector<uint8_t>& par);
bool off_supported(int p);
size_t off_lds_bytes(int p, bool bf3, int W, int ncand);
size_t off_moment_bytes(int p, bool bf3, uint32_t nn);
int off_table_words(int p);
void launch_off_moments(int p, bool bf3, const float* M, int ncp, uint32_t nn, const int4* geo, const double4* box,
                        void* out, hipStream_t st);
void launch_off_dtable(int p, bool bf3, const double4* box, double eps2, int level0, int nlev, int E, uint32_t* out,
                       hipStream_t st);
// the level-l targets below the level-(l-2) cells [goff, goff + gcount); lvl = l - 2
void launch_m2l_offset_level(int p, bool bf3, const FmmParams& F, const OffsetParams& O, uint32_t goff,
                             uint32_t gcount, int lvl, hipStream_t st);

struct M2LPlan {
    int mode;  // 0: plain (one thread per target, tensor per pair); 1: grouped
    int mfma;  // grouped, same-level pairs: 0 FP32 VALU; 1 MFMA bf16x3; 2 MFMA FP32 (by offset, dense operator)
    OffsetParams off;  // with mfma != 0
    uint32_t num_nodes;
    uint32_t level_off[MAX_LEVELS + 2];
    uint32_t level_count[MAX_LEVELS + 1];
    size_t max_lds;  // LDS bytes per workgroup: the same-level table is held in LDS if it fits
    int nsys;        // systems: with more than one, each target reads its own system's table
    // counts read on the device (a captured step): num_nodes and level_count
    // then only size the grids, and level_off is not used
    bool dev;
};
// tensors at R = -(a + b e) h_l / 2^shift for e in [0, E)^3 (offs null) or
// for the nrows offsets e of offs (8 bits per axis; ~0: a row of zeros),
// levels level0 .. level0+nlev-1 of systems 0 .. nsys-1 (h_l from each
// system's root size), into out (rows of F.dcp floats, system-major, then
// level)
void launch_dtable(int q, const FmmParams& F, float* out, int level0, int nlev, int E, int a, int b, int shift,
                   int nsys, const uint32_t* offs, int nrows, hipStream_t st);
void launch_m2l(int p, int q, const FmmParams& F, const M2LPlan& plan, hipStream_t st);
// The grouped M2L in parts (when F carries the tables): the same-level kernel
// of one level (level < 0: every level), and the rest (initialisation, the
// other-level and deferred pairs), which must follow all same-level kernels.
enum : int { M2L_LEVELS = 1, M2L_REST = 2 };
void launch_m2l_parts(int p, int q, const FmmParams& F, const M2LPlan& plan, int parts, int level, hipStream_t st);

// integrator updates: mode 0 = KDK first half (kick dt_k, drift dt_d),
// 1 = Euler (x += v dt; v += a dt), 2 = kick-drift (v += a dt; x += v dt)
void launch_integrate(const State& S, uint32_t n, int mode, double dt_k, double dt_d, hipStream_t st);

// direct sums
void launch_direct_f32(const State& S, uint32_t n, const uint32_t* sys_off, int nsys, const double4* box,
                       float G, float eps2, int compat_law, double kick_dt, hipStream_t st);
void launch_direct_fp64(const State& S, const uint32_t* sys_off, int nsys, const uint32_t* targets, uint32_t nt,
                        double G, double eps2, double* acc, double* pot, hipStream_t st);
void launch_find_slots(const uint32_t* id, uint32_t n, const int64_t* targets, uint32_t nt, uint32_t* slots,
                       hipStream_t st);

void launch_diagnostics(const State& S, uint32_t n, double* partial, uint32_t nblk, hipStream_t st);
uint32_t diagnostics_blocks(uint32_t n);

// adjoint (vjp.hip): g = [gx | gy | gz | gm] in sorted order, FP64.  In
// order: gather the cotangents (pw); the near field (raw sums into g) and the
// transposed far-field chain (dipole P2M, M2M, M2L, L2L into Fw) are
// independent; the far-field join adds the rest to g.
void launch_vjp_gather(const double* w_in, const uint32_t* id, float4* pw, uint32_t n, hipStream_t st);
// reverse leapfrog: w = c vbar (after vbar += dt xbar when update); xbar[id] +=
// g, mbar[id] += gm from a VJP in sorted order
void launch_adj_kick(double* xb, double* vb, double* w, size_t m, double dt, double c, bool update, hipStream_t st);
void launch_adj_accum(const double* g, const uint32_t* id, uint32_t n, double* xb, double* mb, hipStream_t st);
void launch_fill_f64(double* p, size_t m, double v, hipStream_t st);
void launch_vjp_near(int p, const FmmParams& F, const float4* pw, uint32_t n, uint32_t num_leaves, double* g,
                     hipStream_t st);
void launch_vjp_chain(int p, int q, const M2LPlan& plan, const FmmParams& F, const FmmParams& Fw, const float4* pw,
                      uint32_t num_leaves, const uint32_t* level_count, const uint32_t* level_off, int depth,
                      hipStream_t st);
void launch_vjp_far(int p, const FmmParams& F, const FmmParams& Fw, const float4* pw, uint32_t n, double* g,
                    hipStream_t st);

void launch_inverse_perm(const uint32_t* id, uint32_t* slot_of, uint32_t n, hipStream_t st);

}  // namespace nbh
