/*
 * oracle.h -- CPU restatement of the nbody FMM force pipeline, for TESTS ONLY.
 *
 * This library is test infrastructure.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product (libnbody_hip.so, HipSimulation) never links or calls it.
 *
 * Everything here is FP64 unless stated, written for clarity rather than speed,
 * and follows the reference's semantics where the reference defines them:
 *
 *   physics / sign convention ....... src/field.cl:17-47   (softened field, F=-1)
 *   force = charge * field .......... src/force.cl:4-19
 *   multipole centre = cell centre .. src/moment.cl:27
 *   dual-tree pair splitting + MAC .. src/interaction.cl:42-92
 *   pair classification ............. src/open_cl_simulation.cpp:252-265
 *   root seed pair (root,root) ...... include/nbody/open_cl_simulation.h:127-133
 *   explicit Euler integration ...... src/open_cl_simulation.cpp:600-607
 *   NaiveSimulation::step ........... src/naive_simulation.cpp:7-46
 *
 * The reference stops at the monopole; the Cartesian Taylor expansion of
 * order p (P2M/M2M/M2L/L2L/L2P) is the textbook one, restated in oracle.cpp
 * with its derivation, and pinned in tests by closed forms and convergence to
 * the direct sum.
 */
#ifndef NBODY_ORACLE_H_
#define NBODY_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- direct sum --------------------------------------------------------- */
/* acc[3*nt], pot[nt]:  a_i = G * sum_j q_j (x_j - x_i) / (|x_j-x_i|^2+eps^2)^1.5,
 * pot_i = -G * sum_j q_j / sqrt(|x_j-x_i|^2+eps^2), j != self[i] (self[i] < 0
 * means no exclusion).  Targets xt[3*nt], sources xs[3*ns] AoS.  The caller
 * multiplies by q_i/m_i to get accelerations (force.cl:4-19).  OpenMP. */
void orc_direct(size_t ns, const double* xs, const double* qs, size_t nt,
                const double* xt, const int64_t* self, double G, double eps,
                double* acc, double* pot);

/* Vector-Jacobian product of a_i(x, m) (q == m, gravity) with cotangent w:
 * gx[3n] = d(sum_i w_i . a_i)/dx,  gm[n] = d(sum_i w_i . a_i)/dm.  FP64. */
void orc_direct_vjp(size_t n, const double* x, const double* m, const double* w,
                    double G, double eps, double* gx, double* gm);

/* ---- Morton keys -------------------------------------------------------- */
/* Per-system bounding cube (lo[3], size) and 63-bit (or sys-prefixed) keys.
 * box_out[4*s + 0..2] = lo, box_out[4*s+3] = size.  bits = bits per axis. */
void orc_morton(size_t n, const double* x, size_t nsys, const uint32_t* sys_off,
                int bits, uint64_t* keys, double* box_out);

/* ---- tree --------------------------------------------------------------- */
typedef struct orc_tree orc_tree;
/* keys must be sorted ascending (per system contiguous).  Builds the octree:
 * a cell splits iff count > cap and level < max_depth. */
orc_tree* orc_tree_build(size_t n, const uint64_t* sorted_keys, size_t nsys,
                         const uint32_t* sys_off, int bits, int cap, int max_depth);
void orc_tree_free(orc_tree*);
size_t orc_tree_num_nodes(const orc_tree*);
int orc_tree_depth(const orc_tree*);
/* Per node (canonical BFS order = (level, system, Morton prefix)):
 * level, sys, ix, iy, iz, begin, count, first_child (-1), nchild, parent (-1). */
void orc_tree_nodes(const orc_tree*, int32_t* level, int32_t* sys, int32_t* ix,
                    int32_t* iy, int32_t* iz, int32_t* begin, int32_t* count,
                    int32_t* first_child, int32_t* nchild, int32_t* parent);

/* ---- dual-tree walk (reference MAC) ------------------------------------- */
/* theta = theta_num / theta_den exactly.  Restates interaction.cl:42-92 as a
 * recursion over unordered pairs from (root, root) per system and returns
 * directed lists sorted by source id per target. */
typedef struct orc_lists orc_lists;
orc_lists* orc_walk(const orc_tree*, int theta_num, int theta_den);
void orc_lists_free(orc_lists*);
/* sizes: [0]=m2l entries, [1]=p2p entries, [2]=unordered m2l pairs,
 *        [3]=unordered p2p pairs, [4]=pairs visited (split) */
void orc_lists_sizes(const orc_lists*, int64_t* sizes);
/* CSR over nodes: off[nnodes+1], src[...] */
void orc_lists_get(const orc_lists*, int64_t* m2l_off, int32_t* m2l_src,
                   int64_t* p2p_off, int32_t* p2p_src);

/* ---- Cartesian Taylor expansion primitives (order p) --------------------- */
int orc_ncoef(int p);
/* multi-index table: mi[3*k + 0..2] for coefficient k */
void orc_multi_indices(int p, int32_t* mi);
/* D_n(R) = d^n/dR^n (|R|^2+eps^2)^(-1/2), |n| <= p */
void orc_derivs(int p, const double* R, double eps, double* D);
/* M_n = sum_j q_j (x_j - z)^n / n! */
void orc_p2m(int p, const double* z, size_t n, const double* x, const double* q, double* M);
/* shift multipole from child centre zc to parent centre zp */
void orc_m2m(int p, const double* zc, const double* zp, const double* Mc, double* Mp);
/* L_k += sum (-1)^|n| D_{k+n}(zA - zB) M_n over |k|,|n| <= p.  qc = q | 16
 * keeps every term with |k|+|n| <= q; qc = q (the low-degree set) keeps
 * |k|+|n| <= p, or |k|+|n| <= q and min(|k|,|n|) <= q-p (q >= p).  q == p is
 * the usual total-degree truncation either way; the low set at q == p+1 adds
 * only the degree-(p+1) terms with |n| = 1 or |k| = 1.  Both sets are
 * symmetric in (k, n). */
void orc_m2l(int p, int qc, const double* zA, const double* zB, double eps, const double* M, double* L);
/* shift local expansion from parent zp to child zc (accumulates into Lc) */
void orc_l2l(int p, const double* zp, const double* zc, const double* Lp, double* Lc);
/* evaluate: phi_L(x) = sum_k L_k a^k/k!, g = sum_k L_{k+e} a^k/k! (a = x - z) */
void orc_l2p(int p, const double* z, const double* L, const double* x, double* phi, double* g);

/* ---- full FMM force evaluation (FP64), same tree/lists as the product ---- */
/* x[3n] in sorted order, q[n]; returns acc (without the q_i/m_i factor) and
 * potential per particle, in the same order.  box[4*nsys] from orc_morton.
 * qm2l = qc | qc_same << 8: M2L truncation (orc_m2l's qc) of pairs of different
 * levels, and of pairs of one level (qc_same; 0 = the same qc). */
void orc_fmm(const orc_tree*, const orc_lists*, int p, int qm2l, size_t n, const double* x,
             const double* q, const double* box, double G, double eps,
             double* acc, double* pot);

/* VJP of orc_fmm for gravity (q == m): gx[3n], gm[n] of sum_i w_i . acc_i,
 * with the tree and lists held fixed (exact transpose of the FMM). */
void orc_fmm_vjp(const orc_tree*, const orc_lists*, int p, int qm2l, size_t n, const double* x, const double* m,
                 const double* w, const double* box, double G, double eps, double* gx, double* gm);

/* ---- integrators over the FP64 direct sum ------------------------------- */
/* integrator: 0 = KDK leapfrog, 1 = explicit Euler (open_cl_simulation.cpp:
 * 600-607), 2 = kick-drift (naive_simulation.cpp:28-41).  In-place, nsteps. */
void orc_integrate_direct(size_t n, double* x, double* v, const double* m,
                          const double* q, double G, double eps, double dt,
                          int nsteps, int integrator);

/* ---- NaiveSimulation bug-for-bug FP32 restatement (compat) --------------- */
/* p[12*n] float: {px,py,pz,pw, vx,vy,vz,vw, mass, charge, pad, pad} per
 * particle, like the reference Particle layout.  time in/out. */
void orc_naive_compat_steps(size_t n, float* p, float force_constant, float dt,
                            int nsteps, float* time);

#ifdef __cplusplus
}
#endif
#endif
