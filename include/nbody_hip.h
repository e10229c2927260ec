/*
 * nbody_hip.h -- C ABI of libnbody_hip.so, the MI355X (gfx950) FMM force
 * pipeline behind nbody::HipSimulation.
 *
 * The reference has no C ABI: its only boundary is the C++ interface
 * nbody::Simulation (include/nbody/simulation.h:6-37) implemented by
 * NaiveSimulation (include/nbody/naive_simulation.h:12-41) and
 * OpenClSimulation (include/nbody/open_cl_simulation.h:19-203).  Each entry
 * point below names the reference interface or function it replaces.
 *
 * Conventions
 *   - plain pointers and sizes only; no exceptions cross this boundary;
 *   - every function returns 0 (NBH_OK) or a negative NBH_E* status, and the
 *     handle's message is available from nbh_last_error();
 *   - positions/velocities are FP64 AoS xyz triples; masses and charges FP64;
 *   - particle arrays passed in and out are in the caller's input order;
 *   - calls on one handle are not thread-safe; distinct handles are
 *     independent.  Each handle owns one HIP stream on its device.
 */
#ifndef NBODY_HIP_H_
#define NBODY_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBH_ABI_VERSION 3

/* status codes */
#define NBH_OK 0
#define NBH_E_INVALID (-1)     /* bad argument or configuration */
#define NBH_E_HIP (-2)         /* HIP runtime error (no device, launch failure ...) */
#define NBH_E_NOMEM (-3)       /* device allocation failed */
#define NBH_E_NONFINITE (-4)   /* NaN/Inf in the particle state */
#define NBH_E_INTERNAL (-5)    /* an internal consistency check failed */
#define NBH_E_OUTSIDE (-6)     /* fixed_root: a particle lies outside the root box */

/* solver */
#define NBH_SOLVER_FMM 0       /* replaces OpenClSimulation's pipeline */
#define NBH_SOLVER_DIRECT 1    /* O(N^2) on the GPU; replaces NaiveSimulation */

/* far-field (M2L) kernel */
#define NBH_M2L_PLAIN 0        /* per-pair FP32 VALU, derivative tensor per pair */
#define NBH_M2L_GROUPED 1      /* FP32 VALU, grouped by translation offset */
#define NBH_M2L_BF16X3 2       /* grouped; same-level pairs by translation offset on MFMA with the dense
                                  operator (every |k|, |n| <= p), bf16x3 split products, FP32 accumulate */
#define NBH_M2L_MFMA_F32 3     /* the same with FP32 products */

/* M2L term set at truncation q = m2l_order (k, n of order <= p) */
#define NBH_M2L_TERMS_ALL 0    /* every term with |k| + |n| <= q */
#define NBH_M2L_TERMS_LOW 1    /* |k| + |n| <= p, and up to q only those with min(|k|, |n|) <= q - p:
                                  at q = p + 1 the degree-(p + 1) terms of the dipole (|n| = 1) and
                                  their mirror images (|k| = 1); 300 instead of 420 terms at p = 4 */

/* near-field (P2P) kernel */
#define NBH_P2P_LEAF 0         /* one wavefront per target leaf (plain) */
#define NBH_P2P_LANE 1         /* one lane per target particle, 64 consecutive particles per wavefront */

/* integrator */
#define NBH_INT_KDK 0          /* kick-drift-kick leapfrog, FP64 state */
#define NBH_INT_EULER 1        /* explicit Euler as open_cl_simulation.cpp:600-607 */
#define NBH_INT_KICK_DRIFT 2   /* kick then drift as naive_simulation.cpp:28-41 */

/* force law (direct solver only) */
#define NBH_LAW_SOFTENED 0     /* field.cl:17-32 (Plummer softened, exact separation) */
#define NBH_LAW_COMPAT_NAIVE 1 /* naive_simulation.cpp:10-26: separation (dx,dx,dx), no eps */

typedef struct nbh_config {
    uint32_t struct_size;   /* = sizeof(nbh_config), checked by nbh_create */
    uint32_t solver;        /* NBH_SOLVER_* */
    double G;               /* gravitational constant; reference FORCE_CONSTANT = -G (field.cl:7-9) */
    double eps;             /* Plummer softening (reference PARTICLE_RADIUS, field.cl:3-5) */
    double theta;           /* opening ratio (reference NODE_APPROX_RATIO, interaction.cl:3-5);
                               must equal num/den with den <= 64 */
    double dt;              /* time step (OpenClSimulation timeStep) */
    uint32_t leaf_capacity; /* max particles per leaf cell (reference hard-codes 8,
                               open_cl_simulation.cpp:41-47) */
    uint32_t order;         /* Cartesian Taylor order p, 1..6 */
    uint32_t max_depth;     /* deepest octree level, <= bits per axis (21 for one system) */
    uint32_t m2l;           /* NBH_M2L_* */
    uint32_t integrator;    /* NBH_INT_* */
    uint32_t force_law;     /* NBH_LAW_* (direct solver) */
    int32_t device;         /* HIP device ordinal */
    uint32_t fixed_root;    /* 0: root cube = bounding cube each step; 1: [root_lo, root_lo+root_size]
                               (a particle outside it fails with NBH_E_OUTSIDE) */
    double root_lo[3];
    double root_size;
    uint32_t p2p;           /* NBH_P2P_* */
    uint32_t m2l_order;     /* M2L truncation q, p or p + 1 (0 = p + 1): the terms (k, n) with
                               |k|, |n| <= p and |k| + |n| <= q, as m2l_terms selects */
    uint32_t m2l_terms;     /* NBH_M2L_TERMS_*: which terms of degree p + 1 < |k| + |n| <= q
                               the M2L keeps (no effect at q = p) */
    uint32_t p2p_chunk;     /* NBH_P2P_LANE tile shape: source particles staged in LDS per
                               wavefront round, 64 or 128 (0 = 64); nbh_autotune picks it */
    uint32_t graph;         /* 1: capture each KDK step in a HIP graph and replay it (no host
                               round trip inside the step; FMM with NBH_M2L_GROUPED and
                               NBH_P2P_LANE, timing off); 0: off */
} nbh_config;

typedef struct nbh_handle nbh_handle;

/* Per-step statistics of the last force evaluation (FMM). */
typedef struct nbh_stats {
    uint64_t num_nodes, num_leaves, depth;
    uint64_t m2l_pairs;      /* directed M2L list entries */
    uint64_t p2p_pairs;      /* directed P2P leaf-pair list entries */
    uint64_t p2p_interactions; /* directed particle-particle interactions (self excluded) */
    uint64_t split_pairs;    /* pairs opened by the walk */
    double ms_tree, ms_moments, ms_walk, ms_m2l, ms_l2l, ms_p2p, ms_total; /* if timing enabled */
    uint64_t m2l_same_order; /* M2L truncation q of the pairs of one level: 2p (every term
                                |k|, |n| <= p) when they run on the matrix cores, else m2l_order */
    uint64_t m2l_same_pairs;  /* directed M2L entries between cells of one level */
    uint64_t m2l_same_launches; /* timing: same-level M2L launches bracketed by events */
    double ms_m2l_same;       /* timing: sum of their durations, each on its own stream */
    uint64_t graph_captures;  /* cfg.graph: graphs captured, steps replayed from one, and replayed */
    uint64_t graph_replays;   /* steps that outgrew the captured shapes (rolled back, run */
    uint64_t graph_fallbacks; /* uncaptured; the next step captures anew); cumulative */
} nbh_stats;

int nbh_abi_version(void);
size_t nbh_config_size(void);

/* Fills the reference's defaults: G=1 (FORCE_CONSTANT=-1), eps=0.01,
 * theta=0.5, leaf capacity 8, p=4, KDK, FMM, grouped M2L, per-lane P2P. */
int nbh_config_default(nbh_config* cfg);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int nbh_device_count(void);

/* Replaces the OpenClSimulation / NaiveSimulation constructors
 * (open_cl_simulation.cpp:15-51, naive_simulation.h:25-33): copies the
 * particles to the device.  pos, vel: 3n doubles; mass, charge: n doubles.
 * n >= 1. */
int nbh_create(const nbh_config* cfg, size_t n, const double* pos, const double* vel,
               const double* mass, const double* charge, nbh_handle** out);

/* Ensemble of nsys independent systems in one handle (no counterpart in the
 * reference): sys_off[nsys+1] splits the particle arrays into systems that
 * never interact; every FMM phase processes all systems in one launch. */
int nbh_create_ensemble(const nbh_config* cfg, size_t nsys, const uint32_t* sys_off,
                        const double* pos, const double* vel, const double* mass,
                        const double* charge, nbh_handle** out);

void nbh_destroy(nbh_handle* h);

/* Message of the last failing call on h (or of the last failing nbh_create
 * when h is NULL).  Never NULL. */
const char* nbh_last_error(const nbh_handle* h);

/* Replaces Simulation::step() (simulation.h:34; open_cl_simulation.cpp:70-106,
 * naive_simulation.cpp:7-46): advances nsteps steps of dt and returns the new
 * time in *t_out (may be NULL). */
int nbh_step(nbh_handle* h, int nsteps, double* t_out);

/* Replaces Simulation::particles() (simulation.h:35; open_cl_simulation.cpp:
 * 53-68, naive_simulation.h:37-39), in input order.  Any pointer may be NULL. */
int nbh_get_state(nbh_handle* h, double* pos, double* vel, double* mass,
                  double* charge, double* t);
int nbh_set_state(nbh_handle* h, const double* pos, const double* vel, double t);

/* One force evaluation at the current state (no integration), input order:
 * acc[3n] = (q_i/m_i) * field (force.cl:4-19 divided by mass), pot[n] =
 * potential per unit charge.  Either may be NULL. */
int nbh_compute_forces(nbh_handle* h, float* acc, float* pot);

/* Autotuner (no counterpart in the reference, which hard-codes a leaf
 * capacity of 8 with a FIXME, open_cl_simulation.cpp:41-47): for every leaf
 * capacity in caps[0..ncaps) and P2P tile chunk (64 or 128) in
 * chunks[0..nchunks) (NULL or 0: the handle's current value), runs one
 * untimed and `evals` timed force evaluations at the current state, keeps
 * the fastest pair on the handle and reports it in *out (may be NULL);
 * ms_table (may be NULL) receives ncaps x nchunks mean ms per evaluation,
 * chunk index fastest.  Positions and velocities are unchanged; the
 * accelerations are those of the chosen configuration.  FMM solver only. */
typedef struct nbh_tune_result {
    uint32_t leaf_capacity;
    uint32_t p2p_chunk;
    double ms_per_eval;
    uint32_t tried;
} nbh_tune_result;
int nbh_autotune(nbh_handle* h, const uint32_t* caps, size_t ncaps, const uint32_t* chunks, size_t nchunks,
                 int evals, nbh_tune_result* out, double* ms_table);

/* Reference check: FP64 direct sum on the GPU for the given target particle
 * indices (input order) over all sources of their system; acc[3*nt], pot[nt]
 * include the q_i/m_i factor like nbh_compute_forces. */
int nbh_direct_fp64(nbh_handle* h, size_t nt, const int64_t* targets, double* acc, double* pot);

/* nbh_direct_fp64 without device allocation (no counterpart in the
 * reference).  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * per-call target list, slots and FP64 results live in a fixed scratch array
 * of the library on each GPU (44 bytes per target, 2.75 MiB, held until the process
 * ends), so nt is at most NBH_NOALLOC_MAX_TARGETS (larger nt fails with
 * NBH_E_INVALID).  Same targets, same kernel and results as nbh_direct_fp64.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (from several threads or handles) race. */
#define NBH_NOALLOC_MAX_TARGETS 65536
int nbh_direct_fp64_noalloc(nbh_handle* h, size_t nt, const int64_t* targets, double* acc, double* pot);

/* Energy and momentum of the current state (FP64, on the GPU):
 * out[0] kinetic, out[1] potential, out[2..4] linear momentum,
 * out[5..7] angular momentum.  Steps evaluate accelerations only (as the
 * reference's step, open_cl_simulation.cpp:81-98): when the last evaluation
 * was a step's, this call runs one evaluation with potentials at the current
 * positions first, leaving the accelerations the integrator holds as they
 * were. */
int nbh_diagnostics(nbh_handle* h, double* out);

int nbh_get_stats(nbh_handle* h, nbh_stats* s);
/* Per-phase hipEvent timing of force evaluations: 0 off, 1 on, 2 on with the
 * whole evaluation serialized on the handle's main stream (each kernel then
 * runs alone, for standalone kernel durations; results are unchanged up to
 * the FP32 summation order of the far-field parts). */
int nbh_set_timing(nbh_handle* h, int enable);

/* ---- adjoint (no counterpart in the reference) --------------------------- */
/* Vector-Jacobian product of the accelerations a(x, m) of the current state
 * (gravity, charge == mass) with cotangent w[3n]: gx[3n] = (da/dx)^T w,
 * gm[n] = (da/dm)^T w, input order.  Uses the tree and M2L/P2P lists of a
 * force evaluation at the current positions (FP32 kernels). */
int nbh_force_vjp(nbh_handle* h, const double* w, double* gx, double* gm);

/* Reverse-mode gradient of a KDK trajectory, kept on the device (no
 * counterpart in the reference).  nbh_adjoint_forward runs nsteps steps from
 * the current state (like nbh_step) and checkpoints the full state every
 * ckpt_every steps on the device.  nbh_adjoint_backward then takes the
 * cotangents xbar[3n], vbar[3n] of the final positions and velocities and
 * returns the gradient of that scalar with respect to the initial positions
 * gx[3n], velocities gv[3n] and masses gm[n] (gravity, charge == mass; input
 * order): it recomputes each segment from its checkpoint and runs the
 * transposed leapfrog with one force VJP (nbh_force_vjp's kernels) per state.
 * The handle ends at the final state.  Needs the FMM solver and KDK. */
int nbh_adjoint_forward(nbh_handle* h, int nsteps, int ckpt_every);
int nbh_adjoint_backward(nbh_handle* h, const double* xbar, const double* vbar, double* gx, double* gv,
                         double* gm);

/* ---- tree / list export for parity tests (read-only views) -------------- */
/* sizes[0]=nodes, [1]=leaves, [2]=depth, [3]=m2l entries, [4]=p2p entries,
 * [5]=number of systems, of the last force evaluation. */
int nbh_debug_sizes(nbh_handle* h, int64_t* sizes);
/* keys[n] (sorted), perm: id[n] (input index of each sorted slot),
 * box[4*nsys] (lo xyz, size) */
int nbh_debug_keys(nbh_handle* h, uint64_t* keys, uint32_t* ids, double* box);
/* per node: level, sys, ix, iy, iz, begin, count, first_child, nchild, parent */
int nbh_debug_tree(nbh_handle* h, int32_t* level, int32_t* sys, int32_t* ix, int32_t* iy,
                   int32_t* iz, int32_t* begin, int32_t* count, int32_t* first_child,
                   int32_t* nchild, int32_t* parent);
/* CSR lists: off[nodes+1] (int64), src (int32) */
int nbh_debug_lists(nbh_handle* h, int64_t* m2l_off, int32_t* m2l_src, int64_t* p2p_off,
                    int32_t* p2p_src);
/* multipole and local coefficients per node: [nodes * ncoef] float */
int nbh_debug_expansions(nbh_handle* h, float* M, float* L);

/* ---- seeded initial conditions (host-side, no GPU needed) ---------------- */
/* Uniform cube [0,1)^3 positions, zero velocities, masses = charges = 1/n. */
int nbh_ic_uniform(uint64_t seed, size_t n, double* pos, double* vel, double* mass);
/* Plummer sphere (scale a, total mass M, G): radii from the cumulative mass
 * profile truncated at rmax, velocities by von Neumann rejection from the
 * distribution function (Aarseth, Henon & Wielen 1974).  Masses M/n. */
int nbh_ic_plummer(uint64_t seed, size_t n, double a, double M, double G, double rmax,
                   double* pos, double* vel, double* mass);

#ifdef __cplusplus
}
#endif
#endif /* NBODY_HIP_H_ */
