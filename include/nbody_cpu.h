/* nbody_cpu.h -- C ABI of the multi-core CPU backend (libnbody_cpu.so).
 *
 * The FP64 direct-sum engine of include/nbody/cpu_simulation.h behind plain
 * pointers, for the Python harness and other FFI callers.  It mirrors the
 * reference's NaiveSimulation (naive_simulation.h:25-33: particles, force
 * constant, time step; step() advancing and returning the time) with the
 * softened force law of field.cl:17-32 and a choice of integrator.  Every
 * function returns 0 or a negative status; ncpu_last_error() holds the
 * message of the last failure on the calling thread.
 */
#ifndef NBODY_CPU_H_
#define NBODY_CPU_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ncpu_handle ncpu_handle;

/* pos/vel: 3n FP64 (AoS), mass/charge: n (charge NULL: charge = mass);
 * integrator 0 KDK, 1 explicit Euler, 2 kick-drift; threads 0 = default. */
int ncpu_create(size_t n, const double* pos, const double* vel, const double* mass, const double* charge,
                double dt, double G, double eps, int integrator, int threads, ncpu_handle** out);
int ncpu_step(ncpu_handle* h, int nsteps, double* t);
int ncpu_get_state(ncpu_handle* h, double* pos, double* vel, double* t);
int ncpu_set_state(ncpu_handle* h, const double* pos, const double* vel, double t);
/* accelerations (3n) and potentials (n) at the current positions */
int ncpu_forces(ncpu_handle* h, double* acc, double* pot);
/* Forces from the FP64 fast multipole method instead of the direct sum (the
 * FMM of OpenClSimulation, open_cl_simulation.h:194-198, with the reference's
 * traversal and opening criterion interaction.cl:42-92): Taylor order 1..8,
 * M2L truncation q (0: order + 1) and term set (0: every term of degree <= q;
 * 1: beyond degree p only those with min(|k|, |n|) <= q - p), opening ratio
 * theta (NODE_APPROX_RATIO, interaction.cl:3-5; num/den with num, den <= 64),
 * leaf capacity.  order 0 switches back to the direct sum. */
int ncpu_set_fmm(ncpu_handle* h, int order, int m2l_order, int m2l_terms, double theta, int leaf_capacity);
/* statistics of the last FMM evaluation: nodes, leaves, depth, unordered M2L
 * pairs, unordered P2P pairs, visited cell pairs */
int ncpu_fmm_stats(ncpu_handle* h, long long* out6);
/* seconds per phase of the last FMM evaluation: tree, walk, upward pass,
 * M2L, downward pass, leaves (L2P + P2P) */
int ncpu_fmm_times(ncpu_handle* h, double* out6);
void ncpu_destroy(ncpu_handle* h);
const char* ncpu_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
