// cpu_capi.cpp -- the C ABI of include/nbody_cpu.h over nbody::CpuEngine.
#include <algorithm>
#include <exception>
#include <string>

#include "nbody/cpu_simulation.h"
#include "nbody_cpu.h"

struct ncpu_handle {
	nbody::CpuEngine engine;
};

namespace {
thread_local std::string g_error;

int fail(const char* what) {
	g_error = what;
	return -1;
}
}  // namespace

extern "C" {

int ncpu_create(size_t n, const double* pos, const double* vel, const double* mass, const double* charge,
                double dt, double G, double eps, int integrator, int threads, ncpu_handle** out) {
	if (!out || (n && (!pos || !vel || !mass))) return fail("pos, vel, mass and out are required");
	if (integrator < 0 || integrator > 2) return fail("integrator must be 0 (KDK), 1 (Euler) or 2 (kick-drift)");
	try {
		nbody::CpuOptions opt;
		opt.G = G;
		opt.eps = eps;
		opt.integrator = integrator;
		opt.threads = threads;
		*out = new ncpu_handle{nbody::CpuEngine(n, pos, vel, mass, charge, dt, opt)};
	} catch (std::exception const& e) {
		return fail(e.what());
	}
	return 0;
}

int ncpu_step(ncpu_handle* h, int nsteps, double* t) {
	if (!h || nsteps < 0) return fail("invalid handle or step count");
	double tt = h->engine.time();
	for (int s = 0; s < nsteps; ++s) tt = h->engine.step();
	if (t) *t = tt;
	return 0;
}

int ncpu_get_state(ncpu_handle* h, double* pos, double* vel, double* t) {
	if (!h) return fail("invalid handle");
	if (pos) std::copy(h->engine.x().begin(), h->engine.x().end(), pos);
	if (vel) std::copy(h->engine.v().begin(), h->engine.v().end(), vel);
	if (t) *t = h->engine.time();
	return 0;
}

int ncpu_set_state(ncpu_handle* h, const double* pos, const double* vel, double t) {
	if (!h || !pos || !vel) return fail("handle, pos and vel are required");
	h->engine.set_state(pos, vel, t);
	return 0;
}

int ncpu_forces(ncpu_handle* h, double* acc, double* pot) {
	if (!h) return fail("invalid handle");
	h->engine.forces();
	if (acc) std::copy(h->engine.a().begin(), h->engine.a().end(), acc);
	if (pot) std::copy(h->engine.pot().begin(), h->engine.pot().end(), pot);
	return 0;
}

void ncpu_destroy(ncpu_handle* h) { delete h; }

const char* ncpu_last_error(void) { return g_error.c_str(); }

}  // extern "C"
