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
	try {
		for (int s = 0; s < nsteps; ++s) tt = h->engine.step();
	} catch (std::exception const& e) {
		return fail(e.what());
	}
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
	try {
		h->engine.forces();
	} catch (std::exception const& e) {
		return fail(e.what());
	}
	if (acc) std::copy(h->engine.a().begin(), h->engine.a().end(), acc);
	if (pot) std::copy(h->engine.pot().begin(), h->engine.pot().end(), pot);
	return 0;
}

int ncpu_set_fmm(ncpu_handle* h, int order, int m2l_order, int m2l_terms, double theta, int leaf_capacity) {
	if (!h) return fail("invalid handle");
	try {
		if (order == 0) {
			h->engine.set_method(0);
		} else {
			nbody::FmmOptions o;
			o.order = order;
			o.m2l_order = m2l_order;
			if (m2l_terms != 0 && m2l_terms != 1) return fail("ncpu_set_fmm: m2l_terms must be 0 (all) or 1 (low)");
			o.m2l_low_terms = m2l_terms == 1;
			o.theta = theta;
			o.leaf_capacity = leaf_capacity;
			h->engine.set_method(1, o);
		}
	} catch (std::exception const& e) {
		return fail(e.what());
	}
	return 0;
}

int ncpu_fmm_stats(ncpu_handle* h, long long* out6) {
	if (!h || !out6) return fail("handle and output are required");
	const nbody::CpuFmm* f = h->engine.fmm();
	if (!f) return fail("the handle uses the direct sum");
	const nbody::CpuFmm::Stats& s = f->stats();
	out6[0] = (long long)s.nodes;
	out6[1] = (long long)s.leaves;
	out6[2] = (long long)s.depth;
	out6[3] = (long long)s.m2l_unordered;
	out6[4] = (long long)s.p2p_unordered;
	out6[5] = (long long)s.visited;
	return 0;
}

int ncpu_fmm_times(ncpu_handle* h, double* out6) {
	if (!h || !out6) return fail("handle and output are required");
	const nbody::CpuFmm* f = h->engine.fmm();
	if (!f) return fail("the handle uses the direct sum");
	const nbody::CpuFmm::Stats& s = f->stats();
	const double t[6] = {s.t_tree, s.t_walk, s.t_up, s.t_m2l, s.t_down, s.t_leaves};
	std::copy(t, t + 6, out6);
	return 0;
}

void ncpu_destroy(ncpu_handle* h) { delete h; }

const char* ncpu_last_error(void) { return g_error.c_str(); }

}  // extern "C"
