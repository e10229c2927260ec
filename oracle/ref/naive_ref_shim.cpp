// naive_ref_shim.cpp -- C entry point around the reference's NaiveSimulation,
// compiled together with /root/reference/src/naive_simulation.cpp into
// oracle/_ref/libnaive_ref.so (see oracle/Makefile).  Test infrastructure only:
// used to write golden vectors (tests/golden/) and as the timed CPU baseline.
//
// Particle layout of the flat float array (12 floats per particle) mirrors
// include/nbody/simulation.h:16-32 with vector_t = cl_float4
// (include/nbody/device/types.h:40-64): position xyzw, velocity xyzw, mass,
// charge, 2 pad.
#include <cstddef>
#include <vector>

#include "nbody/naive_simulation.h"

extern "C" int ref_naive_run(std::size_t n, float* p, float force_constant,
                             float dt, int nsteps, float* time_out) {
	using Sim = nbody::NaiveSimulation;
	std::vector<Sim::Particle> ps;
	ps.reserve(n);
	for (std::size_t i = 0; i < n; ++i) {
		const float* q = p + 12 * i;
		Sim::Vector x = {q[0], q[1], q[2], q[3]};
		Sim::Vector v = {q[4], q[5], q[6], q[7]};
		ps.push_back(Sim::Particle(x, v, q[8], q[9]));
	}
	Sim sim(ps, force_constant, dt);
	float t = 0;
	for (int s = 0; s < nsteps; ++s) t = sim.step();
	std::vector<Sim::Particle> out = sim.particles();
	for (std::size_t i = 0; i < n; ++i) {
		float* q = p + 12 * i;
		for (int k = 0; k < 4; ++k) q[k] = out[i].position[k];
		for (int k = 0; k < 4; ++k) q[4 + k] = out[i].velocity[k];
		q[8] = out[i].mass;
		q[9] = out[i].charge;
	}
	if (time_out) *time_out = t;
	return 0;
}
