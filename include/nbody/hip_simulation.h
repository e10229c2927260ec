// hip_simulation.h -- nbody::HipSimulation, the MI355X FMM backend.
//
// Drops in for OpenClSimulation (include/nbody/open_cl_simulation.h:194-198):
// the constructor keeps the reference's (bounds, particles, timeStep, log)
// prefix, step() and particles() keep their meaning, and failures throw
// std::runtime_error as the reference does (open_cl_simulation.cpp:630,648,
// 767).  Header-only over the C ABI in nbody_hip.h; link -lnbody_hip.
//
// Differences from OpenClSimulation, by design: particles() returns the input
// order (the reference returns octree order, open_cl_simulation.cpp:53-68),
// the root cube is the particles' bounding cube each step unless
// cfg.fixed_root is set (then [0, max(bounds))), and the integrator defaults
// to KDK leapfrog (NBH_INT_EULER reproduces open_cl_simulation.cpp:600-607).
#ifndef NBODY_HIP_SIMULATION_H_
#define NBODY_HIP_SIMULATION_H_

#include <algorithm>
#include <cstddef>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbody/simulation.h"
#include "nbody_hip.h"

namespace nbody {

// Four floats with an array subscript, laid out like the reference's
// device::vector_t over cl_float4 (include/nbody/device/types.h:40-64).
struct alignas(16) hip_vector_t {
	float v[4] = {0, 0, 0, 0};
	hip_vector_t() = default;
	hip_vector_t(std::initializer_list<float> list) {
		std::size_t i = 0;
		for (float f : list) {
			if (i < 4) v[i++] = f;
		}
	}
	float& operator[](unsigned i) { return v[i]; }
	float const& operator[](unsigned i) const { return v[i]; }
};

class HipSimulation final : public Simulation<float, hip_vector_t> {
public:
	static nbh_config defaultConfig() {
		nbh_config c;
		nbh_config_default(&c);
		return c;
	}

	HipSimulation(Vector bounds, std::vector<Particle> particles, Scalar timeStep, std::ostream& log,
	              nbh_config config = defaultConfig()) :
			_log(log), _n(particles.size()) {
		config.dt = timeStep;
		if (config.fixed_root) {
			config.root_size = std::max(bounds[0], std::max(bounds[1], bounds[2]));
		}
		std::vector<double> pos(3 * _n), vel(3 * _n), mass(_n), charge(_n);
		for (std::size_t i = 0; i < _n; ++i) {
			for (int k = 0; k < 3; ++k) {
				pos[3 * i + k] = particles[i].position[k];
				vel[3 * i + k] = particles[i].velocity[k];
			}
			mass[i] = particles[i].mass;
			charge[i] = particles[i].charge;
		}
		int rc = nbh_create(&config, _n, pos.data(), vel.data(), mass.data(), charge.data(), &_h);
		if (rc != NBH_OK) {
			throw std::runtime_error(std::string("HipSimulation: ") + nbh_last_error(nullptr));
		}
		_log << "HipSimulation: " << _n << " particles on HIP device " << config.device << ".\n";
	}

	// No `override`: when the reference's include/ comes first on the include
	// path, "nbody/simulation.h" is the reference's header, whose Simulation
	// has no virtual destructor (simulation.h:6-37).  Over this repository's
	// header the destructor is virtual all the same.
	~HipSimulation() { nbh_destroy(_h); }
	HipSimulation(HipSimulation const&) = delete;
	HipSimulation& operator=(HipSimulation const&) = delete;

	Scalar step() override {
		double t = 0;
		check(nbh_step(_h, 1, &t), "step");
		return static_cast<Scalar>(t);
	}

	std::vector<Particle> particles() const override {
		std::vector<double> pos(3 * _n), vel(3 * _n), mass(_n), charge(_n);
		check(nbh_get_state(_h, pos.data(), vel.data(), mass.data(), charge.data(), nullptr), "particles");
		std::vector<Particle> out;
		out.reserve(_n);
		for (std::size_t i = 0; i < _n; ++i) {
			Vector x{(float)pos[3 * i], (float)pos[3 * i + 1], (float)pos[3 * i + 2], 0.f};
			Vector v{(float)vel[3 * i], (float)vel[3 * i + 1], (float)vel[3 * i + 2], 0.f};
			out.emplace_back(x, v, (Scalar)mass[i], (Scalar)charge[i]);
		}
		return out;
	}

	nbh_handle* handle() const { return _h; }

private:
	void check(int rc, const char* what) const {
		if (rc != NBH_OK) {
			throw std::runtime_error(std::string("HipSimulation::") + what + ": " + nbh_last_error(_h));
		}
	}

	std::ostream& _log;
	std::size_t _n;
	nbh_handle* _h = nullptr;
};

}  // namespace nbody

#endif
