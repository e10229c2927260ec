// nbody_cli.cpp -- command-line driver over nbody::HipSimulation (or the
// multi-core nbody::CpuSimulation: --backend cpu for its direct sum, cpu-fmm
// for its FP64 fast multipole method with the same theta, order and leaf
// capacity options).
//
// Plays the role of the reference's src/main.cpp: builds the initial
// particles, steps the simulation to a final time and appends one CSV row
// "t,x0,y0,z0,x1,y1,z1,..." per step (main.cpp:77-95, the format
// script/animation.py reads), with seeded initial conditions instead of
// std::rand().  Usage:
//   nbody_cli [--ic uniform|plummer] [--n N] [--seed S] [--dt DT] [--t-end T]
//             [--csv FILE] [--theta TH] [--order P] [--leaf-cap C] [--eps E]
//             [--backend hip|cpu|cpu-fmm]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "nbody/cpu_simulation.h"
#include "nbody/hip_simulation.h"

// the backend is picked at run time; both implement the same interface
using Simulation = nbody::Simulation<float, nbody::hip_vector_t>;
using HipSim = nbody::HipSimulation;
using CpuSim = nbody::CpuSimulation<float, nbody::hip_vector_t>;

int main(int argc, char** argv) {
	std::string ic = "uniform", csv, backend = "hip";
	std::size_t n = 1000;
	unsigned long long seed = 0;
	double dt = 1e-3, t_end = 0.01;
	nbh_config cfg = HipSim::defaultConfig();
	for (int i = 1; i + 1 < argc; i += 2) {
		std::string a = argv[i];
		const char* v = argv[i + 1];
		if (a == "--ic") ic = v;
		else if (a == "--n") n = std::strtoull(v, nullptr, 10);
		else if (a == "--seed") seed = std::strtoull(v, nullptr, 10);
		else if (a == "--dt") dt = std::atof(v);
		else if (a == "--t-end") t_end = std::atof(v);
		else if (a == "--csv") csv = v;
		else if (a == "--theta") cfg.theta = std::atof(v);
		else if (a == "--order") cfg.order = (unsigned)std::atoi(v);
		else if (a == "--leaf-cap") cfg.leaf_capacity = (unsigned)std::atoi(v);
		else if (a == "--eps") cfg.eps = std::atof(v);
		else if (a == "--backend") backend = v;
		else {
			std::cerr << "unknown option " << a << "\n";
			return 2;
		}
	}
	try {
		std::vector<double> pos(3 * n), vel(3 * n), mass(n);
		int rc = ic == "plummer" ? nbh_ic_plummer(seed, n, 1.0, 1.0, 1.0, 20.0, pos.data(), vel.data(), mass.data())
		                         : nbh_ic_uniform(seed, n, pos.data(), vel.data(), mass.data());
		if (rc != NBH_OK) throw std::runtime_error("initial conditions failed");
		std::vector<Simulation::Particle> particles;
		particles.reserve(n);
		for (std::size_t i = 0; i < n; ++i) {
			Simulation::Vector x{(float)pos[3 * i], (float)pos[3 * i + 1], (float)pos[3 * i + 2], 0.f};
			Simulation::Vector v{(float)vel[3 * i], (float)vel[3 * i + 1], (float)vel[3 * i + 2], 0.f};
			particles.emplace_back(x, v, (float)mass[i], (float)mass[i]);
		}
		Simulation::Vector bounds{1.f, 1.f, 1.f, 0.f};
		std::unique_ptr<Simulation> sim;
		if (backend == "cpu" || backend == "cpu-fmm") {
			nbody::CpuOptions opt;
			opt.G = cfg.G;
			opt.eps = cfg.eps;
			if (backend == "cpu-fmm") {
				opt.method = 1;
				opt.fmm.order = (int)cfg.order;
				opt.fmm.theta = cfg.theta;
				opt.fmm.leaf_capacity = (int)cfg.leaf_capacity;
			}
			sim.reset(new CpuSim(bounds, particles, (float)dt, std::cout, opt));
		} else if (backend == "hip") {
			sim.reset(new HipSim(bounds, particles, (float)dt, std::cout, cfg));
		} else {
			std::cerr << "unknown backend " << backend << "\n";
			return 2;
		}
		std::ofstream out;
		if (!csv.empty()) out.open(csv);
		Simulation::Scalar t = 0;
		const long steps = (long)(t_end / dt + 0.5);
		for (long s = 0; s < steps; ++s) {
			t = sim->step();
			if (out) {
				std::vector<Simulation::Particle> ps = sim->particles();
				out << t;
				for (auto const& p : ps) out << "," << p.position[0] << "," << p.position[1] << "," << p.position[2];
				out << "\n";
			}
		}
		std::cout << "Finished at t=" << t << " after " << steps << " steps.\n";
	} catch (std::exception const& e) {
		std::cerr << "error: " << e.what() << "\n";
		return 1;
	}
	return 0;
}
