// cpu_simulation.h -- nbody::CpuSimulation, a multi-core CPU backend behind the
// reference's Simulation interface (include/nbody/simulation.h:6-37).
//
// It drops in beside NaiveSimulation (naive_simulation.h:25-33) and
// OpenClSimulation (open_cl_simulation.h:194-198): same particle type, step()
// returns the new time, particles() returns a copy in input order.  Forces are
// the FP64 direct sum of Plummer-softened gravity with the sign convention of
// field.cl:17-32 (a_i = (q_i/m_i) G sum_j q_j (x_j - x_i) / (|x_j - x_i|^2 +
// eps^2)^(3/2)), parallel over targets with OpenMP when compiled with
// -fopenmp.  State is FP64; the interface's float Particle is converted on the
// way in and out.  Integrators as the HIP backend: KDK leapfrog (default),
// kick-drift (naive_simulation.cpp:28-41) or explicit Euler
// (open_cl_simulation.cpp:600-607).  With method = 1 the forces come from the
// FP64 fast multipole method of nbody/cpu_fmm.h (the HIP library's tree,
// lists and expansions) instead of the direct sum.
#ifndef NBODY_CPU_SIMULATION_H_
#define NBODY_CPU_SIMULATION_H_

#include <cmath>
#include <cstddef>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <vector>

#if defined(_OPENMP)
#include <omp.h>
#endif

#include "nbody/cpu_fmm.h"
#include "nbody/simulation.h"

namespace nbody {

struct CpuOptions {
	double G = 1.0;       // reference FORCE_CONSTANT = -G (field.cl:7-9)
	double eps = 0.01;    // reference PARTICLE_RADIUS (field.cl:3-5)
	int integrator = 0;   // 0 KDK, 1 explicit Euler, 2 kick-drift
	int threads = 0;      // OpenMP threads, 0 = runtime default
	int method = 0;       // 0 direct sum, 1 fast multipole method
	FmmOptions fmm;       // with method = 1
};

// The FP64 engine, independent of the particle type (used by the C ABI too).
class CpuEngine {
public:
	CpuEngine(std::size_t n, const double* pos, const double* vel, const double* mass, const double* charge,
	          double dt, CpuOptions opt) :
			_n(n), _dt(dt), _opt(opt), _x(pos, pos + 3 * n), _v(vel, vel + 3 * n), _m(mass, mass + n),
			_q(charge ? std::vector<double>(charge, charge + n) : std::vector<double>(mass, mass + n)),
			_a(3 * n, 0.0), _pot(n, 0.0) {
		if (!(dt > 0) || !(opt.eps >= 0)) throw std::runtime_error("CpuSimulation: dt > 0 and eps >= 0 required");
		for (std::size_t i = 0; i < n; ++i) {
			if (!(_m[i] > 0)) throw std::runtime_error("CpuSimulation: masses must be positive");
		}
		set_method(opt.method, opt.fmm);
	}

	// switches the force method (0 direct sum, 1 FMM with the given options)
	void set_method(int method, FmmOptions fmm = FmmOptions()) {
		if (method == 1) _fmm = std::make_shared<CpuFmm>(fmm);
		else if (method == 0) _fmm.reset();
		else throw std::runtime_error("CpuSimulation: method must be 0 (direct) or 1 (FMM)");
		_opt.method = method;
		_opt.fmm = fmm;
		_acc_valid = false;
	}

	const CpuFmm* fmm() const { return _fmm.get(); }

	// accelerations and potentials at the current positions
	void forces() {
		if (_fmm) {
			forces_fmm();
			return;
		}
		const std::size_t n = _n;
		const double e2 = _opt.eps * _opt.eps, G = _opt.G;
		const double* x = _x.data();
		const double* q = _q.data();
		double* a = _a.data();
		double* pot = _pot.data();
		const double* m = _m.data();
#if defined(_OPENMP)
		if (_opt.threads > 0) omp_set_num_threads(_opt.threads);
#pragma omp parallel for schedule(static)
#endif
		for (long long ii = 0; ii < (long long)n; ++ii) {
			const std::size_t i = (std::size_t)ii;
			const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
			double ax = 0, ay = 0, az = 0, ph = 0;
			for (std::size_t j = 0; j < n; ++j) {
				if (j == i) continue;
				const double rx = x[3 * j] - xi, ry = x[3 * j + 1] - yi, rz = x[3 * j + 2] - zi;
				const double inv = 1.0 / std::sqrt(rx * rx + ry * ry + rz * rz + e2);
				const double w = q[j] * inv * inv * inv;
				ax += w * rx;
				ay += w * ry;
				az += w * rz;
				ph += q[j] * inv;
			}
			const double f = G * q[i] / m[i];
			a[3 * i] = f * ax;
			a[3 * i + 1] = f * ay;
			a[3 * i + 2] = f * az;
			pot[i] = -G * ph;
		}
		_acc_valid = true;
	}

	void forces_fmm() {
#if defined(_OPENMP)
		if (_opt.threads > 0) omp_set_num_threads(_opt.threads);
#endif
		_fmm->evaluate(_n, _x.data(), _q.data(), _opt.eps, _a.data(), _pot.data());
		const double G = _opt.G;
		for (std::size_t i = 0; i < _n; ++i) {
			const double f = G * _q[i] / _m[i];
			_a[3 * i] *= f;
			_a[3 * i + 1] *= f;
			_a[3 * i + 2] *= f;
			_pot[i] *= -G;
		}
		_acc_valid = true;
	}

	double step() {
		const double dt = _dt;
		const std::size_t n3 = 3 * _n;
		switch (_opt.integrator) {
			case 1:  // Euler: x += v dt; v += a(x_old) dt
				forces();
				for (std::size_t k = 0; k < n3; ++k) {
					_x[k] += _v[k] * dt;
					_v[k] += _a[k] * dt;
				}
				_acc_valid = false;
				break;
			case 2:  // kick-drift: v += a dt; x += v dt
				forces();
				for (std::size_t k = 0; k < n3; ++k) {
					_v[k] += _a[k] * dt;
					_x[k] += _v[k] * dt;
				}
				_acc_valid = false;
				break;
			default:  // KDK
				if (!_acc_valid) forces();
				for (std::size_t k = 0; k < n3; ++k) {
					_v[k] += 0.5 * dt * _a[k];
					_x[k] += dt * _v[k];
				}
				forces();
				for (std::size_t k = 0; k < n3; ++k) _v[k] += 0.5 * dt * _a[k];
				break;
		}
		_time += dt;
		return _time;
	}

	std::size_t size() const { return _n; }
	double time() const { return _time; }
	const std::vector<double>& x() const { return _x; }
	const std::vector<double>& v() const { return _v; }
	const std::vector<double>& m() const { return _m; }
	const std::vector<double>& q() const { return _q; }
	const std::vector<double>& a() const { return _a; }
	const std::vector<double>& pot() const { return _pot; }
	void set_state(const double* pos, const double* vel, double t) {
		_x.assign(pos, pos + 3 * _n);
		_v.assign(vel, vel + 3 * _n);
		_time = t;
		_acc_valid = false;
	}

private:
	std::size_t _n;
	double _dt;
	CpuOptions _opt;
	std::vector<double> _x, _v, _m, _q, _a, _pot;
	double _time = 0;
	bool _acc_valid = false;
	std::shared_ptr<CpuFmm> _fmm;
};

template <typename TScalar, typename TVector>
class CpuSimulation final : public Simulation<TScalar, TVector> {
public:
	using Base = Simulation<TScalar, TVector>;
	using typename Base::Particle;
	using Scalar = TScalar;
	using Vector = TVector;

	// OpenClSimulation's (bounds, particles, timeStep, log) prefix; bounds are
	// not needed by a direct sum and only logged.
	CpuSimulation(Vector bounds, std::vector<Particle> particles, Scalar timeStep, std::ostream& log,
	              CpuOptions opt = CpuOptions()) :
			_engine(make_engine(particles, timeStep, opt)) {
		log << "CpuSimulation: " << particles.size() << " particles, box (" << bounds[0] << ", " << bounds[1]
		    << ", " << bounds[2] << ").\n";
	}

	Scalar step() override { return (Scalar)_engine.step(); }

	std::vector<Particle> particles() const override {
		std::vector<Particle> out;
		out.reserve(_engine.size());
		const auto& x = _engine.x();
		const auto& v = _engine.v();
		for (std::size_t i = 0; i < _engine.size(); ++i) {
			Vector p, u;
			for (int k = 0; k < 3; ++k) {
				p[k] = (float)x[3 * i + k];
				u[k] = (float)v[3 * i + k];
			}
			out.push_back(Particle(p, u, (Scalar)_engine.m()[i], (Scalar)_engine.q()[i]));
		}
		return out;
	}

	CpuEngine& engine() { return _engine; }

private:
	static CpuEngine make_engine(const std::vector<Particle>& ps, Scalar dt, CpuOptions opt) {
		const std::size_t n = ps.size();
		std::vector<double> x(3 * n), v(3 * n), m(n), q(n);
		for (std::size_t i = 0; i < n; ++i) {
			for (int k = 0; k < 3; ++k) {
				x[3 * i + k] = ps[i].position[k];
				v[3 * i + k] = ps[i].velocity[k];
			}
			m[i] = ps[i].mass;
			q[i] = ps[i].charge;
		}
		return CpuEngine(n, x.data(), v.data(), m.data(), q.data(), dt, opt);
	}

	CpuEngine _engine;
};

}  // namespace nbody

#endif
