// simulation.h -- the backend interface of the nbody simulations.
//
// Same shape as the reference's nbody::Simulation (include/nbody/simulation.h:
// 6-37): Scalar/Vector aliases, a Particle record {position, velocity, mass,
// charge} with a four-argument constructor, step() returning the new time and
// particles() returning a copy.  Adds the virtual destructor the reference
// lacks and includes <vector> itself.
#ifndef NBODY_SIMULATION_H_
#define NBODY_SIMULATION_H_

#include <vector>

namespace nbody {

template <typename TScalar, typename TVector>
class Simulation {
public:
	using Scalar = TScalar;
	using Vector = TVector;

	struct Particle final {
		Vector position;
		Vector velocity;
		Scalar mass;
		Scalar charge;

		Particle(Vector position, Vector velocity, Scalar mass, Scalar charge) :
				position(position), velocity(velocity), mass(mass), charge(charge) {
		}
	};

	virtual ~Simulation() = default;
	virtual Scalar step() = 0;
	virtual std::vector<Particle> particles() const = 0;
};

}  // namespace nbody

#endif
