// ics.cpp -- seeded initial conditions (host side, no GPU needed).
//
// The reference draws its initial conditions from unseeded std::rand()
// (main.cpp:16-66, 118-122), which cannot be reproduced.  These generators are
// counter-free and seeded: SplitMix64 (Steele, Lea & Flood 2014) gives the
// uniform stream, and uniforms are (x >> 11) * 2^-53 in [0, 1).
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../../include/nbody_hip.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double open01() {  // (0, 1)
        double u;
        do { u = uniform(); } while (u <= 0.0);
        return u;
    }
};

void isotropic(SplitMix64& r, double mag, double* out) {
    const double c = 2.0 * r.uniform() - 1.0;          // cos(theta)
    const double ph = 2.0 * M_PI * r.uniform();
    const double s = std::sqrt(std::fmax(0.0, 1.0 - c * c));
    out[0] = mag * s * std::cos(ph);
    out[1] = mag * s * std::sin(ph);
    out[2] = mag * c;
}

}  // namespace

extern "C" int nbh_ic_uniform(uint64_t seed, size_t n, double* pos, double* vel, double* mass) {
    if (!pos || n == 0) return NBH_E_INVALID;
    SplitMix64 r(seed);
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = r.uniform();
    if (vel)
        for (size_t i = 0; i < 3 * n; ++i) vel[i] = 0.0;
    if (mass)
        for (size_t i = 0; i < n; ++i) mass[i] = 1.0 / (double)n;
    return NBH_OK;
}

// Plummer model (Aarseth, Henon & Wielen 1974, A&A 37, 183): in units
// G = M = a = 1, r = (X^(-2/3) - 1)^(-1/2) from the cumulative mass X,
// truncated at rmax by redrawing; speed v = q v_esc with v_esc =
// sqrt(2) (1 + r^2)^(-1/4) and q drawn by von Neumann rejection from
// g(q) = q^2 (1 - q^2)^(7/2) (maximum < 0.1).  Scaled to (a, M, G).
extern "C" int nbh_ic_plummer(uint64_t seed, size_t n, double a, double M, double G, double rmax, double* pos,
                              double* vel, double* mass) {
    if (!pos || !vel || n == 0 || !(a > 0) || !(M > 0) || !(G > 0) || !(rmax > 0)) return NBH_E_INVALID;
    SplitMix64 r(seed);
    const double rlim = rmax / a;
    const double vscale = std::sqrt(G * M / a);
    for (size_t i = 0; i < n; ++i) {
        double rad;
        do {
            const double X = r.open01();
            rad = 1.0 / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
        } while (!(rad <= rlim));
        isotropic(r, rad * a, pos + 3 * i);
        double q, g;
        do {
            q = r.uniform();
            g = 0.1 * r.uniform();
        } while (g >= q * q * std::pow(1.0 - q * q, 3.5));
        const double v = q * std::sqrt(2.0) * std::pow(1.0 + rad * rad, -0.25);
        isotropic(r, v * vscale, vel + 3 * i);
        if (mass) mass[i] = M / (double)n;
    }
    return NBH_OK;
}
