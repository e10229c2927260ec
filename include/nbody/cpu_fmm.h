// cpu_fmm.h -- FP64 fast multipole force evaluation on the CPU (OpenMP): the
// FMM method of nbody::CpuSimulation.
//
// The method of the HIP library's force pipeline, on host cores:
//   * a Morton-ordered octree over a cube around the particles (21-bit
//     integer coordinates per axis, x the highest bit of each octal digit); a
//     cell splits while it holds more than leaf_capacity particles, which
//     replaces the host octree of open_cl_simulation.cpp:43-47;
//   * the reference's dual-tree traversal (find_interactions,
//     interaction.cl:42-92, seeded with the root pair as in
//     open_cl_simulation.h:127-133) over unordered cell pairs, with its
//     opening criterion 0.75 (d_a + d_b)^2 / |c_b - c_a|^2 < theta^2
//     (interaction.cl:63-79, NODE_APPROX_RATIO interaction.cl:3-5) evaluated
//     exactly in integers for theta = num/den;
//   * Cartesian Taylor expansions of order p about the cell centres: P2M
//     (moment.cl), M2M, M2L (terms of total degree <= p, and up to degree q
//     those with min(|k|, |n|) <= q - p), L2L, L2P, and the
//     softened pair sum of field.cl:17-32 for the near field.
// Everything is FP64.  The tree and the interaction lists are the HIP
// library's for the same positions; forces differ from it by its FP32
// arithmetic only.  Parallel over cells and leaves with OpenMP when compiled
// with -fopenmp.
#ifndef NBODY_CPU_FMM_H_
#define NBODY_CPU_FMM_H_

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#if defined(_OPENMP)
#include <omp.h>
#if defined(__GLIBCXX__)
#include <parallel/algorithm>  // libstdc++'s parallel sort
#define NBODY_FMM_PARALLEL_SORT 1
#endif
#endif

namespace nbody {

struct FmmOptions {
	int order = 4;           // Taylor order p, 1..8
	int m2l_order = 0;       // M2L truncation q, p <= q <= 2p (0: p + 1): terms of degree <= q
	double theta = 0.5;      // opening ratio; must equal num/den with num, den <= 64
	int leaf_capacity = 32;  // particles per leaf cell
	bool m2l_low_terms = false;  // beyond degree p only the M2L terms with min(|k|, |n|) <= q - p
};

namespace fmm_detail {

constexpr int BITS = 21;  // integer coordinate bits per axis
constexpr int MAX_DEPTH = BITS;

inline int ncoef(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

// coefficient index of the multi-index (x, y, z): by total degree, then by
// descending x, then by descending y
inline int mindex(int x, int y, int z) {
	const int d = x + y + z, a = y + z;
	return d * (d + 1) * (d + 2) / 6 + a * (a + 1) / 2 + (a - y);
}

inline uint64_t spread3(uint32_t v) {
	uint64_t x = v & 0x1fffffu;
	x = (x | x << 32) & 0x1f00000000ffffull;
	x = (x | x << 16) & 0x1f0000ff0000ffull;
	x = (x | x << 8) & 0x100f00f00f00f00full;
	x = (x | x << 4) & 0x10c30c30c30c30c3ull;
	x = (x | x << 2) & 0x1249249249249249ull;
	return x;
}

inline bool rational(double theta, int* num, int* den) {
	for (int d = 1; d <= 64; ++d) {
		const double nf = theta * d;
		const double r = std::floor(nf + 0.5);
		if (r >= 1 && r <= 64 && std::fabs(nf - r) <= 1e-12 * d) {
			*num = (int)r;
			*den = d;
			return true;
		}
	}
	return false;
}

// Index tables of one (p, q): the multi-indices up to degree q, the monomial
// recurrence, and the term lists of the translations.
struct Tables {
	int p = 0, q = 0, nc = 0, nq = 0;
	std::vector<int> mx, my, mz, deg;
	std::vector<int> mono_prev, mono_axis;  // mono[k] = mono[prev] * a[axis] / exponent
	std::vector<double> mono_div;
	std::vector<int> m2l_k, m2l_n, m2l_s;   // L_k += D_s (-1)^|n| M_n, grouped by k
	std::vector<int> m2l_kb;                // terms of k: [m2l_kb[k], m2l_kb[k + 1])
	std::vector<int> m2m_n, m2m_m, m2m_r;   // M'_n += M_m d^r / r!
	std::vector<int> l2l_j, l2l_m, l2l_s;   // L'_j += L_s d^m / m!   (s = j + m)
	std::vector<int> grad_s[3];             // index of k + e_axis (or -1)
	std::vector<double> fact;               // x! y! z!

	// low: beyond degree p only the terms with min(|k|, |n|) <= q - p
	Tables(int p_, int q_, bool low) : p(p_), q(q_), nc(ncoef(p_)), nq(ncoef(q_)) {
		for (int d = 0; d <= q; ++d)
			for (int x = d; x >= 0; --x)
				for (int y = d - x; y >= 0; --y) {
					mx.push_back(x);
					my.push_back(y);
					mz.push_back(d - x - y);
					deg.push_back(d);
				}
		fact.resize(nq);
		mono_prev.assign(nq, 0);
		mono_axis.assign(nq, 0);
		mono_div.assign(nq, 1.0);
		for (int k = 0; k < nq; ++k) {
			auto f = [](int v) {
				double r = 1;
				for (int i = 2; i <= v; ++i) r *= i;
				return r;
			};
			fact[k] = f(mx[k]) * f(my[k]) * f(mz[k]);
			if (k == 0) continue;
			if (mx[k] > 0) {
				mono_prev[k] = mindex(mx[k] - 1, my[k], mz[k]);
				mono_axis[k] = 0;
				mono_div[k] = 1.0 / mx[k];
			} else if (my[k] > 0) {
				mono_prev[k] = mindex(mx[k], my[k] - 1, mz[k]);
				mono_axis[k] = 1;
				mono_div[k] = 1.0 / my[k];
			} else {
				mono_prev[k] = mindex(mx[k], my[k], mz[k] - 1);
				mono_axis[k] = 2;
				mono_div[k] = 1.0 / mz[k];
			}
		}
		for (int k = 0; k < nc; ++k) {
			m2l_kb.push_back((int)m2l_k.size());
			for (int n = 0; n < nc; ++n)
				if (deg[k] + deg[n] <= p ||
				    (deg[k] + deg[n] <= q && (!low || std::min(deg[k], deg[n]) <= q - p))) {
					m2l_k.push_back(k);
					m2l_n.push_back(n);
					m2l_s.push_back(mindex(mx[k] + mx[n], my[k] + my[n], mz[k] + mz[n]));
				}
		}
		m2l_kb.push_back((int)m2l_k.size());
		for (int n = 0; n < nc; ++n)
			for (int a = 0; a <= mx[n]; ++a)
				for (int b = 0; b <= my[n]; ++b)
					for (int c = 0; c <= mz[n]; ++c) {
						m2m_n.push_back(n);
						m2m_m.push_back(mindex(a, b, c));
						m2m_r.push_back(mindex(mx[n] - a, my[n] - b, mz[n] - c));
					}
		for (int j = 0; j < nc; ++j)
			for (int m = 0; m < nc; ++m)
				if (deg[j] + deg[m] <= p) {
					l2l_j.push_back(j);
					l2l_m.push_back(m);
					l2l_s.push_back(mindex(mx[j] + mx[m], my[j] + my[m], mz[j] + mz[m]));
				}
		for (int a = 0; a < 3; ++a) {
			grad_s[a].assign(nc, -1);
			for (int k = 0; k < nc; ++k)
				if (deg[k] < p)
					grad_s[a][k] = mindex(mx[k] + (a == 0), my[k] + (a == 1), mz[k] + (a == 2));
		}
	}

	// out[k] = v^k / k! for k < count
	void monomials(const double v[3], double* out, int count) const {
		out[0] = 1.0;
		for (int k = 1; k < count; ++k) out[k] = out[mono_prev[k]] * v[mono_axis[k]] * mono_div[k];
	}

	// D[k] = d^k f(R), f = (|R|^2 + eps^2)^(-1/2), k < nq.  With a_n = D_n / n!
	// the Taylor coefficients of f satisfy (s = |R|^2 + eps^2)
	//   s |n| a_n = -(2|n| - 1) sum_i R_i a_{n - e_i} - (|n| - 1) sum_i a_{n - 2 e_i}.
	void derivs(const double R[3], double eps2, double* D) const {
		const double s = R[0] * R[0] + R[1] * R[1] + R[2] * R[2] + eps2;
		const double inv_s = 1.0 / s;
		D[0] = 1.0 / std::sqrt(s);
		for (int k = 1; k < nq; ++k) {
			const int e[3] = {mx[k], my[k], mz[k]};
			const int d = deg[k];
			double acc1 = 0, acc2 = 0;
			for (int a = 0; a < 3; ++a) {
				if (e[a] >= 1) {
					int f[3] = {e[0], e[1], e[2]};
					f[a] -= 1;
					acc1 += R[a] * D[mindex(f[0], f[1], f[2])];
				}
				if (e[a] >= 2) {
					int f[3] = {e[0], e[1], e[2]};
					f[a] -= 2;
					acc2 += D[mindex(f[0], f[1], f[2])];
				}
			}
			D[k] = -((2 * d - 1) * acc1 + (d - 1) * acc2) * inv_s / d;
		}
		for (int k = 0; k < nq; ++k) D[k] *= fact[k];
	}
};

// One batch of four M2L sources in SIMD layout: acc[4k + j] += sum over the
// terms (n, s) of k of Db[4s + j] Mb[4n + j].  The four lanes are independent
// sums, so the compiler maps them onto vector registers.
#define NBODY_FMM_CONTRACT4_BODY                                                   \
	const int nc = T.nc;                                                           \
	for (int k = 0; k < nc; ++k) {                                                 \
		double a0 = 0, a1 = 0, a2 = 0, a3 = 0;                                     \
		for (int t = T.m2l_kb[k]; t < T.m2l_kb[k + 1]; ++t) {                      \
			const double* d = Db + 4 * T.m2l_s[t];                                 \
			const double* m = Mb + 4 * T.m2l_n[t];                                 \
			a0 += d[0] * m[0];                                                     \
			a1 += d[1] * m[1];                                                     \
			a2 += d[2] * m[2];                                                     \
			a3 += d[3] * m[3];                                                     \
		}                                                                          \
		acc[4 * k] += a0;                                                          \
		acc[4 * k + 1] += a1;                                                      \
		acc[4 * k + 2] += a2;                                                      \
		acc[4 * k + 3] += a3;                                                      \
	}

inline void contract4_generic(const Tables& T, const double* Db, const double* Mb, double* acc) {
	NBODY_FMM_CONTRACT4_BODY
}

#if defined(__x86_64__) && defined(__GNUC__)
// the same loop compiled for AVX2 + FMA, chosen at run time where the CPU has them
__attribute__((target("avx2,fma"))) inline void contract4_avx2(const Tables& T, const double* Db,
                                                               const double* Mb, double* acc) {
	NBODY_FMM_CONTRACT4_BODY
}
inline bool have_avx2() {
	static const bool v = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
	return v;
}
#endif

inline void contract4(const Tables& T, const double* Db, const double* Mb, double* acc) {
#if defined(__x86_64__) && defined(__GNUC__)
	if (have_avx2()) {
		contract4_avx2(T, Db, Mb, acc);
		return;
	}
#endif
	contract4_generic(T, Db, Mb, acc);
}
#undef NBODY_FMM_CONTRACT4_BODY

// CSR of directed entries (target -> sources) from unordered pairs, in pair
// order per target (the same for any thread count): each thread counts and
// then fills the entries of its own slice of the pairs.
inline void build_csr(const std::vector<std::pair<int32_t, int32_t>>& pairs, std::size_t nn,
                      std::vector<int64_t>& off, std::vector<int32_t>& src) {
	int nt = 1;
#if defined(_OPENMP)
	nt = omp_get_max_threads();
#endif
	const std::size_t np = pairs.size(), chunk = (np + nt - 1) / nt;
	std::vector<int64_t> cnt((std::size_t)nt * nn, 0);
#if defined(_OPENMP)
#pragma omp parallel for schedule(static, 1)
#endif
	for (int t = 0; t < nt; ++t) {
		int64_t* c = &cnt[(std::size_t)t * nn];
		for (std::size_t i = (std::size_t)t * chunk; i < std::min(np, (std::size_t)(t + 1) * chunk); ++i) {
			++c[pairs[i].first];
			if (pairs[i].second != pairs[i].first) ++c[pairs[i].second];
		}
	}
	off.assign(nn + 1, 0);
	for (std::size_t v = 0; v < nn; ++v) {
		int64_t pos = off[v];
		for (int t = 0; t < nt; ++t) {
			const int64_t c = cnt[(std::size_t)t * nn + v];
			cnt[(std::size_t)t * nn + v] = pos;
			pos += c;
		}
		off[v + 1] = pos;
	}
	src.resize((std::size_t)off[nn]);
#if defined(_OPENMP)
#pragma omp parallel for schedule(static, 1)
#endif
	for (int t = 0; t < nt; ++t) {
		int64_t* c = &cnt[(std::size_t)t * nn];
		for (std::size_t i = (std::size_t)t * chunk; i < std::min(np, (std::size_t)(t + 1) * chunk); ++i) {
			src[(std::size_t)c[pairs[i].first]++] = pairs[i].second;
			if (pairs[i].second != pairs[i].first) src[(std::size_t)c[pairs[i].second]++] = pairs[i].first;
		}
	}
}

}  // namespace fmm_detail

class CpuFmm {
public:
	struct Stats {
		std::size_t nodes = 0, leaves = 0, depth = 0;
		std::size_t m2l_unordered = 0, p2p_unordered = 0, visited = 0;
		// seconds: tree (keys, sort, cells), walk (lists), upward (P2M, M2M),
		// M2L (with the offset tables), downward (L2L), leaves (L2P + P2P)
		double t_tree = 0, t_walk = 0, t_up = 0, t_m2l = 0, t_down = 0, t_leaves = 0;
	};

	explicit CpuFmm(FmmOptions opt = FmmOptions()) : _opt(opt) {
		if (opt.order < 1 || opt.order > 8) throw std::runtime_error("CpuFmm: order must be in 1..8");
		_q = opt.m2l_order ? opt.m2l_order : opt.order + 1;
		if (_q < opt.order || _q > 2 * opt.order) throw std::runtime_error("CpuFmm: m2l_order must be in [p, 2p]");
		if (!(opt.theta > 0) || !fmm_detail::rational(opt.theta, &_num, &_den))
			throw std::runtime_error("CpuFmm: theta must be > 0 and equal num/den with num, den <= 64");
		if (opt.leaf_capacity < 1) throw std::runtime_error("CpuFmm: leaf_capacity must be >= 1");
		_tab.reset(new fmm_detail::Tables(opt.order, _q, opt.m2l_low_terms));
	}

	// grad (3n, AoS) and phi (n) of Phi(x_i) = sum_{j != i} q_j / sqrt(|x_j - x_i|^2 + eps^2):
	// grad_i = sum_j q_j (x_j - x_i) / s^(3/2) (far field from the expansions).
	void evaluate(std::size_t n, const double* x, const double* q, double eps, double* grad, double* phi) {
		_stats = Stats();
		if (n == 0) return;
		auto t0 = now();
		build_tree(n, x);
		gather(n, x, q);
		auto t1 = now();
		walk();
		auto t2 = now();
		_stats.t_tree = t1 - t0;
		_stats.t_walk = t2 - t1;
		expansions(eps);
		auto t3 = now();
		near_and_far(eps, grad, phi);
		_stats.t_leaves = now() - t3;
	}

	const Stats& stats() const { return _stats; }
	const FmmOptions& options() const { return _opt; }

private:
	struct Node {
		int32_t level, ix, iy, iz, begin, count, first_child, nchild;
	};

	static double now() {
		return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
	}

	// ---- tree
	void build_tree(std::size_t n, const double* x) {
		using namespace fmm_detail;
		double lo[3] = {x[0], x[1], x[2]}, hi[3] = {x[0], x[1], x[2]};
		for (std::size_t i = 0; i < n; ++i)
			for (int a = 0; a < 3; ++a) {
				const double v = x[3 * i + a];
				if (!std::isfinite(v)) throw std::runtime_error("CpuFmm: non-finite position");
				lo[a] = std::min(lo[a], v);
				hi[a] = std::max(hi[a], v);
			}
		double size = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
		if (!(size > 0)) size = 1.0;
		for (int a = 0; a < 3; ++a) _lo[a] = lo[a];
		_size = size;
		const double inv = std::ldexp(1.0, BITS) / size, qmax = std::ldexp(1.0, BITS) - 1.0;
		std::vector<std::pair<uint64_t, uint32_t>> kp(n);
#if defined(_OPENMP)
#pragma omp parallel for schedule(static)
#endif
		for (long long ii = 0; ii < (long long)n; ++ii) {
			const std::size_t i = (std::size_t)ii;
			uint32_t c[3];
			for (int a = 0; a < 3; ++a) c[a] = (uint32_t)std::min(std::floor((x[3 * i + a] - lo[a]) * inv), qmax);
			kp[i] = std::make_pair((spread3(c[0]) << 2) | (spread3(c[1]) << 1) | spread3(c[2]), (uint32_t)i);
		}
#if defined(NBODY_FMM_PARALLEL_SORT)
		__gnu_parallel::sort(kp.begin(), kp.end());
#else
		std::sort(kp.begin(), kp.end());
#endif
		_keys.resize(n);
		_perm.resize(n);
		for (std::size_t i = 0; i < n; ++i) {
			_keys[i] = kp[i].first;
			_perm[i] = kp[i].second;
		}
		// breadth-first: a level's cells in Morton order, children contiguous
		_nodes.clear();
		_level_off.assign(1, 0);
		_nodes.push_back(Node{0, 0, 0, 0, 0, (int32_t)n, -1, 0});
		const int cap = _opt.leaf_capacity;
		for (int l = 0;; ++l) {
			const std::size_t b = _level_off[l], e = _nodes.size();
			_level_off.push_back(e);
			if (l == MAX_DEPTH) break;
			for (std::size_t c = b; c < e; ++c) {
				if (_nodes[c].count <= cap) continue;
				const int shift = 3 * (BITS - l - 1);
				const Node par = _nodes[c];
				const uint64_t base = _keys[par.begin] >> (shift + 3) << (shift + 3);
				int32_t i0 = par.begin;
				const int32_t end = par.begin + par.count;
				_nodes[c].first_child = (int32_t)_nodes.size();
				for (uint64_t d = 0; d < 8 && i0 < end; ++d) {
					const uint64_t bound = base + ((d + 1) << shift);
					const int32_t i1 =
					    d == 7 ? end : (int32_t)(std::lower_bound(&_keys[i0], &_keys[0] + end, bound) - &_keys[0]);
					if (i1 > i0) {
						_nodes.push_back(Node{l + 1, 2 * par.ix + (int32_t)((d >> 2) & 1),
						                      2 * par.iy + (int32_t)((d >> 1) & 1), 2 * par.iz + (int32_t)(d & 1), i0,
						                      i1 - i0, -1, 0});
						++_nodes[c].nchild;
					}
					i0 = i1;
				}
			}
			if (_nodes.size() == e) break;
		}
		_depth = (int)_level_off.size() - 2;
		_stats.nodes = _nodes.size();
		_stats.depth = (std::size_t)_depth;
		for (const Node& nd : _nodes) _stats.leaves += nd.nchild == 0;
	}

	// integer opening criterion: 3 den^2 (d_a + d_b)^2 < 4 num^2 |c_b - c_a|^2,
	// centres and sides in units of half the finest cell
	bool approx(int a, int b) const {
		using namespace fmm_detail;
		if (a == b) return false;
		const Node& A = _nodes[a];
		const Node& B = _nodes[b];
		const int sa = BITS - A.level, sb = BITS - B.level;
		const int64_t dx = ((2 * (int64_t)B.ix + 1) << sb) - ((2 * (int64_t)A.ix + 1) << sa);
		const int64_t dy = ((2 * (int64_t)B.iy + 1) << sb) - ((2 * (int64_t)A.iy + 1) << sa);
		const int64_t dz = ((2 * (int64_t)B.iz + 1) << sb) - ((2 * (int64_t)A.iz + 1) << sa);
		typedef unsigned __int128 u128;
		const u128 dsq = (u128)(dx * dx) + (u128)(dy * dy) + (u128)(dz * dz);
		const u128 side = ((u128)1 << (sa + 1)) + ((u128)1 << (sb + 1));
		return (u128)3 * (u128)(_den * _den) * side * side < (u128)4 * (u128)(_num * _num) * dsq;
	}

	// ---- dual-tree walk, breadth-first over unordered pairs
	void walk() {
		std::vector<std::pair<int32_t, int32_t>> frontier(1, std::make_pair(0, 0)), m2l, p2p;
		std::size_t visited = 0;
		while (!frontier.empty()) {
			visited += frontier.size();
			const long long nf = (long long)frontier.size();
			int nt = 1;
#if defined(_OPENMP)
			nt = omp_get_max_threads();
#endif
			std::vector<std::vector<std::pair<int32_t, int32_t>>> nx(nt), mm(nt), pp(nt);
#if defined(_OPENMP)
#pragma omp parallel for schedule(static)
#endif
			for (long long f = 0; f < nf; ++f) {
				int t = 0;
#if defined(_OPENMP)
				t = omp_get_thread_num();
#endif
				const int a = frontier[(std::size_t)f].first, b = frontier[(std::size_t)f].second;
				const Node& na = _nodes[a];
				const Node& nb = _nodes[b];
				const int ka = na.nchild ? na.nchild : 1, kb = nb.nchild ? nb.nchild : 1;
				for (int ib = 0; ib < kb; ++ib)
					for (int ia = 0; ia < ka; ++ia) {
						if (a == b && ib > ia) continue;  // interaction.cl:57
						const int A = na.nchild ? na.first_child + ia : a;
						const int B = nb.nchild ? nb.first_child + ib : b;
						if (approx(A, B)) mm[t].push_back(std::make_pair(A, B));
						else if (_nodes[A].nchild || _nodes[B].nchild) nx[t].push_back(std::make_pair(A, B));
						else pp[t].push_back(std::make_pair(A, B));
					}
			}
			frontier.clear();
			for (int t = 0; t < nt; ++t) {
				frontier.insert(frontier.end(), nx[t].begin(), nx[t].end());
				m2l.insert(m2l.end(), mm[t].begin(), mm[t].end());
				p2p.insert(p2p.end(), pp[t].begin(), pp[t].end());
			}
		}
		_stats.visited = visited;
		_stats.m2l_unordered = m2l.size();
		_stats.p2p_unordered = p2p.size();
		fmm_detail::build_csr(m2l, _nodes.size(), _m2l_off, _m2l_src);
		fmm_detail::build_csr(p2p, _nodes.size(), _p2p_off, _p2p_src);
	}

	// particles in Morton order
	void gather(std::size_t n, const double* x, const double* q) {
		_xs.resize(3 * n);
		_qs.resize(n);
		for (std::size_t i = 0; i < n; ++i) {
			const std::size_t j = _perm[i];
			for (int a = 0; a < 3; ++a) _xs[3 * i + a] = x[3 * j + a];
			_qs[i] = q[j];
		}
	}

	void centre(int node, double c[3]) const {
		const Node& nd = _nodes[node];
		const double h = std::ldexp(_size, -nd.level);
		c[0] = _lo[0] + (nd.ix + 0.5) * h;
		c[1] = _lo[1] + (nd.iy + 0.5) * h;
		c[2] = _lo[2] + (nd.iz + 0.5) * h;
	}

	// ---- P2M, M2M, M2L, L2L
	void expansions(double eps) {
		const fmm_detail::Tables& T = *_tab;
		const int nc = T.nc, nq = T.nq;
		const std::size_t nn = _nodes.size();
		_M.assign(nn * nc, 0.0);
		_L.assign(nn * nc, 0.0);
		const double ts = now();
		// P2M at the leaves, M2M upwards level by level
		for (int l = _depth; l >= 0; --l) {
			const long long b = (long long)_level_off[l], e = (long long)_level_off[l + 1];
#if defined(_OPENMP)
#pragma omp parallel for schedule(dynamic, 16)
#endif
			for (long long c = b; c < e; ++c) {
				const Node& nd = _nodes[(std::size_t)c];
				double* M = &_M[(std::size_t)c * nc];
				double z[3];
				centre((int)c, z);
				std::vector<double> mono(nc);
				if (nd.nchild == 0) {
					for (int i = nd.begin; i < nd.begin + nd.count; ++i) {
						const double v[3] = {_xs[3 * i] - z[0], _xs[3 * i + 1] - z[1], _xs[3 * i + 2] - z[2]};
						T.monomials(v, mono.data(), nc);
						for (int k = 0; k < nc; ++k) M[k] += _qs[i] * mono[k];
					}
				} else {
					for (int ch = nd.first_child; ch < nd.first_child + nd.nchild; ++ch) {
						double zc[3];
						centre(ch, zc);
						const double v[3] = {zc[0] - z[0], zc[1] - z[1], zc[2] - z[2]};
						T.monomials(v, mono.data(), nc);
						const double* Mc = &_M[(std::size_t)ch * nc];
						for (std::size_t t = 0; t < T.m2m_n.size(); ++t) M[T.m2m_n[t]] += Mc[T.m2m_m[t]] * mono[T.m2m_r[t]];
					}
				}
			}
		}
		const double tm = now();
		_stats.t_up = tm - ts;
		// same-level pairs: one tensor per integer offset and level, R = o h_l
		const double eps2 = eps * eps;
		std::vector<int> kmax(_depth + 1, -1);
#if defined(_OPENMP)
#pragma omp parallel
#endif
		{
			std::vector<int> km(_depth + 1, -1);
#if defined(_OPENMP)
#pragma omp for schedule(dynamic, 64) nowait
#endif
			for (long long a = 0; a < (long long)nn; ++a) {
				const Node& A = _nodes[(std::size_t)a];
				for (int64_t e = _m2l_off[(std::size_t)a]; e < _m2l_off[(std::size_t)a + 1]; ++e) {
					const Node& B = _nodes[(std::size_t)_m2l_src[(std::size_t)e]];
					if (A.level != B.level) continue;
					const int m =
					    std::max(std::abs(A.ix - B.ix), std::max(std::abs(A.iy - B.iy), std::abs(A.iz - B.iz)));
					km[A.level] = std::max(km[A.level], m);
				}
			}
#if defined(_OPENMP)
#pragma omp critical
#endif
			for (int l = 0; l <= _depth; ++l) kmax[l] = std::max(kmax[l], km[l]);
		}
		std::vector<std::vector<double>> dtab(_depth + 1);
		for (int l = 0; l <= _depth; ++l) {
			const int K = kmax[l];
			if (K < 0) continue;
			const int W = 2 * K + 1;
			dtab[l].assign((std::size_t)W * W * W * nq, 0.0);
			const double h = std::ldexp(_size, -l);
#if defined(_OPENMP)
#pragma omp parallel for schedule(static)
#endif
			for (long long o = 0; o < (long long)W * W * W; ++o) {
				const int ox = (int)(o / (W * W)) - K, oy = (int)(o / W % W) - K, oz = (int)(o % W) - K;
				const double R[3] = {ox * h, oy * h, oz * h};
				T.derivs(R, eps2, &dtab[l][(std::size_t)o * nq]);
			}
		}
		// M2L: per target, sources in SIMD-friendly batches of four
#if defined(_OPENMP)
#pragma omp parallel for schedule(dynamic, 8)
#endif
		for (long long a = 0; a < (long long)nn; ++a) {
			const int64_t e0 = _m2l_off[(std::size_t)a], e1 = _m2l_off[(std::size_t)a + 1];
			if (e0 == e1) continue;
			const Node& A = _nodes[(std::size_t)a];
			double zA[3];
			centre((int)a, zA);
			std::vector<double> Db((std::size_t)nq * 4), Mb((std::size_t)nc * 4), acc((std::size_t)nc * 4, 0.0), Dt(nq);
			for (int64_t e = e0; e < e1; e += 4) {
				for (int j = 0; j < 4; ++j) {
					if (e + j >= e1) {
						for (int s = 0; s < nq; ++s) Db[(std::size_t)s * 4 + j] = 0.0;
						for (int k = 0; k < nc; ++k) Mb[(std::size_t)k * 4 + j] = 0.0;
						continue;
					}
					const int b = _m2l_src[(std::size_t)(e + j)];
					const Node& B = _nodes[(std::size_t)b];
					const double* D;
					if (B.level == A.level) {
						const int K = kmax[A.level], W = 2 * K + 1;
						const int o = ((A.ix - B.ix + K) * W + (A.iy - B.iy + K)) * W + (A.iz - B.iz + K);
						D = &dtab[A.level][(std::size_t)o * nq];
					} else {
						double zB[3];
						centre(b, zB);
						const double R[3] = {zA[0] - zB[0], zA[1] - zB[1], zA[2] - zB[2]};
						T.derivs(R, eps2, Dt.data());
						D = Dt.data();
					}
					for (int s = 0; s < nq; ++s) Db[(std::size_t)s * 4 + j] = D[s];
					const double* M = &_M[(std::size_t)b * nc];
					for (int k = 0; k < nc; ++k) Mb[(std::size_t)k * 4 + j] = (T.deg[k] & 1) ? -M[k] : M[k];
				}
				fmm_detail::contract4(T, Db.data(), Mb.data(), acc.data());
			}
			double* L = &_L[(std::size_t)a * nc];
			for (int k = 0; k < nc; ++k) L[k] = acc[(std::size_t)k * 4] + acc[(std::size_t)k * 4 + 1] +
			                                   acc[(std::size_t)k * 4 + 2] + acc[(std::size_t)k * 4 + 3];
		}
		const double td = now();
		_stats.t_m2l = td - tm;
		// L2L downwards level by level
		for (int l = 0; l < _depth; ++l) {
			const long long b = (long long)_level_off[l], e = (long long)_level_off[l + 1];
#if defined(_OPENMP)
#pragma omp parallel for schedule(dynamic, 16)
#endif
			for (long long c = b; c < e; ++c) {
				const Node& nd = _nodes[(std::size_t)c];
				if (!nd.nchild) continue;
				double z[3];
				centre((int)c, z);
				std::vector<double> mono(nc);
				const double* Lp = &_L[(std::size_t)c * nc];
				for (int ch = nd.first_child; ch < nd.first_child + nd.nchild; ++ch) {
					double zc[3];
					centre(ch, zc);
					const double v[3] = {zc[0] - z[0], zc[1] - z[1], zc[2] - z[2]};
					T.monomials(v, mono.data(), nc);
					double* Lc = &_L[(std::size_t)ch * nc];
					for (std::size_t t = 0; t < T.l2l_j.size(); ++t) Lc[T.l2l_j[t]] += Lp[T.l2l_s[t]] * mono[T.l2l_m[t]];
				}
			}
		}
		_stats.t_down = now() - td;
	}

	// ---- L2P and P2P per leaf; results in input order
	void near_and_far(double eps, double* grad, double* phi) {
		const fmm_detail::Tables& T = *_tab;
		const int nc = T.nc;
		const double eps2 = eps * eps;
		const long long nn = (long long)_nodes.size();
#if defined(_OPENMP)
#pragma omp parallel for schedule(dynamic, 8)
#endif
		for (long long a = 0; a < nn; ++a) {
			const Node& A = _nodes[(std::size_t)a];
			if (A.nchild) continue;
			double z[3];
			centre((int)a, z);
			std::vector<double> mono(nc);
			const double* L = &_L[(std::size_t)a * nc];
			for (int i = A.begin; i < A.begin + A.count; ++i) {
				const double xi = _xs[3 * i], yi = _xs[3 * i + 1], zi = _xs[3 * i + 2];
				const double v[3] = {xi - z[0], yi - z[1], zi - z[2]};
				T.monomials(v, mono.data(), nc);
				double ph = 0, gx = 0, gy = 0, gz = 0;
				for (int k = 0; k < nc; ++k) {
					ph += L[k] * mono[k];
					if (T.grad_s[0][k] >= 0) {
						gx += L[T.grad_s[0][k]] * mono[k];
						gy += L[T.grad_s[1][k]] * mono[k];
						gz += L[T.grad_s[2][k]] * mono[k];
					}
				}
				for (int64_t e = _p2p_off[(std::size_t)a]; e < _p2p_off[(std::size_t)a + 1]; ++e) {
					const Node& B = _nodes[(std::size_t)_p2p_src[(std::size_t)e]];
					for (int j = B.begin; j < B.begin + B.count; ++j) {
						if (j == i) continue;
						const double rx = _xs[3 * j] - xi, ry = _xs[3 * j + 1] - yi, rz = _xs[3 * j + 2] - zi;
						const double inv = 1.0 / std::sqrt(rx * rx + ry * ry + rz * rz + eps2);
						const double w = _qs[j] * inv;
						const double w3 = w * inv * inv;
						gx += w3 * rx;
						gy += w3 * ry;
						gz += w3 * rz;
						ph += w;
					}
				}
				const std::size_t o = _perm[(std::size_t)i];
				grad[3 * o] = gx;
				grad[3 * o + 1] = gy;
				grad[3 * o + 2] = gz;
				phi[o] = ph;
			}
		}
	}

	FmmOptions _opt;
	int _q = 5, _num = 1, _den = 2;
	std::shared_ptr<const fmm_detail::Tables> _tab;
	double _lo[3] = {0, 0, 0}, _size = 1;
	std::vector<uint64_t> _keys;
	std::vector<uint32_t> _perm;
	std::vector<Node> _nodes;
	std::vector<std::size_t> _level_off;
	int _depth = 0;
	std::vector<int64_t> _m2l_off, _p2p_off;
	std::vector<int32_t> _m2l_src, _p2p_src;
	std::vector<double> _xs, _qs, _M, _L;
	Stats _stats;
};

}  // namespace nbody

#endif
