// SuperTerrain+ heightfield hot path -- CPU ORACLE (test infrastructure only).
//
// This file is the checker, never the product. Only the tests, the smoke check and the
// CPU-baseline leg of the benchmark may load it. It restates the reference algorithm in
// single-threaded C++, templated on the arithmetic type:
//   * Real = float  : the same FP32 operation order as the reference, no FMA contraction
//                     (build with -ffp-contract=off). Used for bit-exact comparisons.
//   * Real = double : an FP64 restatement with the same constants, permutation table and
//                     raindrop sequence. Used for the 1e-5 / RMS tolerance comparisons.
//
// Every function cites the reference file:line it follows (paths relative to the reference
// checkout root). Pinning of each function (what checks it besides this restatement) is
// listed in the header comment of tests/test_oracle.py.
//
// Third-party semantics restated here (not in the reference tree):
//   glm::mix(x, y, a) = x * (1 - a) + y * a ; glm::length(v) = sqrt(v.x*v.x + v.y*v.y)
//   (glm, version unpinned by the reference's CMake). std::shuffle is libstdc++'s.
//   Philox4x32-10 is written from Salmon et al., "Parallel random numbers: as easy as
//   1, 2, 3", SC'11 (not cuRAND's stream layout: parity with cuRAND's drop positions is
//   unpinned by design).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

// ------------------------------------------------------------------------------------
// Permutation + gradient table
// SuperTerrain+/SuperAlgorithm+/Host/Private/STPPermutationGenerator.cpp:40-94
// ------------------------------------------------------------------------------------
template <class Engine>
void shuffleTable(const uint64_t seed, uint8_t* const perm) {
	Engine rng(seed);
	for (int i = 0; i < 256; i++) {
		perm[i] = static_cast<uint8_t>(i);
	}
	std::shuffle(perm, perm + 256, rng);
	std::copy(perm, perm + 256, perm + 256);
}

template <class Real>
void gradientTable(const uint32_t distribution, const double offset, Real* const grad) {
	// STPPermutationGenerator.cpp:64-73. The angle ACCUMULATES by `step` (not k*step); the
	// counter bound stops the one-past-the-end write the reference can make for some N.
	const double TwoPi = 2.0 * 3.14159265358979323846264338327950288;
	const double step = TwoPi / (distribution * 1.0);
	double angle = 0.0;
	for (uint32_t counter = 0u; angle < TwoPi && counter < distribution; angle += step, counter++) {
		const double offset_angle = angle + offset;
		grad[counter * 2u] = static_cast<Real>(std::cos(offset_angle));
		grad[counter * 2u + 1u] = static_cast<Real>(std::sin(offset_angle));
	}
}

// ------------------------------------------------------------------------------------
// Simplex noise
// SuperTerrain+/SuperAlgorithm+/Device/Private/STPSimplexNoise.cu:5-110
// ------------------------------------------------------------------------------------
template <class Real> struct SimplexConst;
template <> struct SimplexConst<float> {
	static constexpr float F2 = 0.3660254038f, G2 = 0.2113248654f, H2 = -1.0f + 2.0f * 0.2113248654f;
};
template <> struct SimplexConst<double> {
	static constexpr double F2 = 0.3660254038, G2 = 0.2113248654, H2 = -1.0 + 2.0 * 0.2113248654;
};

// STPSimplexNoise.cu:14-16 -- NOT floor: floori(0) == -1 and floori(-1.0) == -2.
template <class Real>
inline int floori(const Real v) {
	return v > Real(0) ? static_cast<int>(v) : static_cast<int>(v - Real(1));
}

template <class Real>
struct Permutation {
	const uint8_t* perm;
	const Real* grad;
	uint32_t n;
};

// STPSimplexNoise.cu:18-82
template <class Real>
Real simplex2D(const Permutation<Real>& p, const Real x, const Real y) {
	using C = SimplexConst<Real>;
	Real dx[3], dy[3];
	const Real s = (x + y) * C::F2;
	const int i = floori<Real>(x + s), j = floori<Real>(y + s);
	const Real t = static_cast<Real>(i + j) * C::G2;
	const Real X0 = static_cast<Real>(i) - t, Y0 = static_cast<Real>(j) - t;
	dx[0] = x - X0;
	dy[0] = y - Y0;
	const unsigned int i1 = (dx[0] > dy[0]) ? 1u : 0u, j1 = 1u - i1;
	dx[1] = dx[0] - static_cast<Real>(i1) + C::G2;
	dy[1] = dy[0] - static_cast<Real>(j1) + C::G2;
	dx[2] = dx[0] + C::H2;
	dy[2] = dy[0] + C::H2;

	const unsigned int ii = static_cast<unsigned int>(i) & 255u, jj = static_cast<unsigned int>(j) & 255u;
	unsigned int g[3];
	g[0] = p.perm[ii + p.perm[jj]] % p.n;
	g[1] = p.perm[ii + i1 + p.perm[jj + j1]] % p.n;
	g[2] = p.perm[ii + 1u + p.perm[jj + 1u]] % p.n;

	Real corner[3];
	for (int v = 0; v < 3; v++) {
		Real w = Real(0.5) - dx[v] * dx[v] - dy[v] * dy[v];
		if (w <= Real(0)) {
			corner[v] = Real(0);
		} else {
			w *= w;
			corner[v] = w * w * (p.grad[g[v] * 2u] * dx[v] + p.grad[g[v] * 2u + 1u] * dy[v]);
		}
	}
	return Real(70) * (corner[0] + corner[1] + corner[2]);
}

// STPSimplexNoise.cuh:29-52
template <class Real>
struct Fractal {
	Real Persistence, Lacunarity;
	uint32_t Octave;
	Real Scale, OffsetX, OffsetY, HalfX, HalfY, Amplitude, Frequency;
};

template <class Real>
inline Real saturate(const Real v) {
	return v > Real(1) ? Real(1) : (v > Real(0) ? v : Real(0));
}

// STPSimplexNoise.cu:84-110
template <class Real>
Real simplex2DFractal(const Permutation<Real>& p, const Real x, const Real y, const Fractal<Real>& d) {
	Real fractal = Real(0), amplitude = d.Amplitude, frequency = d.Frequency, absRange = Real(0);
	for (uint32_t o = 0u; o < d.Octave; o++) {
		const Real sx = ((x - d.HalfX) + d.OffsetX) / d.Scale * frequency,
			sy = ((y - d.HalfY) + d.OffsetY) / d.Scale * frequency;
		fractal += simplex2D<Real>(p, sx, sy) * amplitude;
		absRange += amplitude;
		amplitude *= d.Persistence;
		frequency *= d.Lacunarity;
	}
	return saturate<Real>((fractal + absRange) / (Real(2) * absRange));
}

// ------------------------------------------------------------------------------------
// Erosion brush
// SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPHeightfieldGenerator.cpp:44-99
// ------------------------------------------------------------------------------------
template <class Real>
uint32_t erosionBrush(const uint32_t dimensionX, const uint32_t erosion_radius, int32_t* const index,
	Real* const weight, const uint32_t capacity) {
	const int radius = static_cast<int>(erosion_radius), radiusSqr = radius * radius;
	double weightSum = 0.0;
	std::vector<int32_t> idx;
	std::vector<double> w;
	for (int by = -radius; by <= radius; by++) {
		for (int bx = -radius; bx <= radius; bx++) {
			const double sqrDst = static_cast<double>(bx * bx + by * by);
			if (sqrDst < radiusSqr) {
				const double cw = 1.0 - std::sqrt(sqrDst) / radius;
				weightSum += cw;
				idx.push_back(by * static_cast<int>(dimensionX) + bx);
				// the reference stores the un-normalised weight as float before normalising
				// (STPHeightfieldGenerator.cpp:69,73-75); the FP64 restatement keeps it double.
				w.push_back(std::is_same<Real, float>::value ? static_cast<double>(static_cast<float>(cw)) : cw);
			}
		}
	}
	const double factor = 1.0 / weightSum;
	const uint32_t n = static_cast<uint32_t>(idx.size());
	for (uint32_t k = 0u; k < n && k < capacity; k++) {
		index[k] = idx[k];
		weight[k] = static_cast<Real>(w[k] * factor);
	}
	return n;
}

// ------------------------------------------------------------------------------------
// Philox4x32-10, from Salmon, Moraes, Dror, Shaw, SC'11 (Random123 paper).
// ------------------------------------------------------------------------------------
inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
	uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
	uint32_t k0 = key_in[0], k1 = key_in[1];
	for (int r = 0; r < 10; r++) {
		if (r > 0) {
			k0 += 0x9E3779B9u;
			k1 += 0xBB67AE85u;
		}
		const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
		const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
		const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
		const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
		c0 = hi1 ^ c1 ^ k0;
		c1 = lo1;
		c2 = hi0 ^ c3 ^ k1;
		c3 = lo0;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

// Raindrop spawn uniform in (0, 1]: ((word >> 8) + 1) * 2^-24, exact in FP32 and FP64.
template <class Real>
inline Real unitUniform(const uint32_t word) {
	return static_cast<Real>((word >> 8) + 1u) * static_cast<Real>(1.0 / 16777216.0);
}

// ------------------------------------------------------------------------------------
// Hydraulic erosion, serialized (drop d runs to completion before drop d+1 starts).
// SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:80-108 (spawn)
// SuperTerrain+/SuperTerrain+/Private/GPGPU/STPRainDrop.cu:18-164 (drop)
// ------------------------------------------------------------------------------------
struct RainDropSetting {  // the float fields of STPRainDropSetting.h:15-53 (+ MaxLifetime)
	float Inertia, SedimentCapacityFactor, minSedimentCapacity, initWaterVolume, minWaterVolume, Friction,
		initSpeed, ErodeSpeed, DepositSpeed, EvaporateSpeed, Gravity;
	uint32_t ErosionBrushRadius, MaxLifetime;
};

struct ErosionStats {
	uint64_t steps, erode_steps, deposit_steps, killed;
};

template <class Real>
struct Drop {
	Real px, py, dx, dy, speed, volume, sediment;
};

// STPRainDrop.cu:23-51 -> (height, gradX, gradY)
template <class Real>
inline void heightGradient(const Real* map, const uint32_t W, const Real px, const Real py, Real out[3]) {
	const uint32_t rx = static_cast<uint32_t>(px), ry = static_cast<uint32_t>(py);
	const Real cx = px - static_cast<Real>(rx), cy = py - static_cast<Real>(ry);
	const uint32_t b = ry * W + rx;
	const Real hx = map[b], hy = map[b + 1], hz = map[b + W], hw = map[b + W + 1];
	out[0] = hx * (1 - cx) * (1 - cy) + hy * cx * (1 - cy) + hz * (1 - cx) * cy + hw * cx * cy;
	// glm::mix(a, b, t) = a * (1 - t) + b * t
	out[1] = (hy - hx) * (1 - cy) + (hw - hz) * cy;
	out[2] = (hz - hx) * (1 - cx) + (hw - hy) * cx;
}

template <class Real>
void erodeSerial(Real* const map, const uint32_t totalX, const uint32_t totalY, const uint32_t nnX,
	const uint32_t nnY, const uint32_t mapX, const uint32_t mapY, const RainDropSetting& s, const int32_t* bIdx,
	const Real* bW, const uint32_t bN, const uint32_t dropCount, const uint64_t seed, const uint32_t streamId,
	Real* const ledger, ErosionStats* const stats) {
	const Real baseX = static_cast<Real>(mapX - 1u), baseY = static_cast<Real>(mapY - 1u);
	const Real rangeX = static_cast<Real>((nnX / 2u) * mapX), rangeY = static_cast<Real>((nnY / 2u) * mapY);
	const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
	const Real radius = static_cast<Real>(s.ErosionBrushRadius);
	const Real limX = Real(1) * static_cast<Real>(totalX) - radius, limY = Real(1) * static_cast<Real>(totalY) - radius;
	const Real inertia = s.Inertia, capK = s.SedimentCapacityFactor, minCap = s.minSedimentCapacity,
		dep = s.DepositSpeed, ero = s.ErodeSpeed, grav = s.Gravity, fric = s.Friction, evap = s.EvaporateSpeed,
		minWater = s.minWaterVolume;
	ErosionStats st = {0, 0, 0, 0};

	for (uint32_t d = 0u; d < dropCount; d++) {
		const uint32_t ctr[4] = {d, streamId, 0u, 0u};
		uint32_t rnd[4];
		philox4x32_10(ctr, key, rnd);
		Drop<Real> dr;
		dr.px = unitUniform<Real>(rnd[0]) * baseX + rangeX;
		dr.py = unitUniform<Real>(rnd[1]) * baseY + rangeY;
		dr.dx = Real(0);
		dr.dy = Real(0);
		dr.speed = s.initSpeed;
		dr.volume = s.initWaterVolume;
		dr.sediment = Real(0);
		Real lost = Real(0);
		// Spawn guard (build-defined): a spawn whose 2x2 bilinear cell leaves the map dies
		// before its first step (the reference would read out of bounds there).
		if (!(static_cast<uint32_t>(dr.px) + 1u <= totalX - 1u && static_cast<uint32_t>(dr.py) + 1u <= totalY - 1u)) {
			dr.volume = Real(0);
			st.killed++;
		}

		for (uint32_t life = 0u; (s.MaxLifetime == 0u || life < s.MaxLifetime) && dr.volume >= minWater; life++) {
			st.steps++;
			const uint32_t cellX = static_cast<uint32_t>(dr.px), cellY = static_cast<uint32_t>(dr.py);
			const uint32_t mapIndex = cellY * totalX + cellX;
			const Real ox = dr.px - static_cast<Real>(static_cast<int>(dr.px)),
				oy = dr.py - static_cast<Real>(static_cast<int>(dr.py));
			Real hg[3];
			heightGradient<Real>(map, totalX, dr.px, dr.py, hg);
			// Direction = mix(-grad, Direction, Inertia)
			dr.dx = (-hg[1]) * (1 - inertia) + dr.dx * inertia;
			dr.dy = (-hg[2]) * (1 - inertia) + dr.dy * inertia;
			const Real length = std::sqrt(dr.dx * dr.dx + dr.dy * dr.dy);
			if (length != Real(0)) {
				dr.dx /= length;
				dr.dy /= length;
			}
			dr.px += dr.dx;
			dr.py += dr.dy;
			// STPRainDrop.cu:96-105, negated so that a NaN position also dies; plus the footprint
			// guard: the old cell must keep its brush (reach r-1) and 2x2 cell (reach 1) inside.
			const uint32_t reach = s.ErosionBrushRadius > 2u ? s.ErosionBrushRadius - 1u : 1u;
			const bool cellOk = cellX + 1u >= s.ErosionBrushRadius && cellY + 1u >= s.ErosionBrushRadius
				&& cellX + reach <= totalX - 1u && cellY + reach <= totalY - 1u;
			if (!(!(dr.dx == Real(0) && dr.dy == Real(0)) && dr.px >= radius && dr.px < limX && dr.py >= radius
					&& dr.py < limY)
				|| !cellOk) {
				lost = dr.sediment;
				dr.volume = Real(0);
				dr.sediment = Real(0);
				st.killed++;
				break;
			}
			Real nh[3];
			heightGradient<Real>(map, totalX, dr.px, dr.py, nh);
			const Real deltaHeight = nh[0] - hg[0];
			const Real cap = std::fmax(-deltaHeight * dr.speed * dr.volume * capK, minCap);
			if (dr.sediment > cap || deltaHeight > Real(0)) {
				const Real amount = (deltaHeight > Real(0)) ? std::fmin(deltaHeight, dr.sediment)
															: (dr.sediment - cap) * dep;
				dr.sediment -= amount;
				map[mapIndex] += amount * (Real(1) - ox) * (Real(1) - oy);
				map[mapIndex + 1] += amount * ox * (Real(1) - oy);
				map[mapIndex + totalX] += amount * (Real(1) - ox) * oy;
				map[mapIndex + totalX + 1] += amount * ox * oy;
				st.deposit_steps++;
			} else {
				const Real amount = std::fmin((cap - dr.sediment) * ero, -deltaHeight);
				for (uint32_t k = 0u; k < bN; k++) {
					const uint32_t ei = mapIndex + static_cast<uint32_t>(bIdx[k]);
					const Real wa = amount * bW[k];
					const Real ds = (map[ei] < wa) ? map[ei] : wa;
					map[ei] -= ds;
					dr.sediment += ds;
				}
				st.erode_steps++;
			}
			dr.speed = std::sqrt(std::fmax(Real(0), dr.speed * dr.speed + deltaHeight * grav));
			dr.speed *= Real(1) - fric;
			dr.volume *= (Real(1) - evap);
		}
		if (ledger) {
			ledger[d] = lost + dr.sediment;
		}
	}
	if (stats) {
		*stats = st;
	}
}

// ------------------------------------------------------------------------------------
// Format: SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:112-122
// ------------------------------------------------------------------------------------
template <class Real>
void formatHeightmap(const Real* in, uint16_t* out, const uint64_t n) {
	for (uint64_t i = 0; i < n; i++) {
		out[i] = static_cast<uint16_t>(std::nearbyint(saturate<Real>(in[i]) * static_cast<Real>(65535)));
	}
}

// ------------------------------------------------------------------------------------
// Single histogram filter, reference accumulator semantics (insertion-ordered bins).
// SuperTerrain+/SuperAlgorithm+/Host/Private/STPSingleHistogramFilter.cpp:368-902
// ------------------------------------------------------------------------------------
struct Accumulator {
	static constexpr uint32_t NoEntry = 0xFFFFFFFFu;
	std::vector<uint32_t> dict;                     // sample -> index into bins
	std::vector<std::pair<uint16_t, uint32_t>> bins; // (item, count) in insertion order

	// STPSingleHistogramFilter.cpp:400-422
	void increment(const uint16_t sample, const uint32_t count) {
		if (dict.size() <= sample) {
			dict.resize(static_cast<size_t>(sample) + 1u, NoEntry);
		}
		uint32_t& bi = dict[sample];
		if (bi == NoEntry) {
			bins.emplace_back(sample, count);
			bi = static_cast<uint32_t>(bins.size() - 1u);
		} else {
			bins[bi].second += count;
		}
	}
	// STPSingleHistogramFilter.cpp:433-450
	void decrement(const uint16_t sample, const uint32_t count) {
		uint32_t& bi = dict[sample];
		uint32_t& q = bins[bi].second;
		if (q <= count) {
			bins.erase(bins.begin() + bi);
			for (size_t k = bi; k < bins.size(); k++) {
				dict[bins[k].first]--;
			}
			bi = NoEntry;
		} else {
			q -= count;
		}
	}
	void clear() {
		dict.clear();
		bins.clear();
	}
};

struct UBin {
	uint16_t item;
	uint32_t count;
};

// Returns 0 on success, 1 on a domain error (STPSingleHistogramFilter.cpp:874-880).
// The reference derives the window start from the neighbourhood (start = Map * (NN / 2),
// STPSingleHistogramFilter.cpp:881-892); this form takes the start directly so that a
// halo-padded tile (start = (r, r)) can be filtered as well.
template <class Real>
int singleHistogramFilterWindow(const uint16_t* sample, const uint32_t totalX, const uint32_t startX,
	const uint32_t startY, const uint32_t mapX, const uint32_t mapY, const uint32_t radius,
	std::vector<uint16_t>& outItem, std::vector<Real>& outWeight, std::vector<uint32_t>& outOffset) {
	if (!(radius > 0u && (radius & 1u) == 0u)) {
		return 1;
	}
	if (!(radius <= startX && radius <= startY)) {
		return 1;
	}
	const uint32_t wStart = startX - radius, wEnd = wStart + mapX + 2u * radius;
	Accumulator acc;
	// vertical pass, STPSingleHistogramFilter.cpp:522-563; column-major output of height mapY
	std::vector<UBin> vBin;
	std::vector<uint32_t> vOff;
	const auto emitV = [&]() {
		vOff.push_back(static_cast<uint32_t>(vBin.size()));
		for (const auto& b : acc.bins) {
			vBin.push_back(UBin {b.first, b.second});
		}
	};
	for (uint32_t i = wStart; i < wEnd; i++) {
		const uint32_t ti = i + startY * totalX;
		uint32_t ui = ti - radius * totalX, di = ti + (radius + 1u) * totalX;
		for (int j = -static_cast<int>(radius); j <= static_cast<int>(radius); j++) {
			acc.increment(sample[ti + j * static_cast<int>(totalX)], 1u);
		}
		emitV();
		for (uint32_t j = 1u; j < mapY; j++) {
			acc.increment(sample[di], 1u);
			acc.decrement(sample[ui], 1u);
			emitV();
			di += totalX;
			ui += totalX;
		}
		acc.clear();
	}
	vOff.push_back(static_cast<uint32_t>(vBin.size()));

	// horizontal pass, STPSingleHistogramFilter.cpp:608-679; row-major output
	outItem.clear();
	outWeight.clear();
	outOffset.clear();
	const auto addColumn = [&](const uint32_t col, const bool inc) {
		for (uint32_t b = vOff[col]; b < vOff[col + 1u]; b++) {
			if (inc) {
				acc.increment(vBin[b].item, vBin[b].count);
			} else {
				acc.decrement(vBin[b].item, vBin[b].count);
			}
		}
	};
	const auto emitH = [&]() {
		// copyToBuffer<true>, STPSingleHistogramFilter.cpp:483-506
		outOffset.push_back(static_cast<uint32_t>(outItem.size()));
		uint32_t sum = 0u;
		for (const auto& b : acc.bins) {
			sum += b.second;
		}
		const Real norm = Real(1) / static_cast<Real>(sum);
		for (const auto& b : acc.bins) {
			outItem.push_back(b.first);
			outWeight.push_back(static_cast<Real>(b.second) * norm);
		}
	};
	for (uint32_t i = 0u; i < mapY; i++) {
		const uint32_t ti = i + radius * mapY;
		uint32_t li = i, ri = ti + (radius + 1u) * mapY;
		for (int j = -static_cast<int>(radius); j <= static_cast<int>(radius); j++) {
			addColumn(static_cast<uint32_t>(static_cast<int>(ti) + j * static_cast<int>(mapY)), true);
		}
		emitH();
		for (uint32_t j = 1u; j < mapX; j++) {
			addColumn(ri, true);
			addColumn(li, false);
			emitH();
			ri += mapY;
			li += mapY;
		}
		acc.clear();
	}
	outOffset.push_back(static_cast<uint32_t>(outItem.size()));
	return 0;
}

template <class Real>
int singleHistogramFilter(const uint16_t* sample, const uint32_t totalX, const uint32_t nnX, const uint32_t nnY,
	const uint32_t mapX, const uint32_t mapY, const uint32_t radius, std::vector<uint16_t>& outItem,
	std::vector<Real>& outWeight, std::vector<uint32_t>& outOffset) {
	return singleHistogramFilterWindow<Real>(sample, totalX, mapX * (nnX / 2u), mapY * (nnY / 2u), mapX, mapY, radius,
		outItem, outWeight, outOffset);
}

// ------------------------------------------------------------------------------------
// Multi-biome blend: SuperDemo+/Script/STPMultiHeightGenerator.cu:35-71 with
// HalfDimension = float(MapSize) / 2 (SuperDemo+/World/Biomes/STPCommonCompiler.cpp:236).
// ------------------------------------------------------------------------------------
struct BiomeProperty {  // SuperDemo+/World/Biomes/STPBiomeProperty.hpp:10-27
	float Scale;
	uint32_t Octave;
	float Persistence, Lacunarity, Depth, Variation;
};

template <class Real>
void multiBiomeHeightmap(const Permutation<Real>& p, const BiomeProperty* table, const uint16_t* binItem,
	const Real* binWeight, const uint32_t* offset, const uint32_t mapX, const uint32_t mapY, const float offX,
	const float offY, const uint32_t step, Real* out) {
	// `step`: world units per pixel (1 = the reference's sampling); pixel x samples x * step
	const Real halfX = static_cast<Real>(static_cast<float>(mapX * step) / 2.0f),
		halfY = static_cast<Real>(static_cast<float>(mapY * step) / 2.0f);
	for (uint32_t y = 0u; y < mapY; y++) {
		for (uint32_t x = 0u; x < mapX; x++) {
			const uint32_t idx = x + y * mapX;
			Real h = Real(0);
			for (uint32_t b = offset[idx]; b < offset[idx + 1u]; b++) {
				const BiomeProperty& bp = table[binItem[b]];
				Fractal<Real> f;
				f.Persistence = bp.Persistence;
				f.Lacunarity = bp.Lacunarity;
				f.Octave = bp.Octave;
				f.Scale = bp.Scale;
				f.OffsetX = offX;
				f.OffsetY = offY;
				f.HalfX = halfX;
				f.HalfY = halfY;
				f.Amplitude = Real(1);
				f.Frequency = Real(1);
				const Real n = simplex2DFractal<Real>(p, Real(1) * static_cast<Real>(x * step),
					Real(1) * static_cast<Real>(y * step), f)
					* static_cast<Real>(bp.Variation) + static_cast<Real>(bp.Depth);
				h += binWeight[b] * n;
			}
			out[idx] = h;
		}
	}
}

// ------------------------------------------------------------------------------------
// Voronoi biome map (build-defined synthetic input; no reference counterpart).
// Sites: one per cell of a cellSize grid, jittered by Philox(key = seed,
// ctr = (cellX, cellY, 0x564F524F, 0)); biome = word2 % biomeCount; nearest site by
// exact int64 squared distance over the 3x3 neighbouring cells, first wins on ties.
// ------------------------------------------------------------------------------------
inline int64_t floorDiv(const int64_t a, const int64_t b) {
	int64_t q = a / b;
	if ((a % b != 0) && ((a < 0) != (b < 0))) {
		q--;
	}
	return q;
}

uint16_t voronoiBiome(const uint64_t seed, const uint32_t biomeCount, const uint32_t cellSize, const int64_t wx,
	const int64_t wy) {
	const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
	const int64_t cx = floorDiv(wx, cellSize), cy = floorDiv(wy, cellSize);
	int64_t best = INT64_MAX;
	uint16_t biome = 0u;
	for (int dy = -1; dy <= 1; dy++) {
		for (int dx = -1; dx <= 1; dx++) {
			const int64_t ncx = cx + dx, ncy = cy + dy;
			const uint32_t ctr[4] = {static_cast<uint32_t>(ncx), static_cast<uint32_t>(ncy), 0x564F524Fu, 0u};
			uint32_t r[4];
			philox4x32_10(ctr, key, r);
			const int64_t sx = ncx * cellSize + (r[0] % cellSize), sy = ncy * cellSize + (r[1] % cellSize);
			const int64_t d = (wx - sx) * (wx - sx) + (wy - sy) * (wy - sy);
			if (d < best) {
				best = d;
				biome = static_cast<uint16_t>(r[2] % biomeCount);
			}
		}
	}
	return biome;
}

} // namespace

// ======================================================================================
// C entry points (ctypes). All buffers are host memory, caller-owned.
// ======================================================================================
extern "C" {

struct orc_fractal_desc {
	float Persistence, Lacunarity;
	uint32_t Octave;
	float Scale, OffsetX, OffsetY, HalfX, HalfY, Amplitude, Frequency;
};

int orc_permutation(const uint64_t seed, const uint32_t distribution, const double offset, const int engine,
	uint8_t* perm512, float* grad_f32, double* grad_f64) {
	if (distribution == 0u || !(offset >= 0.0) || !(offset < 2.0 * 3.14159265358979323846264338327950288)) {
		return 1; // STPAlgorithmEnvironment.cpp:13-17
	}
	if (engine == 0) {
		shuffleTable<std::mt19937_64>(seed, perm512);
	} else if (engine == 1) {
		shuffleTable<std::mt19937>(seed, perm512);
	} else {
		return 2;
	}
	if (grad_f32) {
		gradientTable<float>(distribution, offset, grad_f32);
	}
	if (grad_f64) {
		gradientTable<double>(distribution, offset, grad_f64);
	}
	return 0;
}

int orc_floori_f32(const float v) {
	return floori<float>(v);
}

float orc_simplex2d_f32(const uint8_t* perm, const float* grad, const uint32_t n, const float x, const float y) {
	return simplex2D<float>(Permutation<float> {perm, grad, n}, x, y);
}

double orc_simplex2d_f64(const uint8_t* perm, const double* grad, const uint32_t n, const double x, const double y) {
	return simplex2D<double>(Permutation<double> {perm, grad, n}, x, y);
}

} // extern "C"

template <class Real>
static void fractalMap(const uint8_t* perm, const Real* grad, const uint32_t n, const orc_fractal_desc* d,
	const uint32_t W, const uint32_t H, Real* out) {
	Fractal<Real> f;
	f.Persistence = d->Persistence;
	f.Lacunarity = d->Lacunarity;
	f.Octave = d->Octave;
	f.Scale = d->Scale;
	f.OffsetX = d->OffsetX;
	f.OffsetY = d->OffsetY;
	f.HalfX = d->HalfX;
	f.HalfY = d->HalfY;
	f.Amplitude = d->Amplitude;
	f.Frequency = d->Frequency;
	const Permutation<Real> p {perm, grad, n};
	for (uint32_t y = 0u; y < H; y++) {
		for (uint32_t x = 0u; x < W; x++) {
			out[x + y * W] = simplex2DFractal<Real>(p, Real(1) * static_cast<Real>(x), Real(1) * static_cast<Real>(y), f);
		}
	}
}

extern "C" {

void orc_fractal_map_f32(const uint8_t* perm, const float* grad, const uint32_t n, const orc_fractal_desc* d,
	const uint32_t W, const uint32_t H, float* out) {
	fractalMap<float>(perm, grad, n, d, W, H, out);
}

void orc_fractal_map_f64(const uint8_t* perm, const double* grad, const uint32_t n, const orc_fractal_desc* d,
	const uint32_t W, const uint32_t H, double* out) {
	fractalMap<double>(perm, grad, n, d, W, H, out);
}

uint32_t orc_brush_f32(const uint32_t dimX, const uint32_t radius, int32_t* idx, float* w, const uint32_t cap) {
	return erosionBrush<float>(dimX, radius, idx, w, cap);
}

uint32_t orc_brush_f64(const uint32_t dimX, const uint32_t radius, int32_t* idx, double* w, const uint32_t cap) {
	return erosionBrush<double>(dimX, radius, idx, w, cap);
}

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
	philox4x32_10(ctr, key, out);
}

struct orc_raindrop_setting {
	float Inertia, SedimentCapacityFactor, minSedimentCapacity, initWaterVolume, minWaterVolume, Friction,
		initSpeed, ErodeSpeed, DepositSpeed, EvaporateSpeed, Gravity;
	uint32_t ErosionBrushRadius, MaxLifetime;
};

void orc_erode_f32(float* map, const uint32_t totalX, const uint32_t totalY, const uint32_t nnX,
	const uint32_t nnY, const uint32_t mapX, const uint32_t mapY, const orc_raindrop_setting* s, const int32_t* bIdx,
	const float* bW, const uint32_t bN, const uint32_t drops, const uint64_t seed, const uint32_t streamId,
	float* ledger, uint64_t* stats4) {
	RainDropSetting rs;
	std::memcpy(&rs, s, sizeof(rs));
	ErosionStats st;
	erodeSerial<float>(map, totalX, totalY, nnX, nnY, mapX, mapY, rs, bIdx, bW, bN, drops, seed, streamId, ledger, &st);
	if (stats4) {
		stats4[0] = st.steps;
		stats4[1] = st.erode_steps;
		stats4[2] = st.deposit_steps;
		stats4[3] = st.killed;
	}
}

void orc_erode_f64(double* map, const uint32_t totalX, const uint32_t totalY, const uint32_t nnX,
	const uint32_t nnY, const uint32_t mapX, const uint32_t mapY, const orc_raindrop_setting* s, const int32_t* bIdx,
	const double* bW, const uint32_t bN, const uint32_t drops, const uint64_t seed, const uint32_t streamId,
	double* ledger, uint64_t* stats4) {
	RainDropSetting rs;
	std::memcpy(&rs, s, sizeof(rs));
	ErosionStats st;
	erodeSerial<double>(map, totalX, totalY, nnX, nnY, mapX, mapY, rs, bIdx, bW, bN, drops, seed, streamId, ledger, &st);
	if (stats4) {
		stats4[0] = st.steps;
		stats4[1] = st.erode_steps;
		stats4[2] = st.deposit_steps;
		stats4[3] = st.killed;
	}
}

void orc_format_f32(const float* in, uint16_t* out, const uint64_t n) {
	formatHeightmap<float>(in, out, n);
}

void orc_format_f64(const double* in, uint16_t* out, const uint64_t n) {
	formatHeightmap<double>(in, out, n);
}

// SHF: outputs are sized by the caller: offset[mapX*mapY+1], item/weight[cap].
// Returns the bin total (>= 0) or -1 on a domain error, -2 if cap is too small.
int64_t orc_shf_f32(const uint16_t* sample, const uint32_t totalX, const uint32_t nnX, const uint32_t nnY,
	const uint32_t mapX, const uint32_t mapY, const uint32_t radius, uint16_t* item, float* weight, uint32_t* offset,
	const uint64_t cap) {
	std::vector<uint16_t> it;
	std::vector<float> w;
	std::vector<uint32_t> off;
	if (singleHistogramFilter<float>(sample, totalX, nnX, nnY, mapX, mapY, radius, it, w, off) != 0) {
		return -1;
	}
	if (it.size() > cap) {
		return -2;
	}
	std::copy(it.begin(), it.end(), item);
	std::copy(w.begin(), w.end(), weight);
	std::copy(off.begin(), off.end(), offset);
	return static_cast<int64_t>(it.size());
}

int64_t orc_shf_window_f32(const uint16_t* sample, const uint32_t totalX, const uint32_t startX, const uint32_t startY,
	const uint32_t mapX, const uint32_t mapY, const uint32_t radius, uint16_t* item, float* weight, uint32_t* offset,
	const uint64_t cap) {
	std::vector<uint16_t> it;
	std::vector<float> w;
	std::vector<uint32_t> off;
	if (singleHistogramFilterWindow<float>(sample, totalX, startX, startY, mapX, mapY, radius, it, w, off) != 0) {
		return -1;
	}
	if (it.size() > cap) {
		return -2;
	}
	std::copy(it.begin(), it.end(), item);
	std::copy(w.begin(), w.end(), weight);
	std::copy(off.begin(), off.end(), offset);
	return static_cast<int64_t>(it.size());
}

void orc_nn_merge(const void* const* chunks, void* merged, const uint32_t elemSize, const uint32_t mapX,
	const uint32_t mapY, const uint32_t nnX, const uint32_t nnY) {
	// STPNearestNeighbourTextureBuffer.cpp:97-110 with STPChunk.cpp:71-73
	const uint32_t totalX = mapX * nnX;
	for (uint32_t i = 0u; i < nnX * nnY; i++) {
		const uint32_t lx = i % nnX, ly = i / nnX;
		for (uint32_t py = 0u; py < mapY; py++) {
			for (uint32_t px = 0u; px < mapX; px++) {
				const uint64_t g = static_cast<uint64_t>(lx * mapX + px) + static_cast<uint64_t>(ly * mapY + py) * totalX;
				std::memcpy(static_cast<char*>(merged) + g * elemSize,
					static_cast<const char*>(chunks[i]) + (static_cast<uint64_t>(px) + py * mapX) * elemSize, elemSize);
			}
		}
	}
}

void orc_nn_unmerge(const void* merged, void* const* chunks, const uint32_t elemSize, const uint32_t mapX,
	const uint32_t mapY, const uint32_t nnX, const uint32_t nnY) {
	const uint32_t totalX = mapX * nnX;
	for (uint32_t i = 0u; i < nnX * nnY; i++) {
		const uint32_t lx = i % nnX, ly = i / nnX;
		for (uint32_t py = 0u; py < mapY; py++) {
			for (uint32_t px = 0u; px < mapX; px++) {
				const uint64_t g = static_cast<uint64_t>(lx * mapX + px) + static_cast<uint64_t>(ly * mapY + py) * totalX;
				std::memcpy(static_cast<char*>(chunks[i]) + (static_cast<uint64_t>(px) + py * mapX) * elemSize,
					static_cast<const char*>(merged) + g * elemSize, elemSize);
			}
		}
	}
}

struct orc_biome_property {
	float Scale;
	uint32_t Octave;
	float Persistence, Lacunarity, Depth, Variation;
};

void orc_multi_biome_f32(const uint8_t* perm, const float* grad, const uint32_t n, const orc_biome_property* table,
	const uint16_t* item, const float* weight, const uint32_t* offset, const uint32_t mapX, const uint32_t mapY,
	const float offX, const float offY, const uint32_t step, float* out) {
	multiBiomeHeightmap<float>(Permutation<float> {perm, grad, n}, reinterpret_cast<const BiomeProperty*>(table), item,
		weight, offset, mapX, mapY, offX, offY, step, out);
}

void orc_multi_biome_f64(const uint8_t* perm, const double* grad, const uint32_t n, const orc_biome_property* table,
	const uint16_t* item, const double* weight, const uint32_t* offset, const uint32_t mapX, const uint32_t mapY,
	const float offX, const float offY, const uint32_t step, double* out) {
	multiBiomeHeightmap<double>(Permutation<double> {perm, grad, n}, reinterpret_cast<const BiomeProperty*>(table),
		item, weight, offset, mapX, mapY, offX, offY, step, out);
}

void orc_voronoi_map(const uint64_t seed, const uint32_t biomeCount, const uint32_t cellSize, const int64_t originX,
	const int64_t originY, const uint32_t W, const uint32_t H, const uint32_t step, uint16_t* out) {
	// pixel (x, y) samples world (origin + x * step, origin + y * step)
	for (uint32_t y = 0u; y < H; y++) {
		for (uint32_t x = 0u; x < W; x++) {
			out[x + static_cast<uint64_t>(y) * W] = voronoiBiome(seed, biomeCount, cellSize,
				originX + static_cast<int64_t>(x) * step, originY + static_cast<int64_t>(y) * step);
		}
	}
}

// Chunk math, SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPChunk.cpp:63-108
void orc_chunk_world_coord(const double px, const double pz, const uint32_t csX, const uint32_t csY,
	const double scX, const double scY, int32_t* out2) {
	const double sx = static_cast<double>(csX) * scX, sy = static_cast<double>(csY) * scY;
	out2[0] = static_cast<int32_t>(std::floor(px / sx)) * static_cast<int32_t>(csX);
	out2[1] = static_cast<int32_t>(std::floor(pz / sy)) * static_cast<int32_t>(csY);
}

void orc_chunk_map_offset(const int32_t cx, const int32_t cy, const uint32_t csX, const uint32_t csY,
	const uint32_t msX, const uint32_t msY, const double moX, const double moY, double* out2) {
	out2[0] = static_cast<double>(cx) / static_cast<double>(csX) * static_cast<double>(msX) + moX;
	out2[1] = static_cast<double>(cy) / static_cast<double>(csY) * static_cast<double>(msY) + moY;
}

// Whole-pipeline CPU throughput baseline: each worker runs noise -> erosion -> format on
// its own chunks (FP32, drops serialized per chunk). Returns wall seconds.
double orc_pipeline_threads(const uint8_t* perm, const float* grad, const uint32_t n, const orc_fractal_desc* d,
	const uint32_t chunkCount, const uint32_t side, const orc_raindrop_setting* s, const uint32_t drops,
	const uint64_t seed, const uint32_t threads) {
	RainDropSetting rs;
	std::memcpy(&rs, s, sizeof(rs));
	int32_t bIdx[1024];
	float bW[1024];
	const uint32_t bN = erosionBrush<float>(side, rs.ErosionBrushRadius, bIdx, bW, 1024u);
	const auto t0 = std::chrono::steady_clock::now();
	std::vector<std::thread> pool;
	for (uint32_t t = 0u; t < threads; t++) {
		pool.emplace_back([&, t]() {
			std::vector<float> map(static_cast<size_t>(side) * side);
			std::vector<uint16_t> fixed(map.size());
			for (uint32_t c = t; c < chunkCount; c += threads) {
				orc_fractal_desc dd = *d;
				dd.OffsetX = static_cast<float>(c) * side;
				fractalMap<float>(perm, grad, n, &dd, side, side, map.data());
				erodeSerial<float>(map.data(), side, side, 1u, 1u, side, side, rs, bIdx, bW, bN, drops, seed, c, nullptr,
					nullptr);
				formatHeightmap<float>(map.data(), fixed.data(), map.size());
			}
		});
	}
	for (auto& th : pool) {
		th.join();
	}
	return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // extern "C"

// ------------------------------------------------------------------------------------
// Splatmap
// SuperDemo+/Script/STPSplatmapGenerator.cu:65-137 (constants :2-18), the rule lookups of
// SuperTerrain+/SuperAlgorithm+/Device/Private/STPTextureSplatRuleWrapper.cu:56-126 and the rule
// order of SuperTerrain+/SuperTerrain+/Private/World/Diversity/STPTextureDatabase.cpp:357-411.
// The heightfield is read as CUDA's linear-filtered, clamped, normalised UINT16 texture at
// unnormalised coordinates (STPWorldPipeline.cpp:682-696): texel centres at +0.5, value / 65535.
// ------------------------------------------------------------------------------------
struct orc_altitude_rule {
	uint16_t Sample;
	float UpperBound;
	uint32_t Region;
};
struct orc_gradient_rule {
	uint16_t Sample;
	float minGradient, maxGradient, LowerBound, UpperBound;
	uint32_t Region;
};
struct orc_splat_chunk {
	uint32_t LocalX, LocalY;
	float OffsetX, OffsetY;
};

static float splatTexel(const uint16_t* const m, const uint32_t W, const uint32_t H, int64_t x, int64_t y) {
	x = std::min<int64_t>(std::max<int64_t>(x, 0), static_cast<int64_t>(W) - 1);
	y = std::min<int64_t>(std::max<int64_t>(y, 0), static_cast<int64_t>(H) - 1);
	return static_cast<float>(m[static_cast<size_t>(y) * W + static_cast<size_t>(x)]) / 65535.0f;
}

static float splatLinear(const uint16_t* const m, const uint32_t W, const uint32_t H, const float u, const float v) {
	const float xb = u - 0.5f, yb = v - 0.5f, fx = std::floor(xb), fy = std::floor(yb);
	const float a = xb - fx, b = yb - fy;
	const int64_t i = static_cast<int64_t>(fx), j = static_cast<int64_t>(fy);
	const float t00 = splatTexel(m, W, H, i, j), t10 = splatTexel(m, W, H, i + 1, j), t01 = splatTexel(m, W, H, i, j + 1),
		t11 = splatTexel(m, W, H, i + 1, j + 1);
	return (1.0f - a) * ((1.0f - b) * t00 + b * t01) + a * ((1.0f - b) * t10 + b * t11);
}

extern "C" {

void orc_splatmap(const uint8_t* perm, const float* grad, const uint32_t n, const orc_altitude_rule* alt,
	const uint32_t altCount, const orc_gradient_rule* gra, const uint32_t graCount, const uint16_t* biome,
	const uint16_t* height, uint8_t* splat, const uint32_t W, const uint32_t H, const uint32_t mapX, const uint32_t mapY,
	const orc_splat_chunk* chunks, const uint32_t count) {
	constexpr float GradientBias = 5.5f, KernelRadius = 2.5f, NoiseScale = 91.5f, NoiseContribution = 0.15f,
		HeightFactor = 875.5f;
	constexpr uint32_t NoRegion = 0xFFFFFFFFu;
	// the database order: altitude rules by (sample, upper bound), gradient rules by sample
	std::vector<orc_altitude_rule> a(alt, alt + altCount);
	std::vector<orc_gradient_rule> g(gra, gra + graCount);
	std::stable_sort(a.begin(), a.end(), [](const orc_altitude_rule& p, const orc_altitude_rule& q) {
		return p.Sample != q.Sample ? p.Sample < q.Sample : p.UpperBound < q.UpperBound;
	});
	std::stable_sort(g.begin(), g.end(), [](const orc_gradient_rule& p, const orc_gradient_rule& q) {
		return p.Sample < q.Sample;
	});
	std::vector<uint16_t> samples;
	for (const auto& r : a) {
		samples.push_back(r.Sample);
	}
	for (const auto& r : g) {
		samples.push_back(r.Sample);
	}
	std::sort(samples.begin(), samples.end());
	samples.erase(std::unique(samples.begin(), samples.end()), samples.end());

	const Permutation<float> p {perm, grad, n};
	Fractal<float> f;
	f.Persistence = 0.7f;
	f.Lacunarity = 1.88f;
	f.Octave = 3u;
	f.Scale = NoiseScale;
	f.HalfX = static_cast<float>(mapX) / 2.0f;
	f.HalfY = static_cast<float>(mapY) / 2.0f;
	f.Amplitude = 1.0f;
	f.Frequency = 1.0f;
	const int kernel[4][2] = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
	for (uint32_t c = 0u; c < count; c++) {
		f.OffsetX = chunks[c].OffsetX;
		f.OffsetY = chunks[c].OffsetY;
		for (uint32_t y = 0u; y < mapY; y++) {
			for (uint32_t x = 0u; x < mapX; x++) {
				const uint32_t sx = x + mapX * chunks[c].LocalX, sy = y + mapY * chunks[c].LocalY;
				float cell[4];
				for (int i = 0; i < 4; i++) {
					cell[i] = splatLinear(height, W, H, static_cast<float>(sx) + static_cast<float>(kernel[i][0]) * KernelRadius,
						static_cast<float>(sy) + static_cast<float>(kernel[i][1]) * KernelRadius) * HeightFactor;
				}
				const float g0 = cell[0] - cell[3], g1 = GradientBias, g2 = cell[1] - cell[2];
				const float slope = 1.0f - g1 * (1.0f / std::sqrt(g0 * g0 + g1 * g1 + g2 * g2));
				const float noise = simplex2DFractal<float>(p, 1.0f * static_cast<float>(x), 1.0f * static_cast<float>(y), f)
					* NoiseContribution;
				const uint16_t s = biome[static_cast<size_t>(sy) * W + sx];
				const float h = saturate<float>(splatLinear(height, W, H, static_cast<float>(sx), static_cast<float>(sy)) + noise);
				uint32_t region = NoRegion;
				// the first sample with rules that is not less than s
				size_t k = 0u;
				while (k < samples.size() && samples[k] < s) {
					k++;
				}
				if (k < samples.size()) {
					const uint16_t t = samples[k];
					// gradient rules first, in order; then the altitude rule with the smallest upper bound >= h
					for (const auto& r : g) {
						if (r.Sample == t && slope >= r.minGradient && slope <= r.maxGradient && h >= r.LowerBound
							&& h <= r.UpperBound) {
							region = r.Region;
							break;
						}
					}
					if (region == NoRegion) {
						for (const auto& r : a) {
							if (r.Sample == t && !(r.UpperBound < h)) {
								region = r.Region;
								break;
							}
						}
					}
				}
				splat[static_cast<size_t>(sy) * W + sx] = static_cast<uint8_t>(region);
			}
		}
	}
}

} // extern "C"

// ------------------------------------------------------------------------------------
// Biome layer chain: the reference's evaluation, pixel by pixel through cached recursion.
// STPBiomeFactory.cpp:24-41 (map layout, y = 0), STPLayer.cpp:47-186 (key hash, MurmurHash3
// finaliser index, cache, seeds, local sampler), SuperDemo+/World/Layers/*.h (sample rules),
// STPBiomeRegistry.cpp:61-141 (registry, predicates, precipitation type, CAS).
// Third-party semantics: std::hash<uint64_t> is libstdc++'s identity (param 0) or the MSVC
// library's FNV-1a over the 8 little-endian bytes (param 1).
// ------------------------------------------------------------------------------------

struct orc_layer_node {
	uint32_t Kind, Param;
	int32_t Parent[2];
	uint64_t Salt, CacheSize;
};

namespace {
namespace layer {

using Seed = uint64_t;
using Sample = uint16_t;

namespace K {
enum : uint32_t {
	Continent, Scale, Land, Island, Voronoi, BaseBiome, ClimateSingle, ClimateModerate, ClimateExtreme, DeepOcean, EaseEdge,
	EdgeBiome, Hills, Noise, NoiseToRiver, OceanNoise, OceanTemperate, OceanExtreme, OceanTransition, OceanMix, RiverErode,
	RiverMix, SmoothScale, KindCount
};
}
// registration order of STPBiomeRegistry.cpp:66-104
enum B : uint32_t {
	Ocean, DeepOcean, WarmOcean, LukewarmOcean, ColdOcean, FrozenOcean, DeepWarmOcean, DeepLukewarmOcean, DeepColdOcean,
	DeepFrozenOcean, River, FrozenRiver, Plains, Desert, Mountain, Forest, Taiga, SnowyTaiga, SnowyTundra, Jungle, Savannah,
	Swamp, Badlands, DesertHills, TaigaHills, WoodedMountain, SnowyMountain, ForestHills, JungleHills, SavannahPlateau,
	SwampHills, BadlandsPlateau, Beach, SnowyBeach, StoneShore, BCount
};

Seed mixSeed(Seed s, const long long fac) {
	s *= s * 6364136223846793005ull + 1442695040888963407ull;
	s += static_cast<Seed>(fac);
	return s;
}

struct LocalSampler {
	Seed LayerSeed, LocalSeed;
	Sample nextValue(const Sample range) {
		const Sample v = static_cast<Sample>((this->LocalSeed >> 24u) % static_cast<Seed>(range));
		this->LocalSeed = mixSeed(this->LocalSeed, static_cast<long long>(this->LayerSeed));
		return v;
	}
	Sample choose(const Sample a, const Sample b) {
		return this->nextValue(2u) == 0u ? a : b;
	}
	Sample choose(const Sample a, const Sample b, const Sample c, const Sample d) {
		const Sample i = this->nextValue(4u);
		return i == 0u ? a : i == 1u ? b : i == 2u ? c : d;
	}
};

struct Chain {
	struct Node {
		uint32_t Kind, Param;
		int32_t Parent[2];
		Seed LayerSeed, VoronoiSeed;
		std::vector<unsigned long long> Key;
		std::vector<Sample> Value;
	};
	std::vector<Node> N;
	Sample Id[BCount];
	const uint8_t* Type; // precipitation type per id: 0 none, 1 rain, 2 snow

	Sample id(const B b) const { return this->Id[b]; }
	bool shallowOcean(const Sample v) const {
		return v == id(Ocean) || v == id(FrozenOcean) || v == id(WarmOcean) || v == id(LukewarmOcean) || v == id(ColdOcean);
	}
	bool ocean(const Sample v) const {
		return shallowOcean(v) || v == id(DeepOcean) || v == id(DeepWarmOcean) || v == id(DeepLukewarmOcean)
			|| v == id(DeepColdOcean) || v == id(DeepFrozenOcean);
	}
	bool snowy(const Sample v) const { return this->Type[v] == 2u; }

	static Seed seedLocal(const Seed layer, const int x, const int z) {
		Seed s = mixSeed(layer, x);
		s = mixSeed(s, z);
		s = mixSeed(s, x);
		return mixSeed(s, z);
	}
	LocalSampler rng(const Node& n, const int x, const int z) const { return LocalSampler {n.LayerSeed, seedLocal(n.LayerSeed, x, z)}; }

	// STPLayer::retrieve through STPLayerCache::get
	Sample retrieve(const int32_t n, const int x, const int y, const int z) {
		Node& node = this->N[static_cast<size_t>(n)];
		if (node.Key.empty()) {
			return this->sample(n, x, y, z);
		}
		unsigned long long key = static_cast<unsigned long long>(x) & ((1ull << 28u) - 1ull);
		key |= (static_cast<unsigned long long>(z) & ((1ull << 28u) - 1ull)) << 28u;
		key |= (static_cast<unsigned long long>(y) & 0xFFull) << 56u;
		unsigned long long h = key;
		h ^= h >> 33u;
		h *= 0xFF51AFD7ED558CCDull;
		h ^= h >> 33u;
		h *= 0xC4CEB9FE1A85EC53ull;
		h ^= h >> 33u;
		const size_t index = static_cast<size_t>(h & (node.Key.size() - 1u));
		if (node.Key[index] == key) {
			return node.Value[index];
		}
		const Sample s = this->sample(n, x, y, z);
		Node& again = this->N[static_cast<size_t>(n)];
		again.Key[index] = key;
		again.Value[index] = s;
		return s;
	}

	Sample cross(const Node& n, const Sample c, const Sample north, const Sample e, const Sample s, const Sample w,
		const Seed local) {
		LocalSampler r {n.LayerSeed, local};
		switch (n.Kind) {
		case K::Island:
			return shallowOcean(c) && shallowOcean(north) && shallowOcean(e) && shallowOcean(s) && shallowOcean(w)
				&& r.nextValue(2u) == 0u ? id(Plains) : c;
		case K::ClimateModerate: {
			const auto mf = [this](const Sample v) { return v == id(Mountain) || v == id(Forest); };
			if (c == id(Plains) && (mf(north) || mf(e) || mf(s) || mf(w))) {
				return id(Desert);
			}
			return c;
		}
		case K::ClimateExtreme: {
			const auto pd = [this](const Sample v) { return v == id(Plains) || v == id(Desert); };
			if (c != id(Forest) || (!pd(north) && !pd(e) && !pd(s) && !pd(w))) {
				return c;
			}
			return id(Mountain);
		}
		case K::DeepOcean: {
			if (!shallowOcean(c)) {
				return c;
			}
			Sample i = 0u;
			i += shallowOcean(north);
			i += shallowOcean(e);
			i += shallowOcean(s);
			i += shallowOcean(w);
			if (i > 3u) {
				if (c == id(Ocean)) return id(DeepOcean);
				if (c == id(WarmOcean)) return id(DeepWarmOcean);
				if (c == id(LukewarmOcean)) return id(DeepLukewarmOcean);
				if (c == id(ColdOcean)) return id(DeepColdOcean);
				if (c == id(FrozenOcean)) return id(DeepFrozenOcean);
			}
			return c;
		}
		case K::EaseEdge: {
			const auto any = [&](const Sample v) { return v == north || v == e || v == s || v == w; };
			if (c == id(Desert) && any(id(SnowyTundra))) {
				return id(WoodedMountain);
			}
			if (c == id(Swamp) && (any(id(Desert)) || any(id(SnowyTundra)) || any(id(SnowyTaiga)))) {
				return id(Plains);
			}
			return c;
		}
		case K::EdgeBiome: {
			if (ocean(c)) {
				return c;
			}
			const bool snow = snowy(c);
			if (ocean(north) || ocean(e) || ocean(s) || ocean(w)) {
				if (snow) return id(SnowyBeach);
				if (c == id(Mountain)) return id(StoneShore);
				return id(Beach);
			}
			return c;
		}
		case K::NoiseToRiver: {
			const auto f = [](const Sample v) -> Sample { return v >= 2u ? static_cast<Sample>(2u + (v & 1u)) : v; };
			const Sample i = f(c);
			return i == f(north) && i == f(e) && i == f(s) && i == f(w) ? Sample {65535u} : id(River);
		}
		case K::OceanTemperate: {
			const auto none = [&](const Sample v) { return north != v && e != v && s != v && w != v; };
			if (c != id(LukewarmOcean) || none(id(ColdOcean))) return c;
			if (c != id(ColdOcean) || none(id(LukewarmOcean))) return c;
			return id(Ocean);
		}
		case K::OceanExtreme: {
			const auto any = [&](const Sample v) { return north == v || e == v || s == v || w == v; };
			if (c == id(WarmOcean) && any(id(FrozenOcean))) return id(LukewarmOcean);
			if (c == id(FrozenOcean) && any(id(WarmOcean))) return id(ColdOcean);
			return c;
		}
		case K::RiverErode:
			return north == id(River) || e == id(River) || s == id(River) || w == id(River) ? id(River) : c;
		case K::SmoothScale: {
			const bool xm = w == e, zm = north == s;
			if (xm == zm) {
				return xm ? r.choose(w, north) : c;
			}
			return xm ? w : north;
		}
		default: return c;
		}
	}

	Sample land(const Node& n, const Sample center, const Sample ne, const Sample se, const Sample sw, const Sample nw,
		const Seed local) {
		LocalSampler rng {n.LayerSeed, local};
		const auto so = [this](const Sample v) { return shallowOcean(v); };
		const auto cas = [](const Sample a, const Sample b, const Sample f) { return a == b ? b : f; };
		if (!so(center) || (so(sw) && so(se) && so(ne) && so(nw))) {
			if (so(center) || (!so(sw) && !so(se) && !so(ne) && !so(nw)) || rng.nextValue(5u) != 0u) {
				return center;
			}
			if (so(nw)) return cas(center, id(Forest), nw);
			if (so(sw)) return cas(center, id(Forest), sw);
			if (so(ne)) return cas(center, id(Forest), ne);
			if (so(se)) return cas(center, id(Forest), se);
			return center;
		}
		Sample i = 1u, j = 1u;
		if (!so(nw) && rng.nextValue(i++) == 0u) j = nw;
		if (!so(ne) && rng.nextValue(i++) == 0u) j = ne;
		if (!so(sw) && rng.nextValue(i++) == 0u) j = sw;
		if (!so(se) && rng.nextValue(i) == 0u) j = se;
		if (rng.nextValue(3u) == 0u) {
			return j;
		}
		return j == id(Forest) ? id(Forest) : center;
	}

	static double voronoiDist(const Seed seed, const int x, const int y, const int z, const double xf, const double yf,
		const double zf) {
		const auto distribute = [](const Seed s) {
			const double d = static_cast<double>(static_cast<unsigned int>((s >> 24u) % 1024u)) / 1024.0;
			return (d - 0.5) * 0.9;
		};
		Seed m = mixSeed(seed, x);
		m = mixSeed(m, y);
		m = mixSeed(m, z);
		m = mixSeed(m, x);
		m = mixSeed(m, y);
		m = mixSeed(m, z);
		const double d = distribute(m);
		m = mixSeed(m, static_cast<long long>(seed));
		const double e = distribute(m);
		m = mixSeed(m, static_cast<long long>(seed));
		const double f = distribute(m);
		return (zf + f) * (zf + f) + (yf + e) * (yf + e) + (xf + d) * (xf + d);
	}

	Sample sample(const int32_t ni, const int x, const int y, const int z) {
		const Node n = this->N[static_cast<size_t>(ni)]; // a copy: recursion may touch the vector's caches
		const int32_t p0 = n.Parent[0], p1 = n.Parent[1];
		switch (n.Kind) {
		case K::Continent: {
			LocalSampler r = rng(n, x, z);
			return r.nextValue(10u) == 0u ? id(Plains) : id(Ocean);
		}
		case K::OceanNoise: {
			LocalSampler r = rng(n, x, z);
			const Sample i = r.nextValue(3u);
			return i == 0u ? id(FrozenOcean) : i == 1u ? id(WarmOcean) : id(Ocean);
		}
		case K::Scale: {
			const Sample i = retrieve(p0, x >> 1, y, z >> 1);
			const int xb = x & 1, zb = z & 1;
			LocalSampler r = rng(n, x & -2, z & -2);
			if (xb == 0 && zb == 0) {
				return i;
			}
			const Sample l = retrieve(p0, x >> 1, y, (z + 1) >> 1);
			const Sample m = r.choose(i, l);
			if (xb == 0) {
				return m;
			}
			const Sample nn = retrieve(p0, (x + 1) >> 1, y, z >> 1);
			const Sample o = r.choose(i, nn);
			if (zb == 0) {
				return o;
			}
			const Sample p = retrieve(p0, (x + 1) >> 1, y, (z + 1) >> 1);
			// the rule on (center, e, s, se) = (i, n, l, p)
			const Sample center = i, e = nn, s = l, se = p;
			const Sample ret = r.choose(center, e, s, se);
			if (n.Param != 0u) return ret;
			if (e == s && e == se) return e;
			if (center == e && (center == se || s != se)) return center;
			if (center == s && (center == se || e != se)) return center;
			if (center == se && e != s) return center;
			if (e == s && center != se) return e;
			if (e == se && center != s) return e;
			if (s == se && center != e) return s;
			return ret;
		}
		case K::Land:
			return land(n, retrieve(p0, x, y, z), retrieve(p0, x + 1, y, z - 1), retrieve(p0, x + 1, y, z + 1),
				retrieve(p0, x - 1, y, z + 1), retrieve(p0, x - 1, y, z - 1), seedLocal(n.LayerSeed, x, z));
		case K::Voronoi: {
			const int ijk[3] = {x - 2, y - 2, z - 2};
			const int lmn[3] = {ijk[0] >> 2, ijk[1] >> 2, ijk[2] >> 2};
			const double def[3] = {static_cast<double>(ijk[0] & 3) / 4.0, static_cast<double>(ijk[1] & 3) / 4.0,
				static_cast<double>(ijk[2] & 3) / 4.0};
			double ds[8];
			for (unsigned int c = 0u; c < 8u; c++) {
				const bool bl[3] = {(c & 4u) == 0u, (c & 2u) == 0u, (c & 1u) == 0u};
				ds[c] = voronoiDist(n.VoronoiSeed, bl[0] ? lmn[0] : lmn[0] + 1, bl[1] ? lmn[1] : lmn[1] + 1,
					bl[2] ? lmn[2] : lmn[2] + 1, bl[0] ? def[0] : def[0] - 1.0, bl[1] ? def[1] : def[1] - 1.0,
					bl[2] ? def[2] : def[2] - 1.0);
			}
			unsigned int index = 0u;
			double min = ds[0];
			for (unsigned int c = 1u; c < 8u; c++) {
				if (ds[c] >= min) continue;
				index = c;
				min = ds[c];
			}
			return retrieve(p0, (index & 4u) == 0u ? lmn[0] : lmn[0] + 1, 0, (index & 1u) == 0u ? lmn[2] : lmn[2] + 1);
		}
		case K::BaseBiome: {
			LocalSampler r = rng(n, x, z);
			const Sample c = retrieve(p0, x, y, z);
			if (ocean(c)) return c;
			if (c == id(Plains)) {
				const Sample l[4] = {id(Desert), id(Savannah), id(Plains), id(Badlands)};
				return l[r.nextValue(4u)];
			}
			if (c == id(Desert)) {
				const Sample l[5] = {id(Jungle), id(Forest), id(Mountain), id(Plains), id(Swamp)};
				return l[r.nextValue(5u)];
			}
			if (c == id(Mountain)) {
				const Sample l[4] = {id(Plains), id(Mountain), id(Forest), id(Taiga)};
				return l[r.nextValue(4u)];
			}
			if (c == id(Forest)) {
				const Sample l[3] = {id(SnowyTaiga), id(SnowyTundra), id(SnowyMountain)};
				return l[r.nextValue(3u)];
			}
			return c;
		}
		case K::ClimateSingle: {
			const Sample v = retrieve(p0, x, y, z);
			LocalSampler r = rng(n, x, z);
			if (shallowOcean(v)) return v;
			const Sample i = r.nextValue(6u);
			return i == 0u ? id(Forest) : i == 1u ? id(Mountain) : id(Plains);
		}
		case K::Noise: {
			LocalSampler r = rng(n, x, z);
			const Sample v = retrieve(p0, x, y, z);
			return shallowOcean(v) ? v : static_cast<Sample>(r.nextValue(static_cast<Sample>(65535u / 2u)) + 2u);
		}
		case K::OceanTransition: {
			const Sample v = retrieve(p0, x, y, z);
			if (!ocean(v)) return v;
			for (int rx = -8; rx <= 8; rx += 4) {
				for (int rz = -8; rz <= 8; rz += 4) {
					if (ocean(retrieve(p0, x + rx, y, z + rz))) continue;
					if (v == id(WarmOcean)) return id(LukewarmOcean);
					if (v == id(FrozenOcean)) return id(ColdOcean);
				}
			}
			return v;
		}
		case K::OceanMix: {
			const Sample l = retrieve(p0, x, y, z);
			if (!ocean(l)) return l;
			return retrieve(p1, x, y, z);
		}
		case K::RiverMix: {
			const Sample l = retrieve(p0, x, y, z), rv = retrieve(p1, x, y, z);
			if (ocean(l)) return l;
			if (rv == id(River) && snowy(l)) return id(FrozenRiver);
			return l;
		}
		case K::Hills: {
			LocalSampler r = rng(n, x, z);
			const Sample lv = retrieve(p0, x, y, z), nv = retrieve(p1, x, y, z);
			const Sample is_hill = static_cast<Sample>((nv - 2) % 29);
			if (r.nextValue(3u) == 0u || is_hill == 0u) {
				Sample l = lv;
				if (lv == id(Desert)) l = id(DesertHills);
				else if (lv == id(Taiga)) l = id(TaigaHills);
				else if (lv == id(Mountain)) l = id(WoodedMountain);
				else if (lv == id(SnowyTundra) || lv == id(SnowyTaiga)) l = id(SnowyMountain);
				else if (lv == id(Plains)) l = r.nextValue(3u) == 0u ? id(ForestHills) : id(Forest);
				else if (lv == id(Forest)) l = id(ForestHills);
				else if (lv == id(Jungle)) l = id(JungleHills);
				else if (lv == id(Savannah)) l = id(SavannahPlateau);
				else if (lv == id(Swamp)) l = id(SwampHills);
				else if (lv == id(Badlands)) l = id(BadlandsPlateau);
				else if (lv == id(Ocean)) l = id(DeepOcean);
				else if (lv == id(WarmOcean)) l = id(DeepWarmOcean);
				else if (lv == id(LukewarmOcean)) l = id(DeepLukewarmOcean);
				else if (lv == id(ColdOcean)) l = id(DeepColdOcean);
				else if (lv == id(FrozenOcean)) l = id(DeepFrozenOcean);
				if (ocean(lv) && !shallowOcean(lv) && r.nextValue(3u) == 0u) {
					l = r.nextValue(2u) == 0u ? id(Plains) : id(Forest);
				}
				if (l != lv) {
					unsigned char m = 0u;
					if (lv == retrieve(p0, x, y, z - 1)) m++;
					if (lv == retrieve(p0, x + 1, y, z)) m++;
					if (lv == retrieve(p0, x - 1, y, z)) m++;
					if (lv == retrieve(p0, x, y, z + 1)) m++;
					if (m >= 3u) return l;
				}
			}
			return lv;
		}
		default:
			return cross(n, retrieve(p0, x, y, z), retrieve(p0, x, y, z - 1), retrieve(p0, x + 1, y, z),
				retrieve(p0, x, y, z + 1), retrieve(p0, x - 1, y, z), seedLocal(n.LayerSeed, x, z));
		}
	}
};

} // namespace layer
} // namespace

extern "C" {

// STPLayer::mixSeed, for the pins of the layer RNG
uint64_t orc_layer_mix_seed(const uint64_t s, const int64_t fac) {
	return layer::mixSeed(s, static_cast<long long>(fac));
}

// Biome maps of the chain's root (the last node) at each origin, width x height, the maps one after
// another; one production line (caches kept across maps, as the reference's pooled layers).
// Returns 0, or 1 for a malformed chain (parents after the node, cache not a power of 2).
int orc_layer_biomemap(const orc_layer_node* nodes, const uint32_t count, const uint64_t global, const uint16_t* ids,
	const float* temperature, const float* precipitation, const int32_t* origin, const uint32_t maps, const uint32_t width,
	const uint32_t height, uint16_t* out) {
	using namespace layer;
	Chain ch;
	std::vector<uint8_t> type(65536u, 0u), seen(65536u, 0u);
	for (uint32_t b = 0u; b < BCount; b++) {
		ch.Id[b] = ids[b];
		if (!seen[ids[b]]) {
			seen[ids[b]] = 1u;
			type[ids[b]] = precipitation[b] < 1.0f ? 0u : (temperature[b] < 1.0f ? 2u : 1u);
		}
	}
	ch.Type = type.data();
	for (uint32_t i = 0u; i < count; i++) {
		const orc_layer_node& d = nodes[i];
		if (d.Kind >= K::KindCount || d.Parent[0] >= static_cast<int32_t>(i) || d.Parent[1] >= static_cast<int32_t>(i)
			|| (d.CacheSize & (d.CacheSize - 1u)) != 0u) {
			return 1;
		}
		Chain::Node n;
		n.Kind = d.Kind;
		n.Param = d.Param;
		n.Parent[0] = d.Parent[0];
		n.Parent[1] = d.Parent[1];
		// STPLayer.cpp:143-151
		Seed mid = mixSeed(d.Salt, static_cast<long long>(d.Salt));
		mid = mixSeed(mid, static_cast<long long>(mid));
		mid = mixSeed(mid, static_cast<long long>(mid));
		Seed s = mixSeed(global, static_cast<long long>(mid));
		s = mixSeed(s, static_cast<long long>(mid));
		n.LayerSeed = mixSeed(s, static_cast<long long>(mid));
		Seed h = global;
		if (d.Param == 1u) {
			h = 14695981039346656037ull;
			for (unsigned int k = 0u; k < 8u; k++) {
				h ^= (global >> (8u * k)) & 0xFFu;
				h *= 1099511628211ull;
			}
		}
		n.VoronoiSeed = h;
		if (d.CacheSize) {
			n.Key.assign(d.CacheSize, ~0ull);
			n.Value.assign(d.CacheSize, 0u);
		}
		ch.N.push_back(std::move(n));
	}
	const int32_t root = static_cast<int32_t>(count) - 1;
	for (uint32_t m = 0u; m < maps; m++) {
		for (uint32_t z = 0u; z < height; z++) {
			for (uint32_t x = 0u; x < width; x++) {
				out[static_cast<size_t>(m) * width * height + x + static_cast<size_t>(z) * width] =
					ch.retrieve(root, static_cast<int>(x) + origin[2 * m], 0, static_cast<int>(z) + origin[2 * m + 1]);
			}
		}
	}
	return 0;
}

} // extern "C"
