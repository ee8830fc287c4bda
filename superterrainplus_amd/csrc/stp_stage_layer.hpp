// The sample rule of every demo layer kind (SuperDemo+/World/Layers/*.h), as functions of parent
// values already read; the layer RNG and registry are in stp/STPLayerDevice.hpp.
#pragma once
#ifndef STP_AMD_STAGE_LAYER_HPP
#define STP_AMD_STAGE_LAYER_HPP

#include <stp/STPBiomeLayer.hpp>

namespace SuperTerrainPlus::STPStage::Layer {

	using namespace STPLayerDevice;

	// ---- rules of the single-parent and cross layers, on parent values --------------------------

	// STPScaleLayer.h:26-62: the 2 x 2 parents (centre, east, south, south-east) of an odd-odd cell.
	__device__ inline Sample scaleRule(const Sample center, const Sample e, const Sample s, const Sample se, Sampler& rng,
		const bool fuzzy) {
		const Sample ret = rng.choose(center, e, s, se);
		if (fuzzy) {
			return ret;
		}
		if (e == s && e == se) {
			return e;
		}
		if (center == e && (center == se || s != se)) {
			return center;
		}
		if (center == s && (center == se || e != se)) {
			return center;
		}
		if (center == se && e != s) {
			return center;
		}
		if (e == s && center != se) {
			return e;
		}
		if (e == se && center != s) {
			return e;
		}
		if (s == se && center != e) {
			return s;
		}
		return ret;
	}

	// STPLandLayer.h:21-71 (X cross: north-east, south-east, south-west, north-west).
	__device__ inline Sample land(const Registry& r, const Sample center, const Sample ne, const Sample se, const Sample sw,
		const Sample nw, Sampler rng) {
		const auto so = [&r](const Sample v) { return r.shallowOcean(v); };
		const Sample forest = r[STPNamedBiome::Forest];
		if (!so(center) || (so(sw) && so(se) && so(ne) && so(nw))) {
			if (so(center) || (!so(sw) && !so(se) && !so(ne) && !so(nw)) || rng.next<5u>() != 0u) {
				return center;
			}
			// STPBiomeRegistry::CAS(center, Forest, fallback)
			const auto cas = [center, forest](const Sample fallback) { return center == forest ? forest : fallback; };
			if (so(nw)) {
				return cas(nw);
			}
			if (so(sw)) {
				return cas(sw);
			}
			if (so(ne)) {
				return cas(ne);
			}
			if (so(se)) {
				return cas(se);
			}
			return center;
		}
		// every land neighbour takes the place with chance 1/i, i counting the land neighbours seen
		unsigned int i = 1u;
		Sample j = 1u;
		const auto draw = [&rng](const unsigned int n) {
			switch (n) {
			case 1u: return rng.next<1u>();
			case 2u: return rng.next<2u>();
			case 3u: return rng.next<3u>();
			default: return rng.next<4u>();
			}
		};
		if (!so(nw) && draw(i++) == 0u) {
			j = nw;
		}
		if (!so(ne) && draw(i++) == 0u) {
			j = ne;
		}
		if (!so(sw) && draw(i++) == 0u) {
			j = sw;
		}
		if (!so(se) && draw(i) == 0u) {
			j = se;
		}
		if (rng.next<3u>() == 0u) {
			return j;
		}
		return j == forest ? forest : center;
	}

	// STPIslandLayer.h:18-25 (cross: north, east, south, west).
	__device__ inline Sample island(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w, Sampler rng) {
		return r.shallowOcean(c) && r.shallowOcean(n) && r.shallowOcean(e) && r.shallowOcean(s) && r.shallowOcean(w)
			&& rng.next<2u>() == 0u ? r[STPNamedBiome::Plains] : c;
	}

	// STPBaseBiomeLayer.h:55-79
	__device__ inline Sample baseBiome(const Registry& r, const Sample climate, Sampler rng) {
		if (r.ocean(climate)) {
			return climate;
		}
		if (climate == r[STPNamedBiome::Plains]) {
			const STPNamedBiome dry[4] = {STPNamedBiome::Desert, STPNamedBiome::Savannah, STPNamedBiome::Plains,
				STPNamedBiome::Badlands};
			return r[dry[rng.next<4u>()]];
		}
		if (climate == r[STPNamedBiome::Desert]) {
			const STPNamedBiome temperate[5] = {STPNamedBiome::Jungle, STPNamedBiome::Forest, STPNamedBiome::Mountain,
				STPNamedBiome::Plains, STPNamedBiome::Swamp};
			return r[temperate[rng.next<5u>()]];
		}
		if (climate == r[STPNamedBiome::Mountain]) {
			const STPNamedBiome cool[4] = {STPNamedBiome::Plains, STPNamedBiome::Mountain, STPNamedBiome::Forest,
				STPNamedBiome::Taiga};
			return r[cool[rng.next<4u>()]];
		}
		if (climate == r[STPNamedBiome::Forest]) {
			const STPNamedBiome snowy[3] = {STPNamedBiome::SnowyTaiga, STPNamedBiome::SnowyTundra, STPNamedBiome::SnowyMountain};
			return r[snowy[rng.next<3u>()]];
		}
		return climate;
	}

	// STPClimateLayer.h:30-47
	__device__ inline Sample climateSingle(const Registry& r, const Sample v, Sampler rng) {
		if (r.shallowOcean(v)) {
			return v;
		}
		switch (rng.next<6u>()) {
		case 0u: return r[STPNamedBiome::Forest];
		case 1u: return r[STPNamedBiome::Mountain];
		default: return r[STPNamedBiome::Plains];
		}
	}

	// STPClimateLayer.h:62-73
	__device__ inline Sample climateModerate(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		const auto mf = [&r](const Sample v) { return v == r[STPNamedBiome::Mountain] || v == r[STPNamedBiome::Forest]; };
		return c == r[STPNamedBiome::Plains] && (mf(n) || mf(e) || mf(s) || mf(w)) ? r[STPNamedBiome::Desert] : c;
	}

	// STPClimateLayer.h:88-98
	__device__ inline Sample climateExtreme(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		const auto pd = [&r](const Sample v) { return v == r[STPNamedBiome::Plains] || v == r[STPNamedBiome::Desert]; };
		return c != r[STPNamedBiome::Forest] || !(pd(n) || pd(e) || pd(s) || pd(w)) ? c : r[STPNamedBiome::Mountain];
	}

	// STPDeepOceanLayer.h:21-58
	__device__ inline Sample deepOcean(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		if (!r.shallowOcean(c)) {
			return c;
		}
		const unsigned int i = r.shallowOcean(n) + r.shallowOcean(e) + r.shallowOcean(s) + r.shallowOcean(w);
		if (i > 3u) {
			if (c == r[STPNamedBiome::Ocean]) {
				return r[STPNamedBiome::DeepOcean];
			}
			if (c == r[STPNamedBiome::WarmOcean]) {
				return r[STPNamedBiome::DeepWarmOcean];
			}
			if (c == r[STPNamedBiome::LukewarmOcean]) {
				return r[STPNamedBiome::DeepLukewarmOcean];
			}
			if (c == r[STPNamedBiome::ColdOcean]) {
				return r[STPNamedBiome::DeepColdOcean];
			}
			if (c == r[STPNamedBiome::FrozenOcean]) {
				return r[STPNamedBiome::DeepFrozenOcean];
			}
		}
		return c;
	}

	// STPEaseEdgeLayer.h:31-44
	__device__ inline Sample easeEdge(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		const auto any = [n, e, s, w](const Sample id) { return id == n || id == e || id == s || id == w; };
		if (c == r[STPNamedBiome::Desert] && any(r[STPNamedBiome::SnowyTundra])) {
			return r[STPNamedBiome::WoodedMountain];
		}
		if (c == r[STPNamedBiome::Swamp]
			&& (any(r[STPNamedBiome::Desert]) || any(r[STPNamedBiome::SnowyTundra]) || any(r[STPNamedBiome::SnowyTaiga]))) {
			return r[STPNamedBiome::Plains];
		}
		return c;
	}

	// STPEdgeBiomeLayer.h:21-44
	__device__ inline Sample edgeBiome(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		if (r.ocean(c)) {
			return c;
		}
		const bool snowy = r.snowy(c);
		if (r.ocean(n) || r.ocean(e) || r.ocean(s) || r.ocean(w)) {
			if (snowy) {
				return r[STPNamedBiome::SnowyBeach];
			}
			if (c == r[STPNamedBiome::Mountain]) {
				return r[STPNamedBiome::StoneShore];
			}
			return r[STPNamedBiome::Beach];
		}
		return c;
	}

	// STPNoiseLayer.h:19-26
	__device__ inline Sample noise(const Registry& r, const Sample v, Sampler rng) {
		return r.shallowOcean(v) ? v : static_cast<Sample>(rng.next<65535u / 2u>() + 2u);
	}

	// STPNoiseToRiverLayer.h:16-34
	__device__ inline Sample noiseToRiver(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		const auto filter = [](const Sample v) -> Sample { return v >= 2u ? static_cast<Sample>(2u + (v & 1u)) : v; };
		const Sample i = filter(c);
		return i == filter(n) && i == filter(e) && i == filter(s) && i == filter(w) ? Sample {65535u} : r[STPNamedBiome::River];
	}

	// STPOceanTemperatureLayer.h:24-41
	__device__ inline Sample oceanNoise(const Registry& r, Sampler rng) {
		switch (rng.next<3u>()) {
		case 0u: return r[STPNamedBiome::FrozenOcean];
		case 1u: return r[STPNamedBiome::WarmOcean];
		default: return r[STPNamedBiome::Ocean];
		}
	}

	// STPOceanTemperatureLayer.h:55-76
	__device__ inline Sample oceanTemperate(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		const auto none = [n, e, s, w](const Sample id) { return n != id && e != id && s != id && w != id; };
		if (c != r[STPNamedBiome::LukewarmOcean] || none(r[STPNamedBiome::ColdOcean])) {
			return c;
		}
		if (c != r[STPNamedBiome::ColdOcean] || none(r[STPNamedBiome::LukewarmOcean])) {
			return c;
		}
		return r[STPNamedBiome::Ocean];
	}

	// STPOceanTemperatureLayer.h:90-109
	__device__ inline Sample oceanExtreme(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		const auto any = [n, e, s, w](const Sample id) { return n == id || e == id || s == id || w == id; };
		if (c == r[STPNamedBiome::WarmOcean] && any(r[STPNamedBiome::FrozenOcean])) {
			return r[STPNamedBiome::LukewarmOcean];
		}
		if (c == r[STPNamedBiome::FrozenOcean] && any(r[STPNamedBiome::WarmOcean])) {
			return r[STPNamedBiome::ColdOcean];
		}
		return c;
	}

	// STPRiverErodeLayer.h:18-28
	__device__ inline Sample riverErode(const Registry& r, const Sample c, const Sample n, const Sample e, const Sample s,
		const Sample w) {
		const Sample river = r[STPNamedBiome::River];
		return n == river || e == river || s == river || w == river ? river : c;
	}

	// STPSmoothScaleLayer.h:18-31
	__device__ inline Sample smoothScale(const Sample c, const Sample n, const Sample e, const Sample s, const Sample w,
		Sampler rng) {
		const bool xMatch = w == e, zMatch = n == s;
		if (xMatch == zMatch) {
			return xMatch ? rng.choose(w, n) : c;
		}
		return xMatch ? w : n;
	}

	// STPVoronoiLayer.h:18-20: a jitter coordinate from the top bits of a mixed seed.
	__device__ inline double voronoiJitter(const Seed s) {
		const double d = static_cast<double>(static_cast<unsigned int>((s >> 24u) % 1024u)) / 1024.0;
		return (d - 0.5) * 0.9;
	}

} // namespace SuperTerrainPlus::STPStage::Layer

#endif // STP_AMD_STAGE_LAYER_HPP
