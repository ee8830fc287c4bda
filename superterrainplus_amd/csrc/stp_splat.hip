// Splatmap kernel and rule table (stp_splat.hpp).
//
// Roofline: per pixel 6 UINT16 heightfield reads (4 gradient samples + the centre, each 1, 2 or 4
// texels), a biome read and a byte written, served mostly by L1/L2 (neighbouring pixels share
// texels); 3 noise octaves of ~66 FLOP and two short rule searches. FP32 VALU bound.
#include "stp_splat.hpp"
#include "stp_stage_noise.hpp"

#include <algorithm>

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPStage;
namespace STPTI = SuperTerrainPlus::STPDiversity::STPTextureInformation;

namespace {

	// the demo script's constants (SuperDemo+/Script/STPSplatmapGenerator.cu:2-18)
	constexpr float GradientBias = 5.5f, KernelRadius = 2.5f, NoiseScale = 91.5f, NoiseContribution = 0.15f,
		HeightFactor = 875.5f, Persistence = 0.7f, Lacunarity = 1.88f;
	constexpr unsigned int NoiseOctave = 3u;

	constexpr unsigned int SplatX = 64u, SplatY = 4u, MaxChunkPerLaunch = 64u;

	struct ChunkList {
		STPSplatChunk Chunk[MaxChunkPerLaunch];
	};

	// A UINT16 texel read as a normalised float, clamped to the area (the heightfield texture's
	// clamp address mode and normalised-float read mode, STPWorldPipeline.cpp:682-696).
	__device__ inline float texel(const unsigned short* const map, const uint2 area, const int x, const int y) {
		const int cx = min(max(x, 0), static_cast<int>(area.x) - 1), cy = min(max(y, 0), static_cast<int>(area.y) - 1);
		return static_cast<float>(map[static_cast<size_t>(cy) * area.x + cx]) / 65535.0f;
	}

	// Linear filtering at unnormalised coordinates (texel centres at +0.5): the bilinear mix of the
	// 2 x 2 texels around (u - 0.5, v - 0.5). At the script's sample points every weight is 0, 1/2
	// or 1/4.
	__device__ inline float sampleLinear(const unsigned short* const map, const uint2 area, const float u, const float v) {
		const float xb = u - 0.5f, yb = v - 0.5f, fx = floorf(xb), fy = floorf(yb);
		const float a = xb - fx, b = yb - fy;
		const int i = static_cast<int>(fx), j = static_cast<int>(fy);
		const float t00 = texel(map, area, i, j), t10 = texel(map, area, i + 1, j), t01 = texel(map, area, i, j + 1),
			t11 = texel(map, area, i + 1, j + 1);
		return (1.0f - a) * ((1.0f - b) * t00 + b * t01) + a * ((1.0f - b) * t10 + b * t11);
	}

	// generateTextureSplatmap (SuperDemo+/Script/STPSplatmapGenerator.cu:65-137), one pixel of chunk
	// blockIdx.z per thread and row step.
	__global__ void __launch_bounds__(SplatX * SplatY) splatmapKERNEL(const STPAlgorithm::STPPermutation perm,
		const STPTI::STPSplatRuleDatabase database, const unsigned short* const biome, const unsigned short* const height,
		unsigned char* const splat, const uint2 area, const uint2 map, const float2 half, const ChunkList chunks,
		const unsigned int count) {
		__shared__ Noise::STPStagedPermutation table;
		Noise::stage(perm, table);
		const unsigned int x = blockIdx.x * SplatX + threadIdx.x, y = blockIdx.y * SplatY + threadIdx.y, z = blockIdx.z;
		if (x >= map.x || y >= map.y || z >= count) {
			return;
		}
		const STPSplatChunk& local = chunks.Chunk[z];
		const unsigned int sx = x + map.x * local.LocalX, sy = y + map.y * local.LocalY;
		// heightfield gradient from a cross of samples KernelRadius away: top, left, right, bottom
		constexpr int Kernel[4][2] = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
		float cell[4];
#pragma unroll
		for (unsigned int i = 0u; i < 4u; i++) {
			cell[i] = sampleLinear(height, area, static_cast<float>(sx) + static_cast<float>(Kernel[i][0]) * KernelRadius,
				static_cast<float>(sy) + static_cast<float>(Kernel[i][1]) * KernelRadius) * HeightFactor;
		}
		const float g0 = cell[0] - cell[3], g1 = GradientBias, g2 = cell[1] - cell[2];
		const float slope = 1.0f - g1 * (1.0f / sqrtf(g0 * g0 + g1 * g1 + g2 * g2));

		Noise::STPFractalSimplexInformation info = {};
		info.Persistence = Persistence;
		info.Lacunarity = Lacunarity;
		info.Octave = NoiseOctave;
		info.Scale = NoiseScale;
		info.Offset = make_float2(local.OffsetX, local.OffsetY);
		info.HalfDimension = half;
		info.Amplitude = 1.0f;
		info.Frequency = 1.0f;
		const float noise = Noise::simplex2DFractal(table, 1.0f * static_cast<float>(x), 1.0f * static_cast<float>(y), info)
			* NoiseContribution;

		const unsigned short sample = biome[static_cast<size_t>(sy) * area.x + sx];
		float h = sampleLinear(height, area, static_cast<float>(sx), static_cast<float>(sy)) + noise;
		h = h > 1.0f ? 1.0f : (h > 0.0f ? h : 0.0f);

		const STPAlgorithm::STPTextureSplatRuleWrapper rule(database);
		const auto* const registry = rule.findSplatRegistry(sample);
		// a gradient rule outweighs an altitude rule
		unsigned int region = rule.gradientRegion(registry, slope, h);
		if (region == STPAlgorithm::STPTextureSplatRuleWrapper::NoRegion) {
			region = rule.altitudeRegion(registry, h);
		}
		splat[static_cast<size_t>(sy) * area.x + sx] = static_cast<unsigned char>(region);
	}

	template <class T>
	STPSmartDeviceMemory::STPDevice<T> upload(const std::vector<T>& v) {
		auto d = STPSmartDeviceMemory::makeDevice<T>(std::max<size_t>(v.size(), 1u));
		if (!v.empty()) {
			STP_CHECK_HIP(hipMemcpy(d.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
		}
		return d;
	}

} // namespace

STPSplatRuleTable::STPSplatRuleTable(const std::vector<STPAltitudeRule>& altitude, const std::vector<STPGradientRule>& gradient) {
	// STPTextureDatabase.cpp:357-411: altitudes by (sample, upper bound), gradients by sample
	std::vector<STPAltitudeRule> alt = altitude;
	std::vector<STPGradientRule> gra = gradient;
	for (const STPAltitudeRule& a : alt) {
		if (!(a.UpperBound == a.UpperBound)) {
			STP_THROW(STPInvalidArgument, "altitude upper bound is NaN");
		}
	}
	for (const STPGradientRule& g : gra) {
		if (!(g.minGradient <= g.maxGradient && g.LowerBound <= g.UpperBound)) {
			STP_THROW(STPInvalidArgument, "gradient rule needs minGradient <= maxGradient and LowerBound <= UpperBound");
		}
	}
	std::stable_sort(alt.begin(), alt.end(), [](const auto& p, const auto& q) {
		return p.Sample != q.Sample ? p.Sample < q.Sample : p.UpperBound < q.UpperBound;
	});
	std::stable_sort(gra.begin(), gra.end(), [](const auto& p, const auto& q) { return p.Sample < q.Sample; });
	std::vector<unsigned short> dictionary;
	for (const auto& a : alt) {
		dictionary.push_back(a.Sample);
	}
	for (const auto& g : gra) {
		dictionary.push_back(g.Sample);
	}
	std::sort(dictionary.begin(), dictionary.end());
	dictionary.erase(std::unique(dictionary.begin(), dictionary.end()), dictionary.end());
	// STPTextureFactory.cpp:137-176: one registry per sample, entries into the sorted node arrays
	std::vector<STPTI::STPSplatRegistry> registry;
	std::vector<STPTI::STPAltitudeNode> altNode;
	std::vector<STPTI::STPGradientNode> graNode;
	size_t ai = 0u, gi = 0u;
	for (const unsigned short s : dictionary) {
		STPTI::STPSplatRegistry r {};
		r.AltitudeEntry = static_cast<unsigned int>(altNode.size());
		for (; ai < alt.size() && alt[ai].Sample == s; ai++) {
			STPTI::STPAltitudeNode n {};
			n.Reference.RegionIndex = alt[ai].Region;
			n.UpperBound = alt[ai].UpperBound;
			altNode.push_back(n);
		}
		r.AltitudeSize = static_cast<unsigned int>(altNode.size()) - r.AltitudeEntry;
		r.GradientEntry = static_cast<unsigned int>(graNode.size());
		for (; gi < gra.size() && gra[gi].Sample == s; gi++) {
			STPTI::STPGradientNode n {};
			n.Reference.RegionIndex = gra[gi].Region;
			n.minGradient = gra[gi].minGradient;
			n.maxGradient = gra[gi].maxGradient;
			n.LowerBound = gra[gi].LowerBound;
			n.UpperBound = gra[gi].UpperBound;
			graNode.push_back(n);
		}
		r.GradientSize = static_cast<unsigned int>(graNode.size()) - r.GradientEntry;
		registry.push_back(r);
	}
	this->Count = static_cast<unsigned int>(dictionary.size());
	this->Dictionary = upload(dictionary);
	this->Registry = upload(registry);
	this->Altitude = upload(altNode);
	this->Gradient = upload(graNode);
}

STPTI::STPSplatRuleDatabase STPSplatRuleTable::view() const noexcept {
	return STPTI::STPSplatRuleDatabase {this->Dictionary.get(), this->Count, this->Registry.get(), this->Altitude.get(),
		this->Gradient.get()};
}

void STPHeightfieldKernel::splatmap(const STPAlgorithm::STPPermutation& perm, const STPTI::STPSplatRuleDatabase& database,
	const unsigned short* const biome, const unsigned short* const height, unsigned char* const splat, const uint2 area,
	const uint2 map, const float2 half, const std::vector<STPSplatChunk>& chunks, const hipStream_t stream) {
	if (!biome || !height || !splat) {
		STP_THROW(STPInvalidArgument, "null map");
	}
	if (map.x == 0u || map.y == 0u || area.x == 0u || area.y == 0u) {
		STP_THROW(STPInvalidArgument, "empty map");
	}
	for (const STPSplatChunk& c : chunks) {
		if ((static_cast<unsigned long long>(c.LocalX) + 1ull) * map.x > area.x
			|| (static_cast<unsigned long long>(c.LocalY) + 1ull) * map.y > area.y) {
			STP_THROW(STPInvalidArgument, "a splat chunk lies outside the rendered area");
		}
	}
	for (size_t b = 0u; b < chunks.size(); b += MaxChunkPerLaunch) {
		const size_t e = std::min(chunks.size(), b + MaxChunkPerLaunch);
		ChunkList list {};
		std::copy(chunks.begin() + b, chunks.begin() + e, list.Chunk);
		const dim3 grid((map.x + SplatX - 1u) / SplatX, (map.y + SplatY - 1u) / SplatY, static_cast<unsigned int>(e - b));
		splatmapKERNEL<<<grid, dim3(SplatX, SplatY), 0, stream>>>(perm, database, biome, height, splat, area, map, half,
			list, static_cast<unsigned int>(e - b));
		STP_CHECK_HIP(hipGetLastError());
	}
}
