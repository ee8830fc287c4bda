// Splatmap generation: the texture region of every pixel of the rendered area's chunks, from the
// area's biome map and UINT16 heightfield, by the splat rules of the texture system.
//
// The reference runs the demo's runtime-compiled script SuperDemo+/Script/STPSplatmapGenerator.cu:65-137
// (constants :2-18) through STPTextureFactory (SuperTerrain+/SuperTerrain+/Private/World/Diversity/STPTextureFactory.cpp:137-176
// builds the rule database) from the world pipeline's generateSplatmap
// (SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:671-707).
#pragma once
#ifndef STP_AMD_SPLAT_HPP
#define STP_AMD_SPLAT_HPP

#include "stp_internal.hpp"
#include <stp/STPTextureSplat.hpp>

#include <vector>

namespace SuperTerrainPlus {

	// An altitude rule: for sample `Sample`, altitudes up to UpperBound take region `Region`.
	struct STPAltitudeRule {
		unsigned short Sample;
		float UpperBound;
		unsigned int Region;
	};
	// A gradient rule: for sample `Sample`, gradients in [minGradient, maxGradient] at altitudes in
	// [LowerBound, UpperBound] take region `Region`.
	struct STPGradientRule {
		unsigned short Sample;
		float minGradient, maxGradient, LowerBound, UpperBound;
		unsigned int Region;
	};
	// A chunk of the rendered area to splat: its position in the area (in chunks) and its map offset
	// (STPTextureInformation::STPSplatGeneratorInformation::STPLocalChunkInformation).
	struct STPSplatChunk {
		unsigned int LocalX, LocalY;
		float OffsetX, OffsetY;
	};

	// The rule database in device memory, laid out as the reference's factory does: the samples
	// with rules ascending, altitude nodes by (sample, upper bound), gradient nodes by sample in
	// the order given.
	class STPSplatRuleTable {
	public:
		STPSplatRuleTable(const std::vector<STPAltitudeRule>& altitude, const std::vector<STPGradientRule>& gradient);
		STPDiversity::STPTextureInformation::STPSplatRuleDatabase view() const noexcept;

	private:
		STPSmartDeviceMemory::STPDevice<unsigned short> Dictionary;
		STPSmartDeviceMemory::STPDevice<STPDiversity::STPTextureInformation::STPSplatRegistry> Registry;
		STPSmartDeviceMemory::STPDevice<STPDiversity::STPTextureInformation::STPAltitudeNode> Altitude;
		STPSmartDeviceMemory::STPDevice<STPDiversity::STPTextureInformation::STPGradientNode> Gradient;
		unsigned int Count = 0u;
	};

	namespace STPHeightfieldKernel {
		// Writes the region (u8, 255 = none) of every pixel of each listed chunk into `splat`;
		// `biome` and `height` (UINT16) cover the rendered area (area.x x area.y, row-major), chunk
		// c's pixels are those of map-sized tile (LocalX, LocalY). `half` is the noise's half
		// dimension (the script's HalfDimension: the map size / 2).
		void splatmap(const STPAlgorithm::STPPermutation& perm,
			const STPDiversity::STPTextureInformation::STPSplatRuleDatabase& database, const unsigned short* biome,
			const unsigned short* height, unsigned char* splat, uint2 area, uint2 map, float2 half,
			const std::vector<STPSplatChunk>& chunks, hipStream_t stream);
	}

} // namespace SuperTerrainPlus

#endif // STP_AMD_SPLAT_HPP
