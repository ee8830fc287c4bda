// Per-biome noise parameters of the multi-biome heightfield generator.
//
// Follows SuperDemo+/World/Biomes/STPBiomeProperty.hpp:10-27. Header-only and free of standard
// headers so that runtime-compiled (hipRTC) diversity generators can include it.
#pragma once
#ifndef STP_AMD_BIOME_PROPERTY_HPP
#define STP_AMD_BIOME_PROPERTY_HPP

namespace SuperTerrainPlus {

	struct STPBiomeProperty {
		float Scale;
		unsigned int Octave;
		float Persistence;
		float Lacunarity;
		float Depth;
		float Variation;
	};

} // namespace SuperTerrainPlus

#endif // STP_AMD_BIOME_PROPERTY_HPP
