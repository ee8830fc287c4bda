// Splat rules of the texture system and their device lookup: the rule database of
// STPTextureInformation (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Diversity/Texture/STPTextureInformation.hpp:40-109)
// and STPTextureSplatRuleWrapper
// (SuperTerrain+/SuperAlgorithm+/Device/Public/SuperAlgorithm+Device/STPTextureSplatRuleWrapper.cuh:14-60,
// SuperTerrain+/SuperAlgorithm+/Device/Private/STPTextureSplatRuleWrapper.cu:56-126).
//
// Header-only and hipRTC-safe (no standard headers), so runtime-compiled splatmap scripts include
// it by name as the reference's scripts include the device library.
#pragma once
#ifndef STP_AMD_TEXTURE_SPLAT_HPP
#define STP_AMD_TEXTURE_SPLAT_HPP

namespace SuperTerrainPlus {

	namespace STPDiversity::STPTextureInformation {

		// the texture region a rule activates
		struct STPStructureNode {
			struct {
				unsigned int RegionIndex;
			} Reference;
		};
		// active while the altitude is at most UpperBound (nodes of a sample sorted by UpperBound)
		struct STPAltitudeNode : public STPStructureNode {
			float UpperBound;
		};
		// active while minGradient <= gradient <= maxGradient and LowerBound <= altitude <= UpperBound
		struct STPGradientNode : public STPStructureNode {
			float minGradient, maxGradient, LowerBound, UpperBound;
		};
		// a sample's nodes: [AltitudeEntry, + AltitudeSize) and [GradientEntry, + GradientSize)
		struct STPSplatRegistry {
			unsigned int AltitudeEntry, AltitudeSize, GradientEntry, GradientSize;
		};
		// SplatRegistryDictionary: the samples with rules, ascending; registry i belongs to sample i
		struct STPSplatRuleDatabase {
			const unsigned short* SplatRegistryDictionary;
			unsigned int DictionaryEntryCount;
			const STPSplatRegistry* SplatRegistry;
			const STPAltitudeNode* AltitudeRegistry;
			const STPGradientNode* GradientRegistry;
		};

		// The chunks of the rendered area a splatmap launch covers (STPTextureInformation.hpp:114-135):
		// each chunk's position in the area (in chunks) and its map offset; device array.
		struct STPSplatGeneratorInformation {
			struct STPLocalChunkInformation {
				unsigned int LocalChunkCoordinateX, LocalChunkCoordinateY;
				float ChunkMapOffsetX, ChunkMapOffsetY;
			};
			const STPLocalChunkInformation* RequestingLocalInfo;
			unsigned int LocalCount;
		};

	} // namespace STPDiversity::STPTextureInformation

	namespace STPAlgorithm {

		class STPTextureSplatRuleWrapper {
		public:
			typedef STPDiversity::STPTextureInformation::STPSplatRuleDatabase STPSplatRuleDatabase;
			typedef STPDiversity::STPTextureInformation::STPSplatRegistry STPSplatRegistry;
			static constexpr unsigned int NoRegion = 0xFFFFFFFFu;

			__device__ explicit STPTextureSplatRuleWrapper(const STPSplatRuleDatabase& database) : SplatRule(database) {}

			// The registry of the first dictionary sample not less than `sample` (a lower-bound
			// binary search, as the reference: a sample without rules takes the next larger
			// sample's), or null past the end.
			__device__ const STPSplatRegistry* findSplatRegistry(const unsigned short sample) const {
				unsigned int first = 0u, count = this->SplatRule.DictionaryEntryCount;
				while (count > 0u) {
					const unsigned int step = count / 2u;
					if (this->SplatRule.SplatRegistryDictionary[first + step] < sample) {
						first += step + 1u;
						count -= step + 1u;
					} else {
						count = step;
					}
				}
				return first == this->SplatRule.DictionaryEntryCount ? nullptr : this->SplatRule.SplatRegistry + first;
			}

			// Region of the first altitude node with UpperBound >= alt (lower-bound search), or NoRegion.
			__device__ unsigned int altitudeRegion(const STPSplatRegistry* const registry, const float alt) const {
				if (!registry) {
					return NoRegion;
				}
				const auto* const node = this->SplatRule.AltitudeRegistry + registry->AltitudeEntry;
				unsigned int first = 0u, count = registry->AltitudeSize;
				while (count > 0u) {
					const unsigned int step = count / 2u;
					if (node[first + step].UpperBound < alt) {
						first += step + 1u;
						count -= step + 1u;
					} else {
						count = step;
					}
				}
				return first == registry->AltitudeSize ? NoRegion : node[first].Reference.RegionIndex;
			}

			// Region of the first gradient node, in order, whose ranges hold (gra, alt), or NoRegion.
			__device__ unsigned int gradientRegion(const STPSplatRegistry* const registry, const float gra, const float alt) const {
				if (!registry) {
					return NoRegion;
				}
				const auto* const node = this->SplatRule.GradientRegistry;
				const unsigned int begin = registry->GradientEntry, end = begin + registry->GradientSize;
				for (unsigned int i = begin; i < end; i++) {
					if (gra >= node[i].minGradient && gra <= node[i].maxGradient && alt >= node[i].LowerBound
						&& alt <= node[i].UpperBound) {
						return node[i].Reference.RegionIndex;
					}
				}
				return NoRegion;
			}

		private:
			const STPSplatRuleDatabase& SplatRule;
		};

	} // namespace STPAlgorithm

} // namespace SuperTerrainPlus

#endif // STP_AMD_TEXTURE_SPLAT_HPP
