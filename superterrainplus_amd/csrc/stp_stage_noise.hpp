// Device stage functions of the noise side, shared by the per-chunk kernels
// (stp_kernels_noise.hip) and the batched chunk-table kernels (stp_batch.hip).
#pragma once
#ifndef STP_AMD_STAGE_NOISE_HPP
#define STP_AMD_STAGE_NOISE_HPP

#include "stp_internal.hpp"
#include <stp/STPRandom.hpp>

namespace SuperTerrainPlus::STPStage {

	namespace Noise = SuperTerrainPlus::STPAlgorithm::STPSimplexNoise;

	struct BiomeTable {
		STPBiomeProperty Property[STPHeightfieldKernel::MaxBiome];
	};

	// Host: validates a biome table and packs it for a kernel argument (<= 64 biomes).
	inline BiomeTable packBiomeTable(const STPBiomeProperty* const table, const unsigned int count) {
		if (count == 0u || count > STPHeightfieldKernel::MaxBiome || !table) {
			STP_THROW(STPInvalidArgument, "biome count must be in [1, 64]");
		}
		BiomeTable t = {};
		for (unsigned int i = 0u; i < count; i++) {
			if (table[i].Octave == 0u || !(table[i].Scale != 0.0f)) {
				STP_THROW(STPInvalidArgument, "biome property needs Octave > 0 and Scale != 0");
			}
			t.Property[i] = table[i];
		}
		return t;
	}

	// sampleSimplexNoise (SuperDemo+/Script/STPMultiHeightGenerator.cu:57-71): fractal * Variation + Depth.
	__device__ inline float biomeSample(const Noise::STPStagedPermutation& table, const STPBiomeProperty& bp,
		const float x, const float y, const float2 offset, const float2 half) {
		Noise::STPFractalSimplexInformation f;
		f.Persistence = bp.Persistence;
		f.Lacunarity = bp.Lacunarity;
		f.Octave = bp.Octave;
		f.Scale = bp.Scale;
		f.Offset = offset;
		f.HalfDimension = half;
		f.Amplitude = 1.0f;
		f.Frequency = 1.0f;
		return Noise::simplex2DFractal(table, x, y, f) * bp.Variation + bp.Depth;
	}

	// One pixel of the dense-histogram blend (STPMultiHeightGenerator.cu:35-55): sum over the bins in
	// bin order; a bin naming a biome outside the table contributes nothing.
	__device__ inline float blendPixel(const Noise::STPStagedPermutation& table, const STPBiomeProperty* const prop,
		const unsigned int biome_count, const unsigned int n, const unsigned short* const item,
		const float* const weight, const unsigned int x, const unsigned int y, const float2 offset, const float2 half) {
		float h = 0.0f;
		for (unsigned int b = 0u; b < n; b++) {
			const unsigned short it = item[b];
			if (it < biome_count) {
				h += weight[b] * biomeSample(table, prop[it], 1.0f * static_cast<float>(x), 1.0f * static_cast<float>(y), offset, half);
			}
		}
		return h;
	}

	// blendPixel over packed slots (item | count << 3): weight = count * (1 / window sum), the
	// horizontal pass's product
	__device__ inline float blendPacked(const Noise::STPStagedPermutation& table, const STPBiomeProperty* const prop,
		const unsigned int biome_count, const unsigned int n, const unsigned short* const slot, const float norm,
		const unsigned int x, const unsigned int y, const float2 offset, const float2 half) {
		float h = 0.0f;
		for (unsigned int b = 0u; b < n; b++) {
			const unsigned int s = slot[b], it = s & 7u;
			if (it < biome_count) {
				h += static_cast<float>(s >> 3u) * norm
					* biomeSample(table, prop[it], 1.0f * static_cast<float>(x), 1.0f * static_cast<float>(y), offset, half);
			}
		}
		return h;
	}

	__device__ inline long long floorDiv(const long long a, const long long b) {
		long long q = a / b;
		if ((a % b != 0) && ((a < 0) != (b < 0))) {
			q--;
		}
		return q;
	}

	// Synthetic biome map: nearest jittered site over the 3x3 cell neighbourhood, exact int64
	// distances, first site wins on ties. Site (cx, cy): Philox(key = seed,
	// ctr = (cx, cy, 0x564F524F, 0)); offset = (w0 % cell, w1 % cell); biome = w2 % count.
	__device__ inline unsigned short voronoiBiome(const unsigned long long seed, const unsigned int biome_count,
		const unsigned int cell, const long long wx, const long long wy) {
		const long long cx = floorDiv(wx, cell), cy = floorDiv(wy, cell);
		long long best = 0x7FFFFFFFFFFFFFFFll;
		unsigned int biome = 0u;
		for (int dy = -1; dy <= 1; dy++) {
			for (int dx = -1; dx <= 1; dx++) {
				const long long ncx = cx + dx, ncy = cy + dy;
				const STPRandom::Philox4 r = STPRandom::philox4x32_10(static_cast<unsigned int>(ncx),
					static_cast<unsigned int>(ncy), 0x564F524Fu, 0u, static_cast<unsigned int>(seed),
					static_cast<unsigned int>(seed >> 32));
				const long long sx = ncx * cell + (r.v[0] % cell), sy = ncy * cell + (r.v[1] % cell);
				const long long d = (wx - sx) * (wx - sx) + (wy - sy) * (wy - sy);
				if (d < best) {
					best = d;
					biome = r.v[2] % biome_count;
				}
			}
		}
		return static_cast<unsigned short>(biome);
	}

} // namespace SuperTerrainPlus::STPStage

#endif // STP_AMD_STAGE_NOISE_HPP
