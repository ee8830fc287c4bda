// Noise-side kernels: single-biome fractal heightmap, multi-biome blend, Voronoi biome map.
//
// Roofline: FP32 VALU. One pixel-octave is 66 algorithmic FLOP (53 in simplex2D, 13 in the
// octave loop; a division counts 1, an FMA 2). Memory traffic is 4 B/px written (+ the
// histogram read for the blend), so these kernels are compute-bound. The permutation table is
// staged in LDS once per workgroup (4.5 KB) and every thread produces PixelPerThread pixels
// so the staging cost is amortised.
#include "stp_stage_noise.hpp"

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPAlgorithm;
using namespace SuperTerrainPlus::STPStage;

namespace {

	constexpr unsigned int TileX = 64u, TileY = 4u, PixelPerThread = 2u;

	// Single biome: h = fractal * Variation + Depth (STPMultiHeightGenerator.cu:57-71 with one bin).
	__global__ void __launch_bounds__(TileX * TileY) simplexHeightmapKERNEL(const STPPermutation perm,
		const Noise::STPFractalSimplexInformation info, const float variation, const float depth, float* const height,
		const unsigned int width, const unsigned int rows) {
		__shared__ Noise::STPStagedPermutation table;
		Noise::stage(perm, table);

		const unsigned int x = blockIdx.x * TileX + threadIdx.x;
		const unsigned int y0 = (blockIdx.y * TileY + threadIdx.y) * PixelPerThread;
		if (x >= width) {
			return;
		}
#pragma unroll
		for (unsigned int k = 0u; k < PixelPerThread; k++) {
			const unsigned int y = y0 + k;
			if (y < rows) {
				const float n = Noise::simplex2DFractal(table, 1.0f * static_cast<float>(x), 1.0f * static_cast<float>(y), info);
				height[x + static_cast<size_t>(y) * width] = n * variation + depth;
			}
		}
	}

	// generateMultiBiomeHeightmap (SuperDemo+/Script/STPMultiHeightGenerator.cu:35-55): sums
	// weight * (fractal * Variation + Depth) over the pixel's histogram bins, in bin order.
	__global__ void __launch_bounds__(TileX * TileY) multiBiomeKERNEL(const STPPermutation perm, const BiomeTable biomes,
		const unsigned int biome_count, const STPSingleHistogram histogram, float* const height, const uint2 map_size,
		const float2 offset) {
		__shared__ Noise::STPStagedPermutation table;
		__shared__ STPBiomeProperty prop[STPHeightfieldKernel::MaxBiome];
		const unsigned int tid = threadIdx.x + threadIdx.y * TileX;
		if (tid < biome_count) {
			prop[tid] = biomes.Property[tid];
		}
		Noise::stage(perm, table);

		const float2 half = make_float2(static_cast<float>(map_size.x) / 2.0f, static_cast<float>(map_size.y) / 2.0f);
		const unsigned int x = blockIdx.x * TileX + threadIdx.x;
		const unsigned int y0 = (blockIdx.y * TileY + threadIdx.y) * PixelPerThread;
		if (x >= map_size.x) {
			return;
		}
		for (unsigned int k = 0u; k < PixelPerThread; k++) {
			const unsigned int y = y0 + k;
			if (y >= map_size.y) {
				break;
			}
			const unsigned int index = x + y * map_size.x;
			float h = 0.0f;
			STPSingleHistogramWrapper::iterate(histogram, index, [&](const unsigned short item, const float weight) {
				// a bin naming a biome outside the table contributes nothing
				if (item < biome_count) {
					h += weight * biomeSample(table, prop[item], 1.0f * static_cast<float>(x), 1.0f * static_cast<float>(y), offset, half);
				}
			});
			height[index] = h;
		}
	}

	// The same blend over the dense fixed-stride histogram the device filter produces
	// (bin_count[p], bin_item[p * stride + s], bin_weight[p * stride + s]).
	// Packed: a slot is item | count << 3 and the weight count * norm (no weight array)
	template <bool Packed>
	__global__ void __launch_bounds__(TileX * TileY) multiBiomeDenseKERNEL(const STPPermutation perm,
		const BiomeTable biomes, const unsigned int biome_count, const unsigned char* const bin_count,
		const unsigned short* const bin_item, const float* const bin_weight, const float norm, const unsigned int stride,
		float* const height, const uint2 map_size, const float2 offset) {
		__shared__ Noise::STPStagedPermutation table;
		__shared__ STPBiomeProperty prop[STPHeightfieldKernel::MaxBiome];
		const unsigned int tid = threadIdx.x + threadIdx.y * TileX;
		if (tid < biome_count) {
			prop[tid] = biomes.Property[tid];
		}
		Noise::stage(perm, table);

		const float2 half = make_float2(static_cast<float>(map_size.x) / 2.0f, static_cast<float>(map_size.y) / 2.0f);
		const unsigned int x = blockIdx.x * TileX + threadIdx.x;
		const unsigned int y0 = (blockIdx.y * TileY + threadIdx.y) * PixelPerThread;
		if (x >= map_size.x) {
			return;
		}
		for (unsigned int k = 0u; k < PixelPerThread; k++) {
			const unsigned int y = y0 + k;
			if (y >= map_size.y) {
				break;
			}
			const size_t index = x + static_cast<size_t>(y) * map_size.x;
			if constexpr (Packed) {
				height[index] = blendPacked(table, prop, biome_count, bin_count[index], bin_item + index * stride, norm,
					x, y, offset, half);
			} else {
				height[index] = blendPixel(table, prop, biome_count, bin_count[index], bin_item + index * stride,
					bin_weight + index * stride, x, y, offset, half);
			}
		}
	}

	// Synthetic biome map: nearest jittered site over the 3x3 cell neighbourhood, exact int64
	// distances, first site wins on ties. Site (cx, cy): Philox(key = seed,
	// ctr = (cx, cy, 0x564F524F, 0)); offset = (w0 % cell, w1 % cell); biome = w2 % count.
	__global__ void __launch_bounds__(256) voronoiKERNEL(const unsigned long long seed, const unsigned int biome_count,
		const unsigned int cell, const long long origin_x, const long long origin_y, STPSample_t* const out,
		const unsigned int width, const unsigned int rows, const unsigned int pitch) {
		const unsigned int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
		if (x >= width || y >= rows) {
			return;
		}
		out[x + static_cast<size_t>(y) * pitch] = voronoiBiome(seed, biome_count, cell, origin_x + x, origin_y + y);
	}

	inline dim3 tileGrid(const unsigned int width, const unsigned int rows) {
		return dim3((width + TileX - 1u) / TileX, (rows + TileY * PixelPerThread - 1u) / (TileY * PixelPerThread));
	}

	void checkPermutation(const STPPermutation& perm) {
		if (!perm.Permutation || !perm.Gradient2D || perm.Gradient2DSize == 0u) {
			STP_THROW(STPInvalidArgument, "permutation table is empty");
		}
	}

} // namespace

void STPHeightfieldKernel::simplexHeightmap(const STPPermutation& perm, const Noise::STPFractalSimplexInformation& info,
	const float variation, const float depth, float* const height, const unsigned int width, const unsigned int rows,
	const hipStream_t stream) {
	checkPermutation(perm);
	if (!height || width == 0u || rows == 0u) {
		STP_THROW(STPInvalidArgument, "heightmap output is empty");
	}
	if (info.Octave == 0u) {
		STP_THROW(STPInvalidArgument, "octave count must be positive");
	}
	simplexHeightmapKERNEL<<<tileGrid(width, rows), dim3(TileX, TileY), 0, stream>>>(perm, info, variation, depth, height,
		width, rows);
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::multiBiomeHeightmap(const STPPermutation& perm, const STPBiomeProperty* const table_host,
	const unsigned int biome_count, const STPSingleHistogram& histogram, float* const height, const uint2 map_size,
	const float2 offset, const hipStream_t stream) {
	checkPermutation(perm);
	const BiomeTable t = packBiomeTable(table_host, biome_count);
	if (!histogram.Bin || !histogram.HistogramStartOffset || !height) {
		STP_THROW(STPInvalidArgument, "histogram or heightmap pointer is null");
	}
	multiBiomeKERNEL<<<tileGrid(map_size.x, map_size.y), dim3(TileX, TileY), 0, stream>>>(perm, t, biome_count, histogram,
		height, map_size, offset);
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::multiBiomeHeightmapDense(const STPPermutation& perm, const STPBiomeProperty* const table_host,
	const unsigned int biome_count, const unsigned char* const bin_count, const unsigned short* const bin_item,
	const float* const bin_weight, const unsigned int bin_stride, float* const height, const uint2 map_size,
	const float2 offset, const hipStream_t stream) {
	checkPermutation(perm);
	const BiomeTable t = packBiomeTable(table_host, biome_count);
	multiBiomeDenseKERNEL<false><<<tileGrid(map_size.x, map_size.y), dim3(TileX, TileY), 0, stream>>>(perm, t, biome_count,
		bin_count, bin_item, bin_weight, 0.0f, bin_stride, height, map_size, offset);
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::multiBiomeHeightmapPacked(const STPPermutation& perm, const STPBiomeProperty* const table_host,
	const unsigned int biome_count, const unsigned char* const bin_count, const unsigned short* const bin_slot,
	const float norm, const unsigned int bin_stride, float* const height, const uint2 map_size, const float2 offset,
	const hipStream_t stream) {
	checkPermutation(perm);
	const BiomeTable t = packBiomeTable(table_host, biome_count);
	if (!bin_count || !bin_slot || !height) {
		STP_THROW(STPInvalidArgument, "histogram or heightmap pointer is null");
	}
	multiBiomeDenseKERNEL<true><<<tileGrid(map_size.x, map_size.y), dim3(TileX, TileY), 0, stream>>>(perm, t, biome_count,
		bin_count, bin_slot, nullptr, norm, bin_stride, height, map_size, offset);
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::voronoiBiomemap(const unsigned long long seed, const unsigned int biome_count,
	const unsigned int cell_size, const long long origin_x, const long long origin_y, STPSample_t* const out,
	const unsigned int width, const unsigned int rows, const unsigned int pitch, const hipStream_t stream) {
	if (biome_count == 0u || cell_size == 0u || !out || pitch < width) {
		STP_THROW(STPInvalidArgument, "invalid Voronoi biome map arguments");
	}
	voronoiKERNEL<<<dim3((width + 255u) / 256u, rows), dim3(256u), 0, stream>>>(seed, biome_count, cell_size, origin_x,
		origin_y, out, width, rows, pitch);
	STP_CHECK_HIP(hipGetLastError());
}
