// Internal declarations shared by the HIP kernels, the C++ host classes and the C ABI.
#pragma once
#ifndef STP_AMD_INTERNAL_HPP
#define STP_AMD_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include <stp/SuperTerrainPlus.hpp>

namespace SuperTerrainPlus {

#define STP_THROW(TYPE, MSG) throw ::SuperTerrainPlus::STPException::TYPE((MSG), __FILE__, __LINE__)
#define STP_ASSERT_ENV(EXPR, WHO)                                                                           \
	do {                                                                                                 \
		if (!(EXPR)) STP_THROW(STPInvalidEnvironment, std::string(WHO) + " validation failed: " #EXPR);      \
	} while (0)
#define STP_ASSERT_DOMAIN(EXPR, MSG)                                                                        \
	do {                                                                                                 \
		if (!(EXPR)) STP_THROW(STPNumericDomainError, MSG);                                                  \
	} while (0)
// STPDeviceErrorHandler.hpp:58 (STP_CHECK_CUDA), HIP edition.
// The runtime keeps the last error per thread until hipGetLastError() reads it; it is read here so
// that an error reported by one call is not reported again by the next launch check.
#define STP_CHECK_HIP(EXPR)                                                                                 \
	do {                                                                                                 \
		const hipError_t stp_err_ = (EXPR);                                                              \
		if (stp_err_ != hipSuccess) {                                                                    \
			(void)hipGetLastError();                                                                     \
			STP_THROW(STPHIPError, std::string(#EXPR) + " -> " + hipGetErrorString(stp_err_));             \
		}                                                                                                \
	} while (0)

	// The float/integer fields of STPRainDropSetting.h:15-53 in the order the kernels read them,
	// plus MaxLifetime (0 = unlimited = the reference's behaviour).
	struct STPRainDropParameter {
		unsigned long long Seed;
		unsigned int RainDropCount;
		float Inertia, SedimentCapacityFactor, minSedimentCapacity, initWaterVolume, minWaterVolume, Friction, initSpeed,
			ErodeSpeed, DepositSpeed, EvaporateSpeed, Gravity;
		unsigned int ErosionBrushRadius, MaxLifetime;
	};

	// generateErosionBrush (STPHeightfieldGenerator.cpp:44-99): brush taps (row-major over the
	// (2r+1)^2 square, strictly inside the radius) as linear offsets in a map of width dimension_x.
	namespace STPErosionBrushGenerator {
		void generate(unsigned int dimension_x, unsigned int radius, std::vector<int>& index, std::vector<float>& weight);
		// The same taps as (dx, dy) offsets, independent of the map width.
		void generateOffsets(unsigned int radius, std::vector<int2>& offset, std::vector<float>& weight);
	}

	namespace STPHeightfieldKernel {

		// Erosion scheduling. Serialized: drops run one after another in index order, which is
		// bit-exact against the FP32 oracle. Parallel: one lane per drop with FP32 atomics, as the
		// reference's release build (STPRainDrop.cu:123-137,150-155).
		enum class STPErosionMode : int { Parallel = 0, Serialized = 1 };

		constexpr unsigned int MaxBiome = 64u;
		constexpr unsigned int MaxNeighbour = 64u;
		constexpr unsigned int MaxBinCapacity = 64u;

		// --- noise / blend (stp_kernels_noise.hip)
		void simplexHeightmap(const STPAlgorithm::STPPermutation& perm,
			const STPAlgorithm::STPSimplexNoise::STPFractalSimplexInformation& info, float variation, float depth,
			float* height, unsigned int width, unsigned int rows, hipStream_t stream);
		void multiBiomeHeightmap(const STPAlgorithm::STPPermutation& perm, const STPBiomeProperty* table_host,
			unsigned int biome_count, const STPAlgorithm::STPSingleHistogram& histogram, float* height, uint2 map_size,
			float2 offset, hipStream_t stream);
		void multiBiomeHeightmapDense(const STPAlgorithm::STPPermutation& perm, const STPBiomeProperty* table_host,
			unsigned int biome_count, const unsigned char* bin_count, const unsigned short* bin_item,
			const float* bin_weight, unsigned int bin_stride, float* height, uint2 map_size, float2 offset,
			hipStream_t stream);
		// the dense blend over packed slots (item | count << 3, weight = count * norm)
		void multiBiomeHeightmapPacked(const STPAlgorithm::STPPermutation& perm, const STPBiomeProperty* table_host,
			unsigned int biome_count, const unsigned char* bin_count, const unsigned short* bin_slot, float norm,
			unsigned int bin_stride, float* height, uint2 map_size, float2 offset, hipStream_t stream);
		void voronoiBiomemap(unsigned long long seed, unsigned int biome_count, unsigned int cell_size, long long origin_x,
			long long origin_y, STPSample_t* out, unsigned int width, unsigned int rows, unsigned int pitch,
			hipStream_t stream);

		// --- erosion / format (stp_kernels_erosion.hip)
		// gate (parallel mode only, device words): when not null the drops run only if gate[0] or
		// gate[1] is nonzero, so the decision can be made on the device by earlier work on `stream`
		void hydraulicErosion(float* height, const STPNearestNeighbourInformation& nn, const STPRainDropParameter& rd,
			const int* brush_index, const float* brush_weight, unsigned int brush_size, unsigned int stream_id,
			STPErosionMode mode, float* ledger, uint2* drop_info, hipStream_t stream, const unsigned int* gate = nullptr);
		void formatHeightmap(const float* input, STPHeightFixed_t* output, unsigned long long count, hipStream_t stream);
		// gate[0] = 1 if a height of map[0, count) lies outside [low, high] or is NaN, else 0; gate[1] = 0
		void heightGate(const float* map, unsigned long long count, float low, float high, unsigned int* gate,
			hipStream_t stream);
		// gate[1] = *status
		void copyGateStatus(const unsigned int* status, unsigned int* gate, hipStream_t stream);
		// map[i] = source[i] for i < count, unless gate[0] or gate[1] is nonzero
		void selectUngated(const unsigned int* gate, const float* source, float* map, unsigned long long count,
			hipStream_t stream);

		// --- free-slip merge (stp_kernels_nn.hip)
		void mergeNeighbour(const void* const* chunks, void* merged, unsigned int element_size,
			const STPNearestNeighbourInformation& nn, hipStream_t stream);
		void unmergeNeighbour(const void* merged, void* const* chunks, unsigned int element_size,
			const STPNearestNeighbourInformation& nn, hipStream_t stream);
		// unmerge of the FP32 map into `height` with each pixel's UINT16 format into `fixed` (one pass)
		void unmergeFormatNeighbour(const float* merged, float* const* height, STPHeightFixed_t* const* fixed,
			const STPNearestNeighbourInformation& nn, hipStream_t stream);

		// --- single histogram filter (stp_kernels_shf.hip)
		struct STPFilterWorkspace {
			// sizes in elements, all device memory owned by the caller of filter()
			unsigned int* VerticalBin;     // packed item | count << 16, [row][slot][column]
			unsigned char* VerticalCount;  // [row][column]
			unsigned char* BinCount;       // per output pixel, row-major
			unsigned short* BinItem;       // [pixel][slot]
			float* BinWeight;              // [pixel][slot]
			unsigned int* Offset;          // map pixels + 1
			unsigned int* BlockSum;        // scan scratch
			STPAlgorithm::STPSingleHistogram::STPBin* Bin; // jagged output, capacity pixels * stride
			unsigned int* Overflow;        // one word, set when a histogram needs more bins than the stride
		};
		// Bytes of device workspace filter() needs for the given problem.
		size_t filterWorkspaceBytes(uint2 map_size, unsigned int radius, unsigned int bin_capacity);
		STPFilterWorkspace carveFilterWorkspace(void* base, uint2 map_size, unsigned int radius, unsigned int bin_capacity);
		// Checks the filter domain of a diversity neighbourhood (even radius, window inside the
		// merged map) and returns the window's first pixel (the centre chunk's origin).
		uint2 checkFilterDomain(const STPNearestNeighbourInformation& nn, unsigned int radius);
		// Runs vertical + horizontal passes; writes dense bins and (if jagged) the jagged output.
		void singleHistogramFilter(const STPSample_t* sample, unsigned int pitch, uint2 start, uint2 map_size,
			unsigned int radius, unsigned int bin_capacity, bool jagged, const STPFilterWorkspace& ws, hipStream_t stream);
		// The reference's jagged histogram (STPSingleHistogram: bins of all pixels back to back, and
		// each pixel's first bin) of a dense one: per pixel count[p] slots of bin_capacity, holding
		// items with `weight` the weights, or (weight null) packed item | count << 3 with weight
		// count * norm. The jagged arrays live in `workspace` (jaggedWorkspaceBytes).
		size_t jaggedWorkspaceBytes(unsigned long long pixels, unsigned int bin_capacity);
		STPAlgorithm::STPSingleHistogram jaggedHistogram(const unsigned char* count, const unsigned short* item,
			const float* weight, float norm, unsigned int bin_capacity, unsigned int pixels, void* workspace,
			hipStream_t stream);

	} // namespace STPHeightfieldKernel
} // namespace SuperTerrainPlus

#endif // STP_AMD_INTERNAL_HPP
