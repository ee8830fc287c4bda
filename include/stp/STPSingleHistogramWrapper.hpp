// Single-histogram output layout and its device iterator.
//
// STPSingleHistogram follows SuperTerrain+/SuperAlgorithm+/Host/Public/SuperAlgorithm+Host/
// STPSingleHistogram.hpp:15-50 (jagged bins + per-pixel start offsets, one extra offset at the
// end holding the bin total). The iterator follows SuperTerrain+/SuperAlgorithm+/Device/Public/
// SuperAlgorithm+Device/STPSingleHistogramWrapper.inl:6-19. Header-only and free of standard
// headers so that hipRTC plugins can include it.
#pragma once
#ifndef STP_AMD_SINGLE_HISTOGRAM_WRAPPER_HPP
#define STP_AMD_SINGLE_HISTOGRAM_WRAPPER_HPP

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace SuperTerrainPlus::STPAlgorithm {

	struct STPSingleHistogram {
		template <typename WT>
		struct STPGenericBin {
			unsigned short Item; // STPSample_t
			WT Weight;
		};
		typedef STPGenericBin<float> STPBin; // 8 bytes: item, 2 pad bytes, weight

		const STPBin* Bin;
		const unsigned int* HistogramStartOffset;
	};

	namespace STPSingleHistogramWrapper {

		// Calls function(item, weight) for every bin of pixel `pixel_index`, in bin order.
		template <class Func>
		__device__ inline void iterate(const STPSingleHistogram& histogram, const unsigned int pixel_index, Func&& function) {
			const unsigned int begin = histogram.HistogramStartOffset[pixel_index],
				end = histogram.HistogramStartOffset[pixel_index + 1u];
			for (unsigned int b = begin; b < end; b++) {
				const STPSingleHistogram::STPBin bin = histogram.Bin[b];
				function(bin.Item, bin.Weight);
			}
		}

	} // namespace STPSingleHistogramWrapper
} // namespace SuperTerrainPlus::STPAlgorithm

#endif // STP_AMD_SINGLE_HISTOGRAM_WRAPPER_HPP
