// Single histogram filter on the GPU.
//
// Behaviour follows SuperTerrain+/SuperAlgorithm+/Host/Private/STPSingleHistogramFilter.cpp:
//   vertical pass   :522-563 (one sliding window of 2r+1 rows per column),
//   horizontal pass :608-679 (one sliding window of 2r+1 column-histograms per row),
//   accumulator     :368-468 (bins kept in insertion order; a bin whose count reaches zero is
//                   erased and the later bins shift down; the entering line is added before the
//                   leaving line is removed),
//   normalisation   :483-506 (weight = count * (1 / sum of counts)).
// Keeping those rules makes the bin ORDER identical to the reference, which matters because
// the blend sums weight * noise in bin order.
//
// MI355X design: one lane per line (column for the vertical pass, row for the horizontal pass),
// each lane's accumulator lives in LDS as `cap` slots strided by the workgroup width, so slot k of
// all 64 lanes of a wave sits in 64 consecutive dwords (conflict-free). The vertical pass reads
// 64 consecutive samples per wave per row (coalesced) and writes its per-cell bins as
// [row][slot][column] (coalesced). Lines are independent, so the passes scale with the number
// of lines of all chunks in flight. The dense output [pixel][slot] feeds the fused blend
// directly; the jagged reference layout is produced by a scan + scatter.
#include "stp_internal.hpp"
#include "stp_stage_shf.hpp"

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPHeightfieldKernel;

namespace {

	constexpr unsigned int LineBlock = STPStage::ShfLineBlock;
	constexpr unsigned int ScanBlock = 256u, ScanItems = 8u, ScanTile = ScanBlock * ScanItems;

	// ------------------------------------------------------------ line passes
	__global__ void __launch_bounds__(LineBlock) verticalKERNEL(const STPSample_t* const sample, const unsigned int pitch,
		const uint2 start, const uint2 map, const unsigned int radius, const unsigned int vcap, unsigned int* const vbin,
		unsigned char* const vcount, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		const unsigned int c = blockIdx.x * LineBlock + threadIdx.x;
		if (c >= map.x + 2u * radius) {
			return;
		}
		if (STPStage::shfVerticalLine(sample, pitch, start, map, radius, vcap, lds + threadIdx.x, vbin, vcount, c)) {
			atomicOr(overflow, 1u);
		}
	}

	__global__ void __launch_bounds__(LineBlock) horizontalKERNEL(const uint2 map, const unsigned int radius,
		const unsigned int vcap, const unsigned int cap, const unsigned int* const vbin,
		const unsigned char* const vcount, unsigned char* const bin_count, unsigned short* const bin_item,
		float* const bin_weight, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		const unsigned int y = blockIdx.x * LineBlock + threadIdx.x;
		if (y >= map.y) {
			return;
		}
		if (STPStage::shfHorizontalLine(map, radius, vcap, cap, vbin, vcount, lds + threadIdx.x, bin_count, bin_item,
				bin_weight, y)) {
			atomicOr(overflow, 1u);
		}
	}

	// ------------------------------------------------------------ scan + scatter to jagged
	__device__ inline unsigned int blockExclusiveScan(const unsigned int v, unsigned int* const waveTotal,
		unsigned int& blockTotal) {
		const unsigned int lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
		unsigned int inc = v;
#pragma unroll
		for (unsigned int d = 1u; d < 64u; d <<= 1) {
			const unsigned int o = __shfl_up(inc, d, 64);
			if (lane >= d) {
				inc += o;
			}
		}
		if (lane == 63u) {
			waveTotal[wave] = inc;
		}
		__syncthreads();
		unsigned int waveBase = 0u;
		blockTotal = 0u;
		for (unsigned int w = 0u; w < ScanBlock / 64u; w++) {
			const unsigned int t = waveTotal[w];
			if (w < wave) {
				waveBase += t;
			}
			blockTotal += t;
		}
		__syncthreads();
		return waveBase + inc - v;
	}

	__global__ void __launch_bounds__(ScanBlock) tileSumKERNEL(const unsigned char* const count, const unsigned int n,
		unsigned int* const tile_sum) {
		__shared__ unsigned int waveTotal[ScanBlock / 64u];
		const size_t base = static_cast<size_t>(blockIdx.x) * ScanTile + threadIdx.x * ScanItems;
		unsigned int s = 0u;
		for (unsigned int k = 0u; k < ScanItems; k++) {
			if (base + k < n) {
				s += count[base + k];
			}
		}
		unsigned int total;
		blockExclusiveScan(s, waveTotal, total);
		if (threadIdx.x == 0u) {
			tile_sum[blockIdx.x] = total;
		}
	}

	__global__ void __launch_bounds__(ScanBlock) tileScanKERNEL(unsigned int* const tile_sum, const unsigned int tiles) {
		__shared__ unsigned int waveTotal[ScanBlock / 64u];
		unsigned int carry = 0u;
		for (unsigned int base = 0u; base < tiles; base += ScanBlock) {
			const unsigned int i = base + threadIdx.x;
			const unsigned int v = i < tiles ? tile_sum[i] : 0u;
			unsigned int total;
			const unsigned int ex = blockExclusiveScan(v, waveTotal, total);
			if (i < tiles) {
				tile_sum[i] = carry + ex;
			}
			carry += total;
		}
	}

	// Packed: a slot is item | count << 3 and the weight count * norm (no weight array)
	template <bool Packed>
	__global__ void __launch_bounds__(ScanBlock) scatterKERNEL(const unsigned char* const count,
		const unsigned short* const item, const float* const weight, const float norm, const unsigned int cap,
		const unsigned int n, const unsigned int* const tile_base, unsigned int* const offset,
		STPAlgorithm::STPSingleHistogram::STPBin* const bin) {
		__shared__ unsigned int waveTotal[ScanBlock / 64u];
		const size_t base = static_cast<size_t>(blockIdx.x) * ScanTile + threadIdx.x * ScanItems;
		unsigned int local[ScanItems];
		unsigned int s = 0u;
		for (unsigned int k = 0u; k < ScanItems; k++) {
			local[k] = (base + k < n) ? count[base + k] : 0u;
			s += local[k];
		}
		unsigned int total;
		unsigned int o = tile_base[blockIdx.x] + blockExclusiveScan(s, waveTotal, total);
		for (unsigned int k = 0u; k < ScanItems; k++) {
			const size_t p = base + k;
			if (p < n) {
				offset[p] = o;
				for (unsigned int b = 0u; b < local[k]; b++) {
					STPAlgorithm::STPSingleHistogram::STPBin out;
					if constexpr (Packed) {
						const unsigned int slot = item[p * cap + b];
						out.Item = static_cast<unsigned short>(slot & 7u);
						out.Weight = static_cast<float>(slot >> 3u) * norm;
					} else {
						out.Item = item[p * cap + b];
						out.Weight = weight[p * cap + b];
					}
					bin[o + b] = out;
				}
				o += local[k];
				if (p + 1u == n) {
					offset[n] = o;
				}
			}
		}
	}

	__global__ void zeroWordKERNEL(unsigned int* const word) {
		*word = 0u;
	}

	inline size_t align256(const size_t v) {
		return (v + 255u) & ~static_cast<size_t>(255u);
	}

	struct Layout {
		size_t vbin, vcount, count, item, weight, offset, blocksum, bin, overflow, total;
	};

	Layout layout(const uint2 map, const unsigned int radius, const unsigned int cap) {
		const size_t ncols = map.x + 2ull * radius, pixels = static_cast<size_t>(map.x) * map.y;
		const unsigned int vcap = cap < 2u * radius + 1u ? cap : 2u * radius + 1u;
		const size_t tiles = (pixels + ScanTile - 1u) / ScanTile;
		Layout l;
		size_t at = 0u;
		const auto take = [&at](const size_t bytes) {
			const size_t here = at;
			at += align256(bytes);
			return here;
		};
		l.vbin = take(ncols * map.y * vcap * sizeof(unsigned int));
		l.vcount = take(ncols * map.y);
		l.count = take(pixels);
		l.item = take(pixels * cap * sizeof(unsigned short));
		l.weight = take(pixels * cap * sizeof(float));
		l.offset = take((pixels + 1u) * sizeof(unsigned int));
		l.blocksum = take(tiles * sizeof(unsigned int));
		l.bin = take(pixels * cap * sizeof(STPAlgorithm::STPSingleHistogram::STPBin));
		l.overflow = take(sizeof(unsigned int));
		l.total = at;
		return l;
	}

} // namespace

size_t STPHeightfieldKernel::filterWorkspaceBytes(const uint2 map_size, const unsigned int radius,
	const unsigned int bin_capacity) {
	return layout(map_size, radius, bin_capacity).total;
}

STPHeightfieldKernel::STPFilterWorkspace STPHeightfieldKernel::carveFilterWorkspace(void* const base,
	const uint2 map_size, const unsigned int radius, const unsigned int bin_capacity) {
	const Layout l = layout(map_size, radius, bin_capacity);
	char* const b = static_cast<char*>(base);
	STPFilterWorkspace ws;
	ws.VerticalBin = reinterpret_cast<unsigned int*>(b + l.vbin);
	ws.VerticalCount = reinterpret_cast<unsigned char*>(b + l.vcount);
	ws.BinCount = reinterpret_cast<unsigned char*>(b + l.count);
	ws.BinItem = reinterpret_cast<unsigned short*>(b + l.item);
	ws.BinWeight = reinterpret_cast<float*>(b + l.weight);
	ws.Offset = reinterpret_cast<unsigned int*>(b + l.offset);
	ws.BlockSum = reinterpret_cast<unsigned int*>(b + l.blocksum);
	ws.Bin = reinterpret_cast<STPAlgorithm::STPSingleHistogram::STPBin*>(b + l.bin);
	ws.Overflow = reinterpret_cast<unsigned int*>(b + l.overflow);
	return ws;
}

void STPHeightfieldKernel::singleHistogramFilter(const STPSample_t* const sample, const unsigned int pitch,
	const uint2 start, const uint2 map_size, const unsigned int radius, const unsigned int bin_capacity,
	const bool jagged, const STPFilterWorkspace& ws, const hipStream_t stream) {
	// STPSingleHistogramFilter.cpp:874-880 (domain) -- the caller has already checked that the
	// (map + 2r)^2 window around `start` lies inside the sample map.
	STP_ASSERT_DOMAIN(radius > 0u && (radius & 1u) == 0u, "radius should be an positive even number");
	STP_ASSERT_DOMAIN(radius <= start.x && radius <= start.y,
		"radius is too large and will overflow nearest neighbour boundary");
	if (bin_capacity == 0u || bin_capacity > MaxBinCapacity) {
		STP_THROW(STPInvalidArgument, "histogram bin capacity must be in [1, 64]");
	}
	if (!sample || map_size.x == 0u || map_size.y == 0u) {
		STP_THROW(STPInvalidArgument, "empty sample map");
	}
	const unsigned int ncols = map_size.x + 2u * radius;
	const unsigned int vcap = bin_capacity < 2u * radius + 1u ? bin_capacity : 2u * radius + 1u;
	zeroWordKERNEL<<<1, 1, 0, stream>>>(ws.Overflow);
	verticalKERNEL<<<(ncols + LineBlock - 1u) / LineBlock, LineBlock, vcap * LineBlock * sizeof(unsigned int), stream>>>(
		sample, pitch, start, map_size, radius, vcap, ws.VerticalBin, ws.VerticalCount, ws.Overflow);
	STP_CHECK_HIP(hipGetLastError());
	horizontalKERNEL<<<(map_size.y + LineBlock - 1u) / LineBlock, LineBlock,
		2u * bin_capacity * LineBlock * sizeof(unsigned int), stream>>>(map_size, radius, vcap, bin_capacity,
		ws.VerticalBin, ws.VerticalCount, ws.BinCount, ws.BinItem, ws.BinWeight, ws.Overflow);
	STP_CHECK_HIP(hipGetLastError());
	if (jagged) {
		const unsigned int pixels = map_size.x * map_size.y;
		const unsigned int tiles = (pixels + ScanTile - 1u) / ScanTile;
		tileSumKERNEL<<<tiles, ScanBlock, 0, stream>>>(ws.BinCount, pixels, ws.BlockSum);
		tileScanKERNEL<<<1, ScanBlock, 0, stream>>>(ws.BlockSum, tiles);
		scatterKERNEL<false><<<tiles, ScanBlock, 0, stream>>>(ws.BinCount, ws.BinItem, ws.BinWeight, 0.0f, bin_capacity,
			pixels, ws.BlockSum, ws.Offset, ws.Bin);
		STP_CHECK_HIP(hipGetLastError());
	}
}

size_t STPHeightfieldKernel::jaggedWorkspaceBytes(const unsigned long long pixels, const unsigned int bin_capacity) {
	const unsigned long long tiles = (pixels + ScanTile - 1u) / ScanTile;
	return align256((pixels + 1ull) * sizeof(unsigned int)) + align256(tiles * sizeof(unsigned int))
		+ align256(pixels * bin_capacity * sizeof(STPAlgorithm::STPSingleHistogram::STPBin));
}

STPAlgorithm::STPSingleHistogram STPHeightfieldKernel::jaggedHistogram(const unsigned char* const count,
	const unsigned short* const item, const float* const weight, const float norm, const unsigned int bin_capacity,
	const unsigned int pixels, void* const workspace, const hipStream_t stream) {
	if (!count || !item || !workspace || pixels == 0u || bin_capacity == 0u || bin_capacity > MaxBinCapacity) {
		STP_THROW(STPInvalidArgument, "invalid dense histogram");
	}
	const unsigned int tiles = (pixels + ScanTile - 1u) / ScanTile;
	char* const b = static_cast<char*>(workspace);
	unsigned int* const offset = reinterpret_cast<unsigned int*>(b);
	unsigned int* const blocksum = reinterpret_cast<unsigned int*>(b + align256((pixels + 1ull) * sizeof(unsigned int)));
	auto* const bin = reinterpret_cast<STPAlgorithm::STPSingleHistogram::STPBin*>(
		b + align256((pixels + 1ull) * sizeof(unsigned int)) + align256(static_cast<size_t>(tiles) * sizeof(unsigned int)));
	tileSumKERNEL<<<tiles, ScanBlock, 0, stream>>>(count, pixels, blocksum);
	tileScanKERNEL<<<1, ScanBlock, 0, stream>>>(blocksum, tiles);
	if (weight) {
		scatterKERNEL<false><<<tiles, ScanBlock, 0, stream>>>(count, item, weight, 0.0f, bin_capacity, pixels, blocksum,
			offset, bin);
	} else {
		scatterKERNEL<true><<<tiles, ScanBlock, 0, stream>>>(count, item, nullptr, norm, bin_capacity, pixels, blocksum,
			offset, bin);
	}
	STP_CHECK_HIP(hipGetLastError());
	return STPAlgorithm::STPSingleHistogram {bin, offset};
}
