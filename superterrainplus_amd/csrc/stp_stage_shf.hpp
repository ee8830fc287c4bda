// Device stage functions of the single histogram filter, shared by the per-chunk kernels
// (stp_kernels_shf.hip) and the batched chunk-table kernels (stp_batch.hip).
//
// Behaviour follows SuperTerrain+/SuperAlgorithm+/Host/Private/STPSingleHistogramFilter.cpp:
//   vertical pass   :522-563 (one sliding window of 2r+1 rows per column),
//   horizontal pass :608-679 (one sliding window of 2r+1 column-histograms per row),
//   accumulator     :368-468 (bins kept in insertion order; a bin whose count reaches zero is
//                   erased and the later bins shift down; the entering line is added before the
//                   leaving line is removed),
//   normalisation   :483-506 (weight = count * (1 / sum of counts)).
// Keeping those rules makes the bin ORDER identical to the reference, which matters because the
// blend sums weight * noise in bin order.
//
// One lane owns one line; its accumulator lives in LDS as `cap` slots strided by the workgroup
// width (ShfLineBlock), so slot k of the 64 lanes of a wave is 64 consecutive dwords.
#pragma once
#ifndef STP_AMD_STAGE_SHF_HPP
#define STP_AMD_STAGE_SHF_HPP

#include <hip/hip_runtime.h>

namespace SuperTerrainPlus::STPStage {

	constexpr unsigned int ShfLineBlock = 64u;

	struct VerticalAcc {
		unsigned int* slot; // packed item | count << 16, slot k at slot[k * ShfLineBlock]
		unsigned int n;
		unsigned int cap;
		bool overflow;

		__device__ inline void increment(const unsigned int item) {
			for (unsigned int k = 0u; k < n; k++) {
				const unsigned int e = slot[k * ShfLineBlock];
				if ((e & 0xFFFFu) == item) {
					slot[k * ShfLineBlock] = e + 0x10000u;
					return;
				}
			}
			if (n < cap) {
				slot[n * ShfLineBlock] = item | 0x10000u;
				n++;
			} else {
				overflow = true;
			}
		}
		__device__ inline void decrement(const unsigned int item) {
			for (unsigned int k = 0u; k < n; k++) {
				const unsigned int e = slot[k * ShfLineBlock];
				if ((e & 0xFFFFu) == item) {
					if ((e >> 16) <= 1u) {
						for (unsigned int m = k; m + 1u < n; m++) {
							slot[m * ShfLineBlock] = slot[(m + 1u) * ShfLineBlock];
						}
						n--;
					} else {
						slot[k * ShfLineBlock] = e - 0x10000u;
					}
					return;
				}
			}
		}
	};

	// Insertion-ordered bins held in registers, for a compile-time slot bound MaxCap: the same
	// rules as VerticalAcc / HorizontalAcc (search, append at the end, erase at zero with the later
	// bins shifting down), with every slot index static so the bins never leave registers. `cap`
	// (<= MaxCap) is the run-time capacity that decides overflow.
	template <unsigned int MaxCap>
	struct RegisterAcc {
		unsigned int Item[MaxCap], Count[MaxCap];
		unsigned int n, cap;
		bool overflow;

		__device__ explicit RegisterAcc(const unsigned int capacity) : n(0u), cap(capacity), overflow(false) {
#pragma unroll
			for (unsigned int k = 0u; k < MaxCap; k++) {
				this->Item[k] = 0xFFFFFFFFu;
				this->Count[k] = 0u;
			}
		}
		__device__ inline void increment(const unsigned int it, const unsigned int c) {
			bool found = false;
#pragma unroll
			for (unsigned int k = 0u; k < MaxCap; k++) {
				if (k < this->n && this->Item[k] == it) {
					this->Count[k] += c;
					found = true;
				}
			}
			if (found) {
				return;
			}
			if (this->n < this->cap) {
#pragma unroll
				for (unsigned int k = 0u; k < MaxCap; k++) {
					if (k == this->n) {
						this->Item[k] = it;
						this->Count[k] = c;
					}
				}
				this->n++;
			} else {
				this->overflow = true;
			}
		}
		__device__ inline void decrement(const unsigned int it, const unsigned int c) {
			unsigned int pos = MaxCap, q = 0u;
#pragma unroll
			for (unsigned int k = 0u; k < MaxCap; k++) {
				if (pos == MaxCap && k < this->n && this->Item[k] == it) {
					pos = k;
					q = this->Count[k];
				}
			}
			if (pos == MaxCap) {
				return;
			}
			if (q <= c) {
#pragma unroll
				for (unsigned int k = 0u; k + 1u < MaxCap; k++) {
					if (k >= pos) {
						this->Item[k] = this->Item[k + 1u];
						this->Count[k] = this->Count[k + 1u];
					}
				}
				this->n--;
			} else {
#pragma unroll
				for (unsigned int k = 0u; k < MaxCap; k++) {
					if (k == pos) {
						this->Count[k] = q - c;
					}
				}
			}
		}
		// f(slot, item, count) for every bin, in order
		template <class F>
		__device__ inline void forEach(F&& f) const {
#pragma unroll
			for (unsigned int k = 0u; k < MaxCap; k++) {
				if (k < this->n) {
					f(k, this->Item[k], this->Count[k]);
				}
			}
		}
	};

	// The LDS accumulators behind the same interface (any capacity).
	struct LdsVerticalAcc {
		VerticalAcc a;
		__device__ LdsVerticalAcc(unsigned int* const slot, const unsigned int capacity) : a {slot, 0u, capacity, false} {}
		__device__ inline void increment(const unsigned int it, unsigned int) { this->a.increment(it); }
		__device__ inline void decrement(const unsigned int it, unsigned int) { this->a.decrement(it); }
		template <class F>
		__device__ inline void forEach(F&& f) const {
			for (unsigned int k = 0u; k < this->a.n; k++) {
				const unsigned int e = this->a.slot[k * ShfLineBlock];
				f(k, e & 0xFFFFu, e >> 16);
			}
		}
		__device__ unsigned int size() const { return this->a.n; }
		__device__ bool overflowed() const { return this->a.overflow; }
	};

	template <unsigned int MaxCap>
	struct RegisterLineAcc : RegisterAcc<MaxCap> {
		__device__ explicit RegisterLineAcc(const unsigned int capacity) : RegisterAcc<MaxCap>(capacity) {}
		__device__ unsigned int size() const { return this->n; }
		__device__ bool overflowed() const { return this->overflow; }
	};

	// Vertical pass over column `c` (0 <= c < map.x + 2r) of the window whose central map starts at
	// `start` in a sample map with row pitch `pitch`. Writes vcount[y][c] and vbin[y][slot][c].
	// Returns true when a histogram needed more than `vcap` bins.
	__device__ inline bool shfVerticalLine(const unsigned short* const sample, const unsigned int pitch,
		const uint2 start, const uint2 map, const unsigned int radius, const unsigned int vcap, unsigned int* const lds,
		unsigned int* const vbin, unsigned char* const vcount, const unsigned int c) {
		const unsigned int ncols = map.x + 2u * radius;
		VerticalAcc acc {lds, 0u, vcap, false};
		const unsigned short* const col = sample + (start.x - radius + c);
		const size_t top = start.y - radius;
		for (unsigned int j = 0u; j <= 2u * radius; j++) {
			acc.increment(col[(top + j) * pitch]);
		}
		for (unsigned int y = 0u; y < map.y; y++) {
			if (y > 0u) {
				acc.increment(col[(static_cast<size_t>(start.y) + y + radius) * pitch]);
				acc.decrement(col[(static_cast<size_t>(start.y) + y - radius - 1u) * pitch]);
			}
			vcount[static_cast<size_t>(y) * ncols + c] = static_cast<unsigned char>(acc.n);
			for (unsigned int k = 0u; k < acc.n; k++) {
				vbin[(static_cast<size_t>(y) * vcap + k) * ncols + c] = acc.slot[k * ShfLineBlock];
			}
		}
		return acc.overflow;
	}

	struct HorizontalAcc {
		unsigned int* item; // slot k at item[k * ShfLineBlock]
		unsigned int* count;
		unsigned int n;
		unsigned int cap;
		bool overflow;

		__device__ inline void increment(const unsigned int it, const unsigned int c) {
			for (unsigned int k = 0u; k < n; k++) {
				if (item[k * ShfLineBlock] == it) {
					count[k * ShfLineBlock] += c;
					return;
				}
			}
			if (n < cap) {
				item[n * ShfLineBlock] = it;
				count[n * ShfLineBlock] = c;
				n++;
			} else {
				overflow = true;
			}
		}
		__device__ inline void decrement(const unsigned int it, const unsigned int c) {
			for (unsigned int k = 0u; k < n; k++) {
				if (item[k * ShfLineBlock] == it) {
					const unsigned int q = count[k * ShfLineBlock];
					if (q <= c) {
						for (unsigned int m = k; m + 1u < n; m++) {
							item[m * ShfLineBlock] = item[(m + 1u) * ShfLineBlock];
							count[m * ShfLineBlock] = count[(m + 1u) * ShfLineBlock];
						}
						n--;
					} else {
						count[k * ShfLineBlock] = q - c;
					}
					return;
				}
			}
		}
	};

	struct LdsHorizontalAcc {
		HorizontalAcc a;
		__device__ LdsHorizontalAcc(unsigned int* const item, unsigned int* const count, const unsigned int capacity) :
			a {item, count, 0u, capacity, false} {}
		__device__ inline void increment(const unsigned int it, const unsigned int c) { this->a.increment(it, c); }
		__device__ inline void decrement(const unsigned int it, const unsigned int c) { this->a.decrement(it, c); }
		template <class F>
		__device__ inline void forEach(F&& f) const {
			for (unsigned int k = 0u; k < this->a.n; k++) {
				f(k, this->a.item[k * ShfLineBlock], this->a.count[k * ShfLineBlock]);
			}
		}
		__device__ unsigned int size() const { return this->a.n; }
		__device__ bool overflowed() const { return this->a.overflow; }
	};

	// Horizontal pass over row `y` (0 <= y < map.y) of the vertical output. Writes the dense
	// histogram of every pixel of the row: bin_count[p], bin_item/bin_weight[p * cap + slot].
	// `lds` holds 2 * cap * ShfLineBlock dwords starting at this lane's column.
	__device__ inline bool shfHorizontalLine(const uint2 map, const unsigned int radius, const unsigned int vcap,
		const unsigned int cap, const unsigned int* const vbin, const unsigned char* const vcount,
		unsigned int* const lds, unsigned char* const bin_count, unsigned short* const bin_item,
		float* const bin_weight, const unsigned int y) {
		const unsigned int ncols = map.x + 2u * radius;
		HorizontalAcc acc {lds, lds + cap * ShfLineBlock, 0u, cap, false};
		const unsigned char* const rowCount = vcount + static_cast<size_t>(y) * ncols;
		const unsigned int* const rowBin = vbin + static_cast<size_t>(y) * vcap * ncols;
		const auto add = [&](const unsigned int column, const bool inc) {
			const unsigned int nc = rowCount[column];
			for (unsigned int k = 0u; k < nc; k++) {
				const unsigned int e = rowBin[static_cast<size_t>(k) * ncols + column];
				if (inc) {
					acc.increment(e & 0xFFFFu, e >> 16);
				} else {
					acc.decrement(e & 0xFFFFu, e >> 16);
				}
			}
		};
		for (unsigned int c = 0u; c <= 2u * radius; c++) {
			add(c, true);
		}
		for (unsigned int x = 0u; x < map.x; x++) {
			if (x > 0u) {
				add(x + 2u * radius, true);
				add(x - 1u, false);
			}
			const size_t p = static_cast<size_t>(y) * map.x + x;
			unsigned int sum = 0u;
			for (unsigned int k = 0u; k < acc.n; k++) {
				sum += acc.count[k * ShfLineBlock];
			}
			const float norm = 1.0f / static_cast<float>(sum);
			bin_count[p] = static_cast<unsigned char>(acc.n);
			for (unsigned int k = 0u; k < acc.n; k++) {
				bin_item[p * cap + k] = static_cast<unsigned short>(acc.item[k * ShfLineBlock]);
				bin_weight[p * cap + k] = static_cast<float>(acc.count[k * ShfLineBlock]) * norm;
			}
		}
		return acc.overflow;
	}

} // namespace SuperTerrainPlus::STPStage

#endif // STP_AMD_STAGE_SHF_HPP
