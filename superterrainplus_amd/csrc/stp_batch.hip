// Batched heightfield generation: many chunks of uneven size through the whole path
// (biome map -> single histogram filter -> multi-biome noise blend -> erosion -> UINT16) with
// one launch per stage for ALL chunks, captured once into a hipGraph and replayed.
//
// No reference counterpart: the reference generates one chunk per generate()/erode() call, each
// ending with a stream synchronisation (SuperTerrain+/SuperTerrain+/Private/World/Chunk/
// STPHeightfieldGenerator.cpp:158-197), and every chunk has the same map size
// (STPChunkSetting.h:22). The "plain" mode here reproduces that calling pattern with the same
// kernels (one chunk per launch, host synchronisation after each chunk) so the two can be
// compared on equal work.
//
// Work decomposition: each stage has a tile table {chunk, tile} built once on the host; a
// workgroup reads its entry and processes one tile of one chunk, so uneven chunks share a
// single launch without padding and the grid is large enough to fill all 256 CUs. All chunk
// data lives in resident arenas (heights, UINT16 copies, halo biome maps, filter workspace).
#include "stp_batch.hpp"
#include "stp_stage_erosion.hpp"
#include "stp_stage_noise.hpp"
#include "stp_stage_shf.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPStage;

namespace SuperTerrainPlus::STPBatch {

	constexpr unsigned int BiomeTile = 256u;
	constexpr unsigned int BlendX = 64u, BlendY = 4u, BlendPixel = 4u;
	constexpr unsigned int ErosionBlock = 256u;

	// Synthetic biome map over a 32 x 32 tile of the halo map, BiomeTile threads with
	// BiomeSide / BiomeRows samples each. The Voronoi sites of every grid cell the tile's samples
	// look at (their cells +- 1) are drawn once into LDS, relative to the tile's first world sample
	// and packed as two 16-bit integers; each sample then takes the nearest of its 3 x 3 cells in
	// the order and with the tie rule of voronoiBiome. While the tile's extent in world units stays
	// below 2^15 every offset fits 16 bits and every squared distance is an exact 32-bit integer, so
	// the result is voronoiBiome's; larger extents (very coarse levels of detail, huge cells) call
	// voronoiBiome per sample.
	constexpr unsigned int BiomeSide = 32u, BiomeRows = BiomeTile / BiomeSide, BiomeCells = 1024u;
	__global__ void __launch_bounds__(BiomeTile) biomeKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const unsigned long long seed, const unsigned int count, const unsigned int cell, const unsigned int radius,
		unsigned short* const biome) {
		__shared__ unsigned int site[BiomeCells];
		__shared__ unsigned short siteBiome[BiomeCells];
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int w = c.Side + 2u * radius, tilesX = (w + BiomeSide - 1u) / BiomeSide;
		const unsigned int x0 = (t.y % tilesX) * BiomeSide, y0 = (t.y / tilesX) * BiomeSide;
		const unsigned int tx = threadIdx.x % BiomeSide, ty = threadIdx.x / BiomeSide;
		// halo pixel (x, y) is chunk pixel (x - r, y - r), at world Origin + pixel * Step
		const long long step = c.Step;
		const long long wx0 = static_cast<long long>(c.OriginX) + (static_cast<long long>(x0) - radius) * step,
			wy0 = static_cast<long long>(c.OriginY) + (static_cast<long long>(y0) - radius) * step;
		const long long span = static_cast<long long>(BiomeSide - 1u) * step;
		const long long cx0 = floorDiv(wx0, cell) - 1, cy0 = floorDiv(wy0, cell) - 1;
		const long long nx = floorDiv(wx0 + span, cell) + 1 - cx0 + 1, ny = floorDiv(wy0 + span, cell) + 1 - cy0 + 1;
		const unsigned int x = x0 + tx;
		// every offset below lies within (-(2 cell + span), 2 cell + span)
		if (nx * ny > static_cast<long long>(BiomeCells) || 2ll * cell + span >= 32768ll) {
			for (unsigned int j = 0u; j < BiomeSide; j += BiomeRows) {
				const unsigned int y = y0 + ty + j;
				if (x < w && y < w) {
					biome[c.Halo + static_cast<size_t>(y) * w + x] = voronoiBiome(seed, count, cell,
						wx0 + static_cast<long long>(tx) * step, wy0 + static_cast<long long>(ty + j) * step);
				}
			}
			return;
		}
		const unsigned int cells = static_cast<unsigned int>(nx * ny), cols = static_cast<unsigned int>(nx);
		for (unsigned int k = threadIdx.x; k < cells; k += BiomeTile) {
			const long long ncx = cx0 + static_cast<long long>(k % cols), ncy = cy0 + static_cast<long long>(k / cols);
			const STPRandom::Philox4 r = STPRandom::philox4x32_10(static_cast<unsigned int>(ncx),
				static_cast<unsigned int>(ncy), 0x564F524Fu, 0u, static_cast<unsigned int>(seed),
				static_cast<unsigned int>(seed >> 32));
			// site relative to the tile's first world sample
			const int sx = static_cast<int>(ncx * cell + (r.v[0] % cell) - wx0),
				sy = static_cast<int>(ncy * cell + (r.v[1] % cell) - wy0);
			site[k] = (static_cast<unsigned int>(sx) & 0xFFFFu) | (static_cast<unsigned int>(sy) << 16);
			siteBiome[k] = static_cast<unsigned short>(r.v[2] % count);
		}
		__syncthreads();
		if (x >= w) {
			return;
		}
		// 32-bit from here: sample offsets from the tile's first sample, and from cell (cx0, cy0)
		const int istep = static_cast<int>(step);
		const int px = static_cast<int>(tx) * istep;
		const unsigned int lx = (static_cast<unsigned int>(wx0 - cx0 * cell) + static_cast<unsigned int>(px)) / cell;
		const unsigned int ey0 = static_cast<unsigned int>(wy0 - cy0 * cell);
		for (unsigned int j = 0u; j < BiomeSide; j += BiomeRows) {
			const unsigned int y = y0 + ty + j;
			if (y >= w) {
				break;
			}
			const int py = static_cast<int>(ty + j) * istep;
			const unsigned int ly = (ey0 + static_cast<unsigned int>(py)) / cell;
			int best = 0x7FFFFFFF;
			unsigned int bk = 0u;
#pragma unroll
			for (int dy = -1; dy <= 1; dy++) {
#pragma unroll
				for (int dx = -1; dx <= 1; dx++) {
					const unsigned int k = (ly + dy) * cols + (lx + dx);
					const unsigned int sk = site[k];
					const int ex = px - static_cast<int>(static_cast<short>(sk & 0xFFFFu)),
						ey = py - (static_cast<int>(sk) >> 16);
					const int d = __mul24(ex, ex) + __mul24(ey, ey);
					if (d < best) {
						best = d;
						bk = k;
					}
				}
			}
			biome[c.Halo + static_cast<size_t>(y) * w + x] = siteBiome[bk];
		}
	}

	// ---- single histogram filter, batched form
	//
	// The vertical pass runs one lane per column and the horizontal pass one lane per row
	// (stp_stage_shf.hpp holds the accumulators and the order rules). Between the two passes the
	// column histograms are stored TRANSPOSED, [column][slot][row] (the item passes: [column][row]
	// slot groups, itemSlots), so that the horizontal pass, whose lanes are consecutive rows, reads
	// them with coalesced loads. Both passes go through
	// LDS so that every global access of a wave is a contiguous run: the vertical pass stages
	// `rows` output rows and writes them column by column; the horizontal pass loads blocks of
	// `block` columns (entering and leaving) before sliding over them, and stages its dense
	// histogram output for row-contiguous stores.

	constexpr unsigned int VPrefetch = 16u; // halo rows fetched at once by the vertical pass
	// Filter line segments. Every line longer than LongSegment outputs is split into segments of
	// at most LongSegment, so that long lines do not set the pass's duration; and a binding whose
	// lines number fewer than SegmentTarget (a few large chunks) splits them further, into up to
	// MaxSegments segments of at least MinSegment outputs, so that about SegmentTarget segments
	// keep the GPU busy.
	constexpr unsigned long long SegmentTarget = 131072ull;
	constexpr unsigned int MaxSegments = 32u, MinSegment = 64u, LongSegment = 128u;
	constexpr unsigned int HItemPrefetch = 8u; // columns fetched at once by the biome-indexed horizontal pass
	// columns fetched at once by the register horizontal pass (its registers scale with MaxCap)
	__host__ __device__ constexpr unsigned int horizontalPrefetch(const unsigned int maxCap) {
		return maxCap >= 16u ? 2u : (maxCap >= 8u ? 4u : 8u);
	}

	__host__ __device__ inline unsigned int verticalStageRows(const unsigned int vcap) {
		const unsigned int r = 12288u / (vcap * 65u);
		return r < 1u ? 1u : (r > 32u ? 32u : r);
	}

	__host__ __device__ inline unsigned int verticalLdsWords(const unsigned int vcap) {
		const unsigned int r = verticalStageRows(vcap);
		return vcap * ShfLineBlock + vcap * r * 65u + r * 65u;
	}

	__host__ __device__ inline unsigned int horizontalBlock(const unsigned int vcap, const unsigned int cap) {
		const unsigned int perColumn = 2u * (vcap + 1u) * ShfLineBlock + (1u + 2u * cap) * 65u, acc = 2u * cap * ShfLineBlock;
		const unsigned int b = acc < 12288u ? (12288u - acc) / perColumn : 0u;
		return b < 1u ? 1u : (b > 8u ? 8u : b);
	}

	// Register accumulator bound of the vertical (up to 16 slots) and horizontal (up to 8: its
	// prefetched columns also live in registers) passes; 0 = LDS accumulator.
	__host__ __device__ constexpr unsigned int registerCapacity(const unsigned int cap) {
		return cap <= 4u ? 4u : (cap <= 8u ? 8u : (cap <= 16u ? 16u : 0u));
	}
	__host__ __device__ constexpr unsigned int horizontalRegisterCapacity(const unsigned int cap) {
		return cap <= 4u ? 4u : (cap <= 8u ? 8u : 0u);
	}

	// LDS words of horizontalKERNEL<horizontalRegisterCapacity(cap)>
	__host__ __device__ inline unsigned int horizontalLdsWords(const unsigned int vcap, const unsigned int cap) {
		if (horizontalRegisterCapacity(cap) > 0u) {
			return horizontalPrefetch(horizontalRegisterCapacity(cap)) * (1u + 2u * cap) * 65u;
		}
		const unsigned int b = horizontalBlock(vcap, cap);
		return 2u * cap * ShfLineBlock + 2u * b * (vcap + 1u) * ShfLineBlock + b * (1u + 2u * cap) * 65u;
	}

	// Accumulator of one line: in registers for a capacity up to MaxCap (4, 8, 16), in LDS
	// (MaxCap = 0) for any other.
	template <unsigned int MaxCap>
	struct LineAcc {
		using Vertical = RegisterLineAcc<MaxCap>;
		using Horizontal = RegisterLineAcc<MaxCap>;
		__device__ static Vertical vertical(unsigned int*, const unsigned int cap) { return Vertical(cap); }
		__device__ static Horizontal horizontal(unsigned int*, const unsigned int cap) { return Horizontal(cap); }
	};
	template <>
	struct LineAcc<0u> {
		using Vertical = LdsVerticalAcc;
		using Horizontal = LdsHorizontalAcc;
		__device__ static Vertical vertical(unsigned int* const lds, const unsigned int cap) { return Vertical(lds, cap); }
		__device__ static Horizontal horizontal(unsigned int* const lds, const unsigned int cap) {
			return Horizontal(lds, lds + cap * ShfLineBlock, cap);
		}
	};

	// ExternalBiome batches: the stage-0 tile's halo samples copied from the caller's map. Halo
	// sample (x, y) of chunk c is origin[(OriginY + y - r) * pitch + OriginX + x - r]; a sample >= limit
	// (the item path's biome count) is stored as 0, which keeps the per-biome registers of the
	// filter passes in range, and raises flag[2].
	__global__ void __launch_bounds__(BiomeTile) externalBiomeKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const STPSample_t* const origin, const unsigned int pitch, const unsigned int radius, const unsigned int limit,
		unsigned short* const biome, unsigned int* const flag) {
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int w = c.Side + 2u * radius, tilesX = (w + BiomeSide - 1u) / BiomeSide;
		const unsigned int x = (t.y % tilesX) * BiomeSide + threadIdx.x % BiomeSide,
			y0 = (t.y / tilesX) * BiomeSide + threadIdx.x / BiomeSide;
		bool bad = false;
		if (x < w) {
			const long long col = static_cast<long long>(c.OriginX) + x - radius;
			for (unsigned int j = 0u; j < BiomeSide; j += BiomeRows) {
				const unsigned int y = y0 + j;
				if (y >= w) {
					break;
				}
				const long long row = static_cast<long long>(c.OriginY) + y - radius;
				unsigned int v = origin[row * static_cast<long long>(pitch) + col];
				if (v >= limit) {
					v = 0u;
					bad = true;
				}
				biome[c.Halo + static_cast<size_t>(y) * w + x] = static_cast<unsigned short>(v);
			}
		}
		if (bad) {
			atomicOr(flag + 2, 1u);
		}
	}

	// vcount[col * side + y], vbin[(col * vcap + k) * side + y] (per chunk at Line, Line * vcap)
	template <unsigned int MaxCap>
	__global__ void __launch_bounds__(ShfLineBlock) verticalKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const unsigned int radius, const unsigned int vcap, const unsigned short* const biome, unsigned int* const vbin,
		unsigned char* const vcount, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		const unsigned int R = verticalStageRows(vcap);
		unsigned int* const stage = lds + vcap * ShfLineBlock; // [k][row][65]
		unsigned int* const stageN = stage + vcap * R * 65u;   // [row][65]
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int lane = threadIdx.x, side = c.Side, ncols = side + 2u * radius;
		const unsigned int col = t.y * ShfLineBlock + lane;
		const bool active = col < ncols;
		// halo map (side + 2r)^2 starting at (-r, -r): output row y covers halo rows y .. y + 2r
		auto acc = LineAcc<MaxCap>::vertical(lds + lane, vcap);
		// halo rows are fetched VPrefetch at a time (clamped indices, so every load is in bounds
		// and none waits for another); out-of-range rows are dropped after the load
		const size_t lastRow = side + 2u * radius - 1u;
		const unsigned short* const colc = biome + c.Halo + min(col, ncols - 1u);
		const auto fetch = [&](unsigned int (&v)[VPrefetch], const long long row0) {
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				const long long r = row0 + i;
				v[i] = colc[static_cast<size_t>(r < 0 ? 0 : (r > static_cast<long long>(lastRow) ? lastRow : r)) * ncols];
			}
		};
		for (unsigned int j0 = 0u; j0 <= 2u * radius; j0 += VPrefetch) {
			unsigned int v[VPrefetch];
			fetch(v, j0);
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				if (active && j0 + i <= 2u * radius) {
					acc.increment(v[i], 1u);
				}
			}
		}
		unsigned int* const outBin = vbin + c.Line * vcap;
		unsigned char* const outN = vcount + c.Line;
		for (unsigned int y0 = 0u; y0 < side; y0 += R) {
			const unsigned int rows = min(R, side - y0);
			for (unsigned int b0 = 0u; b0 < rows; b0 += VPrefetch) {
				unsigned int enter[VPrefetch], leave[VPrefetch];
				fetch(enter, static_cast<long long>(y0 + b0) + 2 * static_cast<long long>(radius));
				fetch(leave, static_cast<long long>(y0 + b0) - 1);
#pragma unroll
				for (unsigned int i = 0u; i < VPrefetch; i++) {
					const unsigned int yy = b0 + i, y = y0 + yy;
					if (active && yy < rows) {
						if (y > 0u) {
							acc.increment(enter[i], 1u);
							acc.decrement(leave[i], 1u);
						}
						stageN[yy * 65u + lane] = acc.size();
						acc.forEach([&](const unsigned int k, const unsigned int item, const unsigned int count) {
							stage[(k * R + yy) * 65u + lane] = item | (count << 16);
						});
					}
				}
			}
			__syncthreads();
			// consecutive lanes take consecutive rows of one column
			for (unsigned int i = lane; i < ShfLineBlock * rows; i += ShfLineBlock) {
				const unsigned int cc = i / rows, yy = i - cc * rows;
				const unsigned int cg = t.y * ShfLineBlock + cc;
				if (cg < ncols) {
					const unsigned int n = stageN[yy * 65u + cc];
					const size_t y = static_cast<size_t>(y0) + yy;
					outN[static_cast<size_t>(cg) * side + y] = static_cast<unsigned char>(n);
					for (unsigned int k = 0u; k < n; k++) {
						outBin[(static_cast<size_t>(cg) * vcap + k) * side + y] = stage[(k * R + yy) * 65u + cc];
					}
				}
			}
			__syncthreads();
		}
		if (acc.overflowed()) {
			atomicOr(overflow, 1u);
		}
	}

	// Dense histogram of every pixel: bin_count[p], bin_item/bin_weight[p * cap + slot].
	// Virtual step s runs from -2r to side - 1: column s + 2r enters, column s - 1 leaves (s >= 1),
	// pixel s is emitted (s >= 0) -- the same sequence as shfHorizontalLine.
	template <unsigned int MaxCap>
	__global__ void __launch_bounds__(ShfLineBlock) horizontalKERNEL(const uint2* const tiles,
		const Chunk* const chunks, const unsigned int radius, const unsigned int vcap, const unsigned int cap,
		const unsigned int* const vbin, const unsigned char* const vcount, unsigned char* const bin_count,
		unsigned short* const bin_item, float* const bin_weight, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		// LDS path: [accumulator][entering block][leaving block][output]; register path: [output]
		constexpr unsigned int HPrefetch = horizontalPrefetch(MaxCap);
		const unsigned int B = MaxCap > 0u ? HPrefetch : horizontalBlock(vcap, cap);
		unsigned int* const enterN = lds + 2u * cap * ShfLineBlock; // [col][lane]
		unsigned int* const enterBin = enterN + B * ShfLineBlock;   // [col][k][lane]
		unsigned int* const leaveN = enterBin + B * vcap * ShfLineBlock;
		unsigned int* const leaveBin = leaveN + B * ShfLineBlock;
		unsigned int* const outN = MaxCap > 0u ? lds : leaveBin + B * vcap * ShfLineBlock; // [px][65]
		unsigned int* const outItem = outN + B * 65u;                                        // [px][k][65]
		float* const outW = reinterpret_cast<float*>(outItem + B * cap * 65u);
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int lane = threadIdx.x, side = c.Side, ncols = side + 2u * radius;
		const unsigned int y0 = t.y * ShfLineBlock, y = y0 + lane;
		const bool active = y < side;
		const unsigned int* const inBin = vbin + c.Line * vcap;
		const unsigned char* const inN = vcount + c.Line;
		auto acc = LineAcc<MaxCap>::horizontal(lds + lane, cap);
		const int first = -2 * static_cast<int>(radius);
		// stage one pixel's histogram for the row-contiguous stores
		const auto emit = [&](const unsigned int j) {
			unsigned int sum = 0u;
			acc.forEach([&](unsigned int, unsigned int, const unsigned int count) { sum += count; });
			const float norm = 1.0f / static_cast<float>(sum);
			outN[j * 65u + lane] = acc.size();
			acc.forEach([&](const unsigned int k, const unsigned int item, const unsigned int count) {
				outItem[(j * cap + k) * 65u + lane] = item;
				outW[(j * cap + k) * 65u + lane] = static_cast<float>(count) * norm;
			});
		};
		// row-contiguous stores of the staged pixels [xa, sEnd) of the block starting at s0
		const auto flush = [&](const int s0, const int sEnd) {
			const int xa = max(s0, 0);
			if (sEnd > xa) {
				const unsigned int npx = static_cast<unsigned int>(sEnd - xa), jb = static_cast<unsigned int>(xa - s0);
				const unsigned int rows = min(ShfLineBlock, side - y0);
				for (unsigned int i = lane; i < rows * npx; i += ShfLineBlock) {
					const unsigned int r = i / npx, px = i - r * npx;
					const size_t p = static_cast<size_t>(y0 + r) * side + static_cast<size_t>(xa) + px;
					bin_count[c.Pixel + p] = static_cast<unsigned char>(outN[(jb + px) * 65u + r]);
				}
				for (unsigned int i = lane; i < rows * npx * cap; i += ShfLineBlock) {
					const unsigned int r = i / (npx * cap), q = i - r * npx * cap, px = q / cap, k = q - px * cap;
					const size_t p = static_cast<size_t>(y0 + r) * side + static_cast<size_t>(xa) + px;
					const unsigned int at = ((jb + px) * cap + k) * 65u + r;
					bin_item[(c.Pixel + p) * cap + k] = static_cast<unsigned short>(outItem[at]);
					bin_weight[(c.Pixel + p) * cap + k] = outW[at];
				}
			}
		};
		if constexpr (MaxCap > 0u) {
			// register path: HPrefetch columns (counts and every slot) are fetched with clamped,
			// in-bounds indices before the accumulator slides over them
			const unsigned int yc = min(y, side - 1u);
			const auto colIndex = [&](const int col) {
				return static_cast<size_t>(col < 0 ? 0 : (col >= static_cast<int>(ncols) ? static_cast<int>(ncols) - 1 : col));
			};
			for (int s0 = first; s0 < static_cast<int>(side); s0 += static_cast<int>(HPrefetch)) {
				unsigned int eN[HPrefetch], lN[HPrefetch], eB[HPrefetch][MaxCap], lB[HPrefetch][MaxCap];
#pragma unroll
				for (unsigned int j = 0u; j < HPrefetch; j++) {
					const size_t ce = colIndex(s0 + 2 * static_cast<int>(radius) + static_cast<int>(j)),
						cl = colIndex(s0 - 1 + static_cast<int>(j));
					eN[j] = inN[ce * side + yc];
					lN[j] = inN[cl * side + yc];
#pragma unroll
					for (unsigned int k = 0u; k < MaxCap; k++) {
						eB[j][k] = k < vcap ? inBin[(ce * vcap + k) * side + yc] : 0u;
						lB[j][k] = k < vcap ? inBin[(cl * vcap + k) * side + yc] : 0u;
					}
				}
				const int sEnd = min(s0 + static_cast<int>(HPrefetch), static_cast<int>(side));
#pragma unroll
				for (unsigned int j = 0u; j < HPrefetch; j++) {
					const int s = s0 + static_cast<int>(j);
					if (active && s < sEnd) {
#pragma unroll
						for (unsigned int k = 0u; k < MaxCap; k++) {
							if (k < eN[j]) {
								acc.increment(eB[j][k] & 0xFFFFu, eB[j][k] >> 16);
							}
						}
						if (s >= 1) {
#pragma unroll
							for (unsigned int k = 0u; k < MaxCap; k++) {
								if (k < lN[j]) {
									acc.decrement(lB[j][k] & 0xFFFFu, lB[j][k] >> 16);
								}
							}
						}
						if (s >= 0) {
							emit(j);
						}
					}
				}
				__syncthreads();
				flush(s0, sEnd);
				__syncthreads();
			}
		} else {
			// LDS path (any capacity): blocks of B columns staged in LDS
			for (int s0 = first; s0 < static_cast<int>(side); s0 += static_cast<int>(B)) {
				for (unsigned int j = 0u; j < B; j++) {
					const int ce = s0 + 2 * static_cast<int>(radius) + static_cast<int>(j), cl = s0 - 1 + static_cast<int>(j);
					if (active && ce < static_cast<int>(ncols)) {
						const size_t col = static_cast<size_t>(ce);
						enterN[j * ShfLineBlock + lane] = inN[col * side + y];
						for (unsigned int k = 0u; k < vcap; k++) {
							enterBin[(j * vcap + k) * ShfLineBlock + lane] = inBin[(col * vcap + k) * side + y];
						}
					}
					if (active && cl >= 0 && cl < static_cast<int>(ncols)) {
						const size_t col = static_cast<size_t>(cl);
						leaveN[j * ShfLineBlock + lane] = inN[col * side + y];
						for (unsigned int k = 0u; k < vcap; k++) {
							leaveBin[(j * vcap + k) * ShfLineBlock + lane] = inBin[(col * vcap + k) * side + y];
						}
					}
				}
				__syncthreads();
				const int sEnd = min(s0 + static_cast<int>(B), static_cast<int>(side));
				if (active) {
					for (int s = s0; s < sEnd; s++) {
						const unsigned int j = static_cast<unsigned int>(s - s0);
						const unsigned int ne = enterN[j * ShfLineBlock + lane];
						for (unsigned int k = 0u; k < ne; k++) {
							const unsigned int e = enterBin[(j * vcap + k) * ShfLineBlock + lane];
							acc.increment(e & 0xFFFFu, e >> 16);
						}
						if (s >= 1) {
							const unsigned int nl = leaveN[j * ShfLineBlock + lane];
							for (unsigned int k = 0u; k < nl; k++) {
								const unsigned int e = leaveBin[(j * vcap + k) * ShfLineBlock + lane];
								acc.decrement(e & 0xFFFFu, e >> 16);
							}
						}
						if (s >= 0) {
							emit(j);
						}
					}
				}
				__syncthreads();
				flush(s0, sEnd);
				__syncthreads();
			}
		}
		if (acc.overflowed()) {
			atomicOr(overflow, 1u);
		}
	}

	// ---- single histogram filter for maps of at most MaxB biomes (ids < MaxB)
	//
	// The same filter, stated without the accumulator. For one line and one biome, call an
	// occurrence a position whose value (vertical pass) or column histogram (horizontal pass)
	// holds the biome, and a run start an occurrence with no occurrence in the 2r + 1 positions
	// before it. The accumulator of the reference (STPSingleHistogramFilter.cpp:400-450) then
	// holds, for the window ending at position e, exactly the biomes with an occurrence in the
	// window, counted over it, in the order of (their latest run start <= e, their slot in that
	// run start's column histogram): a biome is appended when its run's first occurrence enters,
	// stays while the run lasts, and erasures keep the relative order of the others. So one lane
	// keeps, per biome, the window count, the last occurrence and the latest run-start key, all in
	// registers, and sorts the present biomes by key when it emits -- O(MaxB) register work per
	// step instead of searching and shifting bins.

	// ascending sorting network over keys, every index static: odd-even transposition, N rounds of
	// adjacent compare-exchanges as min / max
	template <unsigned int N>
	__device__ inline void sortKeys(unsigned int (&key)[N]) {
#pragma unroll
		for (unsigned int round = 0u; round < N; round++) {
#pragma unroll
			for (unsigned int a = round & 1u; a + 1u < N; a += 2u) {
				const unsigned int lo = min(key[a], key[a + 1u]), hi = max(key[a], key[a + 1u]);
				key[a] = lo;
				key[a + 1u] = hi;
			}
		}
	}

	constexpr unsigned int NoRun = 0xFFFFFFFFu;
	constexpr unsigned int ItemStageRows = 16u;
	static_assert(ItemStageRows == VPrefetch, "the vertical item pass stages the rows it fetches at once");
	// Column histograms of the item path: per (column, row) a group of itemSlots(vcap) 16-bit
	// slots at vbin[(Line + col * side + y) * slots], slot k = item | count << 8 of the window's
	// k-th bin (item < 8, count <= 2r + 1 <= 255) and 0 past its last bin (a bin's count is >= 1),
	// so one 8- or 16-byte load gives a lane a column's whole histogram and its bin count.
	__host__ __device__ constexpr unsigned int itemSlots(const unsigned int vcap) { return vcap <= 4u ? 4u : 8u; }
	// The vertical item pass stages a row's slots two to an LDS word (slots 2p, 2p + 1 in the low
	// and high halves, as in the group): itemSlots(vcap) / 2 words per row and line.
	__host__ __device__ constexpr unsigned int verticalItemWords(const unsigned int vcap) {
		return itemSlots(vcap) / 2u * ItemStageRows * 65u;
	}
	// slot k of a group's words
	__device__ inline unsigned int itemSlot(const unsigned int* const w, const unsigned int k) {
		return (w[k / 2u] >> (16u * (k & 1u))) & 0xFFFFu;
	}
	// the first L slots of the group at g (8- or 16-byte aligned: the group's own size)
	template <unsigned int L>
	__device__ inline void loadItemGroup(const unsigned int* const g, unsigned int (&w)[L / 2u]) {
		static_assert(L == 4u || L == 8u, "slot groups of 4 or 8");
		if constexpr (L == 4u) {
			const uint2 v = *reinterpret_cast<const uint2*>(g);
			w[0] = v.x;
			w[1] = v.y;
		} else {
			const uint4 v = *reinterpret_cast<const uint4*>(g);
			w[0] = v.x;
			w[1] = v.y;
			w[2] = v.z;
			w[3] = v.w;
		}
	}

	// Per-biome window state of one line.
	template <unsigned int MaxB>
	struct ItemLine {
		unsigned int Count[MaxB], Key[MaxB];
		int Last[MaxB];
		__device__ ItemLine() {
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				this->Count[i] = 0u;
				this->Key[i] = NoRun;
				this->Last[i] = -0x40000000;
			}
		}
		// the state carried into segment `seg` of `line` (left there by segmentCarryKERNEL)
		__device__ inline void resume(const unsigned int* const state, const Chunk& c, const unsigned int lines,
			const unsigned int seg, const unsigned int line);
		// biome `it` occurs at position `pos` with count `c`, as slot `slot` of that position
		// Straight-line selects over the biomes: written as branches, the compiler turns the updates
		// into a branch tree on `it`, which a wave whose lanes hold different biomes walks in full.
		__device__ inline void enter(const unsigned int it, const int pos, const unsigned int c, const unsigned int slot,
			const unsigned int radius) {
			const int gap = pos - 2 * static_cast<int>(radius) - 1;
			const unsigned int key = static_cast<unsigned int>(pos) * 64u + slot;
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				const bool hit = i == it;
				this->Key[i] = hit && this->Last[i] < gap ? key : this->Key[i];
				this->Last[i] = hit ? pos : this->Last[i];
				this->Count[i] += hit ? c : 0u;
			}
		}
		__device__ inline void leave(const unsigned int it, const unsigned int c) {
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				this->Count[i] -= i == it ? c : 0u;
			}
		}
		// biome `it` occurs again at `pos` with its count and run unchanged
		__device__ inline void touch(const unsigned int it, const int pos) {
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				this->Last[i] = i == it ? pos : this->Last[i];
			}
		}
		// Present biomes in order: item[k], count[k] for k < returned n. A run key is below 2^21
		// (position * 64 + slot, positions below 2^15), so the biome id rides in its low 3 bits and
		// the keys sort alone, with min / max (distinct biomes never share a key).
		// CountBits (8 or 16) bounds the window counts: they are packed 32 / CountBits to a word and
		// each sorted slot's count is a bit-field extract at its biome id.
		template <unsigned int CountBits>
		__device__ inline unsigned int order(unsigned int (&item)[MaxB], unsigned int (&count)[MaxB]) const {
			static_assert(CountBits == 8u || CountBits == 16u, "counts packed as bytes or halves");
			constexpr unsigned int PerWord = 32u / CountBits, Words = (MaxB + PerWord - 1u) / PerWord;
			unsigned int key[MaxB], packed[Words] = {}, n = 0u;
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				const bool present = this->Count[i] > 0u;
				key[i] = present ? (this->Key[i] << 3) | i : NoRun;
				n += present ? 1u : 0u;
				packed[i / PerWord] |= this->Count[i] << (CountBits * (i % PerWord));
			}
			sortKeys(key);
#pragma unroll
			for (unsigned int k = 0u; k < MaxB; k++) {
				const unsigned int it = key[k] & 7u;
				unsigned int w = packed[0];
#pragma unroll
				for (unsigned int q = 1u; q < Words; q++) {
					w = it / PerWord == q ? packed[q] : w;
				}
				item[k] = it;
				count[k] = __builtin_amdgcn_ubfe(w, CountBits * (it % PerWord), CountBits);
			}
			return n;
		}
	};

	// ---- line segments
	//
	// A filter line of a large chunk can be split into segments of outputs that run in parallel.
	// The state an unsplit line holds before position e is, per biome, its last occurrence and
	// its latest run-start key (the window counts restart from the segment's own first window).
	// A block of positions changes that state as its run summary says: the key of the biome's
	// first occurrence in the block, the key of its latest run start after that first occurrence
	// (the gap before it lies inside the block) and its last occurrence. So a summary pass over the
	// blocks, a fold of each line's summaries in order (segmentCarryKERNEL) and a resumed filter
	// per segment give the unsplit line's result exactly.

	// Outputs [x, y) of segment `seg` of chunk c's lines, segments of at most SegLen outputs. A chunk
	// with a free-slip halo (Crop > 0) splits its regions [0, Crop), [Crop, Side - Crop) and
	// [Side - Crop, Side) apart, and each region's first segment starts from an empty state, as a
	// line of its own (segmentRestarts): its centre's lines begin where the lines of the chunk
	// generated alone begin, so the filter's bin order, and the blend, are the same there.
	__device__ inline uint3 segmentRegions(const Chunk& c) {
		const unsigned int len = c.SegLen;
		return make_uint3(len, (c.Crop + len - 1u) / len, (c.Side - 2u * c.Crop + len - 1u) / len);
	}
	__device__ inline uint2 segmentRange(const Chunk& c, const unsigned int seg) {
		const unsigned int len = c.SegLen;
		if (c.Crop == 0u) {
			const unsigned int a = min(seg * len, c.Side);
			return make_uint2(a, min(a + len, c.Side));
		}
		const uint3 r = segmentRegions(c);
		unsigned int base, end, k;
		if (seg < r.y) {
			base = 0u;
			end = c.Crop;
			k = seg;
		} else if (seg < r.y + r.z) {
			base = c.Crop;
			end = c.Side - c.Crop;
			k = seg - r.y;
		} else {
			base = c.Side - c.Crop;
			end = c.Side;
			k = seg - r.y - r.z;
		}
		const unsigned int a = min(base + k * len, end);
		return make_uint2(a, min(a + len, end));
	}
	// a segment whose carried state is empty: every line's first, and each region's first in a halo chunk
	__device__ inline bool segmentRestarts(const Chunk& c, const unsigned int seg) {
		if (seg == 0u) {
			return true;
		}
		if (c.Crop == 0u) {
			return false;
		}
		const uint3 r = segmentRegions(c);
		return seg == r.y || seg == r.y + r.z;
	}

	// word of (segment, field, biome, line) in a chunk's part of the segment-state arena; fields
	// 0, 1, 2 hold the summary (first key, inner key, last occurrence), then 0, 1 the carried state
	// (last occurrence, key)
	template <unsigned int MaxB>
	__device__ inline size_t segmentWord(const Chunk& c, const unsigned int lines, const unsigned int seg,
		const unsigned int field, const unsigned int b, const unsigned int line) {
		return static_cast<size_t>(c.SegLine) * 3u * MaxB + ((static_cast<size_t>(seg) * 3u + field) * MaxB + b) * lines
			+ line;
	}

	template <unsigned int MaxB>
	struct RunSummary {
		unsigned int First[MaxB], Inner[MaxB];
		int Last[MaxB];
		__device__ RunSummary() {
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				this->First[i] = NoRun;
				this->Inner[i] = NoRun;
				this->Last[i] = -0x40000000;
			}
		}
		// straight-line selects, as ItemLine::enter
		__device__ inline void occur(const unsigned int it, const int pos, const unsigned int slot, const unsigned int radius) {
			const int gap = pos - 2 * static_cast<int>(radius) - 1;
			const unsigned int key = static_cast<unsigned int>(pos) * 64u + slot;
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				const bool hit = i == it, fresh = this->First[i] == NoRun;
				this->Inner[i] = hit && !fresh && this->Last[i] < gap ? key : this->Inner[i];
				this->First[i] = hit && fresh ? key : this->First[i];
				this->Last[i] = hit ? pos : this->Last[i];
			}
		}
		__device__ inline void store(unsigned int* const state, const Chunk& c, const unsigned int lines,
			const unsigned int seg, const unsigned int line) const {
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				state[segmentWord<MaxB>(c, lines, seg, 0u, i, line)] = this->First[i];
				state[segmentWord<MaxB>(c, lines, seg, 1u, i, line)] = this->Inner[i];
				state[segmentWord<MaxB>(c, lines, seg, 2u, i, line)] = static_cast<unsigned int>(this->Last[i]);
			}
		}
	};

	template <unsigned int MaxB>
	__device__ inline void ItemLine<MaxB>::resume(const unsigned int* const state, const Chunk& c,
		const unsigned int lines, const unsigned int seg, const unsigned int line) {
#pragma unroll
		for (unsigned int i = 0u; i < MaxB; i++) {
			this->Last[i] = static_cast<int>(state[segmentWord<MaxB>(c, lines, seg, 0u, i, line)]);
			this->Key[i] = state[segmentWord<MaxB>(c, lines, seg, 1u, i, line)];
		}
	}

	// Folds every line's segment summaries in order and leaves in each segment's slot the state
	// carried into it. `extra` = 2r for the vertical pass (Side + 2r lines), 0 for the horizontal.
	// The first segment's line blocks do the work; the other blocks of the table return at once.
	template <unsigned int MaxB>
	__global__ void __launch_bounds__(ShfLineBlock) segmentCarryKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const unsigned int radius, const unsigned int extra, unsigned int* const state) {
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int lines = c.Side + extra, nb = (lines + ShfLineBlock - 1u) / ShfLineBlock;
		const unsigned int line = t.y * ShfLineBlock + threadIdx.x;
		if (t.y >= nb || c.Segs < 2u || line >= lines) {
			return;
		}
		int last[MaxB];
		unsigned int key[MaxB];
#pragma unroll
		for (unsigned int i = 0u; i < MaxB; i++) {
			last[i] = -0x40000000;
			key[i] = NoRun;
		}
		for (unsigned int g = 0u; g < c.Segs; g++) {
			if (segmentRestarts(c, g)) {
#pragma unroll
				for (unsigned int i = 0u; i < MaxB; i++) {
					last[i] = -0x40000000;
					key[i] = NoRun;
				}
			}
#pragma unroll
			for (unsigned int i = 0u; i < MaxB; i++) {
				const size_t w0 = segmentWord<MaxB>(c, lines, g, 0u, i, line), w1 = segmentWord<MaxB>(c, lines, g, 1u, i, line);
				unsigned int first = NoRun, inner = NoRun, slast = 0u;
				if (g + 1u < c.Segs) {
					first = state[w0];
					inner = state[w1];
					slast = state[segmentWord<MaxB>(c, lines, g, 2u, i, line)];
				}
				state[w0] = static_cast<unsigned int>(last[i]);
				state[w1] = key[i];
				if (first != NoRun) {
					if (inner != NoRun) {
						key[i] = inner;
					} else if (last[i] < static_cast<int>(first >> 6) - 2 * static_cast<int>(radius) - 1) {
						key[i] = first;
					}
					last[i] = static_cast<int>(slast);
				}
			}
		}
	}

	// run summaries of the vertical pass's segments (all but each line's last)
	template <unsigned int MaxB>
	__global__ void __launch_bounds__(ShfLineBlock) verticalSummaryKERNEL(const uint2* const tiles,
		const Chunk* const chunks, const unsigned int radius, const unsigned short* const biome,
		unsigned int* const state) {
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int side = c.Side, ncols = side + 2u * radius, nb = (ncols + ShfLineBlock - 1u) / ShfLineBlock;
		const unsigned int seg = t.y / nb, col = (t.y % nb) * ShfLineBlock + threadIdx.x;
		if (seg + 1u >= c.Segs) {
			return;
		}
		const uint2 yr = segmentRange(c, seg);
		const bool active = col < ncols;
		const unsigned short* const colc = biome + c.Halo + min(col, ncols - 1u);
		RunSummary<MaxB> sum;
		for (unsigned int y0 = yr.x; y0 < yr.y; y0 += VPrefetch) {
			unsigned int v[VPrefetch];
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				v[i] = colc[static_cast<size_t>(min(y0 + i, yr.y - 1u)) * ncols];
			}
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				if (active && y0 + i < yr.y) {
					sum.occur(v[i], static_cast<int>(y0 + i), 0u, radius);
				}
			}
		}
		if (active) {
			sum.store(state, c, ncols, seg, col);
		}
	}

	// run summaries of the horizontal pass's segments (all but each line's last)
	// L: slots a lane loads per column group, min(MaxB, itemSlots(vcap)) (a window holds at most
	// MaxB biomes, so the slots past MaxB are 0)
	template <unsigned int MaxB, unsigned int L>
	__global__ void __launch_bounds__(ShfLineBlock) horizontalSummaryKERNEL(const uint2* const tiles,
		const Chunk* const chunks, const unsigned int radius, const unsigned int vcap, const unsigned int* const vbin,
		unsigned int* const state) {
		constexpr unsigned int P = HItemPrefetch;
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int side = c.Side, nb = (side + ShfLineBlock - 1u) / ShfLineBlock;
		const unsigned int seg = t.y / nb, y = (t.y % nb) * ShfLineBlock + threadIdx.x;
		if (seg + 1u >= c.Segs) {
			return;
		}
		const uint2 xr = segmentRange(c, seg);
		const bool active = y < side;
		const unsigned int yc = min(y, side - 1u);
		const unsigned int gw = itemSlots(vcap) / 2u; // words per group
		// the lane's row of the chunk's groups; a column steps by side * gw words (wave-uniform)
		const unsigned int* const inRow = vbin + (c.Line + yc) * gw;
		const size_t colStride = static_cast<size_t>(side) * gw;
		RunSummary<MaxB> sum;
		for (unsigned int x0 = xr.x; x0 < xr.y; x0 += P) {
			unsigned int b[P][L / 2u];
#pragma unroll
			for (unsigned int j = 0u; j < P; j++) {
				const size_t col = min(x0 + j, xr.y - 1u);
				loadItemGroup<L>(inRow + col * colStride, b[j]);
			}
#pragma unroll
			for (unsigned int j = 0u; j < P; j++) {
				if (active && x0 + j < xr.y) {
#pragma unroll
					for (unsigned int k = 0u; k < L; k++) {
						const unsigned int v = itemSlot(b[j], k);
						if (v != 0u) {
							sum.occur(v & 0xFFu, static_cast<int>(x0 + j), k, radius);
						}
					}
				}
			}
		}
		if (active) {
			sum.store(state, c, side, seg, y);
		}
	}

	// vertical pass: the slot group of (col, y) at vbin + (Line + col * side + y) * itemSlots(vcap) / 2
	// words (itemSlots)
	template <unsigned int MaxB>
	__global__ void __launch_bounds__(ShfLineBlock) verticalItemKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const unsigned int radius, const unsigned int vcap, const unsigned short* const biome, unsigned int* const vbin,
		const unsigned int* const state, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		constexpr unsigned int R = ItemStageRows;
		// [slot pair][row][65] (verticalItemWords)
		unsigned int* const stage = lds;
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int lane = threadIdx.x, side = c.Side, ncols = side + 2u * radius;
		const unsigned int nb = (ncols + ShfLineBlock - 1u) / ShfLineBlock, seg = t.y / nb;
		const unsigned int col = (t.y % nb) * ShfLineBlock + lane;
		const bool active = col < ncols;
		// this block's segment of the outputs; an unsplit line is the one segment [0, side)
		const uint2 yr = segmentRange(c, seg);
		const size_t lastRow = side + 2u * radius - 1u;
		const unsigned short* const colc = biome + c.Halo + min(col, ncols - 1u);
		const auto fetch = [&](unsigned int (&v)[VPrefetch], const long long row0) {
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				const long long r = row0 + i;
				v[i] = colc[static_cast<size_t>(r < 0 ? 0 : (r > static_cast<long long>(lastRow) ? lastRow : r)) * ncols];
			}
		};
		ItemLine<MaxB> line;
		if (seg > 0u && active) {
			line.resume(state, c, ncols, seg, col);
		}
		bool over = false;
		for (unsigned int j0 = 0u; j0 <= 2u * radius; j0 += VPrefetch) {
			unsigned int v[VPrefetch];
			fetch(v, static_cast<long long>(yr.x) + j0);
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				if (active && j0 + i <= 2u * radius) {
					line.enter(v[i], static_cast<int>(yr.x + j0 + i), 1u, 0u, radius);
				}
			}
		}
		const unsigned int gw = itemSlots(vcap) / 2u; // words per group
		unsigned int* const out = vbin + c.Line * gw;
		// software pipelined: the rows of the next block are fetched before this block is processed
		unsigned int enter[VPrefetch], leave[VPrefetch];
		fetch(enter, static_cast<long long>(yr.x) + 2 * static_cast<long long>(radius));
		fetch(leave, static_cast<long long>(yr.x) - 1);
		for (unsigned int y0 = yr.x; y0 < yr.y; y0 += R) {
			const unsigned int rows = min(R, yr.y - y0);
			unsigned int nextEnter[VPrefetch] = {}, nextLeave[VPrefetch] = {};
			if (y0 + R < yr.y) {
				fetch(nextEnter, static_cast<long long>(y0 + R) + 2 * static_cast<long long>(radius));
				fetch(nextLeave, static_cast<long long>(y0 + R) - 1);
			}
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				const unsigned int y = y0 + i;
				if (active && i < rows) {
					if (y > yr.x) {
						line.enter(enter[i], static_cast<int>(y + 2u * radius), 1u, 0u, radius);
						line.leave(leave[i], 1u);
					}
					unsigned int item[MaxB], count[MaxB];
					// a column window's counts are at most 2r + 1 <= 255
					const unsigned int n = line.template order<8u>(item, count);
					over = over || n > vcap;
					const unsigned int kept = min(n, vcap);
#pragma unroll
					for (unsigned int p = 0u; p < (MaxB + 1u) / 2u; p++) {
						if (p < gw) {
							const unsigned int k0 = 2u * p, k1 = k0 + 1u;
							const unsigned int lo = k0 < kept ? (item[k0] | (count[k0] << 8)) : 0u;
							const unsigned int hi = k1 < kept ? (item[k1] | (count[k1] << 8)) : 0u;
							stage[(p * R + i) * 65u + lane] = lo | (hi << 16);
						}
					}
				}
			}
			__syncthreads();
			// lane -> (column, row): consecutive lanes store the consecutive groups of one column
			// (16 rows of 8 or 16 bytes), slots 0 .. 7 of a group past MaxB stay 0
			for (unsigned int i = lane; i < ShfLineBlock * R; i += ShfLineBlock) {
				const unsigned int cc = i / R, yy = i % R;
				const unsigned int cg = (t.y % nb) * ShfLineBlock + cc;
				if (cg < ncols && yy < rows) {
					unsigned int w[4] = {};
#pragma unroll
					for (unsigned int p = 0u; p < (MaxB + 1u) / 2u; p++) {
						if (p < gw) {
							w[p] = stage[(p * R + yy) * 65u + cc];
						}
					}
					unsigned int* const g = out + (static_cast<size_t>(cg) * side + y0 + yy) * gw;
					if (gw == 2u) {
						*reinterpret_cast<uint2*>(g) = make_uint2(w[0], w[1]);
					} else {
						*reinterpret_cast<uint4*>(g) = make_uint4(w[0], w[1], w[2], w[3]);
					}
				}
			}
			__syncthreads();
#pragma unroll
			for (unsigned int i = 0u; i < VPrefetch; i++) {
				enter[i] = nextEnter[i];
				leave[i] = nextLeave[i];
			}
		}
		if (over) {
			atomicOr(overflow, 1u);
		}
	}

	// LDS words per staged pixel of horizontalItemKERNEL: the head word, the 1 / sum word unless
	// Packed (whose weights the blend rebuilds), then the counts two to a word
	__host__ __device__ constexpr unsigned int horizontalCountWord(const bool packed) { return packed ? 1u : 2u; }
	__host__ __device__ constexpr unsigned int horizontalItemWords(const unsigned int cap, const bool packed) {
		return horizontalCountWord(packed) + (cap + 1u) / 2u;
	}

	// horizontal pass, dense output as horizontalKERNEL; Packed: a slot holds item | count << 3 in
	// bin_item and bin_weight is not written (the weight is count * (1 / (2r + 1)^2), the window's
	// sum being (2r + 1)^2 at every pixel)
	// L as horizontalSummaryKERNEL
	template <unsigned int MaxB, unsigned int L, bool Packed>
	__global__ void __launch_bounds__(ShfLineBlock) horizontalItemKERNEL(const uint2* const tiles,
		const Chunk* const chunks, const unsigned int radius, const unsigned int vcap, const unsigned int cap,
		const unsigned int* const vbin, unsigned char* const bin_count, unsigned short* const bin_item,
		float* const bin_weight, const unsigned int* const state, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		constexpr unsigned int P = HItemPrefetch;
		// staged pixels, [px][word][65]: word 0 = bin count | items (3 bits each) << 4, word 1 = the
		// weight factor 1 / sum (not Packed), then the counts, two 16-bit halves per word
		constexpr unsigned int C = horizontalCountWord(Packed);
		const unsigned int words = horizontalItemWords(cap, Packed);
		unsigned int* const stage = lds;
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int lane = threadIdx.x, side = c.Side, ncols = side + 2u * radius;
		const unsigned int nb = (side + ShfLineBlock - 1u) / ShfLineBlock, seg = t.y / nb;
		const unsigned int y0 = (t.y % nb) * ShfLineBlock, y = y0 + lane;
		const bool active = y < side;
		// this block's segment of the outputs; an unsplit line is the one segment [0, side)
		const uint2 xr = segmentRange(c, seg);
		const int xBegin = static_cast<int>(xr.x), xEnd = static_cast<int>(xr.y);
		const unsigned int gw = itemSlots(vcap) / 2u; // words per group
		const unsigned int yc = min(y, side - 1u);
		// the lane's row of the chunk's groups; a column steps by side * gw words (wave-uniform)
		const unsigned int* const inRow = vbin + (c.Line + yc) * gw;
		const size_t colStride = static_cast<size_t>(side) * gw;
		const auto colIndex = [&](const int col) {
			return static_cast<size_t>(col < 0 ? 0 : (col >= static_cast<int>(ncols) ? static_cast<int>(ncols) - 1 : col));
		};
		ItemLine<MaxB> line;
		if (seg > 0u && active) {
			line.resume(state, c, side, seg, y);
		}
		bool over = false;
		// the staged words of the lane's last pixel: bin count | items, 1 / sum (unpacked only), counts
		unsigned int outHead = 0u, outPacked[(MaxB + 1u) / 2u] = {};
		float outNorm = 0.0f;
		const int first = xBegin - 2 * static_cast<int>(radius);
		for (int s0 = first; s0 < xEnd; s0 += static_cast<int>(P)) {
			// column histograms of the block: entering s0 + 2r + j, leaving s0 - 1 + j
			unsigned int eB[P][L / 2u], lB[P][L / 2u];
#pragma unroll
			for (unsigned int j = 0u; j < P; j++) {
				const size_t ce = colIndex(s0 + 2 * static_cast<int>(radius) + static_cast<int>(j)),
					cl = colIndex(s0 - 1 + static_cast<int>(j));
				loadItemGroup<L>(inRow + ce * colStride, eB[j]);
				loadItemGroup<L>(inRow + cl * colStride, lB[j]);
			}
			const int sEnd = min(s0 + static_cast<int>(P), xEnd);
#pragma unroll
			for (unsigned int j = 0u; j < P; j++) {
				const int s = s0 + static_cast<int>(j);
				if (s < sEnd) {
					const int ce = s + 2 * static_cast<int>(radius);
					// An entering column group equal to the leaving one (inside a biome's area) leaves
					// every count and run key as it was (each of its items occurs in the leaving column,
					// at s - 1, no further back than the gap a new run needs): only the items' last
					// occurrences move, and the pixel's bins are the previous pixel's. Taken when every
					// lane of the wave can.
					bool same = true;
#pragma unroll
					for (unsigned int w = 0u; w < L / 2u; w++) {
						same = same && eB[j][w] == lB[j][w];
					}
					if (s >= xBegin + 1 && __all(!active || same)) {
						if (active) {
#pragma unroll
							for (unsigned int k = 0u; k < L; k++) {
								const unsigned int v = itemSlot(eB[j], k);
								if (v != 0u) {
									line.touch(v & 0xFFu, ce);
								}
							}
						}
					} else if (active) {
#pragma unroll
						for (unsigned int k = 0u; k < L; k++) {
							const unsigned int v = itemSlot(eB[j], k);
							if (v != 0u) {
								line.enter(v & 0xFFu, ce, v >> 8, k, radius);
							}
						}
						if (s >= xBegin + 1) {
#pragma unroll
							for (unsigned int k = 0u; k < L; k++) {
								const unsigned int v = itemSlot(lB[j], k);
								if (v != 0u) {
									line.leave(v & 0xFFu, v >> 8);
								}
							}
						}
						if (s >= xBegin) {
							unsigned int item[MaxB], count[MaxB];
							// a window's counts are at most (2r + 1)^2 <= 65025
							const unsigned int n = line.template order<16u>(item, count);
							over = over || n > cap;
							if (!Packed) {
								unsigned int sum = 0u;
#pragma unroll
								for (unsigned int k = 0u; k < MaxB; k++) {
									sum += k < n ? count[k] : 0u;
								}
								outNorm = 1.0f / static_cast<float>(sum);
							}
							// every slot packed: the store phase reads slot k only for k below the
							// staged bin count min(n, cap) (items 3 bits, counts 16 bits each)
							outHead = min(n, cap);
#pragma unroll
							for (unsigned int w = 0u; w < (MaxB + 1u) / 2u; w++) {
								outPacked[w] = 0u;
							}
#pragma unroll
							for (unsigned int k = 0u; k < MaxB; k++) {
								outHead |= item[k] << (4u + 3u * k);
								outPacked[k / 2u] |= count[k] << (16u * (k & 1u));
							}
						}
					}
					if (active && s >= xBegin) {
						unsigned int* const at = stage + j * words * 65u + lane;
						at[0] = outHead;
						if (!Packed) {
							at[65] = __float_as_uint(outNorm);
						}
#pragma unroll
						for (unsigned int w = 0u; w < (MaxB + 1u) / 2u; w++) {
							if (2u * w < cap) {
								at[(C + w) * 65u] = outPacked[w];
							}
						}
					}
				}
			}
			__syncthreads();
			// row-contiguous stores of the staged pixels [xa, sEnd)
			const int xa = max(s0, xBegin);
			if (sEnd > xa) {
				const unsigned int npx = static_cast<unsigned int>(sEnd - xa), jb = static_cast<unsigned int>(xa - s0);
				const unsigned int rows = min(ShfLineBlock, side - y0);
				if (cap == 4u || cap == 8u) {
					// lane -> (row, pixel): one pixel's slots in 8- and 16-byte stores; slots past the
					// bin count hold zeros (they are never read)
					constexpr unsigned int RowsPerPass = ShfLineBlock / P;
					const unsigned int px = lane % P, lr = lane / P;
					for (unsigned int r0 = 0u; r0 < rows; r0 += RowsPerPass) {
						const unsigned int r = r0 + lr;
						if (r < rows && px < npx) {
							const size_t p = static_cast<size_t>(y0 + r) * side + static_cast<size_t>(xa) + px;
							const unsigned int* const at = stage + (jb + px) * words * 65u + r;
							const unsigned int head = at[0], n = head & 15u;
							const float norm = Packed ? 0.0f : __uint_as_float(at[65]);
							bin_count[c.Pixel + p] = static_cast<unsigned char>(n);
							for (unsigned int q = 0u; q < cap; q += 4u) {
								const unsigned int lo = at[(C + q / 2u) * 65u], hi = at[(C + 1u + q / 2u) * 65u];
								const unsigned int cnt[4] = {lo & 0xFFFFu, lo >> 16, hi & 0xFFFFu, hi >> 16};
								unsigned short it[4];
								float w[4];
#pragma unroll
								for (unsigned int k = 0u; k < 4u; k++) {
									const bool used = q + k < n;
									const unsigned int item = (head >> (4u + 3u * (q + k))) & 7u;
									it[k] = used ? static_cast<unsigned short>(Packed ? item | cnt[k] << 3u : item) : 0u;
									// the same product as the reference's count * (1 / sum)
									w[k] = used ? static_cast<float>(cnt[k]) * norm : 0.0f;
								}
								*reinterpret_cast<ushort4*>(bin_item + (c.Pixel + p) * cap + q) = make_ushort4(it[0], it[1], it[2], it[3]);
								if (!Packed) {
									*reinterpret_cast<float4*>(bin_weight + (c.Pixel + p) * cap + q) = make_float4(w[0], w[1], w[2], w[3]);
								}
							}
						}
					}
				} else {
					const unsigned int width = npx * cap;
					// lane -> (pixel, slot) of one row; rows advance by the lanes that fit
					const unsigned int perRow = width, rowsPerPass = max(ShfLineBlock / perRow, 1u);
					const unsigned int lr = lane / perRow, lq = lane - lr * perRow, px = lq / cap, k = lq - px * cap;
					for (unsigned int r0 = 0u; r0 < rows; r0 += rowsPerPass) {
						const unsigned int r = r0 + lr;
						if (lr < rowsPerPass && r < rows && lq < width) {
							const size_t p = static_cast<size_t>(y0 + r) * side + static_cast<size_t>(xa) + px;
							const unsigned int* const at = stage + (jb + px) * words * 65u + r;
							const unsigned int head = at[0], n = head & 15u;
							// slots past the pixel's bin count are never read: not stored
							if (k < n) {
								const unsigned int count = (at[(C + k / 2u) * 65u] >> (16u * (k & 1u))) & 0xFFFFu;
								const unsigned int item = (head >> (4u + 3u * k)) & 7u;
								bin_item[(c.Pixel + p) * cap + k] = static_cast<unsigned short>(Packed ? item | count << 3u : item);
								if (!Packed) {
									// the same product as the reference's count * (1 / sum)
									bin_weight[(c.Pixel + p) * cap + k] = static_cast<float>(count) * __uint_as_float(at[65]);
								}
							}
							if (k == 0u) {
								bin_count[c.Pixel + p] = static_cast<unsigned char>(n);
							}
						}
					}
				}
			}
			__syncthreads();
		}
		if (over) {
			atomicOr(overflow, 1u);
		}
	}

	template <bool Packed>
	__global__ void __launch_bounds__(BlendX * BlendY) blendKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const STPAlgorithm::STPPermutation perm, const BiomeTable biomes, const unsigned int biome_count,
		const unsigned int cap, const unsigned int window, const unsigned char* const bin_count,
		const unsigned short* const bin_item, const float* const bin_weight, float* const height) {
		__shared__ Noise::STPStagedPermutation table;
		__shared__ STPBiomeProperty prop[STPHeightfieldKernel::MaxBiome];
		const unsigned int tid = threadIdx.x + threadIdx.y * BlendX;
		if (tid < biome_count) {
			prop[tid] = biomes.Property[tid];
		}
		Noise::stage(perm, table);
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int tilesX = (c.Side + BlendX - 1u) / BlendX;
		// The workgroup covers 64 x 16 pixels; wave w (threadIdx.y) takes the 16 x 16 square of columns
		// 16w .. 16w + 15, and iteration k its 8 x 8 quarter (k % 2, k / 2) (lane l: column l % 8, row
		// l / 8). A wave runs each iteration for as long as its lane with the most bins: a compact
		// block straddles the edge of a biome boundary's band of 2-bin pixels less often (a 64-pixel
		// row took 13% more bin iterations than a 16 x 4 block on the batched workload, which took
		// 1.2% more blend time than an 8 x 8 block).
		static_assert(BlendX == 64u && BlendY == 4u && BlendPixel == 4u, "the 8 x 8 wave blocks assume 64 x 16 tiles");
		const unsigned int xb = (t.y % tilesX) * BlendX + threadIdx.y * 16u + (threadIdx.x & 7u);
		const unsigned int yb = (t.y / tilesX) * (BlendY * BlendPixel) + (threadIdx.x >> 3u);
		const float half = static_cast<float>(c.Side * c.Step) / 2.0f;
		// pixel x of the chunk samples world coordinate Origin + x * Step
		const float2 offset = make_float2(static_cast<float>(c.OriginX) + half, static_cast<float>(c.OriginY) + half);
		for (unsigned int k = 0u; k < BlendPixel; k++) {
			const unsigned int x = xb + 8u * (k & 1u), y = yb + 8u * (k >> 1u);
			if (x < c.Side && y < c.Side) {
				const unsigned long long i = c.Pixel + static_cast<unsigned long long>(y) * c.Side + x;
				if (Packed) {
					// the horizontal pass's 1 / sum, the window (2r + 1)^2 being every pixel's sum
					height[i] = blendPacked(table, prop, biome_count, bin_count[i], bin_item + i * cap,
						1.0f / static_cast<float>(window), x * c.Step, y * c.Step, offset, make_float2(half, half));
				} else {
					height[i] = blendPixel(table, prop, biome_count, bin_count[i], bin_item + i * cap, bin_weight + i * cap,
						x * c.Step, y * c.Step, offset, make_float2(half, half));
				}
			}
		}
	}

	// the erosion arguments of chunk c: drops spawned in its spawn map, a map of Side
	__device__ inline ErosionArgs chunkArgs(ErosionArgs a, const Chunk& c) {
		a.MapSize = make_uint2(c.SpawnSide, c.SpawnSide);
		a.SpawnOrigin = make_uint2(c.SpawnOrigin, c.SpawnOrigin);
		a.Total = make_uint2(c.Side, c.Side);
		a.StreamId = c.StreamId;
		a.Rain.RainDropCount = c.Drops;
		return a;
	}

	__global__ void __launch_bounds__(ErosionBlock) erosionKERNEL(const uint2* const tiles, const Chunk* const chunks,
		ErosionArgs args, const int2* const brush_offset, const float* const brush_weight, float* const height,
		float* const ledger, uint2* const drop_info) {
		extern __shared__ int2 brush[];
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		for (unsigned int k = threadIdx.x; k < args.BrushSize; k += blockDim.x) {
			const int2 o = brush_offset[k];
			brush[k] = make_int2(o.y * static_cast<int>(c.Side) + o.x, __float_as_int(brush_weight[k]));
		}
		__syncthreads();
		const unsigned int drop = t.y * ErosionBlock + threadIdx.x;
		if (drop >= c.Drops) {
			return;
		}
		erodeDropAnyBrush(GlobalSurface {height + c.Pixel, c.Side}, chunkArgs(args, c), brush, drop, drop, ledger + c.DropBase,
			drop_info + c.DropBase);
	}

	// ---- LDS-tiled erosion in colour batches
	//
	// A chunk is cut into core tiles of `tile` x `tile` cells. Every drop belongs to the tile its
	// spawn cell lies in (spawn positions depend only on the seed, the chunk's stream id and the
	// drop index). One workgroup per tile copies the tile plus a halo of `halo` cells (clipped to
	// the chunk) from HBM into LDS, runs the tile's drops on that copy with LDS atomics, and
	// stores the window back. The halo is MaxLifetime + max(r, 2) + 1 cells: a drop moves at most
	// one cell per step, so no access of a drop can leave the window (the window guard would kill
	// it otherwise).
	//
	// Windows of neighbouring tiles overlap, so tiles run in four colour batches by the parity of
	// their tile coordinates (the tile-level form of the conflict-free chunk batches of
	// SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:339-404). With tile >= 2 halo
	// the windows of one colour are disjoint: a batch stores its windows with plain coalesced
	// stores, no atomics, and every later batch reads the terrain the earlier ones eroded.

	constexpr unsigned int MaxTiledTaps = 128u;
	// tiled erosion splits a binding into two stream groups from this many tiles on (about two
	// rounds of one workgroup per CU for each group's colour launches)
	constexpr unsigned int SplitErosionTiles = 2048u;
	// A graph run of a binding of at least this many pixels runs PipelineGroups chunk groups' whole
	// pipelines as concurrent chains: one group's latency-bound filter passes share the GPU with
	// another group's issue-bound blend and erosion, and each fills the others' launch tails
	// (batched workload: one chain 4.72 ms per run, two 4.37, three 4.32, four 4.31, six 4.63, eight
	// 4.60: beyond the four hardware queues of a process chains wait on each other; four is also best
	// on 64 chunks of 512^2).
	constexpr unsigned long long PipelinePixels = 8ull << 20;
	constexpr unsigned int PipelineGroups = 4u;
	// tile sides the autotuner tries, largest first (the default is the largest that fits)
	constexpr unsigned int ErosionTileCandidates[] = {128u, 112u, 96u, 80u, 72u};

	__device__ inline unsigned int dropTile(const DropState& d, const unsigned int side, const unsigned int tile) {
		const unsigned int tx = (side + tile - 1u) / tile;
		const unsigned int x = min(static_cast<unsigned int>(d.px) / tile, tx - 1u),
			y = min(static_cast<unsigned int>(d.py) / tile, tx - 1u);
		return y * tx + x;
	}


	// drop -> rank among the drops of its spawn tile. The workgroup's drops (one chunk) are
	// counted per tile in LDS first; one global atomic per tile then reserves the workgroup's
	// ranks, so a counter sees one atomic per workgroup instead of one per drop.
	constexpr unsigned int BinLds = 4096u;
	// A workgroup counts the drops of up to BinGroups consecutive drop-table entries (ErosionBlock
	// drops each) per spawn tile: per run of entries of one chunk, an LDS histogram, then one global
	// atomic per tile and run for the base of the run's ranks. Wider workgroup coverage keeps the
	// number of global atomics on a tile's counter low (a 1024^2 chunk has 1024 entries).
	constexpr unsigned int BinGroups = 16u;
	__global__ void __launch_bounds__(ErosionBlock) binCountKERNEL(const uint2* const tiles, const unsigned int entries,
		const Chunk* const chunks, const ErosionArgs args, const unsigned int tile, unsigned int* const tile_drops,
		unsigned int* const rank) {
		__shared__ unsigned int local[BinLds];
		const unsigned int e0 = blockIdx.x * BinGroups, e1 = min(e0 + BinGroups, entries);
		unsigned int myTile[BinGroups], myRank[BinGroups];
		unsigned int e = e0;
		while (e < e1) {
			// the run [e, r) of entries of one chunk (uniform over the workgroup)
			const unsigned int chunk = tiles[e].x;
			unsigned int r = e + 1u;
			while (r < e1 && tiles[r].x == chunk) {
				r++;
			}
			const Chunk c = chunks[chunk];
			const unsigned int tx = (c.Side + tile - 1u) / tile, ntile = tx * tx;
			const bool aggregate = ntile <= BinLds;
			if (aggregate) {
				for (unsigned int i = threadIdx.x; i < ntile; i += ErosionBlock) {
					local[i] = 0u;
				}
			}
			__syncthreads();
#pragma unroll
			for (unsigned int g = 0u; g < BinGroups; g++) {
				const unsigned int k = e0 + g;
				const unsigned int drop = tiles[min(k, e1 - 1u)].y * ErosionBlock + threadIdx.x;
				if (k >= e && k < r && drop < c.Drops) {
					const unsigned int my = dropTile(spawn(chunkArgs(args, c), drop), c.Side, tile);
					if (aggregate) {
						myTile[g] = my;
						myRank[g] = atomicAdd(local + my, 1u);
					} else {
						rank[c.DropBase + drop] = atomicAdd(tile_drops + c.TileBase + my, 1u);
					}
				}
			}
			if (aggregate) {
				__syncthreads();
				for (unsigned int i = threadIdx.x; i < ntile; i += ErosionBlock) {
					const unsigned int n = local[i];
					if (n != 0u) {
						local[i] = atomicAdd(tile_drops + c.TileBase + i, n);
					}
				}
				__syncthreads();
#pragma unroll
				for (unsigned int g = 0u; g < BinGroups; g++) {
					const unsigned int k = e0 + g;
					const unsigned int drop = tiles[min(k, e1 - 1u)].y * ErosionBlock + threadIdx.x;
					if (k >= e && k < r && drop < c.Drops) {
						rank[c.DropBase + drop] = local[myTile[g]] + myRank[g];
					}
				}
				// the next run clears the histogram
				__syncthreads();
			}
			e = r;
		}
	}

	// first slot of every tile of a chunk: DropBase + exclusive prefix sum of the tile counts
	constexpr unsigned int ScanBlock = 256u;
	__global__ void __launch_bounds__(ScanBlock) binScanKERNEL(const Chunk* const chunks, const unsigned int first,
		const unsigned int tile, const unsigned int* const tile_drops, unsigned int* const tile_first) {
		__shared__ unsigned int part[ScanBlock / 64u];
		const Chunk c = chunks[first + blockIdx.x];
		const unsigned int tx = (c.Side + tile - 1u) / tile, n = tx * tx;
		const unsigned int lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
		unsigned int carry = c.DropBase;
		for (unsigned int base = 0u; base < n; base += ScanBlock) {
			const unsigned int i = base + threadIdx.x;
			const unsigned int v = i < n ? tile_drops[c.TileBase + i] : 0u;
			unsigned int s = v;
			for (unsigned int o = 1u; o < 64u; o <<= 1) {
				const unsigned int u = __shfl_up(s, o, 64);
				if (lane >= o) {
					s += u;
				}
			}
			if (lane == 63u) {
				part[wave] = s;
			}
			__syncthreads();
			unsigned int before = carry, total = 0u;
			for (unsigned int w = 0u; w < ScanBlock / 64u; w++) {
				before += w < wave ? part[w] : 0u;
				total += part[w];
			}
			if (i < n) {
				tile_first[c.TileBase + i] = before + s - v;
			}
			carry += total;
			__syncthreads();
		}
	}

	// tile-sorted drop list
	__global__ void __launch_bounds__(ErosionBlock) binScatterKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const ErosionArgs args, const unsigned int tile, const unsigned int* const tile_first,
		const unsigned int* const rank, unsigned int* const bin) {
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int drop = t.y * ErosionBlock + threadIdx.x;
		if (drop >= c.Drops) {
			return;
		}
		const DropState d = spawn(chunkArgs(args, c), drop);
		bin[tile_first[c.TileBase + dropTile(d, c.Side, tile)] + rank[c.DropBase + drop]] = drop;
	}

	// Window cells in flight per thread in the tiled erosion's load and store-back loops.
	constexpr unsigned int WindowBatch = 16u;
	// (row, column) of a window cell, stepped by a fixed number of cells along the row-major window.
	struct WindowCursor {
		unsigned int X, Y;
	};
	struct WindowWalk {
		unsigned int Width, StepX, StepY;
		// width and step are the same on every lane: the quotient and remainder are kept scalar
		__device__ WindowWalk(const unsigned int width, const unsigned int step) :
			Width(width), StepX(__builtin_amdgcn_readfirstlane(step % width)),
			StepY(__builtin_amdgcn_readfirstlane(step / width)) {}
		__device__ WindowCursor start(const unsigned int i) const {
			const unsigned int y = i / this->Width;
			return WindowCursor {i - y * this->Width, y};
		}
		__device__ WindowCursor next(const WindowCursor c) const {
			const unsigned int x = c.X + this->StepX;
			const bool wrap = x >= this->Width;
			return WindowCursor {wrap ? x - this->Width : x, c.Y + this->StepY + (wrap ? 1u : 0u)};
		}
	};

	// Taps: the brush size the drops are compiled for, 0 for any (a 25-tap drop alone holds 85
	// VGPRs instead of the 115 every brush size needs)
	template <unsigned int Threads, unsigned int Taps>
	__device__ __forceinline__ void erodeTile(const uint2* const tiles, const Chunk* const chunks,
		const ErosionArgs args, const unsigned int tile, const unsigned int halo, const int2* const brush_offset,
		const float* const brush_weight, const unsigned int* const tile_drops, const unsigned int* const tile_first,
		const unsigned int* const bin, float* const height, float* const ledger, uint2* const drop_info,
		unsigned int* const status) {
		extern __shared__ int window[];
		__shared__ int2 brush[MaxTiledTaps];
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int g = c.TileBase + t.y, n = tile_drops[g];
		if (n == 0u) {
			// a tile without drops leaves its cells alone (tiles outside a neighbourhood's centre map)
			return;
		}
		const unsigned int side = c.Side, tilesX = (side + tile - 1u) / tile;
		const unsigned int x0 = (t.y % tilesX) * tile, y0 = (t.y / tilesX) * tile;
		LocalSurface s;
		s.Lds = window;
		s.X0 = x0 > halo ? x0 - halo : 0u;
		s.Y0 = y0 > halo ? y0 - halo : 0u;
		s.RW = min(x0 + tile + halo, side) - s.X0;
		s.RH = min(y0 + tile + halo, side) - s.Y0;
		float* const map = height + c.Pixel;
		const unsigned int cells = s.RW * s.RH;
		// A thread's window cells i = threadIdx.x + j Threads, walked as (row, column) with the
		// row and column steps of Threads cells: no division per cell. Each batch has WindowBatch
		// loads in flight before the first is used (a load-then-use loop waits out one HBM round
		// trip per cell; the workgroup holds its CU alone, so nothing else hides that wait).
		const WindowWalk walk(s.RW, Threads);
		const WindowCursor last = walk.start(cells - 1u);
		for (unsigned int base = threadIdx.x; base < cells; base += WindowBatch * Threads) {
			float v[WindowBatch];
			WindowCursor at = walk.start(base);
#pragma unroll
			for (unsigned int j = 0u; j < WindowBatch; j++) {
				// past the last cell, reload the last cell (in bounds, never stored)
				const WindowCursor k = base + j * Threads < cells ? at : last;
				v[j] = map[(s.Y0 + k.Y) * side + s.X0 + k.X];
				at = walk.next(at);
			}
#pragma unroll
			for (unsigned int j = 0u; j < WindowBatch; j++) {
				const unsigned int i = base + j * Threads;
				if (i < cells) {
					window[i] = toFixedPoint(v[j]);
				}
			}
		}
		for (unsigned int k = threadIdx.x; k < args.BrushSize; k += Threads) {
			const int2 o = brush_offset[k];
			brush[k] = make_int2(o.y * static_cast<int>(s.RW) + o.x, __float_as_int(brush_weight[k] * LocalSurface::WeightScale));
		}
		__syncthreads();
		const ErosionArgs a = chunkArgs(args, c);
		const unsigned int first = tile_first[g];
		for (unsigned int k = threadIdx.x; k < n; k += Threads) {
			// results in tile order (consecutive lanes write consecutive words), within the chunk's range
			if constexpr (Taps == 0u) {
				erodeDropAnyBrush(s, a, brush, bin[first + k], first + k, ledger, drop_info);
			} else {
				erodeDrop<LocalSurface, Taps>(s, a, brush, bin[first + k], first + k, ledger, drop_info);
			}
		}
		__syncthreads();
		// No other workgroup of this colour batch touches the window, so the heights in HBM are
		// still the ones the window was made from: a cell whose fixed-point value is unchanged
		// keeps its exact FP32 height. A value q >= WindowExact is stored without that test: a
		// height h >= 2^-5 has a float spacing of at least 2^-28, so h * 2^28 is an integer of at
		// most 24 significant bits and fromFixedPoint(toFixedPoint(h)) == h bit for bit. Only the
		// cells below read the map again.
		bool outside = false;
		for (unsigned int base = threadIdx.x; base < cells; base += WindowBatch * Threads) {
			int q[WindowBatch];
			float old[WindowBatch];
			unsigned int at[WindowBatch];
			WindowCursor k = walk.start(base);
#pragma unroll
			for (unsigned int j = 0u; j < WindowBatch; j++) {
				const bool in = base + j * Threads < cells;
				const WindowCursor w = in ? k : last;
				at[j] = (s.Y0 + w.Y) * side + s.X0 + w.X;
				q[j] = window[in ? base + j * Threads : cells - 1u];
				k = walk.next(k);
			}
#pragma unroll
			for (unsigned int j = 0u; j < WindowBatch; j++) {
				old[j] = q[j] < WindowExact ? map[at[j]] : 0.0f;
			}
#pragma unroll
			for (unsigned int j = 0u; j < WindowBatch; j++) {
				if (base + j * Threads < cells) {
					if (q[j] >= WindowExact || q[j] != toFixedPoint(old[j])) {
						map[at[j]] = fromFixedPoint(q[j]);
					}
					outside |= q[j] < WindowLowest || q[j] > WindowHighest;
				}
			}
		}
		if (outside) {
			*status = 1u;
		}
	}

#define STP_ERODE_TILE_PARAMS const uint2* const tiles, const Chunk* const chunks, const ErosionArgs args, \
	const unsigned int tile, const unsigned int halo, const int2* const brush_offset, const float* const brush_weight, \
	const unsigned int* const tile_drops, const unsigned int* const tile_first, const unsigned int* const bin, \
	float* const height, float* const ledger, uint2* const drop_info, unsigned int* const status
#define STP_ERODE_TILE_ARGS tiles, chunks, args, tile, halo, brush_offset, brush_weight, tile_drops, tile_first, bin, \
	height, ledger, drop_info, status

	template <unsigned int Threads, unsigned int Taps>
	__global__ void __launch_bounds__(Threads) erodeTileKERNEL(STP_ERODE_TILE_PARAMS) {
		erodeTile<Threads, Taps>(STP_ERODE_TILE_ARGS);
	}

	// 768 threads: two workgroups of a tile-72 window (78 KB each) share a CU only at six waves per
	// SIMD, which caps a wave at 80 VGPRs (the 25-tap drop fits them)
	template <unsigned int Taps>
	__global__ void __launch_bounds__(768) __attribute__((amdgpu_waves_per_eu(6))) erodeTile768KERNEL(
		STP_ERODE_TILE_PARAMS) {
		erodeTile<768u, Taps>(STP_ERODE_TILE_ARGS);
	}

#undef STP_ERODE_TILE_ARGS
#undef STP_ERODE_TILE_PARAMS

	// chunk blockIdx.y's outputs -> out[chunk] (heights), out[chunks + chunk] (UINT16): its map
	// without the Crop pixels of free-slip halo on every side, (Side - 2 Crop)^2 row-major
	__global__ void __launch_bounds__(256) copyOutKERNEL(const Chunk* const chunks, const unsigned int count,
		const float* const height, const unsigned short* const fixed, void* const* const out) {
		const Chunk c = chunks[blockIdx.y];
		float* const h = static_cast<float*>(out[blockIdx.y]);
		unsigned short* const f = static_cast<unsigned short*>(out[count + blockIdx.y]);
		const unsigned int side = c.Side - 2u * c.Crop;
		const unsigned long long n = static_cast<unsigned long long>(side) * side;
		for (unsigned long long i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
			const unsigned long long y = i / side, x = i - y * side;
			const unsigned long long from = c.Pixel + (y + c.Crop) * c.Side + x + c.Crop;
			h[i] = height[from];
			f[i] = fixed[from];
		}
	}

	__global__ void __launch_bounds__(256) zeroRangeKERNEL(unsigned int* const word, const unsigned int n) {
		const unsigned int i = blockIdx.x * 256u + threadIdx.x;
		if (i < n) {
			word[i] = 0u;
		}
	}

	__global__ void zeroKERNEL(unsigned int* const word, const unsigned int n) {
		if (threadIdx.x < n) {
			word[threadIdx.x] = 0u;
		}
	}

	// Sums of the per-drop counters: {steps, erode steps, deposit steps, killed}.
	__global__ void __launch_bounds__(256) counterKERNEL(const uint2* const info, const unsigned long long n,
		unsigned long long* const out) {
		unsigned long long s[4] = {0ull, 0ull, 0ull, 0ull};
		for (unsigned long long i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
			const uint2 v = info[i];
			s[0] += v.x & 0xFFFFu;
			s[1] += v.x >> 16;
			s[2] += v.y & 0xFFFFu;
			s[3] += v.y >> 16;
		}
		for (int k = 0; k < 4; k++) {
			for (int o = 32; o > 0; o >>= 1) {
				s[k] += __shfl_down(s[k], o, 64);
			}
		}
		if ((threadIdx.x & 63u) == 0u) {
			for (int k = 0; k < 4; k++) {
				atomicAdd(out + k, s[k]);
			}
		}
	}

	__global__ void zero64KERNEL(unsigned long long* const p) {
		if (threadIdx.x < 5u) {
			p[threadIdx.x] = 0ull;
		}
	}

	// Sum of the per-pixel bin counts (the histogram bins the blend evaluates) of chunk blockIdx.y's
	// Side^2 pixels (not the padding after them).
	__global__ void __launch_bounds__(256) binTotalKERNEL(const Chunk* const chunks, const unsigned char* const count,
		unsigned long long* const out) {
		const Chunk c = chunks[blockIdx.y];
		const unsigned long long n = static_cast<unsigned long long>(c.Side) * c.Side;
		unsigned long long s = 0ull;
		for (unsigned long long i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
			s += count[c.Pixel + i];
		}
		for (int o = 32; o > 0; o >>= 1) {
			s += __shfl_down(s, o, 64);
		}
		if ((threadIdx.x & 63u) == 0u) {
			atomicAdd(out, s);
		}
	}

} // namespace SuperTerrainPlus::STPBatch

using namespace SuperTerrainPlus::STPBatch;

namespace {
	template <class T>
	STPSmartDeviceMemory::STPDevice<T> upload(const std::vector<T>& v) {
		auto d = STPSmartDeviceMemory::makeDevice<T>(v.size() ? v.size() : 1u);
		if (!v.empty()) {
			STP_CHECK_HIP(hipMemcpy(d.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
		}
		return d;
	}
} // namespace

STPBatchCapacity STPHeightfieldBatch::capacityOf(const std::vector<STPBatchChunk>& chunks, const float drops_per_pixel) {
	STPBatchCapacity cap {0ull, 0ull, static_cast<unsigned int>(chunks.size()), 0u};
	for (const STPBatchChunk& c : chunks) {
		const unsigned long long px = static_cast<unsigned long long>(c.Side) * c.Side;
		cap.Pixels += (px + 3ull) & ~3ull;
		cap.Drops += static_cast<unsigned long long>(static_cast<double>(px) * drops_per_pixel);
		cap.MaxSide = std::max(cap.MaxSide, c.Side);
	}
	return cap;
}

STPHeightfieldBatch::STPHeightfieldBatch(const STPEnvironment::STPSimplexNoiseSetting& noise,
	const std::vector<STPBiomeProperty>& biome, const STPEnvironment::STPRainDropSetting& rain,
	const STPBatchSetting& setting, const std::vector<STPBatchChunk>& chunks) :
	STPHeightfieldBatch(noise, biome, rain, setting, capacityOf(chunks, setting.DropsPerPixel)) {
	if (chunks.empty()) {
		STP_THROW(STPInvalidArgument, "empty chunk list");
	}
	this->bind(chunks, {}, nullptr);
	STP_CHECK_HIP(hipStreamSynchronize(nullptr));
}

STPHeightfieldBatch::STPHeightfieldBatch(const STPEnvironment::STPSimplexNoiseSetting& noise,
	const std::vector<STPBiomeProperty>& biome, const STPEnvironment::STPRainDropSetting& rain,
	const STPBatchSetting& setting, const STPBatchCapacity& capacity) :
	Setting(setting), Capacity(capacity), Permutation(STPAlgorithm::STPPermutationGenerator::generate(noise)),
	Biome(packBiomeTable(biome.data(), static_cast<unsigned int>(biome.size()))),
	BiomeCount(static_cast<unsigned int>(biome.size())),
	WindowHeights(std::all_of(biome.begin(), biome.end(), [](const STPBiomeProperty& b) {
		// a biome's heights are fractal * Variation + Depth with fractal in [0, 1]
		return b.Depth >= 0.0f && b.Variation >= 0.0f && b.Depth + b.Variation <= STPStage::WindowHeightLimit;
	})) {
	this->init(rain);
}

void STPHeightfieldBatch::init(const STPEnvironment::STPRainDropSetting& rain) {
	const STPBatchSetting& setting = this->Setting;
	STPEnvironment::STPRainDropSetting r = rain;
	r.RainDropCount = 1u; // per chunk counts come from DropsPerPixel
	r.validate();
	const unsigned int radius = setting.Radius;
	STP_ASSERT_DOMAIN(radius > 0u && (radius & 1u) == 0u, "radius should be an positive even number");
	if (setting.BinCapacity == 0u || setting.BinCapacity > STPHeightfieldKernel::MaxBinCapacity) {
		STP_THROW(STPInvalidArgument, "histogram bin capacity must be in [1, 64]");
	}
	if (setting.BiomeCount == 0u || setting.BiomeCell == 0u || !(setting.DropsPerPixel >= 0.0f)) {
		STP_THROW(STPInvalidArgument, "invalid batch setting");
	}
	if (setting.ExternalBiome && setting.BiomeLayers) {
		STP_THROW(STPInvalidArgument, "a batch's biome map comes from a layer chain or an external map, not both");
	}
	const STPBatchCapacity& cap = this->Capacity;
	if (cap.Chunks == 0u || cap.Pixels == 0ull || cap.MaxSide < 2u * r.ErosionBrushRadius + 2u || cap.MaxSide > 16384u) {
		STP_THROW(STPInvalidArgument, "invalid batch capacity");
	}
	if (cap.Drops >= 0xFFFFFFFFull) {
		STP_THROW(STPInvalidArgument, "too many drops in one batch");
	}
	this->VCap = std::min(setting.BinCapacity, 2u * radius + 1u);
	this->Rain.Seed = r.Seed;
	this->Rain.RainDropCount = 0u;
	this->Rain.Inertia = r.Inertia;
	this->Rain.SedimentCapacityFactor = r.SedimentCapacityFactor;
	this->Rain.minSedimentCapacity = r.minSedimentCapacity;
	this->Rain.initWaterVolume = r.initWaterVolume;
	this->Rain.minWaterVolume = r.minWaterVolume;
	this->Rain.Friction = r.Friction;
	this->Rain.initSpeed = r.initSpeed;
	this->Rain.ErodeSpeed = r.ErodeSpeed;
	this->Rain.DepositSpeed = r.DepositSpeed;
	this->Rain.EvaporateSpeed = r.EvaporateSpeed;
	this->Rain.Gravity = r.Gravity;
	this->Rain.ErosionBrushRadius = r.ErosionBrushRadius;
	this->Rain.MaxLifetime = r.MaxLifetime;

	std::vector<int2> offset;
	std::vector<float> weight;
	STPErosionBrushGenerator::generateOffsets(r.ErosionBrushRadius, offset, weight);
	this->BrushSize = static_cast<unsigned int>(offset.size());
	this->BrushOffset = upload(offset);
	this->BrushWeight = upload(weight);
	for (unsigned int k = 0u; shapedRadius(this->BrushSize) != 0 && k < this->BrushSize; k++) {
		const int d2 = offset[k].x * offset[k].x + offset[k].y * offset[k].y;
		const float w = weight[k] * STPStage::FixedScale;
		if (w != shapedWeight(static_cast<int>(r.ErosionBrushRadius), d2)) {
			STP_THROW(STPNumericDomainError, "erosion brush weight differs from the tiled kernel's table");
		}
	}

	// resident arenas, sized by the capacity: (side + 2r)^2 <= side^2 + 4r side + 4r^2 per chunk
	const unsigned long long pixel = cap.Pixels, drops = std::max<unsigned long long>(cap.Drops, 1ull);
	const unsigned long long r2 = 2ull * radius;
	const unsigned long long sideSum = std::min<unsigned long long>(pixel, static_cast<unsigned long long>(cap.Chunks) * cap.MaxSide);
	const unsigned long long halo = pixel + 2ull * r2 * sideSum + r2 * r2 * cap.Chunks;
	const unsigned long long line = pixel + r2 * sideSum;
	const unsigned int binCap = setting.BinCapacity;
	this->Height = STPSmartDeviceMemory::makeDevice<float>(pixel);
	this->Fixed = STPSmartDeviceMemory::makeDevice<unsigned short>(pixel);
	this->BiomeMap = STPSmartDeviceMemory::makeDevice<unsigned short>(halo);
	// the register path's slots are words, the item path's groups itemSlots(VCap) 16-bit slots
	this->VBin = STPSmartDeviceMemory::makeDevice<unsigned int>(
		line * std::max(this->VCap, this->itemPath() ? itemSlots(this->VCap) / 2u : 0u));
	// the item path reads a column's bin count off its slot group
	this->VCount = STPSmartDeviceMemory::makeDevice<unsigned char>(this->itemPath() ? 1ull : line);
	// slots of 3 words for each of at most 8 biomes: a few-line binding holds at most
	// SegmentTarget slots, the long-line split at most sum (side / LongSegment + 1) (side + 2r)
	// (a chunk with a free-slip halo splits its lines into three regions: at most 2 more segments)
	this->SegCapacity = SegmentTarget + pixel / LongSegment + sideSum + sideSum * r2 / LongSegment + r2 * cap.Chunks + 2ull
		+ 2ull * (sideSum + r2 * cap.Chunks);
	this->SegState = STPSmartDeviceMemory::makeDevice<unsigned int>(this->SegCapacity * 3ull * 8ull);
	this->BinCount = STPSmartDeviceMemory::makeDevice<unsigned char>(pixel);
	this->BinItem = STPSmartDeviceMemory::makeDevice<unsigned short>(pixel * binCap);
	this->BinWeight = STPSmartDeviceMemory::makeDevice<float>(pixel * binCap);
	this->DropRank = STPSmartDeviceMemory::makeDevice<unsigned int>(drops);
	this->DropBin = STPSmartDeviceMemory::makeDevice<unsigned int>(drops);
	this->Ledger = STPSmartDeviceMemory::makeDevice<float>(drops);
	this->DropInfo = STPSmartDeviceMemory::makeDevice<uint2>(drops);
	this->Flag = STPSmartDeviceMemory::makeDevice<unsigned int>(4u);
	// a batch that only ever runs some stages (an eroder: stages 4..6) never clears the others' words
	STP_CHECK_HIP(hipMemset(this->Flag.get(), 0, 4u * sizeof(unsigned int)));
	// 0..3 the drop counters, 4 the bin total
	this->Counter = STPSmartDeviceMemory::makeDevice<unsigned long long>(5u);
	STP_CHECK_HIP(hipMemset(this->Fixed.get(), 0, pixel * sizeof(unsigned short)));
	STP_CHECK_HIP(hipMemset(this->Height.get(), 0, pixel * sizeof(float)));
	{
		const int vb = static_cast<int>(verticalLdsWords(this->VCap) * sizeof(unsigned int)),
			hb = static_cast<int>(horizontalLdsWords(this->VCap, setting.BinCapacity) * sizeof(unsigned int));
		const void* const vk[] = {reinterpret_cast<const void*>(&verticalKERNEL<0u>),
			reinterpret_cast<const void*>(&verticalKERNEL<4u>), reinterpret_cast<const void*>(&verticalKERNEL<8u>),
			reinterpret_cast<const void*>(&verticalKERNEL<16u>)};
		const void* const hk[] = {reinterpret_cast<const void*>(&horizontalKERNEL<0u>),
			reinterpret_cast<const void*>(&horizontalKERNEL<4u>), reinterpret_cast<const void*>(&horizontalKERNEL<8u>)};
		for (const void* k : vk) {
			STP_CHECK_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, vb));
		}
		for (const void* k : hk) {
			STP_CHECK_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, hb));
		}
	}
	this->Tuning = this->defaultTuning();
	this->ErosionHalo = this->Tuning.ErosionTile ? this->Rain.MaxLifetime + std::max(this->Rain.ErosionBrushRadius, 2u) + 1u : 0u;
	this->ErosionLds = this->erosionLdsBytes(this->Tuning.ErosionTile);
	for (hipEvent_t& e : this->Event) {
		STP_CHECK_HIP(hipEventCreate(&e));
	}
	STP_CHECK_HIP(hipEventCreateWithFlags(&this->TableCopied, hipEventDisableTiming));
	STP_CHECK_HIP(hipEventCreateWithFlags(&this->OutCopied, hipEventDisableTiming));
	STP_CHECK_HIP(hipStreamCreateWithFlags(&this->CaptureStream, hipStreamNonBlocking));
	STP_CHECK_HIP(hipEventCreateWithFlags(&this->Fork, hipEventDisableTiming));
	for (int k = 0; k < SideStreams; k++) {
		STP_CHECK_HIP(hipStreamCreateWithFlags(&this->Side[k], hipStreamNonBlocking));
		STP_CHECK_HIP(hipEventCreateWithFlags(&this->Join[k], hipEventDisableTiming));
	}
}

void STPHeightfieldBatch::bind(const std::vector<STPBatchChunk>& chunks, const std::vector<unsigned int>& stream_id,
	const hipStream_t stream) {
	if (chunks.empty() || chunks.size() > this->Capacity.Chunks) {
		STP_THROW(STPInvalidArgument, "the chunk list is empty or exceeds the batch capacity");
	}
	if (!stream_id.empty() && stream_id.size() != chunks.size()) {
		STP_THROW(STPInvalidArgument, "one stream id per chunk");
	}
	const unsigned int radius = this->Setting.Radius;
	std::vector<Chunk> rec;
	rec.reserve(chunks.size());
	unsigned long long pixel = 0ull, halo = 0ull, line = 0ull, drops = 0ull;
	for (size_t i = 0u; i < chunks.size(); i++) {
		const STPBatchChunk& in = chunks[i];
		if (in.Side < 2u * this->Rain.ErosionBrushRadius + 2u || in.Side > this->Capacity.MaxSide) {
			STP_THROW(STPInvalidArgument, "chunk side must be in [2r + 2, the capacity's largest side]");
		}
		const unsigned int step = in.Step == 0u ? 1u : in.Step;
		if ((step & (step - 1u)) != 0u || step > 65536u) {
			STP_THROW(STPInvalidArgument, "chunk step must be a power of two up to 65536");
		}
		const unsigned int nn = in.Neighbour == 0u ? 1u : in.Neighbour;
		if ((nn & 1u) == 0u || in.Side % nn != 0u || in.Side / nn < 2u * this->Rain.ErosionBrushRadius + 2u) {
			STP_THROW(STPInvalidArgument, "a neighbourhood is odd, divides the side and keeps maps of at least 2r + 2");
		}
		const unsigned int slip = in.ErosionHalo;
		if (slip != 0u && (nn != 1u || in.Side <= 2u * slip
			|| in.Side - 2u * slip < 2u * this->Rain.ErosionBrushRadius + 2u)) {
			STP_THROW(STPInvalidArgument, "a free-slip halo needs Neighbour 1 and a centre map of at least 2r + 2");
		}
		Chunk c;
		c.Side = in.Side;
		c.OriginX = in.OriginX;
		c.OriginY = in.OriginY;
		// a neighbourhood spawns in its centre map (STPHeightfieldKernel.cu:93-102), a halo chunk in
		// its centre, a plain chunk everywhere
		c.SpawnSide = slip != 0u ? in.Side - 2u * slip : in.Side / nn;
		c.SpawnOrigin = slip != 0u ? slip : (nn / 2u) * c.SpawnSide;
		c.Crop = slip;
		const unsigned long long px = static_cast<unsigned long long>(in.Side) * in.Side,
			spawnPx = static_cast<unsigned long long>(c.SpawnSide) * c.SpawnSide;
		c.Drops = in.Drops != STPBatchChunk::AutoDrops ? in.Drops
													   : static_cast<unsigned int>(static_cast<double>(spawnPx) * this->Setting.DropsPerPixel);
		c.Pixel = pixel;
		c.Halo = halo;
		c.Line = line;
		c.DropBase = static_cast<unsigned int>(drops);
		c.StreamId = stream_id.empty() ? static_cast<unsigned int>(i) : stream_id[i];
		c.TileBase = 0u;
		c.Step = step;
		rec.push_back(c);
		// keep every chunk's first pixel 16-byte aligned in the float arena
		pixel += (px + 3ull) & ~3ull;
		halo += static_cast<unsigned long long>(in.Side + 2u * radius) * (in.Side + 2u * radius);
		line += static_cast<unsigned long long>(in.Side + 2u * radius) * in.Side;
		drops += c.Drops;
	}
	if (pixel > this->Capacity.Pixels || drops > std::max<unsigned long long>(this->Capacity.Drops, 1ull)) {
		STP_THROW(STPInvalidArgument, "the chunk list exceeds the batch capacity");
	}
	// filter line segments: the largest power-of-two factor with lines * factor <= SegmentTarget,
	// so the split binding fits the segment-state arena
	unsigned long long lines = 0ull;
	for (const Chunk& c : rec) {
		lines += c.Side + 2ull * radius;
	}
	unsigned int factor = 1u;
	while (this->itemPath() && factor < MaxSegments && lines * factor * 2ull <= SegmentTarget) {
		factor *= 2u;
	}
	unsigned long long slots = 0ull;
	this->Segmented = false;
	for (Chunk& c : rec) {
		const unsigned int few = factor > 1u ? std::max(1u, std::min(factor, c.Side / MinSegment)) : 1u;
		c.Segs = this->itemPath() ? std::max(few, (c.Side + LongSegment - 1u) / LongSegment) : 1u;
		c.SegLen = (c.Side + c.Segs - 1u) / c.Segs;
		if (c.Crop != 0u) {
			// halo, centre and halo split apart (segmentRange)
			if (!this->itemPath()) {
				STP_THROW(STPInvalidArgument, "a free-slip halo needs the filter's item path (at most 8 biomes, a bin "
					"capacity of at most 8, a radius of at most 127)");
			}
			const unsigned int len = c.SegLen;
			c.Segs = 2u * ((c.Crop + len - 1u) / len) + (c.Side - 2u * c.Crop + len - 1u) / len;
		}
		c.SegLine = slots;
		if (c.Segs > 1u) {
			slots += static_cast<unsigned long long>(c.Segs) * (c.Side + 2u * radius);
			this->Segmented = true;
		}
	}
	if (slots > this->SegCapacity) {
		STP_THROW(STPCapacityError, "filter line segments exceed the segment-state arena");
	}
	this->releaseGraphs();
	if (this->Setting.BiomeLayers) {
		// halo sample (x, y) of a chunk is at world Origin + (x - r, y - r) * Step
		std::vector<STPDiversity::STPLayerMap> maps(rec.size());
		for (size_t i = 0u; i < rec.size(); i++) {
			const Chunk& c = rec[i];
			const unsigned int w = c.Side + 2u * radius;
			const long long ox = c.OriginX - static_cast<long long>(radius) * c.Step,
				oy = c.OriginY - static_cast<long long>(radius) * c.Step;
			if (ox < INT32_MIN || oy < INT32_MIN) {
				STP_THROW(STPInvalidArgument, "a chunk's biome halo leaves the int coordinate range of the biome layers");
			}
			maps[i] = STPDiversity::STPLayerMap {make_int2(static_cast<int>(ox), static_cast<int>(oy)), make_uint2(w, w), c.Step,
				w, c.Halo};
		}
		if (!this->LayerPlan) {
			this->LayerPlan = std::make_unique<STPDiversity::STPLayerPlan>();
		}
		this->Setting.BiomeLayers->plan(*this->LayerPlan, maps, stream);
	}
	this->Record = std::move(rec);
	this->Pixels = pixel;
	this->Drops = drops;
	this->buildTileTables(stream);
}

void STPHeightfieldBatch::copyOut(const std::vector<float*>& height, const std::vector<unsigned short*>& fixed,
	const hipStream_t stream) {
	const size_t n = this->Record.size();
	if (height.size() != n || fixed.size() != n || n > 65535u) {
		STP_THROW(STPInvalidArgument, "one height and one UINT16 destination per bound chunk (at most 65535)");
	}
	for (size_t i = 0u; i < n; i++) {
		if (!height[i] || !fixed[i]) {
			STP_THROW(STPInvalidArgument, "null copy destination");
		}
	}
	if (2u * n > this->OutCapacity) {
		STP_CHECK_HIP(hipStreamSynchronize(stream));
		this->OutCapacity = 2u * n + n / 2u + 2u;
		this->OutTable = STPSmartDeviceMemory::makeDevice<void*>(this->OutCapacity);
		this->OutStaging = STPSmartDeviceMemory::makeHost<void*>(this->OutCapacity);
	}
	// the staging buffer may still feed the previous copy
	STP_CHECK_HIP(hipEventSynchronize(this->OutCopied));
	for (size_t i = 0u; i < n; i++) {
		this->OutStaging[i] = height[i];
		this->OutStaging[n + i] = fixed[i];
	}
	STP_CHECK_HIP(hipMemcpyAsync(this->OutTable.get(), this->OutStaging.get(), 2u * n * sizeof(void*),
		hipMemcpyHostToDevice, stream));
	STP_CHECK_HIP(hipEventRecord(this->OutCopied, stream));
	copyOutKERNEL<<<dim3(64u, static_cast<unsigned int>(n)), 256u, 0, stream>>>(this->ChunkRecord,
		static_cast<unsigned int>(n), this->Height.get(), this->Fixed.get(), this->OutTable.get());
	STP_CHECK_HIP(hipGetLastError());
}

float* STPHeightfieldBatch::heightOf(const unsigned int i) const {
	if (i >= this->Record.size()) {
		STP_THROW(STPInvalidArgument, "chunk index out of range");
	}
	return this->Height.get() + this->Record[i].Pixel;
}

const unsigned char* STPHeightfieldBatch::binCountOf(const unsigned int i) const {
	if (i >= this->Record.size()) {
		STP_THROW(STPInvalidArgument, "chunk index out of range");
	}
	return this->BinCount.get() + this->Record[i].Pixel;
}

const unsigned short* STPHeightfieldBatch::binItemOf(const unsigned int i) const {
	if (i >= this->Record.size()) {
		STP_THROW(STPInvalidArgument, "chunk index out of range");
	}
	return this->BinItem.get() + this->Record[i].Pixel * this->Setting.BinCapacity;
}

const float* STPHeightfieldBatch::binWeightOf(const unsigned int i) const {
	if (i >= this->Record.size()) {
		STP_THROW(STPInvalidArgument, "chunk index out of range");
	}
	return this->BinWeight.get() + this->Record[i].Pixel * this->Setting.BinCapacity;
}

std::unique_ptr<STPHeightfieldBatch> STPHeightfieldBatch::makeFilter(const std::vector<STPBiomeProperty>& biome,
	const unsigned int side, const unsigned int radius, const unsigned int bin_capacity) {
	if (biome.empty() || biome.size() > 8u || bin_capacity == 0u || bin_capacity > 8u || radius == 0u || radius > 126u
		|| (radius & 1u) != 0u || side < 6u || side > 16384u) {
		return nullptr;
	}
	const STPEnvironment::STPSimplexNoiseSetting noise {0ull, 8u, 0.0};
	STPEnvironment::STPRainDropSetting rain {};
	rain.RainDropCount = 1u;
	rain.Inertia = 0.05f;
	rain.SedimentCapacityFactor = 4.0f;
	rain.minSedimentCapacity = 0.01f;
	rain.initWaterVolume = 1.0f;
	rain.minWaterVolume = 0.01f;
	rain.initSpeed = 1.0f;
	rain.ErodeSpeed = 0.3f;
	rain.DepositSpeed = 0.3f;
	rain.EvaporateSpeed = 0.01f;
	rain.Gravity = 4.0f;
	rain.ErosionBrushRadius = 2u;
	rain.MaxLifetime = 30u;
	STPBatchSetting setting {};
	setting.BiomeSeed = 0ull;
	setting.BiomeCount = static_cast<unsigned int>(biome.size());
	setting.BiomeCell = 1u;
	setting.Radius = radius;
	setting.BinCapacity = bin_capacity;
	setting.DropsPerPixel = 0.0f;
	setting.ExternalBiome = true;
	STPBatchCapacity cap {};
	cap.Pixels = static_cast<unsigned long long>(side) * side + 4ull;
	cap.Drops = 1ull;
	cap.Chunks = 1u;
	cap.MaxSide = side;
	return std::make_unique<STPHeightfieldBatch>(noise, biome, rain, setting, cap);
}

void STPHeightfieldBatch::filterExternal(const STPSample_t* const origin, const unsigned int pitch, const hipStream_t stream) {
	if (this->Capacity.Chunks != 1u || !this->Setting.ExternalBiome) {
		STP_THROW(STPInvalidArgument, "filterExternal is for one-chunk batches with ExternalBiome");
	}
	this->bind({STPBatchChunk {this->Capacity.MaxSide, 0, 0, 1u, 1u}}, {}, stream);
	this->setBiomeSource(origin, pitch);
	this->runStages(0, 3, stream);
}

void STPHeightfieldBatch::setBiomeSource(const STPSample_t* const origin, const unsigned int pitch) {
	if (!this->Setting.ExternalBiome) {
		STP_THROW(STPInvalidArgument, "a biome source is for batches with ExternalBiome");
	}
	if (!origin || pitch == 0u) {
		STP_THROW(STPInvalidArgument, "null biome source or zero pitch");
	}
	if (origin != this->BiomeSource || pitch != this->BiomeSourcePitch) {
		// captured stage graphs hold the previous source
		this->releaseGraphs();
	}
	this->BiomeSource = origin;
	this->BiomeSourcePitch = pitch;
}

unsigned short* STPHeightfieldBatch::fixedOf(const unsigned int i) const {
	if (i >= this->Record.size()) {
		STP_THROW(STPInvalidArgument, "chunk index out of range");
	}
	return this->Fixed.get() + this->Record[i].Pixel;
}

void STPHeightfieldBatch::releaseGraphs() noexcept {
	for (int k = 0; k < StageCount; k++) {
		if (this->StageExec[k]) {
			(void)hipGraphExecDestroy(this->StageExec[k]);
			this->StageExec[k] = nullptr;
		}
		if (this->StageGraph[k]) {
			(void)hipGraphDestroy(this->StageGraph[k]);
			this->StageGraph[k] = nullptr;
		}
	}
	if (this->GraphExec) {
		(void)hipGraphExecDestroy(this->GraphExec);
		this->GraphExec = nullptr;
	}
	if (this->Graph) {
		(void)hipGraphDestroy(this->Graph);
		this->Graph = nullptr;
	}
}

namespace {
	template <class Kernel>
	void allowLds(Kernel* const kernel, const unsigned int bytes) {
		STP_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
			hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
	}
	template <unsigned int Threads>
	void allowLds(const unsigned int bytes) {
		if constexpr (Threads == 768u) {
			allowLds(&erodeTile768KERNEL<0u>, bytes);
			allowLds(&erodeTile768KERNEL<25u>, bytes);
		} else {
			allowLds(&erodeTileKERNEL<Threads, 0u>, bytes);
			allowLds(&erodeTileKERNEL<Threads, 25u>, bytes);
		}
	}
} // namespace

bool STPHeightfieldBatch::packedBins() const noexcept {
	// a slot holds item (3 bits) | count << 3 in 16 bits: counts below 2^13
	const unsigned int w = 2u * this->Setting.Radius + 1u;
	return this->itemPath() && w * w < 8192u;
}

bool STPHeightfieldBatch::itemPath() const noexcept {
	// biome ids come from the synthetic biome map or a layer chain that replaces ids >= BiomeCount by
	// 0, so they are < BiomeCount; a column count is at
	// most 2r + 1, which the 16-bit slots hold in 8 bits
	return this->Setting.BiomeCount <= 8u && this->Setting.BinCapacity <= 8u && this->Setting.Radius <= 127u;
}

ErosionArgs STPHeightfieldBatch::erosionArgs() const {
	// per-chunk fields (map size, stream id, drop count) are filled in by the kernels
	ErosionArgs args;
	args.Rain = this->Rain;
	args.MapSize = make_uint2(0u, 0u);
	args.SpawnOrigin = make_uint2(0u, 0u);
	args.Total = args.MapSize;
	args.BrushSize = this->BrushSize;
	args.StreamId = 0u;
	return args;
}

unsigned int STPHeightfieldBatch::erosionLdsBytes(const unsigned int tile) const {
	const unsigned int r = this->Rain.ErosionBrushRadius, life = this->Rain.MaxLifetime;
	if (tile == 0u || life == 0u || this->BrushSize > MaxTiledTaps || !this->WindowHeights) {
		return 0u;
	}
	const unsigned long long halo = static_cast<unsigned long long>(life) + std::max(r, 2u) + 1u;
	if (tile < 2ull * halo) {
		return 0u; // the windows of one colour batch would overlap
	}
	const unsigned long long side = std::min<unsigned long long>(this->Capacity.MaxSide, tile + 2ull * halo);
	int device = 0, limit = 0;
	STP_CHECK_HIP(hipGetDevice(&device));
	STP_CHECK_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
	const unsigned long long bytes = side * side * sizeof(float);
	const unsigned long long room = static_cast<unsigned long long>(limit) - MaxTiledTaps * sizeof(int2);
	return bytes <= room ? static_cast<unsigned int>(bytes) : 0u;
}

STPBatchTuning STPHeightfieldBatch::defaultTuning() const {
	for (const unsigned int tile : ErosionTileCandidates) {
		if (this->erosionLdsBytes(tile) != 0u) {
			return STPBatchTuning {tile, 1024u};
		}
	}
	return STPBatchTuning {0u, ErosionBlock};
}

STPBatchTuning STPHeightfieldBatch::latencyTuning() const {
	STPBatchTuning best = this->defaultTuning();
	for (const unsigned int tile : ErosionTileCandidates) {
		if (this->erosionLdsBytes(tile) != 0u) {
			best = STPBatchTuning {tile, 1024u};
		}
	}
	return best;
}

void STPHeightfieldBatch::setTuning(const STPBatchTuning& tuning) {
	const unsigned int th = tuning.ErosionThreads;
	if (th != 256u && th != 512u && th != 768u && th != 1024u) {
		STP_THROW(STPInvalidArgument, "erosion threads must be 256, 512, 768 or 1024");
	}
	if (tuning.ErosionTile != 0u
		&& (tuning.ErosionTile < 16u || tuning.ErosionTile > 4096u || this->erosionLdsBytes(tuning.ErosionTile) == 0u)) {
		STP_THROW(STPInvalidArgument,
			"tiled erosion needs a tile side in [16, 4096], MaxLifetime > 0, a window that fits in LDS and biome "
			"heights (Depth to Depth + Variation) within [0, 4]");
	}
	this->releaseGraphs();
	this->Tuning = tuning;
	this->ErosionHalo = tuning.ErosionTile ? this->Rain.MaxLifetime + std::max(this->Rain.ErosionBrushRadius, 2u) + 1u : 0u;
	this->ErosionLds = this->erosionLdsBytes(tuning.ErosionTile);
	if (!this->Record.empty()) {
		this->buildTileTables(nullptr);
		STP_CHECK_HIP(hipStreamSynchronize(nullptr));
	}
}

void STPHeightfieldBatch::buildTileTables(const hipStream_t stream) {
	const unsigned int radius = this->Setting.Radius, tile = this->Tuning.ErosionTile;
	unsigned int tiles = 0u;
	for (Chunk& c : this->Record) {
		c.TileBase = tiles;
		if (tile != 0u) {
			const unsigned int tx = (c.Side + tile - 1u) / tile;
			tiles += tx * tx;
		}
	}
	this->ErosionTiles = tiles;
	if (tile != 0u) {
		switch (this->Tuning.ErosionThreads) {
		case 256u:
			allowLds<256u>(this->ErosionLds);
			break;
		case 512u:
			allowLds<512u>(this->ErosionLds);
			break;
		case 768u:
			allowLds<768u>(this->ErosionLds);
			break;
		default:
			allowLds<1024u>(this->ErosionLds);
			break;
		}
	}

	// tile tables, chunk-major so that one chunk's tiles are a contiguous range (plain mode)
	std::vector<uint2> table[TableCount];
	this->Range.assign(this->Record.size(), STPTileRange {});
	for (unsigned int i = 0u; i < this->Record.size(); i++) {
		const Chunk& c = this->Record[i];
		STPTileRange& range = this->Range[i];
		const unsigned int w = c.Side + 2u * radius;
		const unsigned int blend =
			((c.Side + BlendX - 1u) / BlendX) * ((c.Side + BlendY * BlendPixel - 1u) / (BlendY * BlendPixel));
		const unsigned int tx = tile ? (c.Side + tile - 1u) / tile : 0u;
		const unsigned int bt = (w + BiomeSide - 1u) / BiomeSide;
		// filter tables: line blocks x segments, segment-major
		const unsigned int n[5] = {bt * bt, ((w + ShfLineBlock - 1u) / ShfLineBlock) * c.Segs,
			((c.Side + ShfLineBlock - 1u) / ShfLineBlock) * c.Segs, blend, (c.Drops + ErosionBlock - 1u) / ErosionBlock};
		for (int s = 0; s < 5; s++) {
			range.Begin[s] = static_cast<unsigned int>(table[s].size());
			for (unsigned int k = 0u; k < n[s]; k++) {
				table[s].push_back(make_uint2(i, k));
			}
			range.End[s] = static_cast<unsigned int>(table[s].size());
		}
		// erosion tiles by colour: (tile x parity) + 2 (tile y parity), rotated by the chunk's stream
		// word. Any relabelling keeps same-colour tiles apart; the rotation spreads the tiles of chunks
		// with few tiles (a one-tile chunk has colour 0 before it) over the four launches, and it
		// depends on the chunk alone, so a chunk's tiles run in the same order in any batch.
		const unsigned int rotation = c.StreamId & 3u;
		for (int colour = 0; colour < 4; colour++) {
			const int s = 5 + colour;
			range.Begin[s] = static_cast<unsigned int>(table[s].size());
			for (unsigned int k = 0u; k < tx * tx; k++) {
				const unsigned int parity = ((k % tx) & 1u) + 2u * ((k / tx) & 1u);
				if (static_cast<int>((parity + rotation) & 3u) == colour) {
					table[s].push_back(make_uint2(i, k));
				}
			}
			range.End[s] = static_cast<unsigned int>(table[s].size());
		}
	}

	// one arena: [chunk records][table 0] ... [table 8], 256-byte aligned parts
	const auto align = [](const size_t b) { return (b + 255u) & ~static_cast<size_t>(255u); };
	size_t offset[TableCount + 1];
	size_t bytes = align(this->Record.size() * sizeof(Chunk));
	for (int s = 0; s < TableCount; s++) {
		offset[s] = bytes;
		bytes += align(std::max<size_t>(table[s].size(), 1u) * sizeof(uint2));
	}
	if (bytes > this->TableBytes || tiles + 1u > this->TileCapacity) {
		// growing: the previous tables may still be read by work on `stream`
		STP_CHECK_HIP(hipStreamSynchronize(stream));
		if (bytes > this->TableBytes) {
			this->TableBytes = bytes + bytes / 4u;
			this->TableArena = STPSmartDeviceMemory::makeDevice<unsigned char>(this->TableBytes);
			this->TableStaging = STPSmartDeviceMemory::makeHost<unsigned char>(this->TableBytes);
		}
		if (tiles + 1u > this->TileCapacity) {
			this->TileCapacity = tiles + 1u + (tiles + 1u) / 4u;
			this->TileDrops = STPSmartDeviceMemory::makeDevice<unsigned int>(this->TileCapacity);
			this->TileFirst = STPSmartDeviceMemory::makeDevice<unsigned int>(this->TileCapacity);
		}
	}
	// the staging buffer may still feed the previous binding's copy
	STP_CHECK_HIP(hipEventSynchronize(this->TableCopied));
	unsigned char* const host = this->TableStaging.get();
	std::memcpy(host, this->Record.data(), this->Record.size() * sizeof(Chunk));
	this->ChunkRecord = reinterpret_cast<Chunk*>(this->TableArena.get());
	for (int s = 0; s < TableCount; s++) {
		if (!table[s].empty()) {
			std::memcpy(host + offset[s], table[s].data(), table[s].size() * sizeof(uint2));
		}
		this->TileCount[s] = static_cast<unsigned int>(table[s].size());
		this->Tile[s] = reinterpret_cast<uint2*>(this->TableArena.get() + offset[s]);
	}
	STP_CHECK_HIP(hipMemcpyAsync(this->TableArena.get(), host, bytes, hipMemcpyHostToDevice, stream));
	STP_CHECK_HIP(hipEventRecord(this->TableCopied, stream));
}

STPBatchTuning STPHeightfieldBatch::autotune(const hipStream_t stream, const unsigned int repeat) {
	const bool timing = this->Timing;
	std::vector<STPBatchTuning> candidate;
	for (const unsigned int tile : ErosionTileCandidates) {
		if (this->erosionLdsBytes(tile) == 0u) {
			continue;
		}
		for (const unsigned int th : {256u, 512u, 768u, 1024u}) {
			// at 768 threads a drop gets 80 VGPRs, which the 25-tap drop fits and the drop of every
			// brush size does not (it spills)
			if (th != 768u || this->BrushSize == 25u) {
				candidate.push_back(STPBatchTuning {tile, th});
			}
		}
	}
	if (candidate.empty()) {
		candidate.push_back(STPBatchTuning {0u, ErosionBlock});
	}
	// Each shape is scored on whole graph runs (the mode a caller runs, with four concurrent chunk
	// chains on large bindings): the tuning changes only the erosion binning and erosion, but how
	// well they share the GPU with the other chains' stages depends on the shape.
	STPBatchTuning best = candidate.front();
	float bestMs = 0.0f;
	hipEvent_t begin = nullptr, end = nullptr;
	this->Timing = false;
	try {
		STP_CHECK_HIP(hipEventCreate(&begin));
		STP_CHECK_HIP(hipEventCreate(&end));
		for (const STPBatchTuning& t : candidate) {
			this->setTuning(t);
			// the first run captures the graph and warms up
			this->run(STPRunMode::Graph, stream);
			STP_CHECK_HIP(hipEventRecord(begin, stream));
			for (unsigned int k = 0u; k < std::max(repeat, 1u); k++) {
				this->run(STPRunMode::Graph, stream);
			}
			STP_CHECK_HIP(hipEventRecord(end, stream));
			STP_CHECK_HIP(hipEventSynchronize(end));
			float ms = 0.0f;
			STP_CHECK_HIP(hipEventElapsedTime(&ms, begin, end));
			if (bestMs == 0.0f || ms < bestMs) {
				bestMs = ms;
				best = t;
			}
		}
	} catch (...) {
		(void)hipStreamSynchronize(stream);
		if (begin) {
			(void)hipEventDestroy(begin);
		}
		if (end) {
			(void)hipEventDestroy(end);
		}
		this->Timing = timing;
		throw;
	}
	(void)hipEventDestroy(begin);
	(void)hipEventDestroy(end);
	this->Timing = timing;
	this->setTuning(best);
	return best;
}

STPHeightfieldBatch::~STPHeightfieldBatch() {
	this->releaseGraphs();
	for (hipEvent_t e : this->Event) {
		(void)hipEventDestroy(e);
	}
	if (this->TableCopied) {
		(void)hipEventDestroy(this->TableCopied);
	}
	if (this->OutCopied) {
		(void)hipEventDestroy(this->OutCopied);
	}
	if (this->CaptureStream) {
		(void)hipStreamDestroy(this->CaptureStream);
	}
	for (int k = 0; k < SideStreams; k++) {
		if (this->Side[k]) {
			(void)hipStreamSynchronize(this->Side[k]);
			(void)hipStreamDestroy(this->Side[k]);
		}
		if (this->Join[k]) {
			(void)hipEventDestroy(this->Join[k]);
		}
	}
	if (this->Fork) {
		(void)hipEventDestroy(this->Fork);
	}
}

std::vector<unsigned int> STPHeightfieldBatch::pipelineGroups() const {
	static_assert(PipelineGroups >= 1u && PipelineGroups <= SideStreams + 1u, "a chain per side stream");
	const unsigned int n = static_cast<unsigned int>(this->Record.size());
	const unsigned int groups = this->Pixels < PipelinePixels ? 1u : std::min(n, PipelineGroups);
	std::vector<unsigned int> bound {0u};
	unsigned long long acc = 0ull;
	for (unsigned int i = 0u; i < n && bound.size() < groups; i++) {
		acc += static_cast<unsigned long long>(this->Record[i].Side) * this->Record[i].Side;
		// close group g once it holds at least g / groups of the pixels, leaving a chunk for each later group
		if (acc * groups >= this->Pixels * bound.size() && n - (i + 1u) >= groups - bound.size()) {
			bound.push_back(i + 1u);
		}
	}
	bound.push_back(n);
	return bound;
}

void STPHeightfieldBatch::enqueueStage(const hipStream_t stream, const int stage, const unsigned int first,
	const unsigned int last, const bool clear_flag, const bool split_erosion) {
	const unsigned int radius = this->Setting.Radius, cap = this->Setting.BinCapacity;
	const auto range = [&](const int s) {
		const unsigned int b = this->Range[first].Begin[s], e = this->Range[last - 1u].End[s];
		return std::make_pair(this->Tile[s] + b, e - b);
	};
	switch (stage) {
	case 0: {
		if (clear_flag) {
			zeroKERNEL<<<1, 64, 0, stream>>>(this->Flag.get(), 4u);
		}
		if (this->Setting.BiomeLayers) {
			this->Setting.BiomeLayers->run(*this->LayerPlan, this->BiomeMap.get(), first, last, this->Setting.BiomeCount, stream);
			break;
		}
		if (this->Setting.ExternalBiome) {
			if (!this->BiomeSource) {
				STP_THROW(STPInvalidArgument, "an external-biome batch runs stage 0 after setBiomeSource");
			}
			const auto [t, n] = range(0);
			if (n) {
				externalBiomeKERNEL<<<n, BiomeTile, 0, stream>>>(t, this->ChunkRecord, this->BiomeSource,
					this->BiomeSourcePitch, radius, this->itemPath() ? this->Setting.BiomeCount : 0x10000u,
					this->BiomeMap.get(), this->Flag.get());
			}
			break;
		}
		const auto [t, n] = range(0);
		if (n) {
			biomeKERNEL<<<n, BiomeTile, 0, stream>>>(t, this->ChunkRecord, this->Setting.BiomeSeed,
				this->Setting.BiomeCount, this->Setting.BiomeCell, radius, this->BiomeMap.get());
		}
		break;
	}
	case 1: {
		const auto tn = range(1);
		const uint2* const t = tn.first;
		const unsigned int n = tn.second;
		if (n) {
			const auto launch = [&](auto kernel) {
				kernel<<<n, ShfLineBlock, verticalLdsWords(this->VCap) * sizeof(unsigned int), stream>>>(t,
					this->ChunkRecord, radius, this->VCap, this->BiomeMap.get(), this->VBin.get(),
					this->VCount.get(), this->Flag.get());
			};
			if (this->itemPath()) {
				const unsigned int words = verticalItemWords(this->VCap);
				unsigned int* const state = this->SegState.get();
				// the item path's column histograms are groups of 16-bit slots (itemSlots)
				const auto item = [&](auto summary, auto carry, auto filter) {
					if (this->Segmented) {
						summary<<<n, ShfLineBlock, 0, stream>>>(t, this->ChunkRecord, radius, this->BiomeMap.get(), state);
						carry<<<n, ShfLineBlock, 0, stream>>>(t, this->ChunkRecord, radius, 2u * radius, state);
					}
					filter<<<n, ShfLineBlock, words * sizeof(unsigned int), stream>>>(t, this->ChunkRecord, radius,
						this->VCap, this->BiomeMap.get(), this->VBin.get(), state, this->Flag.get());
				};
				if (this->Setting.BiomeCount <= 4u) {
					item(verticalSummaryKERNEL<4u>, segmentCarryKERNEL<4u>, verticalItemKERNEL<4u>);
				} else {
					item(verticalSummaryKERNEL<8u>, segmentCarryKERNEL<8u>, verticalItemKERNEL<8u>);
				}
				break;
			}
			switch (registerCapacity(this->VCap)) {
			case 4u:
				launch(verticalKERNEL<4u>);
				break;
			case 8u:
				launch(verticalKERNEL<8u>);
				break;
			case 16u:
				launch(verticalKERNEL<16u>);
				break;
			default:
				launch(verticalKERNEL<0u>);
				break;
			}
		}
		break;
	}
	case 2: {
		const auto tn = range(2);
		const uint2* const t = tn.first;
		const unsigned int n = tn.second;
		if (n) {
			const auto launch = [&](auto kernel) {
				kernel<<<n, ShfLineBlock, horizontalLdsWords(this->VCap, cap) * sizeof(unsigned int), stream>>>(t,
					this->ChunkRecord, radius, this->VCap, cap, this->VBin.get(), this->VCount.get(),
					this->BinCount.get(), this->BinItem.get(), this->BinWeight.get(), this->Flag.get());
			};
			if (this->itemPath()) {
				const unsigned int words = HItemPrefetch * horizontalItemWords(cap, this->packedBins()) * 65u;
				unsigned int* const state = this->SegState.get();
				const unsigned int* const vbin = this->VBin.get();
				const auto item = [&](auto summary, auto carry, auto filter) {
					if (this->Segmented) {
						summary<<<n, ShfLineBlock, 0, stream>>>(t, this->ChunkRecord, radius, this->VCap, vbin, state);
						carry<<<n, ShfLineBlock, 0, stream>>>(t, this->ChunkRecord, radius, 0u, state);
					}
					filter<<<n, ShfLineBlock, words * sizeof(unsigned int), stream>>>(t, this->ChunkRecord, radius,
						this->VCap, cap, vbin, this->BinCount.get(), this->BinItem.get(), this->BinWeight.get(), state,
						this->Flag.get());
				};
				// slots loaded per column group: with at most 4 biomes a window holds at most 4 bins,
				// and a group holds 4 slots when vcap <= 4
				const bool packed = this->packedBins();
				if (this->Setting.BiomeCount <= 4u) {
					if (packed) {
						item(horizontalSummaryKERNEL<4u, 4u>, segmentCarryKERNEL<4u>, horizontalItemKERNEL<4u, 4u, true>);
					} else {
						item(horizontalSummaryKERNEL<4u, 4u>, segmentCarryKERNEL<4u>, horizontalItemKERNEL<4u, 4u, false>);
					}
				} else if (itemSlots(this->VCap) == 4u) {
					if (packed) {
						item(horizontalSummaryKERNEL<8u, 4u>, segmentCarryKERNEL<8u>, horizontalItemKERNEL<8u, 4u, true>);
					} else {
						item(horizontalSummaryKERNEL<8u, 4u>, segmentCarryKERNEL<8u>, horizontalItemKERNEL<8u, 4u, false>);
					}
				} else if (packed) {
					item(horizontalSummaryKERNEL<8u, 8u>, segmentCarryKERNEL<8u>, horizontalItemKERNEL<8u, 8u, true>);
				} else {
					item(horizontalSummaryKERNEL<8u, 8u>, segmentCarryKERNEL<8u>, horizontalItemKERNEL<8u, 8u, false>);
				}
				break;
			}
			switch (horizontalRegisterCapacity(cap)) {
			case 4u:
				launch(horizontalKERNEL<4u>);
				break;
			case 8u:
				launch(horizontalKERNEL<8u>);
				break;
			default:
				launch(horizontalKERNEL<0u>);
				break;
			}
		}
		break;
	}
	case 3: {
		const auto tn = range(3);
		const uint2* const t = tn.first;
		const unsigned int n = tn.second;
		if (n) {
			const unsigned int window = (2u * radius + 1u) * (2u * radius + 1u);
			const auto launch = [&](auto kernel) {
				kernel<<<n, dim3(BlendX, BlendY), 0, stream>>>(t, this->ChunkRecord, this->Permutation.Permutation,
					this->Biome, this->BiomeCount, cap, window, this->BinCount.get(), this->BinItem.get(),
					this->BinWeight.get(), this->Height.get());
			};
			if (this->packedBins()) {
				launch(blendKERNEL<true>);
			} else {
				launch(blendKERNEL<false>);
			}
		}
		break;
	}
	case 4: {
		if (clear_flag) {
			// the erosion window status (a caller running stages 4..6 only, on heights it wrote)
			zeroKERNEL<<<1, 64, 0, stream>>>(this->Flag.get() + 1, 1u);
		}
		// tiled erosion: bin the drops by spawn tile
		const unsigned int tile = this->Tuning.ErosionTile;
		const auto [t, n] = range(4);
		if (tile == 0u || n == 0u) {
			// nothing to bin; one trivial launch keeps the stage's graph non-empty
			zeroRangeKERNEL<<<1, 256, 0, stream>>>(this->TileDrops.get() + this->ErosionTiles, 1u);
			break;
		}
		const ErosionArgs args = this->erosionArgs();
		const unsigned int tb = this->Record[first].TileBase,
			te = last < this->Record.size() ? this->Record[last].TileBase : this->ErosionTiles;
		if (te > tb) {
			zeroRangeKERNEL<<<(te - tb + 255u) / 256u, 256u, 0, stream>>>(this->TileDrops.get() + tb, te - tb);
		}
		binCountKERNEL<<<(n + BinGroups - 1u) / BinGroups, ErosionBlock, 0, stream>>>(t, n, this->ChunkRecord, args, tile,
			this->TileDrops.get(), this->DropRank.get());
		binScanKERNEL<<<last - first, ScanBlock, 0, stream>>>(this->ChunkRecord, first, tile,
			this->TileDrops.get(), this->TileFirst.get());
		binScatterKERNEL<<<n, ErosionBlock, 0, stream>>>(t, this->ChunkRecord, args, tile, this->TileFirst.get(),
			this->DropRank.get(), this->DropBin.get());
		break;
	}
	case 5: {
		const auto [t, n] = range(4);
		if (n == 0u) {
			break;
		}
		const ErosionArgs args = this->erosionArgs();
		const unsigned int tile = this->Tuning.ErosionTile;
		if (tile == 0u) {
			erosionKERNEL<<<n, ErosionBlock, this->BrushSize * sizeof(int2), stream>>>(t, this->ChunkRecord, args,
				this->BrushOffset.get(), this->BrushWeight.get(), this->Height.get(), this->Ledger.get(),
				this->DropInfo.get());
			break;
		}
		// the four colour launches of chunks [f, l) on stream `on`
		const auto colours = [&](const hipStream_t on, const unsigned int f, const unsigned int l) {
			for (int colour = 0; colour < 4; colour++) {
				const unsigned int b = this->Range[f].Begin[5 + colour], e = this->Range[l - 1u].End[5 + colour];
				const uint2* const tt = this->Tile[5 + colour] + b;
				const unsigned int nt = e - b;
				if (nt == 0u) {
					continue;
				}
				const auto launch = [&](auto kernel, const unsigned int threads) {
					kernel<<<nt, threads, this->ErosionLds, on>>>(tt, this->ChunkRecord, args, tile,
						this->ErosionHalo, this->BrushOffset.get(), this->BrushWeight.get(), this->TileDrops.get(),
						this->TileFirst.get(), this->DropBin.get(), this->Height.get(), this->Ledger.get(),
						this->DropInfo.get(), this->Flag.get() + 1);
				};
				const bool taps25 = args.BrushSize == 25u;
				switch (this->Tuning.ErosionThreads) {
				case 256u:
					taps25 ? launch(erodeTileKERNEL<256u, 25u>, 256u) : launch(erodeTileKERNEL<256u, 0u>, 256u);
					break;
				case 512u:
					taps25 ? launch(erodeTileKERNEL<512u, 25u>, 512u) : launch(erodeTileKERNEL<512u, 0u>, 512u);
					break;
				case 768u:
					taps25 ? launch(erodeTile768KERNEL<25u>, 768u) : launch(erodeTile768KERNEL<0u>, 768u);
					break;
				default:
					taps25 ? launch(erodeTileKERNEL<1024u, 25u>, 1024u) : launch(erodeTileKERNEL<1024u, 0u>, 1024u);
					break;
				}
			}
		};
		// Chunks are separate maps, so two groups of chunks (about half the tiles each) erode on two
		// streams: a colour launch of one group fills the CUs that the other group's launch leaves
		// idle in its last round of tiles, instead of every launch waiting out its own last round.
		const auto tilesOf = [&](const unsigned int i) {
			unsigned int t = 0u;
			for (int colour = 0; colour < 4; colour++) {
				t += this->Range[i].End[5 + colour] - this->Range[i].Begin[5 + colour];
			}
			return t;
		};
		unsigned int total = 0u;
		for (unsigned int i = first; i < last; i++) {
			total += tilesOf(i);
		}
		unsigned int mid = first, acc = 0u;
		while (mid < last && 2u * (acc + tilesOf(mid)) <= total) {
			acc += tilesOf(mid);
			mid++;
		}
		if (!split_erosion || mid == first || mid == last || total < SplitErosionTiles) {
			colours(stream, first, last);
			break;
		}
		STP_CHECK_HIP(hipEventRecord(this->Fork, stream));
		STP_CHECK_HIP(hipStreamWaitEvent(this->Side[0], this->Fork, 0));
		colours(stream, first, mid);
		colours(this->Side[0], mid, last);
		STP_CHECK_HIP(hipEventRecord(this->Join[0], this->Side[0]));
		STP_CHECK_HIP(hipStreamWaitEvent(stream, this->Join[0], 0));
		break;
	}
	case 6: {
		const unsigned long long b = this->Record[first].Pixel;
		const Chunk& l = this->Record[last - 1u];
		const unsigned long long e = l.Pixel + static_cast<unsigned long long>(l.Side) * l.Side;
		STPHeightfieldKernel::formatHeightmap(this->Height.get() + b, this->Fixed.get() + b, e - b, stream);
		break;
	}
	default:
		STP_THROW(STPInvalidArgument, "unknown stage");
	}
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldBatch::enqueue(const hipStream_t stream, const unsigned int first, const unsigned int last,
	const bool timed, const bool clear_flag) {
	for (int s = 0; s < StageCount; s++) {
		if (timed) {
			STP_CHECK_HIP(hipEventRecord(this->Event[s], stream));
		}
		this->enqueueStage(stream, s, first, last, clear_flag);
	}
	if (timed) {
		STP_CHECK_HIP(hipEventRecord(this->Event[StageCount], stream));
	}
}

hipGraphExec_t STPHeightfieldBatch::capture(const int stage_begin, const int stage_end, hipGraph_t& graph) {
	const unsigned int n = static_cast<unsigned int>(this->Record.size());
	const std::vector<unsigned int> bound = stage_begin == 0 && stage_end == StageCount ? this->pipelineGroups()
																					 : std::vector<unsigned int> {0u, n};
	const int groups = static_cast<int>(bound.size()) - 1;
	STP_CHECK_HIP(hipStreamBeginCapture(this->CaptureStream, hipStreamCaptureModeThreadLocal));
	bool ok = true;
	try {
		if (groups > 1) {
			// the flags every stage 0 and 4 would clear, cleared once for all chains
			zeroKERNEL<<<1, 64, 0, this->CaptureStream>>>(this->Flag.get(), 4u);
			STP_CHECK_HIP(hipGetLastError());
			STP_CHECK_HIP(hipEventRecord(this->Fork, this->CaptureStream));
			for (int g = 1; g < groups; g++) {
				STP_CHECK_HIP(hipStreamWaitEvent(this->Side[g - 1], this->Fork, 0));
			}
			for (int g = 0; g < groups; g++) {
				const hipStream_t on = g == 0 ? this->CaptureStream : this->Side[g - 1];
				for (int s = 0; s < StageCount; s++) {
					this->enqueueStage(on, s, bound[g], bound[g + 1], false, false);
				}
			}
			for (int g = 1; g < groups; g++) {
				STP_CHECK_HIP(hipEventRecord(this->Join[g - 1], this->Side[g - 1]));
				STP_CHECK_HIP(hipStreamWaitEvent(this->CaptureStream, this->Join[g - 1], 0));
			}
		} else {
			for (int s = stage_begin; s < stage_end; s++) {
				this->enqueueStage(this->CaptureStream, s, 0u, n, true);
			}
		}
	} catch (...) {
		ok = false;
	}
	hipGraph_t g = nullptr;
	const hipError_t end = hipStreamEndCapture(this->CaptureStream, &g);
	if (!ok || end != hipSuccess) {
		if (g) {
			(void)hipGraphDestroy(g);
		}
		(void)hipGetLastError();
		STP_THROW(STPHIPError, "graph capture of the batch failed");
	}
	hipGraphExec_t exec = nullptr;
	const hipError_t inst = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
	if (inst != hipSuccess) {
		(void)hipGraphDestroy(g);
		(void)hipGetLastError();
		STP_THROW(STPHIPError, std::string("graph instantiation failed: ") + hipGetErrorString(inst));
	}
	graph = g;
	return exec;
}

void STPHeightfieldBatch::run(const STPRunMode mode, const hipStream_t stream) {
	const unsigned int n = static_cast<unsigned int>(this->Record.size());
	this->LastTimed = false;
	switch (mode) {
	case STPRunMode::Graph:
		if (this->Timing) {
			// one graph per stage; the events are recorded on the stream between the graph launches
			// (an event captured inside a graph cannot be queried afterwards)
			if (!this->StageExec[0]) {
				for (int s = 0; s < StageCount; s++) {
					this->StageExec[s] = this->capture(s, s + 1, this->StageGraph[s]);
				}
			}
			for (int s = 0; s < StageCount; s++) {
				STP_CHECK_HIP(hipEventRecord(this->Event[s], stream));
				STP_CHECK_HIP(hipGraphLaunch(this->StageExec[s], stream));
			}
			STP_CHECK_HIP(hipEventRecord(this->Event[StageCount], stream));
			this->LastTimed = true;
		} else {
			if (!this->GraphExec) {
				this->GraphExec = this->capture(0, StageCount, this->Graph);
			}
			STP_CHECK_HIP(hipGraphLaunch(this->GraphExec, stream));
		}
		break;
	case STPRunMode::Eager:
		this->enqueue(stream, 0u, n, this->Timing, true);
		this->LastTimed = this->Timing;
		break;
	case STPRunMode::Plain: {
		this->Latency.assign(n, 0.0);
		zeroKERNEL<<<1, 64, 0, stream>>>(this->Flag.get(), 4u);
		STP_CHECK_HIP(hipGetLastError());
		for (unsigned int i = 0u; i < n; i++) {
			const auto t0 = std::chrono::steady_clock::now();
			this->enqueue(stream, i, i + 1u, false, false);
			STP_CHECK_HIP(hipStreamSynchronize(stream));
			this->Latency[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		}
		break;
	}
	default:
		STP_THROW(STPInvalidArgument, "unknown batch run mode");
	}
}

void STPHeightfieldBatch::runStages(const int first, const int last, const hipStream_t stream) {
	if (first < 0 || last > StageCount || first >= last) {
		STP_THROW(STPInvalidArgument, "stage range outside [0, 7)");
	}
	this->LastTimed = false;
	for (int s = first; s < last; s++) {
		this->enqueueStage(stream, s, 0u, static_cast<unsigned int>(this->Record.size()), true);
	}
}

void STPHeightfieldBatch::setTiming(const bool enable) {
	this->Timing = enable;
}

std::vector<float> STPHeightfieldBatch::stageTimes() const {
	if (!this->LastTimed) {
		STP_THROW(STPInvalidArgument, "the last run was not timed");
	}
	std::vector<float> ms(StageCount);
	for (int k = 0; k < StageCount; k++) {
		STP_CHECK_HIP(hipEventElapsedTime(&ms[k], this->Event[k], this->Event[k + 1]));
	}
	return ms;
}

std::array<unsigned long long, 4> STPHeightfieldBatch::counters(const hipStream_t stream) {
	zero64KERNEL<<<1, 64, 0, stream>>>(this->Counter.get());
	counterKERNEL<<<512, 256, 0, stream>>>(this->DropInfo.get(), this->Drops, this->Counter.get());
	STP_CHECK_HIP(hipGetLastError());
	std::array<unsigned long long, 4> out {};
	STP_CHECK_HIP(hipMemcpyAsync(out.data(), this->Counter.get(), sizeof(out), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipStreamSynchronize(stream));
	return out;
}

unsigned long long STPHeightfieldBatch::binTotal(const hipStream_t stream) {
	zero64KERNEL<<<1, 64, 0, stream>>>(this->Counter.get());
	if (!this->Record.empty()) {
		binTotalKERNEL<<<dim3(16u, this->chunkCount()), 256, 0, stream>>>(this->ChunkRecord, this->BinCount.get(),
			this->Counter.get() + 4);
	}
	STP_CHECK_HIP(hipGetLastError());
	unsigned long long out = 0ull;
	STP_CHECK_HIP(hipMemcpyAsync(&out, this->Counter.get() + 4, sizeof(out), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipStreamSynchronize(stream));
	return out;
}

void STPHeightfieldBatch::check(const hipStream_t stream) {
	const unsigned int flag = this->status(stream);
	if (flag & 1u) {
		STP_THROW(STPCapacityError, "a pixel histogram holds more distinct biomes than the bin capacity");
	}
	if (flag & 2u) {
		STP_THROW(STPCapacityError, "an eroded height left the range of the LDS-tiled erosion window ([-4, 7.5]); "
			"erode these chunks with the global-atomic kernel (erosion tile 0)");
	}
	if (flag & 4u) {
		STP_THROW(STPInvalidArgument, "an external biome sample is not below the biome count (it was read as biome 0)");
	}
}

unsigned int STPHeightfieldBatch::status(const hipStream_t stream) {
	unsigned int flag[3] {};
	STP_CHECK_HIP(hipMemcpyAsync(flag, this->Flag.get(), sizeof(flag), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipStreamSynchronize(stream));
	return (flag[0] != 0u ? 1u : 0u) | (flag[1] != 0u ? 2u : 0u) | (flag[2] != 0u ? 4u : 0u);
}

void STPHeightfieldBatch::readChunk(const unsigned int i, float* const height, unsigned short* const fixed,
	unsigned short* const halo, float* const ledger) const {
	if (i >= this->Record.size()) {
		STP_THROW(STPInvalidArgument, "chunk index out of range");
	}
	const Chunk& c = this->Record[i];
	const size_t px = static_cast<size_t>(c.Side) * c.Side;
	if (height) {
		STP_CHECK_HIP(hipMemcpy(height, this->Height.get() + c.Pixel, px * sizeof(float), hipMemcpyDeviceToHost));
	}
	if (fixed) {
		STP_CHECK_HIP(hipMemcpy(fixed, this->Fixed.get() + c.Pixel, px * sizeof(unsigned short), hipMemcpyDeviceToHost));
	}
	if (halo) {
		const size_t w = c.Side + 2u * this->Setting.Radius;
		STP_CHECK_HIP(hipMemcpy(halo, this->BiomeMap.get() + c.Halo, w * w * sizeof(unsigned short), hipMemcpyDeviceToHost));
	}
	if (ledger && c.Drops) {
		STP_CHECK_HIP(hipMemcpy(ledger, this->Ledger.get() + c.DropBase, c.Drops * sizeof(float), hipMemcpyDeviceToHost));
	}
}

unsigned long long STPHeightfieldBatch::pixels() const noexcept {
	unsigned long long p = 0ull;
	for (const Chunk& c : this->Record) {
		p += static_cast<unsigned long long>(c.Side) * c.Side;
	}
	return p;
}
