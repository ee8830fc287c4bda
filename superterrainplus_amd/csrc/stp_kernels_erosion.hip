// Free-slip particle hydraulic erosion and FP32 -> UINT16 formatting (per-chunk launches).
//
// Behaviour follows SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:80-122
// and SuperTerrain+/SuperTerrain+/Private/GPGPU/STPRainDrop.cu:18-164 (drop logic in
// stp_stage_erosion.hpp).
//
// MI355X design:
//  * Parallel mode: one lane per drop. Heights are read with relaxed agent-scope atomic loads,
//    which bypass the per-CU vector L1, so a drop sees other CUs' updates as soon as L2 has
//    them (the reference's plain loads can hit a stale L1 line). Updates are hardware FP32
//    atomic adds executed at L2 (no CAS loop). The brush (index, weight) is staged in LDS and
//    read with a wave-uniform address, i.e. an LDS broadcast; the heights under the brush are
//    fetched in one group before the atomics are issued.
//  * Serialized mode: one wavefront walks the drops in index order; all lanes carry the same
//    drop state, the brush taps are spread over the lanes, and the sediment sum is taken in tap
//    order with lane reads, so the result is bit-exact against the FP32 oracle.
//  Roofline: L2 atomics + dependent-latency chains; per drop-step the algorithmic traffic is
//  32 B of height reads, plus 25 x (4 B read + 8 B atomic) for an erode step at r = 3, or
//  4 x 8 B atomics for a deposit step.
#include "stp_stage_erosion.hpp"

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPStage;

namespace {

	constexpr unsigned int ParallelBlock = 256u;

	__global__ void __launch_bounds__(ParallelBlock) erosionParallelKERNEL(float* const map, const ErosionArgs a,
		const int* const brush_index, const float* const brush_weight, float* const ledger, uint2* const drop_info,
		const unsigned int* const gate) {
		extern __shared__ int2 brush[]; // (index, weight bits)
		if (gate && (gate[0] | gate[1]) == 0u) {
			// another path eroded the map: the whole grid leaves at once
			return;
		}
		for (unsigned int k = threadIdx.x; k < a.BrushSize; k += blockDim.x) {
			brush[k] = make_int2(brush_index[k], __float_as_int(brush_weight[k]));
		}
		__syncthreads();
		const unsigned int drop = blockIdx.x * blockDim.x + threadIdx.x;
		if (drop < a.Rain.RainDropCount) {
			erodeDropAnyBrush(GlobalSurface {map, a.Total.x}, a, brush, drop, drop, ledger, drop_info);
		}
	}

	constexpr unsigned int SerialLanes = 64u;

	__global__ void __launch_bounds__(SerialLanes) erosionSerialKERNEL(float* const map, const ErosionArgs a,
		const int* const brush_index, const float* const brush_weight, float* const ledger, uint2* const drop_info) {
		const unsigned int lane = threadIdx.x;
		const unsigned int W = a.Total.x;
		for (unsigned int drop = 0u; drop < a.Rain.RainDropCount; drop++) {
			DropState d = spawn(a, drop);
			float lost = 0.0f;
			unsigned int steps = 0u, erodeSteps = 0u, depositSteps = 0u, killed = 0u;
			if (!spawnInside(a, d)) {
				d.volume = 0.0f;
				killed = 1u;
			}
			for (unsigned int life = 0u; alive(a, d, life); life++) {
				steps++;
				const unsigned int cellX = static_cast<unsigned int>(d.px), cellY = static_cast<unsigned int>(d.py);
				const unsigned int mapIndex = cellY * W + cellX;
				const float ox = d.px - static_cast<float>(static_cast<int>(d.px)),
					oy = d.py - static_cast<float>(static_cast<int>(d.py));
				float3 hg;
				if (!moveDrop<false>(a, map, d, hg) || !cellInside(a, cellX, cellY)) {
					lost = d.sediment;
					d.volume = 0.0f;
					d.sediment = 0.0f;
					killed = 1u;
					break;
				}
				const float deltaHeight = heightGradient<false>(map, W, d.px, d.py).x - hg.x;
				const float cap = fmaxf(-deltaHeight * d.speed * d.volume * a.Rain.SedimentCapacityFactor,
					a.Rain.minSedimentCapacity);
				if (d.sediment > cap || deltaHeight > 0.0f) {
					const float amount = (deltaHeight > 0.0f) ? fminf(deltaHeight, d.sediment)
															  : (d.sediment - cap) * a.Rain.DepositSpeed;
					d.sediment -= amount;
					// four distinct cells, one lane each
					if (lane < 4u) {
						const unsigned int cell = mapIndex + (lane & 1u) + ((lane >> 1) ? W : 0u);
						const float fx = (lane & 1u) ? ox : (1.0f - ox), fy = (lane >> 1) ? oy : (1.0f - oy);
						map[cell] = map[cell] + amount * fx * fy;
					}
					depositSteps++;
				} else {
					const float amount = fminf((cap - d.sediment) * a.Rain.ErodeSpeed, -deltaHeight);
					for (unsigned int base = 0u; base < a.BrushSize; base += SerialLanes) {
						const unsigned int k = base + lane;
						float ds = 0.0f;
						if (k < a.BrushSize) {
							const unsigned int ei = mapIndex + static_cast<unsigned int>(brush_index[k]);
							const float wa = amount * brush_weight[k];
							const float h = map[ei];
							ds = (h < wa) ? h : wa;
							map[ei] = h - ds;
						}
						const unsigned int group = min(SerialLanes, a.BrushSize - base);
						for (unsigned int j = 0u; j < group; j++) {
							d.sediment += __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ds), j));
						}
					}
					erodeSteps++;
				}
				// make this step's writes visible to every lane before the next read
				__syncthreads();
				advance(a, d, deltaHeight);
			}
			if (lane == 0u) {
				if (ledger) {
					ledger[drop] = lost + d.sediment;
				}
				if (drop_info) {
					drop_info[drop] = packDropInfo(steps, erodeSteps, depositSteps, killed);
				}
			}
		}
	}

	__global__ void __launch_bounds__(256) formatKERNEL(const float* const input, STPHeightFixed_t* const output,
		const unsigned long long count) {
		const unsigned long long i4 = (static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x) * 4ull;
		if (i4 + 4ull <= count) {
			const float4 h = *reinterpret_cast<const float4*>(input + i4);
			ushort4 o;
			o.x = toFixed(h.x);
			o.y = toFixed(h.y);
			o.z = toFixed(h.z);
			o.w = toFixed(h.w);
			*reinterpret_cast<ushort4*>(output + i4) = o;
		} else {
			for (unsigned long long i = i4; i < count; i++) {
				output[i] = toFixed(input[i]);
			}
		}
	}

	// gate words: [0] a height out of range, [1] a status word copied from other work
	__global__ void gateClearKERNEL(unsigned int* const gate) {
		gate[0] = 0u;
		gate[1] = 0u;
	}

	constexpr unsigned int GateBlock = 256u, GateItems = 16u;

	__global__ void __launch_bounds__(GateBlock) heightGateKERNEL(const float* const map, const unsigned long long count,
		const float low, const float high, unsigned int* const gate) {
		const unsigned long long first = static_cast<unsigned long long>(blockIdx.x) * GateBlock * GateItems + threadIdx.x;
		bool out = false;
#pragma unroll
		for (unsigned int k = 0u; k < GateItems; k++) {
			const unsigned long long i = first + static_cast<unsigned long long>(k) * GateBlock;
			if (i < count) {
				const float h = map[i];
				// NaN fails both comparisons
				out = out || !(h >= low && h <= high);
			}
		}
		if (__any(out) && (threadIdx.x & 63u) == 0u) {
			atomicOr(gate, 1u);
		}
	}

	__global__ void gateStatusKERNEL(const unsigned int* const status, unsigned int* const gate) { gate[1] = *status; }

	__global__ void __launch_bounds__(256) selectUngatedKERNEL(const unsigned int* const gate, const float* const source,
		float* const map, const unsigned long long count) {
		if ((gate[0] | gate[1]) != 0u) {
			return;
		}
		const unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
		if (i < count) {
			map[i] = source[i];
		}
	}

} // namespace

void STPHeightfieldKernel::heightGate(const float* const map, const unsigned long long count, const float low,
	const float high, unsigned int* const gate, const hipStream_t stream) {
	if (!map || !gate) {
		STP_THROW(STPInvalidArgument, "height gate needs a map and a gate");
	}
	gateClearKERNEL<<<1, 1, 0, stream>>>(gate);
	const unsigned long long per = static_cast<unsigned long long>(GateBlock) * GateItems;
	if (count != 0ull) {
		heightGateKERNEL<<<static_cast<unsigned int>((count + per - 1ull) / per), GateBlock, 0, stream>>>(map, count, low, high,
			gate);
	}
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::copyGateStatus(const unsigned int* const status, unsigned int* const gate, const hipStream_t stream) {
	if (!status || !gate) {
		STP_THROW(STPInvalidArgument, "gate status needs a status word and a gate");
	}
	gateStatusKERNEL<<<1, 1, 0, stream>>>(status, gate);
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::selectUngated(const unsigned int* const gate, const float* const source, float* const map,
	const unsigned long long count, const hipStream_t stream) {
	if (!gate || !source || !map) {
		STP_THROW(STPInvalidArgument, "select needs a gate, a source and a map");
	}
	if (count != 0ull) {
		selectUngatedKERNEL<<<static_cast<unsigned int>((count + 255ull) / 256ull), 256, 0, stream>>>(gate, source, map, count);
	}
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::hydraulicErosion(float* const height, const STPNearestNeighbourInformation& nn,
	const STPRainDropParameter& rd, const int* const brush_index, const float* const brush_weight,
	const unsigned int brush_size, const unsigned int stream_id, const STPErosionMode mode, float* const ledger,
	uint2* const drop_info, const hipStream_t stream, const unsigned int* const gate) {
	if (!height || !brush_index || !brush_weight || brush_size == 0u) {
		STP_THROW(STPInvalidArgument, "erosion needs a heightmap and a non-empty brush");
	}
	if (nn.TotalMapSize.x != nn.MapSize.x * nn.ChunkNearestNeighbour.x
		|| nn.TotalMapSize.y != nn.MapSize.y * nn.ChunkNearestNeighbour.y || nn.MapSize.x < 2u || nn.MapSize.y < 2u) {
		STP_THROW(STPInvalidArgument, "inconsistent nearest-neighbour information");
	}
	const unsigned int r = rd.ErosionBrushRadius;
	if (r == 0u || 2u * r + 2u > nn.TotalMapSize.x || 2u * r + 2u > nn.TotalMapSize.y) {
		STP_THROW(STPInvalidArgument, "brush radius does not fit the merged map");
	}
	ErosionArgs a;
	a.Rain = rd;
	a.MapSize = nn.MapSize;
	a.SpawnOrigin = make_uint2((nn.ChunkNearestNeighbour.x / 2u) * nn.MapSize.x, (nn.ChunkNearestNeighbour.y / 2u) * nn.MapSize.y);
	a.Total = nn.TotalMapSize;
	a.BrushSize = brush_size;
	a.StreamId = stream_id;
	if (rd.RainDropCount == 0u) {
		return;
	}
	if (mode == STPErosionMode::Parallel) {
		const unsigned int grid = (rd.RainDropCount + ParallelBlock - 1u) / ParallelBlock;
		erosionParallelKERNEL<<<grid, ParallelBlock, brush_size * sizeof(int2), stream>>>(height, a, brush_index,
			brush_weight, ledger, drop_info, gate);
	} else {
		if (gate) {
			STP_THROW(STPInvalidArgument, "a gated erosion is parallel");
		}
		erosionSerialKERNEL<<<1, SerialLanes, 0, stream>>>(height, a, brush_index, brush_weight, ledger, drop_info);
	}
	STP_CHECK_HIP(hipGetLastError());
}

void STPHeightfieldKernel::formatHeightmap(const float* const input, STPHeightFixed_t* const output,
	const unsigned long long count, const hipStream_t stream) {
	if (count == 0ull) {
		return;
	}
	if (!input || !output) {
		STP_THROW(STPInvalidArgument, "format needs input and output");
	}
	if ((reinterpret_cast<uintptr_t>(input) & 15u) != 0u || (reinterpret_cast<uintptr_t>(output) & 7u) != 0u) {
		STP_THROW(STPInvalidArgument, "format needs 16-byte aligned input and 8-byte aligned output");
	}
	const unsigned long long threads = (count + 3ull) / 4ull;
	formatKERNEL<<<static_cast<unsigned int>((threads + 255ull) / 256ull), 256u, 0, stream>>>(input, output, count);
	STP_CHECK_HIP(hipGetLastError());
}
