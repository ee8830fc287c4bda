// Free-slip particle hydraulic erosion and FP32 -> UINT16 formatting.
//
// Behaviour follows SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:80-122
// and SuperTerrain+/SuperTerrain+/Private/GPGPU/STPRainDrop.cu:18-164, with two additions:
// MaxLifetime (0 = unlimited, the reference) and a stateless Philox spawn.
//
// MI355X design:
//  * Parallel mode: one lane per drop. Heights are read with relaxed agent-scope atomic loads,
//    which bypass the per-CU vector L1, so a drop sees other CUs' updates as soon as L2 has
//    them (the reference's plain loads can hit a stale L1 line). Updates are hardware FP32
//    atomic adds executed at L2 (no CAS loop). The brush (index, weight) is staged in LDS and
//    read with a wave-uniform address, i.e. an LDS broadcast.
//  * Serialized mode: one wavefront walks the drops in index order; all lanes carry the same
//    drop state, the brush taps are spread over the lanes, and the sediment sum is taken in tap
//    order with lane reads, so the result is bit-exact against the FP32 oracle.
//  Roofline: L2 atomics + dependent-latency chains; per drop-step the algorithmic traffic is
//  32 B of height reads, plus 25 x (4 B read + 4 B atomic) for an erode step at r = 3, or
//  4 x 4 B atomics for a deposit step.
#include "stp_internal.hpp"
#include <stp/STPRandom.hpp>

using namespace SuperTerrainPlus;

namespace {

	struct ErosionArgs {
		STPRainDropParameter Rain;
		uint2 MapSize;
		uint2 NN;
		uint2 Total;
		unsigned int BrushSize;
		unsigned int StreamId;
	};

	__device__ inline float loadHeight(const float* p) {
		return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}

	template <bool Atomic>
	__device__ inline float readH(const float* map, const unsigned int i) {
		if constexpr (Atomic) {
			return loadHeight(map + i);
		} else {
			return map[i];
		}
	}

	// STPRainDrop.cu:23-51 -> (height, gradient x, gradient y); glm::mix(a, b, t) = a*(1-t) + b*t
	template <bool Atomic>
	__device__ inline float3 heightGradient(const float* map, const unsigned int W, const float px, const float py) {
		const unsigned int rx = static_cast<unsigned int>(px), ry = static_cast<unsigned int>(py);
		const float cx = px - static_cast<float>(rx), cy = py - static_cast<float>(ry);
		const unsigned int b = ry * W + rx;
		const float hx = readH<Atomic>(map, b), hy = readH<Atomic>(map, b + 1u), hz = readH<Atomic>(map, b + W),
			hw = readH<Atomic>(map, b + W + 1u);
		float3 r;
		r.x = hx * (1 - cx) * (1 - cy) + hy * cx * (1 - cy) + hz * (1 - cx) * cy + hw * cx * cy;
		r.y = (hy - hx) * (1 - cy) + (hw - hz) * cy;
		r.z = (hz - hx) * (1 - cx) + (hw - hy) * cx;
		return r;
	}

	struct DropState {
		float px, py, dx, dy, speed, volume, sediment;
	};

	__device__ inline DropState spawn(const ErosionArgs& a, const unsigned int drop) {
		// STPHeightfieldKernel.cu:93-102: centre chunk only, u * (Map - 1) + (NN / 2) * Map
		const STPRandom::Philox4 r = STPRandom::philox4x32_10(drop, a.StreamId, 0u, 0u,
			static_cast<unsigned int>(a.Rain.Seed), static_cast<unsigned int>(a.Rain.Seed >> 32));
		DropState d;
		d.px = STPRandom::unitUniform(r.v[0]) * static_cast<float>(a.MapSize.x - 1u)
			+ static_cast<float>((a.NN.x / 2u) * a.MapSize.x);
		d.py = STPRandom::unitUniform(r.v[1]) * static_cast<float>(a.MapSize.y - 1u)
			+ static_cast<float>((a.NN.y / 2u) * a.MapSize.y);
		d.dx = 0.0f;
		d.dy = 0.0f;
		d.speed = a.Rain.initSpeed;
		d.volume = a.Rain.initWaterVolume;
		d.sediment = 0.0f;
		return d;
	}

	// Moves the drop one cell; returns false when it dies (STPRainDrop.cu:78-105).
	template <bool Atomic>
	__device__ inline bool moveDrop(const ErosionArgs& a, const float* map, DropState& d, float3& hg) {
		const unsigned int W = a.Total.x;
		hg = heightGradient<Atomic>(map, W, d.px, d.py);
		const float inertia = a.Rain.Inertia;
		d.dx = (-hg.y) * (1 - inertia) + d.dx * inertia;
		d.dy = (-hg.z) * (1 - inertia) + d.dy * inertia;
		const float length = sqrtf(d.dx * d.dx + d.dy * d.dy);
		if (length != 0.0f) {
			d.dx /= length;
			d.dy /= length;
		}
		d.px += d.dx;
		d.py += d.dy;
		const float radius = static_cast<float>(a.Rain.ErosionBrushRadius);
		// STPRainDrop.cu:96-105, written so that a NaN position also dies
		return !(d.dx == 0.0f && d.dy == 0.0f) && d.px >= radius && d.px < 1.0f * static_cast<float>(a.Total.x) - radius
			&& d.py >= radius && d.py < 1.0f * static_cast<float>(a.Total.y) - radius;
	}

	// Footprint guard (no reference counterpart): the cell a drop erodes/deposits at must keep
	// its brush (reach r-1) and its 2x2 bilinear cell (reach +1) inside the map. In exact
	// arithmetic the kill test already implies this; it only fails when a unit direction
	// rounds above 1 in magnitude. Spawns whose 2x2 cell leaves the map (u == 1 on a 1x1
	// neighbourhood, which the reference reads out of bounds for) die before their first step.
	__device__ inline bool cellInside(const ErosionArgs& a, const unsigned int cx, const unsigned int cy) {
		const unsigned int r = a.Rain.ErosionBrushRadius, reach = r > 2u ? r - 1u : 1u;
		return cx + 1u >= r && cy + 1u >= r && cx + reach <= a.Total.x - 1u && cy + reach <= a.Total.y - 1u;
	}

	__device__ inline bool spawnInside(const ErosionArgs& a, const DropState& d) {
		return static_cast<unsigned int>(d.px) + 1u <= a.Total.x - 1u && static_cast<unsigned int>(d.py) + 1u <= a.Total.y - 1u;
	}

	__device__ inline void advance(const ErosionArgs& a, DropState& d, const float deltaHeight) {
		d.speed = sqrtf(fmaxf(0.0f, d.speed * d.speed + deltaHeight * a.Rain.Gravity));
		d.speed *= 1.0f - a.Rain.Friction;
		d.volume *= (1.0f - a.Rain.EvaporateSpeed);
	}

	__device__ inline bool alive(const ErosionArgs& a, const DropState& d, const unsigned int life) {
		return (a.Rain.MaxLifetime == 0u || life < a.Rain.MaxLifetime) && d.volume >= a.Rain.minWaterVolume;
	}

	constexpr unsigned int ParallelBlock = 256u;

	// ------------------------------------------------------------------ parallel mode
	__global__ void __launch_bounds__(ParallelBlock) erosionParallelKERNEL(float* const map, const ErosionArgs a,
		const int* const brush_index, const float* const brush_weight, float* const ledger, uint2* const drop_info) {
		extern __shared__ int2 brush[]; // (index, weight bits)
		for (unsigned int k = threadIdx.x; k < a.BrushSize; k += blockDim.x) {
			brush[k] = make_int2(brush_index[k], __float_as_int(brush_weight[k]));
		}
		__syncthreads();

		const unsigned int drop = blockIdx.x * blockDim.x + threadIdx.x;
		if (drop >= a.Rain.RainDropCount) {
			return;
		}
		DropState d = spawn(a, drop);
		const unsigned int W = a.Total.x;
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
			if (!moveDrop<true>(a, map, d, hg) || !cellInside(a, cellX, cellY)) {
				lost = d.sediment;
				d.volume = 0.0f;
				d.sediment = 0.0f;
				killed = 1u;
				break;
			}
			const float deltaHeight = heightGradient<true>(map, W, d.px, d.py).x - hg.x;
			const float cap = fmaxf(-deltaHeight * d.speed * d.volume * a.Rain.SedimentCapacityFactor,
				a.Rain.minSedimentCapacity);
			if (d.sediment > cap || deltaHeight > 0.0f) {
				const float amount = (deltaHeight > 0.0f) ? fminf(deltaHeight, d.sediment)
														  : (d.sediment - cap) * a.Rain.DepositSpeed;
				d.sediment -= amount;
				unsafeAtomicAdd(map + mapIndex, amount * (1.0f - ox) * (1.0f - oy));
				unsafeAtomicAdd(map + mapIndex + 1u, amount * ox * (1.0f - oy));
				unsafeAtomicAdd(map + mapIndex + W, amount * (1.0f - ox) * oy);
				unsafeAtomicAdd(map + mapIndex + W + 1u, amount * ox * oy);
				depositSteps++;
			} else {
				const float amount = fminf((cap - d.sediment) * a.Rain.ErodeSpeed, -deltaHeight);
				for (unsigned int k = 0u; k < a.BrushSize; k++) {
					const int2 b = brush[k];
					const unsigned int ei = mapIndex + static_cast<unsigned int>(b.x);
					const float wa = amount * __int_as_float(b.y);
					const float h = loadHeight(map + ei);
					const float ds = (h < wa) ? h : wa;
					unsafeAtomicAdd(map + ei, -ds);
					d.sediment += ds;
				}
				erodeSteps++;
			}
			advance(a, d, deltaHeight);
		}
		if (ledger) {
			ledger[drop] = lost + d.sediment;
		}
		if (drop_info) {
			drop_info[drop] = make_uint2(steps | (erodeSteps << 16), depositSteps | (killed << 16));
		}
	}

	// ------------------------------------------------------------------ serialized mode
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
					drop_info[drop] = make_uint2(steps | (erodeSteps << 16), depositSteps | (killed << 16));
				}
			}
		}
	}

	// ------------------------------------------------------------------ format
	// STPHeightfieldKernel.cu:112-122: u16 = rint(saturate(h) * 65535), round half to even.
	__device__ inline unsigned short toFixed(const float h) {
		const float s = h > 1.0f ? 1.0f : (h > 0.0f ? h : 0.0f);
		return static_cast<unsigned short>(rintf(s * 65535.0f));
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

} // namespace

void STPHeightfieldKernel::hydraulicErosion(float* const height, const STPNearestNeighbourInformation& nn,
	const STPRainDropParameter& rd, const int* const brush_index, const float* const brush_weight,
	const unsigned int brush_size, const unsigned int stream_id, const STPErosionMode mode, float* const ledger,
	uint2* const drop_info, const hipStream_t stream) {
	if (!height || !brush_index || !brush_weight || brush_size == 0u) {
		STP_THROW(STPInvalidArgument, "erosion needs a heightmap and a non-empty brush");
	}
	if (nn.TotalMapSize.x != nn.MapSize.x * nn.ChunkNearestNeighbour.x
		|| nn.TotalMapSize.y != nn.MapSize.y * nn.ChunkNearestNeighbour.y || nn.MapSize.x < 2u || nn.MapSize.y < 2u) {
		STP_THROW(STPInvalidArgument, "inconsistent nearest-neighbour information");
	}
	// every in-map position a drop can occupy keeps its bilinear 2x2 footprint and brush inside
	// the map: the kill test removes drops closer than `radius` to the merged edge, so the brush
	// radius must cover the brush footprint and the map must be wider than twice the radius.
	const unsigned int r = rd.ErosionBrushRadius;
	if (r == 0u || 2u * r + 2u > nn.TotalMapSize.x || 2u * r + 2u > nn.TotalMapSize.y) {
		STP_THROW(STPInvalidArgument, "brush radius does not fit the merged map");
	}
	ErosionArgs a;
	a.Rain = rd;
	a.MapSize = nn.MapSize;
	a.NN = nn.ChunkNearestNeighbour;
	a.Total = nn.TotalMapSize;
	a.BrushSize = brush_size;
	a.StreamId = stream_id;
	if (rd.RainDropCount == 0u) {
		return;
	}
	if (mode == STPErosionMode::Parallel) {
		const unsigned int grid = (rd.RainDropCount + ParallelBlock - 1u) / ParallelBlock;
		erosionParallelKERNEL<<<grid, ParallelBlock, brush_size * sizeof(int2), stream>>>(height, a, brush_index,
			brush_weight, ledger, drop_info);
	} else {
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
