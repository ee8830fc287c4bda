// Device side of the biome layers, hipRTC-safe: the layer RNG (SuperTerrain+/SuperTerrain+/Private/World/Diversity/STPLayer.cpp:111-178),
// the named biomes of the demo registry, and how one layer of a chain is evaluated over a plan's
// rectangles. The library's layer kernels and runtime-compiled custom layers share it, so a custom
// layer sees exactly what the built-in ones see.
//
// A custom layer (the reference's STPLayer subclass, SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Diversity/STPLayer.h:153-165:
// `virtual STPSample_t sample(int x, int y, int z)`) is a device function
//   __device__ unsigned short sample(const SuperTerrainPlus::STPLayerDevice::STPLayerContext& c, int x, int z);
// reading its parents through c.parent(k, x, z) (the reference's getAscendant(k).retrieve(x, 0, z)),
// with c.sampler(x, z) (createLocalSampler), c.seedLocal(x, z) and c.registry() at hand.
#pragma once
#ifndef STP_LAYER_DEVICE_HPP
#define STP_LAYER_DEVICE_HPP

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace SuperTerrainPlus {

	namespace STPDiversity {
		// The demo's named biomes in the order the registry registers them
		// (SuperDemo+/World/Biomes/STPBiomeRegistry.cpp:66-104); the first biome registered under an id
		// decides that id's precipitation type.
		enum class STPNamedBiome : unsigned int {
			Ocean, DeepOcean, WarmOcean, LukewarmOcean, ColdOcean, FrozenOcean, DeepWarmOcean, DeepLukewarmOcean,
			DeepColdOcean, DeepFrozenOcean, River, FrozenRiver, Plains, Desert, Mountain, Forest, Taiga, SnowyTaiga,
			SnowyTundra, Jungle, Savannah, Swamp, Badlands, DesertHills, TaigaHills, WoodedMountain, SnowyMountain,
			ForestHills, JungleHills, SavannahPlateau, SwampHills, BadlandsPlateau, Beach, SnowyBeach, StoneShore,
			Count
		};
		constexpr unsigned int NamedBiomeCount = static_cast<unsigned int>(STPNamedBiome::Count);
	} // namespace STPDiversity

	namespace STPLayerDevice {

		using Seed = unsigned long long;
		using Sample = unsigned short;
		using STPDiversity::STPNamedBiome;

		// STPLayer::mixSeed (STPLayer.cpp:174-178): s <- s (s A + C) + fac over 64-bit words.
		__host__ __device__ inline Seed mixSeed(Seed s, const long long fac) {
			s *= s * 6364136223846793005ull + 1442695040888963407ull;
			s += static_cast<Seed>(fac);
			return s;
		}

		// STPLayer::seedLayer (STPLayer.cpp:143-151).
		__host__ __device__ inline Seed seedLayer(const Seed global, const Seed salt) {
			Seed mid = mixSeed(salt, static_cast<long long>(salt));
			mid = mixSeed(mid, static_cast<long long>(mid));
			mid = mixSeed(mid, static_cast<long long>(mid));
			Seed s = mixSeed(global, static_cast<long long>(mid));
			s = mixSeed(s, static_cast<long long>(mid));
			return mixSeed(s, static_cast<long long>(mid));
		}

		// STPLayer::seedLocal (STPLayer.cpp:153-159).
		__host__ __device__ inline Seed seedLocal(const Seed layer, const int x, const int z) {
			Seed s = mixSeed(layer, x);
			s = mixSeed(s, z);
			s = mixSeed(s, x);
			return mixSeed(s, z);
		}

		// STPLayer::STPLocalSampler (STPLayer.cpp:107-132): a value is (state >> 24) mod range, then the
		// state advances by mixSeed(state, layer seed). next<R>() takes a compile-time range (the
		// 64-bit remainder becomes a multiply); nextValue(range) any range.
		struct Sampler {
			Seed Layer, Local;

			template <unsigned int R>
			__device__ inline unsigned int next() {
				const unsigned int v = static_cast<unsigned int>((this->Local >> 24u) % static_cast<Seed>(R));
				this->Local = mixSeed(this->Local, static_cast<long long>(this->Layer));
				return v;
			}
			__device__ inline unsigned int nextValue(const Sample range) {
				const unsigned int v = static_cast<unsigned int>((this->Local >> 24u) % static_cast<Seed>(range));
				this->Local = mixSeed(this->Local, static_cast<long long>(this->Layer));
				return v;
			}
			__device__ inline Sample choose(const Sample a, const Sample b) {
				return this->next<2u>() == 0u ? a : b;
			}
			__device__ inline Sample choose(const Sample a, const Sample b, const Sample c, const Sample d) {
				const unsigned int i = this->next<4u>();
				return i == 0u ? a : i == 1u ? b : i == 2u ? c : d;
			}
		};

		__device__ inline Sampler sampler(const Seed layer, const int x, const int z) {
			return Sampler {layer, seedLocal(layer, x, z)};
		}

		// The biome ids the layers compare against, and each id's precipitation type (0 none, 1 rain,
		// 2 snow; STPBiomeRegistry.cpp:125-137) in a 65536-entry table.
		struct Registry {
			Sample Id[STPDiversity::NamedBiomeCount];
			const unsigned char* Precipitation;

			__device__ inline Sample operator[](const STPNamedBiome b) const {
				return this->Id[static_cast<unsigned int>(b)];
			}
			// STPBiomeRegistry::isShallowOcean / isOcean (STPBiomeRegistry.cpp:109-119)
			__device__ inline bool shallowOcean(const Sample v) const {
				const Registry& r = *this;
				return v == r[STPNamedBiome::Ocean] || v == r[STPNamedBiome::FrozenOcean] || v == r[STPNamedBiome::WarmOcean]
					|| v == r[STPNamedBiome::LukewarmOcean] || v == r[STPNamedBiome::ColdOcean];
			}
			__device__ inline bool ocean(const Sample v) const {
				const Registry& r = *this;
				return this->shallowOcean(v) || v == r[STPNamedBiome::DeepOcean] || v == r[STPNamedBiome::DeepWarmOcean]
					|| v == r[STPNamedBiome::DeepLukewarmOcean] || v == r[STPNamedBiome::DeepColdOcean]
					|| v == r[STPNamedBiome::DeepFrozenOcean];
			}
			__device__ inline bool snowy(const Sample v) const {
				return this->Precipitation[v] == 2u;
			}
		};

		// Where map c's samples of a node lie: sample (i, j) is the layer at (X0 + i Step, Z0 + j Step),
		// stored at base[Offset + j Pitch + i]; a node's rectangles are Rect[node * Count + c].
		struct LayerRect {
			int X0, Z0;
			unsigned int W, H, Pitch, Step;
			unsigned long long Offset;
		};
		static_assert(sizeof(LayerRect) == 32u, "rectangle layout");

		constexpr unsigned int TileX = 64u, TileY = 4u;

		// One node's launch over a range of a plan's tiles.
		struct NodeArgs {
			unsigned int Kind, Param, Count, Node, Limit, Root;
			Seed LayerSeed, VoronoiSeed;
			Sample* Out;       // the caller's maps for the root, the plan's buffer otherwise
			const Sample* In;  // the plan's buffer (parents)
			int Parent[2];
			const LayerRect* Rect;
			const uint2* Tile; // (map, tile) of this launch
			Registry Reg;
			unsigned int* Fault;
		};

		// Parent reads of map c, checked against the parent's rectangle: a read outside it sets
		// status bit 0 and gives 0.
		struct Reader {
			const Sample* Data;
			int X0, Z0;
			unsigned int W, H, Pitch;
			unsigned int* Fault;

			__device__ Reader(const NodeArgs& a, const unsigned int k, const unsigned int c) {
				const int p = a.Parent[k];
				const LayerRect r = p >= 0 ? a.Rect[static_cast<unsigned long long>(p) * a.Count + c] : LayerRect {};
				this->Data = a.In + r.Offset;
				this->X0 = r.X0;
				this->Z0 = r.Z0;
				this->W = p >= 0 ? r.W : 0u;
				this->H = p >= 0 ? r.H : 0u;
				this->Pitch = r.Pitch;
				this->Fault = a.Fault;
			}
			__device__ inline Sample operator()(const int x, const int z) const {
				const unsigned int u = static_cast<unsigned int>(x - this->X0), v = static_cast<unsigned int>(z - this->Z0);
				if (u >= this->W || v >= this->H) {
					atomicOr(this->Fault, 1u);
					return 0u;
				}
				return this->Data[static_cast<unsigned long long>(v) * this->Pitch + u];
			}
		};

		// the root's ids index the caller's biome table: one out of range is flagged (status bit 1)
		// and written as 0
		__device__ inline Sample limited(const NodeArgs& a, const Sample v) {
			if (a.Root && v >= a.Limit) {
				atomicOr(a.Fault, 2u);
				return 0u;
			}
			return v;
		}

		// What a custom layer's sample function sees.
		struct STPLayerContext {
			const NodeArgs& Args;
			const Reader& First;
			const Reader& Second;

			// getAscendant(k).retrieve(x, 0, z), k = 0 or 1
			__device__ inline Sample parent(const unsigned int k, const int x, const int z) const {
				return k == 0u ? this->First(x, z) : this->Second(x, z);
			}
			// createLocalSampler(x, z)
			__device__ inline Sampler sampler(const int x, const int z) const {
				return STPLayerDevice::sampler(this->Args.LayerSeed, x, z);
			}
			// createLocalSampler(local_seed)
			__device__ inline Sampler sampler(const Seed local) const {
				return Sampler {this->Args.LayerSeed, local};
			}
			__device__ inline Seed seedLocal(const int x, const int z) const {
				return STPLayerDevice::seedLocal(this->Args.LayerSeed, x, z);
			}
			__device__ inline Seed layerSeed() const {
				return this->Args.LayerSeed;
			}
			__device__ inline const Registry& registry() const {
				return this->Args.Reg;
			}
		};

		// One TileX x TileY tile of a node for one map: every sample in the node's rectangle of the
		// map is fn(a, first parent, second parent, x, z) (y = 0), stored through `limited`.
		template <class Fn>
		__device__ inline void runNode(const NodeArgs& a, Fn&& fn) {
			const uint2 t = a.Tile[blockIdx.x];
			const LayerRect rect = a.Rect[static_cast<unsigned long long>(a.Node) * a.Count + t.x];
			const unsigned int tilesX = (rect.W + TileX - 1u) / TileX;
			const unsigned int lx = (t.y % tilesX) * TileX + threadIdx.x, lz = (t.y / tilesX) * TileY + threadIdx.y;
			if (lx >= rect.W || lz >= rect.H) {
				return;
			}
			const Reader p(a, 0u, t.x), q(a, 1u, t.x);
			const Sample v = fn(a, p, q, rect.X0 + static_cast<int>(lx * rect.Step), rect.Z0 + static_cast<int>(lz * rect.Step));
			a.Out[rect.Offset + static_cast<unsigned long long>(lz) * rect.Pitch + lx] = limited(a, v);
		}

	} // namespace STPLayerDevice

} // namespace SuperTerrainPlus

#endif // STP_LAYER_DEVICE_HPP
