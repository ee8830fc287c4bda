// Biome map generation by a chain of layers (the Minecraft-style layered algorithm) on the GPU.
//
// The reference builds the biome map on the host: an STPBiomeFactory
// (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Diversity/STPBiomeFactory.h:16-93,
// SuperTerrain+/SuperTerrain+/Private/World/Diversity/STPBiomeFactory.cpp:24-41) samples the root of a
// tree of STPLayer objects (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Diversity/STPLayer.h,
// SuperTerrain+/SuperTerrain+/Private/World/Diversity/STPLayer.cpp) pixel by pixel, each layer
// recursing into its parents through a hash cache. The demo's layers are SuperDemo+/World/Layers/*.h and
// its chain SuperDemo+/World/Layers/STPAllLayers.cpp:55-110.
//
// Here a chain is a list of nodes of built-in layer kinds (the demo's), parents before children, the
// last node the root. The GPU evaluates it bottom-up: every node once over the rectangle of
// coordinates its children read, for every chunk of a batch in one launch.
#pragma once
#ifndef STP_BIOME_LAYER_HPP
#define STP_BIOME_LAYER_HPP

#include <hip/hip_runtime.h>
#include <stp/STPLayerDevice.hpp>

#include <cstdint>
#include <string>
#include <memory>
#include <vector>

namespace SuperTerrainPlus::STPDiversity {

	// The layer kinds of the demo (SuperDemo+/World/Layers). Parents: 0 for Continent and OceanNoise,
	// 2 for Hills (land, noise), OceanMix (land, ocean) and RiverMix (land, river), 1 otherwise.
	enum class STPLayerKind : unsigned int {
		Continent = 0u,       // STPContinentLayer.h
		Scale = 1u,           // STPScaleLayer.h; Param 0 normal, 1 fuzzy
		Land = 2u,            // STPLandLayer.h (X cross)
		Island = 3u,          // STPIslandLayer.h (cross)
		Voronoi = 4u,         // STPVoronoiLayer.h, 2D only; Param 0 std::hash as libstdc++ (identity), 1 as MSVC (FNV-1a)
		BaseBiome = 5u,       // STPBaseBiomeLayer.h
		ClimateSingle = 6u,   // STPClimateLayer.h
		ClimateModerate = 7u,
		ClimateExtreme = 8u,
		DeepOcean = 9u,       // STPDeepOceanLayer.h
		EaseEdge = 10u,       // STPEaseEdgeLayer.h
		EdgeBiome = 11u,      // STPEdgeBiomeLayer.h
		Hills = 12u,          // STPHillsLayer.h
		Noise = 13u,          // STPNoiseLayer.h
		NoiseToRiver = 14u,   // STPNoiseToRiverLayer.h
		OceanNoise = 15u,     // STPOceanTemperatureLayer.h
		OceanTemperate = 16u,
		OceanExtreme = 17u,
		OceanTransition = 18u,
		OceanMix = 19u,
		RiverErode = 20u,     // STPRiverErodeLayer.h
		RiverMix = 21u,       // STPRiverMixLayer.h
		SmoothScale = 22u,    // STPSmoothScaleLayer.h
		Custom = 23u,         // a runtime-compiled layer (STPCustomLayer); Param = its index
		Count = 24u
	};

	// A layer of the caller's own (an STPLayer subclass, STPLayer.h:153-165), compiled at run time
	// (hipRTC). Source defines
	//   __device__ unsigned short sample(const SuperTerrainPlus::STPLayerDevice::STPLayerContext& c, int x, int z);
	// with the library's device headers (stp/STPLayerDevice.hpp and the others) available. The
	// footprint declares which parent coordinates a sample at (x, z) may read: from ((x + Lo) >> Shift,
	// (z + Lo) >> Shift) to ((x + Hi) >> Shift, (z + Hi) >> Shift), for both parents (Parents = 1 or 2).
	// A read outside it gives 0 and status bit 0. E.g. a cross layer: Lo -1, Hi 1, Shift 0; a zoom:
	// Lo 0, Hi 1, Shift 1.
	struct STPCustomLayer {
		std::string Source;
		int Lo = 0, Hi = 0;
		unsigned int Shift = 0u;
		unsigned int Parents = 1u;
		std::vector<std::string> Options;
	};

	// One layer: its kind, salt (STPLayer::Salt), parents (node indices, -1 for none), the kind's
	// parameter, and the reference's cache size (0 or a power of 2; validated as the reference does,
	// without effect here: the GPU evaluates whole rectangles).
	struct STPLayerNode {
		unsigned int Kind;
		unsigned long long Salt;
		int Parent[2];
		unsigned int Param;
		unsigned long long CacheSize;
	};

	// Ids, temperatures and precipitations of the named biomes (STPBiome::ID/Temperature/Precipitation).
	struct STPBiomeRegistry {
		unsigned short ID[NamedBiomeCount];
		float Temperature[NamedBiomeCount];
		float Precipitation[NamedBiomeCount];

		// The demo's registry after loading SuperDemo+/Biome.ini: ocean 0, plains 1, desert 2, forest 3
		// (STPTerrainParaLoader.cpp:271-294); every other biome is never loaded and keeps the zero
		// of its static storage, so its id is 0 (the ocean's).
		static STPBiomeRegistry demo() noexcept;
	};

	// A list of layers; the last one added is the root.
	class STPLayerProgram {
	public:
		std::vector<STPLayerNode> Node;

		// Appends a node, returns its index. Parents must be added before.
		int add(STPLayerKind kind, unsigned long long salt, int parent0 = -1, int parent1 = -1, unsigned int param = 0u,
			unsigned long long cache_size = 0u);

		// The demo's chain (STPAllLayers.cpp:55-110): Continent, fuzzy Scale, Land, Scale, 3 Land, Island,
		// 3 Scale, 3 Voronoi, with the demo's salts and cache sizes.
		static STPLayerProgram demo();
	};

	// A map the factory writes: sample (i, j) is the root layer at world
	// (Origin.x + i * Step, 0, Origin.y + j * Step), stored at out[Offset + j * Pitch + i].
	struct STPLayerMap {
		int2 Origin;
		uint2 Size;
		unsigned int Step;
		unsigned int Pitch;
		unsigned long long Offset;
	};

	class STPLayerBiomeFactory;

	// The rectangles every layer evaluates for a list of maps, their tiles and buffers, in device
	// memory; rebuilt in place for a new list (its memory grows, never shrinks).
	class STPLayerPlan {
	public:
		STPLayerPlan();
		~STPLayerPlan();
		STPLayerPlan(const STPLayerPlan&) = delete;
		STPLayerPlan& operator=(const STPLayerPlan&) = delete;

		unsigned int count() const noexcept;

		struct STPImplementation;
		std::unique_ptr<STPImplementation> Impl;
	};

	// The biome factory on the GPU (STPBiomeFactory): biome maps of BiomeDimension pixels whose pixel
	// (x, y) is the root layer sampled at world (x + offset.x, 0, y + offset.y).
	class STPLayerBiomeFactory {
	public:
		const uint2 BiomeDimension;

		// Throws STPNumericDomainError for an empty dimension or a cache size that is not a power of 2,
		// STPInvalidArgument for a malformed chain, STPCompilationError for a custom layer that does not
		// compile (with the compiler log).
		STPLayerBiomeFactory(const STPLayerProgram& program, unsigned long long global_seed, const STPBiomeRegistry& registry,
			uint2 dimension, const std::vector<STPCustomLayer>& custom = {});
		~STPLayerBiomeFactory();

		STPLayerBiomeFactory(const STPLayerBiomeFactory&) = delete;
		STPLayerBiomeFactory& operator=(const STPLayerBiomeFactory&) = delete;

		// STPBiomeFactory::operator(): one map into host memory, synchronous.
		void operator()(unsigned short* biomemap, int2 offset);
		// `count` maps into device memory (map i at biomemap + i * dimension.x * dimension.y), offsets in
		// host memory; asynchronous on `stream`.
		void generateDevice(unsigned short* biomemap, const int2* offset, unsigned int count, hipStream_t stream);
		// The same over arbitrary rectangles: map i has origin[i] and size (width, height), row pitch
		// `pitch` elements, map stride `stride` elements.
		void generateDevice(unsigned short* biomemap, const int2* origin, unsigned int count, unsigned int width,
			unsigned int height, size_t pitch, size_t stride, hipStream_t stream);

		// Builds `plan` for `maps` (uploads on `stream`). Throws STPInvalidArgument for empty maps or
		// rectangles beyond 2^32 samples or the int coordinate range.
		void plan(STPLayerPlan& plan, const std::vector<STPLayerMap>& maps, hipStream_t stream) const;
		// Writes maps [first, last) of `plan` into `out`. A root id >= `limit` is written as 0 and
		// flagged (status bit 1). Launches only, so a graph may capture it.
		void run(const STPLayerPlan& plan, unsigned short* out, unsigned int first, unsigned int last, unsigned int limit,
			hipStream_t stream) const;
		// Bit 0: a layer read outside its parent's rectangle (an internal error); bit 1: a root id at
		// or beyond the limit. Reading clears them; synchronous.
		unsigned int status();

		struct STPImplementation;

	private:
		std::unique_ptr<STPImplementation> Impl;
	};

} // namespace SuperTerrainPlus::STPDiversity

#endif // STP_BIOME_LAYER_HPP
