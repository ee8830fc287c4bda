// The reference's world pipeline contract on resident device state: a requested chunk is complete
// once it has been eroded as the centre of its free-slip erosion neighbourhood, which needs the
// heightmap of every chunk of that neighbourhood first; centres whose neighbourhoods overlap are
// never eroded at the same time.
//
// SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp: requestChunk (:640-663) ->
// erodeHeightmap (:485-559: heightmaps of the neighbourhoods, conflict-free schedule, erosion of
// the centres in batches) -> generateHeightmap (:437-476) -> generateBiomemap (:411-430); the
// schedule is scheduleChunk (:339-404). Here the biome map is procedural and sampled with its
// filter halo inside the heightmap batch, the heightmaps of many chunks come from one batch, and
// one batch erodes every centre of a conflict-free group at once, on merged neighbourhood maps.
#pragma once
#ifndef STP_AMD_WORLD_HPP
#define STP_AMD_WORLD_HPP

#include "stp_batch.hpp"

#include <atomic>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace SuperTerrainPlus {

	struct STPWorldSetting {
		unsigned int MapSize;          // chunk side in pixels (STPChunkSetting::MapSize)
		unsigned int ErosionNeighbour; // odd: chunks per side of the erosion neighbourhood (STPChunkSetting::ErosionNearestNeighbour)
		unsigned int GenerateChunks;   // heightmaps generated per batch
		unsigned int ErodeChunks;      // neighbourhoods eroded per batch
	};

	class STPHeightfieldWorld {
	public:
		// chunk state: no map, heightmap generated (not yet eroded as a centre), complete
		enum class STPChunkState : unsigned int { None = 0u, Heightmap = 1u, Complete = 2u };

		STPHeightfieldWorld(const STPEnvironment::STPSimplexNoiseSetting& noise, const std::vector<STPBiomeProperty>& biome,
			const STPEnvironment::STPRainDropSetting& rain, const STPBatchSetting& batch, const STPWorldSetting& setting);
		STPHeightfieldWorld(const STPHeightfieldWorld&) = delete;
		STPHeightfieldWorld& operator=(const STPHeightfieldWorld&) = delete;
		~STPHeightfieldWorld();

		// Makes every requested chunk complete and returns the chunks whose maps changed: the
		// union of the eroded neighbourhoods, in first-seen order (the reference's
		// RequireHeightmapReload). Synchronous.
		std::vector<int2> request(const std::vector<int2>& chunks);
		// Copies a resident chunk (heightmap or complete) to host memory; false if it has no map.
		bool read(int2 chunk, float* height, unsigned short* fixed);
		// The rendered area around `centre` ((2 radius + 1)^2 chunks, made complete first), as the
		// reference's rendering buffer holds it (STPWorldPipeline.cpp:671-707): the chunks' UINT16
		// heightmaps and their biome map side by side in device buffers of side
		// (2 radius + 1) * MapSize, row-major. Synchronous.
		void renderArea(int2 centre, unsigned int radius, unsigned short* biome, unsigned short* height);
		STPChunkState state(int2 chunk) const;
		unsigned long long residentBytes() const noexcept { return this->Resident; }
		// Chunk records (stp_chunk_record.hpp) of the listed chunks, each with a map, one after
		// another into host memory of `capacity` bytes: `maps` selects the biome map (the one the
		// heightmap sampled), FP32 and UINT16 heights. Returns the bytes written. Synchronous.
		unsigned long long exportRecords(const std::vector<int2>& chunks, unsigned int maps, unsigned char* out,
			unsigned long long capacity);

		// STPWorldPipeline::load (STPWorldPipeline.cpp:1504-1633): the rendered area around `centre`
		// ((2 radius + 1)^2 chunks, made complete) is assembled on a worker thread into the back
		// buffer while the caller renders from the front buffer. BackBufferBusy: the worker is still
		// loading (this call is ignored), or a changed centre's load has just started (a finished
		// load that was not swapped is discarded); Swapped: the load of this centre finished and the
		// buffers were swapped; Unchanged: the same centre and nothing finished since. Every other
		// member waits for a running load.
		enum class STPWorldLoadStatus : unsigned int { Unchanged = 0u, BackBufferBusy = 1u, Swapped = 2u };
		STPWorldLoadStatus load(int2 centre, unsigned int radius);
		// The front buffer: device biome and UINT16 height maps of the last swapped load, `side` =
		// (2 radius + 1) MapSize pixels a side, row-major (null and 0 before the first swap); valid
		// until the next Swapped load.
		void front(const unsigned short** biome, const unsigned short** height, unsigned int* side) const;

		// Conflict-free schedule: batch number of each centre, first fit in the given order; two
		// centres share a batch only if their neighbourhoods (nn x nn chunks) are disjoint.
		static std::vector<unsigned int> schedule(const std::vector<int2>& centres, unsigned int nn);
		// Philox stream word of the drops of the neighbourhood centred at a chunk
		static unsigned int streamWord(int2 chunk) noexcept;

	private:
		struct Key {
			int X, Y;
			bool operator==(const Key& k) const noexcept { return X == k.X && Y == k.Y; }
		};
		struct KeyHash {
			size_t operator()(const Key& k) const noexcept;
		};
		struct Entry {
			float* Height = nullptr;
			unsigned short* Fixed = nullptr;
			STPChunkState State = STPChunkState::None;
		};

		STPWorldSetting Setting;
		STPBatchSetting Batch;
		std::unique_ptr<STPHeightfieldBatch> Generator, Eroder;
		std::unordered_map<Key, Entry, KeyHash> Chunk;
		hipStream_t Stream = nullptr;
		hipMemPool_t Pool = nullptr;
		unsigned long long Resident = 0ull;
		// chunk record export: the byte stream as it leaves the device, its pack table and checks
		STPSmartDeviceMemory::STPDevice<unsigned char> ExportStream;
		unsigned long long ExportCapacity = 0ull;
		STPSmartDeviceMemory::STPDevice<unsigned char> ExportTable;
		STPSmartDeviceMemory::STPHost<unsigned char> ExportStaging;
		size_t ExportTableCapacity = 0u;
		std::unique_ptr<STPDiversity::STPLayerPlan> ExportPlan;

		// the asynchronous loader: the worker holds Lock for a whole load, every member takes it
		struct STPAreaBuffer {
			STPSmartDeviceMemory::STPDevice<unsigned short> Biome, Height;
			unsigned int Side = 0u;
		};
		mutable std::recursive_mutex Lock;
		std::thread Loader;
		std::atomic<unsigned int> LoaderState {0u}; // 0 idle, 1 loading, 2 finished and not swapped
		std::exception_ptr LoaderError;
		int Device = 0;
		bool Pending = false; // a load was started (of PendingCentre, PendingRadius)
		int2 PendingCentre {0, 0};
		unsigned int PendingRadius = 0u;
		STPAreaBuffer Front, Back;

		void generate(const std::vector<Key>& chunks);
		void erode(const std::vector<Key>& centres);
		// joins a finished loader thread; rethrows its error
		void joinLoader();
	};

} // namespace SuperTerrainPlus

#endif // STP_AMD_WORLD_HPP
