// Host class of the batched heightfield path (implementation in stp_batch.hip).
#pragma once
#ifndef STP_AMD_BATCH_HPP
#define STP_AMD_BATCH_HPP

#include "stp_stage_noise.hpp"

#include <array>
#include <vector>

namespace SuperTerrainPlus {

	// Shape of the batched workload shared by every chunk.
	struct STPBatchSetting {
		unsigned long long BiomeSeed; // synthetic Voronoi biome map
		unsigned int BiomeCount;
		unsigned int BiomeCell;
		unsigned int Radius;          // single histogram filter radius (even)
		unsigned int BinCapacity;     // most distinct biomes one pixel's histogram may hold
		float DropsPerPixel;          // raindrops of a chunk = side^2 * DropsPerPixel
	};

	struct STPBatchChunk {
		unsigned int Side; // square chunk, pixels per side
		int OriginX, OriginY;
	};

	namespace STPBatch {
		// Device record of one chunk (square, `Side` pixels).
		struct Chunk {
			unsigned int Side;
			int OriginX, OriginY;      // world pixel of the chunk's first pixel
			unsigned int Drops;
			unsigned long long Pixel;  // first pixel in the height / fixed / histogram arenas
			unsigned long long Halo;   // first sample in the halo biome arena ((Side + 2r)^2)
			unsigned long long Line;   // first cell in the vertical filter arenas ((Side + 2r) * Side)
			unsigned int DropBase;     // first drop in the ledger / drop-info arenas
			unsigned int StreamId;     // Philox counter word of this chunk's drops
		};
	}

	class STPHeightfieldBatch {
	public:
		// Graph: one captured hipGraph replay per run. Eager: the same launches without a graph.
		// Plain: one chunk per launch sequence, the host waits for each chunk (reference pattern).
		enum class STPRunMode : int { Graph = 0, Eager = 1, Plain = 2 };
		static constexpr int StageCount = 6; // biome, vertical, horizontal, blend, erosion, format

		STPHeightfieldBatch(const STPEnvironment::STPSimplexNoiseSetting& noise, const std::vector<STPBiomeProperty>& biome,
			const STPEnvironment::STPRainDropSetting& rain, const STPBatchSetting& setting,
			const std::vector<STPBatchChunk>& chunks);
		STPHeightfieldBatch(const STPHeightfieldBatch&) = delete;
		STPHeightfieldBatch& operator=(const STPHeightfieldBatch&) = delete;
		~STPHeightfieldBatch();

		void run(STPRunMode mode, hipStream_t stream);
		// Record HIP events between the stages of later runs. Timed graph runs launch one graph per
		// stage with the events recorded on the stream between them.
		void setTiming(bool enable);
		// Per-stage milliseconds of the last timed run (after the stream has been synchronised).
		std::vector<float> stageTimes() const;
		// Host wall milliseconds of each chunk in the last plain run.
		const std::vector<double>& latency() const noexcept { return this->Latency; }
		// {steps, erode steps, deposit steps, killed drops} of the last run (synchronises).
		std::array<unsigned long long, 4> counters(hipStream_t stream);
		// Throws STPCapacityError if a histogram overflowed its bin capacity (synchronises).
		void check(hipStream_t stream);
		void readChunk(unsigned int i, float* height, unsigned short* fixed, unsigned short* halo, float* ledger) const;

		unsigned long long pixels() const noexcept;
		unsigned long long drops() const noexcept { return this->Drops; }
		unsigned int chunkCount() const noexcept { return static_cast<unsigned int>(this->Record.size()); }

	private:
		struct STPTileRange {
			unsigned int Begin[5], End[5];
		};

		STPBatchSetting Setting;
		STPAlgorithm::STPPermutationGenerator::STPPermutationTable Permutation;
		STPStage::BiomeTable Biome;
		unsigned int BiomeCount;
		STPRainDropParameter Rain {};
		unsigned int VCap = 0u, BrushSize = 0u;
		std::vector<STPBatch::Chunk> Record;
		std::vector<STPTileRange> Range;
		std::array<unsigned int, 5> TileCount {};
		unsigned long long Pixels = 0ull, Drops = 0ull;

		STPSmartDeviceMemory::STPDevice<STPBatch::Chunk> ChunkRecord;
		STPSmartDeviceMemory::STPDevice<uint2> Tile[5];
		STPSmartDeviceMemory::STPDevice<int2> BrushOffset;
		STPSmartDeviceMemory::STPDevice<float> BrushWeight;
		STPSmartDeviceMemory::STPDevice<float> Height;
		STPSmartDeviceMemory::STPDevice<unsigned short> Fixed, BiomeMap;
		STPSmartDeviceMemory::STPDevice<unsigned int> VBin;
		STPSmartDeviceMemory::STPDevice<unsigned char> VCount, BinCount;
		STPSmartDeviceMemory::STPDevice<unsigned short> BinItem;
		STPSmartDeviceMemory::STPDevice<float> BinWeight, Ledger;
		STPSmartDeviceMemory::STPDevice<uint2> DropInfo;
		STPSmartDeviceMemory::STPDevice<unsigned int> Flag;
		STPSmartDeviceMemory::STPDevice<unsigned long long> Counter;

		hipStream_t CaptureStream = nullptr;
		hipGraph_t Graph = nullptr;
		hipGraphExec_t GraphExec = nullptr;
		hipGraph_t StageGraph[StageCount] {};
		hipGraphExec_t StageExec[StageCount] {};
		bool Timing = false, LastTimed = false;
		hipEvent_t Event[StageCount + 1] {};
		std::vector<double> Latency;

		void enqueueStage(hipStream_t stream, int stage, unsigned int first, unsigned int last, bool clear_flag);
		void enqueue(hipStream_t stream, unsigned int first, unsigned int last, bool timed, bool clear_flag);
		hipGraphExec_t capture(int stage_begin, int stage_end, hipGraph_t& graph);
	};

} // namespace SuperTerrainPlus

#endif // STP_AMD_BATCH_HPP
