// Host class of the batched heightfield path (implementation in stp_batch.hip).
#pragma once
#ifndef STP_AMD_BATCH_HPP
#define STP_AMD_BATCH_HPP

#include "stp_stage_erosion.hpp"
#include "stp_stage_noise.hpp"

#include <array>
#include <memory>
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
		// null: the synthetic Voronoi biome map; otherwise every chunk's biome map and filter halo come
		// from this layer chain (ids >= BiomeCount are written as 0 and flagged by the factory)
		STPDiversity::STPLayerBiomeFactory* BiomeLayers = nullptr;
		// true (and no layer chain): every chunk's biome map and filter halo are copied from the map
		// set by STPHeightfieldBatch::setBiomeSource, one sample per pixel (Step is not applied)
		bool ExternalBiome = false;
	};

	struct STPBatchChunk {
		unsigned int Side; // square chunk, pixels per side
		int OriginX, OriginY;
		unsigned int Step = 1u; // world units per pixel (level of detail); 0 is read as 1
		// Free-slip neighbourhood (odd): the chunk is the merged map of Neighbour x Neighbour chunks
		// of side Side / Neighbour, and its drops spawn in the centre one, as the reference's
		// erosion of a neighbourhood (STPHeightfieldKernel.cu:93-102). 1 = a chunk on its own.
		unsigned int Neighbour = 1u;
		// Free-slip halo (Neighbour 1 only): the chunk is a centre map of side Side - 2 ErosionHalo
		// with ErosionHalo pixels of its neighbours' terrain on every side, and its drops spawn in
		// the centre. With ErosionHalo >= MaxLifetime + max(r, 2) + 1 no drop reaches the halo's
		// border, so the centre erodes as the centre of a neighbourhood of uneroded chunks would.
		// copyOut copies the centre map. The filter lines restart at the centre's edges, so before
		// erosion the centre is bit-exact with the chunk generated alone (needs the filter's item
		// path: at most 8 biomes and a bin capacity of at most 8). 0 = no halo.
		unsigned int ErosionHalo = 0u;
		// Raindrops of the chunk; AutoDrops: the spawn area's pixels x DropsPerPixel.
		static constexpr unsigned int AutoDrops = 0xFFFFFFFFu;
		unsigned int Drops = AutoDrops;
	};

	// Arena sizes of a batch that is bound to different chunk lists over its lifetime.
	struct STPBatchCapacity {
		unsigned long long Pixels; // sum of side^2 (each chunk rounded up to 4 pixels)
		unsigned long long Drops;
		unsigned int Chunks;
		unsigned int MaxSide;
	};

	// Launch shapes of the erosion stage (chosen by STPHeightfieldBatch::autotune or set by hand).
	struct STPBatchTuning {
		// Core tile side of the LDS-tiled erosion: each workgroup erodes the drops that spawn in
		// one tile on an LDS copy of the tile plus a halo the drops cannot leave within their
		// lifetime (MaxLifetime + max(r, 2) + 1 cells); tiles run in four colour batches, which
		// needs tile >= 2 halo. 0 = every drop works on the heightmap in HBM with global atomics.
		unsigned int ErosionTile;
		unsigned int ErosionThreads; // threads of one tile workgroup: 256, 512 or 1024
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
			unsigned int DropBase;     // first drop in the ledger / drop-info / drop-bin arenas
			unsigned int StreamId;     // Philox counter word of this chunk's drops
			unsigned int TileBase;     // first erosion tile of the chunk in the tile-bin arenas
			unsigned int Step;         // world units per pixel
			unsigned int SpawnSide;    // drops spawn in the square [SpawnOrigin, SpawnOrigin + SpawnSide)^2
			unsigned int SpawnOrigin;  //   (the centre map of a neighbourhood, the centre of a halo chunk)
			unsigned int Crop;         // pixels copyOut trims on every side (the free-slip halo)
			unsigned int Segs;         // filter lines are split into Segs segments of outputs
			unsigned int SegLen;       //   of at most SegLen outputs (a halo chunk splits halo, centre, halo apart)
			unsigned long long SegLine; // first slot of the chunk in the segment-state arena (Segs * (Side + 2r) slots)
		};
	}

	class STPHeightfieldBatch {
	public:
		// Graph: one captured hipGraph replay per run. Eager: the same launches without a graph.
		// Plain: one chunk per launch sequence, the host waits for each chunk (reference pattern).
		enum class STPRunMode : int { Graph = 0, Eager = 1, Plain = 2 };
		// biome, vertical filter, horizontal filter, blend, erosion binning, erosion, format
		static constexpr int StageCount = 7;

		// A batch sized for and bound to `chunks` (chunk i's drops use stream id i).
		STPHeightfieldBatch(const STPEnvironment::STPSimplexNoiseSetting& noise, const std::vector<STPBiomeProperty>& biome,
			const STPEnvironment::STPRainDropSetting& rain, const STPBatchSetting& setting,
			const std::vector<STPBatchChunk>& chunks);
		// An unbound batch with arenas of the given capacity; bind() a chunk list before running.
		STPHeightfieldBatch(const STPEnvironment::STPSimplexNoiseSetting& noise, const std::vector<STPBiomeProperty>& biome,
			const STPEnvironment::STPRainDropSetting& rain, const STPBatchSetting& setting,
			const STPBatchCapacity& capacity);

		// Binds a chunk list within the capacity (stream_id: the Philox stream word of each chunk's
		// drops; empty = chunk index). The chunk and tile tables are copied asynchronously on
		// `stream`; the batch must not be running.
		void bind(const std::vector<STPBatchChunk>& chunks, const std::vector<unsigned int>& stream_id, hipStream_t stream);
		static STPBatchCapacity capacityOf(const std::vector<STPBatchChunk>& chunks, float drops_per_pixel);
		const STPBatchCapacity& capacity() const noexcept { return this->Capacity; }
		// Device addresses of chunk i's outputs (valid until the next bind).
		float* heightOf(unsigned int i) const;
		// Device addresses of chunk i's filter output (valid until the next bind): per pixel a bin count
		// and BinCapacity slots. With packedHistogram() a slot is item | count << 3 and the weight is
		// count * (1 / (2r + 1)^2) (no weight array); otherwise slots hold items and binWeightOf the
		// weights.
		const unsigned char* binCountOf(unsigned int i) const;
		const unsigned short* binItemOf(unsigned int i) const;
		const float* binWeightOf(unsigned int i) const;
		bool packedHistogram() const noexcept { return this->packedBins(); }
		// The external biome map of ExternalBiome batches: halo sample (x, y) of chunk c, for x, y in
		// [0, Side + 2 Radius), is origin[(c.OriginY + y - Radius) * pitch + c.OriginX + x - Radius]
		// (device memory; the caller keeps every such sample inside its allocation). On the item path
		// a sample >= BiomeCount is read as 0 and sets status bit 4.
		void setBiomeSource(const STPSample_t* origin, unsigned int pitch);
		// A one-chunk batch that runs the filter passes alone on square maps of `side` (ExternalBiome;
		// its noise, biome-map and erosion settings are unused), for the per-chunk generators; null
		// where the item path does not apply (more than 8 biomes or bins, a radius above 126).
		static std::unique_ptr<STPHeightfieldBatch> makeFilter(const std::vector<STPBiomeProperty>& biome,
			unsigned int side, unsigned int radius, unsigned int bin_capacity);
		// Binds the one chunk and enqueues stages 0-2 (halo copy, vertical and horizontal passes) on
		// `stream`: the chunk's pixel (0, 0) is origin[0] of a sample map of row pitch `pitch`, whose
		// samples within Radius of the chunk must exist. Read the histograms with binCountOf(0) etc.
		void filterExternal(const STPSample_t* origin, unsigned int pitch, hipStream_t stream);
		unsigned short* fixedOf(unsigned int i) const;
		// Enqueues copies of every bound chunk's outputs to height[i] / fixed[i] (device memory, one
		// pointer per chunk) on `stream`, in one launch.
		void copyOut(const std::vector<float*>& height, const std::vector<unsigned short*>& fixed, hipStream_t stream);
		// Enqueues stages [first, last) of every chunk on `stream` without a graph (eager). A caller
		// that writes the heights itself (heightOf) runs the erosion binning, erosion and format
		// stages only (4, 7).
		void runStages(int first, int last, hipStream_t stream);
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
		// Sum over the bound chunks' pixels of the last run's histogram bin counts (synchronises)
		unsigned long long binTotal(hipStream_t stream);
		// Throws STPCapacityError if a histogram overflowed its bin capacity or a tiled erosion
		// window cell left its range (synchronises).
		void check(hipStream_t stream);
		// Status bits of the runs since the flags were last cleared (construction and every run of
		// stage 0 clear them; a run of stage 4 clears bit 2): 1 a histogram overflowed its bin capacity, 2 a
		// tiled erosion window cell left [WindowLowest, WindowHighest], 4 an external biome sample was not
		// below BiomeCount on the item path (synchronises).
		unsigned int status(hipStream_t stream);
		// Device address of the status word of bit 2 (nonzero when a tiled erosion window cell left its
		// range; cleared by every run of stage 4), for work that decides on the device.
		const unsigned int* windowStatus() const noexcept { return this->Flag.get() + 1; }
		void readChunk(unsigned int i, float* height, unsigned short* fixed, unsigned short* halo, float* ledger) const;

		// Switches the erosion launch shape (throws STPInvalidArgument if it cannot run here: the
		// tiled erosion needs MaxLifetime > 0 and a window that fits in LDS). Captured graphs are
		// rebuilt at the next graph run.
		void setTuning(const STPBatchTuning& tuning);
		const STPBatchTuning& tuning() const noexcept { return this->Tuning; }
		// Times every candidate launch shape on this batch (graph runs on `stream`, the whole
		// pipeline) and keeps the fastest. Returns the chosen shape.
		STPBatchTuning autotune(hipStream_t stream, unsigned int repeat = 2u);
		// The shape a new batch starts with: the largest tile whose window fits in LDS.
		STPBatchTuning defaultTuning() const;
		// The smallest valid tile with 1024 threads: the most workgroups per colour batch, which
		// shortens a small batch (a few chunks) at some cost in throughput.
		STPBatchTuning latencyTuning() const;

		unsigned long long pixels() const noexcept;
		unsigned long long drops() const noexcept { return this->Drops; }
		unsigned int chunkCount() const noexcept { return static_cast<unsigned int>(this->Record.size()); }

	private:
		// tile tables: 0 biome, 1 vertical filter, 2 horizontal filter, 3 blend, 4 drops,
		// 5..8 erosion tiles of colour 0..3
		static constexpr int TableCount = 9;
		struct STPTileRange {
			unsigned int Begin[TableCount], End[TableCount];
		};

		STPBatchSetting Setting;
		STPBatchCapacity Capacity {};
		STPAlgorithm::STPPermutationGenerator::STPPermutationTable Permutation;
		STPStage::BiomeTable Biome;
		unsigned int BiomeCount;
		// every biome's heights lie in [0, WindowHeightLimit], so the LDS-tiled erosion may run
		bool WindowHeights;
		STPRainDropParameter Rain {};
		unsigned int VCap = 0u, BrushSize = 0u;
		std::vector<STPBatch::Chunk> Record;
		std::vector<STPTileRange> Range;
		std::array<unsigned int, TableCount> TileCount {};
		STPBatchTuning Tuning {};
		unsigned int ErosionTiles = 0u, ErosionHalo = 0u, ErosionLds = 0u;
		unsigned long long Pixels = 0ull, Drops = 0ull;

		// chunk records and tile tables, packed in one device arena (filled from pinned staging)
		STPSmartDeviceMemory::STPDevice<unsigned char> TableArena;
		STPSmartDeviceMemory::STPHost<unsigned char> TableStaging;
		size_t TableBytes = 0u;
		STPBatch::Chunk* ChunkRecord = nullptr;
		uint2* Tile[TableCount] {};
		unsigned int TileCapacity = 0u;
		// destination pointers of copyOut (device) and their pinned staging
		STPSmartDeviceMemory::STPDevice<void*> OutTable;
		STPSmartDeviceMemory::STPHost<void*> OutStaging;
		size_t OutCapacity = 0u;
		STPSmartDeviceMemory::STPDevice<int2> BrushOffset;
		STPSmartDeviceMemory::STPDevice<float> BrushWeight;
		STPSmartDeviceMemory::STPDevice<float> Height;
		// tiled erosion: per drop its rank in its spawn tile and the tile-sorted drop list; per
		// tile its drop count and first slot
		STPSmartDeviceMemory::STPDevice<unsigned int> DropRank, DropBin, TileDrops, TileFirst;
		STPSmartDeviceMemory::STPDevice<unsigned short> Fixed, BiomeMap;
		STPSmartDeviceMemory::STPDevice<unsigned int> VBin;
		// filter line segments: per slot (segment of a line) and biome, the run summary of the
		// segment and then the state carried into it (3 words per biome, at most 8 biomes)
		STPSmartDeviceMemory::STPDevice<unsigned int> SegState;
		unsigned long long SegCapacity = 0ull; // slots
		bool Segmented = false;                // some chunk of the binding has Segs > 1
		STPSmartDeviceMemory::STPDevice<unsigned char> VCount, BinCount;
		STPSmartDeviceMemory::STPDevice<unsigned short> BinItem;
		STPSmartDeviceMemory::STPDevice<float> BinWeight, Ledger;
		STPSmartDeviceMemory::STPDevice<uint2> DropInfo;
		STPSmartDeviceMemory::STPDevice<unsigned int> Flag;
		const STPSample_t* BiomeSource = nullptr;
		unsigned int BiomeSourcePitch = 0u;
		STPSmartDeviceMemory::STPDevice<unsigned long long> Counter;
		// the layer chain's plan of the bound chunks' halo windows (biome maps from a layer chain)
		std::unique_ptr<STPDiversity::STPLayerPlan> LayerPlan;

		hipStream_t CaptureStream = nullptr;
		// Work forked from a stage's stream and joined back into it with these events: the tiled
		// erosion's second chunk group (Side[0]), and the chains of a graph run's chunk groups after
		// the first (Side[0..PipelineGroups - 2]).
		static constexpr int SideStreams = 3;
		hipStream_t Side[SideStreams] {};
		hipEvent_t Fork = nullptr, Join[SideStreams] {};
		hipGraph_t Graph = nullptr;
		hipGraphExec_t GraphExec = nullptr;
		hipGraph_t StageGraph[StageCount] {};
		hipGraphExec_t StageExec[StageCount] {};
		bool Timing = false, LastTimed = false;
		hipEvent_t Event[StageCount + 1] {};
		hipEvent_t TableCopied = nullptr, OutCopied = nullptr;
		std::vector<double> Latency;

		// split_erosion: the tiled erosion of a large range may fork its second chunk group onto Side[0]
		void enqueueStage(hipStream_t stream, int stage, unsigned int first, unsigned int last, bool clear_flag,
			bool split_erosion = true);
		// The chunk groups (about equal pixels each, [bound[g], bound[g + 1])) whose whole pipelines a
		// graph run captures as concurrent chains; a single group where the binding is too small.
		std::vector<unsigned int> pipelineGroups() const;
		void enqueue(hipStream_t stream, unsigned int first, unsigned int last, bool timed, bool clear_flag);
		hipGraphExec_t capture(int stage_begin, int stage_end, hipGraph_t& graph);
		void releaseGraphs() noexcept;
		void init(const STPEnvironment::STPRainDropSetting& rain);
		void buildTileTables(hipStream_t stream);
		// Window side bound and LDS bytes of the tiled erosion for a tile side (0 if impossible).
		unsigned int erosionLdsBytes(unsigned int tile) const;
		STPStage::ErosionArgs erosionArgs() const;
		// the filter passes with per-biome registers (at most 8 biomes and 8 bins per pixel)
		bool itemPath() const noexcept;
		// the item path with each histogram slot packed as item | count << 3 (no weight array)
		bool packedBins() const noexcept;
	};

} // namespace SuperTerrainPlus

#endif // STP_AMD_BATCH_HPP
