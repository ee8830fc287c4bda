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

#include <chrono>

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPStage;

namespace SuperTerrainPlus::STPBatch {

	constexpr unsigned int BiomeTile = 256u;
	constexpr unsigned int BlendX = 64u, BlendY = 4u, BlendPixel = 2u;
	constexpr unsigned int ErosionBlock = 256u;

	__global__ void __launch_bounds__(BiomeTile) biomeKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const unsigned long long seed, const unsigned int count, const unsigned int cell, const unsigned int radius,
		unsigned short* const biome) {
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int w = c.Side + 2u * radius;
		const unsigned int p = t.y * BiomeTile + threadIdx.x;
		if (p >= w * w) {
			return;
		}
		const unsigned int x = p % w, y = p / w;
		biome[c.Halo + p] = voronoiBiome(seed, count, cell, static_cast<long long>(c.OriginX) - radius + x,
			static_cast<long long>(c.OriginY) - radius + y);
	}

	__global__ void __launch_bounds__(ShfLineBlock) verticalKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const unsigned int radius, const unsigned int vcap, const unsigned short* const biome, unsigned int* const vbin,
		unsigned char* const vcount, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int col = t.y * ShfLineBlock + threadIdx.x, ncols = c.Side + 2u * radius;
		if (col >= ncols) {
			return;
		}
		if (shfVerticalLine(biome + c.Halo, ncols, make_uint2(radius, radius), make_uint2(c.Side, c.Side), radius, vcap,
				lds + threadIdx.x, vbin + c.Line * vcap, vcount + c.Line, col)) {
			atomicOr(overflow, 1u);
		}
	}

	__global__ void __launch_bounds__(ShfLineBlock) horizontalKERNEL(const uint2* const tiles,
		const Chunk* const chunks, const unsigned int radius, const unsigned int vcap, const unsigned int cap,
		const unsigned int* const vbin, const unsigned char* const vcount, unsigned char* const bin_count,
		unsigned short* const bin_item, float* const bin_weight, unsigned int* const overflow) {
		extern __shared__ unsigned int lds[];
		const uint2 t = tiles[blockIdx.x];
		const Chunk c = chunks[t.x];
		const unsigned int row = t.y * ShfLineBlock + threadIdx.x;
		if (row >= c.Side) {
			return;
		}
		if (shfHorizontalLine(make_uint2(c.Side, c.Side), radius, vcap, cap, vbin + c.Line * vcap, vcount + c.Line,
				lds + threadIdx.x, bin_count + c.Pixel, bin_item + c.Pixel * cap, bin_weight + c.Pixel * cap, row)) {
			atomicOr(overflow, 1u);
		}
	}

	__global__ void __launch_bounds__(BlendX * BlendY) blendKERNEL(const uint2* const tiles, const Chunk* const chunks,
		const STPAlgorithm::STPPermutation perm, const BiomeTable biomes, const unsigned int biome_count,
		const unsigned int cap, const unsigned char* const bin_count, const unsigned short* const bin_item,
		const float* const bin_weight, float* const height) {
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
		const unsigned int x = (t.y % tilesX) * BlendX + threadIdx.x;
		const unsigned int y0 = ((t.y / tilesX) * BlendY + threadIdx.y) * BlendPixel;
		if (x >= c.Side) {
			return;
		}
		const float half = static_cast<float>(c.Side) / 2.0f;
		// pixel x of the chunk samples world coordinate Origin + x
		const float2 offset = make_float2(static_cast<float>(c.OriginX) + half, static_cast<float>(c.OriginY) + half);
		for (unsigned int k = 0u; k < BlendPixel; k++) {
			const unsigned int y = y0 + k;
			if (y < c.Side) {
				const unsigned long long i = c.Pixel + static_cast<unsigned long long>(y) * c.Side + x;
				height[i] = blendPixel(table, prop, biome_count, bin_count[i], bin_item + i * cap, bin_weight + i * cap, x, y,
					offset, make_float2(half, half));
			}
		}
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
		args.MapSize = make_uint2(c.Side, c.Side);
		args.NN = make_uint2(1u, 1u);
		args.Total = args.MapSize;
		args.StreamId = c.StreamId;
		args.Rain.RainDropCount = c.Drops;
		erodeDropParallel(height + c.Pixel, args, brush, drop, ledger + c.DropBase, drop_info + c.DropBase);
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
		if (threadIdx.x < 4u) {
			p[threadIdx.x] = 0ull;
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

STPHeightfieldBatch::STPHeightfieldBatch(const STPEnvironment::STPSimplexNoiseSetting& noise,
	const std::vector<STPBiomeProperty>& biome, const STPEnvironment::STPRainDropSetting& rain,
	const STPBatchSetting& setting, const std::vector<STPBatchChunk>& chunks) :
	Setting(setting), Permutation(STPAlgorithm::STPPermutationGenerator::generate(noise)),
	Biome(packBiomeTable(biome.data(), static_cast<unsigned int>(biome.size()))),
	BiomeCount(static_cast<unsigned int>(biome.size())) {
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
	if (chunks.empty()) {
		STP_THROW(STPInvalidArgument, "empty chunk list");
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

	// chunk records and arena offsets
	std::vector<Chunk> rec;
	unsigned long long pixel = 0ull, halo = 0ull, line = 0ull, drops = 0ull;
	for (size_t i = 0u; i < chunks.size(); i++) {
		const STPBatchChunk& in = chunks[i];
		if (in.Side < 2u * r.ErosionBrushRadius + 2u || in.Side > 16384u) {
			STP_THROW(STPInvalidArgument, "chunk side must be in [2r + 2, 16384]");
		}
		Chunk c;
		c.Side = in.Side;
		c.OriginX = in.OriginX;
		c.OriginY = in.OriginY;
		const unsigned long long px = static_cast<unsigned long long>(in.Side) * in.Side;
		c.Drops = static_cast<unsigned int>(static_cast<double>(px) * setting.DropsPerPixel);
		c.Pixel = pixel;
		c.Halo = halo;
		c.Line = line;
		c.DropBase = static_cast<unsigned int>(drops);
		c.StreamId = static_cast<unsigned int>(i);
		rec.push_back(c);
		// keep every chunk's first pixel 16-byte aligned in the float arena
		pixel += (px + 3ull) & ~3ull;
		halo += static_cast<unsigned long long>(in.Side + 2u * radius) * (in.Side + 2u * radius);
		line += static_cast<unsigned long long>(in.Side + 2u * radius) * in.Side;
		drops += c.Drops;
	}
	if (drops >= 0xFFFFFFFFull) {
		STP_THROW(STPInvalidArgument, "too many drops in one batch");
	}
	this->Record = rec;
	this->Pixels = pixel;
	this->Drops = drops;
	this->ChunkRecord = upload(rec);

	// tile tables, chunk-major so that one chunk's tiles are a contiguous range (plain mode)
	std::vector<uint2> tb, tv, th, tl, te;
	for (unsigned int i = 0u; i < rec.size(); i++) {
		const Chunk& c = rec[i];
		STPTileRange range;
		const unsigned int w = c.Side + 2u * radius;
		range.Begin[0] = static_cast<unsigned int>(tb.size());
		for (unsigned int k = 0u; k < (w * w + BiomeTile - 1u) / BiomeTile; k++) tb.push_back(make_uint2(i, k));
		range.Begin[1] = static_cast<unsigned int>(tv.size());
		for (unsigned int k = 0u; k < (w + ShfLineBlock - 1u) / ShfLineBlock; k++) tv.push_back(make_uint2(i, k));
		range.Begin[2] = static_cast<unsigned int>(th.size());
		for (unsigned int k = 0u; k < (c.Side + ShfLineBlock - 1u) / ShfLineBlock; k++) th.push_back(make_uint2(i, k));
		range.Begin[3] = static_cast<unsigned int>(tl.size());
		const unsigned int blend = ((c.Side + BlendX - 1u) / BlendX) * ((c.Side + BlendY * BlendPixel - 1u) / (BlendY * BlendPixel));
		for (unsigned int k = 0u; k < blend; k++) tl.push_back(make_uint2(i, k));
		range.Begin[4] = static_cast<unsigned int>(te.size());
		for (unsigned int k = 0u; k < (c.Drops + ErosionBlock - 1u) / ErosionBlock; k++) te.push_back(make_uint2(i, k));
		this->Range.push_back(range);
	}
	const size_t counts[5] = {tb.size(), tv.size(), th.size(), tl.size(), te.size()};
	for (unsigned int i = 0u; i < rec.size(); i++) {
		for (int s = 0; s < 5; s++) {
			this->Range[i].End[s] = i + 1u < rec.size() ? this->Range[i + 1u].Begin[s] : static_cast<unsigned int>(counts[s]);
		}
	}
	this->TileCount = {static_cast<unsigned int>(tb.size()), static_cast<unsigned int>(tv.size()),
		static_cast<unsigned int>(th.size()), static_cast<unsigned int>(tl.size()), static_cast<unsigned int>(te.size())};
	this->Tile[0] = upload(tb);
	this->Tile[1] = upload(tv);
	this->Tile[2] = upload(th);
	this->Tile[3] = upload(tl);
	this->Tile[4] = upload(te);

	// resident arenas
	const unsigned int cap = setting.BinCapacity;
	this->Height = STPSmartDeviceMemory::makeDevice<float>(pixel);
	this->Fixed = STPSmartDeviceMemory::makeDevice<unsigned short>(pixel);
	this->BiomeMap = STPSmartDeviceMemory::makeDevice<unsigned short>(halo);
	this->VBin = STPSmartDeviceMemory::makeDevice<unsigned int>(line * this->VCap);
	this->VCount = STPSmartDeviceMemory::makeDevice<unsigned char>(line);
	this->BinCount = STPSmartDeviceMemory::makeDevice<unsigned char>(pixel);
	this->BinItem = STPSmartDeviceMemory::makeDevice<unsigned short>(pixel * cap);
	this->BinWeight = STPSmartDeviceMemory::makeDevice<float>(pixel * cap);
	this->Ledger = STPSmartDeviceMemory::makeDevice<float>(drops ? drops : 1u);
	this->DropInfo = STPSmartDeviceMemory::makeDevice<uint2>(drops ? drops : 1u);
	this->Flag = STPSmartDeviceMemory::makeDevice<unsigned int>(4u);
	this->Counter = STPSmartDeviceMemory::makeDevice<unsigned long long>(4u);
	STP_CHECK_HIP(hipMemset(this->Fixed.get(), 0, pixel * sizeof(unsigned short)));
	STP_CHECK_HIP(hipMemset(this->Height.get(), 0, pixel * sizeof(float)));
	for (hipEvent_t& e : this->Event) {
		STP_CHECK_HIP(hipEventCreate(&e));
	}
	STP_CHECK_HIP(hipStreamCreateWithFlags(&this->CaptureStream, hipStreamNonBlocking));
}

STPHeightfieldBatch::~STPHeightfieldBatch() {
	for (int k = 0; k < StageCount; k++) {
		if (this->StageExec[k]) {
			(void)hipGraphExecDestroy(this->StageExec[k]);
		}
		if (this->StageGraph[k]) {
			(void)hipGraphDestroy(this->StageGraph[k]);
		}
	}
	if (this->GraphExec) {
		(void)hipGraphExecDestroy(this->GraphExec);
	}
	if (this->Graph) {
		(void)hipGraphDestroy(this->Graph);
	}
	for (hipEvent_t e : this->Event) {
		(void)hipEventDestroy(e);
	}
	if (this->CaptureStream) {
		(void)hipStreamDestroy(this->CaptureStream);
	}
}

void STPHeightfieldBatch::enqueueStage(const hipStream_t stream, const int stage, const unsigned int first,
	const unsigned int last, const bool clear_flag) {
	const unsigned int radius = this->Setting.Radius, cap = this->Setting.BinCapacity;
	const auto range = [&](const int s) {
		const unsigned int b = this->Range[first].Begin[s], e = this->Range[last - 1u].End[s];
		return std::make_pair(this->Tile[s].get() + b, e - b);
	};
	switch (stage) {
	case 0: {
		if (clear_flag) {
			zeroKERNEL<<<1, 64, 0, stream>>>(this->Flag.get(), 4u);
		}
		const auto [t, n] = range(0);
		if (n) {
			biomeKERNEL<<<n, BiomeTile, 0, stream>>>(t, this->ChunkRecord.get(), this->Setting.BiomeSeed,
				this->Setting.BiomeCount, this->Setting.BiomeCell, radius, this->BiomeMap.get());
		}
		break;
	}
	case 1: {
		const auto [t, n] = range(1);
		if (n) {
			verticalKERNEL<<<n, ShfLineBlock, this->VCap * ShfLineBlock * sizeof(unsigned int), stream>>>(t,
				this->ChunkRecord.get(), radius, this->VCap, this->BiomeMap.get(), this->VBin.get(), this->VCount.get(),
				this->Flag.get());
		}
		break;
	}
	case 2: {
		const auto [t, n] = range(2);
		if (n) {
			horizontalKERNEL<<<n, ShfLineBlock, 2u * cap * ShfLineBlock * sizeof(unsigned int), stream>>>(t,
				this->ChunkRecord.get(), radius, this->VCap, cap, this->VBin.get(), this->VCount.get(), this->BinCount.get(),
				this->BinItem.get(), this->BinWeight.get(), this->Flag.get());
		}
		break;
	}
	case 3: {
		const auto [t, n] = range(3);
		if (n) {
			blendKERNEL<<<n, dim3(BlendX, BlendY), 0, stream>>>(t, this->ChunkRecord.get(), this->Permutation.Permutation,
				this->Biome, this->BiomeCount, cap, this->BinCount.get(), this->BinItem.get(), this->BinWeight.get(),
				this->Height.get());
		}
		break;
	}
	case 4: {
		const auto [t, n] = range(4);
		if (n) {
			ErosionArgs args;
			args.Rain = this->Rain;
			args.MapSize = make_uint2(0u, 0u);
			args.NN = make_uint2(1u, 1u);
			args.Total = args.MapSize;
			args.BrushSize = this->BrushSize;
			args.StreamId = 0u;
			erosionKERNEL<<<n, ErosionBlock, this->BrushSize * sizeof(int2), stream>>>(t, this->ChunkRecord.get(), args,
				this->BrushOffset.get(), this->BrushWeight.get(), this->Height.get(), this->Ledger.get(),
				this->DropInfo.get());
		}
		break;
	}
	case 5: {
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
	STP_CHECK_HIP(hipStreamBeginCapture(this->CaptureStream, hipStreamCaptureModeThreadLocal));
	bool ok = true;
	try {
		for (int s = stage_begin; s < stage_end; s++) {
			this->enqueueStage(this->CaptureStream, s, 0u, n, true);
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

void STPHeightfieldBatch::setTiming(const bool enable) {
	this->Timing = enable;
}

std::vector<float> STPHeightfieldBatch::stageTimes() const {
	if (!this->LastTimed) {
		STP_THROW(STPInvalidArgument, "the last run was not timed");
	}
	std::vector<float> ms(6);
	for (int k = 0; k < 6; k++) {
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

void STPHeightfieldBatch::check(const hipStream_t stream) {
	unsigned int flag = 0u;
	STP_CHECK_HIP(hipMemcpyAsync(&flag, this->Flag.get(), sizeof(flag), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipStreamSynchronize(stream));
	if (flag != 0u) {
		STP_THROW(STPCapacityError, "a pixel histogram holds more distinct biomes than the bin capacity");
	}
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
