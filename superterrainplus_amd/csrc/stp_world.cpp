// STPHeightfieldWorld: the world pipeline's generation chain on resident device state (see
// stp_world.hpp for the reference functions it follows).
#include "stp_world.hpp"
#include "stp_chunk_record.hpp"
#include "stp_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <unordered_set>

using namespace SuperTerrainPlus;

namespace {
	// neighbour slots passed by value to one gather or scatter launch (kernel arguments stay
	// below 4 KiB)
	constexpr unsigned int MaxSlots = 128u, MaxMerged = 16u;

	struct NeighbourhoodTable {
		float* ChunkHeight[MaxSlots];
		unsigned short* ChunkFixed[MaxSlots];
		float* MergedHeight[MaxMerged];
		unsigned short* MergedFixed[MaxMerged];
	};

	// Merged pixel (gx, gy) of neighbourhood blockIdx.z is pixel (gx % map, gy % map) of its chunk
	// (gx / map, gy / map), chunks in row-major order
	// (SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPNearestNeighbourTextureBuffer.cpp:97-110).
	// Gather copies the chunks' heights into the merged map; scatter copies the merged heights and
	// UINT16 values back. One block row covers 256 consecutive merged pixels.
	template <bool Gather>
	__global__ void __launch_bounds__(256) neighbourhoodKERNEL(const NeighbourhoodTable table, const unsigned int map,
		const unsigned int nn) {
		const unsigned int i = blockIdx.z, side = map * nn;
		const unsigned int gx = blockIdx.x * 256u + threadIdx.x, gy = blockIdx.y;
		if (gx >= side) {
			return;
		}
		const unsigned int k = i * nn * nn + (gy / map) * nn + gx / map;
		const size_t l = static_cast<size_t>(gy % map) * map + gx % map, g = static_cast<size_t>(gy) * side + gx;
		if constexpr (Gather) {
			table.MergedHeight[i][g] = table.ChunkHeight[k][l];
		} else {
			table.ChunkHeight[k][l] = table.MergedHeight[i][g];
			table.ChunkFixed[k][l] = table.MergedFixed[i][g];
		}
	}

	// FP32 heights, then UINT16, 256-byte aligned parts
	unsigned long long heightBytes(const unsigned int map) {
		return ((static_cast<unsigned long long>(map) * map * sizeof(float)) + 255ull) & ~255ull;
	}
	unsigned long long chunkBytes(const unsigned int map) {
		return heightBytes(map) + ((static_cast<unsigned long long>(map) * map * sizeof(unsigned short) + 255ull) & ~255ull);
	}
} // namespace

size_t STPHeightfieldWorld::KeyHash::operator()(const Key& k) const noexcept {
	unsigned long long h = static_cast<unsigned int>(k.X) * 0x9E3779B97F4A7C15ull;
	h ^= static_cast<unsigned long long>(static_cast<unsigned int>(k.Y)) + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
	return static_cast<size_t>(h);
}

unsigned int STPHeightfieldWorld::streamWord(const int2 chunk) noexcept {
	unsigned int h = static_cast<unsigned int>(chunk.x) * 0x85EBCA6Bu;
	h ^= static_cast<unsigned int>(chunk.y) * 0xC2B2AE35u + 0x165667B1u + (h << 6) + (h >> 2);
	h ^= h >> 16;
	h *= 0x7FEB352Du;
	h ^= h >> 15;
	return h;
}

std::vector<unsigned int> STPHeightfieldWorld::schedule(const std::vector<int2>& centres, const unsigned int nn) {
	if (nn == 0u || (nn & 1u) == 0u) {
		STP_THROW(STPInvalidArgument, "the erosion neighbourhood is odd");
	}
	// neighbourhoods [X - h, X + h] overlap iff the centres are less than nn apart on both axes
	std::vector<unsigned int> batch(centres.size());
	std::vector<unsigned char> used;
	for (size_t i = 0u; i < centres.size(); i++) {
		used.assign(centres.size() + 1u, 0u);
		for (size_t j = 0u; j < i; j++) {
			const long long dx = static_cast<long long>(centres[i].x) - centres[j].x,
				dy = static_cast<long long>(centres[i].y) - centres[j].y;
			if (std::llabs(dx) < nn && std::llabs(dy) < nn) {
				used[batch[j]] = 1u;
			}
		}
		batch[i] = static_cast<unsigned int>(std::find(used.begin(), used.end(), 0u) - used.begin());
	}
	return batch;
}

STPHeightfieldWorld::STPHeightfieldWorld(const STPEnvironment::STPSimplexNoiseSetting& noise,
	const std::vector<STPBiomeProperty>& biome, const STPEnvironment::STPRainDropSetting& rain,
	const STPBatchSetting& batch, const STPWorldSetting& setting) :
	Setting(setting), Batch(batch) {
	const unsigned int map = setting.MapSize, nn = setting.ErosionNeighbour;
	if (nn == 0u || (nn & 1u) == 0u || nn > 9u) {
		STP_THROW(STPInvalidArgument, "the erosion neighbourhood is odd and at most 9 chunks per side");
	}
	if (map < 2u * rain.ErosionBrushRadius + 2u || map > 4096u || nn * map > 16384u) {
		STP_THROW(STPInvalidArgument, "the map size is in [2r + 2, 4096] and a neighbourhood at most 16384 pixels wide");
	}
	if (setting.GenerateChunks == 0u || setting.ErodeChunks == 0u || setting.GenerateChunks > 65535u
		|| setting.ErodeChunks > 65535u) {
		STP_THROW(STPInvalidArgument, "chunks per batch in [1, 65535]");
	}
	if (static_cast<int>(std::min(MaxSlots / (nn * nn), MaxMerged)) < 1) {
		STP_THROW(STPInvalidArgument, "the erosion neighbourhood is too large");
	}
	// heightmaps: the whole path without drops; erosion: binning, erosion and format of merged
	// neighbourhoods with drops in their centre map
	STPBatchSetting gen = batch;
	gen.DropsPerPixel = 0.0f;
	STPBatchCapacity gc;
	gc.Pixels = static_cast<unsigned long long>(setting.GenerateChunks) * (static_cast<unsigned long long>(map) * map + 4ull);
	gc.Drops = 1ull;
	gc.Chunks = setting.GenerateChunks;
	gc.MaxSide = map;
	this->Generator = std::make_unique<STPHeightfieldBatch>(noise, biome, rain, gen, gc);
	STPBatchCapacity ec;
	const unsigned long long side = static_cast<unsigned long long>(nn) * map;
	ec.Pixels = static_cast<unsigned long long>(setting.ErodeChunks) * (side * side + 4ull);
	ec.Drops = static_cast<unsigned long long>(static_cast<double>(setting.ErodeChunks) * map * map * batch.DropsPerPixel)
		+ setting.ErodeChunks;
	ec.Chunks = setting.ErodeChunks;
	ec.MaxSide = static_cast<unsigned int>(side);
	this->Eroder = std::make_unique<STPHeightfieldBatch>(noise, biome, rain, batch, ec);
	// a conflict-free group holds a few neighbourhoods, whose drops fill only the centre map's
	// tiles: the smallest tile spreads them over the most workgroups
	this->Eroder->setTuning(this->Eroder->latencyTuning());

	STP_CHECK_HIP(hipStreamCreateWithFlags(&this->Stream, hipStreamNonBlocking));
	int device = 0;
	STP_CHECK_HIP(hipGetDevice(&device));
	this->Device = device;
	hipMemPoolProps props {};
	props.allocType = hipMemAllocationTypePinned;
	props.location.type = hipMemLocationTypeDevice;
	props.location.id = device;
	STP_CHECK_HIP(hipMemPoolCreate(&this->Pool, &props));
}

STPHeightfieldWorld::~STPHeightfieldWorld() {
	if (this->Loader.joinable()) {
		this->Loader.join();
	}
	if (this->Stream) {
		(void)hipStreamSynchronize(this->Stream);
		for (auto& [k, e] : this->Chunk) {
			(void)hipFreeAsync(e.Height, this->Stream);
		}
		(void)hipStreamSynchronize(this->Stream);
		(void)hipStreamDestroy(this->Stream);
	}
	if (this->Pool) {
		(void)hipMemPoolDestroy(this->Pool);
	}
}

STPHeightfieldWorld::STPChunkState STPHeightfieldWorld::state(const int2 chunk) const {
	const std::lock_guard<std::recursive_mutex> guard(this->Lock);
	const auto it = this->Chunk.find(Key {chunk.x, chunk.y});
	return it == this->Chunk.end() ? STPChunkState::None : it->second.State;
}

void STPHeightfieldWorld::generate(const std::vector<Key>& chunks) {
	const unsigned int map = this->Setting.MapSize;
	for (size_t b = 0u; b < chunks.size(); b += this->Setting.GenerateChunks) {
		const size_t e = std::min(chunks.size(), b + this->Setting.GenerateChunks);
		std::vector<STPBatchChunk> in;
		std::vector<float*> height;
		std::vector<unsigned short*> fixed;
		for (size_t i = b; i < e; i++) {
			const Key k = chunks[i];
			// chunk (X, Y) starts at world pixel (X, Y) * MapSize (STPChunk.cpp:81-85, no map offset)
			in.push_back(STPBatchChunk {map, k.X * static_cast<int>(map), k.Y * static_cast<int>(map), 1u, 1u});
			Entry& en = this->Chunk[k];
			void* mem = nullptr;
			STP_CHECK_HIP(hipMallocFromPoolAsync(&mem, chunkBytes(map), this->Pool, this->Stream));
			en.Height = static_cast<float*>(mem);
			en.Fixed = reinterpret_cast<unsigned short*>(static_cast<unsigned char*>(mem) + heightBytes(map));
			en.State = STPChunkState::Heightmap;
			this->Resident += chunkBytes(map);
			height.push_back(en.Height);
			fixed.push_back(en.Fixed);
		}
		this->Generator->bind(in, {}, this->Stream);
		this->Generator->run(STPHeightfieldBatch::STPRunMode::Eager, this->Stream);
		this->Generator->copyOut(height, fixed, this->Stream);
	}
}

void STPHeightfieldWorld::erode(const std::vector<Key>& centres) {
	const unsigned int map = this->Setting.MapSize, nn = this->Setting.ErosionNeighbour, h = (nn - 1u) / 2u;
	const unsigned int side = nn * map, slots = nn * nn;
	const unsigned int perLaunch = std::min(MaxSlots / slots, MaxMerged);
	std::vector<STPBatchChunk> in;
	std::vector<unsigned int> stream;
	for (const Key& c : centres) {
		in.push_back(STPBatchChunk {side, (c.X - static_cast<int>(h)) * static_cast<int>(map),
			(c.Y - static_cast<int>(h)) * static_cast<int>(map), 1u, nn});
		stream.push_back(streamWord(make_int2(c.X, c.Y)));
	}
	this->Eroder->bind(in, stream, this->Stream);
	// neighbourhood i's chunks, row-major from (X - h, Y - h)
	const auto tables = [&](const size_t first, const size_t last) {
		NeighbourhoodTable t {};
		for (size_t i = first; i < last; i++) {
			const Key c = centres[i];
			for (unsigned int s = 0u; s < slots; s++) {
				const Key k {c.X - static_cast<int>(h) + static_cast<int>(s % nn), c.Y - static_cast<int>(h) + static_cast<int>(s / nn)};
				const Entry& en = this->Chunk.at(k);
				t.ChunkHeight[(i - first) * slots + s] = en.Height;
				t.ChunkFixed[(i - first) * slots + s] = en.Fixed;
			}
			t.MergedHeight[i - first] = this->Eroder->heightOf(static_cast<unsigned int>(i));
			t.MergedFixed[i - first] = this->Eroder->fixedOf(static_cast<unsigned int>(i));
		}
		return t;
	};
	const auto pass = [&](auto kernel) {
		for (size_t b = 0u; b < centres.size(); b += perLaunch) {
			const size_t e = std::min(centres.size(), b + static_cast<size_t>(perLaunch));
			const dim3 grid((side + 255u) / 256u, side, static_cast<unsigned int>(e - b));
			kernel<<<grid, 256, 0, this->Stream>>>(tables(b, e), map, nn);
			STP_CHECK_HIP(hipGetLastError());
		}
	};
	pass(neighbourhoodKERNEL<true>);
	// erosion binning, erosion and format of the merged maps
	this->Eroder->runStages(4, STPHeightfieldBatch::StageCount, this->Stream);
	pass(neighbourhoodKERNEL<false>);
	// a re-eroded chunk can pile heights up: a window cell out of range is an error, not a result
	this->Eroder->check(this->Stream);
	for (const Key& c : centres) {
		this->Chunk.at(c).State = STPChunkState::Complete;
	}
}

std::vector<int2> STPHeightfieldWorld::request(const std::vector<int2>& chunks) {
	const std::lock_guard<std::recursive_mutex> guard(this->Lock);
	const int h = static_cast<int>((this->Setting.ErosionNeighbour - 1u) / 2u);
	// every chunk of a neighbourhood (and its filter halo) keeps its world pixels within +-2^30, so
	// the 32-bit world origins of generate() and erode() cannot overflow
	const long long reach = (1ll << 30) / static_cast<long long>(this->Setting.MapSize);
	for (const int2 c : chunks) {
		if (std::llabs(static_cast<long long>(c.x)) + h >= reach || std::llabs(static_cast<long long>(c.y)) + h >= reach) {
			STP_THROW(STPInvalidArgument, "chunk (" + std::to_string(c.x) + ", " + std::to_string(c.y)
				+ ") or its erosion neighbourhood is out of the 32-bit world range");
		}
	}
	std::vector<Key> centres;
	std::unordered_set<Key, KeyHash> seen;
	for (const int2 c : chunks) {
		const Key k {c.x, c.y};
		if (this->state(c) != STPChunkState::Complete && seen.insert(k).second) {
			centres.push_back(k);
		}
	}
	if (centres.empty()) {
		return {};
	}
	// every chunk of every neighbourhood needs its heightmap
	std::vector<Key> missing, changed;
	seen.clear();
	for (const Key& c : centres) {
		for (int dy = -h; dy <= h; dy++) {
			for (int dx = -h; dx <= h; dx++) {
				const Key k {c.X + dx, c.Y + dy};
				if (seen.insert(k).second) {
					changed.push_back(k);
					if (this->state(make_int2(k.X, k.Y)) == STPChunkState::None) {
						missing.push_back(k);
					}
				}
			}
		}
	}
	this->generate(missing);
	// centres of one batch have disjoint neighbourhoods; batches run in order
	std::vector<int2> c2;
	for (const Key& c : centres) {
		c2.push_back(make_int2(c.X, c.Y));
	}
	const std::vector<unsigned int> batch = schedule(c2, this->Setting.ErosionNeighbour);
	const unsigned int batches = *std::max_element(batch.begin(), batch.end()) + 1u;
	for (unsigned int b = 0u; b < batches; b++) {
		std::vector<Key> group;
		for (size_t i = 0u; i < centres.size(); i++) {
			if (batch[i] == b) {
				group.push_back(centres[i]);
			}
		}
		for (size_t g = 0u; g < group.size(); g += this->Setting.ErodeChunks) {
			this->erode(std::vector<Key>(group.begin() + g,
				group.begin() + std::min(group.size(), g + this->Setting.ErodeChunks)));
		}
	}
	STP_CHECK_HIP(hipStreamSynchronize(this->Stream));
	std::vector<int2> out;
	for (const Key& k : changed) {
		out.push_back(make_int2(k.X, k.Y));
	}
	return out;
}

bool STPHeightfieldWorld::read(const int2 chunk, float* const height, unsigned short* const fixed) {
	const std::lock_guard<std::recursive_mutex> guard(this->Lock);
	const auto it = this->Chunk.find(Key {chunk.x, chunk.y});
	if (it == this->Chunk.end()) {
		return false;
	}
	const size_t px = static_cast<size_t>(this->Setting.MapSize) * this->Setting.MapSize;
	if (height) {
		STP_CHECK_HIP(hipMemcpyAsync(height, it->second.Height, px * sizeof(float), hipMemcpyDeviceToHost, this->Stream));
	}
	if (fixed) {
		STP_CHECK_HIP(hipMemcpyAsync(fixed, it->second.Fixed, px * sizeof(unsigned short), hipMemcpyDeviceToHost, this->Stream));
	}
	STP_CHECK_HIP(hipStreamSynchronize(this->Stream));
	return true;
}

void STPHeightfieldWorld::renderArea(const int2 centre, const unsigned int radius, unsigned short* const biome,
	unsigned short* const height) {
	const std::lock_guard<std::recursive_mutex> guard(this->Lock);
	if (!biome || !height) {
		STP_THROW(STPInvalidArgument, "null area buffer");
	}
	const unsigned int map = this->Setting.MapSize, n = 2u * radius + 1u;
	if (radius > 64u || static_cast<unsigned long long>(n) * map > 65536ull) {
		STP_THROW(STPInvalidArgument, "the rendered area is at most 129 chunks and 65536 pixels wide");
	}
	const int r = static_cast<int>(radius);
	std::vector<int2> chunks;
	for (int dy = -r; dy <= r; dy++) {
		for (int dx = -r; dx <= r; dx++) {
			chunks.push_back(make_int2(centre.x + dx, centre.y + dy));
		}
	}
	this->request(chunks);
	const size_t side = static_cast<size_t>(n) * map;
	for (size_t i = 0u; i < chunks.size(); i++) {
		const Entry& e = this->Chunk.at(Key {chunks[i].x, chunks[i].y});
		const size_t lx = i % n, ly = i / n;
		STP_CHECK_HIP(hipMemcpy2DAsync(height + ly * map * side + lx * map, side * sizeof(unsigned short), e.Fixed,
			map * sizeof(unsigned short), map * sizeof(unsigned short), map, hipMemcpyDeviceToDevice, this->Stream));
	}
	// the biome map the heightmap batch sampled: chunk (X, Y) starts at world pixel (X, Y) * MapSize
	const long long ox = static_cast<long long>(centre.x - r) * map, oy = static_cast<long long>(centre.y - r) * map;
	if (this->Batch.BiomeLayers) {
		if (ox < INT32_MIN || oy < INT32_MIN || ox + static_cast<long long>(side) > INT32_MAX
			|| oy + static_cast<long long>(side) > INT32_MAX) {
			STP_THROW(STPInvalidArgument, "the rendered area leaves the int coordinate range of the biome layers");
		}
		const int2 origin = make_int2(static_cast<int>(ox), static_cast<int>(oy));
		this->Batch.BiomeLayers->generateDevice(biome, &origin, 1u, static_cast<unsigned int>(side),
			static_cast<unsigned int>(side), side, side * side, this->Stream);
	} else {
		STPHeightfieldKernel::voronoiBiomemap(this->Batch.BiomeSeed, this->Batch.BiomeCount, this->Batch.BiomeCell, ox, oy,
			biome, static_cast<unsigned int>(side), static_cast<unsigned int>(side), static_cast<unsigned int>(side),
			this->Stream);
	}
	STP_CHECK_HIP(hipStreamSynchronize(this->Stream));
}


void STPHeightfieldWorld::joinLoader() {
	if (this->Loader.joinable()) {
		this->Loader.join();
	}
	if (this->LoaderError) {
		std::exception_ptr e = this->LoaderError;
		this->LoaderError = nullptr;
		std::rethrow_exception(e);
	}
}

STPHeightfieldWorld::STPWorldLoadStatus STPHeightfieldWorld::load(const int2 centre, const unsigned int radius) {
	if (this->LoaderState.load() == 1u) {
		// STPWorldPipeline.cpp:1510-1514: the loader works on the back buffer, the request is ignored
		return STPWorldLoadStatus::BackBufferBusy;
	}
	const std::lock_guard<std::recursive_mutex> guard(this->Lock);
	if (this->Pending && centre.x == this->PendingCentre.x && centre.y == this->PendingCentre.y
		&& radius == this->PendingRadius) {
		// :1529-1567: the same centre; a finished load is swapped in
		if (this->LoaderState.load() == 2u) {
			this->LoaderState = 0u;
			this->joinLoader();
			std::swap(this->Front, this->Back);
			return STPWorldLoadStatus::Swapped;
		}
		return STPWorldLoadStatus::Unchanged;
	}
	// :1519-1528, 1618-1633: a changed centre; a finished load that was not swapped is discarded
	if (this->LoaderState.load() == 2u) {
		this->LoaderState = 0u;
		this->joinLoader();
	}
	const unsigned int map = this->Setting.MapSize, n = 2u * radius + 1u;
	if (radius > 64u || static_cast<unsigned long long>(n) * map > 65536ull) {
		STP_THROW(STPInvalidArgument, "the rendered area is at most 129 chunks and 65536 pixels wide");
	}
	const unsigned int side = n * map;
	if (this->Back.Side != side) {
		const size_t px = static_cast<size_t>(side) * side;
		this->Back.Biome = STPSmartDeviceMemory::makeDevice<unsigned short>(px);
		this->Back.Height = STPSmartDeviceMemory::makeDevice<unsigned short>(px);
		this->Back.Side = side;
	}
	this->Pending = true;
	this->PendingCentre = centre;
	this->PendingRadius = radius;
	this->LoaderState = 1u;
	// the worker takes the lock once this call has released it, and holds it for the whole load
	this->Loader = std::thread([this, centre, radius]() {
		{
			const std::lock_guard<std::recursive_mutex> hold(this->Lock);
			try {
				STP_CHECK_HIP(hipSetDevice(this->Device));
				this->renderArea(centre, radius, this->Back.Biome.get(), this->Back.Height.get());
			} catch (...) {
				this->LoaderError = std::current_exception();
			}
		}
		this->LoaderState = 2u;
	});
	return STPWorldLoadStatus::BackBufferBusy;
}

void STPHeightfieldWorld::front(const unsigned short** const biome, const unsigned short** const height,
	unsigned int* const side) const {
	const std::lock_guard<std::recursive_mutex> guard(this->Lock);
	if (biome) {
		*biome = this->Front.Biome.get();
	}
	if (height) {
		*height = this->Front.Height.get();
	}
	if (side) {
		*side = this->Front.Side;
	}
}

unsigned long long STPHeightfieldWorld::exportRecords(const std::vector<int2>& chunks, const unsigned int maps,
	unsigned char* const out, const unsigned long long capacity) {
	const std::lock_guard<std::recursive_mutex> guard(this->Lock);
	namespace R = STPChunkRecord;
	if (maps == 0u || (maps & ~R::MapAll) != 0u) {
		STP_THROW(STPInvalidArgument, "maps: a non-empty combination of biome (1), FP32 height (2), UINT16 height (4)");
	}
	const unsigned int m = this->Setting.MapSize;
	if (m > R::MaxSide) {
		STP_THROW(STPInvalidArgument, "chunk records hold maps of at most 65536 pixels per side");
	}
	const R::Layout l(m, m, maps);
	const size_t n = chunks.size();
	const unsigned long long record = l.record(), total = record * n;
	if (n == 0u) {
		return 0u;
	}
	if (!out || total > capacity) {
		STP_THROW(STPInvalidArgument, "the export needs " + std::to_string(total) + " bytes of host memory");
	}
	// pack table: where each record's payload lies, and the chunk's resident maps
	const size_t entryBytes = n * sizeof(R::PackEntry), checkAt = (entryBytes + 255u) & ~size_t {255u},
		tableBytes = checkAt + n * sizeof(unsigned int);
	std::vector<const Entry*> entry(n);
	for (size_t i = 0u; i < n; i++) {
		const auto it = this->Chunk.find(Key {chunks[i].x, chunks[i].y});
		if (it == this->Chunk.end() || it->second.State == STPChunkState::None) {
			STP_THROW(STPInvalidArgument, "chunk (" + std::to_string(chunks[i].x) + ", " + std::to_string(chunks[i].y)
				+ ") has no map to export");
		}
		const long long ox = static_cast<long long>(chunks[i].x) * m, oy = static_cast<long long>(chunks[i].y) * m;
		if (ox < INT32_MIN || oy < INT32_MIN || ox + m - 1 > INT32_MAX || oy + m - 1 > INT32_MAX) {
			STP_THROW(STPInvalidArgument, "a chunk's world pixels leave the 32-bit range of the record");
		}
		entry[i] = &it->second;
	}
	if (total > this->ExportCapacity || tableBytes > this->ExportTableCapacity) {
		STP_CHECK_HIP(hipStreamSynchronize(this->Stream));
		if (total > this->ExportCapacity) {
			this->ExportCapacity = total + total / 4u;
			this->ExportStream = STPSmartDeviceMemory::makeDevice<unsigned char>(this->ExportCapacity);
		}
		if (tableBytes > this->ExportTableCapacity) {
			this->ExportTableCapacity = tableBytes + tableBytes / 2u;
			this->ExportTable = STPSmartDeviceMemory::makeDevice<unsigned char>(this->ExportTableCapacity);
			this->ExportStaging = STPSmartDeviceMemory::makeHost<unsigned char>(this->ExportTableCapacity);
		}
	}
	auto* const staged = reinterpret_cast<R::PackEntry*>(this->ExportStaging.get());
	for (size_t i = 0u; i < n; i++) {
		staged[i] = R::PackEntry {entry[i]->Height, entry[i]->Fixed, i * record + R::HeaderBytes, l.Height, l.Fixed, l.Payload};
	}
	unsigned char* const bytes = this->ExportStream.get();
	auto* const table = reinterpret_cast<R::PackEntry*>(this->ExportTable.get());
	auto* const check = reinterpret_cast<unsigned int*>(this->ExportTable.get() + checkAt);
	STP_CHECK_HIP(hipMemcpyAsync(table, staged, entryBytes, hipMemcpyHostToDevice, this->Stream));
	STP_CHECK_HIP(hipMemsetAsync(check, 0, n * sizeof(unsigned int), this->Stream));
	// padding is zero, so the checks do not depend on stale bytes
	STP_CHECK_HIP(hipMemsetAsync(bytes, 0, total, this->Stream));
	if (maps & R::MapBiome) {
		// the biome map the heightmap batch sampled: chunk (X, Y) starts at world pixel (X, Y) * MapSize
		if (this->Batch.BiomeLayers) {
			std::vector<STPDiversity::STPLayerMap> area(n);
			for (size_t i = 0u; i < n; i++) {
				area[i] = STPDiversity::STPLayerMap {make_int2(chunks[i].x * static_cast<int>(m), chunks[i].y * static_cast<int>(m)),
					make_uint2(m, m), 1u, m, (i * record + R::HeaderBytes + l.Biome) / 2u};
			}
			if (!this->ExportPlan) {
				this->ExportPlan = std::make_unique<STPDiversity::STPLayerPlan>();
			}
			this->Batch.BiomeLayers->plan(*this->ExportPlan, area, this->Stream);
			this->Batch.BiomeLayers->run(*this->ExportPlan, reinterpret_cast<unsigned short*>(bytes), 0u,
				static_cast<unsigned int>(n), this->Batch.BiomeCount, this->Stream);
		} else {
			for (size_t i = 0u; i < n; i++) {
				STPHeightfieldKernel::voronoiBiomemap(this->Batch.BiomeSeed, this->Batch.BiomeCount, this->Batch.BiomeCell,
					static_cast<long long>(chunks[i].x) * m, static_cast<long long>(chunks[i].y) * m,
					reinterpret_cast<unsigned short*>(bytes + i * record + R::HeaderBytes + l.Biome), m, m, m, this->Stream);
			}
		}
	}
	R::pack(table, static_cast<unsigned int>(n), static_cast<unsigned long long>(m) * m, maps, bytes, this->Stream);
	R::check(table, static_cast<unsigned int>(n), l.Payload, bytes, check, this->Stream);
	STP_CHECK_HIP(hipMemcpyAsync(out, bytes, total, hipMemcpyDeviceToHost, this->Stream));
	STP_CHECK_HIP(hipMemcpyAsync(this->ExportStaging.get(), check, n * sizeof(unsigned int), hipMemcpyDeviceToHost,
		this->Stream));
	STP_CHECK_HIP(hipStreamSynchronize(this->Stream));
	const auto* const sum = reinterpret_cast<const unsigned int*>(this->ExportStaging.get());
	for (size_t i = 0u; i < n; i++) {
		// our states: heightmap generated -> HeightmapReady (2), eroded and formatted -> Complete (3)
		const unsigned int completeness = entry[i]->State == STPChunkState::Complete ? 3u : 2u;
		R::writeHeader(out + i * record, chunks[i].x, chunks[i].y, chunks[i].x * static_cast<int>(m),
			chunks[i].y * static_cast<int>(m), m, m, completeness, maps, l.Payload, sum[i]);
	}
	return total;
}
