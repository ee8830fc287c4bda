// Serving loop over the batched heightfield path (declarations and rationale in stp_server.hpp).
//
// Scheduling:
//  * a request for a resident chunk completes at the next poll; a request for a chunk that a
//    running batch is producing waits for that batch;
//  * every idle engine takes the oldest queued requests that fit its capacity (pixels, chunks,
//    drops) and the memory budget, binds them, runs the whole path eagerly on its own stream and
//    copies each chunk's outputs into a resident buffer; engines run concurrently, so requests
//    that arrive while one batch runs join the next one (continuous batching);
//  * resident chunks are evicted least recently used first when a new chunk would exceed the
//    memory budget; a chunk that a batch is producing is never evicted.
// A chunk's drops use a Philox stream word made from its grid coordinate and level of detail, so
// its result does not depend on which batch it runs in.
#include "stp_server.hpp"

#include <algorithm>
#include <thread>

using namespace SuperTerrainPlus;

size_t STPHeightfieldServer::KeyHash::operator()(const Key& k) const noexcept {
	unsigned long long h = static_cast<unsigned int>(k.X) * 0x9E3779B97F4A7C15ull;
	h ^= (static_cast<unsigned long long>(static_cast<unsigned int>(k.Y)) + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2));
	h ^= (static_cast<unsigned long long>(k.Side) + 0x94D049BB133111EBull + (h << 6) + (h >> 2));
	return static_cast<size_t>(h);
}

unsigned int STPHeightfieldServer::streamWord(const int x, const int y, const unsigned int side) noexcept {
	unsigned int h = static_cast<unsigned int>(x) * 0x85EBCA6Bu;
	h ^= static_cast<unsigned int>(y) * 0xC2B2AE35u + 0x27D4EB2Fu + (h << 6) + (h >> 2);
	h ^= side * 0x165667B1u + (h << 6) + (h >> 2);
	h ^= h >> 16;
	h *= 0x7FEB352Du;
	h ^= h >> 15;
	return h;
}

namespace {
	bool powerOfTwo(const unsigned int v) { return v != 0u && (v & (v - 1u)) == 0u; }
} // namespace

unsigned long long STPHeightfieldServer::bytesOf(const unsigned int side) noexcept {
	const unsigned long long px = static_cast<unsigned long long>(side) * side;
	// FP32 heights, then UINT16, each part 256-byte aligned
	return ((px * sizeof(float) + 255ull) & ~255ull) + ((px * sizeof(unsigned short) + 255ull) & ~255ull);
}

STPHeightfieldServer::STPHeightfieldServer(const STPEnvironment::STPSimplexNoiseSetting& noise,
	const std::vector<STPBiomeProperty>& biome, const STPEnvironment::STPRainDropSetting& rain,
	const STPBatchSetting& batch, const STPServerSetting& setting) :
	Setting(setting), BatchSetting(batch) {
	if (setting.Engines == 0u || setting.Engines > 8u || setting.BatchChunks == 0u || setting.BatchChunks > 65535u
		|| setting.BatchPixels < static_cast<unsigned long long>(setting.MaxSide) * setting.MaxSide
		|| !powerOfTwo(setting.MaxSide) || !powerOfTwo(setting.ChunkWorldSize) || setting.ChunkWorldSize < setting.MaxSide) {
		STP_THROW(STPInvalidArgument,
			"server setting: 1-8 engines, 1-65535 chunks per batch, a batch must hold one largest chunk, "
			"MaxSide and ChunkWorldSize powers of two with MaxSide <= ChunkWorldSize");
	}
	// a pixel's histogram holds at most BiomeCount biomes, so it can never overflow and the
	// server needs no overflow readback between batches
	if (batch.BinCapacity < batch.BiomeCount) {
		STP_THROW(STPInvalidArgument, "the server needs BinCapacity >= BiomeCount");
	}
	// every chunk is produced with a free-slip halo the drops cannot cross within their lifetime
	// (no halo without a lifetime bound)
	this->ErosionHalo = rain.MaxLifetime > 0u ? rain.MaxLifetime + std::max(rain.ErosionBrushRadius, 2u) + 1u : 0u;
	if (static_cast<unsigned long long>(this->ErosionHalo) * setting.ChunkWorldSize >= (1ull << 30)) {
		STP_THROW(STPInvalidArgument, "the free-slip halo (MaxLifetime + max(r, 2) + 1 pixels) spans more than 2^30 "
			"world units at the coarsest level of detail");
	}
	const unsigned long long largest = static_cast<unsigned long long>(setting.MaxSide + 2u * this->ErosionHalo)
		* (setting.MaxSide + 2u * this->ErosionHalo);
	STPBatchCapacity cap;
	cap.Pixels = std::max(setting.BatchPixels, largest) + 4ull * setting.BatchChunks;
	cap.Drops = static_cast<unsigned long long>(static_cast<double>(setting.BatchPixels) * batch.DropsPerPixel) + setting.BatchChunks;
	cap.Chunks = setting.BatchChunks;
	cap.MaxSide = setting.MaxSide + 2u * this->ErosionHalo;

	size_t freeBefore = 0u, freeAfter = 0u, total = 0u;
	STP_CHECK_HIP(hipMemGetInfo(&freeBefore, &total));
	this->Engines.resize(setting.Engines);
	for (Engine& e : this->Engines) {
		e.Batch = std::make_unique<STPHeightfieldBatch>(noise, biome, rain, batch, cap);
		// a serving batch holds a few chunks: the smallest erosion tile spreads them over the most
		// workgroups
		e.Batch->setTuning(e.Batch->latencyTuning());
		// default priority, non-blocking with respect to the null stream
		STP_CHECK_HIP(hipStreamCreateWithFlags(&e.Stream, hipStreamNonBlocking));
		STP_CHECK_HIP(hipEventCreateWithFlags(&e.Done, hipEventDisableTiming));
	}
	STP_CHECK_HIP(hipMemGetInfo(&freeAfter, &total));
	this->EngineBytes = freeBefore > freeAfter ? freeBefore - freeAfter : 0u;
	if (this->EngineBytes >= setting.MemoryBudget) {
		STP_THROW(STPInvalidArgument, "the memory budget does not even hold the batch engines");
	}

	int device = 0;
	STP_CHECK_HIP(hipGetDevice(&device));
	hipMemPoolProps props {};
	props.allocType = hipMemAllocationTypePinned;
	props.location.type = hipMemLocationTypeDevice;
	props.location.id = device;
	STP_CHECK_HIP(hipMemPoolCreate(&this->Pool, &props));
	// keep freed chunk buffers in the pool up to the budget, so steady-state allocation is cheap
	unsigned long long threshold = setting.MemoryBudget;
	STP_CHECK_HIP(hipMemPoolSetAttribute(this->Pool, hipMemPoolAttrReleaseThreshold, &threshold));
}

STPHeightfieldServer::~STPHeightfieldServer() {
	for (Engine& e : this->Engines) {
		if (e.Stream) {
			(void)hipStreamSynchronize(e.Stream);
		}
	}
	if (!this->Engines.empty() && this->Engines.front().Stream) {
		for (auto& [k, e] : this->Cache) {
			(void)hipFreeAsync(e.Height, this->Engines.front().Stream);
		}
		(void)hipStreamSynchronize(this->Engines.front().Stream);
	}
	for (Engine& e : this->Engines) {
		if (e.Done) {
			(void)hipEventDestroy(e.Done);
		}
		if (e.Stream) {
			(void)hipStreamDestroy(e.Stream);
		}
	}
	if (this->Pool) {
		(void)hipMemPoolDestroy(this->Pool);
	}
}

unsigned long long STPHeightfieldServer::submit(const STPChunkRequest& request) {
	const unsigned int side = request.Side;
	if (!powerOfTwo(side) || side > this->Setting.MaxSide || side < 2u * this->BatchSetting.Radius
		|| side < 16u) {
		STP_THROW(STPInvalidArgument, "chunk side must be a power of two in [max(16, 2 r), MaxSide]");
	}
	// world origins are 32-bit
	const long long reach = (1ll << 30) / static_cast<long long>(this->Setting.ChunkWorldSize);
	if (request.X <= -reach || request.X >= reach || request.Y <= -reach || request.Y >= reach) {
		STP_THROW(STPInvalidArgument, "chunk coordinate out of the 32-bit world range");
	}
	const unsigned long long ticket = this->NextTicket++;
	this->Queue.push_back(Pending {ticket, Key {request.X, request.Y, side}, Clock::now()});
	return ticket;
}

void STPHeightfieldServer::complete(Entry& e, const Clock::time_point now) {
	for (const Entry::Waiter& w : e.Waiting) {
		this->Completed.push_back(
			STPServerCompletion {w.Ticket, std::chrono::duration<double, std::milli>(now - w.Arrival).count(), w.Generated});
	}
	e.Waiting.clear();
}

bool STPHeightfieldServer::makeRoom(const unsigned long long bytes) {
	const unsigned long long budget = this->Setting.MemoryBudget - this->EngineBytes;
	if (bytes > budget) {
		return false;
	}
	auto it = this->Lru.end();
	while (this->Resident + bytes > budget && it != this->Lru.begin()) {
		--it;
		auto found = this->Cache.find(*it);
		if (!found->second.Ready) {
			continue; // being produced
		}
		// stream-ordered free: every batch that wrote or read this buffer has completed
		STP_CHECK_HIP(hipFreeAsync(found->second.Height, this->Engines.front().Stream));
		this->Resident -= found->second.Bytes;
		this->Evictions++;
		this->Cache.erase(found);
		it = this->Lru.erase(it);
	}
	return this->Resident + bytes <= budget;
}

void STPHeightfieldServer::retire() {
	const Clock::time_point now = Clock::now();
	for (Engine& e : this->Engines) {
		if (!e.Busy) {
			continue;
		}
		const hipError_t q = hipEventQuery(e.Done);
		if (q == hipErrorNotReady) {
			(void)hipGetLastError();
			continue;
		}
		STP_CHECK_HIP(q);
		for (const Key& k : e.Chunks) {
			Entry& entry = this->Cache.at(k);
			entry.Ready = true;
			this->complete(entry, now);
		}
		e.Chunks.clear();
		e.Busy = false;
	}
}

void STPHeightfieldServer::launch() {
	const Clock::time_point now = Clock::now();
	// resident or in-production chunks: complete now, or wait for their batch
	for (auto it = this->Queue.begin(); it != this->Queue.end();) {
		const auto found = this->Cache.find(it->K);
		if (found == this->Cache.end()) {
			++it;
			continue;
		}
		Entry& e = found->second;
		this->Lru.splice(this->Lru.begin(), this->Lru, e.Lru);
		e.Waiting.push_back(Entry::Waiter {it->Ticket, it->Arrival, false});
		if (e.Ready) {
			this->Hits++;
			this->complete(e, now);
		}
		it = this->Queue.erase(it);
	}
	// new chunks, oldest first, onto idle engines
	for (Engine& engine : this->Engines) {
		if (engine.Busy || this->Queue.empty()) {
			continue;
		}
		const STPBatchCapacity& cap = engine.Batch->capacity();
		std::vector<STPBatchChunk> chunks;
		std::vector<unsigned int> streamId;
		std::vector<float*> outHeight;
		std::vector<unsigned short*> outFixed;
		unsigned long long pixels = 0ull, drops = 0ull;
		while (!this->Queue.empty() && chunks.size() < cap.Chunks) {
			const Pending p = this->Queue.front();
			const auto found = this->Cache.find(p.K);
			if (found != this->Cache.end()) {
				// the same chunk asked twice in this batch
				found->second.Waiting.push_back(Entry::Waiter {p.Ticket, p.Arrival, false});
				this->Queue.pop_front();
				continue;
			}
			// the batch produces the chunk with its halo (`produced` pixels per side); the resident
			// entry holds the chunk alone (`centre` pixels), as bytesOf lays it out
			const unsigned int halo = this->ErosionHalo, produced = p.K.Side + 2u * halo;
			const unsigned long long px = static_cast<unsigned long long>(produced) * produced,
				centre = static_cast<unsigned long long>(p.K.Side) * p.K.Side;
			const unsigned long long d = static_cast<unsigned long long>(static_cast<double>(centre) * this->BatchSetting.DropsPerPixel);
			if (pixels + ((px + 3ull) & ~3ull) > cap.Pixels || drops + d > cap.Drops) {
				break;
			}
			const unsigned long long bytes = bytesOf(p.K.Side);
			if (!this->makeRoom(bytes)) {
				break; // the budget is held by chunks in production: wait for them
			}
			void* mem = nullptr;
			STP_CHECK_HIP(hipMallocFromPoolAsync(&mem, bytes, this->Pool, engine.Stream));
			Entry e;
			e.Height = static_cast<float*>(mem);
			e.Fixed = reinterpret_cast<unsigned short*>(static_cast<unsigned char*>(mem)
				+ ((centre * sizeof(float) + 255ull) & ~255ull));
			e.Bytes = bytes;
			e.Waiting.push_back(Entry::Waiter {p.Ticket, p.Arrival, true});
			this->Lru.push_front(p.K);
			e.Lru = this->Lru.begin();
			this->Resident += bytes;
			this->PeakResident = std::max(this->PeakResident, this->Resident);
			outHeight.push_back(e.Height);
			outFixed.push_back(e.Fixed);
			this->Cache.emplace(p.K, std::move(e));

			const unsigned int step = this->Setting.ChunkWorldSize / p.K.Side;
			// submit() keeps |X|, |Y| below 2^30 / ChunkWorldSize and the constructor the halo's world
			// units (halo * step <= halo * ChunkWorldSize) below 2^30, so the origins fit in 32 bits
			const long long world = static_cast<long long>(this->Setting.ChunkWorldSize),
				back = static_cast<long long>(halo) * step;
			// the chunk with its halo: the centre is bit-exact with the chunk alone before erosion
			// (same sample coordinates), and erodes as the centre of its uneroded neighbours
			chunks.push_back(STPBatchChunk {produced, static_cast<int>(p.K.X * world - back),
				static_cast<int>(p.K.Y * world - back), step, 1u, halo});
			streamId.push_back(streamWord(p.K.X, p.K.Y, p.K.Side));
			engine.Chunks.push_back(p.K);
			pixels += (px + 3ull) & ~3ull;
			drops += d;
			this->Queue.pop_front();
		}
		if (chunks.empty()) {
			continue;
		}
		engine.Batch->bind(chunks, streamId, engine.Stream);
		engine.Batch->run(STPHeightfieldBatch::STPRunMode::Eager, engine.Stream);
		engine.Batch->copyOut(outHeight, outFixed, engine.Stream);
		STP_CHECK_HIP(hipEventRecord(engine.Done, engine.Stream));
		engine.Busy = true;
		this->Batches++;
		this->BatchedChunks += chunks.size();
	}
}

void STPHeightfieldServer::poll() {
	this->retire();
	this->launch();
}

std::vector<STPServerCompletion> STPHeightfieldServer::drain() {
	std::vector<STPServerCompletion> out;
	out.swap(this->Completed);
	return out;
}

bool STPHeightfieldServer::read(const STPChunkRequest& request, float* const height, unsigned short* const fixed) {
	const auto found = this->Cache.find(Key {request.X, request.Y, request.Side});
	if (found == this->Cache.end() || !found->second.Ready) {
		return false;
	}
	const unsigned long long px = static_cast<unsigned long long>(request.Side) * request.Side;
	if (height) {
		STP_CHECK_HIP(hipMemcpy(height, found->second.Height, px * sizeof(float), hipMemcpyDeviceToHost));
	}
	if (fixed) {
		STP_CHECK_HIP(hipMemcpy(fixed, found->second.Fixed, px * sizeof(unsigned short), hipMemcpyDeviceToHost));
	}
	return true;
}

void STPHeightfieldServer::flush() {
	for (;;) {
		this->poll();
		bool busy = !this->Queue.empty();
		for (const Engine& e : this->Engines) {
			busy = busy || e.Busy;
		}
		if (!busy) {
			return;
		}
		std::this_thread::yield();
	}
}

STPServerStats STPHeightfieldServer::replay(const std::vector<STPChunkRequest>& request,
	const std::vector<double>& arrival_s, const double deadline_ms, std::vector<double>& latency_ms) {
	if (request.size() != arrival_s.size()) {
		STP_THROW(STPInvalidArgument, "one arrival time per request");
	}
	for (size_t i = 1u; i < arrival_s.size(); i++) {
		if (arrival_s[i] < arrival_s[i - 1u]) {
			STP_THROW(STPInvalidArgument, "arrival times must not decrease");
		}
	}
	const size_t n = request.size();
	latency_ms.assign(n, -1.0);
	std::vector<char> generated(n, 0);
	std::unordered_map<unsigned long long, size_t> index;
	const unsigned long long hits0 = this->Hits, batches0 = this->Batches, chunks0 = this->BatchedChunks,
		evict0 = this->Evictions;
	this->PeakResident = this->Resident;
	const Clock::time_point t0 = Clock::now();
	size_t next = 0u, done = 0u;
	unsigned long long pixels = 0ull, generatedPixels = 0ull;
	Clock::time_point last = t0;
	while (done < n) {
		const Clock::time_point now = Clock::now();
		const double elapsed = std::chrono::duration<double>(now - t0).count();
		// every request that is due, stamped with its scheduled arrival
		while (next < n && arrival_s[next] <= elapsed) {
			const STPChunkRequest& r = request[next];
			const unsigned long long ticket = this->submit(r);
			this->Queue.back().Arrival =
				t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[next]));
			index.emplace(ticket, next);
			next++;
		}
		this->poll();
		for (const STPServerCompletion& c : this->drain()) {
			const auto it = index.find(c.Ticket);
			if (it == index.end()) {
				continue; // submitted outside this replay
			}
			latency_ms[it->second] = c.Ms;
			const unsigned long long px = static_cast<unsigned long long>(request[it->second].Side) * request[it->second].Side;
			pixels += px;
			if (c.Generated) {
				generated[it->second] = 1;
				generatedPixels += px;
			}
			done++;
			last = Clock::now();
		}
		if (done < n) {
			std::this_thread::yield();
		}
	}
	STPServerStats s {};
	std::vector<double> sorted(latency_ms), sortedGenerated;
	for (size_t i = 0u; i < n; i++) {
		if (generated[i]) {
			sortedGenerated.push_back(latency_ms[i]);
		}
	}
	std::sort(sorted.begin(), sorted.end());
	std::sort(sortedGenerated.begin(), sortedGenerated.end());
	// linear interpolation between closest ranks
	const auto pct = [](const std::vector<double>& v, const double q) {
		const double pos = q * static_cast<double>(v.size() - 1u);
		const size_t lo = static_cast<size_t>(pos);
		const size_t hi = std::min(lo + 1u, v.size() - 1u);
		return v[lo] + (v[hi] - v[lo]) * (pos - static_cast<double>(lo));
	};
	if (!sortedGenerated.empty()) {
		s.GeneratedP50 = pct(sortedGenerated, 0.5);
		s.GeneratedP99 = pct(sortedGenerated, 0.99);
		s.GeneratedP999 = pct(sortedGenerated, 0.999);
		s.GeneratedMax = sortedGenerated.back();
	}
	s.Generated = sortedGenerated.size();
	if (!sorted.empty()) {
		s.P50 = pct(sorted, 0.5);
		s.P99 = pct(sorted, 0.99);
		s.P999 = pct(sorted, 0.999);
		s.Max = sorted.back();
		double sum = 0.0;
		size_t miss = 0u;
		for (const double v : sorted) {
			sum += v;
			miss += v > deadline_ms ? 1u : 0u;
		}
		s.Mean = sum / static_cast<double>(sorted.size());
		s.DeadlineMissFraction = static_cast<double>(miss) / static_cast<double>(sorted.size());
	}
	const double wall = std::chrono::duration<double>(last - t0).count();
	s.MpixelPerSecond = wall > 0.0 ? static_cast<double>(pixels) / wall / 1e6 : 0.0;
	s.GeneratedMpixelPerSecond = wall > 0.0 ? static_cast<double>(generatedPixels) / wall / 1e6 : 0.0;
	s.Requests = n;
	s.CacheHits = this->Hits - hits0;
	s.Batches = this->Batches - batches0;
	s.Evictions = this->Evictions - evict0;
	s.MeanBatchChunks = s.Batches ? static_cast<double>(this->BatchedChunks - chunks0) / static_cast<double>(s.Batches) : 0.0;
	s.PeakResidentBytes = this->PeakResident;
	return s;
}
