// Serving loop over the batched heightfield path: chunk requests arrive one by one, results stay
// resident in device memory, and requests that arrive while batches are running join the next
// batch (continuous batching).
//
// No reference counterpart. The nearest relative is the reference world pipeline's asynchronous
// chunk loader (SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:1504-1633), which
// hands chunks to a 5-thread pool that calls STPHeightfieldGenerator one chunk at a time, and its
// 100 MB memory pool (SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPHeightfieldGenerator.cpp:148).
#pragma once
#ifndef STP_AMD_SERVER_HPP
#define STP_AMD_SERVER_HPP

#include "stp_batch.hpp"

#include <chrono>
#include <deque>
#include <list>
#include <memory>
#include <unordered_map>
#include <vector>

namespace SuperTerrainPlus {

	struct STPServerSetting {
		unsigned long long MemoryBudget; // device bytes for the batch engines plus the resident chunks
		unsigned long long BatchPixels;  // pixel capacity of one batch
		unsigned int BatchChunks;        // chunk capacity of one batch
		unsigned int MaxSide;            // largest chunk side (pixels)
		unsigned int Engines;            // batches in flight at once
		unsigned int ChunkWorldSize;     // world units of one grid chunk at every level of detail
	};

	// Chunk (X, Y) of the grid at `Side` pixels per side: world origin (X, Y) * ChunkWorldSize,
	// ChunkWorldSize / Side world units per pixel.
	struct STPChunkRequest {
		int X, Y;
		unsigned int Side;
	};

	struct STPServerStats {
		double P50, P99, P999, Mean, Max; // request latency, milliseconds
		double MpixelPerSecond;           // pixels of completed requests / wall time of the replay
		double DeadlineMissFraction;      // requests completed after the deadline
		unsigned long long Requests, CacheHits, Batches, Evictions;
		double MeanBatchChunks;
		unsigned long long PeakResidentBytes;
		// the requests whose chunk was generated for them (not resident, not already in production)
		unsigned long long Generated;
		double GeneratedP50, GeneratedP99, GeneratedP999, GeneratedMax; // their latency, milliseconds
		double GeneratedMpixelPerSecond;  // generated pixels / wall time of the replay
	};

	// A completed request: its ticket, latency and whether its chunk was generated for it.
	struct STPServerCompletion {
		unsigned long long Ticket;
		double Ms;
		bool Generated;
	};

	class STPHeightfieldServer {
	public:
		STPHeightfieldServer(const STPEnvironment::STPSimplexNoiseSetting& noise, const std::vector<STPBiomeProperty>& biome,
			const STPEnvironment::STPRainDropSetting& rain, const STPBatchSetting& batch, const STPServerSetting& setting);
		STPHeightfieldServer(const STPHeightfieldServer&) = delete;
		STPHeightfieldServer& operator=(const STPHeightfieldServer&) = delete;
		~STPHeightfieldServer();

		// Queues a request (arrival = now); returns its ticket. Requests for a resident chunk complete
		// at the next poll.
		unsigned long long submit(const STPChunkRequest& request);
		// One scheduler pass, never blocks on the GPU: retires finished batches, admits queued
		// requests within the memory budget (evicting least recently used chunks) and launches them
		// on idle engines.
		void poll();
		// Tickets completed since the last call, with their latency in milliseconds.
		std::vector<STPServerCompletion> drain();
		// Copies a resident chunk to host memory (either pointer may be null); false if the chunk is
		// not resident. Synchronous.
		bool read(const STPChunkRequest& request, float* height, unsigned short* fixed);
		// Waits until every queued request has completed.
		void flush();
		// Replays a trace in real time: request i is submitted `arrival_s[i]` seconds after the start.
		// latency_ms[i] receives each request's latency.
		STPServerStats replay(const std::vector<STPChunkRequest>& request, const std::vector<double>& arrival_s,
			double deadline_ms, std::vector<double>& latency_ms);

		// Philox stream word of a chunk's drops: a hash of (X, Y, Side), independent of the batch it
		// runs in.
		static unsigned int streamWord(int x, int y, unsigned int side) noexcept;

		unsigned long long residentBytes() const noexcept { return this->Resident; }
		unsigned long long engineBytes() const noexcept { return this->EngineBytes; }

	private:
		using Clock = std::chrono::steady_clock;
		struct Key {
			int X, Y;
			unsigned int Side;
			bool operator==(const Key& k) const noexcept { return X == k.X && Y == k.Y && Side == k.Side; }
		};
		struct KeyHash {
			size_t operator()(const Key& k) const noexcept;
		};
		struct Entry {
			float* Height = nullptr;
			unsigned short* Fixed = nullptr;
			unsigned long long Bytes = 0ull;
			bool Ready = false;
			std::list<Key>::iterator Lru;
			struct Waiter {
				unsigned long long Ticket;
				Clock::time_point Arrival;
				bool Generated; // the request this chunk is produced for
			};
			std::vector<Waiter> Waiting;
		};
		struct Engine {
			std::unique_ptr<STPHeightfieldBatch> Batch;
			hipStream_t Stream = nullptr;
			hipEvent_t Done = nullptr;
			std::vector<Key> Chunks;
			bool Busy = false;
		};
		struct Pending {
			unsigned long long Ticket;
			Key K;
			Clock::time_point Arrival;
		};

		STPServerSetting Setting;
		STPBatchSetting BatchSetting;
		std::vector<Engine> Engines;
		std::unordered_map<Key, Entry, KeyHash> Cache;
		std::list<Key> Lru; // front = most recently used
		std::deque<Pending> Queue;
		std::vector<STPServerCompletion> Completed;
		hipMemPool_t Pool = nullptr;
		unsigned long long NextTicket = 0ull, Resident = 0ull, PeakResident = 0ull, EngineBytes = 0ull;
		unsigned long long Hits = 0ull, Batches = 0ull, Evictions = 0ull, BatchedChunks = 0ull;
		// pixels of free-slip halo every chunk is produced with (MaxLifetime + max(r, 2) + 1; 0 without
		// a lifetime bound)
		unsigned int ErosionHalo = 0u;

		static unsigned long long bytesOf(unsigned int side) noexcept;
		void complete(Entry& e, Clock::time_point now);
		bool makeRoom(unsigned long long bytes);
		void retire();
		void launch();
	};

} // namespace SuperTerrainPlus

#endif // STP_AMD_SERVER_HPP
