// Host classes of the path: nearest-neighbour texture buffer, single histogram filter front-end,
// heightfield generator and the two built-in diversity generators.
#include "stp_batch.hpp"
#include "stp_internal.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPAlgorithm;
using SuperTerrainPlus::STPHeightfieldKernel::STPFilterWorkspace;

// ==================================================================== nearest-neighbour buffer
// Behaviour: STPNearestNeighbourTextureBuffer.cpp:18-151. Packing uses the device copy engines
// (pitched 2-D copies straight between each chunk and its slot in the merged map) or, when the
// chunks are already in HBM, one gather/scatter kernel; a 1x1 neighbourhood of device-resident
// chunks is used in place without any copy.
namespace {
	template <typename T>
	unsigned int neighbourCount(const STPNearestNeighbourInformation& nn) {
		return nn.ChunkNearestNeighbour.x * nn.ChunkNearestNeighbour.y;
	}
}

template <typename T, STPNearestNeighbourTextureBufferMemoryMode MM>
STPNearestNeighbourTextureBuffer<T, MM>::STPNearestNeighbourTextureBuffer(TextureType* const* const texture,
	const STPNearestNeighbourInformation& nn_info, const STPDeviceMemoryOperator& mem, const STPChunkLocation chunk_location) :
	NeighbourInfo(nn_info), DeviceMemInfo(mem), ChunkLocation(chunk_location), NeighbourTexture(texture) {
	if (!texture) {
		STP_THROW(STPInvalidArgument, "null texture pointer array");
	}
}

template <typename T, STPNearestNeighbourTextureBufferMemoryMode MM>
STPNearestNeighbourTextureBuffer<T, MM>::STPMergedBuffer::STPMergedBuffer(const STPNearestNeighbourTextureBuffer& buffer,
	const STPMemoryLocation location) : Location(location), Main(buffer) {
	const STPNearestNeighbourInformation& nn = this->Main.NeighbourInfo;
	const size_t bytes = static_cast<size_t>(nn.TotalMapSize.x) * nn.TotalMapSize.y * sizeof(T);
	const auto [pool, stream] = this->Main.DeviceMemInfo;
	const bool aliased = this->Location == STPMemoryLocation::DeviceMemory && this->Main.ChunkLocation == STPChunkLocation::Device
		&& neighbourCount<T>(nn) == 1u;
	const auto deviceBuffer = [&]() {
		void* p = nullptr;
		if (pool) {
			STP_CHECK_HIP(hipMallocFromPoolAsync(&p, bytes, pool, stream));
		} else {
			STP_CHECK_HIP(hipMallocAsync(&p, bytes, stream));
		}
		return static_cast<MutableT*>(p);
	};
	if (this->Location == STPMemoryLocation::HostMemory) {
		this->HostMem = static_cast<MutableT*>(STPSmartDeviceMemory::allocatePinned(bytes));
	} else if (aliased) {
		this->DeviceMem = const_cast<MutableT*>(this->Main.NeighbourTexture[0]);
	} else {
		this->DeviceMem = deviceBuffer();
		if (this->Main.ChunkLocation == STPChunkLocation::Host) {
			this->Staging = deviceBuffer();
		}
	}
	if constexpr (MM != STPNearestNeighbourTextureBufferMemoryMode::WriteOnly) {
		if (!aliased) {
			this->copy(true);
		}
	}
}

template <typename T, STPNearestNeighbourTextureBufferMemoryMode MM>
STPNearestNeighbourTextureBuffer<T, MM>::STPMergedBuffer::~STPMergedBuffer() noexcept(false) {
	const STPNearestNeighbourInformation& nn = this->Main.NeighbourInfo;
	const hipStream_t stream = this->Main.DeviceMemInfo.second;
	const bool aliased = this->Location == STPMemoryLocation::DeviceMemory && this->Main.ChunkLocation == STPChunkLocation::Device
		&& neighbourCount<T>(nn) == 1u;
	hipError_t err = hipSuccess;
	std::string what;
	try {
		if constexpr (MM != STPNearestNeighbourTextureBufferMemoryMode::ReadOnly) {
			if (!aliased) {
				this->copy(false);
			}
		}
	} catch (const std::exception& e) {
		what = e.what();
	}
	if (this->Location == STPMemoryLocation::HostMemory) {
		err = hipStreamSynchronize(stream);
		(void)hipHostFree(this->HostMem);
	} else {
		if (!aliased) {
			const hipError_t e = hipFreeAsync(this->DeviceMem, stream);
			err = err == hipSuccess ? e : err;
		}
		if (this->Staging) {
			const hipError_t e = hipFreeAsync(this->Staging, stream);
			err = err == hipSuccess ? e : err;
		}
		if (this->Main.ChunkLocation == STPChunkLocation::Host) {
			// host chunks must hold the result when the buffer dies (reference semantics)
			const hipError_t e = hipStreamSynchronize(stream);
			err = err == hipSuccess ? e : err;
		}
	}
	if (!what.empty() && std::uncaught_exceptions() == 0) {
		STP_THROW(STPHIPError, what);
	}
	if (err != hipSuccess && std::uncaught_exceptions() == 0) {
		STP_THROW(STPHIPError, std::string("merged buffer release: ") + hipGetErrorString(err));
	}
}

template <typename T, STPNearestNeighbourTextureBufferMemoryMode MM>
void STPNearestNeighbourTextureBuffer<T, MM>::STPMergedBuffer::copy(const bool pack) {
	const STPNearestNeighbourInformation& nn = this->Main.NeighbourInfo;
	const hipStream_t stream = this->Main.DeviceMemInfo.second;
	const unsigned int count = neighbourCount<T>(nn);
	const bool deviceChunks = this->Main.ChunkLocation == STPChunkLocation::Device;
	if (this->Location == STPMemoryLocation::DeviceMemory && deviceChunks) {
		if (pack) {
			STPHeightfieldKernel::mergeNeighbour(reinterpret_cast<const void* const*>(this->Main.NeighbourTexture),
				this->DeviceMem, sizeof(T), nn, stream);
		} else {
			STPHeightfieldKernel::unmergeNeighbour(this->DeviceMem,
				reinterpret_cast<void* const*>(const_cast<MutableT* const*>(this->Main.NeighbourTexture)), sizeof(T), nn,
				stream);
		}
		return;
	}
	if (this->Location == STPMemoryLocation::DeviceMemory) {
		// host chunks: one contiguous copy per chunk between the chunk and its slot in the staging
		// buffer, and the device gather / scatter between the staging buffer and the merged map (a
		// pitched copy from pageable host memory moves row by row)
		const size_t elements = static_cast<size_t>(nn.MapSize.x) * nn.MapSize.y;
		std::vector<void*> slot(count);
		for (unsigned int i = 0u; i < count; i++) {
			slot[i] = this->Staging + i * elements;
		}
		if (pack) {
			for (unsigned int i = 0u; i < count; i++) {
				STP_CHECK_HIP(hipMemcpyAsync(slot[i], this->Main.NeighbourTexture[i], elements * sizeof(T),
					hipMemcpyHostToDevice, stream));
			}
			STPHeightfieldKernel::mergeNeighbour(slot.data(), this->DeviceMem, sizeof(T), nn, stream);
		} else {
			STPHeightfieldKernel::unmergeNeighbour(this->DeviceMem, slot.data(), sizeof(T), nn, stream);
			for (unsigned int i = 0u; i < count; i++) {
				STP_CHECK_HIP(hipMemcpyAsync(const_cast<MutableT*>(this->Main.NeighbourTexture[i]), slot[i],
					elements * sizeof(T), hipMemcpyDeviceToHost, stream));
			}
		}
		return;
	}
	MutableT* const merged = this->HostMem;
	const size_t row = nn.MapSize.x * sizeof(T), total_row = row * nn.ChunkNearestNeighbour.x;
	for (unsigned int i = 0u; i < count; i++) {
		const uint2 local = STPChunk::calcLocalChunkCoordinate(i, nn.ChunkNearestNeighbour);
		const size_t offset = static_cast<size_t>(local.x) * nn.MapSize.x
			+ static_cast<size_t>(local.y) * nn.MapSize.y * nn.TotalMapSize.x;
		MutableT* const chunk = const_cast<MutableT*>(this->Main.NeighbourTexture[i]);
		if (pack) {
			STP_CHECK_HIP(hipMemcpy2DAsync(merged + offset, total_row, chunk, row, row, nn.MapSize.y, hipMemcpyDefault, stream));
		} else {
			STP_CHECK_HIP(hipMemcpy2DAsync(chunk, row, merged + offset, total_row, row, nn.MapSize.y, hipMemcpyDefault, stream));
		}
	}
	if (pack && this->Location == STPMemoryLocation::HostMemory) {
		// host memory is usable right after construction (reference semantics)
		STP_CHECK_HIP(hipStreamSynchronize(stream));
	}
}

template <typename T, STPNearestNeighbourTextureBufferMemoryMode MM>
typename STPNearestNeighbourTextureBuffer<T, MM>::STPMergedBuffer::MutableT*
STPNearestNeighbourTextureBuffer<T, MM>::STPMergedBuffer::getHost() const noexcept {
	return this->HostMem;
}

template <typename T, STPNearestNeighbourTextureBufferMemoryMode MM>
typename STPNearestNeighbourTextureBuffer<T, MM>::STPMergedBuffer::MutableT*
STPNearestNeighbourTextureBuffer<T, MM>::STPMergedBuffer::getDevice() const noexcept {
	return this->DeviceMem;
}

template class SuperTerrainPlus::STPNearestNeighbourTextureBuffer<STPHeightFloat_t, STPNearestNeighbourTextureBufferMemoryMode::WriteOnly>;
template class SuperTerrainPlus::STPNearestNeighbourTextureBuffer<STPHeightFloat_t, STPNearestNeighbourTextureBufferMemoryMode::ReadWrite>;
template class SuperTerrainPlus::STPNearestNeighbourTextureBuffer<STPSample_t, STPNearestNeighbourTextureBufferMemoryMode::ReadOnly>;
template class SuperTerrainPlus::STPNearestNeighbourTextureBuffer<STPHeightFixed_t, STPNearestNeighbourTextureBufferMemoryMode::WriteOnly>;

// ==================================================================== single histogram filter
struct STPSingleHistogramFilter::STPFilterBuffer::STPImpl {
	STPExecutionType Type;
	unsigned int Capacity;
	STPSmartDeviceMemory::STPDevice<char> Workspace;
	size_t WorkspaceBytes = 0u;
	STPFilterWorkspace Carved {};
	STPSmartDeviceMemory::STPDevice<STPSample_t> Staging;
	size_t StagingCount = 0u;
	STPSmartDeviceMemory::STPHost<STPSingleHistogram::STPBin> HostBin;
	size_t HostBinCapacity = 0u;
	STPSmartDeviceMemory::STPHost<unsigned int> HostOffset;
	size_t HostOffsetCapacity = 0u;
	STPSmartDeviceMemory::STPHost<unsigned int> Flag; // {overflow, bin total}
	STPSingleHistogram Result {nullptr, nullptr};
	size_t BinCount = 0u, OffsetCount = 0u;
	hipStream_t OwnStream = nullptr;

	~STPImpl() {
		if (this->OwnStream) {
			(void)hipStreamDestroy(this->OwnStream);
		}
	}
};

STPSingleHistogramFilter::STPFilterBuffer::STPFilterBuffer(const STPExecutionType type, const unsigned int bin_capacity) :
	Impl(std::make_unique<STPImpl>()) {
	if (type != STPExecutionType::Serial && type != STPExecutionType::Parallel) {
		STP_THROW(STPInvalidArgument, "invalid execution type"); // STPInvalidEnum in the reference
	}
	if (bin_capacity == 0u || bin_capacity > STPHeightfieldKernel::MaxBinCapacity) {
		STP_THROW(STPInvalidArgument, "histogram bin capacity must be in [1, 64]");
	}
	this->Impl->Type = type;
	this->Impl->Capacity = bin_capacity;
	this->Impl->Flag = STPSmartDeviceMemory::makeHost<unsigned int>(2u);
}

STPSingleHistogramFilter::STPFilterBuffer::STPFilterBuffer(STPFilterBuffer&&) noexcept = default;
STPSingleHistogramFilter::STPFilterBuffer& STPSingleHistogramFilter::STPFilterBuffer::operator=(STPFilterBuffer&&) noexcept = default;
STPSingleHistogramFilter::STPFilterBuffer::~STPFilterBuffer() = default;

STPSingleHistogram STPSingleHistogramFilter::STPFilterBuffer::readHistogram() const {
	return this->Impl->Result;
}

STPSingleHistogramFilter::STPFilterBuffer::STPHistogramSize STPSingleHistogramFilter::STPFilterBuffer::size() const {
	return {this->Impl->BinCount, this->Impl->OffsetCount};
}

STPSingleHistogramFilter::STPFilterBuffer::STPExecutionType STPSingleHistogramFilter::STPFilterBuffer::type() const noexcept {
	return this->Impl->Type;
}

unsigned int STPSingleHistogramFilter::STPFilterBuffer::binCapacity() const noexcept {
	return this->Impl->Capacity;
}

// STPSingleHistogramFilter.cpp:874-880, and the window must lie inside the merged map.
uint2 STPHeightfieldKernel::checkFilterDomain(const STPNearestNeighbourInformation& nn, const unsigned int radius) {
	STP_ASSERT_DOMAIN(radius > 0u && (radius & 0x01u) == 0x00u, "radius should be an positive even number");
	const uint2 start = make_uint2(nn.MapSize.x * (nn.ChunkNearestNeighbour.x / 2u), nn.MapSize.y * (nn.ChunkNearestNeighbour.y / 2u));
	STP_ASSERT_DOMAIN(radius <= start.x && radius <= start.y, "radius is too large and will overflow nearest neighbour boundary");
	if (nn.TotalMapSize.x != nn.MapSize.x * nn.ChunkNearestNeighbour.x
		|| nn.TotalMapSize.y != nn.MapSize.y * nn.ChunkNearestNeighbour.y || nn.MapSize.x == 0u || nn.MapSize.y == 0u) {
		STP_THROW(STPInvalidArgument, "inconsistent nearest-neighbour information");
	}
	// the (map + 2r)^2 window must lie inside the merged map (fails for even neighbourhoods)
	STP_ASSERT_DOMAIN(static_cast<size_t>(start.x) + nn.MapSize.x + radius <= nn.TotalMapSize.x
			&& static_cast<size_t>(start.y) + nn.MapSize.y + radius <= nn.TotalMapSize.y,
		"the filter window leaves the nearest-neighbour map (the neighbourhood must be odd)");
	return start;
}

STPSingleHistogram STPSingleHistogramFilter::filterDevice(const STPSample_t* const samplemap,
	const STPNearestNeighbourInformation& nn, STPFilterBuffer& filter_buffer, const unsigned int radius,
	const hipStream_t stream) {
	const uint2 start = STPHeightfieldKernel::checkFilterDomain(nn, radius);
	STPFilterBuffer::STPImpl& b = *filter_buffer.Impl;
	const size_t need = STPHeightfieldKernel::filterWorkspaceBytes(nn.MapSize, radius, b.Capacity);
	if (need > b.WorkspaceBytes) {
		STP_CHECK_HIP(hipStreamSynchronize(stream));
		b.Workspace.reset();
		b.Workspace = STPSmartDeviceMemory::makeDevice<char>(need);
		b.WorkspaceBytes = need;
	}
	b.Carved = STPHeightfieldKernel::carveFilterWorkspace(b.Workspace.get(), nn.MapSize, radius, b.Capacity);
	STPHeightfieldKernel::singleHistogramFilter(samplemap, nn.TotalMapSize.x, start, nn.MapSize, radius, b.Capacity, true,
		b.Carved, stream);
	const size_t pixels = static_cast<size_t>(nn.MapSize.x) * nn.MapSize.y;
	STP_CHECK_HIP(hipMemcpyAsync(b.Flag.get(), b.Carved.Overflow, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipMemcpyAsync(b.Flag.get() + 1, b.Carved.Offset + pixels, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipStreamSynchronize(stream));
	if (b.Flag[0] != 0u) {
		STP_THROW(STPCapacityError, "a pixel histogram holds more distinct items than the filter buffer's bin capacity ("
			+ std::to_string(b.Capacity) + ")");
	}
	b.BinCount = b.Flag[1];
	b.OffsetCount = pixels + 1u;
	b.Result = STPSingleHistogram {b.Carved.Bin, b.Carved.Offset};
	return b.Result;
}

STPSingleHistogram STPSingleHistogramFilter::operator()(const STPSample_t* const samplemap,
	const STPNearestNeighbourInformation& nn, STPFilterBuffer& filter_buffer, const unsigned int radius) {
	STPHeightfieldKernel::checkFilterDomain(nn, radius);
	STPFilterBuffer::STPImpl& b = *filter_buffer.Impl;
	if (!b.OwnStream) {
		STP_CHECK_HIP(hipStreamCreateWithFlags(&b.OwnStream, hipStreamNonBlocking));
	}
	const size_t samples = static_cast<size_t>(nn.TotalMapSize.x) * nn.TotalMapSize.y;
	if (samples > b.StagingCount) {
		b.Staging = STPSmartDeviceMemory::makeDevice<STPSample_t>(samples);
		b.StagingCount = samples;
	}
	STP_CHECK_HIP(hipMemcpyAsync(b.Staging.get(), samplemap, samples * sizeof(STPSample_t), hipMemcpyHostToDevice, b.OwnStream));
	const STPSingleHistogram dev = this->filterDevice(b.Staging.get(), nn, filter_buffer, radius, b.OwnStream);
	if (b.BinCount > b.HostBinCapacity) {
		b.HostBin = STPSmartDeviceMemory::makeHost<STPSingleHistogram::STPBin>(b.BinCount);
		b.HostBinCapacity = b.BinCount;
	}
	if (b.OffsetCount > b.HostOffsetCapacity) {
		b.HostOffset = STPSmartDeviceMemory::makeHost<unsigned int>(b.OffsetCount);
		b.HostOffsetCapacity = b.OffsetCount;
	}
	STP_CHECK_HIP(hipMemcpyAsync(b.HostBin.get(), dev.Bin, b.BinCount * sizeof(STPSingleHistogram::STPBin),
		hipMemcpyDeviceToHost, b.OwnStream));
	STP_CHECK_HIP(hipMemcpyAsync(b.HostOffset.get(), dev.HistogramStartOffset, b.OffsetCount * sizeof(unsigned int),
		hipMemcpyDeviceToHost, b.OwnStream));
	STP_CHECK_HIP(hipStreamSynchronize(b.OwnStream));
	b.Result = STPSingleHistogram {b.HostBin.get(), b.HostOffset.get()};
	return b.Result;
}

// ==================================================================== heightfield generator
namespace {
	STPNearestNeighbourInformation neighbourInfo(const uint2 map, const uint2 nn) {
		return STPNearestNeighbourInformation {map, nn, make_uint2(map.x * nn.x, map.y * nn.y)};
	}

	STPRainDropParameter toParameter(const STPEnvironment::STPRainDropSetting& s) {
		STPRainDropParameter p;
		p.Seed = s.Seed;
		p.RainDropCount = s.RainDropCount;
		p.Inertia = s.Inertia;
		p.SedimentCapacityFactor = s.SedimentCapacityFactor;
		p.minSedimentCapacity = s.minSedimentCapacity;
		p.initWaterVolume = s.initWaterVolume;
		p.minWaterVolume = s.minWaterVolume;
		p.Friction = s.Friction;
		p.initSpeed = s.initSpeed;
		p.ErodeSpeed = s.ErodeSpeed;
		p.DepositSpeed = s.DepositSpeed;
		p.EvaporateSpeed = s.EvaporateSpeed;
		p.Gravity = s.Gravity;
		p.ErosionBrushRadius = s.ErosionBrushRadius;
		p.MaxLifetime = s.MaxLifetime;
		return p;
	}
} // namespace

struct STPHeightfieldGenerator::STPImpl {
	STPRainDropParameter Rain;
	STPEnvironment::STPRainDropSetting RainSetting;
	STPNearestNeighbourInformation NoNeighbour, DiversityNeighbour, ErosionNeighbour;
	STPDiversityGenerator& Diversity;
	STPSmartDeviceMemory::STPDevice<int> BrushIndex;
	STPSmartDeviceMemory::STPDevice<float> BrushWeight;
	unsigned int BrushSize = 0u;
	hipMemPool_t Pool = nullptr;
	std::mutex StreamLock;
	std::vector<hipStream_t> FreeStream;
	std::atomic<unsigned int> ErodeCounter {0u};
	std::atomic<int> Mode {0};
	// Parallel erosion of a square merged map runs on an erosion-only batch (LDS-tiled windows in
	// colour batches) when the batch has a tiled launch shape for it; every call shares it, in turn.
	std::mutex EroderLock;
	bool EroderChecked = false, EroderUsed = false;
	std::unique_ptr<STPHeightfieldBatch> Eroder;
	STPSmartDeviceMemory::STPDevice<unsigned int> Gate; // 2 words, see STPHeightfieldKernel::heightGate
	hipEvent_t EroderDone = nullptr;                     // the last call's work on the eroder

	STPImpl(const STPEnvironment::STPChunkSetting& chunk, const STPEnvironment::STPRainDropSetting& rain,
		STPDiversityGenerator& diversity) :
		Rain(toParameter(rain)), RainSetting(rain), NoNeighbour(neighbourInfo(chunk.MapSize, make_uint2(1u, 1u))),
		DiversityNeighbour(neighbourInfo(chunk.MapSize, chunk.DiversityNearestNeighbour)),
		ErosionNeighbour(neighbourInfo(chunk.MapSize, chunk.ErosionNearestNeighbour)), Diversity(diversity) {}

	~STPImpl() {
		if (this->EroderDone) {
			// the eroder's arenas may still be in use by the last call's work
			(void)hipEventSynchronize(this->EroderDone);
			(void)hipEventDestroy(this->EroderDone);
		}
		this->Eroder.reset();
		for (hipStream_t s : this->FreeStream) {
			(void)hipStreamDestroy(s);
		}
		if (this->Pool) {
			(void)hipMemPoolDestroy(this->Pool);
		}
	}

	// STPObjectPool.inl:10-29 (stream pool)
	hipStream_t request() {
		{
			std::lock_guard<std::mutex> g(this->StreamLock);
			if (!this->FreeStream.empty()) {
				const hipStream_t s = this->FreeStream.back();
				this->FreeStream.pop_back();
				return s;
			}
		}
		hipStream_t s = nullptr;
		STP_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		return s;
	}
	void release(const hipStream_t s) {
		std::lock_guard<std::mutex> g(this->StreamLock);
		this->FreeStream.push_back(s);
	}

	// The erosion-only batch for this generator's merged map, or null where the tiled erosion does
	// not apply: a merged map that is not square, no lifetime bound (MaxLifetime = 0), or no tile
	// shape whose window fits in LDS. Called with EroderLock held.
	STPHeightfieldBatch* eroder() {
		if (this->EroderChecked) {
			return this->Eroder.get();
		}
		this->EroderChecked = true;
		const STPNearestNeighbourInformation& nn = this->ErosionNeighbour;
		const unsigned int side = nn.TotalMapSize.x;
		if (nn.TotalMapSize.y != side || nn.ChunkNearestNeighbour.x != nn.ChunkNearestNeighbour.y || side > 16384u
			|| this->Rain.MaxLifetime == 0u || this->Rain.RainDropCount == 0u) {
			return nullptr;
		}
		// the batch's noise and biome stages never run: any valid table; a biome in [0, 1] lets the
		// batch choose a tiled shape, and the heights the generator hands over are checked on the
		// device before each erosion (heightGate)
		const STPEnvironment::STPSimplexNoiseSetting noise {0ull, 8u, 0.0};
		const std::vector<STPBiomeProperty> biome {STPBiomeProperty {1.0f, 1u, 0.5f, 2.0f, 0.0f, 1.0f}};
		STPBatchSetting setting {};
		setting.BiomeSeed = 0ull;
		setting.BiomeCount = 1u;
		setting.BiomeCell = 1u;
		setting.Radius = 2u;
		setting.BinCapacity = 1u;
		setting.DropsPerPixel = 0.0f;
		STPBatchCapacity cap {};
		cap.Pixels = static_cast<unsigned long long>(side) * side + 4ull;
		cap.Drops = this->Rain.RainDropCount;
		cap.Chunks = 1u;
		cap.MaxSide = side;
		auto batch = std::make_unique<STPHeightfieldBatch>(noise, biome, this->RainSetting, setting, cap);
		const STPBatchTuning tuning = batch->latencyTuning();
		if (tuning.ErosionTile == 0u) {
			return nullptr;
		}
		batch->setTuning(tuning);
		this->Gate = STPSmartDeviceMemory::makeDevice<unsigned int>(2u);
		STP_CHECK_HIP(hipEventCreateWithFlags(&this->EroderDone, hipEventDisableTiming));
		this->Eroder = std::move(batch);
		return this->Eroder.get();
	}

	// Parallel erosion of the merged map `map` on the generator's erosion-only batch, enqueued on
	// `stream`; false (nothing enqueued) where that batch does not apply. The tiled window holds
	// heights in [0, WindowHeightLimit] (a fixed-point LDS copy), so the batch erodes a copy of the
	// map while heightGate checks the map's range on the device; when the range check or the
	// window's own range flag fails, the global-atomic kernel erodes the map itself instead, and
	// otherwise the copy replaces the map. Both paths run the same drops (seed, stream id, drop
	// index and spawn rule: STPHeightfieldKernel.cu:93-102).
	bool erodeTiled(float* const map, const unsigned int sid, const hipStream_t stream) {
		STPImpl& g = *this;
		const STPNearestNeighbourInformation& nn = g.ErosionNeighbour;
		const std::lock_guard<std::mutex> guard(g.EroderLock);
		STPHeightfieldBatch* const e = g.eroder();
		if (!e) {
			return false;
		}
		if (g.EroderUsed) {
			// the previous call's work may still read the batch's arenas and the gate
			STP_CHECK_HIP(hipStreamWaitEvent(stream, g.EroderDone, 0));
		}
		const unsigned int side = nn.TotalMapSize.x;
		const unsigned long long px = static_cast<unsigned long long>(side) * side;
		STPBatchChunk chunk {side, 0, 0, 1u, nn.ChunkNearestNeighbour.x};
		chunk.Drops = g.Rain.RainDropCount;
		try {
			e->bind({chunk}, {sid}, stream);
			STPHeightfieldKernel::heightGate(map, px, 0.0f, STPStage::WindowHeightLimit, g.Gate.get(), stream);
			STP_CHECK_HIP(hipMemcpyAsync(e->heightOf(0u), map, px * sizeof(float), hipMemcpyDeviceToDevice, stream));
			// drop binning and the tiled erosion
			e->runStages(4, 6, stream);
			STPHeightfieldKernel::copyGateStatus(e->windowStatus(), g.Gate.get(), stream);
			STPHeightfieldKernel::hydraulicErosion(map, nn, g.Rain, g.BrushIndex.get(), g.BrushWeight.get(), g.BrushSize,
				sid, STPHeightfieldKernel::STPErosionMode::Parallel, nullptr, nullptr, stream, g.Gate.get());
			STPHeightfieldKernel::selectUngated(g.Gate.get(), e->heightOf(0u), map, px, stream);
		} catch (...) {
			// whatever was enqueued still orders the next call
			(void)hipEventRecord(g.EroderDone, stream);
			g.EroderUsed = true;
			throw;
		}
		STP_CHECK_HIP(hipEventRecord(g.EroderDone, stream));
		g.EroderUsed = true;
		return true;
	}

	template <class FB, class XB>
	void runErosion(const unsigned int sid, const FB& fbuf, const XB& xbuf) {
		STPImpl& g = *this;
		const typename FB::STPMergedBuffer fm(fbuf, FB::STPMemoryLocation::DeviceMemory);
		const typename XB::STPMergedBuffer xm(xbuf, XB::STPMemoryLocation::DeviceMemory);
		const hipStream_t stream = fbuf.DeviceMemInfo.second;
		const STPNearestNeighbourInformation& nn = g.ErosionNeighbour;
		const bool serial = static_cast<STPHeightfieldGenerator::STPErosionMode>(g.Mode.load())
			== STPHeightfieldGenerator::STPErosionMode::Serialized;
		if (serial || !this->erodeTiled(fm.getDevice(), sid, stream)) {
			STPHeightfieldKernel::hydraulicErosion(fm.getDevice(), nn, g.Rain, g.BrushIndex.get(), g.BrushWeight.get(),
				g.BrushSize, sid,
				serial ? STPHeightfieldKernel::STPErosionMode::Serialized : STPHeightfieldKernel::STPErosionMode::Parallel,
				nullptr, nullptr, stream);
		}
		STPHeightfieldKernel::formatHeightmap(fm.getDevice(), xm.getDevice(),
			static_cast<unsigned long long>(nn.TotalMapSize.x) * nn.TotalMapSize.y, stream);
	}
};

STPHeightfieldGenerator::STPHeightfieldGenerator(const STPGeneratorSetup& setup) {
	if (!setup.ChunkSetting || !setup.RainDropSetting || !setup.DiversityGenerator) {
		STP_THROW(STPInvalidArgument, "generator setup has a null member");
	}
	const STPEnvironment::STPChunkSetting& chunk = *setup.ChunkSetting;
	const STPEnvironment::STPRainDropSetting& rain = *setup.RainDropSetting;
	chunk.validate();
	rain.validate();
	this->Impl = std::make_unique<STPImpl>(chunk, rain, *setup.DiversityGenerator);
	STPImpl& g = *this->Impl;
	std::vector<int> idx;
	std::vector<float> w;
	STPErosionBrushGenerator::generate(chunk.ErosionNearestNeighbour.x * chunk.MapSize.x, rain.ErosionBrushRadius, idx, w);
	g.BrushSize = static_cast<unsigned int>(idx.size());
	g.BrushIndex = STPSmartDeviceMemory::makeDevice<int>(idx.size());
	g.BrushWeight = STPSmartDeviceMemory::makeDevice<float>(w.size());
	STP_CHECK_HIP(hipMemcpy(g.BrushIndex.get(), idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
	STP_CHECK_HIP(hipMemcpy(g.BrushWeight.get(), w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));

	// STPHeightfieldGenerator.cpp:140-149: a device memory pool with a 100 MB release threshold
	int device = 0;
	STP_CHECK_HIP(hipGetDevice(&device));
	hipMemPoolProps props = {};
	props.allocType = hipMemAllocationTypePinned;
	props.handleTypes = hipMemHandleTypeNone;
	props.location.type = hipMemLocationTypeDevice;
	props.location.id = device;
	STP_CHECK_HIP(hipMemPoolCreate(&g.Pool, &props));
	uint64_t threshold = 104857600ull;
	STP_CHECK_HIP(hipMemPoolSetAttribute(g.Pool, hipMemPoolAttrReleaseThreshold, &threshold));
}

STPHeightfieldGenerator::~STPHeightfieldGenerator() = default;

void STPHeightfieldGenerator::setErosionMode(const STPErosionMode mode) noexcept {
	this->Impl->Mode = static_cast<int>(mode);
}

const STPNearestNeighbourInformation& STPHeightfieldGenerator::erosionNeighbour() const noexcept {
	return this->Impl->ErosionNeighbour;
}

void STPHeightfieldGenerator::generate(STPHeightFloat_t* const heightfield, const STPSample_t* const* const biomemap,
	const float2 offset) {
	STPImpl& g = *this->Impl;
	const hipStream_t stream = g.request();
	try {
		{
			const STPNearestNeighbourHeightFloatWTextureBuffer heightmap_buffer(&heightfield, g.NoNeighbour, {g.Pool, stream});
			const STPNearestNeighbourSampleRTextureBuffer samplemap_buffer(biomemap, g.DiversityNeighbour, {g.Pool, stream});
			g.Diversity(heightmap_buffer, samplemap_buffer, offset);
		}
		STP_CHECK_HIP(hipStreamSynchronize(stream));
	} catch (...) {
		(void)hipStreamSynchronize(stream);
		g.release(stream);
		throw;
	}
	g.release(stream);
}

void STPHeightfieldGenerator::generateDevice(STPHeightFloat_t* const heightfield, const STPSample_t* const* const biomemap,
	const float2 offset, const hipStream_t stream) {
	STPImpl& g = *this->Impl;
	using HB = STPNearestNeighbourHeightFloatWTextureBuffer;
	using SB = STPNearestNeighbourSampleRTextureBuffer;
	const HB heightmap_buffer(&heightfield, g.NoNeighbour, {g.Pool, stream}, HB::STPChunkLocation::Device);
	const SB samplemap_buffer(biomemap, g.DiversityNeighbour, {g.Pool, stream}, SB::STPChunkLocation::Device);
	g.Diversity(heightmap_buffer, samplemap_buffer, offset);
}

// ==================================================================== allocation-free erosion
// The merged FP32 map of erodeDeviceNoAlloc lives in a file-scope device array. In host code the
// array's name is not a device address: its address on each device comes from
// hipGetSymbolAddress, once per device with that device current, and only that address is passed
// to kernels.
static __device__ __align__(256) float scratch[STPHeightfieldGenerator::ErosionScratchPixels];

namespace {
	constexpr int ScratchDevices = 64;
	std::mutex ScratchLock;
	float* ScratchAddress[ScratchDevices] = {};

	// the scratch's address on the current device
	float* scratchOnDevice() {
		int device = 0;
		STP_CHECK_HIP(hipGetDevice(&device));
		if (device < 0 || device >= ScratchDevices) {
			STP_THROW(STPInvalidArgument, "device ordinal beyond the scratch table");
		}
		std::lock_guard<std::mutex> guard(ScratchLock);
		if (!ScratchAddress[device]) {
			void* p = nullptr;
			STP_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
			if (!p) {
				STP_THROW(STPHIPError, "no device address for the erosion scratch");
			}
			ScratchAddress[device] = static_cast<float*>(p);
		}
		return ScratchAddress[device];
	}
} // namespace

void STPHeightfieldGenerator::prepareScratch() {
	(void)scratchOnDevice();
}

void STPHeightfieldGenerator::erode(STPHeightFloat_t* const* const heightfield_original,
	STPHeightFixed_t* const* const heightfield_low) {
	STPImpl& g = *this->Impl;
	const hipStream_t stream = g.request();
	try {
		const STPNearestNeighbourHeightFloatRWTextureBuffer fbuf(heightfield_original, g.ErosionNeighbour, {g.Pool, stream});
		const STPNearestNeighbourHeightFixedWTextureBuffer xbuf(heightfield_low, g.ErosionNeighbour, {g.Pool, stream});
		g.runErosion(g.ErodeCounter++, fbuf, xbuf);
		STP_CHECK_HIP(hipStreamSynchronize(stream));
	} catch (...) {
		(void)hipStreamSynchronize(stream);
		g.release(stream);
		throw;
	}
	g.release(stream);
}

void STPHeightfieldGenerator::erodeDevice(STPHeightFloat_t* const* const heightfield_original,
	STPHeightFixed_t* const* const heightfield_low, const hipStream_t stream) {
	STPImpl& g = *this->Impl;
	using FB = STPNearestNeighbourHeightFloatRWTextureBuffer;
	using XB = STPNearestNeighbourHeightFixedWTextureBuffer;
	const FB fbuf(heightfield_original, g.ErosionNeighbour, {g.Pool, stream}, FB::STPChunkLocation::Device);
	const XB xbuf(heightfield_low, g.ErosionNeighbour, {g.Pool, stream}, XB::STPChunkLocation::Device);
	g.runErosion(g.ErodeCounter++, fbuf, xbuf);
}

void STPHeightfieldGenerator::erodeDeviceNoAlloc(STPHeightFloat_t* const* const heightfield_original,
	STPHeightFixed_t* const* const heightfield_low, const hipStream_t stream) {
	STPImpl& g = *this->Impl;
	const STPNearestNeighbourInformation& nn = g.ErosionNeighbour;
	if (!heightfield_original || !heightfield_low) {
		STP_THROW(STPInvalidArgument, "null chunk pointer array");
	}
	const unsigned long long pixels = static_cast<unsigned long long>(nn.TotalMapSize.x) * nn.TotalMapSize.y;
	if (pixels > ErosionScratchPixels) {
		STP_THROW(STPCapacityError, "the merged erosion map holds " + std::to_string(pixels) + " pixels; the "
			"allocation-free erosion holds at most " + std::to_string(ErosionScratchPixels));
	}
	float* const merged = scratchOnDevice();
	// every merged pixel is overwritten by the merge: nothing in the scratch needs clearing
	STPHeightfieldKernel::mergeNeighbour(reinterpret_cast<const void* const*>(heightfield_original), merged,
		sizeof(STPHeightFloat_t), nn, stream);
	STPHeightfieldKernel::hydraulicErosion(merged, nn, g.Rain, g.BrushIndex.get(), g.BrushWeight.get(), g.BrushSize,
		g.ErodeCounter++,
		static_cast<STPErosionMode>(g.Mode.load()) == STPErosionMode::Serialized
			? STPHeightfieldKernel::STPErosionMode::Serialized : STPHeightfieldKernel::STPErosionMode::Parallel,
		nullptr, nullptr, stream);
	STPHeightfieldKernel::unmergeFormatNeighbour(merged, heightfield_original, heightfield_low, nn, stream);
}

// ==================================================================== built-in diversity generators
STPSimplexHeightGenerator::STPSimplexHeightGenerator(const STPEnvironment::STPSimplexNoiseSetting& noise,
	const STPBiomeProperty& biome, const uint2 map_size) :
	Table(STPPermutationGenerator::generate(noise)), Biome(biome), MapSize(map_size) {
	if (biome.Octave == 0u || !(biome.Scale != 0.0f)) {
		STP_THROW(STPInvalidArgument, "biome property needs Octave > 0 and Scale != 0");
	}
}

const STPAlgorithm::STPPermutation& STPSimplexHeightGenerator::permutation() const noexcept {
	return this->Table.Permutation;
}

void STPSimplexHeightGenerator::operator()(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
	const STPNearestNeighbourSampleRTextureBuffer&, const float2 offset) {
	const STPNearestNeighbourInformation& nn = heightmap_buffer.NeighbourInfo;
	if (nn.MapSize.x != this->MapSize.x || nn.MapSize.y != this->MapSize.y) {
		STP_THROW(STPInvalidArgument, "the heightmap size differs from the size this generator was made for");
	}
	const STPNearestNeighbourHeightFloatWTextureBuffer::STPMergedBuffer merged(heightmap_buffer,
		STPNearestNeighbourHeightFloatWTextureBuffer::STPMemoryLocation::DeviceMemory);
	STPSimplexNoise::STPFractalSimplexInformation info;
	info.Persistence = this->Biome.Persistence;
	info.Lacunarity = this->Biome.Lacunarity;
	info.Octave = this->Biome.Octave;
	info.Scale = this->Biome.Scale;
	info.Offset = offset;
	info.HalfDimension = make_float2(static_cast<float>(nn.MapSize.x) / 2.0f, static_cast<float>(nn.MapSize.y) / 2.0f);
	info.Amplitude = 1.0f;
	info.Frequency = 1.0f;
	STPHeightfieldKernel::simplexHeightmap(this->Table.Permutation, info, this->Biome.Variation, this->Biome.Depth,
		merged.getDevice(), nn.TotalMapSize.x, nn.TotalMapSize.y, heightmap_buffer.DeviceMemInfo.second);
}

// The filter engine: a one-chunk batch of the map's side with an external biome map, of which only
// stages 0-2 run (halo copy, vertical and horizontal filter passes with lines split into
// segments); the blend stays the plugin's own (its offset is a float). Calls take it in turn.
struct STPBiomefieldGenerator::STPFilterEngine {
	std::mutex Lock;
	bool Checked = false;
	std::unique_ptr<STPHeightfieldBatch> Batch; // null: the segmented passes do not apply
};

STPBiomefieldGenerator::STPBiomefieldGenerator(const STPEnvironment::STPSimplexNoiseSetting& noise,
	std::vector<STPBiomeProperty> table, const uint2 map_size, const unsigned int interpolation_radius,
	const unsigned int bin_capacity) :
	Table(STPPermutationGenerator::generate(noise)), BiomeTable(std::move(table)), MapSize(map_size),
	InterpolationRadius(interpolation_radius), BinCapacity(bin_capacity), Engine(std::make_unique<STPFilterEngine>()) {
	if (this->BiomeTable.empty() || this->BiomeTable.size() > STPHeightfieldKernel::MaxBiome) {
		STP_THROW(STPInvalidArgument, "biome count must be in [1, 64]");
	}
	if (bin_capacity == 0u || bin_capacity > STPHeightfieldKernel::MaxBinCapacity) {
		STP_THROW(STPInvalidArgument, "histogram bin capacity must be in [1, 64]");
	}
}

const STPAlgorithm::STPPermutation& STPBiomefieldGenerator::permutation() const noexcept {
	return this->Table.Permutation;
}

STPBiomefieldGenerator::~STPBiomefieldGenerator() = default;

bool STPBiomefieldGenerator::filterBySegments(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
	const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, const float2 offset) {
	const unsigned int side = this->MapSize.x, radius = this->InterpolationRadius;
	if (this->MapSize.y != side) {
		return false;
	}
	STPFilterEngine& e = *this->Engine;
	const std::lock_guard<std::mutex> guard(e.Lock);
	if (!e.Checked) {
		e.Checked = true;
		e.Batch = STPHeightfieldBatch::makeFilter(this->BiomeTable, side, radius, this->BinCapacity);
	}
	if (!e.Batch) {
		return false;
	}
	const STPNearestNeighbourInformation& bnn = biomemap_buffer.NeighbourInfo;
	const uint2 start = STPHeightfieldKernel::checkFilterDomain(bnn, radius);
	const hipStream_t stream = heightmap_buffer.DeviceMemInfo.second;
	STPHeightfieldBatch& b = *e.Batch;
	const STPNearestNeighbourHeightFloatWTextureBuffer::STPMergedBuffer height(heightmap_buffer,
		STPNearestNeighbourHeightFloatWTextureBuffer::STPMemoryLocation::DeviceMemory);
	const STPNearestNeighbourSampleRTextureBuffer::STPMergedBuffer biome(biomemap_buffer,
		STPNearestNeighbourSampleRTextureBuffer::STPMemoryLocation::DeviceMemory);
	b.filterExternal(biome.getDevice() + static_cast<size_t>(start.y) * bnn.TotalMapSize.x + start.x, bnn.TotalMapSize.x,
		stream);
	const unsigned int count = static_cast<unsigned int>(this->BiomeTable.size());
	if (b.packedHistogram()) {
		const unsigned int w = 2u * radius + 1u;
		STPHeightfieldKernel::multiBiomeHeightmapPacked(this->Table.Permutation, this->BiomeTable.data(), count,
			b.binCountOf(0u), b.binItemOf(0u), 1.0f / static_cast<float>(w * w), this->BinCapacity, height.getDevice(),
			this->MapSize, offset, stream);
	} else {
		STPHeightfieldKernel::multiBiomeHeightmapDense(this->Table.Permutation, this->BiomeTable.data(), count,
			b.binCountOf(0u), b.binItemOf(0u), b.binWeightOf(0u), this->BinCapacity, height.getDevice(), this->MapSize,
			offset, stream);
	}
	const unsigned int status = b.status(stream);
	if (status & 1u) {
		STP_THROW(STPCapacityError, "a pixel histogram holds more distinct biomes than the bin capacity ("
			+ std::to_string(this->BinCapacity) + ")");
	}
	// a biome id beyond the table: the line-by-line filter keeps it as an item of its own
	return (status & 4u) == 0u;
}

void STPBiomefieldGenerator::operator()(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
	const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, const float2 offset) {
	const STPNearestNeighbourInformation& bnn = biomemap_buffer.NeighbourInfo;
	if (bnn.MapSize.x != this->MapSize.x || bnn.MapSize.y != this->MapSize.y
		|| heightmap_buffer.NeighbourInfo.MapSize.x != this->MapSize.x
		|| heightmap_buffer.NeighbourInfo.MapSize.y != this->MapSize.y) {
		STP_THROW(STPInvalidArgument, "the map size differs from the size this generator was made for");
	}
	if (this->filterBySegments(heightmap_buffer, biomemap_buffer, offset)) {
		return;
	}
	const uint2 start = STPHeightfieldKernel::checkFilterDomain(bnn, this->InterpolationRadius);
	const auto [pool, stream] = heightmap_buffer.DeviceMemInfo;
	const STPNearestNeighbourHeightFloatWTextureBuffer::STPMergedBuffer height(heightmap_buffer,
		STPNearestNeighbourHeightFloatWTextureBuffer::STPMemoryLocation::DeviceMemory);
	const STPNearestNeighbourSampleRTextureBuffer::STPMergedBuffer biome(biomemap_buffer,
		STPNearestNeighbourSampleRTextureBuffer::STPMemoryLocation::DeviceMemory);

	const size_t bytes = STPHeightfieldKernel::filterWorkspaceBytes(bnn.MapSize, this->InterpolationRadius, this->BinCapacity);
	void* ws = nullptr;
	if (pool) {
		STP_CHECK_HIP(hipMallocFromPoolAsync(&ws, bytes, pool, stream));
	} else {
		STP_CHECK_HIP(hipMallocAsync(&ws, bytes, stream));
	}
	const STPFilterWorkspace carved = STPHeightfieldKernel::carveFilterWorkspace(ws, bnn.MapSize, this->InterpolationRadius,
		this->BinCapacity);
	STPHeightfieldKernel::singleHistogramFilter(biome.getDevice(), bnn.TotalMapSize.x, start, bnn.MapSize,
		this->InterpolationRadius, this->BinCapacity, false, carved, stream);
	STPHeightfieldKernel::multiBiomeHeightmapDense(this->Table.Permutation, this->BiomeTable.data(),
		static_cast<unsigned int>(this->BiomeTable.size()), carved.BinCount, carved.BinItem, carved.BinWeight,
		this->BinCapacity, height.getDevice(), heightmap_buffer.NeighbourInfo.MapSize, offset, stream);
	unsigned int overflow = 0u;
	STP_CHECK_HIP(hipMemcpyAsync(&overflow, carved.Overflow, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipFreeAsync(ws, stream));
	STP_CHECK_HIP(hipStreamSynchronize(stream));
	if (overflow != 0u) {
		STP_THROW(STPCapacityError, "a pixel histogram holds more distinct biomes than the bin capacity ("
			+ std::to_string(this->BinCapacity) + ")");
	}
}
