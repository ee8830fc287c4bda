// C++ API of the MI355X-native SuperTerrain+ heightfield path.
//
// Class names, members and call semantics follow the reference's public headers so that code
// written against them ports by swapping includes and the glm vector types for HIP's
// (uint2/float2/double2/double3):
//   STPHeightfieldGenerator  SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Chunk/STPHeightfieldGenerator.h:26-161
//   STPDiversityGenerator    .../World/Chunk/STPDiversityGenerator.hpp:14-37
//   STPNearestNeighbourTextureBuffer  .../World/Chunk/STPNearestNeighbourTextureBuffer.h:20-139
//   STPChunkSetting / STPRainDropSetting  .../Environment/STPChunkSetting.h:15-45, STPRainDropSetting.h:15-53
//   STPSimplexNoiseSetting, STPPermutationGenerator, STPSingleHistogramFilter
//                            SuperTerrain+/SuperAlgorithm+/Host/Public/SuperAlgorithm+Host/*
//   STPChunk (coordinate math)  .../World/Chunk/STPChunk.h:110-157
// Differences are listed where they occur (MaxLifetime, the permutation engine, device-resident
// overloads, HIP types).
#pragma once
#ifndef STP_AMD_SUPERTERRAINPLUS_HPP
#define STP_AMD_SUPERTERRAINPLUS_HPP

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <stp/STPBiomeLayer.hpp>
#include <stp/STPBiomeProperty.hpp>
#include <stp/STPSimplexNoise.hpp>
#include <stp/STPSingleHistogramWrapper.hpp>

namespace SuperTerrainPlus {

	// STPWorldMapPixelFormat.hpp:12-21
	typedef std::uint16_t STPSample_t;
	typedef float STPHeightFloat_t;
	typedef std::uint16_t STPHeightFixed_t;
	typedef std::uint64_t STPSeed_t;

	namespace STPException {

		// Status codes shared with the C ABI (include/stp_amd.h).
		enum STPStatus : int {
			Ok = 0,
			InvalidEnvironment = 1,
			NumericDomain = 2,
			HIPRuntime = 3,
			Capacity = 4,
			Compilation = 5,
			InvalidArgument = 6,
			Internal = 7
		};

		// STPFundamentalException.h:16-81: a description plus the source location.
		class STPFundamentalException : public std::runtime_error {
		public:
			const STPStatus Status;
			STPFundamentalException(const STPStatus status, const std::string& what, const char* file, const int line) :
				std::runtime_error(what + " (" + file + ":" + std::to_string(line) + ")"), Status(status) {}
		};
#define STP_DECLARE_EXCEPTION(NAME, CODE)                                                                  \
	class NAME : public STPFundamentalException {                                                         \
	public:                                                                                              \
		NAME(const std::string& what, const char* file, const int line) :                                \
			STPFundamentalException(CODE, std::string(#NAME ": ") + what, file, line) {}                 \
	};
		STP_DECLARE_EXCEPTION(STPInvalidEnvironment, InvalidEnvironment)
		STP_DECLARE_EXCEPTION(STPNumericDomainError, NumericDomain)
		STP_DECLARE_EXCEPTION(STPHIPError, HIPRuntime)
		STP_DECLARE_EXCEPTION(STPCapacityError, Capacity)
		STP_DECLARE_EXCEPTION(STPCompilationError, Compilation)
		STP_DECLARE_EXCEPTION(STPInvalidArgument, InvalidArgument)
#undef STP_DECLARE_EXCEPTION

	} // namespace STPException

	// STPNearestNeighbourInformation.hpp:13-25
	struct STPNearestNeighbourInformation {
		uint2 MapSize;
		uint2 ChunkNearestNeighbour;
		uint2 TotalMapSize;
	};

	namespace STPEnvironment {

		// STPChunkSetting.h:15-45
		struct STPChunkSetting {
			uint2 ChunkSize;
			uint2 MapSize;
			double3 ChunkOffset;
			double2 ChunkScale;
			double2 MapOffset;
			uint2 DiversityNearestNeighbour;
			uint2 ErosionNearestNeighbour;
			uint2 SplatNearestNeighbour;
			uint2 RenderDistance;

			void validate() const; // STPEnvironment.cpp:17-45
		};

		// STPRainDropSetting.h:15-53, plus MaxLifetime (0 = unlimited, the reference's behaviour).
		struct STPRainDropSetting {
			STPSeed_t Seed;
			unsigned int RainDropCount;
			float Inertia;
			float SedimentCapacityFactor;
			float minSedimentCapacity;
			float initWaterVolume;
			float minWaterVolume;
			float Friction;
			float initSpeed;
			float ErodeSpeed;
			float DepositSpeed;
			float EvaporateSpeed;
			float Gravity;
			unsigned int ErosionBrushRadius;
			unsigned int MaxLifetime = 0u;

			// STPEnvironment.cpp:51-65; additionally rejects drops that could live > 65535 steps.
			void validate() const;
			// Number of steps a drop can live at most (MaxLifetime, or the evaporation bound).
			unsigned int lifetimeBound() const;
		};

		// The engine that shuffles the permutation (the reference hard-codes std::mt19937_64).
		enum class STPPermutationEngine : unsigned int { MT19937_64 = 0u, MT19937 = 1u };

		// STPSimplexNoiseSetting.h:14-27 (+ Engine)
		struct STPSimplexNoiseSetting {
			STPSeed_t Seed;
			unsigned int Distribution;
			double Offset;
			STPPermutationEngine Engine = STPPermutationEngine::MT19937_64;

			void validate() const; // STPAlgorithmEnvironment.cpp:13-17
		};

	} // namespace STPEnvironment

	// RAII over HIP memory (STPSmartDeviceMemory.inl:47-68).
	namespace STPSmartDeviceMemory {
		struct STPDeviceDeleter {
			void operator()(void* p) const noexcept;
		};
		struct STPPinnedDeleter {
			void operator()(void* p) const noexcept;
		};
		template <class T>
		using STPDevice = std::unique_ptr<T[], STPDeviceDeleter>;
		template <class T>
		using STPHost = std::unique_ptr<T[], STPPinnedDeleter>;

		void* allocateDevice(size_t bytes);
		void* allocatePinned(size_t bytes);

		template <class T>
		STPDevice<T> makeDevice(const size_t count) {
			return STPDevice<T>(static_cast<T*>(allocateDevice(sizeof(T) * count)));
		}
		template <class T>
		STPHost<T> makeHost(const size_t count) {
			return STPHost<T>(static_cast<T*>(allocatePinned(sizeof(T) * count)));
		}
	} // namespace STPSmartDeviceMemory

	// STPChunk.h:110-157 coordinate math (pure functions).
	struct STPChunk {
		static int2 calcWorldChunkCoordinate(double3 pointPos, uint2 chunkSize, double2 scale) noexcept;
		static uint2 calcLocalChunkCoordinate(unsigned int chunkID, uint2 chunkRange) noexcept;
		static int2 calcLocalChunkOrigin(int2 centreChunkCoord, uint2 chunkSize, uint2 neighbourSize) noexcept;
		static double2 calcChunkMapOffset(int2 chunkCoord, uint2 chunkSize, uint2 mapSize, double2 mapOffset) noexcept;
		static int2 offsetChunk(int2 chunkCoord, uint2 chunkSize, int2 offset) noexcept;
		static std::vector<int2> calcChunkNeighbourOffset(uint2 chunkSize, uint2 regionSize);
	};

	namespace STPAlgorithm {

		namespace STPPermutationGenerator {

			// STPPermutationGenerator.h:23-43: owns the device tables; Permutation is the device view.
			class STPPermutationTable {
			public:
				STPSmartDeviceMemory::STPDevice<unsigned char> PermutationMemory;
				STPSmartDeviceMemory::STPDevice<float> GradientMemory;
				STPPermutation Permutation;
			};

			// Host tables: permutation[512], gradient[2 * Distribution].
			void generateHost(const STPEnvironment::STPSimplexNoiseSetting& setting, unsigned char* permutation,
				float* gradient);
			STPPermutationTable generate(const STPEnvironment::STPSimplexNoiseSetting& setting);

		} // namespace STPPermutationGenerator

		// STPSingleHistogramFilter.h:27-208 on the GPU. The filter keeps the reference's bin order
		// (see the kernel file for the rules), so the output is identical to the CPU reference.
		class STPSingleHistogramFilter {
		public:
			class STPFilterBuffer {
			public:
				// Kept for source compatibility; both run on the GPU.
				enum class STPExecutionType : unsigned char { Serial = 0x00u, Parallel = 0xFFu };
				typedef std::pair<size_t, size_t> STPHistogramSize; // (bins, offsets)

				explicit STPFilterBuffer(STPExecutionType type = STPExecutionType::Parallel, unsigned int bin_capacity = 16u);
				STPFilterBuffer(STPFilterBuffer&&) noexcept;
				STPFilterBuffer& operator=(STPFilterBuffer&&) noexcept;
				~STPFilterBuffer();

				STPSingleHistogram readHistogram() const; // host view after a host call, device view after a device call
				STPHistogramSize size() const;
				STPExecutionType type() const noexcept;
				unsigned int binCapacity() const noexcept;

			private:
				friend class STPSingleHistogramFilter;
				struct STPImpl;
				std::unique_ptr<STPImpl> Impl;
			};

			STPSingleHistogramFilter() = default;
			STPSingleHistogramFilter(const STPSingleHistogramFilter&) = delete;
			STPSingleHistogramFilter& operator=(const STPSingleHistogramFilter&) = delete;

			// STPSingleHistogramFilter.cpp:870-902: host samplemap in, host (pinned) histogram out.
			STPSingleHistogram operator()(const STPSample_t* samplemap, const STPNearestNeighbourInformation& nn_info,
				STPFilterBuffer& filter_buffer, unsigned int radius);
			// Device samplemap in, device histogram out (synchronises `stream` to check capacity).
			STPSingleHistogram filterDevice(const STPSample_t* samplemap, const STPNearestNeighbourInformation& nn_info,
				STPFilterBuffer& filter_buffer, unsigned int radius, hipStream_t stream);
		};

	} // namespace STPAlgorithm

	// STPNearestNeighbourTextureBuffer.h:20-30
	enum class STPNearestNeighbourTextureBufferMemoryMode : unsigned char { ReadOnly = 0x00u, WriteOnly = 0x01u, ReadWrite = 0x02u };

	// STPNearestNeighbourTextureBuffer.h:43-139. The chunk textures may live in host memory (the
	// reference's only case) or in device memory (MI355X-native resident chunks): ChunkLocation.
	template <typename T, STPNearestNeighbourTextureBufferMemoryMode MM>
	class STPNearestNeighbourTextureBuffer {
	public:
		using TextureType = std::conditional_t<MM == STPNearestNeighbourTextureBufferMemoryMode::ReadOnly, std::add_const_t<T>, T>;
		typedef std::pair<hipMemPool_t, hipStream_t> STPDeviceMemoryOperator;
		enum class STPMemoryLocation : unsigned char { HostMemory = 0x00u, DeviceMemory = 0x01u };
		enum class STPChunkLocation : unsigned char { Host = 0x00u, Device = 0x01u };

		class STPMergedBuffer {
		public:
			using MutableT = std::remove_const_t<T>;
			const STPMemoryLocation Location;

			STPMergedBuffer(const STPNearestNeighbourTextureBuffer& buffer, STPMemoryLocation location);
			STPMergedBuffer(const STPMergedBuffer&) = delete;
			STPMergedBuffer& operator=(const STPMergedBuffer&) = delete;
			~STPMergedBuffer() noexcept(false);

			MutableT* getHost() const noexcept;
			MutableT* getDevice() const noexcept;

		private:
			const STPNearestNeighbourTextureBuffer& Main;
			MutableT* HostMem = nullptr;
			MutableT* DeviceMem = nullptr;
			// host chunks of a device buffer: each chunk's contiguous copy in device memory
			MutableT* Staging = nullptr;
			void copy(bool pack);
		};

		const STPNearestNeighbourInformation& NeighbourInfo;
		const STPDeviceMemoryOperator DeviceMemInfo;
		const STPChunkLocation ChunkLocation;

		STPNearestNeighbourTextureBuffer(TextureType* const* texture, const STPNearestNeighbourInformation& nn_info,
			const STPDeviceMemoryOperator& mem, STPChunkLocation chunk_location = STPChunkLocation::Host);
		STPNearestNeighbourTextureBuffer(const STPNearestNeighbourTextureBuffer&) = delete;
		STPNearestNeighbourTextureBuffer& operator=(const STPNearestNeighbourTextureBuffer&) = delete;

	private:
		TextureType* const* const NeighbourTexture;
	};

	typedef STPNearestNeighbourTextureBuffer<STPHeightFloat_t, STPNearestNeighbourTextureBufferMemoryMode::WriteOnly>
		STPNearestNeighbourHeightFloatWTextureBuffer;
	typedef STPNearestNeighbourTextureBuffer<STPHeightFloat_t, STPNearestNeighbourTextureBufferMemoryMode::ReadWrite>
		STPNearestNeighbourHeightFloatRWTextureBuffer;
	typedef STPNearestNeighbourTextureBuffer<STPSample_t, STPNearestNeighbourTextureBufferMemoryMode::ReadOnly>
		STPNearestNeighbourSampleRTextureBuffer;
	typedef STPNearestNeighbourTextureBuffer<STPHeightFixed_t, STPNearestNeighbourTextureBufferMemoryMode::WriteOnly>
		STPNearestNeighbourHeightFixedWTextureBuffer;

	// STPDiversityGenerator.hpp:14-37
	class STPDiversityGenerator {
	public:
		STPDiversityGenerator() = default;
		virtual ~STPDiversityGenerator() = default;
		virtual void operator()(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
			const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, float2 offset) = 0;
	};

	// STPHeightfieldGenerator.h:26-161
	class STPHeightfieldGenerator {
	public:
		struct STPGeneratorSetup {
			const STPEnvironment::STPChunkSetting* ChunkSetting;
			const STPEnvironment::STPRainDropSetting* RainDropSetting;
			STPDiversityGenerator* DiversityGenerator;
		};
		// Parallel (default) matches the reference's release build; Serialized is bit-exact.
		enum class STPErosionMode : int { Parallel = 0, Serialized = 1 };

		explicit STPHeightfieldGenerator(const STPGeneratorSetup& setup);
		STPHeightfieldGenerator(const STPHeightfieldGenerator&) = delete;
		STPHeightfieldGenerator& operator=(const STPHeightfieldGenerator&) = delete;
		~STPHeightfieldGenerator();

		// STPHeightfieldGenerator.h:145: host chunk maps, synchronous.
		void generate(STPHeightFloat_t* heightfield, const STPSample_t* const* biomemap, float2 offset);
		// STPHeightfieldGenerator.h:159: host chunk maps, synchronous.
		void erode(STPHeightFloat_t* const* heightfield_original, STPHeightFixed_t* const* heightfield_low);

		// Device-resident overloads: chunk maps already in HBM, work enqueued on `stream`.
		void generateDevice(STPHeightFloat_t* heightfield, const STPSample_t* const* biomemap, float2 offset,
			hipStream_t stream);
		void erodeDevice(STPHeightFloat_t* const* heightfield_original, STPHeightFixed_t* const* heightfield_low,
			hipStream_t stream);
		// No reference counterpart; for callers that cannot allocate (code under graph capture,
		// serving loops, harnesses that forbid allocation). erodeDevice with the merged map held in
		// the library's fixed device scratch instead of the generator's memory pool: no allocation
		// and no free. The merged map may hold at most ErosionScratchPixels pixels (STPCapacityError
		// otherwise). Every call on one device shares that scratch: calls that overlap on one
		// device race.
		void erodeDeviceNoAlloc(STPHeightFloat_t* const* heightfield_original, STPHeightFixed_t* const* heightfield_low,
			hipStream_t stream);
		static constexpr unsigned long long ErosionScratchPixels = 4ull << 20;
		// Looks up the scratch's address on the current device (once per device; the library's
		// initialisation calls it, and the first allocation-free call on a device does if needed).
		static void prepareScratch();

		void setErosionMode(STPErosionMode mode) noexcept;
		const STPNearestNeighbourInformation& erosionNeighbour() const noexcept;

	private:
		struct STPImpl;
		std::unique_ptr<STPImpl> Impl;
	};

	// Built-in diversity generator: one biome, h = simplex2DFractal * Variation + Depth.
	class STPSimplexHeightGenerator : public STPDiversityGenerator {
	public:
		STPSimplexHeightGenerator(const STPEnvironment::STPSimplexNoiseSetting& noise, const STPBiomeProperty& biome,
			uint2 map_size);
		void operator()(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
			const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, float2 offset) override;
		const STPAlgorithm::STPPermutation& permutation() const noexcept;

	private:
		STPAlgorithm::STPPermutationGenerator::STPPermutationTable Table;
		STPBiomeProperty Biome;
		uint2 MapSize;
	};

	// Built-in multi-biome diversity generator (SuperDemo+/World/Biomes/STPBiomefieldGenerator.cpp:79-152):
	// device single histogram filter over the diversity neighbourhood, then the blend kernel.
	class STPBiomefieldGenerator : public STPDiversityGenerator {
	public:
		STPBiomefieldGenerator(const STPEnvironment::STPSimplexNoiseSetting& noise, std::vector<STPBiomeProperty> table,
			uint2 map_size, unsigned int interpolation_radius, unsigned int bin_capacity = 16u);
		~STPBiomefieldGenerator() override;
		void operator()(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
			const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, float2 offset) override;
		const STPAlgorithm::STPPermutation& permutation() const noexcept;

	private:
		STPAlgorithm::STPPermutationGenerator::STPPermutationTable Table;
		std::vector<STPBiomeProperty> BiomeTable;
		uint2 MapSize;
		unsigned int InterpolationRadius, BinCapacity;
		// The batched path's filter (lines split into segments that run in parallel) for square maps
		// of at most 8 biomes and bins; false where it does not apply (nothing written then).
		struct STPFilterEngine;
		std::unique_ptr<STPFilterEngine> Engine;
		bool filterBySegments(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
			const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, float2 offset);
	};

} // namespace SuperTerrainPlus

#endif // STP_AMD_SUPERTERRAINPLUS_HPP
