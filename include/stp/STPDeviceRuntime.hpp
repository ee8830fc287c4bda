// Runtime compilation of device code (hipRTC) and runtime linking into a loadable module.
//
// STPDeviceRuntimeBinary follows SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/GPGPU/
// STPDeviceRuntimeBinary.h:26-127 (compile a source with options, name expressions and external
// headers; keep the log and the lowered names), and STPDeviceRuntimeProgram follows
// SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/GPGPU/STPDeviceRuntimeProgram.h:26-94 (link
// compiled programs into one module). hipRTC in place of NVRTC: the binary types are AMDGPU code
// objects and LLVM bitcode in place of PTX and NVVM; scripts test __HIPCC_RTC__ in place of
// __CUDACC_RTC__.
//
// The library's device headers (stp/STPSimplexNoise.hpp, stp/STPSingleHistogramWrapper.hpp,
// stp/STPRandom.hpp) are compiled into the library; libraryHeaders() hands them to a script as
// external headers, so a runtime-compiled diversity generator runs the same device code as the
// built-in one.
#pragma once
#ifndef STP_AMD_DEVICE_RUNTIME_HPP
#define STP_AMD_DEVICE_RUNTIME_HPP

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <stp/SuperTerrainPlus.hpp>
#include <stp/STPTextureSplat.hpp>

#include <memory>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

namespace SuperTerrainPlus {

	namespace STPDeviceRuntimeBinary {

		// STPDeviceRuntimeBinary.h:26-56
		struct STPSourceInformation {
			struct STPSourceArgument {
				std::vector<std::string> StringArgument;
				// Adds one argument; returns *this for chaining.
				STPSourceArgument& operator[](const std::string& arg);
			};
			// Compiler options. The library always adds --offload-arch for the current device
			// (gfx950 when none is visible), -O3 and -ffp-contract=off unless the option list names
			// them already.
			STPSourceArgument Option;
			// Names whose lowered (mangled) names should be recorded: kernels, template
			// instantiations and global variables.
			STPSourceArgument NameExpression;
			// Names of external headers the source includes; each must be a key of the header
			// source map passed to compile().
			STPSourceArgument ExternalHeader;
		};
		// Key: the header name used in #include; value: the header's source.
		typedef std::unordered_map<std::string, std::string> STPExternalHeaderSource;
		// Key: name expression; value: lowered name.
		typedef std::unordered_map<std::string, std::string> STPLoweredName;

		struct STPProgramDeleter {
			void operator()(hiprtcProgram program) const;
		};
		using STPSmartProgram = std::unique_ptr<std::remove_pointer_t<hiprtcProgram>, STPProgramDeleter>;
		using STPProgramData = std::pair<std::unique_ptr<char[]>, size_t>;

		// STPDeviceRuntimeBinary.h:76-101
		struct STPCompilationOutput {
			struct STPCompiledBinary {
				std::string Identifier;
				STPSmartProgram Program;
			} ProgramObject;
			std::string Log;
			// Name expressions that could not be lowered (not found in the source) are left out.
			STPLoweredName LoweredName;
		};

		// Compiles `source_code`. Throws STPException::STPCompilationError (what() holds the
		// compiler log) if compilation fails, STPInvalidArgument if an external header named in
		// `source_info` has no source in `external_header`.
		STPCompilationOutput compile(std::string&& source_name, const std::string& source_code,
			const STPSourceInformation& source_info, const STPExternalHeaderSource& external_header = {});

		// LLVM bitcode of a program compiled with -fgpu-rdc (the NVVM counterpart).
		STPProgramData readBitcode(hiprtcProgram program);
		// AMDGPU code object of a program compiled without -fgpu-rdc (the CUBIN counterpart).
		STPProgramData readCodeObject(hiprtcProgram program);

		// The library's device headers as external headers, keyed by their include names.
		const STPExternalHeaderSource& libraryHeaders();

	} // namespace STPDeviceRuntimeBinary

	namespace STPDeviceRuntimeProgram {

		// STPDeviceRuntimeProgram.h:26-33 (PTX / NVVM)
		enum class STPBinaryType : unsigned char { CodeObject = 0x00u, LLVMBitcode = 0xFFu };

		// STPDeviceRuntimeProgram.h:35-68
		struct STPLinkerInformation {
			struct STPDataJitOption {
				std::vector<hipJitOption> OptionFlag;
				std::vector<void*> OptionValue;
				STPDataJitOption& operator()(hipJitOption flag, void* value);
			};
			STPDataJitOption LinkerOption;
			std::vector<std::tuple<const STPDeviceRuntimeBinary::STPCompilationOutput::STPCompiledBinary*, STPBinaryType,
				STPDataJitOption>>
				DataOption;
			// Bitcode archives (.a of bundled bitcode) to link, with their options.
			std::vector<std::pair<std::string, STPDataJitOption>> ArchiveOption;
			// Options for hipModuleLoadDataEx.
			STPDataJitOption ModuleOption;
		};

		struct STPModuleDeleter {
			void operator()(hipModule_t module) const;
		};
		using STPSmartModule = std::unique_ptr<std::remove_pointer_t<hipModule_t>, STPModuleDeleter>;

		// Links every program into one module and loads it on the current device. One code
		// object and nothing else loads directly; bitcode inputs and archives go through the
		// hipRTC linker. Throws STPException::STPCompilationError on a link failure.
		STPSmartModule link(STPLinkerInformation& linker_info);

	} // namespace STPDeviceRuntimeProgram

	// Multi-biome diversity generator whose blend kernel is compiled at run time
	// (SuperDemo+/World/Biomes/STPBiomefieldGenerator.cpp:44-152 with
	// SuperDemo+/Script/STPMultiHeightGenerator.cu): device single histogram filter over the
	// diversity neighbourhood, then the script's generateMultiBiomeHeightmap kernel. The script
	// contract (globals Dimension, HalfDimension, Permutation, BiomeCount, BiomeTable; kernel
	// generateMultiBiomeHeightmap(float*, STPSingleHistogram, float2)) is the one of the built-in
	// script superterrainplus_amd/rtc/STPMultiHeightGenerator.rtc.
	class STPRuntimeBiomefieldGenerator : public STPDiversityGenerator {
	public:
		// `script`: generator source, empty = the built-in script. `options`: extra compiler
		// options (the library's device headers are always available as external headers).
		STPRuntimeBiomefieldGenerator(const STPEnvironment::STPSimplexNoiseSetting& noise,
			std::vector<STPBiomeProperty> table, uint2 map_size, unsigned int interpolation_radius,
			unsigned int bin_capacity = 16u, const std::string& script = std::string(),
			const std::vector<std::string>& options = {});
		~STPRuntimeBiomefieldGenerator() override;
		void operator()(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
			const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, float2 offset) override;
		// Compiler log of the script (may be empty).
		const std::string& log() const noexcept { return this->Log; }
		// The built-in script's source.
		static const char* builtinScript() noexcept;

	private:
		STPAlgorithm::STPPermutationGenerator::STPPermutationTable Table;
		uint2 MapSize;
		unsigned int InterpolationRadius, BinCapacity;
		std::string Log;
		STPDeviceRuntimeProgram::STPSmartModule Module;
		hipFunction_t Kernel = nullptr;
		// the batched path's filter (lines split into segments) for square maps of at most 8 biomes
		// and bins, whose histograms are handed to the script in the reference's jagged form
		struct STPFilterEngine;
		std::unique_ptr<STPFilterEngine> Engine;
	};

	// The demo's runtime-compiled splatmap generator (SuperDemo+/World/Biomes/STPSplatmapGenerator.cpp:27-100
	// compiling SuperDemo+/Script/STPSplatmapGenerator.cu) on hipRTC: the script reads the rendered
	// area's biome map and UINT16 heightfield from device arrays instead of texture objects.
	class STPRuntimeSplatmapGenerator {
	public:
		using STPChunkInformation = STPDiversity::STPTextureInformation::STPSplatGeneratorInformation::STPLocalChunkInformation;

		// `perm` and `database` are device views that must outlive the generator. `script`: empty =
		// the built-in script; `options`: extra compiler options.
		STPRuntimeSplatmapGenerator(const STPAlgorithm::STPPermutation& perm,
			const STPDiversity::STPTextureInformation::STPSplatRuleDatabase& database, uint2 map_size,
			const std::string& script = std::string(), const std::vector<std::string>& options = {});
		~STPRuntimeSplatmapGenerator();
		STPRuntimeSplatmapGenerator(const STPRuntimeSplatmapGenerator&) = delete;
		STPRuntimeSplatmapGenerator& operator=(const STPRuntimeSplatmapGenerator&) = delete;

		// STPTextureFactory::splat: the region (255 = none) of every pixel of the listed chunks of the
		// rendered area (area.x x area.y, row-major). Asynchronous on `stream`.
		void splat(const unsigned short* biome, const unsigned short* height, unsigned char* splat, uint2 area,
			const std::vector<STPChunkInformation>& chunks, hipStream_t stream);
		const std::string& log() const noexcept { return this->Log; }
		static const char* builtinScript() noexcept;

	private:
		uint2 MapSize;
		std::string Log;
		STPDeviceRuntimeProgram::STPSmartModule Module;
		hipFunction_t Kernel = nullptr;
		// the chunk list of a launch, through pinned staging
		STPSmartDeviceMemory::STPDevice<STPChunkInformation> Chunk;
		STPSmartDeviceMemory::STPHost<STPChunkInformation> Staging;
		size_t Capacity = 0u;
		hipEvent_t Uploaded = nullptr;
	};

} // namespace SuperTerrainPlus

#endif // STP_AMD_DEVICE_RUNTIME_HPP
