// hipRTC runtime compilation and linking, and the runtime-compiled multi-biome diversity
// generator.
//
// Compile: SuperTerrain+/SuperTerrain+/Private/GPGPU/STPDeviceRuntimeBinary.cpp (options, name
// expressions, external headers, log, lowered names; a failed compilation throws with the log).
// Link: SuperTerrain+/SuperTerrain+/Private/GPGPU/STPDeviceRuntimeProgram.cpp (per-input options,
// archives, module options).
// Plugin: SuperDemo+/World/Biomes/STPBiomefieldGenerator.cpp:44-152 compiles
// SuperDemo+/Script/STPMultiHeightGenerator.cu at run time, looks its kernel and globals up by
// lowered name, fills the biome table and launches the kernel after the single histogram filter.
#include "stp_batch.hpp"
#include "stp_internal.hpp"

#include <stp/STPDeviceRuntime.hpp>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <mutex>

// The library's device headers, embedded at build time (generated source).
extern "C" {
extern const char* const STPLibraryHeaderName[];
extern const char* const STPLibraryHeaderSource[];
extern const unsigned int STPLibraryHeaderCount;
}

using namespace SuperTerrainPlus;

namespace {

	void checkRTC(const hiprtcResult r, const char* what) {
		if (r != HIPRTC_SUCCESS) {
			STP_THROW(STPCompilationError, std::string(what) + " -> " + hiprtcGetErrorString(r));
		}
	}

	bool isIdentifier(const std::string& n) {
		if (n.empty() || !(std::isalpha(static_cast<unsigned char>(n[0])) || n[0] == '_')) {
			return false;
		}
		return std::all_of(n.begin(), n.end(), [](const char c) { return std::isalnum(static_cast<unsigned char>(c)) || c == '_'; });
	}

	bool hasOption(const std::vector<std::string>& opt, const char* prefix) {
		return std::any_of(opt.begin(), opt.end(), [prefix](const std::string& o) { return o.rfind(prefix, 0) == 0; });
	}

	// Architecture of the current device, gfx950 when no device is visible (compilation alone
	// needs no GPU).
	std::string deviceArchitecture() {
		int count = 0, device = 0;
		if (hipGetDeviceCount(&count) != hipSuccess || count == 0 || hipGetDevice(&device) != hipSuccess) {
			(void)hipGetLastError();
			return "gfx950";
		}
		hipDeviceProp_t prop;
		if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
			(void)hipGetLastError();
			return "gfx950";
		}
		return prop.gcnArchName;
	}

} // namespace

// ---------------------------------------------------------------- STPDeviceRuntimeBinary

STPDeviceRuntimeBinary::STPSourceInformation::STPSourceArgument&
STPDeviceRuntimeBinary::STPSourceInformation::STPSourceArgument::operator[](const std::string& arg) {
	this->StringArgument.push_back(arg);
	return *this;
}

void STPDeviceRuntimeBinary::STPProgramDeleter::operator()(hiprtcProgram program) const {
	(void)hiprtcDestroyProgram(&program);
}

STPDeviceRuntimeBinary::STPCompilationOutput STPDeviceRuntimeBinary::compile(std::string&& source_name,
	const std::string& source_code, const STPSourceInformation& source_info, const STPExternalHeaderSource& external_header) {
	// external headers named by the source information, in its order
	std::vector<const char*> headerName, headerSource;
	for (const std::string& name : source_info.ExternalHeader.StringArgument) {
		const auto it = external_header.find(name);
		if (it == external_header.end()) {
			STP_THROW(STPInvalidArgument, "external header '" + name + "' has no source");
		}
		headerName.push_back(it->first.c_str());
		headerSource.push_back(it->second.c_str());
	}
	hiprtcProgram raw = nullptr;
	checkRTC(hiprtcCreateProgram(&raw, source_code.c_str(), source_name.c_str(), static_cast<int>(headerName.size()),
				 headerSource.data(), headerName.data()),
		"hiprtcCreateProgram");
	STPSmartProgram program(raw);

	const std::vector<std::string>& names = source_info.NameExpression.StringArgument;
	for (const std::string& n : names) {
		checkRTC(hiprtcAddNameExpression(program.get(), n.c_str()), "hiprtcAddNameExpression");
	}

	std::vector<std::string> opt = source_info.Option.StringArgument;
	if (!hasOption(opt, "--offload-arch") && !hasOption(opt, "--gpu-architecture") && !hasOption(opt, "-arch")) {
		opt.push_back("--offload-arch=" + deviceArchitecture());
	}
	if (!hasOption(opt, "-O")) {
		opt.push_back("-O3");
	}
	if (!hasOption(opt, "-ffp-contract")) {
		opt.push_back("-ffp-contract=off");
	}
	std::vector<const char*> optPtr;
	for (const std::string& o : opt) {
		optPtr.push_back(o.c_str());
	}
	const hiprtcResult compiled = hiprtcCompileProgram(program.get(), static_cast<int>(optPtr.size()), optPtr.data());

	STPCompilationOutput out;
	size_t logSize = 0u;
	checkRTC(hiprtcGetProgramLogSize(program.get(), &logSize), "hiprtcGetProgramLogSize");
	if (logSize > 1u) {
		std::string log(logSize, '\0');
		checkRTC(hiprtcGetProgramLog(program.get(), log.data()), "hiprtcGetProgramLog");
		log.resize(std::strlen(log.c_str()));
		out.Log = std::move(log);
	}
	if (compiled != HIPRTC_SUCCESS) {
		STP_THROW(STPCompilationError, "compilation of '" + source_name + "' failed (" + hiprtcGetErrorString(compiled)
			+ "):\n" + out.Log);
	}
	for (const std::string& n : names) {
		const char* lowered = nullptr;
		if (hiprtcGetLoweredName(program.get(), n.c_str(), &lowered) == HIPRTC_SUCCESS && lowered) {
			out.LoweredName.emplace(n, lowered);
		} else if (isIdentifier(n) && source_code.find(n) != std::string::npos) {
			// hipRTC does not lower global variables of relocatable (-fgpu-rdc) programs; a
			// variable at global scope keeps its name in the code object
			out.LoweredName.emplace(n, n);
		}
	}
	out.ProgramObject.Identifier = std::move(source_name);
	out.ProgramObject.Program = std::move(program);
	return out;
}

STPDeviceRuntimeBinary::STPProgramData STPDeviceRuntimeBinary::readBitcode(hiprtcProgram program) {
	size_t size = 0u;
	checkRTC(hiprtcGetBitcodeSize(program, &size), "hiprtcGetBitcodeSize (the program needs -fgpu-rdc)");
	STPProgramData d(std::make_unique<char[]>(size), size);
	checkRTC(hiprtcGetBitcode(program, d.first.get()), "hiprtcGetBitcode");
	return d;
}

STPDeviceRuntimeBinary::STPProgramData STPDeviceRuntimeBinary::readCodeObject(hiprtcProgram program) {
	size_t size = 0u;
	checkRTC(hiprtcGetCodeSize(program, &size), "hiprtcGetCodeSize");
	if (size == 0u) {
		STP_THROW(STPCompilationError, "the program has no code object (it was compiled with -fgpu-rdc)");
	}
	STPProgramData d(std::make_unique<char[]>(size), size);
	checkRTC(hiprtcGetCode(program, d.first.get()), "hiprtcGetCode");
	return d;
}

const STPDeviceRuntimeBinary::STPExternalHeaderSource& STPDeviceRuntimeBinary::libraryHeaders() {
	static const STPExternalHeaderSource headers = [] {
		STPExternalHeaderSource h;
		for (unsigned int i = 0u; i < STPLibraryHeaderCount; i++) {
			h.emplace(STPLibraryHeaderName[i], STPLibraryHeaderSource[i]);
		}
		return h;
	}();
	return headers;
}

// ---------------------------------------------------------------- STPDeviceRuntimeProgram

STPDeviceRuntimeProgram::STPLinkerInformation::STPDataJitOption&
STPDeviceRuntimeProgram::STPLinkerInformation::STPDataJitOption::operator()(const hipJitOption flag, void* const value) {
	this->OptionFlag.push_back(flag);
	this->OptionValue.push_back(value);
	return *this;
}

void STPDeviceRuntimeProgram::STPModuleDeleter::operator()(hipModule_t module) const {
	(void)hipModuleUnload(module);
}

namespace {
	using STPDataJitOption = STPDeviceRuntimeProgram::STPLinkerInformation::STPDataJitOption;

	void checkOption(const STPDataJitOption& o) {
		if (o.OptionFlag.size() != o.OptionValue.size()) {
			STP_THROW(STPInvalidArgument, "every JIT option flag needs exactly one value");
		}
	}

	struct LinkState {
		hiprtcLinkState State = nullptr;
		~LinkState() {
			if (this->State) {
				(void)hiprtcLinkDestroy(this->State);
			}
		}
	};
} // namespace

STPDeviceRuntimeProgram::STPSmartModule STPDeviceRuntimeProgram::link(STPLinkerInformation& linker_info) {
	checkOption(linker_info.LinkerOption);
	checkOption(linker_info.ModuleOption);
	if (linker_info.DataOption.empty()) {
		STP_THROW(STPInvalidArgument, "nothing to link");
	}
	size_t codeObjects = 0u;
	for (const auto& [binary, type, option] : linker_info.DataOption) {
		if (!binary || !binary->Program) {
			STP_THROW(STPInvalidArgument, "a linker input has no compiled program");
		}
		checkOption(option);
		codeObjects += type == STPBinaryType::CodeObject ? 1u : 0u;
	}
	const auto load = [&linker_info](const void* image) {
		hipModule_t module = nullptr;
		STPDataJitOption& mo = linker_info.ModuleOption;
		const hipError_t r = hipModuleLoadDataEx(&module, image, static_cast<unsigned int>(mo.OptionFlag.size()),
			mo.OptionFlag.data(), mo.OptionValue.data());
		if (r != hipSuccess) {
			(void)hipGetLastError();
			STP_THROW(STPCompilationError, std::string("hipModuleLoadDataEx -> ") + hipGetErrorString(r));
		}
		return STPSmartModule(module);
	};

	if (codeObjects > 0u) {
		// a code object is already linked: it can only be loaded on its own
		if (codeObjects != 1u || linker_info.DataOption.size() != 1u || !linker_info.ArchiveOption.empty()) {
			STP_THROW(STPInvalidArgument,
				"a code object cannot be linked with other inputs; compile every input with -fgpu-rdc for bitcode");
		}
		const STPDeviceRuntimeBinary::STPProgramData code =
			STPDeviceRuntimeBinary::readCodeObject(std::get<0>(linker_info.DataOption.front())->Program.get());
		return load(code.first.get());
	}

	LinkState state;
	STPDataJitOption& lo = linker_info.LinkerOption;
	checkRTC(hiprtcLinkCreate(static_cast<unsigned int>(lo.OptionFlag.size()), lo.OptionFlag.data(), lo.OptionValue.data(),
				 &state.State),
		"hiprtcLinkCreate");
	std::vector<STPDeviceRuntimeBinary::STPProgramData> bitcode; // alive until the link completes
	for (auto& [binary, type, option] : linker_info.DataOption) {
		bitcode.push_back(STPDeviceRuntimeBinary::readBitcode(binary->Program.get()));
		const STPDeviceRuntimeBinary::STPProgramData& bc = bitcode.back();
		checkRTC(hiprtcLinkAddData(state.State, HIPRTC_JIT_INPUT_LLVM_BITCODE, bc.first.get(), bc.second,
					 binary->Identifier.c_str(), static_cast<unsigned int>(option.OptionFlag.size()), option.OptionFlag.data(),
					 option.OptionValue.data()),
			"hiprtcLinkAddData");
	}
	for (auto& [file, option] : linker_info.ArchiveOption) {
		checkOption(option);
		checkRTC(hiprtcLinkAddFile(state.State, HIPRTC_JIT_INPUT_LLVM_ARCHIVES_OF_BUNDLED_BITCODE, file.c_str(),
					 static_cast<unsigned int>(option.OptionFlag.size()), option.OptionFlag.data(), option.OptionValue.data()),
			"hiprtcLinkAddFile");
	}
	void* image = nullptr;
	size_t imageSize = 0u;
	checkRTC(hiprtcLinkComplete(state.State, &image, &imageSize), "hiprtcLinkComplete");
	return load(image);
}

// ---------------------------------------------------------------- STPRuntimeBiomefieldGenerator

extern "C" {
extern const char* const STPLibraryMultiHeightScript;
}

namespace {
	constexpr char KernelName[] = "generateMultiBiomeHeightmap";
	constexpr unsigned int ScriptTileX = 64u, ScriptTileY = 4u;

	struct ScriptGlobal {
		hipDeviceptr_t Pointer = nullptr;
		size_t Bytes = 0u;
	};

	ScriptGlobal scriptGlobal(hipModule_t module, const STPDeviceRuntimeBinary::STPLoweredName& lowered,
		const std::string& name, const size_t need) {
		const auto it = lowered.find(name);
		if (it == lowered.end()) {
			STP_THROW(STPCompilationError, "the generator script does not define '" + name + "'");
		}
		ScriptGlobal g;
		STP_CHECK_HIP(hipModuleGetGlobal(&g.Pointer, &g.Bytes, module, it->second.c_str()));
		if (g.Bytes < need) {
			STP_THROW(STPInvalidArgument, "the script's '" + name + "' holds " + std::to_string(g.Bytes) + " bytes, "
				+ std::to_string(need) + " are needed");
		}
		return g;
	}
} // namespace

const char* STPRuntimeBiomefieldGenerator::builtinScript() noexcept {
	return STPLibraryMultiHeightScript;
}

// The filter engine: a one-chunk batch of the map's side with an external biome map (stages 0-2);
// calls take it in turn.
struct STPRuntimeBiomefieldGenerator::STPFilterEngine {
	std::mutex Lock;
	bool Checked = false;
	std::vector<STPBiomeProperty> Biome;
	std::unique_ptr<STPHeightfieldBatch> Batch; // null: the segmented passes do not apply
};

STPRuntimeBiomefieldGenerator::~STPRuntimeBiomefieldGenerator() = default;

STPRuntimeBiomefieldGenerator::STPRuntimeBiomefieldGenerator(const STPEnvironment::STPSimplexNoiseSetting& noise,
	std::vector<STPBiomeProperty> table, const uint2 map_size, const unsigned int interpolation_radius,
	const unsigned int bin_capacity, const std::string& script, const std::vector<std::string>& options) :
	Table(STPAlgorithm::STPPermutationGenerator::generate(noise)), MapSize(map_size),
	InterpolationRadius(interpolation_radius), BinCapacity(bin_capacity), Engine(std::make_unique<STPFilterEngine>()) {
	if (table.empty() || map_size.x == 0u || map_size.y == 0u) {
		STP_THROW(STPInvalidArgument, "the runtime generator needs a biome table and a map size");
	}
	if (bin_capacity == 0u || bin_capacity > STPHeightfieldKernel::MaxBinCapacity) {
		STP_THROW(STPInvalidArgument, "histogram bin capacity must be in [1, 64]");
	}
	STPDeviceRuntimeBinary::STPSourceInformation info;
	info.Option["-std=c++17"];
	for (const std::string& o : options) {
		info.Option[o];
	}
	info.NameExpression[KernelName]["Dimension"]["HalfDimension"]["Permutation"]["BiomeCount"]["BiomeTable"];
	const STPDeviceRuntimeBinary::STPExternalHeaderSource& headers = STPDeviceRuntimeBinary::libraryHeaders();
	for (const auto& h : headers) {
		info.ExternalHeader[h.first];
	}
	STPDeviceRuntimeBinary::STPCompilationOutput out = STPDeviceRuntimeBinary::compile("STPMultiHeightGenerator",
		script.empty() ? std::string(STPLibraryMultiHeightScript) : script, info, headers);
	this->Log = out.Log;

	STPDeviceRuntimeProgram::STPLinkerInformation link;
	link.DataOption.emplace_back(&out.ProgramObject, STPDeviceRuntimeProgram::STPBinaryType::CodeObject,
		STPDeviceRuntimeProgram::STPLinkerInformation::STPDataJitOption {});
	this->Module = STPDeviceRuntimeProgram::link(link);
	const auto kernel = out.LoweredName.find(KernelName);
	if (kernel == out.LoweredName.end()) {
		STP_THROW(STPCompilationError, std::string("the generator script does not define ") + KernelName);
	}
	STP_CHECK_HIP(hipModuleGetFunction(&this->Kernel, this->Module.get(), kernel->second.c_str()));

	// the globals are constant for the generator's lifetime: filled once here
	const unsigned int count = static_cast<unsigned int>(table.size());
	const float2 half = make_float2(static_cast<float>(map_size.x) / 2.0f, static_cast<float>(map_size.y) / 2.0f);
	const STPAlgorithm::STPPermutation perm = this->Table.Permutation;
	const auto put = [&](const char* name, const void* src, const size_t bytes) {
		const ScriptGlobal g = scriptGlobal(this->Module.get(), out.LoweredName, name, bytes);
		STP_CHECK_HIP(hipMemcpyHtoD(g.Pointer, const_cast<void*>(src), bytes));
	};
	put("Dimension", &map_size, sizeof(uint2));
	put("HalfDimension", &half, sizeof(float2));
	put("Permutation", &perm, sizeof(perm));
	put("BiomeCount", &count, sizeof(unsigned int));
	put("BiomeTable", table.data(), table.size() * sizeof(STPBiomeProperty));
	this->Engine->Biome = std::move(table);
}

void STPRuntimeBiomefieldGenerator::operator()(const STPNearestNeighbourHeightFloatWTextureBuffer& heightmap_buffer,
	const STPNearestNeighbourSampleRTextureBuffer& biomemap_buffer, const float2 offset) {
	const STPNearestNeighbourInformation& bnn = biomemap_buffer.NeighbourInfo;
	if (bnn.MapSize.x != this->MapSize.x || bnn.MapSize.y != this->MapSize.y
		|| heightmap_buffer.NeighbourInfo.MapSize.x != this->MapSize.x
		|| heightmap_buffer.NeighbourInfo.MapSize.y != this->MapSize.y) {
		STP_THROW(STPInvalidArgument, "the map size differs from the size this generator was made for");
	}
	const uint2 start = STPHeightfieldKernel::checkFilterDomain(bnn, this->InterpolationRadius);
	const auto pool = heightmap_buffer.DeviceMemInfo.first;
	const hipStream_t stream = heightmap_buffer.DeviceMemInfo.second;
	const STPNearestNeighbourHeightFloatWTextureBuffer::STPMergedBuffer height(heightmap_buffer,
		STPNearestNeighbourHeightFloatWTextureBuffer::STPMemoryLocation::DeviceMemory);
	const STPNearestNeighbourSampleRTextureBuffer::STPMergedBuffer biome(biomemap_buffer,
		STPNearestNeighbourSampleRTextureBuffer::STPMemoryLocation::DeviceMemory);
	const uint2 m = this->MapSize;
	const auto launchScript = [&](STPAlgorithm::STPSingleHistogram histogram) {
		float* out = height.getDevice();
		float2 off = offset;
		void* args[] = {&out, &histogram, &off};
		STP_CHECK_HIP(hipModuleLaunchKernel(this->Kernel, (m.x + ScriptTileX - 1u) / ScriptTileX,
			(m.y + ScriptTileY - 1u) / ScriptTileY, 1u, ScriptTileX, ScriptTileY, 1u, 0u, stream, args, nullptr));
	};
	const auto allocate = [&](const size_t bytes) {
		void* p = nullptr;
		if (pool) {
			STP_CHECK_HIP(hipMallocFromPoolAsync(&p, bytes, pool, stream));
		} else {
			STP_CHECK_HIP(hipMallocAsync(&p, bytes, stream));
		}
		return p;
	};
	{
		STPFilterEngine& e = *this->Engine;
		const std::lock_guard<std::mutex> guard(e.Lock);
		if (!e.Checked) {
			e.Checked = true;
			if (m.x == m.y) {
				e.Batch = STPHeightfieldBatch::makeFilter(e.Biome, m.x, this->InterpolationRadius, this->BinCapacity);
			}
		}
		if (e.Batch) {
			STPHeightfieldBatch& b = *e.Batch;
			b.filterExternal(biome.getDevice() + static_cast<size_t>(start.y) * bnn.TotalMapSize.x + start.x,
				bnn.TotalMapSize.x, stream);
			const unsigned int pixels = m.x * m.y, w = 2u * this->InterpolationRadius + 1u;
			void* const jagged = allocate(STPHeightfieldKernel::jaggedWorkspaceBytes(pixels, this->BinCapacity));
			try {
				launchScript(STPHeightfieldKernel::jaggedHistogram(b.binCountOf(0u), b.binItemOf(0u),
					b.packedHistogram() ? nullptr : b.binWeightOf(0u), 1.0f / static_cast<float>(w * w), this->BinCapacity,
					pixels, jagged, stream));
			} catch (...) {
				(void)hipFreeAsync(jagged, stream);
				throw;
			}
			STP_CHECK_HIP(hipFreeAsync(jagged, stream));
			const unsigned int status = b.status(stream);
			if (status & 1u) {
				STP_THROW(STPCapacityError, "a pixel histogram holds more distinct biomes than the bin capacity ("
					+ std::to_string(this->BinCapacity) + ")");
			}
			if ((status & 4u) == 0u) {
				return;
			}
			// a biome id beyond the table: the line-by-line filter keeps it as an item of its own
		}
	}

	const size_t bytes =
		STPHeightfieldKernel::filterWorkspaceBytes(bnn.MapSize, this->InterpolationRadius, this->BinCapacity);
	void* const ws = allocate(bytes);
	const STPHeightfieldKernel::STPFilterWorkspace carved =
		STPHeightfieldKernel::carveFilterWorkspace(ws, bnn.MapSize, this->InterpolationRadius, this->BinCapacity);
	STPHeightfieldKernel::singleHistogramFilter(biome.getDevice(), bnn.TotalMapSize.x, start, bnn.MapSize,
		this->InterpolationRadius, this->BinCapacity, true, carved, stream);
	launchScript(STPAlgorithm::STPSingleHistogram {carved.Bin, carved.Offset});
	unsigned int overflow = 0u;
	STP_CHECK_HIP(hipMemcpyAsync(&overflow, carved.Overflow, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
	STP_CHECK_HIP(hipFreeAsync(ws, stream));
	STP_CHECK_HIP(hipStreamSynchronize(stream));
	if (overflow != 0u) {
		STP_THROW(STPCapacityError, "a pixel histogram holds more distinct biomes than the bin capacity ("
			+ std::to_string(this->BinCapacity) + ")");
	}
}

// ---------------------------------------------------------------- STPRuntimeSplatmapGenerator

extern "C" {
extern const char* const STPLibrarySplatmapScript;
}

namespace {
	constexpr char SplatKernelName[] = "generateTextureSplatmap";
	constexpr unsigned int SplatTileX = 64u, SplatTileY = 4u;
} // namespace

const char* STPRuntimeSplatmapGenerator::builtinScript() noexcept {
	return STPLibrarySplatmapScript;
}

STPRuntimeSplatmapGenerator::STPRuntimeSplatmapGenerator(const STPAlgorithm::STPPermutation& perm,
	const STPDiversity::STPTextureInformation::STPSplatRuleDatabase& database, const uint2 map_size,
	const std::string& script, const std::vector<std::string>& options) :
	MapSize(map_size) {
	if (map_size.x == 0u || map_size.y == 0u || !perm.Permutation || !perm.Gradient2D) {
		STP_THROW(STPInvalidArgument, "the runtime splatmap generator needs a map size and a permutation table");
	}
	STPDeviceRuntimeBinary::STPSourceInformation info;
	info.Option["-std=c++17"];
	for (const std::string& o : options) {
		info.Option[o];
	}
	info.NameExpression[SplatKernelName]["Dimension"]["HalfDimension"]["Permutation"]["SplatDatabase"];
	const STPDeviceRuntimeBinary::STPExternalHeaderSource& headers = STPDeviceRuntimeBinary::libraryHeaders();
	for (const auto& h : headers) {
		info.ExternalHeader[h.first];
	}
	STPDeviceRuntimeBinary::STPCompilationOutput out = STPDeviceRuntimeBinary::compile("STPSplatmapGenerator",
		script.empty() ? std::string(STPLibrarySplatmapScript) : script, info, headers);
	this->Log = out.Log;
	STPDeviceRuntimeProgram::STPLinkerInformation link;
	link.DataOption.emplace_back(&out.ProgramObject, STPDeviceRuntimeProgram::STPBinaryType::CodeObject,
		STPDeviceRuntimeProgram::STPLinkerInformation::STPDataJitOption {});
	this->Module = STPDeviceRuntimeProgram::link(link);
	const auto kernel = out.LoweredName.find(SplatKernelName);
	if (kernel == out.LoweredName.end()) {
		STP_THROW(STPCompilationError, std::string("the splatmap script does not define ") + SplatKernelName);
	}
	STP_CHECK_HIP(hipModuleGetFunction(&this->Kernel, this->Module.get(), kernel->second.c_str()));
	// STPSplatmapGenerator.cpp:44-53: the globals are filled once
	const float2 half = make_float2(static_cast<float>(map_size.x) / 2.0f, static_cast<float>(map_size.y) / 2.0f);
	const auto put = [&](const char* name, const void* src, const size_t bytes) {
		const ScriptGlobal g = scriptGlobal(this->Module.get(), out.LoweredName, name, bytes);
		STP_CHECK_HIP(hipMemcpyHtoD(g.Pointer, const_cast<void*>(src), bytes));
	};
	put("Dimension", &map_size, sizeof(uint2));
	put("HalfDimension", &half, sizeof(float2));
	put("Permutation", &perm, sizeof(perm));
	put("SplatDatabase", &database, sizeof(database));
	STP_CHECK_HIP(hipEventCreateWithFlags(&this->Uploaded, hipEventDisableTiming));
}

STPRuntimeSplatmapGenerator::~STPRuntimeSplatmapGenerator() {
	if (this->Uploaded) {
		(void)hipEventSynchronize(this->Uploaded);
		(void)hipEventDestroy(this->Uploaded);
	}
}

void STPRuntimeSplatmapGenerator::splat(const unsigned short* const biome, const unsigned short* const height,
	unsigned char* const splat, const uint2 area, const std::vector<STPChunkInformation>& chunks, const hipStream_t stream) {
	if (!biome || !height || !splat) {
		STP_THROW(STPInvalidArgument, "null map");
	}
	if (area.x == 0u || area.y == 0u) {
		STP_THROW(STPInvalidArgument, "empty rendered area");
	}
	const uint2 m = this->MapSize;
	for (const STPChunkInformation& c : chunks) {
		if ((static_cast<unsigned long long>(c.LocalChunkCoordinateX) + 1ull) * m.x > area.x
			|| (static_cast<unsigned long long>(c.LocalChunkCoordinateY) + 1ull) * m.y > area.y) {
			STP_THROW(STPInvalidArgument, "a splat chunk lies outside the rendered area");
		}
	}
	if (chunks.empty()) {
		return;
	}
	if (chunks.size() > 65535u) {
		STP_THROW(STPInvalidArgument, "at most 65535 chunks per splatmap launch");
	}
	// the staging buffer may still feed the previous upload
	STP_CHECK_HIP(hipEventSynchronize(this->Uploaded));
	if (chunks.size() > this->Capacity) {
		STP_CHECK_HIP(hipDeviceSynchronize());
		this->Capacity = chunks.size() + chunks.size() / 2u;
		this->Chunk = STPSmartDeviceMemory::makeDevice<STPChunkInformation>(this->Capacity);
		this->Staging = STPSmartDeviceMemory::makeHost<STPChunkInformation>(this->Capacity);
	}
	std::copy(chunks.begin(), chunks.end(), this->Staging.get());
	STP_CHECK_HIP(hipMemcpyAsync(this->Chunk.get(), this->Staging.get(), chunks.size() * sizeof(STPChunkInformation),
		hipMemcpyHostToDevice, stream));
	STP_CHECK_HIP(hipEventRecord(this->Uploaded, stream));
	STPDiversity::STPTextureInformation::STPSplatGeneratorInformation info {this->Chunk.get(),
		static_cast<unsigned int>(chunks.size())};
	const unsigned short* b = biome;
	const unsigned short* h = height;
	unsigned char* sp = splat;
	uint2 a = area;
	void* args[] = {&b, &h, &sp, &a, &info};
	STP_CHECK_HIP(hipModuleLaunchKernel(this->Kernel, (m.x + SplatTileX - 1u) / SplatTileX, (m.y + SplatTileY - 1u) / SplatTileY,
		static_cast<unsigned int>(chunks.size()), SplatTileX, SplatTileY, 1u, 0u, stream, args, nullptr));
}
