/*
 * stp_amd.h -- C ABI of the MI355X-native SuperTerrain+ heightfield hot path.
 *
 * The reference exposes this path only as C++ classes (STPHeightfieldGenerator,
 * STPDiversityGenerator, STPSingleHistogramFilter, STPPermutationGenerator, ...). This header is
 * the FFI boundary a harness binds with ctypes/cgo/JNI: plain pointers and sizes, `int` status
 * codes, a thread-local last-error string, no exceptions across the boundary. Every entry
 * point cites the reference interface it replaces (paths relative to the reference checkout).
 *
 * Conventions
 *  - Functions return STP_OK (0) or an STP_ERROR_* code; stp_last_error() describes the last
 *    failure on the calling thread.
 *  - `d_` pointers are device (HBM) pointers, `h_` pointers are host pointers.
 *  - `stream` is a hipStream_t passed as void*; NULL is the default stream. Device entry points
 *    only enqueue work unless documented as synchronous.
 *  - Pixel formats follow STPWorldMapPixelFormat.hpp:12-21: biome sample uint16, height float,
 *    fixed-point height uint16, seed uint64.
 */
#ifndef STP_AMD_H
#define STP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
	STP_OK = 0,
	STP_ERROR_INVALID_ENVIRONMENT = 1, /* STPInvalidEnvironment (STPEnvironment.cpp:17-65) */
	STP_ERROR_NUMERIC_DOMAIN = 2,      /* STPNumericDomainError (STPSingleHistogramFilter.cpp:874-880) */
	STP_ERROR_HIP = 3,                 /* STPCUDAError (STPDeviceErrorHandler.hpp:22-58), HIP edition */
	STP_ERROR_CAPACITY = 4,            /* a histogram needs more bins than the configured capacity */
	STP_ERROR_COMPILATION = 5,         /* runtime compilation failed (STPDeviceRuntimeBinary.cpp:80-104) */
	STP_ERROR_INVALID_ARGUMENT = 6,
	STP_ERROR_INTERNAL = 7
};

/* Description of the last error raised on this thread ("" if none). */
const char* stp_last_error(void);

/* STPEngineInitialiser::initialise (SuperTerrain+/SuperTerrain+/Private/STPEngineInitialiser.cpp:37-58):
 * selects the device and creates its context. */
int stp_engine_initialise(int device);

/* ------------------------------------------------------------------ device memory plumbing
 * For callers without their own allocator (the reference wraps these in STPSmartDeviceMemory,
 * SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/Utility/Memory/STPSmartDeviceMemory.inl:47-68). */
int stp_device_malloc(void** d_ptr, size_t bytes);
int stp_device_free(void* d_ptr);
int stp_host_malloc_pinned(void** h_ptr, size_t bytes);
int stp_host_free_pinned(void* h_ptr);
/* Copy of `bytes` bytes in any direction (unified addressing). Asynchronous on `stream`. */
int stp_memcpy_async(void* dst, const void* src, size_t bytes, void* stream);
int stp_stream_create(void** stream);
int stp_stream_destroy(void* stream);
int stp_stream_synchronize(void* stream);

/* ------------------------------------------------------------------ settings */
/* STPSimplexNoiseSetting (SuperTerrain+/SuperAlgorithm+/Host/Public/SuperAlgorithm+Host/STPSimplexNoiseSetting.h:14-27)
 * + Engine: 0 = std::mt19937_64 (reference), 1 = std::mt19937. */
typedef struct {
	uint64_t Seed;
	uint32_t Distribution;
	double Offset;
	uint32_t Engine;
} stp_simplex_noise_setting;

/* STPRainDropSetting (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/Environment/STPRainDropSetting.h:15-53)
 * + MaxLifetime: steps per drop, 0 = unlimited (reference behaviour). */
typedef struct {
	uint64_t Seed;
	uint32_t RainDropCount;
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
	uint32_t ErosionBrushRadius;
	uint32_t MaxLifetime;
} stp_raindrop_setting;

/* STPChunkSetting (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/Environment/STPChunkSetting.h:15-45). */
typedef struct {
	uint32_t ChunkSize[2];
	uint32_t MapSize[2];
	double ChunkOffset[3];
	double ChunkScale[2];
	double MapOffset[2];
	uint32_t DiversityNearestNeighbour[2];
	uint32_t ErosionNearestNeighbour[2];
	uint32_t SplatNearestNeighbour[2];
	uint32_t RenderDistance[2];
} stp_chunk_setting;

/* STPNearestNeighbourInformation (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Chunk/STPNearestNeighbourInformation.hpp:13-25). */
typedef struct {
	uint32_t MapSize[2];
	uint32_t ChunkNearestNeighbour[2];
	uint32_t TotalMapSize[2];
} stp_nn_information;

/* STPSimplexNoise::STPFractalSimplexInformation (SuperTerrain+/SuperAlgorithm+/Device/Public/SuperAlgorithm+Device/STPSimplexNoise.cuh:29-52). */
typedef struct {
	float Persistence;
	float Lacunarity;
	uint32_t Octave;
	float Scale;
	float Offset[2];
	float HalfDimension[2];
	float Amplitude;
	float Frequency;
} stp_fractal_simplex_information;

/* STPDemo::STPBiomeProperty (SuperDemo+/World/Biomes/STPBiomeProperty.hpp:10-27). */
typedef struct {
	float Scale;
	uint32_t Octave;
	float Persistence;
	float Lacunarity;
	float Depth;
	float Variation;
} stp_biome_property;

/* STPPermutation (SuperTerrain+/SuperAlgorithm+/Host/Public/SuperAlgorithm+Host/STPPermutation.hpp:9-22): device view. */
typedef struct {
	const uint8_t* Permutation;
	const float* Gradient2D;
	uint32_t Gradient2DSize;
} stp_permutation;

/* STPSingleHistogram (SuperTerrain+/SuperAlgorithm+/Host/Public/SuperAlgorithm+Host/STPSingleHistogram.hpp:15-50).
 * Bin is an array of 8-byte records {uint16 Item; 2 pad bytes; float Weight}. */
typedef struct {
	const void* Bin;
	const uint32_t* HistogramStartOffset;
} stp_single_histogram;

/* STP*Setting::validate() (SuperTerrain+/SuperTerrain+/Private/STPEnvironment.cpp:17-65,
 * SuperTerrain+/SuperAlgorithm+/Host/Private/STPAlgorithmEnvironment.cpp:13-17). The rain drop
 * check also rejects settings whose drops could live for more than 65535 steps. */
int stp_simplex_noise_setting_validate(const stp_simplex_noise_setting* setting);
int stp_raindrop_setting_validate(const stp_raindrop_setting* setting);
int stp_chunk_setting_validate(const stp_chunk_setting* setting);

/* ------------------------------------------------------------------ host-side tables */
/* STPPermutationGenerator::generate (SuperTerrain+/SuperAlgorithm+/Host/Private/STPPermutationGenerator.cpp:40-94):
 * h_permutation[512], h_gradient[2 * Distribution] (gradient_capacity floats available). */
int stp_permutation_generate(const stp_simplex_noise_setting* setting, uint8_t* h_permutation, float* h_gradient,
	uint32_t gradient_capacity);

/* Same, but the tables are uploaded to device memory owned by the returned handle;
 * `out_view` receives the device view (STPPermutationGenerator::STPPermutationTable,
 * SuperTerrain+/SuperAlgorithm+/Host/Public/SuperAlgorithm+Host/STPPermutationGenerator.h:23-43). */
typedef struct stp_permutation_table stp_permutation_table;
int stp_permutation_table_create(const stp_simplex_noise_setting* setting, stp_permutation_table** out,
	stp_permutation* out_view);
int stp_permutation_table_destroy(stp_permutation_table* table);

/* generateErosionBrush (SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPHeightfieldGenerator.cpp:44-99). */
int stp_erosion_brush_generate(uint32_t dimension_x, uint32_t radius, int32_t* h_index, float* h_weight,
	uint32_t capacity, uint32_t* brush_size);

/* ------------------------------------------------------------------ chunk math
 * STPChunk (SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPChunk.cpp:63-108). */
int stp_chunk_world_coordinate(const double point[3], const uint32_t chunk_size[2], const double scale[2],
	int32_t out[2]);
int stp_chunk_local_coordinate(uint32_t chunk_id, const uint32_t chunk_range[2], uint32_t out[2]);
int stp_chunk_local_origin(const int32_t centre[2], const uint32_t chunk_size[2], const uint32_t neighbour_size[2],
	int32_t out[2]);
int stp_chunk_map_offset(const int32_t chunk[2], const uint32_t chunk_size[2], const uint32_t map_size[2],
	const double map_offset[2], double out[2]);
int stp_chunk_offset(const int32_t chunk[2], const uint32_t chunk_size[2], const int32_t offset[2], int32_t out[2]);
/* out receives region_size[0] * region_size[1] (x, y) pairs. */
int stp_chunk_neighbour_offset(const uint32_t chunk_size[2], const uint32_t region_size[2], int32_t* out,
	uint32_t capacity);

/* ------------------------------------------------------------------ device kernels (asynchronous) */
/* Single-biome heightmap: h = simplex2DFractal(x, y) * variation + depth, row-major width x height
 * (STPSimplexNoise.cu:84-110 via SuperDemo+/Script/STPMultiHeightGenerator.cu:57-71). */
int stp_simplex_heightmap(const stp_permutation* perm, const stp_fractal_simplex_information* info, float variation,
	float depth, float* d_height, uint32_t width, uint32_t height, void* stream);

/* generateMultiBiomeHeightmap (SuperDemo+/Script/STPMultiHeightGenerator.cu:35-55): blends the biomes
 * of `histogram` (device, jagged) with the host biome table (<= 64 entries). */
int stp_multi_biome_heightmap(const stp_permutation* perm, const stp_biome_property* h_biome_table,
	uint32_t biome_count, const stp_single_histogram* histogram, float* d_height, uint32_t width, uint32_t height,
	float offset_x, float offset_y, void* stream);

/* formatHeightmap (SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:51-57,112-122). */
int stp_format_heightmap(const float* d_height, uint16_t* d_fixed, uint32_t width, uint32_t height, void* stream);

/* hydraulicErosion (SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:37-49,80-108) on a merged
 * free-slip map. mode 0 = parallel (FP32 atomics), 1 = serialized (bit-exact). stream_id is the
 * Philox counter word that distinguishes repeated erosions with one seed. d_ledger (optional,
 * RainDropCount floats) receives each drop's final carried sediment; d_drop_info (optional,
 * 2 x uint32 per drop) receives {steps | erode_steps << 16, deposit_steps | killed << 16}. */
int stp_hydraulic_erosion(float* d_height, const stp_nn_information* nn, const stp_raindrop_setting* setting,
	const int32_t* d_brush_index, const float* d_brush_weight, uint32_t brush_size, uint32_t stream_id, int mode,
	float* d_ledger, uint32_t* d_drop_info, void* stream);

/* STPNearestNeighbourTextureBuffer merge / unmerge (SuperTerrain+/SuperTerrain+/Private/World/Chunk/
 * STPNearestNeighbourTextureBuffer.cpp:69-113) as device gather / scatter. h_chunk_ptrs is a host
 * array of NN.x * NN.y device pointers in row-major neighbour order; element_size is 1, 2 or 4. */
int stp_nn_merge(const void* const* h_chunk_ptrs, void* d_merged, uint32_t element_size, const stp_nn_information* nn,
	void* stream);
int stp_nn_unmerge(const void* d_merged, void* const* h_chunk_ptrs, uint32_t element_size,
	const stp_nn_information* nn, void* stream);

/* Synthetic Voronoi biome map (no reference counterpart; input generator for the multi-biome
 * workloads): nearest of one jittered site per cell_size cell, biome = hash % biome_count. */
int stp_voronoi_biomemap(uint64_t seed, uint32_t biome_count, uint32_t cell_size, int64_t origin_x, int64_t origin_y,
	uint16_t* d_out, uint32_t width, uint32_t height, uint32_t pitch, void* stream);

/* ------------------------------------------------------------------ single histogram filter
 * STPSingleHistogramFilter + STPFilterBuffer (SuperTerrain+/SuperAlgorithm+/Host/Public/SuperAlgorithm+Host/
 * STPSingleHistogramFilter.h:27-208). The buffer owns the device workspace and output. */
typedef struct stp_filter_buffer stp_filter_buffer;
/* bin_capacity (1..64): most distinct items one pixel's histogram may hold. */
int stp_filter_buffer_create(uint32_t bin_capacity, stp_filter_buffer** out);
int stp_filter_buffer_destroy(stp_filter_buffer* buffer);
/* STPSingleHistogramFilter::operator() (STPSingleHistogramFilter.cpp:870-902) on a DEVICE samplemap laid
 * out as the merged nearest-neighbour map of `nn`. Synchronous on `stream` (it checks the
 * capacity flag). `out` receives device pointers valid until the next call on `buffer`;
 * bin_total receives the bin count. */
int stp_single_histogram_filter(stp_filter_buffer* buffer, const uint16_t* d_samplemap, const stp_nn_information* nn,
	uint32_t radius, stp_single_histogram* out, uint64_t* bin_total, void* stream);
/* Same with a HOST samplemap and HOST (pinned) output, the reference's own calling convention. */
int stp_single_histogram_filter_host(stp_filter_buffer* buffer, const uint16_t* h_samplemap,
	const stp_nn_information* nn, uint32_t radius, stp_single_histogram* out, uint64_t* bin_total);

/* ------------------------------------------------------------------ heightfield generator
 * STPHeightfieldGenerator (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Chunk/STPHeightfieldGenerator.h:26-161)
 * with a built-in STPDiversityGenerator (STPDiversityGenerator.hpp:14-37). */
typedef struct stp_diversity_generator stp_diversity_generator;
typedef struct stp_heightfield_generator stp_heightfield_generator;

/* Single-biome plugin: h = simplex2DFractal * variation + depth over one chunk. */
int stp_diversity_simplex_create(const stp_simplex_noise_setting* noise, const stp_biome_property* biome,
	const uint32_t map_size[2], stp_diversity_generator** out);
/* Multi-biome plugin (STPBiomefieldGenerator, SuperDemo+/World/Biomes/STPBiomefieldGenerator.cpp:79-152):
 * device single histogram filter of `radius` over the diversity neighbourhood, then the blend. */
int stp_diversity_biomefield_create(const stp_simplex_noise_setting* noise, const stp_biome_property* h_table,
	uint32_t biome_count, const uint32_t map_size[2], uint32_t radius, uint32_t bin_capacity,
	stp_diversity_generator** out);
/* Multi-biome plugin whose blend kernel is compiled at run time with hipRTC (the reference demo's
 * NVRTC-compiled STPMultiHeightGenerator.cu, SuperDemo+/World/Biomes/STPBiomefieldGenerator.cpp:44-77).
 * `script` NULL = the built-in script; `options` are extra compiler options. A script that does
 * not compile gives STP_ERROR_COMPILATION with the compiler log in stp_last_error(). */
int stp_diversity_runtime_create(const stp_simplex_noise_setting* noise, const stp_biome_property* h_table,
	uint32_t biome_count, const uint32_t map_size[2], uint32_t radius, uint32_t bin_capacity, const char* script,
	const char* const* options, uint32_t option_count, stp_diversity_generator** out);
int stp_diversity_destroy(stp_diversity_generator* generator);

int stp_heightfield_generator_create(const stp_chunk_setting* chunk, const stp_raindrop_setting* raindrop,
	stp_diversity_generator* diversity, stp_heightfield_generator** out);
int stp_heightfield_generator_destroy(stp_heightfield_generator* generator);
/* 0 = parallel (default, as the reference's release build), 1 = serialized. */
int stp_heightfield_generator_set_erosion_mode(stp_heightfield_generator* generator, int mode);

/* STPHeightfieldGenerator::generate (STPHeightfieldGenerator.h:145, .cpp:158-169): HOST heightfield of one
 * chunk, HOST biomemaps of the diversity neighbourhood (row-major neighbour order). Synchronous. */
int stp_heightfield_generate(stp_heightfield_generator* generator, float* h_heightfield,
	const uint16_t* const* h_biomemap, float offset_x, float offset_y);
/* STPHeightfieldGenerator::erode (STPHeightfieldGenerator.h:159, .cpp:171-197): HOST maps of the erosion
 * neighbourhood, eroded in place, plus their UINT16 copies. Synchronous. */
int stp_heightfield_erode(stp_heightfield_generator* generator, float* const* h_heightfield,
	uint16_t* const* h_fixed);
/* Device-resident variants: every pointer is a DEVICE pointer (the pointer arrays are host
 * arrays). Asynchronous on `stream`; the call returns once the work is enqueued. */
int stp_heightfield_generate_device(stp_heightfield_generator* generator, float* d_heightfield,
	const uint16_t* const* d_biomemap, float offset_x, float offset_y, void* stream);
int stp_heightfield_erode_device(stp_heightfield_generator* generator, float* const* d_heightfield,
	uint16_t* const* d_fixed, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation):
 * stp_heightfield_erode_device with the merged free-slip map in the library's fixed device scratch
 * (STP_EROSION_SCRATCH_PIXELS FP32 pixels, 16 MiB per device, held until the process ends) instead
 * of the generator's memory pool. It calls no allocation or free. A merged map larger than the
 * scratch gives STP_ERROR_CAPACITY. Every call on one device shares that device's scratch and
 * leaves it holding its merged map: calls that overlap on one device race. */
#define STP_EROSION_SCRATCH_PIXELS (4u << 20)
int stp_heightfield_erode_device_noalloc(stp_heightfield_generator* generator, float* const* d_heightfield,
	uint16_t* const* d_fixed, void* stream);

/* ------------------------------------------------------------------ runtime compilation (hipRTC)
 * STPDeviceRuntimeBinary::compile (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/GPGPU/
 * STPDeviceRuntimeBinary.h:26-127) and STPDeviceRuntimeProgram::link (.../STPDeviceRuntimeProgram.h:26-94).
 * Compilation needs no GPU; linking and loading a module do. */
typedef struct stp_rtc_program stp_rtc_program;
typedef struct stp_rtc_module stp_rtc_module;
enum { STP_RTC_CODE_OBJECT = 0, STP_RTC_LLVM_BITCODE = 1 };

/* Compiles `source`. Name expressions name kernels, template instantiations and global
 * variables whose lowered names should be kept; header_names/header_sources are external
 * headers the source includes by name; library_headers != 0 also offers the library's device
 * headers (stp/STPSimplexNoise.hpp, stp/STPSingleHistogramWrapper.hpp, stp/STPRandom.hpp,
 * stp/STPBiomeProperty.hpp). --offload-arch (current device, else gfx950), -O3 and
 * -ffp-contract=off are added unless given. A failed compilation returns STP_ERROR_COMPILATION
 * and stp_last_error() holds the compiler log. */
int stp_rtc_compile(const char* name, const char* source, const char* const* options, uint32_t option_count,
	const char* const* name_expressions, uint32_t name_count, const char* const* header_names,
	const char* const* header_sources, uint32_t header_count, int library_headers, stp_rtc_program** out);
int stp_rtc_program_destroy(stp_rtc_program* program);
/* Copies the compiler log (NUL-terminated, truncated to `capacity`); *length = full length. */
int stp_rtc_program_log(const stp_rtc_program* program, char* h_text, uint64_t capacity, uint64_t* length);
/* Lowered name of a registered name expression; STP_ERROR_INVALID_ARGUMENT if it was not found. */
int stp_rtc_lowered_name(const stp_rtc_program* program, const char* name_expression, char* h_text,
	uint64_t capacity, uint64_t* length);
/* Links programs into a module on the current device: one STP_RTC_CODE_OBJECT program alone, or
 * STP_RTC_LLVM_BITCODE programs (compiled with -fgpu-rdc) through the hipRTC linker. */
int stp_rtc_link(stp_rtc_program* const* programs, const int* binary_types, uint32_t count, stp_rtc_module** out);
int stp_rtc_module_destroy(stp_rtc_module* module);
int stp_rtc_module_function(stp_rtc_module* module, const char* lowered_name, void** function);
int stp_rtc_module_global(stp_rtc_module* module, const char* lowered_name, void** d_pointer, uint64_t* bytes);
/* hipModuleLaunchKernel; `args` holds one pointer per kernel argument. Asynchronous. */
int stp_rtc_launch(void* function, const uint32_t grid[3], const uint32_t block[3], uint32_t shared_bytes,
	void* stream, void** args);

/* ------------------------------------------------------------------ batched chunks
 * No reference counterpart (the reference makes one generate()/erode() call per chunk, each ending
 * with a synchronisation, STPHeightfieldGenerator.cpp:158-197). A batch holds many square chunks of
 * uneven size resident in device memory and runs the whole path for all of them: Voronoi biome
 * map with a radius halo -> single histogram filter -> multi-biome blend -> parallel erosion on
 * the chunk alone -> UINT16. */
typedef struct stp_layer_factory stp_layer_factory;

typedef struct {
	uint64_t BiomeSeed;
	uint32_t BiomeCount;
	uint32_t BiomeCell;
	uint32_t Radius;       /* filter radius, even */
	uint32_t BinCapacity;  /* 1..64 */
	float DropsPerPixel;   /* drops of a chunk = side^2 * DropsPerPixel */
	/* NULL: the synthetic Voronoi biome map (BiomeSeed, BiomeCell). Otherwise the biome map of every
	 * chunk and its filter halo comes from this layer chain (the reference's biome factory); ids at
	 * or beyond BiomeCount are written as 0 and flagged (stp_layer_factory_status). The factory
	 * must outlive the batch, server or world. */
	stp_layer_factory* BiomeLayers;
} stp_batch_setting;

typedef struct {
	uint32_t Side;         /* square chunk side in pixels */
	int32_t Origin[2];     /* world coordinate of the chunk's first pixel */
	uint32_t Step;         /* world units per pixel, a power of two (level of detail); 0 means 1 */
} stp_batch_chunk;

typedef struct stp_heightfield_batch stp_heightfield_batch;

/* raindrop->RainDropCount is ignored (per-chunk counts come from DropsPerPixel). */
int stp_batch_create(const stp_simplex_noise_setting* noise, const stp_biome_property* h_table, uint32_t biome_count,
	const stp_raindrop_setting* raindrop, const stp_batch_setting* setting, const stp_batch_chunk* h_chunks,
	uint32_t chunk_count, stp_heightfield_batch** out);
int stp_batch_destroy(stp_heightfield_batch* batch);
/* mode 0: replay of one captured hipGraph (asynchronous; a binding of at least 8 Mi pixels runs as
 * four chunk groups of about equal pixels whose whole pipelines are concurrent chains of the
 * graph; with timing on, one graph per stage with events between, so no stage overlaps another);
 * 1: the same launches without a graph (asynchronous); 2: plain -- one chunk per launch sequence
 * with a host synchronisation after each chunk, the reference's calling pattern (synchronous). */
int stp_batch_run(stp_heightfield_batch* batch, int mode, void* stream);
/* Record HIP events between the 7 stages (biome, vertical filter, horizontal filter, blend,
 * erosion drop binning, erosion, format) of later runs. */
int stp_batch_set_timing(stp_heightfield_batch* batch, int enable);
/* Stage milliseconds of the last timed run; call after the run's stream has completed. */
int stp_batch_stage_times(stp_heightfield_batch* batch, float* h_ms, uint32_t capacity);
/* Host wall milliseconds of each chunk of the last plain run. */
int stp_batch_chunk_latencies(stp_heightfield_batch* batch, double* h_ms, uint32_t capacity);
/* h_out = {pixels, drops, drop steps, erode steps, deposit steps, killed drops} of the last run
 * (synchronises `stream`); fails with STP_ERROR_CAPACITY if a histogram overflowed. */
int stp_batch_counters(stp_heightfield_batch* batch, uint64_t h_out[6], void* stream);
/* No reference counterpart: *h_status = the status bits of the last run (synchronises `stream`):
 * 1 a pixel histogram held more distinct biomes than the bin capacity, 2 a cell of the LDS-tiled
 * erosion's fixed-point window ended outside [-4, 7.5] (heights piled up towards the window's
 * range of [-8, 8)); the results of such a run are not to be used. stp_batch_counters and
 * stp_batch_bin_total return STP_ERROR_CAPACITY for either. */
int stp_batch_status(stp_heightfield_batch* batch, uint32_t* h_status, void* stream);
/* No reference counterpart (a measurement aid): *h_out = the sum over every pixel of the bound
 * chunks of its histogram bin count in the last run, i.e. the (pixel, biome) noise evaluations of
 * the blend (synchronises `stream`; STP_ERROR_CAPACITY if a histogram overflowed). */
int stp_batch_bin_total(stp_heightfield_batch* batch, uint64_t* h_out, void* stream);
/* Launch shape of the erosion stage. ErosionTile > 0: LDS-tiled erosion -- the drops that spawn
 * in one ErosionTile x ErosionTile tile run on an LDS copy of the tile plus a halo of
 * H = MaxLifetime + max(r, 2) + 1 cells, and the tiles run in four colour batches (tile
 * coordinate parities) whose windows are disjoint. Needs MaxLifetime > 0, ErosionTile >= 2 H, a
 * window that fits in LDS, and every biome's heights (Depth to Depth + Variation) within [0, 4]:
 * the window holds heights in 32-bit fixed point, and below 0 the erode update can drive them
 * without bound. A batch outside that range erodes with 0. 0: every drop works on the heightmap in
 * device memory with global atomics. ErosionThreads: threads of one tile workgroup, 256, 512, 768
 * or 1024 (768: two tile-72 windows per CU at six waves per SIMD; offered by the autotuner for the
 * 25-tap brush, whose drop fits the 80 registers per lane that leaves). */
typedef struct {
	uint32_t ErosionTile;
	uint32_t ErosionThreads;
} stp_batch_tuning;

/* STP_ERROR_INVALID_ARGUMENT if the shape cannot run on this batch. */
int stp_batch_set_tuning(stp_heightfield_batch* batch, const stp_batch_tuning* tuning);
int stp_batch_get_tuning(stp_heightfield_batch* batch, stp_batch_tuning* out);
/* Autotuner: times every candidate shape with graph runs of the whole batch on `stream` (the
 * pipeline as stp_batch_run runs it, `repeat` timed runs after one warm-up each), keeps the
 * fastest and writes it to `out` (optional). Synchronous. */
int stp_batch_autotune(stp_heightfield_batch* batch, uint32_t repeat, void* stream, stp_batch_tuning* out);
/* Copies chunk `index`'s outputs to host memory (each pointer optional): heights and UINT16
 * (side^2), halo biome map ((side + 2 * Radius)^2), per-drop final sediment (drops; one value per
 * drop of the chunk, in the order the erosion ran them: drop index for the global-atomic kernel,
 * tile order for the LDS-tiled one). */
int stp_batch_read_chunk(stp_heightfield_batch* batch, uint32_t index, float* h_height, uint16_t* h_fixed,
	uint16_t* h_biome_halo, float* h_ledger);

/* ------------------------------------------------------------------ serving loop
 * No reference counterpart (nearest: the world pipeline's asynchronous chunk loader,
 * SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:1504-1633). Chunk requests arrive
 * one by one; `Engines` batches run concurrently on their own streams; requests that arrive while
 * batches run join the next batch (continuous batching); results stay resident in device memory
 * within MemoryBudget, least recently used evicted first. A chunk's result does not depend on the
 * batch it runs in (its drops use a Philox stream word made from its coordinate and side). */
typedef struct {
	uint64_t MemoryBudget;   /* device bytes for the batch engines plus the resident chunks */
	uint64_t BatchPixels;    /* pixel capacity of one batch (>= MaxSide^2) */
	uint32_t BatchChunks;    /* chunk capacity of one batch (1..65535) */
	uint32_t MaxSide;        /* largest chunk side, a power of two */
	uint32_t Engines;        /* batches in flight (1..8) */
	uint32_t ChunkWorldSize; /* world units of one grid chunk, a power of two >= MaxSide */
} stp_server_setting;

/* Grid chunk (X, Y) at Side pixels per side (a power of two): world origin (X, Y) * ChunkWorldSize,
 * ChunkWorldSize / Side world units per pixel. */
typedef struct {
	int32_t X, Y;
	uint32_t Side;
} stp_chunk_request;

typedef struct {
	double P50, P99, P999, Mean, Max;  /* request latency, milliseconds */
	double MpixelPerSecond;            /* completed pixels / replay wall time */
	double DeadlineMissFraction;
	uint64_t Requests, CacheHits, Batches, Evictions;
	double MeanBatchChunks;
	uint64_t PeakResidentBytes;
	/* the requests whose chunk was generated for them (neither resident nor already in
	 * production when they were admitted): count, latency percentiles (ms) and their pixels per
	 * second of replay wall time; the fields above cover every request, cache hits included */
	uint64_t Generated;
	double GeneratedP50, GeneratedP99, GeneratedP999, GeneratedMax;
	double GeneratedMpixelPerSecond;
} stp_server_stats;

typedef struct stp_heightfield_server stp_heightfield_server;

/* raindrop->RainDropCount is ignored (per-chunk counts come from batch->DropsPerPixel). */
int stp_server_create(const stp_simplex_noise_setting* noise, const stp_biome_property* h_table, uint32_t biome_count,
	const stp_raindrop_setting* raindrop, const stp_batch_setting* batch, const stp_server_setting* setting,
	stp_heightfield_server** out);
int stp_server_destroy(stp_heightfield_server* server);
/* Queues a request (arrival = now) and returns its ticket. */
int stp_server_submit(stp_heightfield_server* server, const stp_chunk_request* request, uint64_t* ticket);
/* One scheduler pass (never waits for the GPU), then returns up to `capacity` completed tickets
 * with their latency in milliseconds; *count = how many were written. */
int stp_server_poll(stp_heightfield_server* server, uint64_t* h_ticket, double* h_latency_ms, uint32_t capacity,
	uint32_t* count);
/* Waits until every queued request has completed (the completions stay available to poll). */
int stp_server_flush(stp_heightfield_server* server);
/* Copies a resident chunk to host memory (either pointer may be NULL); STP_ERROR_INVALID_ARGUMENT if
 * it is not resident. Synchronous. */
int stp_server_read(stp_heightfield_server* server, const stp_chunk_request* request, float* h_height,
	uint16_t* h_fixed);
/* Device bytes held by the batch engines and by the resident chunks. */
int stp_server_memory(stp_heightfield_server* server, uint64_t* engine_bytes, uint64_t* resident_bytes);
/* Replays a trace in real time: request i is submitted h_arrival_s[i] seconds after the start
 * (non-decreasing). h_latency_ms[i] receives each request's latency. Synchronous. */
int stp_server_replay(stp_heightfield_server* server, const stp_chunk_request* h_requests, const double* h_arrival_s,
	uint32_t count, double deadline_ms, double* h_latency_ms, stp_server_stats* out);
/* No reference counterpart (a test aid): the Philox stream word of the drops of chunk (X, Y) at
 * Side pixels, the same in every batch. Every served chunk is produced with a free-slip halo of
 * MaxLifetime + max(r, 2) + 1 pixels of its neighbours' terrain at its level of detail and its
 * drops spawn in the chunk, so it erodes as the centre of its 3 x 3 neighbourhood of uneroded
 * chunks does (no drop reaches beyond the halo); the halo is discarded. */
int stp_server_stream_word(int32_t x, int32_t y, uint32_t side, uint32_t* word);

/* ------------------------------------------------------------------ splatmap
 * The texture region of every pixel of the rendered area's chunks (STPWorldPipeline.cpp:671-707
 * generateSplatmap running the demo script SuperDemo+/Script/STPSplatmapGenerator.cu:65-137; rule
 * lookups STPTextureSplatRuleWrapper, SuperTerrain+/SuperAlgorithm+/Device/Private/STPTextureSplatRuleWrapper.cu:56-126;
 * rule database STPTextureFactory.cpp:137-176). */
typedef struct {
	uint16_t Sample;    /* biome id */
	float UpperBound;   /* altitudes up to this take Region */
	uint32_t Region;
} stp_altitude_rule;

typedef struct {
	uint16_t Sample;
	float minGradient, maxGradient; /* the slope factor's range */
	float LowerBound, UpperBound;   /* the altitude's range */
	uint32_t Region;
} stp_gradient_rule;

/* A chunk of the rendered area (STPSplatGeneratorInformation::STPLocalChunkInformation,
 * STPTextureInformation.hpp:120-128): its tile position in the area and its map offset. */
typedef struct {
	uint32_t LocalX, LocalY;
	float OffsetX, OffsetY;
} stp_splat_chunk;

typedef struct stp_splat_database stp_splat_database;

/* Builds the rule database in device memory: samples with rules ascending, altitude rules by
 * (sample, upper bound), gradient rules by sample in the given order (STPTextureDatabase.cpp:357-411). */
int stp_splat_database_create(const stp_altitude_rule* h_altitude, uint32_t altitude_count,
	const stp_gradient_rule* h_gradient, uint32_t gradient_count, stp_splat_database** out);
int stp_splat_database_destroy(stp_splat_database* database);
/* d_biome (UINT16 biome ids) and d_height (UINT16 heights) cover the rendered area, area_x x area_y
 * row-major; for every listed chunk, its map_x x map_y pixels get their region (255 = none) in
 * d_splat. The noise's half dimension is map / 2. Asynchronous on `stream`. */
int stp_splatmap_generate(const stp_permutation* perm, const stp_splat_database* database, const uint16_t* d_biome,
	const uint16_t* d_height, uint8_t* d_splat, uint32_t area_x, uint32_t area_y, uint32_t map_x, uint32_t map_y,
	const stp_splat_chunk* h_chunks, uint32_t count, void* stream);

/* ------------------------------------------------------------------ biome layers
 * Biome maps from a chain of layers (the reference's STPBiomeFactory over STPLayer trees,
 * SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/Diversity/STPBiomeFactory.h:16-93 and STPLayer.h:20-260,
 * with the demo layers of SuperDemo+/World/Layers and its chain STPAllLayers.cpp:55-110), evaluated on the GPU.
 * A chain is a list of nodes, parents before children, the last node the root. */
typedef struct {
	uint32_t Kind;      /* 0 Continent, 1 Scale, 2 Land, 3 Island, 4 Voronoi, 5 BaseBiome, 6 ClimateSingle,
	                       7 ClimateModerate, 8 ClimateExtreme, 9 DeepOcean, 10 EaseEdge, 11 EdgeBiome, 12 Hills,
	                       13 Noise, 14 NoiseToRiver, 15 OceanNoise, 16 OceanTemperate, 17 OceanExtreme,
	                       18 OceanTransition, 19 OceanMix, 20 RiverErode, 21 RiverMix, 22 SmoothScale,
	                       23 custom (Param = the custom layer's index) */
	uint32_t Param;     /* Scale: 0 normal, 1 fuzzy. Voronoi (2D): seed hash 0 identity (libstdc++), 1 FNV-1a (MSVC) */
	int32_t Parent[2];  /* node indices, -1 for none; Hills (land, noise), OceanMix (land, ocean), RiverMix (land, river) */
	uint64_t Salt;      /* STPLayer::Salt */
	uint64_t CacheSize; /* STPLayer cache size: 0 or a power of 2 (validated; no effect on the GPU) */
} stp_layer_node;

/* Ids, temperatures and precipitations of the demo's 35 named biomes in registration order
 * (STPBiomeRegistry.cpp:66-104): Ocean, DeepOcean, WarmOcean, LukewarmOcean, ColdOcean, FrozenOcean,
 * DeepWarmOcean, DeepLukewarmOcean, DeepColdOcean, DeepFrozenOcean, River, FrozenRiver, Plains, Desert,
 * Mountain, Forest, Taiga, SnowyTaiga, SnowyTundra, Jungle, Savannah, Swamp, Badlands, DesertHills, TaigaHills,
 * WoodedMountain, SnowyMountain, ForestHills, JungleHills, SavannahPlateau, SwampHills, BadlandsPlateau,
 * Beach, SnowyBeach, StoneShore. */
typedef struct {
	uint16_t ID[35];
	float Temperature[35];
	float Precipitation[35];
} stp_biome_registry;

/* A layer of the caller's own (an STPLayer subclass, STPLayer.h:153-165), compiled at run time: Source
 * defines `__device__ unsigned short sample(const SuperTerrainPlus::STPLayerDevice::STPLayerContext& c, int x,
 * int z)` (stp/STPLayerDevice.hpp); a node of Kind 23 with Param = the custom layer's index runs it.
 * A sample at (x, z) may read parent coordinates from ((x + Lo) >> Shift, (z + Lo) >> Shift) to
 * ((x + Hi) >> Shift, (z + Hi) >> Shift) of its 1 or 2 parents; a read outside gives 0 and status bit 0. */
typedef struct {
	const char* Source;
	int32_t Lo, Hi;
	uint32_t Shift;
	uint32_t Parents;
	const char* const* Options;
	uint32_t OptionCount;
} stp_custom_layer;

/* The demo's chain (STPAllLayers.cpp:55-110, 14 nodes) into h_nodes (capacity >= 14). */
int stp_layer_program_demo(stp_layer_node* h_nodes, uint32_t capacity, uint32_t* count);
/* The demo's registry after loading SuperDemo+/Biome.ini (ocean 0, plains 1, desert 2, forest 3, the
 * other biomes id 0). */
int stp_biome_registry_demo(stp_biome_registry* out);
/* STPBiomeFactory(dimension) with the chain as its production line; STP_ERROR_NUMERIC_DOMAIN for an
 * empty dimension or a cache size that is not a power of 2 (STPBiomeFactory.cpp:20-22, STPLayer.cpp:68-73). */
int stp_layer_factory_create(const stp_layer_node* h_nodes, uint32_t count, uint64_t global_seed,
	const stp_biome_registry* registry, uint32_t dimension_x, uint32_t dimension_y, stp_layer_factory** out);
/* As stp_layer_factory_create with custom layers; STP_ERROR_COMPILATION (compiler log in
 * stp_last_error()) for one that does not compile. */
int stp_layer_factory_create_custom(const stp_layer_node* h_nodes, uint32_t count, uint64_t global_seed,
	const stp_biome_registry* registry, uint32_t dimension_x, uint32_t dimension_y, const stp_custom_layer* h_custom,
	uint32_t custom_count, stp_layer_factory** out);
int stp_layer_factory_destroy(stp_layer_factory* factory);
/* STPBiomeFactory::operator()(biomemap, offset): pixel (x, y) is the root sampled at world
 * (x + offset_x, 0, y + offset_y). Synchronous, host memory. */
int stp_layer_factory_generate(stp_layer_factory* factory, uint16_t* h_biomemap, int32_t offset_x, int32_t offset_y);
/* `count` maps into device memory one after another, h_offsets = count (x, y) pairs. Asynchronous
 * on `stream`; one factory is used by one thread and one stream at a time (batches, servers and
 * worlds that share it keep their own layer buffers). */
int stp_layer_factory_generate_device(stp_layer_factory* factory, uint16_t* d_biomemap, const int32_t* h_offsets,
	uint32_t count, void* stream);
/* Bit 0: a layer read outside its rectangle (an internal error); bit 1: an id at or beyond a batch's
 * BiomeCount was replaced by 0. Reading clears them. Synchronous. */
int stp_layer_factory_status(stp_layer_factory* factory, uint32_t* status);

/* The demo's runtime-compiled splatmap generator (SuperDemo+/World/Biomes/STPSplatmapGenerator.cpp:27-100
 * over SuperDemo+/Script/STPSplatmapGenerator.cu) on hipRTC. `script` NULL = the built-in script
 * (superterrainplus_amd/rtc/STPSplatmapGenerator.rtc); `perm` and `database` must outlive it. A
 * script that does not compile gives STP_ERROR_COMPILATION with the compiler log. */
typedef struct stp_splatmap_runtime stp_splatmap_runtime;
int stp_splatmap_runtime_create(const stp_permutation* perm, const stp_splat_database* database, uint32_t map_x,
	uint32_t map_y, const char* script, const char* const* options, uint32_t option_count, stp_splatmap_runtime** out);
int stp_splatmap_runtime_destroy(stp_splatmap_runtime* generator);
/* As stp_splatmap_generate, through the compiled script. Asynchronous on `stream`. */
int stp_splatmap_runtime_generate(stp_splatmap_runtime* generator, const uint16_t* d_biome, const uint16_t* d_height,
	uint8_t* d_splat, uint32_t area_x, uint32_t area_y, const stp_splat_chunk* h_chunks, uint32_t count, void* stream);

/* ------------------------------------------------------------------ world pipeline
 * The reference world pipeline's generation chain (STPWorldPipeline.cpp: requestChunk :640-663,
 * erodeHeightmap :485-559, generateHeightmap :437-476, scheduleChunk :339-404) on resident device
 * state. A requested chunk is complete once it has been eroded as the centre of its
 * ErosionNeighbour x ErosionNeighbour neighbourhood (drops spawn in the centre chunk, the erosion
 * runs free-slip over the merged neighbourhood and writes every chunk of it); that needs the
 * heightmap of every chunk of the neighbourhood first. Centres whose neighbourhoods overlap are
 * eroded in different batches. Chunk (X, Y) starts at world pixel (X, Y) * MapSize. */
typedef struct {
	uint32_t MapSize;          /* chunk side in pixels (STPChunkSetting::MapSize) */
	uint32_t ErosionNeighbour; /* odd, <= 9 (STPChunkSetting::ErosionNearestNeighbour) */
	uint32_t GenerateChunks;   /* heightmaps generated per batch */
	uint32_t ErodeChunks;      /* neighbourhoods eroded per batch */
} stp_world_setting;

typedef struct stp_heightfield_world stp_heightfield_world;

/* batch->DropsPerPixel gives the drops of one erosion: MapSize^2 * DropsPerPixel (the reference's
 * RainDropCount); raindrop->RainDropCount is ignored. */
int stp_world_create(const stp_simplex_noise_setting* noise, const stp_biome_property* h_table, uint32_t biome_count,
	const stp_raindrop_setting* raindrop, const stp_batch_setting* batch, const stp_world_setting* setting,
	stp_heightfield_world** out);
int stp_world_destroy(stp_heightfield_world* world);
/* Makes every requested chunk (h_chunks: count (X, Y) pairs) complete; writes up to `capacity`
 * chunks whose maps changed (the union of the eroded neighbourhoods, the reference's
 * RequireHeightmapReload) to h_changed as (X, Y) pairs and their number to *changed_count.
 * Synchronous. */
int stp_world_request(stp_heightfield_world* world, const int32_t* h_chunks, uint32_t count, int32_t* h_changed,
	uint32_t capacity, uint32_t* changed_count);
/* Copies a chunk with a map (either pointer may be NULL); STP_ERROR_INVALID_ARGUMENT if it has none. */
int stp_world_read(stp_heightfield_world* world, int32_t x, int32_t y, float* h_height, uint16_t* h_fixed);
/* 0 no map, 1 heightmap generated, 2 complete */
int stp_world_state(stp_heightfield_world* world, int32_t x, int32_t y, uint32_t* state);
/* The rendered area around chunk (x, y): its (2 radius + 1)^2 chunks are made complete, then their
 * UINT16 heightmaps and biome map are written side by side into device buffers of side
 * (2 radius + 1) * MapSize, row-major (the reference's rendering buffer, STPWorldPipeline.cpp:671-707;
 * the input of stp_splatmap_generate). Synchronous. */
int stp_world_render_area(stp_heightfield_world* world, int32_t x, int32_t y, uint32_t radius, uint16_t* d_biome,
	uint16_t* d_height);
/* STPWorldPipeline::load (SuperTerrain+/SuperTerrain+/Private/World/STPWorldPipeline.cpp:1504-1633,
 * STPWorldPipeline.h:51-60,139): the rendered area around chunk (x, y) ((2 radius + 1)^2 chunks, made
 * complete, as stp_world_render_area assembles it) is loaded on a worker thread into the back buffer
 * while the caller renders from the front buffer. *status: 0 unchanged (the same centre, nothing
 * finished since), 1 back buffer busy (a load is running and this call is ignored, or a changed
 * centre's load has just started, discarding a finished load that was not swapped), 2 swapped (the
 * load of this centre finished; front and back buffers were swapped). Returns at once; every other
 * stp_world_* call on this world waits for a running load. */
int stp_world_load(stp_heightfield_world* world, int32_t x, int32_t y, uint32_t radius, uint32_t* status);
/* The front buffer: device biome and UINT16 height maps of the last swapped load, *side =
 * (2 radius + 1) * MapSize pixels a side, row-major (NULL and 0 before the first swap); valid until
 * the next load that returns 2. Any pointer may be NULL. */
int stp_world_front(stp_heightfield_world* world, const uint16_t** d_biome, const uint16_t** d_height, uint32_t* side);
/* ------------------------------------------------------------------ chunk records
 * The byte format in which chunk maps leave the device (files, sockets): the reference's map pixel
 * formats (STPWorldMapPixelFormat.hpp:12-21: biome u16, height f32 and u16, row-major per chunk as
 * STPChunk holds them, STPChunk.cpp:17-25) behind a 64-byte little-endian header; the layout is this
 * library's own (the reference keeps chunks in memory). Header: "STPC", u16 version 1, u16 header
 * bytes 64, i32 chunk x, y, i32 world pixel x, y of the first pixel, u32 width, height, u32
 * completeness (STPChunkCompleteness: 0 empty, 1 biome map, 2 heightmap, 3 complete), u32 maps
 * (bit 0 biome, 1 FP32 height, 2 UINT16 height), u64 payload bytes, u32 check, 12 zero bytes. The
 * payload holds the present maps in bit order, each zero-padded to a multiple of 16 bytes;
 * check = sum of the payload's 32-bit words w_i times (2 i + 1), modulo 2^32. */
typedef struct {
	int32_t X, Y, OriginX, OriginY;
	uint32_t Width, Height, Completeness, Maps, Check;
	const uint16_t* BiomeMap;   /* NULL when absent; pointers into the parsed bytes */
	const float* HeightMap;
	const uint16_t* FixedMap;
	uint64_t Bytes;             /* of the whole record */
} stp_chunk_record;

/* Bytes of one record of a width x height chunk with `maps` (width and height in [1, 65536]). */
int stp_chunk_record_bytes(uint32_t width, uint32_t height, uint32_t maps, uint64_t* bytes);
/* Records of the listed chunks (count (X, Y) pairs, each with a map) one after another into h_out
 * (capacity bytes); *written = count * the record size. Synchronous. */
int stp_world_export(stp_heightfield_world* world, const int32_t* h_chunks, uint32_t count, uint32_t maps, void* h_out,
	uint64_t capacity, uint64_t* written);
/* Validates the record at `data` (size bytes available: magic, version, width and height in
 * [1, 65536], payload size, check) and points *record into it; STP_ERROR_INVALID_ARGUMENT on any
 * mismatch. Host only. */
int stp_chunk_record_read(const void* data, uint64_t size, stp_chunk_record* record);

/* The conflict-free schedule the world uses (host only): h_batch[i] = batch of centre i, first
 * fit in the given order; centres in one batch have disjoint neighbourhoods. */
int stp_world_schedule(const int32_t* h_centres, uint32_t count, uint32_t neighbour, uint32_t* h_batch);
/* Philox stream word of the drops of the neighbourhood centred at chunk (X, Y). */
int stp_world_stream_word(int32_t x, int32_t y, uint32_t* word);

#ifdef __cplusplus
}
#endif

#endif /* STP_AMD_H */
