// extern "C" boundary (include/stp_amd.h): converts exceptions into status codes plus a
// thread-local message, and plain C structs into the C++ API types.
#include "stp_batch.hpp"
#include "stp_server.hpp"
#include "stp_splat.hpp"
#include "stp_world.hpp"
#include "stp_chunk_record.hpp"

#include <stp/STPBiomeLayer.hpp>
#include "stp_internal.hpp"

#include <stp/STPDeviceRuntime.hpp>

#include <stp_amd.h>

#include <algorithm>
#include <cstring>
#include <string>

using namespace SuperTerrainPlus;

namespace {

	thread_local std::string LastError;

	template <class F>
	int guarded(F&& f) noexcept {
		try {
			f();
			LastError.clear();
			return STP_OK;
		} catch (const STPException::STPFundamentalException& e) {
			LastError = e.what();
			return static_cast<int>(e.Status);
		} catch (const std::bad_alloc& e) {
			LastError = std::string("out of memory: ") + e.what();
			return STP_ERROR_INTERNAL;
		} catch (const std::exception& e) {
			LastError = e.what();
			return STP_ERROR_INTERNAL;
		} catch (...) {
			LastError = "unknown error";
			return STP_ERROR_INTERNAL;
		}
	}

	void require(const bool cond, const char* what) {
		if (!cond) {
			STP_THROW(STPInvalidArgument, what);
		}
	}

	STPEnvironment::STPSimplexNoiseSetting toNoise(const stp_simplex_noise_setting* s) {
		require(s, "null simplex noise setting");
		STPEnvironment::STPSimplexNoiseSetting n;
		n.Seed = s->Seed;
		n.Distribution = s->Distribution;
		n.Offset = s->Offset;
		n.Engine = static_cast<STPEnvironment::STPPermutationEngine>(s->Engine);
		return n;
	}

	STPBatchSetting toBatchSetting(const stp_batch_setting* s);

	STPEnvironment::STPRainDropSetting toRain(const stp_raindrop_setting* s) {
		require(s, "null raindrop setting");
		STPEnvironment::STPRainDropSetting r;
		r.Seed = s->Seed;
		r.RainDropCount = s->RainDropCount;
		r.Inertia = s->Inertia;
		r.SedimentCapacityFactor = s->SedimentCapacityFactor;
		r.minSedimentCapacity = s->minSedimentCapacity;
		r.initWaterVolume = s->initWaterVolume;
		r.minWaterVolume = s->minWaterVolume;
		r.Friction = s->Friction;
		r.initSpeed = s->initSpeed;
		r.ErodeSpeed = s->ErodeSpeed;
		r.DepositSpeed = s->DepositSpeed;
		r.EvaporateSpeed = s->EvaporateSpeed;
		r.Gravity = s->Gravity;
		r.ErosionBrushRadius = s->ErosionBrushRadius;
		r.MaxLifetime = s->MaxLifetime;
		return r;
	}

	STPRainDropParameter toParam(const STPEnvironment::STPRainDropSetting& s) {
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

	STPEnvironment::STPChunkSetting toChunk(const stp_chunk_setting* s) {
		require(s, "null chunk setting");
		STPEnvironment::STPChunkSetting c;
		c.ChunkSize = make_uint2(s->ChunkSize[0], s->ChunkSize[1]);
		c.MapSize = make_uint2(s->MapSize[0], s->MapSize[1]);
		c.ChunkOffset = make_double3(s->ChunkOffset[0], s->ChunkOffset[1], s->ChunkOffset[2]);
		c.ChunkScale = make_double2(s->ChunkScale[0], s->ChunkScale[1]);
		c.MapOffset = make_double2(s->MapOffset[0], s->MapOffset[1]);
		c.DiversityNearestNeighbour = make_uint2(s->DiversityNearestNeighbour[0], s->DiversityNearestNeighbour[1]);
		c.ErosionNearestNeighbour = make_uint2(s->ErosionNearestNeighbour[0], s->ErosionNearestNeighbour[1]);
		c.SplatNearestNeighbour = make_uint2(s->SplatNearestNeighbour[0], s->SplatNearestNeighbour[1]);
		c.RenderDistance = make_uint2(s->RenderDistance[0], s->RenderDistance[1]);
		return c;
	}

	STPNearestNeighbourInformation toNN(const stp_nn_information* s) {
		require(s, "null nearest-neighbour information");
		STPNearestNeighbourInformation n;
		n.MapSize = make_uint2(s->MapSize[0], s->MapSize[1]);
		n.ChunkNearestNeighbour = make_uint2(s->ChunkNearestNeighbour[0], s->ChunkNearestNeighbour[1]);
		n.TotalMapSize = make_uint2(s->TotalMapSize[0], s->TotalMapSize[1]);
		require(n.TotalMapSize.x == n.MapSize.x * n.ChunkNearestNeighbour.x
				&& n.TotalMapSize.y == n.MapSize.y * n.ChunkNearestNeighbour.y && n.MapSize.x > 0u && n.MapSize.y > 0u,
			"inconsistent nearest-neighbour information");
		return n;
	}

	STPAlgorithm::STPPermutation toPerm(const stp_permutation* p) {
		require(p && p->Permutation && p->Gradient2D && p->Gradient2DSize > 0u, "invalid permutation view");
		return STPAlgorithm::STPPermutation {p->Permutation, p->Gradient2D, p->Gradient2DSize};
	}

	STPBiomeProperty toBiome(const stp_biome_property& b) {
		return STPBiomeProperty {b.Scale, b.Octave, b.Persistence, b.Lacunarity, b.Depth, b.Variation};
	}

	hipStream_t toStream(void* const s) {
		return static_cast<hipStream_t>(s);
	}

} // namespace

struct stp_permutation_table {
	STPAlgorithm::STPPermutationGenerator::STPPermutationTable Table;
};

struct stp_filter_buffer {
	STPAlgorithm::STPSingleHistogramFilter Filter;
	STPAlgorithm::STPSingleHistogramFilter::STPFilterBuffer Buffer;
	explicit stp_filter_buffer(const unsigned int cap) :
		Buffer(STPAlgorithm::STPSingleHistogramFilter::STPFilterBuffer::STPExecutionType::Parallel, cap) {}
};

struct stp_diversity_generator {
	std::unique_ptr<STPDiversityGenerator> Generator;
};

struct stp_rtc_program {
	STPDeviceRuntimeBinary::STPCompilationOutput Output;
};

struct stp_rtc_module {
	STPDeviceRuntimeProgram::STPSmartModule Module;
};

struct stp_splat_database {
	std::unique_ptr<STPSplatRuleTable> Table;
};

struct stp_splatmap_runtime {
	std::unique_ptr<STPRuntimeSplatmapGenerator> Generator;
};

struct stp_layer_factory {
	std::unique_ptr<STPDiversity::STPLayerBiomeFactory> Factory;
};

namespace {
	STPBatchSetting toBatchSetting(const stp_batch_setting* const s) {
		require(s, "null batch setting");
		STPBatchSetting bs;
		bs.BiomeSeed = s->BiomeSeed;
		bs.BiomeCount = s->BiomeCount;
		bs.BiomeCell = s->BiomeCell;
		bs.Radius = s->Radius;
		bs.BinCapacity = s->BinCapacity;
		bs.DropsPerPixel = s->DropsPerPixel;
		bs.BiomeLayers = s->BiomeLayers ? s->BiomeLayers->Factory.get() : nullptr;
		return bs;
	}
} // namespace

struct stp_heightfield_world {
	std::unique_ptr<STPHeightfieldWorld> World;
};

struct stp_heightfield_server {
	std::unique_ptr<STPHeightfieldServer> Server;
	std::vector<STPServerCompletion> Done; // completed, not yet returned by poll
};

struct stp_heightfield_batch {
	std::unique_ptr<STPHeightfieldBatch> Batch;
};

struct stp_heightfield_generator {
	std::unique_ptr<STPHeightfieldGenerator> Generator;
	STPEnvironment::STPChunkSetting Chunk;
};

extern "C" {

const char* stp_last_error(void) {
	return LastError.c_str();
}

int stp_engine_initialise(const int device) {
	return guarded([&] {
		STP_CHECK_HIP(hipSetDevice(device));
		STP_CHECK_HIP(hipFree(nullptr));
		// the allocation-free erosion's scratch address on this device
		STPHeightfieldGenerator::prepareScratch();
	});
}

int stp_device_malloc(void** const d_ptr, const size_t bytes) {
	return guarded([&] {
		require(d_ptr, "null output pointer");
		*d_ptr = STPSmartDeviceMemory::allocateDevice(bytes);
	});
}

int stp_device_free(void* const d_ptr) {
	return guarded([&] { STP_CHECK_HIP(hipFree(d_ptr)); });
}

int stp_host_malloc_pinned(void** const h_ptr, const size_t bytes) {
	return guarded([&] {
		require(h_ptr, "null output pointer");
		*h_ptr = STPSmartDeviceMemory::allocatePinned(bytes);
	});
}

int stp_host_free_pinned(void* const h_ptr) {
	return guarded([&] { STP_CHECK_HIP(hipHostFree(h_ptr)); });
}

int stp_memcpy_async(void* const dst, const void* const src, const size_t bytes, void* const stream) {
	return guarded([&] { STP_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, toStream(stream))); });
}

int stp_stream_create(void** const stream) {
	return guarded([&] {
		require(stream, "null output pointer");
		hipStream_t s = nullptr;
		STP_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
		*stream = s;
	});
}

int stp_stream_destroy(void* const stream) {
	return guarded([&] { STP_CHECK_HIP(hipStreamDestroy(toStream(stream))); });
}

int stp_stream_synchronize(void* const stream) {
	return guarded([&] { STP_CHECK_HIP(hipStreamSynchronize(toStream(stream))); });
}

int stp_simplex_noise_setting_validate(const stp_simplex_noise_setting* const setting) {
	return guarded([&] { toNoise(setting).validate(); });
}

int stp_raindrop_setting_validate(const stp_raindrop_setting* const setting) {
	return guarded([&] { toRain(setting).validate(); });
}

int stp_chunk_setting_validate(const stp_chunk_setting* const setting) {
	return guarded([&] { toChunk(setting).validate(); });
}

int stp_permutation_generate(const stp_simplex_noise_setting* const setting, uint8_t* const h_permutation,
	float* const h_gradient, const uint32_t gradient_capacity) {
	return guarded([&] {
		const STPEnvironment::STPSimplexNoiseSetting n = toNoise(setting);
		n.validate();
		require(h_permutation && h_gradient, "null output table");
		require(gradient_capacity >= 2u * n.Distribution, "gradient capacity must hold 2 * Distribution floats");
		STPAlgorithm::STPPermutationGenerator::generateHost(n, h_permutation, h_gradient);
	});
}

int stp_permutation_table_create(const stp_simplex_noise_setting* const setting, stp_permutation_table** const out,
	stp_permutation* const out_view) {
	return guarded([&] {
		require(out && out_view, "null output");
		auto t = std::make_unique<stp_permutation_table>();
		t->Table = STPAlgorithm::STPPermutationGenerator::generate(toNoise(setting));
		out_view->Permutation = t->Table.Permutation.Permutation;
		out_view->Gradient2D = t->Table.Permutation.Gradient2D;
		out_view->Gradient2DSize = t->Table.Permutation.Gradient2DSize;
		*out = t.release();
	});
}

int stp_permutation_table_destroy(stp_permutation_table* const table) {
	return guarded([&] { delete table; });
}

int stp_erosion_brush_generate(const uint32_t dimension_x, const uint32_t radius, int32_t* const h_index,
	float* const h_weight, const uint32_t capacity, uint32_t* const brush_size) {
	return guarded([&] {
		std::vector<int> idx;
		std::vector<float> w;
		STPErosionBrushGenerator::generate(dimension_x, radius, idx, w);
		if (brush_size) {
			*brush_size = static_cast<uint32_t>(idx.size());
		}
		if (h_index || h_weight) {
			require(capacity >= idx.size(), "brush capacity too small");
			if (h_index) {
				std::memcpy(h_index, idx.data(), idx.size() * sizeof(int));
			}
			if (h_weight) {
				std::memcpy(h_weight, w.data(), w.size() * sizeof(float));
			}
		}
	});
}

int stp_chunk_world_coordinate(const double point[3], const uint32_t chunk_size[2], const double scale[2], int32_t out[2]) {
	return guarded([&] {
		require(point && chunk_size && scale && out, "null argument");
		const int2 r = STPChunk::calcWorldChunkCoordinate(make_double3(point[0], point[1], point[2]),
			make_uint2(chunk_size[0], chunk_size[1]), make_double2(scale[0], scale[1]));
		out[0] = r.x;
		out[1] = r.y;
	});
}

int stp_chunk_local_coordinate(const uint32_t chunk_id, const uint32_t chunk_range[2], uint32_t out[2]) {
	return guarded([&] {
		require(chunk_range && out && chunk_range[0] > 0u, "invalid argument");
		const uint2 r = STPChunk::calcLocalChunkCoordinate(chunk_id, make_uint2(chunk_range[0], chunk_range[1]));
		out[0] = r.x;
		out[1] = r.y;
	});
}

int stp_chunk_local_origin(const int32_t centre[2], const uint32_t chunk_size[2], const uint32_t neighbour_size[2],
	int32_t out[2]) {
	return guarded([&] {
		require(centre && chunk_size && neighbour_size && out, "null argument");
		const int2 r = STPChunk::calcLocalChunkOrigin(make_int2(centre[0], centre[1]), make_uint2(chunk_size[0], chunk_size[1]),
			make_uint2(neighbour_size[0], neighbour_size[1]));
		out[0] = r.x;
		out[1] = r.y;
	});
}

int stp_chunk_map_offset(const int32_t chunk[2], const uint32_t chunk_size[2], const uint32_t map_size[2],
	const double map_offset[2], double out[2]) {
	return guarded([&] {
		require(chunk && chunk_size && map_size && map_offset && out, "null argument");
		const double2 r = STPChunk::calcChunkMapOffset(make_int2(chunk[0], chunk[1]), make_uint2(chunk_size[0], chunk_size[1]),
			make_uint2(map_size[0], map_size[1]), make_double2(map_offset[0], map_offset[1]));
		out[0] = r.x;
		out[1] = r.y;
	});
}

int stp_chunk_offset(const int32_t chunk[2], const uint32_t chunk_size[2], const int32_t offset[2], int32_t out[2]) {
	return guarded([&] {
		require(chunk && chunk_size && offset && out, "null argument");
		const int2 r = STPChunk::offsetChunk(make_int2(chunk[0], chunk[1]), make_uint2(chunk_size[0], chunk_size[1]),
			make_int2(offset[0], offset[1]));
		out[0] = r.x;
		out[1] = r.y;
	});
}

int stp_chunk_neighbour_offset(const uint32_t chunk_size[2], const uint32_t region_size[2], int32_t* const out,
	const uint32_t capacity) {
	return guarded([&] {
		require(chunk_size && region_size && out, "null argument");
		const std::vector<int2> r = STPChunk::calcChunkNeighbourOffset(make_uint2(chunk_size[0], chunk_size[1]),
			make_uint2(region_size[0], region_size[1]));
		require(capacity >= r.size(), "output capacity too small");
		for (size_t i = 0u; i < r.size(); i++) {
			out[2u * i] = r[i].x;
			out[2u * i + 1u] = r[i].y;
		}
	});
}

int stp_simplex_heightmap(const stp_permutation* const perm, const stp_fractal_simplex_information* const info,
	const float variation, const float depth, float* const d_height, const uint32_t width, const uint32_t height,
	void* const stream) {
	return guarded([&] {
		require(info, "null fractal information");
		STPAlgorithm::STPSimplexNoise::STPFractalSimplexInformation f;
		f.Persistence = info->Persistence;
		f.Lacunarity = info->Lacunarity;
		f.Octave = info->Octave;
		f.Scale = info->Scale;
		f.Offset = make_float2(info->Offset[0], info->Offset[1]);
		f.HalfDimension = make_float2(info->HalfDimension[0], info->HalfDimension[1]);
		f.Amplitude = info->Amplitude;
		f.Frequency = info->Frequency;
		STPHeightfieldKernel::simplexHeightmap(toPerm(perm), f, variation, depth, d_height, width, height, toStream(stream));
	});
}

int stp_multi_biome_heightmap(const stp_permutation* const perm, const stp_biome_property* const h_biome_table,
	const uint32_t biome_count, const stp_single_histogram* const histogram, float* const d_height, const uint32_t width,
	const uint32_t height, const float offset_x, const float offset_y, void* const stream) {
	return guarded([&] {
		require(h_biome_table && histogram, "null argument");
		std::vector<STPBiomeProperty> t;
		for (uint32_t i = 0u; i < biome_count; i++) {
			t.push_back(toBiome(h_biome_table[i]));
		}
		const STPAlgorithm::STPSingleHistogram h {static_cast<const STPAlgorithm::STPSingleHistogram::STPBin*>(histogram->Bin),
			histogram->HistogramStartOffset};
		STPHeightfieldKernel::multiBiomeHeightmap(toPerm(perm), t.data(), biome_count, h, d_height,
			make_uint2(width, height), make_float2(offset_x, offset_y), toStream(stream));
	});
}

int stp_format_heightmap(const float* const d_height, uint16_t* const d_fixed, const uint32_t width, const uint32_t height,
	void* const stream) {
	return guarded([&] {
		STPHeightfieldKernel::formatHeightmap(d_height, d_fixed, static_cast<unsigned long long>(width) * height,
			toStream(stream));
	});
}

int stp_hydraulic_erosion(float* const d_height, const stp_nn_information* const nn, const stp_raindrop_setting* const setting,
	const int32_t* const d_brush_index, const float* const d_brush_weight, const uint32_t brush_size,
	const uint32_t stream_id, const int mode, float* const d_ledger, uint32_t* const d_drop_info, void* const stream) {
	return guarded([&] {
		const STPEnvironment::STPRainDropSetting r = toRain(setting);
		r.validate();
		require(mode == 0 || mode == 1, "erosion mode must be 0 (parallel) or 1 (serialized)");
		const STPNearestNeighbourInformation n = toNN(nn);
		// the brush must be the one generated for this map width and radius (its extent is what
		// the in-bounds argument of the kernel relies on)
		std::vector<int> idx;
		std::vector<float> w;
		STPErosionBrushGenerator::generate(n.TotalMapSize.x, r.ErosionBrushRadius, idx, w);
		require(brush_size == idx.size(), "brush size does not match the brush of this radius and map width");
		// the kernel's in-bounds argument holds for this brush's taps only: check the indices
		require(d_brush_index != nullptr && d_brush_weight != nullptr, "null brush");
		std::vector<int> given(idx.size());
		STP_CHECK_HIP(hipMemcpyAsync(given.data(), d_brush_index, idx.size() * sizeof(int), hipMemcpyDeviceToHost,
			toStream(stream)));
		STP_CHECK_HIP(hipStreamSynchronize(toStream(stream)));
		require(given == idx, "brush indices differ from the brush of this radius and map width");
		STPHeightfieldKernel::hydraulicErosion(d_height, n, toParam(r), d_brush_index, d_brush_weight, brush_size,
			stream_id, static_cast<STPHeightfieldKernel::STPErosionMode>(mode), d_ledger,
			reinterpret_cast<uint2*>(d_drop_info), toStream(stream));
	});
}

int stp_nn_merge(const void* const* const h_chunk_ptrs, void* const d_merged, const uint32_t element_size,
	const stp_nn_information* const nn, void* const stream) {
	return guarded([&] {
		require(h_chunk_ptrs, "null chunk pointer array");
		STPHeightfieldKernel::mergeNeighbour(h_chunk_ptrs, d_merged, element_size, toNN(nn), toStream(stream));
	});
}

int stp_nn_unmerge(const void* const d_merged, void* const* const h_chunk_ptrs, const uint32_t element_size,
	const stp_nn_information* const nn, void* const stream) {
	return guarded([&] {
		require(h_chunk_ptrs, "null chunk pointer array");
		STPHeightfieldKernel::unmergeNeighbour(d_merged, h_chunk_ptrs, element_size, toNN(nn), toStream(stream));
	});
}

int stp_voronoi_biomemap(const uint64_t seed, const uint32_t biome_count, const uint32_t cell_size, const int64_t origin_x,
	const int64_t origin_y, uint16_t* const d_out, const uint32_t width, const uint32_t height, const uint32_t pitch,
	void* const stream) {
	return guarded([&] {
		STPHeightfieldKernel::voronoiBiomemap(seed, biome_count, cell_size, origin_x, origin_y, d_out, width, height, pitch,
			toStream(stream));
	});
}

int stp_filter_buffer_create(const uint32_t bin_capacity, stp_filter_buffer** const out) {
	return guarded([&] {
		require(out, "null output");
		*out = new stp_filter_buffer(bin_capacity);
	});
}

int stp_filter_buffer_destroy(stp_filter_buffer* const buffer) {
	return guarded([&] { delete buffer; });
}

int stp_single_histogram_filter(stp_filter_buffer* const buffer, const uint16_t* const d_samplemap,
	const stp_nn_information* const nn, const uint32_t radius, stp_single_histogram* const out, uint64_t* const bin_total,
	void* const stream) {
	return guarded([&] {
		require(buffer && d_samplemap && out, "null argument");
		const STPAlgorithm::STPSingleHistogram h = buffer->Filter.filterDevice(d_samplemap, toNN(nn), buffer->Buffer, radius,
			toStream(stream));
		out->Bin = h.Bin;
		out->HistogramStartOffset = h.HistogramStartOffset;
		if (bin_total) {
			*bin_total = buffer->Buffer.size().first;
		}
	});
}

int stp_single_histogram_filter_host(stp_filter_buffer* const buffer, const uint16_t* const h_samplemap,
	const stp_nn_information* const nn, const uint32_t radius, stp_single_histogram* const out, uint64_t* const bin_total) {
	return guarded([&] {
		require(buffer && h_samplemap && out, "null argument");
		const STPAlgorithm::STPSingleHistogram h = buffer->Filter(h_samplemap, toNN(nn), buffer->Buffer, radius);
		out->Bin = h.Bin;
		out->HistogramStartOffset = h.HistogramStartOffset;
		if (bin_total) {
			*bin_total = buffer->Buffer.size().first;
		}
	});
}

int stp_diversity_simplex_create(const stp_simplex_noise_setting* const noise, const stp_biome_property* const biome,
	const uint32_t map_size[2], stp_diversity_generator** const out) {
	return guarded([&] {
		require(biome && map_size && out, "null argument");
		auto g = std::make_unique<stp_diversity_generator>();
		g->Generator = std::make_unique<STPSimplexHeightGenerator>(toNoise(noise), toBiome(*biome),
			make_uint2(map_size[0], map_size[1]));
		*out = g.release();
	});
}

int stp_diversity_biomefield_create(const stp_simplex_noise_setting* const noise, const stp_biome_property* const h_table,
	const uint32_t biome_count, const uint32_t map_size[2], const uint32_t radius, const uint32_t bin_capacity,
	stp_diversity_generator** const out) {
	return guarded([&] {
		require(h_table && map_size && out, "null argument");
		std::vector<STPBiomeProperty> t;
		for (uint32_t i = 0u; i < biome_count; i++) {
			t.push_back(toBiome(h_table[i]));
		}
		auto g = std::make_unique<stp_diversity_generator>();
		g->Generator = std::make_unique<STPBiomefieldGenerator>(toNoise(noise), std::move(t),
			make_uint2(map_size[0], map_size[1]), radius, bin_capacity);
		*out = g.release();
	});
}

int stp_diversity_runtime_create(const stp_simplex_noise_setting* const noise, const stp_biome_property* const h_table,
	const uint32_t biome_count, const uint32_t map_size[2], const uint32_t radius, const uint32_t bin_capacity,
	const char* const script, const char* const* const options, const uint32_t option_count,
	stp_diversity_generator** const out) {
	return guarded([&] {
		require(h_table && map_size && out && (options || option_count == 0u), "null argument");
		std::vector<STPBiomeProperty> t;
		for (uint32_t i = 0u; i < biome_count; i++) {
			t.push_back(toBiome(h_table[i]));
		}
		std::vector<std::string> opt;
		for (uint32_t i = 0u; i < option_count; i++) {
			require(options[i], "null option");
			opt.emplace_back(options[i]);
		}
		auto g = std::make_unique<stp_diversity_generator>();
		g->Generator = std::make_unique<STPRuntimeBiomefieldGenerator>(toNoise(noise), std::move(t),
			make_uint2(map_size[0], map_size[1]), radius, bin_capacity, script ? std::string(script) : std::string(), opt);
		*out = g.release();
	});
}

int stp_diversity_destroy(stp_diversity_generator* const generator) {
	return guarded([&] { delete generator; });
}

int stp_heightfield_generator_create(const stp_chunk_setting* const chunk, const stp_raindrop_setting* const raindrop,
	stp_diversity_generator* const diversity, stp_heightfield_generator** const out) {
	return guarded([&] {
		require(diversity && diversity->Generator && out, "null argument");
		auto g = std::make_unique<stp_heightfield_generator>();
		g->Chunk = toChunk(chunk);
		const STPEnvironment::STPRainDropSetting r = toRain(raindrop);
		const STPHeightfieldGenerator::STPGeneratorSetup setup {&g->Chunk, &r, diversity->Generator.get()};
		g->Generator = std::make_unique<STPHeightfieldGenerator>(setup);
		*out = g.release();
	});
}

int stp_heightfield_generator_destroy(stp_heightfield_generator* const generator) {
	return guarded([&] { delete generator; });
}

int stp_heightfield_generator_set_erosion_mode(stp_heightfield_generator* const generator, const int mode) {
	return guarded([&] {
		require(generator && (mode == 0 || mode == 1), "invalid argument");
		generator->Generator->setErosionMode(static_cast<STPHeightfieldGenerator::STPErosionMode>(mode));
	});
}

int stp_heightfield_generate(stp_heightfield_generator* const generator, float* const h_heightfield,
	const uint16_t* const* const h_biomemap, const float offset_x, const float offset_y) {
	return guarded([&] {
		require(generator && h_heightfield && h_biomemap, "null argument");
		generator->Generator->generate(h_heightfield, h_biomemap, make_float2(offset_x, offset_y));
	});
}

int stp_heightfield_erode(stp_heightfield_generator* const generator, float* const* const h_heightfield,
	uint16_t* const* const h_fixed) {
	return guarded([&] {
		require(generator && h_heightfield && h_fixed, "null argument");
		generator->Generator->erode(h_heightfield, h_fixed);
	});
}

int stp_heightfield_generate_device(stp_heightfield_generator* const generator, float* const d_heightfield,
	const uint16_t* const* const d_biomemap, const float offset_x, const float offset_y, void* const stream) {
	return guarded([&] {
		require(generator && d_heightfield && d_biomemap, "null argument");
		generator->Generator->generateDevice(d_heightfield, d_biomemap, make_float2(offset_x, offset_y), toStream(stream));
	});
}

int stp_heightfield_erode_device(stp_heightfield_generator* const generator, float* const* const d_heightfield,
	uint16_t* const* const d_fixed, void* const stream) {
	return guarded([&] {
		require(generator && d_heightfield && d_fixed, "null argument");
		generator->Generator->erodeDevice(d_heightfield, d_fixed, toStream(stream));
	});
}

int stp_heightfield_erode_device_noalloc(stp_heightfield_generator* const generator, float* const* const d_heightfield,
	uint16_t* const* const d_fixed, void* const stream) {
	return guarded([&] {
		require(generator && d_heightfield && d_fixed, "null argument");
		generator->Generator->erodeDeviceNoAlloc(d_heightfield, d_fixed, toStream(stream));
	});
}

namespace {
	void copyText(const std::string& text, char* const h_text, const uint64_t capacity, uint64_t* const length) {
		if (length) {
			*length = text.size();
		}
		if (h_text && capacity > 0u) {
			const size_t n = std::min<size_t>(text.size(), capacity - 1u);
			std::memcpy(h_text, text.data(), n);
			h_text[n] = '\0';
		}
	}
} // namespace

int stp_rtc_compile(const char* const name, const char* const source, const char* const* const options,
	const uint32_t option_count, const char* const* const name_expressions, const uint32_t name_count,
	const char* const* const header_names, const char* const* const header_sources, const uint32_t header_count,
	const int library_headers, stp_rtc_program** const out) {
	return guarded([&] {
		require(name && source && out, "null argument");
		require((options || option_count == 0u) && (name_expressions || name_count == 0u)
				&& ((header_names && header_sources) || header_count == 0u),
			"null argument array");
		STPDeviceRuntimeBinary::STPSourceInformation info;
		for (uint32_t i = 0u; i < option_count; i++) {
			require(options[i], "null option");
			info.Option[options[i]];
		}
		for (uint32_t i = 0u; i < name_count; i++) {
			require(name_expressions[i], "null name expression");
			info.NameExpression[name_expressions[i]];
		}
		STPDeviceRuntimeBinary::STPExternalHeaderSource headers;
		for (uint32_t i = 0u; i < header_count; i++) {
			require(header_names[i] && header_sources[i], "null header");
			headers.emplace(header_names[i], header_sources[i]);
			info.ExternalHeader[header_names[i]];
		}
		if (library_headers) {
			for (const auto& h : STPDeviceRuntimeBinary::libraryHeaders()) {
				if (headers.emplace(h.first, h.second).second) {
					info.ExternalHeader[h.first];
				}
			}
		}
		auto p = std::make_unique<stp_rtc_program>();
		p->Output = STPDeviceRuntimeBinary::compile(std::string(name), std::string(source), info, headers);
		*out = p.release();
	});
}

int stp_rtc_program_destroy(stp_rtc_program* const program) {
	return guarded([&] { delete program; });
}

int stp_rtc_program_log(const stp_rtc_program* const program, char* const h_text, const uint64_t capacity,
	uint64_t* const length) {
	return guarded([&] {
		require(program, "null program");
		copyText(program->Output.Log, h_text, capacity, length);
	});
}

int stp_rtc_lowered_name(const stp_rtc_program* const program, const char* const name_expression, char* const h_text,
	const uint64_t capacity, uint64_t* const length) {
	return guarded([&] {
		require(program && name_expression, "null argument");
		const auto it = program->Output.LoweredName.find(name_expression);
		require(it != program->Output.LoweredName.end(), "the name expression has no lowered name");
		copyText(it->second, h_text, capacity, length);
	});
}

int stp_rtc_link(stp_rtc_program* const* const programs, const int* const binary_types, const uint32_t count,
	stp_rtc_module** const out) {
	return guarded([&] {
		require(programs && binary_types && out && count > 0u, "invalid argument");
		STPDeviceRuntimeProgram::STPLinkerInformation info;
		for (uint32_t i = 0u; i < count; i++) {
			require(programs[i], "null program");
			require(binary_types[i] == STP_RTC_CODE_OBJECT || binary_types[i] == STP_RTC_LLVM_BITCODE, "unknown binary type");
			info.DataOption.emplace_back(&programs[i]->Output.ProgramObject,
				binary_types[i] == STP_RTC_CODE_OBJECT ? STPDeviceRuntimeProgram::STPBinaryType::CodeObject
													   : STPDeviceRuntimeProgram::STPBinaryType::LLVMBitcode,
				STPDeviceRuntimeProgram::STPLinkerInformation::STPDataJitOption {});
		}
		auto m = std::make_unique<stp_rtc_module>();
		m->Module = STPDeviceRuntimeProgram::link(info);
		*out = m.release();
	});
}

int stp_rtc_module_destroy(stp_rtc_module* const module) {
	return guarded([&] { delete module; });
}

int stp_rtc_module_function(stp_rtc_module* const module, const char* const lowered_name, void** const function) {
	return guarded([&] {
		require(module && lowered_name && function, "null argument");
		hipFunction_t f = nullptr;
		STP_CHECK_HIP(hipModuleGetFunction(&f, module->Module.get(), lowered_name));
		*function = f;
	});
}

int stp_rtc_module_global(stp_rtc_module* const module, const char* const lowered_name, void** const d_pointer,
	uint64_t* const bytes) {
	return guarded([&] {
		require(module && lowered_name && d_pointer, "null argument");
		hipDeviceptr_t p = nullptr;
		size_t n = 0u;
		STP_CHECK_HIP(hipModuleGetGlobal(&p, &n, module->Module.get(), lowered_name));
		*d_pointer = p;
		if (bytes) {
			*bytes = n;
		}
	});
}

int stp_rtc_launch(void* const function, const uint32_t grid[3], const uint32_t block[3], const uint32_t shared_bytes,
	void* const stream, void** const args) {
	return guarded([&] {
		require(function && grid && block, "null argument");
		STP_CHECK_HIP(hipModuleLaunchKernel(static_cast<hipFunction_t>(function), grid[0], grid[1], grid[2], block[0],
			block[1], block[2], shared_bytes, toStream(stream), args, nullptr));
	});
}

int stp_batch_create(const stp_simplex_noise_setting* const noise, const stp_biome_property* const h_table,
	const uint32_t biome_count, const stp_raindrop_setting* const raindrop, const stp_batch_setting* const setting,
	const stp_batch_chunk* const h_chunks, const uint32_t chunk_count, stp_heightfield_batch** const out) {
	return guarded([&] {
		require(h_table && setting && h_chunks && out, "null argument");
		std::vector<STPBiomeProperty> t;
		for (uint32_t i = 0u; i < biome_count; i++) {
			t.push_back(toBiome(h_table[i]));
		}
		const STPBatchSetting bs = toBatchSetting(setting);
		std::vector<STPBatchChunk> c(chunk_count);
		for (uint32_t i = 0u; i < chunk_count; i++) {
			c[i] = STPBatchChunk {h_chunks[i].Side, h_chunks[i].Origin[0], h_chunks[i].Origin[1], h_chunks[i].Step};
		}
		auto b = std::make_unique<stp_heightfield_batch>();
		b->Batch = std::make_unique<STPHeightfieldBatch>(toNoise(noise), t, toRain(raindrop), bs, c);
		*out = b.release();
	});
}

int stp_batch_destroy(stp_heightfield_batch* const batch) {
	return guarded([&] { delete batch; });
}

int stp_batch_run(stp_heightfield_batch* const batch, const int mode, void* const stream) {
	return guarded([&] {
		require(batch && mode >= 0 && mode <= 2, "invalid argument");
		batch->Batch->run(static_cast<STPHeightfieldBatch::STPRunMode>(mode), toStream(stream));
	});
}

int stp_batch_set_timing(stp_heightfield_batch* const batch, const int enable) {
	return guarded([&] {
		require(batch, "null batch");
		batch->Batch->setTiming(enable != 0);
	});
}

int stp_batch_stage_times(stp_heightfield_batch* const batch, float* const h_ms, const uint32_t capacity) {
	return guarded([&] {
		require(batch && h_ms, "null argument");
		const std::vector<float> t = batch->Batch->stageTimes();
		require(capacity >= t.size(), "capacity must hold 7 stages");
		std::memcpy(h_ms, t.data(), t.size() * sizeof(float));
	});
}

int stp_batch_chunk_latencies(stp_heightfield_batch* const batch, double* const h_ms, const uint32_t capacity) {
	return guarded([&] {
		require(batch && h_ms, "null argument");
		const std::vector<double>& l = batch->Batch->latency();
		require(!l.empty(), "no plain run yet");
		require(capacity >= l.size(), "capacity must hold one latency per chunk");
		std::memcpy(h_ms, l.data(), l.size() * sizeof(double));
	});
}

int stp_batch_counters(stp_heightfield_batch* const batch, uint64_t h_out[6], void* const stream) {
	return guarded([&] {
		require(batch && h_out, "null argument");
		batch->Batch->check(toStream(stream));
		const auto c = batch->Batch->counters(toStream(stream));
		h_out[0] = batch->Batch->pixels();
		h_out[1] = batch->Batch->drops();
		for (int k = 0; k < 4; k++) {
			h_out[2 + k] = c[k];
		}
	});
}

int stp_batch_status(stp_heightfield_batch* const batch, uint32_t* const h_status, void* const stream) {
	return guarded([&] {
		require(batch && h_status, "null argument");
		*h_status = batch->Batch->status(toStream(stream));
	});
}

int stp_batch_bin_total(stp_heightfield_batch* const batch, uint64_t* const h_out, void* const stream) {
	return guarded([&] {
		require(batch && h_out, "null argument");
		batch->Batch->check(toStream(stream));
		*h_out = batch->Batch->binTotal(toStream(stream));
	});
}

int stp_batch_set_tuning(stp_heightfield_batch* const batch, const stp_batch_tuning* const tuning) {
	return guarded([&] {
		require(batch && tuning, "null argument");
		batch->Batch->setTuning(STPBatchTuning {tuning->ErosionTile, tuning->ErosionThreads});
	});
}

int stp_batch_get_tuning(stp_heightfield_batch* const batch, stp_batch_tuning* const out) {
	return guarded([&] {
		require(batch && out, "null argument");
		const STPBatchTuning& t = batch->Batch->tuning();
		out->ErosionTile = t.ErosionTile;
		out->ErosionThreads = t.ErosionThreads;
	});
}

int stp_batch_autotune(stp_heightfield_batch* const batch, const uint32_t repeat, void* const stream,
	stp_batch_tuning* const out) {
	return guarded([&] {
		require(batch && repeat > 0u, "invalid argument");
		const STPBatchTuning t = batch->Batch->autotune(toStream(stream), repeat);
		if (out) {
			out->ErosionTile = t.ErosionTile;
			out->ErosionThreads = t.ErosionThreads;
		}
	});
}

int stp_batch_read_chunk(stp_heightfield_batch* const batch, const uint32_t index, float* const h_height,
	uint16_t* const h_fixed, uint16_t* const h_biome_halo, float* const h_ledger) {
	return guarded([&] {
		require(batch, "null batch");
		batch->Batch->readChunk(index, h_height, h_fixed, h_biome_halo, h_ledger);
	});
}

int stp_server_create(const stp_simplex_noise_setting* const noise, const stp_biome_property* const h_table,
	const uint32_t biome_count, const stp_raindrop_setting* const raindrop, const stp_batch_setting* const batch,
	const stp_server_setting* const setting, stp_heightfield_server** const out) {
	return guarded([&] {
		require(h_table && batch && setting && out, "null argument");
		std::vector<STPBiomeProperty> t;
		for (uint32_t i = 0u; i < biome_count; i++) {
			t.push_back(toBiome(h_table[i]));
		}
		const STPBatchSetting bs = toBatchSetting(batch);
		const STPServerSetting ss {setting->MemoryBudget, setting->BatchPixels, setting->BatchChunks, setting->MaxSide,
			setting->Engines, setting->ChunkWorldSize};
		auto s = std::make_unique<stp_heightfield_server>();
		s->Server = std::make_unique<STPHeightfieldServer>(toNoise(noise), t, toRain(raindrop), bs, ss);
		*out = s.release();
	});
}

int stp_server_destroy(stp_heightfield_server* const server) {
	return guarded([&] { delete server; });
}

int stp_server_submit(stp_heightfield_server* const server, const stp_chunk_request* const request,
	uint64_t* const ticket) {
	return guarded([&] {
		require(server && request, "null argument");
		const unsigned long long t = server->Server->submit(STPChunkRequest {request->X, request->Y, request->Side});
		if (ticket) {
			*ticket = t;
		}
	});
}

int stp_server_poll(stp_heightfield_server* const server, uint64_t* const h_ticket, double* const h_latency_ms,
	const uint32_t capacity, uint32_t* const count) {
	return guarded([&] {
		require(server && count && (capacity == 0u || (h_ticket && h_latency_ms)), "null argument");
		server->Server->poll();
		for (const auto& d : server->Server->drain()) {
			server->Done.push_back(d);
		}
		const uint32_t n = static_cast<uint32_t>(std::min<size_t>(capacity, server->Done.size()));
		for (uint32_t i = 0u; i < n; i++) {
			h_ticket[i] = server->Done[i].Ticket;
			h_latency_ms[i] = server->Done[i].Ms;
		}
		server->Done.erase(server->Done.begin(), server->Done.begin() + n);
		*count = n;
	});
}

int stp_server_flush(stp_heightfield_server* const server) {
	return guarded([&] {
		require(server, "null server");
		server->Server->flush();
		for (const auto& d : server->Server->drain()) {
			server->Done.push_back(d);
		}
	});
}

int stp_server_read(stp_heightfield_server* const server, const stp_chunk_request* const request,
	float* const h_height, uint16_t* const h_fixed) {
	return guarded([&] {
		require(server && request, "null argument");
		require(server->Server->read(STPChunkRequest {request->X, request->Y, request->Side}, h_height, h_fixed),
			"the chunk is not resident");
	});
}

int stp_server_memory(stp_heightfield_server* const server, uint64_t* const engine_bytes,
	uint64_t* const resident_bytes) {
	return guarded([&] {
		require(server, "null server");
		if (engine_bytes) {
			*engine_bytes = server->Server->engineBytes();
		}
		if (resident_bytes) {
			*resident_bytes = server->Server->residentBytes();
		}
	});
}

int stp_server_replay(stp_heightfield_server* const server, const stp_chunk_request* const h_requests,
	const double* const h_arrival_s, const uint32_t count, const double deadline_ms, double* const h_latency_ms,
	stp_server_stats* const out) {
	return guarded([&] {
		require(server && (count == 0u || (h_requests && h_arrival_s && h_latency_ms)), "null argument");
		std::vector<STPChunkRequest> req(count);
		for (uint32_t i = 0u; i < count; i++) {
			req[i] = STPChunkRequest {h_requests[i].X, h_requests[i].Y, h_requests[i].Side};
		}
		std::vector<double> arrival(h_arrival_s, h_arrival_s + count), latency;
		const STPServerStats st = server->Server->replay(req, arrival, deadline_ms, latency);
		std::copy(latency.begin(), latency.end(), h_latency_ms);
		if (out) {
			out->P50 = st.P50;
			out->P99 = st.P99;
			out->P999 = st.P999;
			out->Mean = st.Mean;
			out->Max = st.Max;
			out->MpixelPerSecond = st.MpixelPerSecond;
			out->DeadlineMissFraction = st.DeadlineMissFraction;
			out->Requests = st.Requests;
			out->CacheHits = st.CacheHits;
			out->Batches = st.Batches;
			out->Evictions = st.Evictions;
			out->MeanBatchChunks = st.MeanBatchChunks;
			out->PeakResidentBytes = st.PeakResidentBytes;
			out->Generated = st.Generated;
			out->GeneratedP50 = st.GeneratedP50;
			out->GeneratedP99 = st.GeneratedP99;
			out->GeneratedP999 = st.GeneratedP999;
			out->GeneratedMax = st.GeneratedMax;
			out->GeneratedMpixelPerSecond = st.GeneratedMpixelPerSecond;
		}
	});
}

int stp_splat_database_create(const stp_altitude_rule* const h_altitude, const uint32_t altitude_count,
	const stp_gradient_rule* const h_gradient, const uint32_t gradient_count, stp_splat_database** const out) {
	return guarded([&] {
		require(out && (altitude_count == 0u || h_altitude) && (gradient_count == 0u || h_gradient), "null argument");
		std::vector<STPAltitudeRule> alt(altitude_count);
		for (uint32_t i = 0u; i < altitude_count; i++) {
			alt[i] = STPAltitudeRule {h_altitude[i].Sample, h_altitude[i].UpperBound, h_altitude[i].Region};
		}
		std::vector<STPGradientRule> gra(gradient_count);
		for (uint32_t i = 0u; i < gradient_count; i++) {
			const stp_gradient_rule& g = h_gradient[i];
			gra[i] = STPGradientRule {g.Sample, g.minGradient, g.maxGradient, g.LowerBound, g.UpperBound, g.Region};
		}
		auto d = std::make_unique<stp_splat_database>();
		d->Table = std::make_unique<STPSplatRuleTable>(alt, gra);
		*out = d.release();
	});
}

int stp_splat_database_destroy(stp_splat_database* const database) {
	return guarded([&] { delete database; });
}

int stp_splatmap_generate(const stp_permutation* const perm, const stp_splat_database* const database,
	const uint16_t* const d_biome, const uint16_t* const d_height, uint8_t* const d_splat, const uint32_t area_x,
	const uint32_t area_y, const uint32_t map_x, const uint32_t map_y, const stp_splat_chunk* const h_chunks,
	const uint32_t count, void* const stream) {
	return guarded([&] {
		require(perm && database && (count == 0u || h_chunks), "null argument");
		std::vector<STPSplatChunk> c(count);
		for (uint32_t i = 0u; i < count; i++) {
			c[i] = STPSplatChunk {h_chunks[i].LocalX, h_chunks[i].LocalY, h_chunks[i].OffsetX, h_chunks[i].OffsetY};
		}
		STPHeightfieldKernel::splatmap(toPerm(perm), database->Table->view(), d_biome, d_height, d_splat,
			make_uint2(area_x, area_y), make_uint2(map_x, map_y),
			make_float2(static_cast<float>(map_x) / 2.0f, static_cast<float>(map_y) / 2.0f), c, toStream(stream));
	});
}

int stp_world_create(const stp_simplex_noise_setting* const noise, const stp_biome_property* const h_table,
	const uint32_t biome_count, const stp_raindrop_setting* const raindrop, const stp_batch_setting* const batch,
	const stp_world_setting* const setting, stp_heightfield_world** const out) {
	return guarded([&] {
		require(h_table && batch && setting && out, "null argument");
		std::vector<STPBiomeProperty> t;
		for (uint32_t i = 0u; i < biome_count; i++) {
			t.push_back(toBiome(h_table[i]));
		}
		const STPBatchSetting bs = toBatchSetting(batch);
		const STPWorldSetting ws {setting->MapSize, setting->ErosionNeighbour, setting->GenerateChunks, setting->ErodeChunks};
		auto w = std::make_unique<stp_heightfield_world>();
		w->World = std::make_unique<STPHeightfieldWorld>(toNoise(noise), t, toRain(raindrop), bs, ws);
		*out = w.release();
	});
}

int stp_world_destroy(stp_heightfield_world* const world) {
	return guarded([&] { delete world; });
}

int stp_world_request(stp_heightfield_world* const world, const int32_t* const h_chunks, const uint32_t count,
	int32_t* const h_changed, const uint32_t capacity, uint32_t* const changed_count) {
	return guarded([&] {
		require(world && (count == 0u || h_chunks) && (capacity == 0u || h_changed), "null argument");
		std::vector<int2> req(count);
		for (uint32_t i = 0u; i < count; i++) {
			req[i] = make_int2(h_chunks[2u * i], h_chunks[2u * i + 1u]);
		}
		const std::vector<int2> changed = world->World->request(req);
		const uint32_t n = static_cast<uint32_t>(std::min<size_t>(capacity, changed.size()));
		for (uint32_t i = 0u; i < n; i++) {
			h_changed[2u * i] = changed[i].x;
			h_changed[2u * i + 1u] = changed[i].y;
		}
		if (changed_count) {
			*changed_count = static_cast<uint32_t>(changed.size());
		}
	});
}

int stp_world_read(stp_heightfield_world* const world, const int32_t x, const int32_t y, float* const h_height,
	uint16_t* const h_fixed) {
	return guarded([&] {
		require(world, "null world");
		require(world->World->read(make_int2(x, y), h_height, h_fixed), "the chunk has no map");
	});
}

int stp_world_state(stp_heightfield_world* const world, const int32_t x, const int32_t y, uint32_t* const state) {
	return guarded([&] {
		require(world && state, "null argument");
		*state = static_cast<uint32_t>(world->World->state(make_int2(x, y)));
	});
}

int stp_world_render_area(stp_heightfield_world* const world, const int32_t x, const int32_t y, const uint32_t radius,
	uint16_t* const d_biome, uint16_t* const d_height) {
	return guarded([&] {
		require(world, "null world");
		world->World->renderArea(make_int2(x, y), radius, d_biome, d_height);
	});
}

int stp_world_load(stp_heightfield_world* const world, const int32_t x, const int32_t y, const uint32_t radius,
	uint32_t* const status) {
	return guarded([&] {
		require(world && status, "null argument");
		*status = static_cast<uint32_t>(world->World->load(make_int2(x, y), radius));
	});
}

int stp_world_front(stp_heightfield_world* const world, const uint16_t** const d_biome, const uint16_t** const d_height,
	uint32_t* const side) {
	return guarded([&] {
		require(world, "null world");
		world->World->front(d_biome, d_height, side);
	});
}

int stp_world_schedule(const int32_t* const h_centres, const uint32_t count, const uint32_t neighbour,
	uint32_t* const h_batch) {
	return guarded([&] {
		require(count == 0u || (h_centres && h_batch), "null argument");
		std::vector<int2> c(count);
		for (uint32_t i = 0u; i < count; i++) {
			c[i] = make_int2(h_centres[2u * i], h_centres[2u * i + 1u]);
		}
		const std::vector<unsigned int> b = STPHeightfieldWorld::schedule(c, neighbour);
		std::copy(b.begin(), b.end(), h_batch);
	});
}

int stp_world_stream_word(const int32_t x, const int32_t y, uint32_t* const word) {
	return guarded([&] {
		require(word, "null argument");
		*word = STPHeightfieldWorld::streamWord(make_int2(x, y));
	});
}

int stp_server_stream_word(const int32_t x, const int32_t y, const uint32_t side, uint32_t* const word) {
	return guarded([&] {
		require(word, "null argument");
		*word = STPHeightfieldServer::streamWord(x, y, side);
	});
}

} // extern "C"

/* ---------------------------------------------------------------------------- biome layers */

int stp_layer_program_demo(stp_layer_node* const h_nodes, const uint32_t capacity, uint32_t* const count) {
	return guarded([&] {
		require(count, "null count");
		const STPDiversity::STPLayerProgram p = STPDiversity::STPLayerProgram::demo();
		*count = static_cast<uint32_t>(p.Node.size());
		require(h_nodes && capacity >= p.Node.size(), "node capacity too small");
		for (size_t i = 0u; i < p.Node.size(); i++) {
			const STPDiversity::STPLayerNode& n = p.Node[i];
			h_nodes[i] = stp_layer_node {n.Kind, n.Param, {n.Parent[0], n.Parent[1]}, n.Salt, n.CacheSize};
		}
	});
}

int stp_biome_registry_demo(stp_biome_registry* const out) {
	return guarded([&] {
		require(out, "null registry");
		const STPDiversity::STPBiomeRegistry r = STPDiversity::STPBiomeRegistry::demo();
		static_assert(sizeof(out->ID) == sizeof(r.ID), "registry size");
		std::memcpy(out->ID, r.ID, sizeof r.ID);
		std::memcpy(out->Temperature, r.Temperature, sizeof r.Temperature);
		std::memcpy(out->Precipitation, r.Precipitation, sizeof r.Precipitation);
	});
}

int stp_layer_factory_create(const stp_layer_node* const h_nodes, const uint32_t count, const uint64_t global_seed,
	const stp_biome_registry* const registry, const uint32_t dimension_x, const uint32_t dimension_y,
	stp_layer_factory** const out) {
	return stp_layer_factory_create_custom(h_nodes, count, global_seed, registry, dimension_x, dimension_y, nullptr, 0u, out);
}

int stp_layer_factory_create_custom(const stp_layer_node* const h_nodes, const uint32_t count, const uint64_t global_seed,
	const stp_biome_registry* const registry, const uint32_t dimension_x, const uint32_t dimension_y,
	const stp_custom_layer* const h_custom, const uint32_t custom_count, stp_layer_factory** const out) {
	return guarded([&] {
		require(custom_count == 0u || h_custom, "null custom layers");
		std::vector<STPDiversity::STPCustomLayer> custom(custom_count);
		for (uint32_t i = 0u; i < custom_count; i++) {
			const stp_custom_layer& c = h_custom[i];
			require(c.Source && (c.OptionCount == 0u || c.Options), "a custom layer needs its source");
			custom[i].Source = c.Source;
			custom[i].Lo = c.Lo;
			custom[i].Hi = c.Hi;
			custom[i].Shift = c.Shift;
			custom[i].Parents = c.Parents;
			for (uint32_t k = 0u; k < c.OptionCount; k++) {
				require(c.Options[k], "null option");
				custom[i].Options.emplace_back(c.Options[k]);
			}
		}
		require(out && h_nodes && registry, "null argument");
		STPDiversity::STPLayerProgram p;
		for (uint32_t i = 0u; i < count; i++) {
			const stp_layer_node& n = h_nodes[i];
			p.Node.push_back(STPDiversity::STPLayerNode {n.Kind, n.Salt, {n.Parent[0], n.Parent[1]}, n.Param, n.CacheSize});
		}
		STPDiversity::STPBiomeRegistry r {};
		std::memcpy(r.ID, registry->ID, sizeof r.ID);
		std::memcpy(r.Temperature, registry->Temperature, sizeof r.Temperature);
		std::memcpy(r.Precipitation, registry->Precipitation, sizeof r.Precipitation);
		auto f = std::make_unique<stp_layer_factory>();
		f->Factory = std::make_unique<STPDiversity::STPLayerBiomeFactory>(p, global_seed, r, make_uint2(dimension_x, dimension_y),
			custom);
		*out = f.release();
	});
}

int stp_layer_factory_destroy(stp_layer_factory* const factory) {
	return guarded([&] { delete factory; });
}

int stp_layer_factory_generate(stp_layer_factory* const factory, uint16_t* const h_biomemap, const int32_t offset_x,
	const int32_t offset_y) {
	return guarded([&] {
		require(factory && h_biomemap, "null argument");
		(*factory->Factory)(h_biomemap, make_int2(offset_x, offset_y));
	});
}

int stp_layer_factory_generate_device(stp_layer_factory* const factory, uint16_t* const d_biomemap,
	const int32_t* const h_offsets, const uint32_t count, void* const stream) {
	return guarded([&] {
		require(factory && (count == 0u || (d_biomemap && h_offsets)), "null argument");
		std::vector<int2> o(count);
		for (uint32_t i = 0u; i < count; i++) {
			o[i] = make_int2(h_offsets[2u * i], h_offsets[2u * i + 1u]);
		}
		factory->Factory->generateDevice(d_biomemap, o.data(), count, static_cast<hipStream_t>(stream));
	});
}

int stp_layer_factory_status(stp_layer_factory* const factory, uint32_t* const status) {
	return guarded([&] {
		require(factory && status, "null argument");
		*status = factory->Factory->status();
	});
}

/* ---------------------------------------------------------------------------- chunk records */

int stp_chunk_record_bytes(const uint32_t width, const uint32_t height, const uint32_t maps, uint64_t* const bytes) {
	return guarded([&] {
		require(bytes && width > 0u && height > 0u && width <= STPChunkRecord::MaxSide
			&& height <= STPChunkRecord::MaxSide && maps != 0u && (maps & ~STPChunkRecord::MapAll) == 0u,
			"size in [1, 65536] per side, maps a non-empty combination of 1, 2, 4");
		*bytes = STPChunkRecord::Layout(width, height, maps).record();
	});
}

int stp_world_export(stp_heightfield_world* const world, const int32_t* const h_chunks, const uint32_t count,
	const uint32_t maps, void* const h_out, const uint64_t capacity, uint64_t* const written) {
	return guarded([&] {
		require(world && written && (count == 0u || h_chunks), "null argument");
		std::vector<int2> c(count);
		for (uint32_t i = 0u; i < count; i++) {
			c[i] = make_int2(h_chunks[2u * i], h_chunks[2u * i + 1u]);
		}
		*written = world->World->exportRecords(c, maps, static_cast<unsigned char*>(h_out), capacity);
	});
}

int stp_chunk_record_read(const void* const data, const uint64_t size, stp_chunk_record* const record) {
	return guarded([&] {
		require(record, "null record");
		const STPChunkRecord::View v = STPChunkRecord::read(static_cast<const unsigned char*>(data), size);
		*record = stp_chunk_record {v.X, v.Y, v.OriginX, v.OriginY, v.Width, v.Height, v.Completeness, v.Maps, v.Check,
			v.Biome, v.HeightMap, v.Fixed, v.Bytes};
	});
}

/* ---------------------------------------------------------------------------- runtime splatmap */

int stp_splatmap_runtime_create(const stp_permutation* const perm, const stp_splat_database* const database,
	const uint32_t map_x, const uint32_t map_y, const char* const script, const char* const* const options,
	const uint32_t option_count, stp_splatmap_runtime** const out) {
	return guarded([&] {
		require(perm && database && out && (options || option_count == 0u), "null argument");
		std::vector<std::string> opt;
		for (uint32_t i = 0u; i < option_count; i++) {
			require(options[i], "null option");
			opt.emplace_back(options[i]);
		}
		const STPAlgorithm::STPPermutation p {perm->Permutation, perm->Gradient2D, perm->Gradient2DSize};
		auto g = std::make_unique<stp_splatmap_runtime>();
		g->Generator = std::make_unique<STPRuntimeSplatmapGenerator>(p, database->Table->view(), make_uint2(map_x, map_y),
			script ? std::string(script) : std::string(), opt);
		*out = g.release();
	});
}

int stp_splatmap_runtime_destroy(stp_splatmap_runtime* const generator) {
	return guarded([&] { delete generator; });
}

int stp_splatmap_runtime_generate(stp_splatmap_runtime* const generator, const uint16_t* const d_biome,
	const uint16_t* const d_height, uint8_t* const d_splat, const uint32_t area_x, const uint32_t area_y,
	const stp_splat_chunk* const h_chunks, const uint32_t count, void* const stream) {
	return guarded([&] {
		require(generator && (count == 0u || h_chunks), "null argument");
		std::vector<STPRuntimeSplatmapGenerator::STPChunkInformation> c(count);
		for (uint32_t i = 0u; i < count; i++) {
			c[i] = STPRuntimeSplatmapGenerator::STPChunkInformation {h_chunks[i].LocalX, h_chunks[i].LocalY,
				h_chunks[i].OffsetX, h_chunks[i].OffsetY};
		}
		generator->Generator->splat(d_biome, d_height, d_splat, make_uint2(area_x, area_y), c,
			static_cast<hipStream_t>(stream));
	});
}
