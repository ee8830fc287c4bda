// Free-slip neighbourhood merge / unmerge on the device.
//
// Index mapping (SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPNearestNeighbourTextureBuffer.cpp:97-110
// with STPChunk::calcLocalChunkCoordinate, SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPChunk.cpp:71-73):
// chunk i sits at local (i % NN.x, i / NN.x); its pixel (px, py) lands at merged
// (lx * Map.x + px) + (ly * Map.y + py) * Total.x. The reference packs on the host and copies the
// merged map over PCIe; here the chunks stay in HBM and one gather (or scatter) kernel does it.
// One workgroup row covers 256 consecutive merged pixels, so reads and writes are coalesced
// except at chunk seams. Roofline: HBM, 2 x element size bytes per pixel.
#include "stp_internal.hpp"
#include "stp_stage_erosion.hpp"

using namespace SuperTerrainPlus;

namespace {

	struct ChunkTable {
		const void* Ptr[STPHeightfieldKernel::MaxNeighbour];
	};

	template <typename T, bool Merge>
	__global__ void __launch_bounds__(256) neighbourKERNEL(const ChunkTable chunks, T* const merged_out,
		const T* const merged_in, const STPNearestNeighbourInformation nn) {
		const unsigned int gx = blockIdx.x * blockDim.x + threadIdx.x, gy = blockIdx.y;
		if (gx >= nn.TotalMapSize.x) {
			return;
		}
		const unsigned int lx = gx / nn.MapSize.x, px = gx - lx * nn.MapSize.x;
		const unsigned int ly = gy / nn.MapSize.y, py = gy - ly * nn.MapSize.y;
		const unsigned int chunk = lx + ly * nn.ChunkNearestNeighbour.x;
		const size_t g = gx + static_cast<size_t>(gy) * nn.TotalMapSize.x;
		const size_t l = px + static_cast<size_t>(py) * nn.MapSize.x;
		T* const c = const_cast<T*>(static_cast<const T*>(chunks.Ptr[chunk]));
		if constexpr (Merge) {
			merged_out[g] = c[l];
		} else {
			c[l] = merged_in[g];
		}
	}

	// Unmerge of an FP32 merged map fused with the UINT16 format of each pixel: the same bits as
	// formatting the merged map and unmerging both, in one pass (4 B read, 6 B written per pixel).
	__global__ void __launch_bounds__(256) unmergeFormatKERNEL(const ChunkTable height, const ChunkTable fixed,
		const float* const merged, const STPNearestNeighbourInformation nn) {
		const unsigned int gx = blockIdx.x * blockDim.x + threadIdx.x, gy = blockIdx.y;
		if (gx >= nn.TotalMapSize.x) {
			return;
		}
		const unsigned int lx = gx / nn.MapSize.x, px = gx - lx * nn.MapSize.x;
		const unsigned int ly = gy / nn.MapSize.y, py = gy - ly * nn.MapSize.y;
		const unsigned int chunk = lx + ly * nn.ChunkNearestNeighbour.x;
		const size_t l = px + static_cast<size_t>(py) * nn.MapSize.x;
		const float h = merged[gx + static_cast<size_t>(gy) * nn.TotalMapSize.x];
		static_cast<float*>(const_cast<void*>(height.Ptr[chunk]))[l] = h;
		static_cast<STPHeightFixed_t*>(const_cast<void*>(fixed.Ptr[chunk]))[l] = STPStage::toFixed(h);
	}

	template <bool Merge>
	void launch(const ChunkTable& t, void* const out, const void* const in, const unsigned int element_size,
		const STPNearestNeighbourInformation& nn, const hipStream_t stream) {
		const dim3 grid((nn.TotalMapSize.x + 255u) / 256u, nn.TotalMapSize.y);
		switch (element_size) {
		case 1u:
			neighbourKERNEL<unsigned char, Merge><<<grid, 256, 0, stream>>>(t, static_cast<unsigned char*>(out),
				static_cast<const unsigned char*>(in), nn);
			break;
		case 2u:
			neighbourKERNEL<unsigned short, Merge><<<grid, 256, 0, stream>>>(t, static_cast<unsigned short*>(out),
				static_cast<const unsigned short*>(in), nn);
			break;
		case 4u:
			neighbourKERNEL<unsigned int, Merge><<<grid, 256, 0, stream>>>(t, static_cast<unsigned int*>(out),
				static_cast<const unsigned int*>(in), nn);
			break;
		default:
			STP_THROW(STPInvalidArgument, "neighbour merge supports 1, 2 and 4 byte pixels");
		}
		STP_CHECK_HIP(hipGetLastError());
	}

	ChunkTable table(const void* const* chunks, const STPNearestNeighbourInformation& nn) {
		const unsigned int count = nn.ChunkNearestNeighbour.x * nn.ChunkNearestNeighbour.y;
		if (count == 0u || count > STPHeightfieldKernel::MaxNeighbour) {
			STP_THROW(STPInvalidArgument, "neighbour count must be in [1, 64]");
		}
		if (nn.TotalMapSize.x != nn.MapSize.x * nn.ChunkNearestNeighbour.x
			|| nn.TotalMapSize.y != nn.MapSize.y * nn.ChunkNearestNeighbour.y) {
			STP_THROW(STPInvalidArgument, "inconsistent nearest-neighbour information");
		}
		ChunkTable t = {};
		for (unsigned int i = 0u; i < count; i++) {
			if (!chunks[i]) {
				STP_THROW(STPInvalidArgument, "null neighbour chunk pointer");
			}
			t.Ptr[i] = chunks[i];
		}
		return t;
	}

} // namespace

void STPHeightfieldKernel::mergeNeighbour(const void* const* const chunks, void* const merged,
	const unsigned int element_size, const STPNearestNeighbourInformation& nn, const hipStream_t stream) {
	if (!merged) {
		STP_THROW(STPInvalidArgument, "null merged buffer");
	}
	launch<true>(table(chunks, nn), merged, nullptr, element_size, nn, stream);
}

void STPHeightfieldKernel::unmergeNeighbour(const void* const merged, void* const* const chunks,
	const unsigned int element_size, const STPNearestNeighbourInformation& nn, const hipStream_t stream) {
	if (!merged) {
		STP_THROW(STPInvalidArgument, "null merged buffer");
	}
	launch<false>(table(const_cast<const void* const*>(chunks), nn), nullptr, merged, element_size, nn, stream);
}

void STPHeightfieldKernel::unmergeFormatNeighbour(const float* const merged, float* const* const height,
	STPHeightFixed_t* const* const fixed, const STPNearestNeighbourInformation& nn, const hipStream_t stream) {
	if (!merged || !height || !fixed) {
		STP_THROW(STPInvalidArgument, "null merged buffer or chunk pointer array");
	}
	const ChunkTable h = table(reinterpret_cast<const void* const*>(height), nn),
		f = table(reinterpret_cast<const void* const*>(fixed), nn);
	const dim3 grid((nn.TotalMapSize.x + 255u) / 256u, nn.TotalMapSize.y);
	unmergeFormatKERNEL<<<grid, 256, 0, stream>>>(h, f, merged, nn);
	STP_CHECK_HIP(hipGetLastError());
}
