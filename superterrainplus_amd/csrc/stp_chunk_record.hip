// Chunk record packing and checks on the device, header writing and parsing on the host
// (stp_chunk_record.hpp).
//
// Roofline: both kernels stream bytes through HBM once: the pack reads 6 bytes and writes 6 bytes
// per pixel (FP32 and UINT16 heights), the check reads the payload once (8 bytes per pixel with the
// biome map). HBM bound.
#include "stp_chunk_record.hpp"

#include <cstring>

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPChunkRecord;

namespace {

	constexpr unsigned int PackThreads = 256u, PackBlocks = 64u, CheckThreads = 256u, CheckBlocks = 64u;

	// blockIdx.y = record; the record's pixels are split over PackBlocks blocks
	__global__ void __launch_bounds__(PackThreads) packKERNEL(const PackEntry* const entry, const unsigned long long pixels,
		const unsigned int maps, unsigned char* const out) {
		const PackEntry e = entry[blockIdx.y];
		const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * PackThreads;
		const unsigned long long first = static_cast<unsigned long long>(blockIdx.x) * PackThreads + threadIdx.x;
		if (maps & MapHeight) {
			float* const dst = reinterpret_cast<float*>(out + e.Payload + e.HeightAt);
			for (unsigned long long i = first; i < pixels; i += stride) {
				dst[i] = e.Height[i];
			}
		}
		if (maps & MapFixed) {
			unsigned short* const dst = reinterpret_cast<unsigned short*>(out + e.Payload + e.FixedAt);
			for (unsigned long long i = first; i < pixels; i += stride) {
				dst[i] = e.Fixed[i];
			}
		}
	}

	// blockIdx.y = record; sum of w_i (2 i + 1) over its payload words, modulo 2^32
	__global__ void __launch_bounds__(CheckThreads) checkKERNEL(const PackEntry* const entry, const unsigned char* const out,
		unsigned int* const check) {
		__shared__ unsigned int partial[CheckThreads / 64u];
		const PackEntry e = entry[blockIdx.y];
		const unsigned int* const word = reinterpret_cast<const unsigned int*>(out + e.Payload);
		const unsigned long long words = e.PayloadBytes / 4u, stride = static_cast<unsigned long long>(gridDim.x) * CheckThreads;
		unsigned int sum = 0u;
		for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * CheckThreads + threadIdx.x; i < words;
			i += stride) {
			sum += word[i] * (2u * static_cast<unsigned int>(i) + 1u);
		}
		for (unsigned int o = 32u; o > 0u; o >>= 1u) {
			sum += __shfl_xor(sum, o, 64);
		}
		if ((threadIdx.x & 63u) == 0u) {
			partial[threadIdx.x / 64u] = sum;
		}
		__syncthreads();
		if (threadIdx.x == 0u) {
			unsigned int total = 0u;
			for (unsigned int w = 0u; w < CheckThreads / 64u; w++) {
				total += partial[w];
			}
			atomicAdd(check + blockIdx.y, total);
		}
	}

	template <class T>
	void put(unsigned char* const at, const T v) noexcept {
		std::memcpy(at, &v, sizeof v);
	}
	template <class T>
	T get(const unsigned char* const at) noexcept {
		T v;
		std::memcpy(&v, at, sizeof v);
		return v;
	}

} // namespace

void STPChunkRecord::pack(const PackEntry* const entry, const unsigned int count, const unsigned long long pixels,
	const unsigned int maps, unsigned char* const out, const hipStream_t stream) {
	if (count == 0u || !(maps & (MapHeight | MapFixed))) {
		return;
	}
	packKERNEL<<<dim3(PackBlocks, count), PackThreads, 0, stream>>>(entry, pixels, maps, out);
	STP_CHECK_HIP(hipGetLastError());
}

void STPChunkRecord::check(const PackEntry* const entry, const unsigned int count, const unsigned long long payload,
	const unsigned char* const out, unsigned int* const check, const hipStream_t stream) {
	if (count == 0u || payload == 0u) {
		return;
	}
	checkKERNEL<<<dim3(CheckBlocks, count), CheckThreads, 0, stream>>>(entry, out, check);
	STP_CHECK_HIP(hipGetLastError());
}

void STPChunkRecord::writeHeader(unsigned char* const at, const int x, const int y, const int origin_x, const int origin_y,
	const unsigned int width, const unsigned int height, const unsigned int completeness, const unsigned int maps,
	const unsigned long long payload, const unsigned int check) noexcept {
	std::memset(at, 0, HeaderBytes);
	std::memcpy(at, "STPC", 4u);
	put<unsigned short>(at + 4u, static_cast<unsigned short>(Version));
	put<unsigned short>(at + 6u, static_cast<unsigned short>(HeaderBytes));
	put<int>(at + 8u, x);
	put<int>(at + 12u, y);
	put<int>(at + 16u, origin_x);
	put<int>(at + 20u, origin_y);
	put<unsigned int>(at + 24u, width);
	put<unsigned int>(at + 28u, height);
	put<unsigned int>(at + 32u, completeness);
	put<unsigned int>(at + 36u, maps);
	put<unsigned long long>(at + 40u, payload);
	put<unsigned int>(at + 48u, check);
}

STPChunkRecord::View STPChunkRecord::read(const unsigned char* const data, const unsigned long long size) {
	if (!data || size < HeaderBytes) {
		STP_THROW(STPInvalidArgument, "a chunk record needs at least its 64-byte header");
	}
	if (std::memcmp(data, "STPC", 4u) != 0) {
		STP_THROW(STPInvalidArgument, "not a chunk record (magic)");
	}
	if (get<unsigned short>(data + 4u) != Version || get<unsigned short>(data + 6u) != HeaderBytes) {
		STP_THROW(STPInvalidArgument, "unsupported chunk record version or header size");
	}
	View v {};
	v.X = get<int>(data + 8u);
	v.Y = get<int>(data + 12u);
	v.OriginX = get<int>(data + 16u);
	v.OriginY = get<int>(data + 20u);
	v.Width = get<unsigned int>(data + 24u);
	v.Height = get<unsigned int>(data + 28u);
	v.Completeness = get<unsigned int>(data + 32u);
	v.Maps = get<unsigned int>(data + 36u);
	const unsigned long long payload = get<unsigned long long>(data + 40u);
	v.Check = get<unsigned int>(data + 48u);
	if (v.Maps == 0u || (v.Maps & ~MapAll) != 0u || v.Completeness > 3u || v.Width == 0u || v.Height == 0u
		|| v.Width > MaxSide || v.Height > MaxSide) {
		STP_THROW(STPInvalidArgument, "chunk record header: maps, completeness or size out of range");
	}
	const Layout l(v.Width, v.Height, v.Maps);
	if (payload != l.Payload || size - HeaderBytes < payload) {
		STP_THROW(STPInvalidArgument, "chunk record payload size does not match its maps, or the record is truncated");
	}
	const unsigned char* const p = data + HeaderBytes;
	unsigned int sum = 0u;
	for (unsigned long long i = 0u; i < payload / 4u; i++) {
		sum += get<unsigned int>(p + 4u * i) * (2u * static_cast<unsigned int>(i) + 1u);
	}
	if (sum != v.Check) {
		STP_THROW(STPInvalidArgument, "chunk record check mismatch");
	}
	v.Biome = v.Maps & MapBiome ? reinterpret_cast<const unsigned short*>(p + l.Biome) : nullptr;
	v.HeightMap = v.Maps & MapHeight ? reinterpret_cast<const float*>(p + l.Height) : nullptr;
	v.Fixed = v.Maps & MapFixed ? reinterpret_cast<const unsigned short*>(p + l.Fixed) : nullptr;
	v.Bytes = l.record();
	return v;
}
