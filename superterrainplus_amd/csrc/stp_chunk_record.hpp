// Chunk record: the byte format in which chunk maps leave the device for external consumers (files,
// sockets) without graphics interop. It carries the reference's map pixel formats
// (SuperTerrain+/SuperTerrain+/Public/SuperTerrain+/World/STPWorldMapPixelFormat.hpp:12-21: biome STPSample_t
// = u16, height STPHeightFloat_t = f32 and STPHeightFixed_t = u16, row-major per chunk as STPChunk holds
// them, SuperTerrain+/SuperTerrain+/Private/World/Chunk/STPChunk.cpp:17-25) and the chunk's completeness
// (STPChunk::STPChunkCompleteness numbering). The reference keeps chunks in memory only; the layout
// below is this library's own.
//
// A record, little-endian, 16-byte aligned:
//   header, 64 bytes:
//     0 "STPC"   4 u16 version (1)   6 u16 header bytes (64)   8 i32 chunk x   12 i32 chunk y
//    16 i32 world pixel x of the first pixel   20 i32 world pixel y   24 u32 width   28 u32 height
//    32 u32 completeness (0 empty, 1 biome map, 2 heightmap, 3 complete)   36 u32 maps (bit 0 biome,
//       bit 1 FP32 height, bit 2 UINT16 height)   40 u64 payload bytes   48 u32 check   52..63 zero
//   payload: the present maps in bit order, each row-major and zero-padded to a multiple of 16 bytes.
// check = sum over the payload's 32-bit words w_i of w_i * (2 i + 1), modulo 2^32.
#pragma once
#ifndef STP_AMD_CHUNK_RECORD_HPP
#define STP_AMD_CHUNK_RECORD_HPP

#include "stp_internal.hpp"

#include <vector>

namespace SuperTerrainPlus::STPChunkRecord {

	constexpr unsigned int HeaderBytes = 64u, Version = 1u;
	constexpr unsigned int MapBiome = 1u, MapHeight = 2u, MapFixed = 4u, MapAll = 7u;
	// Largest width and height of a record: with both at most 2^16 every map and payload size
	// (at most 8 bytes per pixel, 2^35 bytes) is exact in 64 bits.
	constexpr unsigned int MaxSide = 65536u;

	inline unsigned long long padded(const unsigned long long bytes) noexcept {
		return (bytes + 15ull) & ~15ull;
	}

	// Byte offsets of the maps inside the payload and the payload size (width and height at most
	// MaxSide; callers check).
	struct Layout {
		unsigned long long Biome = 0u, Height = 0u, Fixed = 0u, Payload = 0u;

		Layout(const unsigned int width, const unsigned int height, const unsigned int maps) noexcept {
			const unsigned long long px = static_cast<unsigned long long>(width) * height;
			unsigned long long at = 0u;
			if (maps & MapBiome) {
				this->Biome = at;
				at += padded(px * 2u);
			}
			if (maps & MapHeight) {
				this->Height = at;
				at += padded(px * 4u);
			}
			if (maps & MapFixed) {
				this->Fixed = at;
				at += padded(px * 2u);
			}
			this->Payload = at;
		}
		unsigned long long record() const noexcept { return HeaderBytes + this->Payload; }
	};

	// Where one record's maps come from and go (device pointers, byte offsets into the stream).
	struct PackEntry {
		const float* Height;
		const unsigned short* Fixed;
		unsigned long long Payload; // offset of the payload in the stream
		unsigned long long HeightAt, FixedAt, PayloadBytes;
	};

	// Copies the FP32 and UINT16 maps (`pixels` each) of every entry (device table) into the stream
	// `out`. Asynchronous on `stream`.
	void pack(const PackEntry* entry, unsigned int count, unsigned long long pixels, unsigned int maps, unsigned char* out,
		hipStream_t stream);
	// Sums every entry's payload check into check[i] (zeroed by the caller). Asynchronous.
	void check(const PackEntry* entry, unsigned int count, unsigned long long payload, const unsigned char* out,
		unsigned int* check, hipStream_t stream);

	// Writes a header at `at`.
	void writeHeader(unsigned char* at, int x, int y, int origin_x, int origin_y, unsigned int width, unsigned int height,
		unsigned int completeness, unsigned int maps, unsigned long long payload, unsigned int check) noexcept;

	// A parsed record (pointers into the caller's bytes).
	struct View {
		int X, Y, OriginX, OriginY;
		unsigned int Width, Height, Completeness, Maps, Check;
		const unsigned short* Biome;
		const float* HeightMap;
		const unsigned short* Fixed;
		unsigned long long Bytes; // of the whole record
	};
	// Validates and parses the record at `data` (`size` bytes available): magic, version, header
	// size, map bits, sizes and the check. Throws STPInvalidArgument on any mismatch.
	View read(const unsigned char* data, unsigned long long size);

} // namespace SuperTerrainPlus::STPChunkRecord

#endif // STP_AMD_CHUNK_RECORD_HPP
