// Counter-based random numbers for raindrop spawning and synthetic biome maps.
//
// Philox4x32-10 written from Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11). Stateless: a drop's position is a pure function of
// (seed, drop index, stream id), so no per-drop generator state has to be initialised or
// stored (the reference keeps one cuRAND state per drop,
// SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:65-75).
#pragma once
#ifndef STP_AMD_RANDOM_HPP
#define STP_AMD_RANDOM_HPP

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace SuperTerrainPlus::STPRandom {

	struct Philox4 {
		unsigned int v[4];
	};

	__host__ __device__ inline Philox4 philox4x32_10(unsigned int c0, unsigned int c1, unsigned int c2, unsigned int c3,
		unsigned int k0, unsigned int k1) {
		for (int r = 0; r < 10; r++) {
			if (r > 0) {
				k0 += 0x9E3779B9u;
				k1 += 0xBB67AE85u;
			}
			const unsigned long long p0 = static_cast<unsigned long long>(0xD2511F53u) * c0,
				p1 = static_cast<unsigned long long>(0xCD9E8D57u) * c2;
			const unsigned int n0 = static_cast<unsigned int>(p1 >> 32) ^ c1 ^ k0,
				n2 = static_cast<unsigned int>(p0 >> 32) ^ c3 ^ k1;
			c1 = static_cast<unsigned int>(p1);
			c3 = static_cast<unsigned int>(p0);
			c0 = n0;
			c2 = n2;
		}
		return Philox4 {{c0, c1, c2, c3}};
	}

	// Uniform in (0, 1]: ((word >> 8) + 1) * 2^-24, exactly representable in FP32.
	__host__ __device__ inline float unitUniform(const unsigned int word) {
		return static_cast<float>((word >> 8) + 1u) * (1.0f / 16777216.0f);
	}

} // namespace SuperTerrainPlus::STPRandom

#endif // STP_AMD_RANDOM_HPP
