// Device-side improved 2D simplex noise for CDNA4 (gfx950).
//
// Header-only on purpose: the built-in generator kernels include it, and runtime-compiled
// (hipRTC) diversity-generator plugins receive it as an external header, so both run the
// same device code. It uses no standard-library headers so that hipRTC can compile it.
//
// Behaviour follows SuperTerrain+/SuperAlgorithm+/Device/Private/STPSimplexNoise.cu:5-110
// (FP32, same operation order). Bit-exactness against the FP32 CPU oracle requires that
// no FMA contraction happens: the library and every hipRTC program are compiled with
// -ffp-contract=off.
//
// MI355X layout: the table a workgroup samples from is staged once in LDS as
//   Perm[512]   : the permutation bytes,
//   GradOf[512] : GradOf[k] = gradient[Perm[k] % N] as float2,
// so a corner costs one LDS byte read and one 8-byte LDS read instead of two byte reads,
// an integer modulo and two 4-byte global reads.
#pragma once
#ifndef STP_AMD_SIMPLEX_NOISE_HPP
#define STP_AMD_SIMPLEX_NOISE_HPP

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace SuperTerrainPlus::STPAlgorithm {

	// STPPermutation.hpp:9-22 (device pointers; same field meaning as the reference).
	struct STPPermutation {
		const unsigned char* Permutation; // 512 entries (256 shuffled, repeated)
		const float* Gradient2D;          // 2 * Gradient2DSize floats
		unsigned int Gradient2DSize;
	};

	namespace STPSimplexNoise {

		// STPSimplexNoise.cuh:29-52
		struct STPFractalSimplexInformation {
			float Persistence;
			float Lacunarity;
			unsigned int Octave;
			float Scale;
			float2 Offset;
			float2 HalfDimension;
			float Amplitude = 1.0f;
			float Frequency = 1.0f;
		};

		// The LDS-resident table. Declare one per workgroup as __shared__ and fill it with
		// stage() before sampling.
		struct STPStagedPermutation {
			unsigned char Perm[512];
			float2 GradOf[512];
		};

		__device__ inline void stage(const STPPermutation& p, STPStagedPermutation& s) {
			const unsigned int tid = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y,
				nthr = blockDim.x * blockDim.y * blockDim.z;
			for (unsigned int k = tid; k < 512u; k += nthr) {
				const unsigned char v = p.Permutation[k];
				const unsigned int g = v % p.Gradient2DSize;
				s.Perm[k] = v;
				s.GradOf[k] = make_float2(p.Gradient2D[2u * g], p.Gradient2D[2u * g + 1u]);
			}
			__syncthreads();
		}

		namespace Detail {
			constexpr float F2 = 0.3660254038f; // 0.5 * (sqrt(3) - 1), same literal as the reference
			constexpr float G2 = 0.2113248654f; // (3 - sqrt(3)) / 6
			constexpr float H2 = -1.0f + 2.0f * G2;

			// STPSimplexNoise.cu:14-16: truncation after shifting non-positive values down by one,
			// so floori(0.0f) == -1 (this differs from floor and is kept on purpose).
			__device__ inline int floori(const float v) {
				return v > 0.0f ? static_cast<int>(v) : static_cast<int>(v - 1.0f);
			}

			__device__ inline float corner(const float dx, const float dy, const float2 g) {
				float w = 0.5f - dx * dx - dy * dy;
				if (w <= 0.0f) {
					return 0.0f;
				}
				w *= w;
				return w * w * (g.x * dx + g.y * dy);
			}

			// Shared skew/unskew; Fetch(jj, ii, i1, j1) returns the three corner gradients.
			template <class Fetch>
			__device__ inline float simplex(const float x, const float y, Fetch&& fetch) {
				const float s = (x + y) * F2;
				const int i = floori(x + s), j = floori(y + s);
				const float t = static_cast<float>(i + j) * G2;
				const float X0 = static_cast<float>(i) - t, Y0 = static_cast<float>(j) - t;
				const float dx0 = x - X0, dy0 = y - Y0;
				const unsigned int i1 = (dx0 > dy0) ? 1u : 0u, j1 = 1u - i1;
				const float dx1 = dx0 - static_cast<float>(i1) + G2, dy1 = dy0 - static_cast<float>(j1) + G2;
				const float dx2 = dx0 + H2, dy2 = dy0 + H2;
				const unsigned int ii = static_cast<unsigned int>(i) & 255u, jj = static_cast<unsigned int>(j) & 255u;
				float2 g0, g1, g2;
				fetch(ii, jj, i1, j1, g0, g1, g2);
				const float c0 = corner(dx0, dy0, g0), c1 = corner(dx1, dy1, g1), c2 = corner(dx2, dy2, g2);
				return 70.0f * (c0 + c1 + c2);
			}
		} // namespace Detail

		// Sample from the LDS-staged table.
		__device__ inline float simplex2D(const STPStagedPermutation& s, const float x, const float y) {
			return Detail::simplex(x, y,
				[&s](const unsigned int ii, const unsigned int jj, const unsigned int i1, const unsigned int j1, float2& g0,
					float2& g1, float2& g2) {
					g0 = s.GradOf[ii + s.Perm[jj]];
					g1 = s.GradOf[ii + i1 + s.Perm[jj + j1]];
					g2 = s.GradOf[ii + 1u + s.Perm[jj + 1u]];
				});
		}

		// Reference-signature overload reading the table from global memory (for plugins that
		// do not stage). STPSimplexNoise.cuh:62.
		__device__ inline float simplex2D(const STPPermutation& p, const float x, const float y) {
			return Detail::simplex(x, y,
				[&p](const unsigned int ii, const unsigned int jj, const unsigned int i1, const unsigned int j1, float2& g0,
					float2& g1, float2& g2) {
					const unsigned char* const perm = p.Permutation;
					const unsigned int n = p.Gradient2DSize;
					const unsigned int k0 = perm[ii + perm[jj]] % n, k1 = perm[ii + i1 + perm[jj + j1]] % n,
						k2 = perm[ii + 1u + perm[jj + 1u]] % n;
					g0 = make_float2(p.Gradient2D[2u * k0], p.Gradient2D[2u * k0 + 1u]);
					g1 = make_float2(p.Gradient2D[2u * k1], p.Gradient2D[2u * k1 + 1u]);
					g2 = make_float2(p.Gradient2D[2u * k2], p.Gradient2D[2u * k2 + 1u]);
				});
		}

		// STPSimplexNoise.cu:84-110. The two divisions by Scale are loop-invariant and are
		// hoisted; the per-octave operation order (divide, then multiply by frequency) is kept.
		template <class Table>
		__device__ inline float simplex2DFractal(const Table& table, const float x, const float y,
			const STPFractalSimplexInformation& desc) {
			const float baseX = ((x - desc.HalfDimension.x) + desc.Offset.x) / desc.Scale,
				baseY = ((y - desc.HalfDimension.y) + desc.Offset.y) / desc.Scale;
			float fractal = 0.0f, amplitude = desc.Amplitude, frequency = desc.Frequency, absRange = 0.0f;
			for (unsigned int o = 0u; o < desc.Octave; o++) {
				fractal += simplex2D(table, baseX * frequency, baseY * frequency) * amplitude;
				absRange += amplitude;
				amplitude *= desc.Persistence;
				frequency *= desc.Lacunarity;
			}
			const float v = (fractal + absRange) / (2.0f * absRange);
			// __saturatef: clamp to [0, 1], NaN -> 0
			return v > 1.0f ? 1.0f : (v > 0.0f ? v : 0.0f);
		}

	} // namespace STPSimplexNoise
} // namespace SuperTerrainPlus::STPAlgorithm

#endif // STP_AMD_SIMPLEX_NOISE_HPP
