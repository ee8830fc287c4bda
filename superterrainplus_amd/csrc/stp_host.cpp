// Host-side pieces of the path: setting validation, permutation + gradient tables, erosion
// brush, chunk coordinate math and the RAII memory helpers.
#include "stp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <random>

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPEnvironment;

// ------------------------------------------------------------------ memory
void STPSmartDeviceMemory::STPDeviceDeleter::operator()(void* const p) const noexcept {
	if (p) {
		(void)hipFree(p);
	}
}

void STPSmartDeviceMemory::STPPinnedDeleter::operator()(void* const p) const noexcept {
	if (p) {
		(void)hipHostFree(p);
	}
}

void* STPSmartDeviceMemory::allocateDevice(const size_t bytes) {
	void* p = nullptr;
	STP_CHECK_HIP(hipMalloc(&p, bytes == 0u ? 1u : bytes));
	return p;
}

void* STPSmartDeviceMemory::allocatePinned(const size_t bytes) {
	void* p = nullptr;
	STP_CHECK_HIP(hipHostMalloc(&p, bytes == 0u ? 1u : bytes, hipHostMallocDefault));
	return p;
}

// ------------------------------------------------------------------ settings
namespace {
	constexpr double TwoPi = 2.0 * 3.14159265358979323846264338327950288;

	bool isOdd(const uint2 v) {
		return (v.x & 1u) == 1u && (v.y & 1u) == 1u;
	}
	bool isPositive(const uint2 v) {
		return v.x > 0u && v.y > 0u;
	}
	bool inRange(const float v, const float lo, const float hi) {
		return v >= lo && v <= hi;
	}
} // namespace

// STPEnvironment.cpp:17-45
void STPChunkSetting::validate() const {
	constexpr const char* who = "STPChunkSetting";
	STP_ASSERT_ENV(isPositive(this->ChunkSize), who);
	STP_ASSERT_ENV(isPositive(this->MapSize), who);
	STP_ASSERT_ENV(this->ChunkScale.x > 0.0 && this->ChunkScale.y > 0.0, who);
	STP_ASSERT_ENV(isPositive(this->DiversityNearestNeighbour), who);
	STP_ASSERT_ENV(isPositive(this->ErosionNearestNeighbour), who);
	STP_ASSERT_ENV(isPositive(this->SplatNearestNeighbour), who);
	STP_ASSERT_ENV(isPositive(this->RenderDistance), who);
	STP_ASSERT_ENV(isOdd(this->DiversityNearestNeighbour), who);
	STP_ASSERT_ENV(isOdd(this->ErosionNearestNeighbour), who);
	STP_ASSERT_ENV(isOdd(this->SplatNearestNeighbour), who);
	STP_ASSERT_ENV(isOdd(this->RenderDistance), who);
}

// STPEnvironment.cpp:51-65
void STPRainDropSetting::validate() const {
	constexpr const char* who = "STPRainDropSetting";
	STP_ASSERT_ENV(this->RainDropCount > 0u, who);
	STP_ASSERT_ENV(inRange(this->Inertia, 0.0f, 1.0f), who);
	STP_ASSERT_ENV(this->SedimentCapacityFactor > 0.0f, who);
	STP_ASSERT_ENV(this->minSedimentCapacity >= 0.0f, who);
	STP_ASSERT_ENV(this->initWaterVolume > 0.0f, who);
	STP_ASSERT_ENV(this->minWaterVolume >= 0.0f, who);
	STP_ASSERT_ENV(inRange(this->Friction, 0.0f, 1.0f), who);
	STP_ASSERT_ENV(this->initSpeed >= 0.0f, who);
	STP_ASSERT_ENV(inRange(this->ErodeSpeed, 0.0f, 1.0f), who);
	STP_ASSERT_ENV(inRange(this->DepositSpeed, 0.0f, 1.0f), who);
	STP_ASSERT_ENV(inRange(this->EvaporateSpeed, 0.0f, 1.0f), who);
	STP_ASSERT_ENV(this->Gravity > 0.0f, who);
	STP_ASSERT_ENV(this->ErosionBrushRadius != 0u, who);
	// a drop must die in bounded time on the device (no reference counterpart)
	STP_ASSERT_ENV(this->lifetimeBound() <= 65535u, who);
}

unsigned int STPRainDropSetting::lifetimeBound() const {
	if (this->MaxLifetime != 0u) {
		return this->MaxLifetime;
	}
	// volume follows exactly the kernel's FP32 recurrence: v *= (1 - evaporate) until v < min
	float v = this->initWaterVolume;
	const float keep = 1.0f - this->EvaporateSpeed;
	unsigned int n = 0u;
	while (v >= this->minWaterVolume) {
		v *= keep;
		if (++n > 65535u) {
			return 65536u;
		}
	}
	return n;
}

// STPAlgorithmEnvironment.cpp:13-17
void STPSimplexNoiseSetting::validate() const {
	constexpr const char* who = "STPSimplexNoiseSetting";
	STP_ASSERT_ENV(this->Distribution != 0u, who);
	STP_ASSERT_ENV(this->Offset >= 0.0, who);
	STP_ASSERT_ENV(this->Offset < TwoPi, who);
	STP_ASSERT_ENV(this->Engine == STPPermutationEngine::MT19937_64 || this->Engine == STPPermutationEngine::MT19937, who);
}

// ------------------------------------------------------------------ permutation
// STPPermutationGenerator.cpp:40-94. std::shuffle is the standard library's (libstdc++ here), as in
// the reference; the gradient angle accumulates by `step` like the reference loop, bounded by
// the counter so the table never receives a (Distribution + 1)-th entry.
void STPAlgorithm::STPPermutationGenerator::generateHost(const STPSimplexNoiseSetting& setting,
	unsigned char* const permutation, float* const gradient) {
	setting.validate();
	for (unsigned int i = 0u; i < 256u; i++) {
		permutation[i] = static_cast<unsigned char>(i);
	}
	if (setting.Engine == STPPermutationEngine::MT19937_64) {
		std::mt19937_64 rng(setting.Seed);
		std::shuffle(permutation, permutation + 256, rng);
	} else {
		std::mt19937 rng(static_cast<std::mt19937::result_type>(setting.Seed));
		std::shuffle(permutation, permutation + 256, rng);
	}
	std::copy(permutation, permutation + 256, permutation + 256);

	const double step = TwoPi / (setting.Distribution * 1.0);
	unsigned int counter = 0u;
	for (double angle = 0.0; angle < TwoPi && counter < setting.Distribution; angle += step, counter++) {
		gradient[2u * counter] = static_cast<float>(std::cos(angle + setting.Offset));
		gradient[2u * counter + 1u] = static_cast<float>(std::sin(angle + setting.Offset));
	}
}

STPAlgorithm::STPPermutationGenerator::STPPermutationTable STPAlgorithm::STPPermutationGenerator::generate(
	const STPSimplexNoiseSetting& setting) {
	std::vector<unsigned char> perm(512u);
	std::vector<float> grad(2u * static_cast<size_t>(setting.Distribution));
	generateHost(setting, perm.data(), grad.data());
	STPPermutationTable table;
	table.PermutationMemory = STPSmartDeviceMemory::makeDevice<unsigned char>(perm.size());
	table.GradientMemory = STPSmartDeviceMemory::makeDevice<float>(grad.size());
	STP_CHECK_HIP(hipMemcpy(table.PermutationMemory.get(), perm.data(), perm.size(), hipMemcpyHostToDevice));
	STP_CHECK_HIP(hipMemcpy(table.GradientMemory.get(), grad.data(), grad.size() * sizeof(float), hipMemcpyHostToDevice));
	table.Permutation = STPPermutation {table.PermutationMemory.get(), table.GradientMemory.get(), setting.Distribution};
	return table;
}

// ------------------------------------------------------------------ chunk math (STPChunk.cpp:63-108)
int2 STPChunk::calcWorldChunkCoordinate(const double3 pointPos, const uint2 chunkSize, const double2 scale) noexcept {
	const double sx = static_cast<double>(chunkSize.x) * scale.x, sy = static_cast<double>(chunkSize.y) * scale.y;
	return make_int2(static_cast<int>(std::floor(pointPos.x / sx)) * static_cast<int>(chunkSize.x),
		static_cast<int>(std::floor(pointPos.z / sy)) * static_cast<int>(chunkSize.y));
}

uint2 STPChunk::calcLocalChunkCoordinate(const unsigned int chunkID, const uint2 chunkRange) noexcept {
	return make_uint2(chunkID % chunkRange.x, chunkID / chunkRange.x);
}

int2 STPChunk::offsetChunk(const int2 chunkCoord, const uint2 chunkSize, const int2 offset) noexcept {
	return make_int2(chunkCoord.x + static_cast<int>(chunkSize.x) * offset.x,
		chunkCoord.y + static_cast<int>(chunkSize.y) * offset.y);
}

int2 STPChunk::calcLocalChunkOrigin(const int2 centreChunkCoord, const uint2 chunkSize, const uint2 neighbourSize) noexcept {
	return STPChunk::offsetChunk(centreChunkCoord, chunkSize,
		make_int2(-static_cast<int>(neighbourSize.x / 2u), -static_cast<int>(neighbourSize.y / 2u)));
}

double2 STPChunk::calcChunkMapOffset(const int2 chunkCoord, const uint2 chunkSize, const uint2 mapSize,
	const double2 mapOffset) noexcept {
	return make_double2(static_cast<double>(chunkCoord.x) / static_cast<double>(chunkSize.x) * static_cast<double>(mapSize.x) + mapOffset.x,
		static_cast<double>(chunkCoord.y) / static_cast<double>(chunkSize.y) * static_cast<double>(mapSize.y) + mapOffset.y);
}

std::vector<int2> STPChunk::calcChunkNeighbourOffset(const uint2 chunkSize, const uint2 regionSize) {
	const unsigned int count = regionSize.x * regionSize.y;
	std::vector<int2> out(count);
	const int2 base = STPChunk::calcLocalChunkOrigin(make_int2(0, 0), chunkSize, regionSize);
	for (unsigned int i = 0u; i < count; i++) {
		const uint2 local = STPChunk::calcLocalChunkCoordinate(i, regionSize);
		out[i] = STPChunk::offsetChunk(base, chunkSize, make_int2(static_cast<int>(local.x), static_cast<int>(local.y)));
	}
	return out;
}

// ------------------------------------------------------------------ erosion brush
// STPHeightfieldGenerator.cpp:44-99: weights 1 - |d| / r in double, stored as float, then
// normalised as float((double)w_float * (1 / sum_double)).
void STPErosionBrushGenerator::generateOffsets(const unsigned int radius, std::vector<int2>& offset,
	std::vector<float>& weight) {
	STP_ASSERT_DOMAIN(radius > 0u, "Erosion brush radius must be a positive integer");
	const int r = static_cast<int>(radius), r2 = r * r;
	double sum = 0.0;
	offset.clear();
	weight.clear();
	for (int by = -r; by <= r; by++) {
		for (int bx = -r; bx <= r; bx++) {
			const double d2 = static_cast<double>(bx * bx + by * by);
			if (d2 < r2) {
				const double w = 1.0 - std::sqrt(d2) / r;
				sum += w;
				offset.push_back(make_int2(bx, by));
				weight.push_back(static_cast<float>(w));
			}
		}
	}
	const double factor = 1.0 / sum;
	for (float& w : weight) {
		w = static_cast<float>(w * factor);
	}
}

void STPErosionBrushGenerator::generate(const unsigned int dimension_x, const unsigned int radius,
	std::vector<int>& index, std::vector<float>& weight) {
	STP_ASSERT_DOMAIN(dimension_x > 0u, "Dimension row count should be positive");
	std::vector<int2> offset;
	generateOffsets(radius, offset, weight);
	index.resize(offset.size());
	for (size_t k = 0u; k < offset.size(); k++) {
		index[k] = offset[k].y * static_cast<int>(dimension_x) + offset[k].x;
	}
}
