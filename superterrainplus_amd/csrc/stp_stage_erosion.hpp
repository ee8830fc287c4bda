// Device stage functions of free-slip hydraulic erosion, shared by the per-chunk kernels
// (stp_kernels_erosion.hip) and the batched chunk-table kernels (stp_batch.hip).
//
// Behaviour follows SuperTerrain+/SuperTerrain+/Private/GPGPU/STPHeightfieldKernel.cu:80-108 (spawn)
// and SuperTerrain+/SuperTerrain+/Private/GPGPU/STPRainDrop.cu:18-164 (drop), plus: MaxLifetime
// (0 = unlimited), a stateless Philox spawn, a NaN-safe kill test and a footprint guard that keeps
// every access inside the map (see cellInside / spawnInside).
#pragma once
#ifndef STP_AMD_STAGE_EROSION_HPP
#define STP_AMD_STAGE_EROSION_HPP

#include "stp_internal.hpp"
#include <stp/STPRandom.hpp>

namespace SuperTerrainPlus::STPStage {

	struct ErosionArgs {
		STPRainDropParameter Rain;
		uint2 MapSize;     // the map drops spawn in (one chunk)
		uint2 SpawnOrigin; // its first cell in the merged map: (NN / 2) * MapSize for a neighbourhood
		uint2 Total;       // merged map
		unsigned int BrushSize;
		unsigned int StreamId;
	};

	__device__ inline float loadHeight(const float* p) {
		// relaxed agent-scope load: served by L2, never by a stale per-CU L1 line
		return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}

	template <bool Atomic>
	__device__ inline float readH(const float* map, const unsigned int i) {
		if constexpr (Atomic) {
			return loadHeight(map + i);
		} else {
			return map[i];
		}
	}

	// STPRainDrop.cu:23-51 -> (height, gradient x, gradient y); glm::mix(a, b, t) = a*(1-t) + b*t
	template <bool Atomic>
	__device__ inline float3 heightGradient(const float* map, const unsigned int W, const float px, const float py) {
		const unsigned int rx = static_cast<unsigned int>(px), ry = static_cast<unsigned int>(py);
		const float cx = px - static_cast<float>(rx), cy = py - static_cast<float>(ry);
		const unsigned int b = ry * W + rx;
		const float hx = readH<Atomic>(map, b), hy = readH<Atomic>(map, b + 1u), hz = readH<Atomic>(map, b + W),
			hw = readH<Atomic>(map, b + W + 1u);
		float3 r;
		r.x = hx * (1 - cx) * (1 - cy) + hy * cx * (1 - cy) + hz * (1 - cx) * cy + hw * cx * cy;
		r.y = (hy - hx) * (1 - cy) + (hw - hz) * cy;
		r.z = (hz - hx) * (1 - cx) + (hw - hy) * cx;
		return r;
	}

	struct DropState {
		float px, py, dx, dy, speed, volume, sediment;
	};

	__device__ inline DropState spawn(const ErosionArgs& a, const unsigned int drop) {
		// STPHeightfieldKernel.cu:93-102: centre chunk only, u * (Map - 1) + (NN / 2) * Map, the
		// second term being the spawn map's origin
		const STPRandom::Philox4 r = STPRandom::philox4x32_10(drop, a.StreamId, 0u, 0u,
			static_cast<unsigned int>(a.Rain.Seed), static_cast<unsigned int>(a.Rain.Seed >> 32));
		DropState d;
		d.px = STPRandom::unitUniform(r.v[0]) * static_cast<float>(a.MapSize.x - 1u) + static_cast<float>(a.SpawnOrigin.x);
		d.py = STPRandom::unitUniform(r.v[1]) * static_cast<float>(a.MapSize.y - 1u) + static_cast<float>(a.SpawnOrigin.y);
		d.dx = 0.0f;
		d.dy = 0.0f;
		d.speed = a.Rain.initSpeed;
		d.volume = a.Rain.initWaterVolume;
		d.sediment = 0.0f;
		return d;
	}

	// Moves the drop one cell; returns false when it dies (STPRainDrop.cu:78-105).
	template <bool Atomic>
	__device__ inline bool moveDrop(const ErosionArgs& a, const float* map, DropState& d, float3& hg) {
		hg = heightGradient<Atomic>(map, a.Total.x, d.px, d.py);
		const float inertia = a.Rain.Inertia;
		d.dx = (-hg.y) * (1 - inertia) + d.dx * inertia;
		d.dy = (-hg.z) * (1 - inertia) + d.dy * inertia;
		const float length = sqrtf(d.dx * d.dx + d.dy * d.dy);
		if (length != 0.0f) {
			d.dx /= length;
			d.dy /= length;
		}
		d.px += d.dx;
		d.py += d.dy;
		const float radius = static_cast<float>(a.Rain.ErosionBrushRadius);
		// STPRainDrop.cu:96-105, written so that a NaN position also dies
		return !(d.dx == 0.0f && d.dy == 0.0f) && d.px >= radius && d.px < 1.0f * static_cast<float>(a.Total.x) - radius
			&& d.py >= radius && d.py < 1.0f * static_cast<float>(a.Total.y) - radius;
	}

	// Footprint guard (no reference counterpart): the cell a drop erodes/deposits at must keep its
	// brush (reach r-1) and its 2x2 cell (reach +1) inside the map. In exact arithmetic the kill
	// test already implies this; it can only fail when a unit direction rounds above 1 in
	// magnitude. Spawns whose 2x2 cell leaves the map (u == 1 on a 1x1 neighbourhood, which the
	// reference reads out of bounds for) die before their first step.
	__device__ inline bool cellInside(const ErosionArgs& a, const unsigned int cx, const unsigned int cy) {
		const unsigned int r = a.Rain.ErosionBrushRadius, reach = r > 2u ? r - 1u : 1u;
		return cx + 1u >= r && cy + 1u >= r && cx + reach <= a.Total.x - 1u && cy + reach <= a.Total.y - 1u;
	}

	// The kill test's bounds on the moved position (STPRainDrop.cu:96-105): x in [r, Total.x - r),
	// y in [r, Total.y - r), as (x lo, x hi, y lo, y hi).
	__device__ inline float4 killBounds(const ErosionArgs& a) {
		const float radius = static_cast<float>(a.Rain.ErosionBrushRadius);
		return make_float4(radius, 1.0f * static_cast<float>(a.Total.x) - radius, radius,
			1.0f * static_cast<float>(a.Total.y) - radius);
	}

	__device__ inline bool spawnInside(const ErosionArgs& a, const DropState& d) {
		return static_cast<unsigned int>(d.px) + 1u <= a.Total.x - 1u
			&& static_cast<unsigned int>(d.py) + 1u <= a.Total.y - 1u;
	}

	// Fast = the hardware square root (1 ulp) of the parallel modes, which are assessed
	// statistically; the serialized kernel keeps the correctly rounded sqrtf.
	template <bool Fast = false>
	__device__ inline void advance(const ErosionArgs& a, DropState& d, const float deltaHeight) {
		const float v = fmaxf(0.0f, d.speed * d.speed + deltaHeight * a.Rain.Gravity);
		if constexpr (Fast) {
			d.speed = __builtin_amdgcn_sqrtf(v);
		} else {
			d.speed = sqrtf(v);
		}
		d.speed *= 1.0f - a.Rain.Friction;
		d.volume *= (1.0f - a.Rain.EvaporateSpeed);
	}

	// A drop's counters as two words: {steps | erode steps << 16, deposit steps | killed << 16}.
	// STPRainDropSetting::validate bounds every drop's lifetime by 65535 steps; each count saturates
	// at 0xFFFF all the same, so that no count can carry into its neighbour's field.
	__device__ inline uint2 packDropInfo(const unsigned int steps, const unsigned int erode_steps,
		const unsigned int deposit_steps, const unsigned int killed) {
		return make_uint2(min(steps, 0xFFFFu) | (min(erode_steps, 0xFFFFu) << 16),
			min(deposit_steps, 0xFFFFu) | (killed << 16));
	}

	__device__ inline bool alive(const ErosionArgs& a, const DropState& d, const unsigned int life) {
		return (a.Rain.MaxLifetime == 0u || life < a.Rain.MaxLifetime) && d.volume >= a.Rain.minWaterVolume;
	}

	// Cell of a brush tap: the (possibly negative) linear tap offset is added to the cell index
	// in 32-bit arithmetic. The footprint guards keep the result inside the surface.
	__device__ inline unsigned int tapIndex(const unsigned int map_index, const int tap) {
		return static_cast<unsigned int>(static_cast<int>(map_index) + tap);
	}

	// Erode update of all taps of a step on a surface with per-tap `take`; returns the sediment
	// taken in height units. Taps > 0: straight-line code, the heights of a group of taps are
	// fetched, then their updates issued, with one wait per group. The taps of one step are
	// distinct cells, so fetching a group's heights before updating leaves the drop's own result
	// unchanged (the reference reads and updates them one by one). Taps == 0 loops over `count`.
	template <class Surface, unsigned int Taps>
	__device__ inline float erodeGrouped(const Surface& s, const unsigned int map_index, const float amount,
		const int2* const brush, const unsigned int count);

	// The height surface a drop reads and updates. `load` gives the FP32 height; `Raw` is the
	// stored form, `raw` reads it. `take` removes min(height, amount) from a cell with one atomic
	// (the reference's erode update, STPRainDrop.cu:150-155); the brush weights it receives amounts
	// with are pre-multiplied by WeightScale, and the sum of what it returns converts to height units
	// with `height`. `add` adds (deposit).
	//
	// GlobalSurface: the whole map in HBM; reads bypass L1, updates are L2 FP32 atomics.
	struct GlobalSurface {
		typedef float Raw;
		float* Map;
		unsigned int W;
		__device__ unsigned int cell(const unsigned int x, const unsigned int y) const { return y * this->W + x; }
		__device__ unsigned int stride() const { return this->W; }
		__device__ float load(const unsigned int i) const { return loadHeight(this->Map + i); }
		__device__ Raw raw(const unsigned int i) const { return loadHeight(this->Map + i); }
		// height = unscaled(i) * Unit
		static constexpr float Unit = 1.0f;
		__device__ float unscaled(const unsigned int i) const { return loadHeight(this->Map + i); }
		static constexpr float WeightScale = 1.0f;
		__device__ float take(const unsigned int i, const Raw h, const float amount) const {
			const float ds = (h < amount) ? h : amount;
			unsafeAtomicAdd(this->Map + i, -ds);
			return ds;
		}
		__device__ static float height(const float taken) { return taken; }
		__device__ void add(const unsigned int i, const float v) const { unsafeAtomicAdd(this->Map + i, v); }
		// the chunk-level guards already keep every access inside the map
		__device__ bool holds(unsigned int, unsigned int, unsigned int, unsigned int) const { return true; }
		// a spawn may start; a step that left cell (cx, cy) for a position inside the kill bounds
		// (STPRainDrop.cu:96-105: [r, Total - r) on both axes) may erode at (cx, cy)
		__device__ bool spawnGuard(const ErosionArgs& a, const DropState& d) const { return spawnInside(a, d); }
		__device__ float4 moveBounds(const ErosionArgs& a) const { return killBounds(a); }
		__device__ bool stepGuard(const ErosionArgs& a, const unsigned int cx, const unsigned int cy) const {
			return cellInside(a, cx, cy);
		}
		// all taps of a step in one group: one HBM/L2 round trip
		static constexpr unsigned int groupOf(const unsigned int taps) { return taps; }
		template <unsigned int Taps>
		__device__ float erode(const unsigned int i, const float amount, const int2* const brush,
			const unsigned int count) const {
			return erodeGrouped<GlobalSurface, Taps>(*this, i, amount, brush, count);
		}
	};

	// LocalSurface: an LDS window [X0, X0 + RW) x [Y0, Y0 + RH) of the map (map coordinates),
	// held in 32-bit fixed point with FixedBits fractional bits. LDS FP32 atomic adds cost ~25x
	// an integer LDS atomic on gfx950 (measured: ~194 vs ~8 cycles per wave instruction and CU),
	// so the window uses integer atomics: every update is exact and order independent, and a
	// height converts to FP32 when read. The window guard kills a drop whose footprint would
	// leave the window.
	constexpr int FixedBits = 28; // heights in [-8, 8), resolution 2^-28
	constexpr float FixedScale = 268435456.0f, FixedStep = 1.0f / 268435456.0f;
	// The heights a window may start from: [0, 4]. From heights >= 0 a take leaves a cell at
	// old - min(old, want) >= 0 (only a concurrent refund can make `old` transiently negative, by
	// at most one tap's amount), and 4 leaves as much again for what the drops pile up. Below 0
	// the reference's update takes negative amounts and feeds them back as negative deposits
	// without bound: FP32 maps follow that, a fixed-point window would wrap.
	constexpr float WindowHeightLimit = 4.0f;
	// A window cell whose final value lies outside [WindowLowest, WindowHighest] (fixed point) has
	// left, or nearly left, the range the window can hold: heights start in [0, 4] and end >= 0 up
	// to rounding, and a cell piled up past 8 wraps to about -8. The tiled kernel flags such a cell
	// in the batch status rather than pass a wrapped height on silently.
	constexpr int WindowLowest = -(4 << FixedBits), WindowHighest = 15 << (FixedBits - 1);
	// 7.5 in fixed point is 15 * 2^27: below INT_MAX (15 << FixedBits would not be)
	static_assert(static_cast<long long>(WindowHighest) == 15ll << (FixedBits - 1), "window bound overflows");
	static_assert(WindowHighest > (7 << FixedBits) && WindowLowest < 0, "window bounds out of order");

	// 2^-4 in fixed point: from here up every fixed-point value that toFixedPoint makes converts
	// back to the height it came from, exactly (the float spacing at 2^-5 is 2^-28)
	constexpr int WindowExact = 1 << (FixedBits - 4);
	static_assert(WindowExact == 1 << 24, "exact store threshold moved");

	__device__ inline int toFixedPoint(const float h) { return __float2int_rn(h * FixedScale); }
	__device__ inline float fromFixedPoint(const int q) { return static_cast<float>(q) * FixedStep; }

	// The brush generator's weights (STPHeightfieldGenerator.cpp:44-99: 1 - |d| / r in double,
	// stored as float, times float(1 / sum) in double) of the radius 2, 3 and 4 brushes, multiplied
	// by FixedScale (exact: a power of two), by squared tap distance. As literals they cost no
	// register and no load; the batch engine checks them bit for bit against the generator when
	// it is built. 0 = not a tap.
	constexpr float shapedWeight(const int R, const int d2) {
		switch (R * 64 + d2) {
		case 2 * 64 + 0:
			return 0x1.eaf122p+25f;
		case 2 * 64 + 1:
			return 0x1.eaf122p+24f;
		case 2 * 64 + 2:
			return 0x1.1f964ep+24f;
		case 3 * 64 + 0:
			return 0x1.b4a8fp+24f;
		case 3 * 64 + 1:
			return 0x1.231b4cp+24f;
		case 3 * 64 + 2:
			return 0x1.cda208p+23f;
		case 3 * 64 + 4:
			return 0x1.231b4cp+23f;
		case 3 * 64 + 5:
			return 0x1.bcc574p+22f;
		case 3 * 64 + 8:
			return 0x1.8f9168p+20f;
		case 4 * 64 + 0:
			return 0x1.e9165ap+23f;
		case 4 * 64 + 1:
			return 0x1.6ed0c4p+23f;
		case 4 * 64 + 2:
			return 0x1.3c2b38p+23f;
		case 4 * 64 + 4:
			return 0x1.e9165ap+22f;
		case 4 * 64 + 5:
			return 0x1.af5bbep+22f;
		case 4 * 64 + 8:
			return 0x1.1e803p+22f;
		case 4 * 64 + 9:
			return 0x1.e9165ap+21f;
		case 4 * 64 + 10:
			return 0x1.99b82ap+21f;
		case 4 * 64 + 13:
			return 0x1.81d6fap+20f;
		default:
			return 0.0f;
		}
	}

	// The brush of radius R as the generator lists it (STPHeightfieldGenerator.cpp:44-99): rows
	// dy = -(R-1) .. R-1 in order, each the contiguous run dx = -h .. h with dx^2 + dy^2 < R^2.
	constexpr int brushRowHalf(const int R, const int dy) {
		int h = 0;
		while ((h + 1) * (h + 1) + dy * dy < R * R) {
			h++;
		}
		return h;
	}
	// brush radius of a tap count with a shaped (compile-time) form; 0 = none
	constexpr int shapedRadius(const unsigned int taps) {
		return taps == 9u ? 2 : (taps == 25u ? 3 : (taps == 45u ? 4 : 0));
	}

	// taps of the brush of radius R
	constexpr int shapedTaps(const int R) {
		int n = 0;
		for (int dy = 1 - R; dy < R; dy++) {
			n += 2 * brushRowHalf(R, dy) + 1;
		}
		return n;
	}

	// The erode update of the brush of radius R around `cell` (LDS, row stride rw), in fixed point.
	// Per tap: want = round(amount * w) and one returning atomic add of -want; the old height it
	// returns gives min(old, want), and a tap with old < want gets the excess back, so the cell ends
	// at old - min(old, want), the reference's erode update (STPRainDrop.cu:150-155). Within a row
	// the taps are consecutive words (one address, immediate offsets).
	//
	// shapedIssue: rows Dy .. R-1, every tap's atomic issued, -want (the atomic's addend) and old
	// kept at index J + j for the row's j-th tap.
	template <int R, int Dy, int J>
	__device__ inline void shapedIssue(int* const cell, const int rw, const float amount, int* const take, int* const old) {
		constexpr int H = brushRowHalf(R, Dy), N = 2 * H + 1;
		// the row's first tap; opaque to the optimiser, so that it keeps one address per row with
		// non-negative immediate offsets instead of one address per tap left of the centre
		// (rw is the window's row stride, the same on every lane: read it as a scalar, whatever
		// register the surrounding code left it in)
		int first = Dy * __builtin_amdgcn_readfirstlane(rw) - H;
		asm("" : "+s"(first));
		int* const p = cell + first;
#pragma unroll
		for (int j = 0; j < N; j++) {
			// -want, rounded the same way: fma(a, -w, -0.5) = -fma(a, w, 0.5) and truncation is odd
			take[J + j] = static_cast<int>(fmaf(amount, -shapedWeight(R, (j - H) * (j - H) + Dy * Dy), -0.5f));
			old[J + j] = atomicAdd(p + j, take[J + j]);
		}
		if constexpr (Dy < R - 1) {
			shapedIssue<R, Dy + 1, J + N>(cell, rw, amount, take, old);
		}
	}

	// Gives back want - old to every tap of rows Dy .. R-1 with old < want; returns what it gave.
	template <int R, int Dy, int J>
	__device__ inline int shapedRefund(int* const cell, const int rw, const int* const take, const int* const old) {
		constexpr int H = brushRowHalf(R, Dy), N = 2 * H + 1;
		int* const p = cell + (Dy * rw - H);
		int back = 0;
#pragma unroll
		for (int j = 0; j < N; j++) {
			const int want = -take[J + j];
			if (old[J + j] < want) {
				atomicAdd(p + j, want - old[J + j]);
				back += want - old[J + j];
			}
		}
		if constexpr (Dy < R - 1) {
			back += shapedRefund<R, Dy + 1, J + N>(cell, rw, take, old);
		}
		return back;
	}

	// The whole brush with every tap's atomic in flight at once (one wait per step; settling one
	// row at a time took 2% longer). Returns the sediment taken. min(old) >= max(want) proves that
	// no tap needs its excess back, so the per-tap refund test runs only when it fails (a cell
	// lower than the amount the brush takes from it: rare).
	template <int R>
	__device__ inline int shapedTakeAll(int* const cell, const int rw, const float amount) {
		constexpr int T = shapedTaps(R);
		int take[T], old[T];
		shapedIssue<R, 1 - R, 0>(cell, rw, amount, take, old);
		// sum = -sum(take), max(want) = -min(take)
		int sum = 0, low = old[0], least = take[0];
#pragma unroll
		for (int j = 0; j < T; j++) {
			sum += take[j];
			low = min(low, old[j]);
			least = min(least, take[j]);
		}
		sum = -sum;
		if (low < -least) {
			sum -= shapedRefund<R, 1 - R, 0>(cell, rw, take, old);
		}
		return sum;
	}

	struct LocalSurface {
		typedef int Raw;
		int* Lds;
		unsigned int X0, Y0, RW, RH;
		__device__ unsigned int cell(const unsigned int x, const unsigned int y) const {
			// window sides are far below 2^24
			return __umul24(y - this->Y0, this->RW) + (x - this->X0);
		}
		__device__ unsigned int stride() const { return this->RW; }
		__device__ float load(const unsigned int i) const { return fromFixedPoint(this->Lds[i]); }
		__device__ Raw raw(const unsigned int i) const { return this->Lds[i]; }
		// height = unscaled(i) * Unit exactly (a power-of-two scale)
		static constexpr float Unit = FixedStep;
		__device__ float unscaled(const unsigned int i) const { return static_cast<float>(this->Lds[i]); }
		// weights arrive pre-scaled to fixed point, so `amount` is already in fixed-point units
		static constexpr float WeightScale = FixedScale;
		__device__ float take(const unsigned int i, const Raw h, const float amount) const {
			const int want = __float2int_rn(amount);
			const int dq = (h < want) ? h : want;
			atomicSub(this->Lds + i, dq);
			return static_cast<float>(dq);
		}
		__device__ static float height(const float taken) { return taken * FixedStep; }
		// deposits are >= 0: round half up
		__device__ void add(const unsigned int i, const float v) const {
			atomicAdd(this->Lds + i, static_cast<int>(fmaf(v, FixedScale, 0.5f)));
		}
		__device__ bool holds(const unsigned int x0, const unsigned int y0, const unsigned int x1, const unsigned int y1) const {
			return x0 >= this->X0 && y0 >= this->Y0 && x1 < this->X0 + this->RW && y1 < this->Y0 + this->RH;
		}
		// The window must hold the footprint (brush reach and 2 x 2 cell) of the cell a drop will
		// erode or deposit at. Checked when the drop arrives at the cell (spawn, or the step that
		// moved it there), which also covers the 2 x 2 cell its gradient reads. Inside the window the
		// footprint is inside the map, so cellInside needs no separate test; the window's halo is
		// wide enough that in exact arithmetic no drop fails this test within its lifetime.
		__device__ bool footprint(const ErosionArgs& a, const unsigned int x, const unsigned int y) const {
			const unsigned int r = a.Rain.ErosionBrushRadius, lo = r - 1u, reach = r > 2u ? r - 1u : 1u;
			return this->holds(x - lo, y - lo, x + reach, y + reach);
		}
		__device__ bool spawnGuard(const ErosionArgs& a, const DropState& d) const {
			const unsigned int x = static_cast<unsigned int>(d.px), y = static_cast<unsigned int>(d.py);
			// cellInside keeps x, y >= r - 1, so the footprint test cannot wrap
			return spawnInside(a, d) && cellInside(a, x, y) && this->footprint(a, x, y);
		}
		// The footprint test of the cell a step moves to, folded into the kill test's bounds: for a
		// position p >= 0 and an integer k, floor(p) >= k <=> p >= k and floor(p) < k <=> p < k, so
		// footprint(floor(px), floor(py)) within the kill bounds is px in [X0 + r - 1, X0 + RW - reach)
		// and likewise for py. The bounds are integers below 2^24 (exact in FP32), and a NaN
		// position fails every comparison.
		__device__ float4 moveBounds(const ErosionArgs& a) const {
			const unsigned int r = a.Rain.ErosionBrushRadius, lo = r - 1u, reach = r > 2u ? r - 1u : 1u;
			const float4 k = killBounds(a);
			return make_float4(fmaxf(k.x, static_cast<float>(this->X0 + lo)), fminf(k.y, static_cast<float>(this->X0 + this->RW - reach)),
				fmaxf(k.z, static_cast<float>(this->Y0 + lo)), fminf(k.w, static_cast<float>(this->Y0 + this->RH - reach)));
		}
		__device__ bool stepGuard(const ErosionArgs&, unsigned int, unsigned int) const { return true; }
		static constexpr unsigned int groupOf(const unsigned int taps) { return taps; }
		// The 9/25/45-tap brushes go through shapedTakeAll. Concurrent drops can see a cell's height
		// transiently lowered by a refunded excess, a race of the same kind as the reference's
		// read-then-update. `amount` is held below 8 (the window's height range) so the fixed-point
		// tap sum stays below 2^31.
		template <unsigned int Taps>
		__device__ float erode(const unsigned int i, const float amount, const int2* const brush,
			const unsigned int count) const {
			constexpr int R = shapedRadius(Taps);
			if constexpr (R == 0) {
				return erodeGrouped<LocalSurface, Taps>(*this, i, amount, brush, count);
			} else {
				const int sum = shapedTakeAll<R>(this->Lds + i, static_cast<int>(this->RW), fminf(amount, 7.5f));
				return static_cast<float>(sum) * FixedStep;
			}
		}
	};

	template <class Surface, unsigned int Taps>
	__device__ inline float erodeGrouped(const Surface& s, const unsigned int map_index, const float amount,
		const int2* const brush, const unsigned int count) {
		float taken = 0.0f;
		if constexpr (Taps > 0u) {
			// the group is a compile-time divisor of Taps
			constexpr unsigned int G = Surface::groupOf(Taps);
#pragma unroll 1
			for (unsigned int base = 0u; base < Taps; base += G) {
				typename Surface::Raw h[G];
#pragma unroll
				for (unsigned int j = 0u; j < G; j++) {
					h[j] = s.raw(tapIndex(map_index, brush[base + j].x));
				}
#pragma unroll
				for (unsigned int j = 0u; j < G; j++) {
					const int2 b = brush[base + j];
					taken += s.take(tapIndex(map_index, b.x), h[j], amount * __int_as_float(b.y));
				}
			}
		} else {
			for (unsigned int j = 0u; j < count; j++) {
				const int2 b = brush[j];
				const unsigned int ei = tapIndex(map_index, b.x);
				taken += s.take(ei, s.raw(ei), amount * __int_as_float(b.y));
			}
		}
		return Surface::height(taken);
	}

	// STPRainDrop.cu:23-51 on a surface, for the parallel modes, in an algebraically equal form
	// with a third of the operations: with top = mix(hx, hy, cx) and bottom = mix(hz, hw, cx) the
	// height is mix(top, bottom, cy), the x slope mix(hy - hx, hw - hz, cy), and the y slope
	// (hz - hx)(1 - cx) + (hw - hy) cx is exactly bottom - top. Its roundings differ from the
	// reference's expanded form by an ulp; parallel erosion is assessed statistically, and the
	// serialized kernel keeps the reference's expression (heightGradient).
	// The heights are read unscaled and the results stay unscaled: the caller applies the surface's
	// Unit, a power of two, where it multiplies or subtracts anyway (exact: scaling by a power of
	// two commutes with rounding, and no value here is small enough to lose bits to it).
	//
	// A position's cell: its corner (X, Y), the offsets in it and its surface index. A drop's
	// position is >= 0 wherever it is read (spawned inside the map, or inside the kill bounds), so
	// the unsigned conversion is the floor.
	struct DropCell {
		unsigned int X, Y, Index;
		float Ox, Oy;
	};
	template <class Surface>
	__device__ inline DropCell dropCell(const Surface& s, const float px, const float py) {
		DropCell c;
		c.X = static_cast<unsigned int>(px);
		c.Y = static_cast<unsigned int>(py);
		c.Ox = px - static_cast<float>(c.X);
		c.Oy = py - static_cast<float>(c.Y);
		c.Index = s.cell(c.X, c.Y);
		return c;
	}
	template <class Surface>
	__device__ inline float3 surfaceGradient(const Surface& s, const DropCell& c) {
		const unsigned int b = c.Index, w = s.stride();
		const float hx = s.unscaled(b), hy = s.unscaled(b + 1u), hz = s.unscaled(b + w), hw = s.unscaled(b + w + 1u);
		const float topSlope = hy - hx, bottomSlope = hw - hz;
		const float top = fmaf(topSlope, c.Ox, hx), bottom = fmaf(bottomSlope, c.Ox, hz);
		float3 r;
		r.x = fmaf(bottom - top, c.Oy, top);
		r.y = fmaf(bottomSlope - topSlope, c.Oy, topSlope);
		r.z = bottom - top;
		return r;
	}

	// One drop in parallel mode on `s`. `brush` holds (linear tap offset for s.stride(), weight
	// bits). Taps > 0 is the compile-time brush size, Taps == 0 a loop over a.BrushSize taps; the
	// surface's `erode` does a step's erode update. The drop's final sediment and counters go to
	// ledger[slot] and drop_info[slot].
	template <class Surface, unsigned int Taps>
	__device__ inline void erodeDrop(const Surface& s, const ErosionArgs& a, const int2* const brush,
		const unsigned int drop, const unsigned int slot, float* const ledger, uint2* const drop_info) {
		DropState d = spawn(a, drop);
		float lost = 0.0f;
		unsigned int steps = 0u, erodeSteps = 0u, depositSteps = 0u, killed = 0u;
		if (!s.spawnGuard(a, d)) {
			d.volume = 0.0f;
			killed = 1u;
		}
		// the kill test's bounds, with the surface's own footprint test folded in where it has one
		const float4 bounds = s.moveBounds(a);
		// the drop's cell, carried from the step that moved the drop there
		DropCell c = dropCell(s, d.px, d.py);
		const float inertia = a.Rain.Inertia, gradientScale = (1 - inertia) * Surface::Unit;
		for (unsigned int life = 0u; alive(a, d, life); life++) {
			steps++;
			// unscaled: (-g Unit)(1 - inertia) = (-g)((1 - inertia) Unit) exactly
			const float3 hg = surfaceGradient(s, c);
			// STPRainDrop.cu:86-105 (direction, move, kill test)
			d.dx = (-hg.y) * gradientScale + d.dx * inertia;
			d.dy = (-hg.z) * gradientScale + d.dy * inertia;
			// parallel mode normalises with the hardware reciprocal square root (1 ulp): its result
			// is assessed statistically; the serialized kernel keeps the reference's division
			const float length2 = d.dx * d.dx + d.dy * d.dy;
			if (length2 != 0.0f) {
				const float inv = __builtin_amdgcn_rsqf(length2);
				d.dx *= inv;
				d.dy *= inv;
			}
			d.px += d.dx;
			d.py += d.dy;
			// one branch: every term is evaluated (a NaN position fails the float tests)
			const bool inMap = ((d.dx != 0.0f) | (d.dy != 0.0f)) & (d.px >= bounds.x) & (d.px < bounds.y)
				& (d.py >= bounds.z) & (d.py < bounds.w);
			if (!(inMap & s.stepGuard(a, c.X, c.Y))) {
				lost = d.sediment;
				d.volume = 0.0f;
				d.sediment = 0.0f;
				killed = 1u;
				break;
			}
			const DropCell next = dropCell(s, d.px, d.py);
			const unsigned int mapIndex = c.Index, W = s.stride();
			const float ox = c.Ox, oy = c.Oy;
			const float deltaHeight = (surfaceGradient(s, next).x - hg.x) * Surface::Unit;
			const float cap = fmaxf(-deltaHeight * d.speed * d.volume * a.Rain.SedimentCapacityFactor,
				a.Rain.minSedimentCapacity);
			if (d.sediment > cap || deltaHeight > 0.0f) {
				const float amount = (deltaHeight > 0.0f) ? fminf(deltaHeight, d.sediment)
														  : (d.sediment - cap) * a.Rain.DepositSpeed;
				d.sediment -= amount;
				s.add(mapIndex, amount * (1.0f - ox) * (1.0f - oy));
				s.add(mapIndex + 1u, amount * ox * (1.0f - oy));
				s.add(mapIndex + W, amount * (1.0f - ox) * oy);
				s.add(mapIndex + W + 1u, amount * ox * oy);
				depositSteps++;
			} else {
				const float amount = fminf((cap - d.sediment) * a.Rain.ErodeSpeed, -deltaHeight);
				d.sediment += s.template erode<Taps>(mapIndex, amount, brush, a.BrushSize);
				erodeSteps++;
			}
			advance<true>(a, d, deltaHeight);
			c = next;
		}
		if (ledger) {
			ledger[slot] = lost + d.sediment;
		}
		if (drop_info) {
			drop_info[slot] = packDropInfo(steps, erodeSteps, depositSteps, killed);
		}
	}

	// Brush sizes with a specialised (straight-line) drop: r = 2, 3, 4.
	template <class Surface>
	__device__ inline void erodeDropAnyBrush(const Surface& s, const ErosionArgs& a, const int2* const brush,
		const unsigned int drop, const unsigned int slot, float* const ledger, uint2* const drop_info) {
		switch (a.BrushSize) {
		case 9u:
			erodeDrop<Surface, 9u>(s, a, brush, drop, slot, ledger, drop_info);
			break;
		case 25u:
			erodeDrop<Surface, 25u>(s, a, brush, drop, slot, ledger, drop_info);
			break;
		case 45u:
			erodeDrop<Surface, 45u>(s, a, brush, drop, slot, ledger, drop_info);
			break;
		default:
			erodeDrop<Surface, 0u>(s, a, brush, drop, slot, ledger, drop_info);
			break;
		}
	}

	// STPHeightfieldKernel.cu:112-122: u16 = rint(saturate(h) * 65535), round half to even.
	__device__ inline unsigned short toFixed(const float h) {
		const float s = h > 1.0f ? 1.0f : (h > 0.0f ? h : 0.0f);
		return static_cast<unsigned short>(rintf(s * 65535.0f));
	}

} // namespace SuperTerrainPlus::STPStage

#endif // STP_AMD_STAGE_EROSION_HPP
