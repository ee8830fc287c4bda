// Biome layer chain on the GPU (STPBiomeLayer.hpp): bottom-up evaluation over rectangles.
//
// A plan walks the chain from the root to the leaves and gives every node, per map, the bounding
// rectangle of the coordinates its children read (a scale layer reads x >> 1 .. (x + 1) >> 1, a
// cross layer one pixel around, a Voronoi layer ((x - 2) >> 2) + {0, 1}, the ocean transition 8
// around), a place in the plan's buffer and a list of tiles. Running the plan is one launch per
// node over the tiles of a range of maps, parents first; the root writes the caller's maps, at
// a step between samples for coarse levels of detail. Every parent read is checked against the
// parent's rectangle (a miss sets a flag and reads 0), so a wrong rectangle is reported instead of
// reading foreign memory.
//
// Roofline: the root is the only large node (a Voronoi layer over the whole map, every lower node
// is 1/16 of the one below it or smaller). Per pixel it reads one parent sample (L2-resident) and
// writes 2 bytes; its work is 8 FP64 distances over jitters computed once per lattice point into
// LDS (8 64-bit seed mixes per lattice point, shared by 16 pixels). Integer/FP64 VALU bound.
#include "stp_internal.hpp"
#include "stp_stage_layer.hpp"

#include <stp/STPDeviceRuntime.hpp>

#include <algorithm>
#include <cstring>

using namespace SuperTerrainPlus;
using namespace SuperTerrainPlus::STPDiversity;
namespace L = SuperTerrainPlus::STPStage::Layer;

namespace {

	using L::Sample;
	using L::Seed;

	constexpr unsigned int LayerX = L::TileX, LayerY = L::TileY;
	// Voronoi tile: 64 x 16 pixels, 4 rows per thread; its lattice in LDS spans 18 x 6 points at the
	// two lattice heights of a 2D map (y = -1, 0)
	constexpr unsigned int VoronoiRows = 16u, LatticeX = 18u, LatticeZ = 6u, LatticePoints = LatticeX * LatticeZ * 2u;

	using L::LayerRect;
	using L::NodeArgs;
	using L::Reader;
	using L::limited;

	// One pixel of a non-Voronoi node (the demo layers' sample(x, y, z) with y = 0).
	__device__ Sample samplePixel(const NodeArgs& a, const Reader& p, const Reader& q, const int x, const int z) {
		const L::Registry& r = a.Reg;
		const Seed seed = a.LayerSeed;
		using K = STPLayerKind;
		switch (static_cast<K>(a.Kind)) {
		case K::Continent:
			// STPContinentLayer.h:16-22
			return L::sampler(seed, x, z).next<10u>() == 0u ? r[STPNamedBiome::Plains] : r[STPNamedBiome::Ocean];
		case K::Scale: {
			// STPScaleLayer.h:73-101
			const Sample i = p(x >> 1, z >> 1);
			const int xb = x & 1, zb = z & 1;
			L::Sampler rng = L::sampler(seed, x & -2, z & -2);
			if (xb == 0 && zb == 0) {
				return i;
			}
			const Sample l = p(x >> 1, (z + 1) >> 1);
			const Sample m = rng.choose(i, l);
			if (xb == 0) {
				return m;
			}
			const Sample n = p((x + 1) >> 1, z >> 1);
			const Sample o = rng.choose(i, n);
			if (zb == 0) {
				return o;
			}
			return L::scaleRule(i, n, l, p((x + 1) >> 1, (z + 1) >> 1), rng, a.Param != 0u);
		}
		case K::Land:
			// STPXCrossLayer.h:24-34
			return L::land(r, p(x, z), p(x + 1, z - 1), p(x + 1, z + 1), p(x - 1, z + 1), p(x - 1, z - 1),
				L::sampler(seed, x, z));
		case K::BaseBiome:
			return L::baseBiome(r, p(x, z), L::sampler(seed, x, z));
		case K::ClimateSingle:
			return L::climateSingle(r, p(x, z), L::sampler(seed, x, z));
		case K::Noise:
			return L::noise(r, p(x, z), L::sampler(seed, x, z));
		case K::OceanNoise:
			return L::oceanNoise(r, L::sampler(seed, x, z));
		case K::OceanMix: {
			// STPOceanTemperatureLayer.h:155-163
			const Sample land = p(x, z);
			return r.ocean(land) ? q(x, z) : land;
		}
		case K::RiverMix: {
			// STPRiverMixLayer.h:19-33
			const Sample land = p(x, z), river = q(x, z);
			if (r.ocean(land)) {
				return land;
			}
			return river == r[STPNamedBiome::River] && r.snowy(land) ? r[STPNamedBiome::FrozenRiver] : land;
		}
		case K::OceanTransition: {
			// STPOceanTemperatureLayer.h:121-143
			const Sample v = p(x, z);
			if (!r.ocean(v)) {
				return v;
			}
			for (int rx = -8; rx <= 8; rx += 4) {
				for (int rz = -8; rz <= 8; rz += 4) {
					if (r.ocean(p(x + rx, z + rz))) {
						continue;
					}
					if (v == r[STPNamedBiome::WarmOcean]) {
						return r[STPNamedBiome::LukewarmOcean];
					}
					if (v == r[STPNamedBiome::FrozenOcean]) {
						return r[STPNamedBiome::ColdOcean];
					}
				}
			}
			return v;
		}
		case K::Hills: {
			// STPHillsLayer.h:24-111
			L::Sampler rng = L::sampler(seed, x, z);
			const Sample land = p(x, z), noise = q(x, z);
			const Sample hill = static_cast<Sample>((static_cast<int>(noise) - 2) % 29);
			if (rng.next<3u>() == 0u || hill == 0u) {
				using B = STPNamedBiome;
				Sample l = land;
				if (land == r[B::Desert]) {
					l = r[B::DesertHills];
				} else if (land == r[B::Taiga]) {
					l = r[B::TaigaHills];
				} else if (land == r[B::Mountain]) {
					l = r[B::WoodedMountain];
				} else if (land == r[B::SnowyTundra] || land == r[B::SnowyTaiga]) {
					l = r[B::SnowyMountain];
				} else if (land == r[B::Plains]) {
					l = rng.next<3u>() == 0u ? r[B::ForestHills] : r[B::Forest];
				} else if (land == r[B::Forest]) {
					l = r[B::ForestHills];
				} else if (land == r[B::Jungle]) {
					l = r[B::JungleHills];
				} else if (land == r[B::Savannah]) {
					l = r[B::SavannahPlateau];
				} else if (land == r[B::Swamp]) {
					l = r[B::SwampHills];
				} else if (land == r[B::Badlands]) {
					l = r[B::BadlandsPlateau];
				} else if (land == r[B::Ocean]) {
					l = r[B::DeepOcean];
				} else if (land == r[B::WarmOcean]) {
					l = r[B::DeepWarmOcean];
				} else if (land == r[B::LukewarmOcean]) {
					l = r[B::DeepLukewarmOcean];
				} else if (land == r[B::ColdOcean]) {
					l = r[B::DeepColdOcean];
				} else if (land == r[B::FrozenOcean]) {
					l = r[B::DeepFrozenOcean];
				}
				if (r.ocean(land) && !r.shallowOcean(land) && rng.next<3u>() == 0u) {
					l = rng.next<2u>() == 0u ? r[B::Plains] : r[B::Forest];
				}
				if (l != land) {
					const unsigned int m = (land == p(x, z - 1)) + (land == p(x + 1, z)) + (land == p(x - 1, z))
						+ (land == p(x, z + 1));
					if (m >= 3u) {
						return l;
					}
				}
			}
			return land;
		}
		default: {
			// the cross layers (STPCrossLayer.h:24-34): centre, north, east, south, west
			const Sample c = p(x, z), n = p(x, z - 1), e = p(x + 1, z), s = p(x, z + 1), w = p(x - 1, z);
			switch (static_cast<K>(a.Kind)) {
			case K::Island: return L::island(r, c, n, e, s, w, L::sampler(seed, x, z));
			case K::ClimateModerate: return L::climateModerate(r, c, n, e, s, w);
			case K::ClimateExtreme: return L::climateExtreme(r, c, n, e, s, w);
			case K::DeepOcean: return L::deepOcean(r, c, n, e, s, w);
			case K::EaseEdge: return L::easeEdge(r, c, n, e, s, w);
			case K::EdgeBiome: return L::edgeBiome(r, c, n, e, s, w);
			case K::NoiseToRiver: return L::noiseToRiver(r, c, n, e, s, w);
			case K::OceanTemperate: return L::oceanTemperate(r, c, n, e, s, w);
			case K::OceanExtreme: return L::oceanExtreme(r, c, n, e, s, w);
			case K::RiverErode: return L::riverErode(r, c, n, e, s, w);
			case K::SmoothScale: return L::smoothScale(c, n, e, s, w, L::sampler(seed, x, z));
			default: return c;
			}
		}
		}
	}

	__global__ void __launch_bounds__(LayerX * LayerY) layerKERNEL(const NodeArgs a) {
		L::runNode(a, [](const NodeArgs& args, const Reader& p, const Reader& q, const int x, const int z) {
			return samplePixel(args, p, q, x, z);
		});
	}

	// The jitter of lattice point (px, py, pz) (STPVoronoiLayer.h:16-43), x, y and z components.
	__device__ inline void voronoiPoint(const Seed seed, const int px, const int py, const int pz, double* const j) {
		Seed m = L::mixSeed(seed, px);
		m = L::mixSeed(m, py);
		m = L::mixSeed(m, pz);
		m = L::mixSeed(m, px);
		m = L::mixSeed(m, py);
		m = L::mixSeed(m, pz);
		j[0] = L::voronoiJitter(m);
		m = L::mixSeed(m, static_cast<long long>(seed));
		j[1] = L::voronoiJitter(m);
		m = L::mixSeed(m, static_cast<long long>(seed));
		j[2] = L::voronoiJitter(m);
	}

	// STPVoronoiLayer.h:59-103 in 2D (y = 0): the nearest of the 8 jittered lattice points around
	// ((x - 2) / 4, -2 / 4, (z - 2) / 4) picks which parent sample the pixel takes. At step 1 the
	// tile's lattice points are jittered once, into LDS, and shared by 16 pixels each; at coarser
	// steps every pixel jitters its 8 points itself.
	__global__ void __launch_bounds__(LayerX * LayerY) voronoiKERNEL(const NodeArgs a) {
		__shared__ double jitter[LatticePoints][3];
		const uint2 t = a.Tile[blockIdx.x];
		const LayerRect rect = a.Rect[static_cast<size_t>(a.Node) * a.Count + t.x];
		const unsigned int tilesX = (rect.W + LayerX - 1u) / LayerX;
		const unsigned int bx = (t.y % tilesX) * LayerX, bz = (t.y / tilesX) * VoronoiRows;
		const bool staged = rect.Step == 1u; // uniform over the block: one map per tile
		const int tx0 = rect.X0 + static_cast<int>(bx), tz0 = rect.Z0 + static_cast<int>(bz);
		const int ax = (tx0 - 2) >> 2, az = (tz0 - 2) >> 2;
		if (staged) {
			const unsigned int tid = threadIdx.y * LayerX + threadIdx.x;
			if (tid < LatticePoints) {
				const int i = static_cast<int>(tid % LatticeX), rest = static_cast<int>(tid / LatticeX);
				voronoiPoint(a.VoronoiSeed, ax + i, rest % 2 - 1, az + rest / 2, jitter[tid]);
			}
			__syncthreads();
		}

		const Reader p(a, 0u, t.x);
		const unsigned int lx = bx + threadIdx.x;
		if (lx >= rect.W) {
			return;
		}
		const int x = rect.X0 + static_cast<int>(lx * rect.Step), ix = x - 2, lmx = ix >> 2;
		const double fx = static_cast<double>(ix & 3) / 4.0;
		for (unsigned int k = 0u; k < VoronoiRows / LayerY; k++) {
			const unsigned int lz = bz + threadIdx.y + k * LayerY;
			if (lz >= rect.H) {
				break;
			}
			const int z = rect.Z0 + static_cast<int>(lz * rect.Step), iz = z - 2, lmz = iz >> 2;
			const double fz = static_cast<double>(iz & 3) / 4.0;
			unsigned int best = 0u;
			double min = 0.0;
#pragma unroll
			for (unsigned int q = 0u; q < 8u; q++) {
				const unsigned int dx = (q >> 2u) & 1u, dy = (q >> 1u) & 1u, dz = q & 1u;
				const double gx = dx ? fx - 1.0 : fx, gy = dy ? 0.5 - 1.0 : 0.5, gz = dz ? fz - 1.0 : fz;
				double jx, jy, jz;
				if (staged) {
					const double* const j = jitter[((lmz + static_cast<int>(dz) - az) * 2 + static_cast<int>(dy))
						* static_cast<int>(LatticeX) + (lmx + static_cast<int>(dx) - ax)];
					jx = j[0];
					jy = j[1];
					jz = j[2];
				} else {
					double own[3];
					voronoiPoint(a.VoronoiSeed, lmx + static_cast<int>(dx), static_cast<int>(dy) - 1, lmz + static_cast<int>(dz), own);
					jx = own[0];
					jy = own[1];
					jz = own[2];
				}
				const double sz = gz + jz, sy = gy + jy, sx = gx + jx;
				const double d = sz * sz + sy * sy + sx * sx;
				if (q == 0u || d < min) {
					best = q;
					min = d;
				}
			}
			a.Out[rect.Offset + static_cast<size_t>(lz) * rect.Pitch + lx] =
				limited(a, p(lmx + static_cast<int>((best >> 2u) & 1u), lmz + static_cast<int>(best & 1u)));
		}
	}

	// Number of parents of each kind.
	constexpr unsigned int parentCount(const STPLayerKind k) {
		switch (k) {
		case STPLayerKind::Continent:
		case STPLayerKind::OceanNoise: return 0u;
		case STPLayerKind::Hills:
		case STPLayerKind::OceanMix:
		case STPLayerKind::RiverMix: return 2u;
		default: return 1u;
		}
	}

	// Inclusive rectangle of layer coordinates; empty when x0 > x1.
	struct Box {
		long long X0 = 1, Z0 = 1, X1 = 0, Z1 = 0;
		bool empty() const noexcept { return this->X0 > this->X1; }
		void join(const Box& b) noexcept {
			if (b.empty()) {
				return;
			}
			if (this->empty()) {
				*this = b;
				return;
			}
			this->X0 = std::min(this->X0, b.X0);
			this->Z0 = std::min(this->Z0, b.Z0);
			this->X1 = std::max(this->X1, b.X1);
			this->Z1 = std::max(this->Z1, b.Z1);
		}
		Box grown(const long long r) const noexcept {
			return Box {this->X0 - r, this->Z0 - r, this->X1 + r, this->Z1 + r};
		}
	};

	// floor division by a power of 2, as the arithmetic shift of an int
	constexpr long long shr(const long long v, const int s) {
		return v >> s;
	}

} // namespace

STPBiomeRegistry STPBiomeRegistry::demo() noexcept {
	STPBiomeRegistry r {};
	const auto set = [&r](const STPNamedBiome b, const unsigned short id, const float t, const float p) {
		const unsigned int i = static_cast<unsigned int>(b);
		r.ID[i] = id;
		r.Temperature[i] = t;
		r.Precipitation[i] = p;
	};
	// SuperDemo+/Biome.ini
	set(STPNamedBiome::Ocean, 0u, 1.5f, 1.5f);
	set(STPNamedBiome::Plains, 1u, 1.5f, 1.5f);
	set(STPNamedBiome::Desert, 2u, 2.8f, 0.2f);
	set(STPNamedBiome::Forest, 3u, 1.5f, 1.8f);
	return r;
}

int STPLayerProgram::add(const STPLayerKind kind, const unsigned long long salt, const int parent0, const int parent1,
	const unsigned int param, const unsigned long long cache_size) {
	this->Node.push_back(STPLayerNode {static_cast<unsigned int>(kind), salt, {parent0, parent1}, param, cache_size});
	return static_cast<int>(this->Node.size()) - 1;
}

STPLayerProgram STPLayerProgram::demo() {
	// STPAllLayers.cpp:53 (cache size 2048) and :82-105 (salts, scale types, Voronoi in 2D)
	constexpr unsigned long long Cache = 2048u;
	using K = STPLayerKind;
	STPLayerProgram p;
	const int a0 = p.add(K::Continent, 23457829ull, -1, -1, 0u, Cache);
	const int b0 = p.add(K::Scale, 875944ull, a0, -1, 1u, Cache);
	const int b1 = p.add(K::Land, 5748329ull, b0, -1, 0u, Cache);
	const int c0 = p.add(K::Scale, 8947358941ull, b1, -1, 0u, Cache);
	const int c1 = p.add(K::Land, 361249673ull, c0, -1, 0u, Cache);
	const int c2 = p.add(K::Land, 8769575ull, c1, -1, 0u, Cache);
	const int c3 = p.add(K::Land, 43562783426564ull, c2, -1, 0u, Cache);
	const int c4 = p.add(K::Island, 74368ull, c3, -1, 0u, Cache);
	const int d0 = p.add(K::Scale, 1ull, c4, -1, 0u, Cache);
	const int d1 = p.add(K::Scale, 2ull, d0, -1, 0u, Cache);
	const int d2 = p.add(K::Scale, 3ull, d1, -1, 0u, Cache);
	const int d3 = p.add(K::Voronoi, 4ull, d2, -1, 0u, Cache);
	const int d4 = p.add(K::Voronoi, 5ull, d3, -1, 0u, Cache);
	p.add(K::Voronoi, 6ull, d4, -1, 0u, 0u);
	return p;
}

struct STPLayerPlan::STPImplementation {
	unsigned int Count = 0u;
	// per node and map, the node's first tile of the map in the tile table; [node * (Count + 1) + map]
	std::vector<unsigned int> TileBegin;
	size_t RectBytes = 0u;
	// rectangles, then tiles, in one upload from one of two pinned staging buffers in turn (a
	// plan built while the previous upload is in flight does not wait for it)
	STPSmartDeviceMemory::STPHost<unsigned char> Staging[2];
	hipEvent_t Uploaded[2] {};
	size_t StagingCapacity[2] {};
	unsigned int Slot = 0u;
	STPSmartDeviceMemory::STPDevice<unsigned char> Table;
	size_t TableCapacity = 0u;
	STPSmartDeviceMemory::STPDevice<Sample> Arena;
	size_t ArenaCapacity = 0u;
	// host work space, kept between builds
	std::vector<Box> Bound;
	std::vector<LayerRect> Rect;
	std::vector<uint2> Tile;

	~STPImplementation() {
		for (hipEvent_t& e : this->Uploaded) {
			if (e) {
				(void)hipEventSynchronize(e);
				(void)hipEventDestroy(e);
			}
		}
	}
};

STPLayerPlan::STPLayerPlan() : Impl(std::make_unique<STPImplementation>()) {
	for (hipEvent_t& e : this->Impl->Uploaded) {
		STP_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
	}
}

STPLayerPlan::~STPLayerPlan() = default;

unsigned int STPLayerPlan::count() const noexcept {
	return this->Impl->Count;
}

namespace {
	// a compiled custom layer
	struct CustomKind {
		int Lo, Hi;
		unsigned int Shift, Parents;
		STPDeviceRuntimeProgram::STPSmartModule Module;
		hipFunction_t Kernel = nullptr;
	};

	constexpr char CustomKernelName[] = "STPCustomLayerKERNEL";
} // namespace

struct STPLayerBiomeFactory::STPImplementation {
	std::vector<CustomKind> Custom;
	std::vector<STPLayerNode> Node;
	std::vector<Seed> LayerSeed, VoronoiSeed;
	std::vector<unsigned char> Reachable;
	L::Registry Reg {};
	STPSmartDeviceMemory::STPDevice<unsigned char> Precipitation;
	STPSmartDeviceMemory::STPDevice<unsigned int> Fault;
	// the plan of generateDevice
	STPLayerPlan Plan;
};

STPLayerBiomeFactory::STPLayerBiomeFactory(const STPLayerProgram& program, const unsigned long long global_seed,
	const STPBiomeRegistry& registry, const uint2 dimension, const std::vector<STPCustomLayer>& custom) :
	BiomeDimension(dimension), Impl(std::make_unique<STPImplementation>()) {
	// STPBiomeFactory.cpp:20-22
	STP_ASSERT_DOMAIN(dimension.x > 0u && dimension.y > 0u,
		"Biomemap should have strictly positive dimension in both vector components");
	STPImplementation& m = *this->Impl;
	m.Node = program.Node;
	const size_t n = m.Node.size();
	if (n == 0u || n > 4096u) {
		STP_THROW(STPInvalidArgument, "a layer chain needs between 1 and 4096 layers");
	}
	// custom layers: the caller's sample function inside the library's node driver, one code object each
	for (size_t k = 0u; k < custom.size(); k++) {
		const STPCustomLayer& cl = custom[k];
		if (cl.Parents < 1u || cl.Parents > 2u || cl.Shift > 16u || cl.Lo > cl.Hi || cl.Lo < -1024 || cl.Hi > 1024) {
			STP_THROW(STPInvalidArgument, "custom layer " + std::to_string(k)
				+ ": 1 or 2 parents, Lo <= Hi within [-1024, 1024], Shift <= 16");
		}
		const std::string source = "#include <stp/STPLayerDevice.hpp>\n#line 1 \"custom layer " + std::to_string(k) + "\"\n"
			+ cl.Source + "\nextern \"C\" __global__ void __launch_bounds__(256) " + CustomKernelName
			+ "(const SuperTerrainPlus::STPLayerDevice::NodeArgs a) {\n"
			  "\tSuperTerrainPlus::STPLayerDevice::runNode(a, [](const SuperTerrainPlus::STPLayerDevice::NodeArgs& args,\n"
			  "\t\tconst SuperTerrainPlus::STPLayerDevice::Reader& p, const SuperTerrainPlus::STPLayerDevice::Reader& q,\n"
			  "\t\tconst int x, const int z) {\n"
			  "\t\treturn sample(SuperTerrainPlus::STPLayerDevice::STPLayerContext {args, p, q}, x, z);\n\t});\n}\n";
		STPDeviceRuntimeBinary::STPSourceInformation info;
		info.Option["-std=c++17"];
		for (const std::string& o : cl.Options) {
			info.Option[o];
		}
		const STPDeviceRuntimeBinary::STPExternalHeaderSource& headers = STPDeviceRuntimeBinary::libraryHeaders();
		for (const auto& h : headers) {
			info.ExternalHeader[h.first];
		}
		STPDeviceRuntimeBinary::STPCompilationOutput out =
			STPDeviceRuntimeBinary::compile("STPCustomLayer" + std::to_string(k), std::string(source), info, headers);
		STPDeviceRuntimeProgram::STPLinkerInformation link;
		link.DataOption.emplace_back(&out.ProgramObject, STPDeviceRuntimeProgram::STPBinaryType::CodeObject,
			STPDeviceRuntimeProgram::STPLinkerInformation::STPDataJitOption {});
		CustomKind kind {cl.Lo, cl.Hi, cl.Shift, cl.Parents, STPDeviceRuntimeProgram::link(link), nullptr};
		STP_CHECK_HIP(hipModuleGetFunction(&kind.Kernel, kind.Module.get(), CustomKernelName));
		m.Custom.push_back(std::move(kind));
	}
	for (size_t i = 0u; i < n; i++) {
		const STPLayerNode& node = m.Node[i];
		// STPLayer.cpp:68-73: a cache capacity is 0 or a power of 2
		STP_ASSERT_DOMAIN(node.CacheSize == 0u || (node.CacheSize & (node.CacheSize - 1u)) == 0u,
			"The capacity of the layer cache must be a power of 2");
		if (node.Kind >= static_cast<unsigned int>(STPLayerKind::Count)) {
			STP_THROW(STPInvalidArgument, "unknown layer kind");
		}
		if (node.Kind == static_cast<unsigned int>(STPLayerKind::Custom) && node.Param >= m.Custom.size()) {
			STP_THROW(STPInvalidArgument, "layer " + std::to_string(i) + " names a custom layer that was not given");
		}
		const unsigned int want = node.Kind == static_cast<unsigned int>(STPLayerKind::Custom) ? m.Custom[node.Param].Parents
			: parentCount(static_cast<STPLayerKind>(node.Kind));
		for (unsigned int k = 0u; k < 2u; k++) {
			const int p = node.Parent[k];
			if (k < want ? (p < 0 || static_cast<size_t>(p) >= i) : p != -1) {
				STP_THROW(STPInvalidArgument, "layer " + std::to_string(i) + " needs " + std::to_string(want)
					+ " parent(s) added before it, and -1 for the others");
			}
		}
		if (node.Kind == static_cast<unsigned int>(STPLayerKind::Voronoi) && node.Param > 1u) {
			STP_THROW(STPInvalidArgument, "Voronoi parameter: 0 (identity seed hash) or 1 (FNV-1a seed hash); 3D is not supported");
		}
		if (node.Kind == static_cast<unsigned int>(STPLayerKind::Scale) && node.Param > 1u) {
			STP_THROW(STPInvalidArgument, "scale parameter: 0 normal, 1 fuzzy");
		}
	}
	// the layer seeds (STPLayer.cpp:136-141) and the Voronoi seed std::hash<STPSeed_t>(global)
	// (STPVoronoiLayer.h:51-54): libstdc++ hashes an integer to itself, the MSVC library by FNV-1a
	// over its 8 bytes
	m.LayerSeed.resize(n);
	m.VoronoiSeed.resize(n);
	for (size_t i = 0u; i < n; i++) {
		m.LayerSeed[i] = L::seedLayer(global_seed, m.Node[i].Salt);
		Seed h = global_seed;
		if (m.Node[i].Param == 1u) {
			h = 14695981039346656037ull;
			for (unsigned int b = 0u; b < 8u; b++) {
				h ^= (global_seed >> (8u * b)) & 0xFFu;
				h *= 1099511628211ull;
			}
		}
		m.VoronoiSeed[i] = h;
	}
	// only layers the root depends on are evaluated
	m.Reachable.assign(n, 0u);
	m.Reachable[n - 1u] = 1u;
	for (size_t i = n; i-- > 0u;) {
		if (m.Reachable[i]) {
			for (const int p : m.Node[i].Parent) {
				if (p >= 0) {
					m.Reachable[static_cast<size_t>(p)] = 1u;
				}
			}
		}
	}
	// registry: ids, and the precipitation type of every id, the first registration of an id
	// deciding (STPBiomeRegistry.cpp:61-107, 125-137); an id nobody registered has none
	std::vector<unsigned char> type(65536u, 0u), seen(65536u, 0u);
	for (unsigned int b = 0u; b < NamedBiomeCount; b++) {
		m.Reg.Id[b] = registry.ID[b];
		const unsigned short id = registry.ID[b];
		if (!seen[id]) {
			seen[id] = 1u;
			type[id] = registry.Precipitation[b] < 1.0f ? 0u : (registry.Temperature[b] < 1.0f ? 2u : 1u);
		}
	}
	m.Precipitation = STPSmartDeviceMemory::makeDevice<unsigned char>(type.size());
	STP_CHECK_HIP(hipMemcpy(m.Precipitation.get(), type.data(), type.size(), hipMemcpyHostToDevice));
	m.Reg.Precipitation = m.Precipitation.get();
	m.Fault = STPSmartDeviceMemory::makeDevice<unsigned int>(1u);
	STP_CHECK_HIP(hipMemset(m.Fault.get(), 0, sizeof(unsigned int)));
}

STPLayerBiomeFactory::~STPLayerBiomeFactory() = default;

void STPLayerBiomeFactory::operator()(unsigned short* const biomemap, const int2 offset) {
	if (!biomemap) {
		STP_THROW(STPInvalidArgument, "null biome map");
	}
	const size_t pixel = static_cast<size_t>(this->BiomeDimension.x) * this->BiomeDimension.y;
	auto out = STPSmartDeviceMemory::makeDevice<unsigned short>(pixel);
	this->generateDevice(out.get(), &offset, 1u, nullptr);
	STP_CHECK_HIP(hipMemcpy(biomemap, out.get(), pixel * sizeof(unsigned short), hipMemcpyDeviceToHost));
	if (this->status() & 1u) {
		throw std::logic_error("biome layer read outside its parent's rectangle");
	}
}

void STPLayerBiomeFactory::generateDevice(unsigned short* const biomemap, const int2* const offset, const unsigned int count,
	const hipStream_t stream) {
	const uint2 d = this->BiomeDimension;
	this->generateDevice(biomemap, offset, count, d.x, d.y, d.x, static_cast<size_t>(d.x) * d.y, stream);
}

void STPLayerBiomeFactory::generateDevice(unsigned short* const biomemap, const int2* const origin, const unsigned int count,
	const unsigned int width, const unsigned int height, const size_t pitch, const size_t stride, const hipStream_t stream) {
	if (count == 0u) {
		return;
	}
	if (!biomemap || !origin) {
		STP_THROW(STPInvalidArgument, "null biome map or origins");
	}
	if (width == 0u || height == 0u || pitch < width || pitch > 0xFFFFFFFFu
		|| (count > 1u && stride < pitch * (height - 1u) + width)) {
		STP_THROW(STPInvalidArgument, "biome map rectangle: width and height positive, pitch >= width, maps not overlapping");
	}
	std::vector<STPLayerMap> maps(count);
	for (unsigned int c = 0u; c < count; c++) {
		maps[c] = STPLayerMap {origin[c], make_uint2(width, height), 1u, static_cast<unsigned int>(pitch), c * stride};
	}
	this->plan(this->Impl->Plan, maps, stream);
	this->run(this->Impl->Plan, biomemap, 0u, count, 65536u, stream);
}

void STPLayerBiomeFactory::plan(STPLayerPlan& plan, const std::vector<STPLayerMap>& maps, const hipStream_t stream) const {
	const STPImplementation& m = *this->Impl;
	STPLayerPlan::STPImplementation& q = *plan.Impl;
	const size_t n = m.Node.size(), root = n - 1u, count = maps.size();
	if (count == 0u || count > 0xFFFFFFFFu) {
		STP_THROW(STPInvalidArgument, "a plan needs at least one map");
	}
	// rectangles: from the root down, every node gets the bounding box of what its children read
	std::vector<Box>& box = q.Bound;
	box.assign(n * count, Box {});
	for (size_t c = 0u; c < count; c++) {
		const STPLayerMap& map = maps[c];
		const unsigned int step = map.Step == 0u ? 1u : map.Step;
		if (map.Size.x == 0u || map.Size.y == 0u || map.Pitch < map.Size.x) {
			STP_THROW(STPInvalidArgument, "a map needs a positive size and a pitch of at least its width");
		}
		box[root * count + c] = Box {map.Origin.x, map.Origin.y,
			map.Origin.x + static_cast<long long>(map.Size.x - 1u) * step, map.Origin.y + static_cast<long long>(map.Size.y - 1u) * step};
	}
	for (size_t i = n; i-- > 0u;) {
		if (!m.Reachable[i]) {
			continue;
		}
		const STPLayerNode& node = m.Node[i];
		const auto kind = static_cast<STPLayerKind>(node.Kind);
		for (size_t c = 0u; c < count; c++) {
			const Box& b = box[i * count + c];
			Box need;
			switch (kind) {
			case STPLayerKind::Continent:
			case STPLayerKind::OceanNoise: continue;
			case STPLayerKind::Scale:
				need = Box {shr(b.X0, 1), shr(b.Z0, 1), shr(b.X1 + 1, 1), shr(b.Z1 + 1, 1)};
				break;
			case STPLayerKind::Voronoi:
				need = Box {shr(b.X0 - 2, 2), shr(b.Z0 - 2, 2), shr(b.X1 - 2, 2) + 1, shr(b.Z1 - 2, 2) + 1};
				break;
			case STPLayerKind::OceanTransition: need = b.grown(8); break;
			case STPLayerKind::Custom: {
				const CustomKind& k = m.Custom[node.Param];
				need = Box {shr(b.X0 + k.Lo, static_cast<int>(k.Shift)), shr(b.Z0 + k.Lo, static_cast<int>(k.Shift)),
					shr(b.X1 + k.Hi, static_cast<int>(k.Shift)), shr(b.Z1 + k.Hi, static_cast<int>(k.Shift))};
				box[static_cast<size_t>(node.Parent[0]) * count + c].join(need);
				if (node.Parent[1] >= 0) {
					box[static_cast<size_t>(node.Parent[1]) * count + c].join(need);
				}
				continue;
			}
			case STPLayerKind::BaseBiome:
			case STPLayerKind::ClimateSingle:
			case STPLayerKind::Noise:
			case STPLayerKind::OceanMix:
			case STPLayerKind::RiverMix: need = b; break;
			default: need = b.grown(1); break; // cross, X cross, and the hills' land parent
			}
			box[static_cast<size_t>(node.Parent[0]) * count + c].join(need);
			if (node.Parent[1] >= 0) {
				// the second parent of hills and the mixes is read at the pixel itself
				box[static_cast<size_t>(node.Parent[1]) * count + c].join(b);
			}
		}
	}
	// any coordinate outside int is outside what the reference's int arithmetic can address
	for (const Box& b : box) {
		if (!b.empty() && (b.X0 < INT32_MIN || b.Z0 < INT32_MIN || b.X1 > INT32_MAX || b.Z1 > INT32_MAX)) {
			STP_THROW(STPInvalidArgument, "biome map coordinates out of the int range");
		}
	}
	// rectangles, buffer places and tiles
	std::vector<LayerRect>& rect = q.Rect;
	rect.assign(n * count, LayerRect {});
	std::vector<uint2>& tile = q.Tile;
	tile.clear();
	q.TileBegin.assign(n * (count + 1u), 0u);
	unsigned long long arena = 0u;
	for (size_t i = 0u; i < n; i++) {
		const bool voronoi = m.Node[i].Kind == static_cast<unsigned int>(STPLayerKind::Voronoi);
		const unsigned int rows = voronoi ? VoronoiRows : LayerY;
		for (size_t c = 0u; c < count; c++) {
			q.TileBegin[i * (count + 1u) + c] = static_cast<unsigned int>(tile.size());
			if (!m.Reachable[i]) {
				continue;
			}
			const Box& b = box[i * count + c];
			LayerRect& r = rect[i * count + c];
			r.X0 = static_cast<int>(b.X0);
			r.Z0 = static_cast<int>(b.Z0);
			if (i == root) {
				const STPLayerMap& map = maps[c];
				r.W = map.Size.x;
				r.H = map.Size.y;
				r.Pitch = map.Pitch;
				r.Step = map.Step == 0u ? 1u : map.Step;
				r.Offset = map.Offset;
			} else {
				const unsigned long long w = static_cast<unsigned long long>(b.X1 - b.X0 + 1),
					h = static_cast<unsigned long long>(b.Z1 - b.Z0 + 1);
				if (w > (1ull << 24u) || h > (1ull << 24u)) {
					STP_THROW(STPInvalidArgument, "a biome layer's rectangle is too large");
				}
				r.W = static_cast<unsigned int>(w);
				r.H = static_cast<unsigned int>(h);
				r.Pitch = r.W;
				r.Step = 1u;
				r.Offset = arena;
				arena += w * h;
			}
			const unsigned long long tiles = static_cast<unsigned long long>((r.W + LayerX - 1u) / LayerX) * ((r.H + rows - 1u) / rows);
			if (arena > (1ull << 32u) || tile.size() + tiles > 0x7FFFFFFFull) {
				STP_THROW(STPInvalidArgument, "a batch of biome maps needs too many layer samples (more than 2^32)");
			}
			for (unsigned long long t = 0u; t < tiles; t++) {
				tile.push_back(make_uint2(static_cast<unsigned int>(c), static_cast<unsigned int>(t)));
			}
		}
		q.TileBegin[i * (count + 1u) + count] = static_cast<unsigned int>(tile.size());
	}

	// upload through the staging buffer whose previous upload has finished; grow the device tables
	// after the device has finished with the old ones
	const size_t rectBytes = rect.size() * sizeof(LayerRect), bytes = rectBytes + tile.size() * sizeof(uint2);
	const unsigned int slot = q.Slot;
	q.Slot ^= 1u;
	STP_CHECK_HIP(hipEventSynchronize(q.Uploaded[slot]));
	if (bytes > q.TableCapacity || arena > q.ArenaCapacity) {
		STP_CHECK_HIP(hipDeviceSynchronize());
		if (bytes > q.TableCapacity) {
			q.TableCapacity = bytes + bytes / 4u;
			q.Table = STPSmartDeviceMemory::makeDevice<unsigned char>(q.TableCapacity);
		}
		if (arena > q.ArenaCapacity) {
			q.ArenaCapacity = static_cast<size_t>(arena + arena / 4u + 1u);
			q.Arena = STPSmartDeviceMemory::makeDevice<Sample>(q.ArenaCapacity);
		}
	}
	if (bytes > q.StagingCapacity[slot]) {
		q.StagingCapacity[slot] = bytes + bytes / 4u;
		q.Staging[slot] = STPSmartDeviceMemory::makeHost<unsigned char>(q.StagingCapacity[slot]);
	}
	unsigned char* const staging = q.Staging[slot].get();
	std::memcpy(staging, rect.data(), rectBytes);
	std::memcpy(staging + rectBytes, tile.data(), tile.size() * sizeof(uint2));
	STP_CHECK_HIP(hipMemcpyAsync(q.Table.get(), staging, bytes, hipMemcpyHostToDevice, stream));
	STP_CHECK_HIP(hipEventRecord(q.Uploaded[slot], stream));
	q.RectBytes = rectBytes;
	q.Count = static_cast<unsigned int>(count);
}

void STPLayerBiomeFactory::run(const STPLayerPlan& plan, unsigned short* const out, const unsigned int first,
	const unsigned int last, const unsigned int limit, const hipStream_t stream) const {
	const STPImplementation& m = *this->Impl;
	const STPLayerPlan::STPImplementation& q = *plan.Impl;
	if (!out || first > last || last > q.Count) {
		STP_THROW(STPInvalidArgument, "maps [first, last) of the plan, into a non-null buffer");
	}
	const size_t n = m.Node.size(), root = n - 1u, count = q.Count;
	const auto* const rects = reinterpret_cast<const LayerRect*>(q.Table.get());
	const auto* const tiles = reinterpret_cast<const uint2*>(q.Table.get() + q.RectBytes);
	for (size_t i = 0u; i < n; i++) {
		if (!m.Reachable[i]) {
			continue;
		}
		const unsigned int tb = q.TileBegin[i * (count + 1u) + first], te = q.TileBegin[i * (count + 1u) + last];
		if (te == tb) {
			continue;
		}
		const STPLayerNode& node = m.Node[i];
		NodeArgs a {};
		a.Kind = node.Kind;
		a.Param = node.Param;
		a.Count = static_cast<unsigned int>(count);
		a.Node = static_cast<unsigned int>(i);
		a.Limit = limit;
		a.Root = i == root;
		a.LayerSeed = m.LayerSeed[i];
		a.VoronoiSeed = m.VoronoiSeed[i];
		a.Out = i == root ? out : q.Arena.get();
		a.In = q.Arena.get();
		a.Parent[0] = node.Parent[0];
		a.Parent[1] = node.Parent[1];
		a.Rect = rects;
		a.Tile = tiles + tb;
		a.Reg = m.Reg;
		a.Fault = m.Fault.get();
		if (node.Kind == static_cast<unsigned int>(STPLayerKind::Custom)) {
			void* args[] = {&a};
			STP_CHECK_HIP(hipModuleLaunchKernel(m.Custom[node.Param].Kernel, te - tb, 1u, 1u, LayerX, LayerY, 1u, 0u, stream,
				args, nullptr));
			continue;
		}
		if (node.Kind == static_cast<unsigned int>(STPLayerKind::Voronoi)) {
			voronoiKERNEL<<<te - tb, dim3(LayerX, LayerY), 0, stream>>>(a);
		} else {
			layerKERNEL<<<te - tb, dim3(LayerX, LayerY), 0, stream>>>(a);
		}
		STP_CHECK_HIP(hipGetLastError());
	}
}

unsigned int STPLayerBiomeFactory::status() {
	unsigned int fault = 0u;
	STP_CHECK_HIP(hipMemcpy(&fault, this->Impl->Fault.get(), sizeof fault, hipMemcpyDeviceToHost));
	if (fault) {
		STP_CHECK_HIP(hipMemset(this->Impl->Fault.get(), 0, sizeof(unsigned int)));
	}
	return fault;
}
