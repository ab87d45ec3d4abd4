// nxb_benchmark.cpp - C++ user of the drop-in API, the counterpart of the
// reference benchmark driver (NexusBVH/Test/src/Main.cpp:66-76): uploads a
// triangle soup and runs BenchmarkBuild(BuildBVH8<Triangle>, warmup, measured).
//
// usage: nxb_benchmark [triangles=1000000] [warmup=5] [measured=10] [prioritizeSpeed=1]
#include <NXB/BVHBuilder.h>

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" void nxbs_uniform(uint32_t n, uint64_t seed, float half_extent, nxb_triangle* out);

int main(int argc, char** argv)
{
	uint32_t n = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 1000000u;
	uint32_t warmup = argc > 2 ? (uint32_t)std::strtoul(argv[2], nullptr, 10) : 5u;
	uint32_t measured = argc > 3 ? (uint32_t)std::strtoul(argv[3], nullptr, 10) : 10u;
	bool speed = argc > 4 ? std::atoi(argv[4]) != 0 : true;

	// Uniform small triangles in the unit cube (libnxb_scenes).
	std::vector<NXB::Triangle> tris(n);
	nxbs_uniform(n, 0x5EED2024ull, 0.005f, reinterpret_cast<nxb_triangle*>(tris.data()));

	NXB::Triangle* d = nullptr;
	if (hipMalloc(&d, sizeof(NXB::Triangle) * n) != hipSuccess ||
	    hipMemcpy(d, tris.data(), sizeof(NXB::Triangle) * n, hipMemcpyHostToDevice) != hipSuccess)
	{
		std::fprintf(stderr, "device upload failed\n");
		return 1;
	}

	NXB::BuildConfig cfg;
	cfg.prioritizeSpeed = speed;
	NXB::BVHBuildMetrics m = NXB::BenchmarkBuild(NXB::BuildBVH8<NXB::Triangle>, warmup, measured, d, n, cfg);
	std::printf("triangles %u  total %.4f ms  %.1f Mtri/s\n", n, m.totalTime, n / (m.totalTime * 1e3));

	NXB::BVH2 b2 = NXB::BuildBVH2<NXB::Triangle>(d, n, cfg);
	NXB::BVH2 h = NXB::ToHost(b2);
	std::printf("BVH2 nodes %u, root %u is %s\n", h.nodeCount, h.nodeCount - 1,
	            h.nodes[h.nodeCount - 1].leftChild != INVALID_IDX ? "internal" : "a leaf");
	NXB::FreeHostBVH(h);
	NXB::FreeDeviceBVH(b2);
	(void)hipFree(d);
	return 0;
}
