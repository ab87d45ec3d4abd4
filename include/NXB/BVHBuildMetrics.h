// NXB/BVHBuildMetrics.h - drop-in for NexusBVH/include/NXB/BVHBuildMetrics.h:7-107.
#pragma once

#include <cstdint>
#include <iostream>
#include <type_traits>
#include <utility>

#include "BVH.h"

namespace NXB
{
	struct BVHBuildMetrics
	{
		BVHBuildMetrics& operator+=(const BVHBuildMetrics& o)
		{
			computeSceneBoundsTime += o.computeSceneBoundsTime;
			computeMortonCodesTime += o.computeMortonCodesTime;
			radixSortTime += o.radixSortTime;
			bvhBuildTime += o.bvhBuildTime;
			bvh8ConversionTime += o.bvh8ConversionTime;
			totalTime += o.totalTime;
			bvh2Cost += o.bvh2Cost;
			bvh8Cost += o.bvh8Cost;
			averageChildPerNode += o.averageChildPerNode;
			return *this;
		}

		BVHBuildMetrics operator/(float d) const
		{
			BVHBuildMetrics r;
			r.computeSceneBoundsTime = computeSceneBoundsTime / d;
			r.computeMortonCodesTime = computeMortonCodesTime / d;
			r.radixSortTime = radixSortTime / d;
			r.bvhBuildTime = bvhBuildTime / d;
			r.bvh8ConversionTime = bvh8ConversionTime / d;
			r.totalTime = totalTime / d;
			r.bvh2Cost = bvh2Cost / d;
			r.bvh8Cost = bvh8Cost / d;
			r.averageChildPerNode = averageChildPerNode / d;
			return r;
		}

		// Timings (ms)
		float computeSceneBoundsTime = 0.0f;
		float computeMortonCodesTime = 0.0f;
		float radixSortTime = 0.0f;
		float bvhBuildTime = 0.0f;
		float bvh8ConversionTime = 0.0f;
		float totalTime = 0.0f;

		// SAH cost
		float bvh2Cost = 0.0f;
		float bvh8Cost = 0.0f;

		// For wide BVHs: average number of children per internal node
		float averageChildPerNode = 0.0f;
	};

	static_assert(sizeof(BVHBuildMetrics) == sizeof(nxb_metrics), "BVHBuildMetrics layout");

	void FreeDeviceBVH(BVH2 deviceBvh);
	void FreeDeviceBVH(BVH8 deviceBvh);

	// Warm up, then average the metrics of measuredIterations builds.
	template<typename Func, typename ...Args>
	BVHBuildMetrics BenchmarkBuild(Func&& func, uint32_t warmupIterations, uint32_t measuredIterations, Args&& ...args)
	{
		BVHBuildMetrics agg = {};
		using ReturnT = decltype(func(std::forward<Args>(args)..., nullptr));

		std::cout << std::endl << "========== BENCHMARKING BVH BUILD ==========" << std::endl << std::endl;
		for (uint32_t i = 0; i < warmupIterations; ++i)
		{
			BVHBuildMetrics dummy;
			auto bvh = std::forward<Func>(func)(std::forward<Args>(args)..., &dummy);
			FreeDeviceBVH(bvh);
		}
		for (uint32_t i = 0; i < measuredIterations; ++i)
		{
			BVHBuildMetrics m = {};
			auto bvh = std::forward<Func>(func)(std::forward<Args>(args)..., &m);
			FreeDeviceBVH(bvh);
			agg += m;
			std::cout << "Iteration " << i << ", total time: " << m.totalTime << " ms" << std::endl;
		}
		agg = agg / static_cast<float>(measuredIterations);

		std::cout << std::endl << "========== BENCHMARK RESULTS ==========" << std::endl << std::endl;
		std::cout << "Scene bounds: " << agg.computeSceneBoundsTime << " ms" << std::endl;
		std::cout << "Morton codes: " << agg.computeMortonCodesTime << " ms" << std::endl;
		std::cout << "Radix sort: " << agg.radixSortTime << " ms" << std::endl;
		std::cout << "BVH2 build time: " << agg.bvhBuildTime << " ms" << std::endl;
		if constexpr (std::is_same_v<ReturnT, BVH8>)
			std::cout << "BVH8 conversion time: " << agg.bvh8ConversionTime << " ms" << std::endl;
		std::cout << "Total BVH build time: " << agg.totalTime << " ms" << std::endl;
		std::cout << std::endl << "BVH2 cost: " << agg.bvh2Cost << std::endl;
		if constexpr (std::is_same_v<ReturnT, BVH8>)
		{
			std::cout << "BVH8 cost: " << agg.bvh8Cost << std::endl;
			std::cout << "Average children per node: " << agg.averageChildPerNode << std::endl;
		}
		std::cout << std::endl << "========== BENCHMARKING DONE ==========" << std::endl << std::endl;
		return agg;
	}
}
