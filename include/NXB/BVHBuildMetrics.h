// NXB/BVHBuildMetrics.h - drop-in for NexusBVH/include/NXB/BVHBuildMetrics.h:7-107.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

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
	struct BuildConfig;

	// What a build moves: primitive count, primitive record size and Morton
	// key width, read off the arguments of a BuildBVH2/BuildBVH8 call.
	struct BuildShape
	{
		uint32_t primCount = 0;
		uint32_t primBytes = 0;   // 36 (Triangle) or 24 (AABB); 0 if unknown
		uint32_t keyBytes = 8;    // 4 or 8
		bool bvh8 = false;        // a BVH8 build (its H-PLOC writes wide records)
	};

	// Stage order of BVHBuildMetrics: scene bounds, Morton codes, radix sort,
	// BVH2 (H-PLOC), BVH8 collapse.
	constexpr int kBuildStages = 5;

	// Algorithmic HBM bytes per primitive of each stage: what the stage has
	// to read and write at least, not what the caches see.
	//   bounds: primitive in, leaf node (32 B) out;
	//   Morton: leaf box (24 B) in, key and index out;
	//   sort:   one read and one write of key + index per 8-bit digit pass
	//           (4 passes for 32-bit keys, 8 for 64-bit) plus the keys read
	//           once for the digit counts;
	//   BVH2:   sorted key and index, the leaf box gathered, the internal node
	//           (32 B; in a BVH8 build its 64-B wide record instead), the
	//           cluster index written back, the parent word reset and its two
	//           exchanges;
	//   BVH8:   one 64-B wide record per primitive, the primitive index, and
	//           per wide node its 80 bytes and its work item written and read
	//           (16 B).
	inline double StageBytesPerPrim(int stage, const BuildShape& s, double bvh8NodesPerPrim)
	{
		const double k = s.keyBytes;
		switch (stage)
		{
		case 0: return s.primBytes + 32.0;
		case 1: return 24.0 + k + 4.0;
		case 2: return 2.0 * (k + 4.0) * (s.keyBytes == 4 ? 4.0 : 8.0) + k;
		case 3: return k + (s.bvh8 ? 108.0 : 76.0);
		case 4: return 68.0 + 96.0 * bvh8NodesPerPrim;
		default: return 0.0;
		}
	}

	namespace detail
	{
		template<typename T> struct PrimRecord { static constexpr uint32_t bytes = 0; };
		template<typename T> struct PrimRecord<T*> { static constexpr uint32_t bytes = sizeof(T); };
		template<typename T> struct PrimRecord<const T*> { static constexpr uint32_t bytes = sizeof(T); };

		// Build arguments are (primitives, primCount[, config]).
		template<typename ...Args>
		BuildShape ShapeOf(const Args& ...args)
		{
			BuildShape shape;
			const auto t = std::tie(args...);
			constexpr size_t n = sizeof...(Args);
			if constexpr (n >= 1)
				shape.primBytes = PrimRecord<std::decay_t<std::tuple_element_t<0, std::tuple<Args...>>>>::bytes;
			if constexpr (n >= 2 && std::is_integral_v<std::decay_t<std::tuple_element_t<1, std::tuple<Args...>>>>)
				shape.primCount = static_cast<uint32_t>(std::get<1>(t));
			if constexpr (n >= 3 && std::is_same_v<std::decay_t<std::tuple_element_t<2, std::tuple<Args...>>>, BuildConfig>)
			{
				const auto& cfg = std::get<2>(t);
				const uint32_t bits = cfg.mortonBits ? cfg.mortonBits : (cfg.prioritizeSpeed ? 32u : 64u);
				shape.keyBytes = bits / 8;
			}
			return shape;
		}

		inline void PrintStage(std::ostream& os, const char* name, const std::vector<BVHBuildMetrics>& runs,
			float BVHBuildMetrics::* field, double bytesPerPrim, uint32_t primCount)
		{
			double sum = 0.0, lo = 0.0;
			for (size_t i = 0; i < runs.size(); ++i)
			{
				const double v = runs[i].*field;
				sum += v;
				lo = i == 0 ? v : std::min(lo, v);
			}
			const double mean = runs.empty() ? 0.0 : sum / runs.size();
			os << std::left << std::setw(16) << name << std::right << std::fixed << std::setprecision(5)
			   << std::setw(10) << mean << std::setw(10) << lo;
			if (mean > 0.0 && primCount > 0)
			{
				const double mtris = primCount / (mean * 1e3);
				const double gbs = bytesPerPrim * primCount / (mean * 1e6);
				os << std::setprecision(1) << std::setw(12) << mtris;
				if (bytesPerPrim > 0.0)
					os << std::setw(10) << gbs << std::setw(9) << 100.0 * gbs / 8000.0 << "%";
			}
			os << "\n";
		}
	}

	// Runs warmupIterations builds, then measuredIterations builds with
	// metrics, and returns the mean of the measured metrics (the reference's
	// BenchmarkBuild contract, BVHBuildMetrics.h:62-107). The report gives,
	// per stage, the mean and the fastest time, the rate in million
	// primitives per second and the algorithmic bandwidth (StageBytesPerPrim)
	// with its share of the MI355X's 8 TB/s, and for the whole build the
	// mean, median and fastest time.
	template<typename Func, typename ...Args>
	BVHBuildMetrics BenchmarkBuild(Func&& func, uint32_t warmupIterations, uint32_t measuredIterations, Args&& ...args)
	{
		using ReturnT = decltype(func(std::forward<Args>(args)..., nullptr));
		constexpr bool wide = std::is_same_v<ReturnT, BVH8>;
		BuildShape shape = detail::ShapeOf(args...);
		shape.bvh8 = wide;

		for (uint32_t i = 0; i < warmupIterations; ++i)
		{
			BVHBuildMetrics scratch;  // with metrics, so the timing path is warm too
			FreeDeviceBVH(func(args..., &scratch));
		}

		std::vector<BVHBuildMetrics> runs;
		runs.reserve(measuredIterations);
		for (uint32_t i = 0; i < measuredIterations; ++i)
		{
			runs.emplace_back();
			FreeDeviceBVH(func(args..., &runs.back()));
		}

		BVHBuildMetrics mean;
		for (const BVHBuildMetrics& r : runs)
			mean += r;
		if (!runs.empty())
			mean = mean / static_cast<float>(runs.size());

		std::vector<float> totals;
		for (const BVHBuildMetrics& r : runs)
			totals.push_back(r.totalTime);
		std::sort(totals.begin(), totals.end());
		const float median = totals.empty() ? 0.0f : totals[totals.size() / 2];
		const float fastest = totals.empty() ? 0.0f : totals.front();

		// BVH8 nodes per primitive from the children per node:
		// children = nodes - 1 + primitives.
		const double nodesPerPrim = wide && mean.averageChildPerNode > 1.0f && shape.primCount > 1
			? (shape.primCount - 1.0) / (mean.averageChildPerNode - 1.0) / shape.primCount : 0.0;

		std::ostringstream os;
		os << "\n--- NXB build benchmark: " << (wide ? "BVH8" : "BVH2") << ", " << shape.primCount
		   << " primitives, " << 8 * shape.keyBytes << "-bit Morton, " << warmupIterations << " warm-up + "
		   << measuredIterations << " measured builds ---\n";
		os << std::left << std::setw(16) << "stage" << std::right << std::setw(10) << "mean ms" << std::setw(10)
		   << "min ms" << std::setw(12) << "Mprim/s" << std::setw(10) << "GB/s" << std::setw(10) << "of HBM" << "\n";
		static const char* const names[kBuildStages] = { "scene bounds", "Morton codes", "radix sort", "BVH2 H-PLOC",
			"BVH8 collapse" };
		static constexpr float BVHBuildMetrics::* fields[kBuildStages] = { &BVHBuildMetrics::computeSceneBoundsTime,
			&BVHBuildMetrics::computeMortonCodesTime, &BVHBuildMetrics::radixSortTime, &BVHBuildMetrics::bvhBuildTime,
			&BVHBuildMetrics::bvh8ConversionTime };
		double bytesTotal = 0.0;
		for (int st = 0; st < (wide ? kBuildStages : kBuildStages - 1); ++st)
		{
			const double b = shape.primBytes ? StageBytesPerPrim(st, shape, nodesPerPrim) : 0.0;
			bytesTotal += b;
			detail::PrintStage(os, names[st], runs, fields[st], b, shape.primCount);
		}
		detail::PrintStage(os, "whole build", runs, &BVHBuildMetrics::totalTime, bytesTotal, shape.primCount);
		os << std::fixed << std::setprecision(4) << "build time median " << median << " ms, fastest " << fastest
		   << " ms\n";
		os << std::setprecision(4) << "SAH: BVH2 " << mean.bvh2Cost;
		if (wide)
			os << ", BVH8 " << mean.bvh8Cost << ", children per BVH8 node " << mean.averageChildPerNode;
		os << "\n";
		std::cout << os.str() << std::flush;
		return mean;
	}
}
