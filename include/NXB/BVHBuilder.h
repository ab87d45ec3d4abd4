// NXB/BVHBuilder.h - drop-in for NexusBVH/include/NXB/BVHBuilder.h:18-55.
//
// Same entry points, signatures and ownership as the reference, implemented
// inline over the C ABI of libnxb.so (include/nxb.h):
//   BuildBVH2<PrimT>, BuildBVH8<PrimT>   (PrimT = Triangle or AABB, device pointer in)
//   ToHost(BVH2), FreeHostBVH(BVH2), FreeDeviceBVH(BVH2), FreeDeviceBVH(BVH8)
// Builds run on the null stream and block, as in the reference.
//
// Errors: like the reference (CudaUtils.h:6-15) a failing call prints the
// status and exits with code 99. Define NXB_THROW_ON_ERROR before including
// this header to get a std::runtime_error instead.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../nxb.h"
#include "AABB.h"
#include "BVH.h"
#include "BVHBuildMetrics.h"
#include "BuildConfig.h"
#include "Triangle.h"

namespace NXB
{
	namespace detail
	{
		inline void check(nxb_status s, const char* what)
		{
			if (s == NXB_SUCCESS)
				return;
#ifdef NXB_THROW_ON_ERROR
			throw std::runtime_error(std::string(what) + ": " + nxb_status_string(s));
#else
			std::fprintf(stderr, "NXB error = %d (%s) in %s\n", (int)s, nxb_status_string(s), what);
			std::exit(99);
#endif
		}

		template <typename PrimT> constexpr nxb_prim_type prim_type();
		template <> constexpr nxb_prim_type prim_type<Triangle>() { return NXB_PRIM_TRIANGLE; }
		template <> constexpr nxb_prim_type prim_type<AABB>() { return NXB_PRIM_AABB; }

		inline nxb_build_config to_c(const BuildConfig& b)
		{
			nxb_build_config c;
			nxb_default_config(&c);
			c.prioritize_speed = b.prioritizeSpeed ? 1u : 0u;
			c.morton_bits = b.mortonBits;
			c.search_radius = b.searchRadius;
			c.merge_mode = b.mergeMode;
			c.radius_policy = b.radiusPolicy;
			c.segment_log2 = b.segmentLog2;
			c.d_radius_table = b.radiusTable;
			c.radius_table_len = b.radiusTableLen;
			c.sort_digit_bits = b.sortDigitBits;
			c.spin_limit = b.spinLimit;
			c.metrics_flags = b.metricsFlags;
			c.d_mlp_params = b.mlpParams;
			c.mlp_min_confidence = b.mlpMinConfidence;
			c.slot_assignment = b.slotAssignment;
			c.hploc_blocks_per_cu = b.hplocBlocksPerCu;
			return c;
		}
	}

	/* \brief Builds a binary BVH from device primitives (AABB or Triangle). */
	template <typename PrimT>
	BVH2 BuildBVH2(PrimT* primitives, uint32_t primCount, BuildConfig buildConfig = BuildConfig(),
	               BVHBuildMetrics* buildMetrics = nullptr)
	{
		nxb_build_config c = detail::to_c(buildConfig);
		nxb_bvh2 out;
		detail::check(nxb_build_bvh2(primitives, detail::prim_type<PrimT>(), primCount, &c, &out,
		                             reinterpret_cast<nxb_metrics*>(buildMetrics), nullptr),
		              "BuildBVH2");
		BVH2 r;
		static_assert(sizeof(r) == sizeof(out), "BVH2 layout");
		std::memcpy(&r, &out, sizeof(r));
		return r;
	}

	/* \brief Builds a compressed wide BVH from device primitives (AABB or Triangle). */
	template <typename PrimT>
	BVH8 BuildBVH8(PrimT* primitives, uint32_t primCount, BuildConfig buildConfig = BuildConfig(),
	               BVHBuildMetrics* buildMetrics = nullptr)
	{
		nxb_build_config c = detail::to_c(buildConfig);
		nxb_bvh8 out;
		detail::check(nxb_build_bvh8(primitives, detail::prim_type<PrimT>(), primCount, &c, &out,
		                             reinterpret_cast<nxb_metrics*>(buildMetrics), nullptr),
		              "BuildBVH8");
		BVH8 r;
		static_assert(sizeof(r) == sizeof(out), "BVH8 layout");
		std::memcpy(&r, &out, sizeof(r));
		return r;
	}

	/* \brief Compressed wide BVH of an indexed device triangle mesh (three
	 * vertex indices per triangle); no counterpart in the reference. Equal to
	 * BuildBVH8 on the expanded triangles. */
	inline BVH8 BuildBVH8Indexed(const float3* vertices, uint32_t vertexCount, const uint32_t* indices,
	                             uint32_t triCount, BuildConfig buildConfig = BuildConfig(),
	                             BVHBuildMetrics* buildMetrics = nullptr)
	{
		nxb_build_config c = detail::to_c(buildConfig);
		nxb_bvh8 out;
		detail::check(nxb_build_bvh8_indexed(reinterpret_cast<const float*>(vertices), vertexCount, indices,
		                                     triCount, &c, &out, reinterpret_cast<nxb_metrics*>(buildMetrics),
		                                     nullptr),
		              "BuildBVH8Indexed");
		BVH8 r;
		std::memcpy(&r, &out, sizeof(r));
		return r;
	}

	/* \brief Binary BVH of an indexed device triangle mesh. */
	inline BVH2 BuildBVH2Indexed(const float3* vertices, uint32_t vertexCount, const uint32_t* indices,
	                             uint32_t triCount, BuildConfig buildConfig = BuildConfig(),
	                             BVHBuildMetrics* buildMetrics = nullptr)
	{
		nxb_build_config c = detail::to_c(buildConfig);
		nxb_bvh2 out;
		detail::check(nxb_build_bvh2_indexed(reinterpret_cast<const float*>(vertices), vertexCount, indices,
		                                     triCount, &c, &out, reinterpret_cast<nxb_metrics*>(buildMetrics),
		                                     nullptr),
		              "BuildBVH2Indexed");
		BVH2 r;
		std::memcpy(&r, &out, sizeof(r));
		return r;
	}

	/* \brief Copies a device BVH2 to host memory (free with FreeHostBVH). */
	inline BVH2 ToHost(BVH2 deviceBvh)
	{
		nxb_bvh2 d, h;
		std::memcpy(&d, &deviceBvh, sizeof(d));
		detail::check(nxb_bvh2_to_host(&d, &h), "ToHost");
		BVH2 r;
		std::memcpy(&r, &h, sizeof(r));
		return r;
	}

	inline void FreeHostBVH(BVH2 hostBvh)
	{
		nxb_bvh2 h;
		std::memcpy(&h, &hostBvh, sizeof(h));
		nxb_free_host_bvh2(&h);
	}

	/* \brief Copies a device BVH8 (nodes and primIdx) to host memory (free
	 * with FreeHostBVH). The reference has no BVH8 ToHost. */
	inline BVH8 ToHost(BVH8 deviceBvh)
	{
		nxb_bvh8 d, h;
		std::memcpy(&d, &deviceBvh, sizeof(d));
		detail::check(nxb_bvh8_to_host(&d, &h), "ToHost");
		BVH8 r;
		std::memcpy(&r, &h, sizeof(r));
		return r;
	}

	inline void FreeHostBVH(BVH8 hostBvh)
	{
		nxb_bvh8 h;
		std::memcpy(&h, &hostBvh, sizeof(h));
		nxb_free_host_bvh8(&h);
	}

	inline void FreeDeviceBVH(BVH2 deviceBvh)
	{
		nxb_bvh2 d;
		std::memcpy(&d, &deviceBvh, sizeof(d));
		detail::check(nxb_free_device_bvh2(&d), "FreeDeviceBVH");
	}

	inline void FreeDeviceBVH(BVH8 deviceBvh)
	{
		nxb_bvh8 d;
		std::memcpy(&d, &deviceBvh, sizeof(d));
		detail::check(nxb_free_device_bvh8(&d), "FreeDeviceBVH");
	}
}
