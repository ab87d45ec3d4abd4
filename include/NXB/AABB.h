// NXB/AABB.h - drop-in for NexusBVH/include/NXB/AABB.h:8-54 (same layout and
// members; 24 bytes). float3 is HIP's float3 (the reference uses CUDA's).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace NXB
{
	namespace detail
	{
		__host__ __device__ inline float3 min3(float3 a, float3 b) { return make_float3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
		__host__ __device__ inline float3 max3(float3 a, float3 b) { return make_float3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
	}

	struct AABB
	{
		AABB() = default;
		__host__ __device__ AABB(float3 v0, float3 v1)
		{
			bMin = detail::min3(v0, v1);
			bMax = detail::max3(v0, v1);
		}
		__host__ __device__ AABB(float3 v0, float3 v1, float3 v2)
		{
			bMin = detail::min3(v0, detail::min3(v1, v2));
			bMax = detail::max3(v0, detail::max3(v1, v2));
		}

		__host__ __device__ void Grow(float3 v)
		{
			bMin = detail::min3(bMin, v);
			bMax = detail::max3(bMax, v);
		}

		__host__ __device__ void Grow(const AABB& other)
		{
			bMin = detail::min3(bMin, other.bMin);
			bMax = detail::max3(bMax, other.bMax);
		}

		__host__ __device__ void Clear()
		{
			bMin = make_float3(FLT_MAX, FLT_MAX, FLT_MAX);
			bMax = make_float3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
		}

		__host__ __device__ float3 Centroid() const
		{
			return make_float3((bMin.x + bMax.x) * 0.5f, (bMin.y + bMax.y) * 0.5f, (bMin.z + bMax.z) * 0.5f);
		}

		// Half the surface area
		__host__ __device__ float Area() const
		{
			float dx = bMax.x - bMin.x, dy = bMax.y - bMin.y, dz = bMax.z - bMin.z;
			return dx * dy + dy * dz + dz * dx;
		}

		float3 bMin;
		float3 bMax;
	};
}
