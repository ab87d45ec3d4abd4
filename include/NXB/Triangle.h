// NXB/Triangle.h - drop-in for NexusBVH/include/NXB/Triangle.h:10-48 (36 bytes).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "AABB.h"

namespace NXB
{
	struct Triangle
	{
		Triangle() = default;

		__host__ __device__ Triangle(float3 pos0, float3 pos1, float3 pos2)
			: v0(pos0), v1(pos1), v2(pos2) { }

		// Same value as the reference (the vertex sum halved, Triangle.h:17-20).
		__host__ __device__ float3 Centroid()
		{
			return make_float3((v0.x + v1.x + v2.x) * 0.5f, (v0.y + v1.y + v2.y) * 0.5f, (v0.z + v1.z + v2.z) * 0.5f);
		}

		__host__ __device__ AABB Bounds()
		{
			return AABB(v0, v1, v2);
		}

		// Normal (not normalized)
		__host__ __device__ float3 Normal() const
		{
			float3 e0 = make_float3(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z);
			float3 e1 = make_float3(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
			return make_float3(e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x);
		}

		__host__ __device__ float Area() const
		{
			float3 n = Normal();
			return 0.5f * sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
		}

		float3 v0, v1, v2;
	};
}
