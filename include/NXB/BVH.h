// NXB/BVH.h - drop-in for NexusBVH/include/NXB/BVH.h:1-97: BVH2 (32-byte
// nodes, leaves 0..n-1 in input order, root 2n-2) and the compressed BVH8
// (80-byte nodes, root 0). Layouts are asserted against include/nxb.h.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../nxb.h"
#include "AABB.h"

#define INVALID_IDX (uint32_t)(-1)

namespace NXB
{
	enum struct PrimType : unsigned char
	{
		AABB,
		TRIANGLE
	};

	struct BVH2
	{
		struct Node
		{
			AABB bounds;

			// leftChild = INVALID_IDX if leaf node
			uint32_t leftChild;

			// rightChild = primIdx if leaf node
			uint32_t rightChild;
		};

		Node* nodes;
		uint32_t nodeCount;
		uint32_t primCount;

		// Root bounds
		AABB bounds;
	};

	// Compressed wide BVH (Ylitie et al. 2017)
	struct BVH8
	{
		struct NodeExplicit
		{
			float3 p;                 // origin of the child grid
			uint8_t e[3];             // biased exponents of the grid scale
			uint8_t imask = 0;        // bit i: slot i holds an inner node
			uint32_t childBaseIdx = 0;
			uint32_t primBaseIdx = 0;
			uint8_t meta[8];
			uint8_t qlox[8], qloy[8], qloz[8];
			uint8_t qhix[8], qhiy[8], qhiz[8];
		};

		struct Node
		{
			float4 p_e_imask;
			float4 childidx_tridx_meta;
			float4 qlox_qloy;
			float4 qloz_qhix;
			float4 qhiy_qhiz;
		};

		Node* nodes;
		uint32_t nodeCount;

		uint32_t* primIdx;
		uint32_t primCount;

		// Root bounds
		AABB bounds;
	};

	static_assert(sizeof(AABB) == sizeof(nxb_aabb), "AABB layout");
	static_assert(sizeof(BVH2::Node) == sizeof(nxb_bvh2_node), "BVH2::Node layout");
	static_assert(offsetof(BVH2::Node, leftChild) == offsetof(nxb_bvh2_node, left), "BVH2::Node layout");
	static_assert(sizeof(BVH8::Node) == sizeof(nxb_bvh8_node), "BVH8::Node layout");
	static_assert(sizeof(BVH8::NodeExplicit) == sizeof(nxb_bvh8_node), "BVH8::NodeExplicit layout");
	static_assert(offsetof(BVH8::NodeExplicit, meta) == offsetof(nxb_bvh8_node, meta), "BVH8 layout");
	static_assert(offsetof(BVH8::NodeExplicit, qhiz) == offsetof(nxb_bvh8_node, qhiz), "BVH8 layout");
	static_assert(sizeof(BVH2) == sizeof(nxb_bvh2) && sizeof(BVH8) == sizeof(nxb_bvh8), "handle layout");
}
