// NXB/BuildConfig.h - drop-in for NexusBVH/include/NXB/BuildConfig.h:6-12,
// extended with the options of this build (defaults = reference behaviour).
#pragma once

#include <cstdint>

#include "BVHBuildMetrics.h"

namespace NXB
{
	struct BuildConfig
	{
		// 64-bit Morton keys unless prioritizeSpeed (then 32-bit keys)
		bool prioritizeSpeed = false;

		uint32_t mortonBits = 0;        // 0: follow prioritizeSpeed; else 32 or 64
		uint32_t searchRadius = 8;      // H-PLOC nearest-neighbour radius, 1..16
		uint32_t mergeMode = NXB_MERGE_WAVE32;   // or NXB_MERGE_WAVE64
		uint32_t radiusPolicy = NXB_RADIUS_FIXED;
		uint32_t segmentLog2 = 12;      // radius-table segment = 2^segmentLog2 sorted primitives
		const uint8_t* radiusTable = nullptr;    // device array, one radius per segment
		uint32_t radiusTableLen = 0;
		uint32_t sortDigitBits = 0;     // 0: automatic
		uint32_t spinLimit = 0;         // 0: default watchdog
		uint32_t metricsFlags = 0;      // NXB_METRICS_TIMES_ONLY: no SAH in the metrics
		const float* mlpParams = nullptr;        // NXB_RADIUS_LEARNED: device array of NXB_MLP_PARAMS floats
		float mlpMinConfidence = 0.5f;  // below this class probability the segment uses radius 8
		uint32_t slotAssignment = NXB_SLOT_GREEDY;  // or NXB_SLOT_AUCTION
		uint32_t hplocBlocksPerCu = 0;  // H-PLOC resident blocks per CU; 0: default (12)
	};
}
