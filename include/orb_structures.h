/*
 * Keypoint and result layout of the ORB path, byte-identical to the
 * reference's orb_structures.h:8-35 (OrbPoint 28 B AoS, OrbData, ScoreType).
 */
#pragma once

#ifdef __cplusplus
#include <cstddef>
namespace orb
{
#endif

	/* One keypoint (orb_structures.h:8-17).  x, y are level-0 pixel
	   coordinates, octave the pyramid level it was detected on. */
	struct OrbPoint
	{
		int x;
		int y;
		int octave;
		float score;
		float angle;
		int match;
		int distance;
	};

	/* Detection result (orb_structures.h:20-26): host and device arrays of
	   OrbPoint with num_pts valid entries. */
	struct OrbData
	{
		int num_pts;
		struct OrbPoint* h_data;
		struct OrbPoint* d_data;
	};

	/* Score used to rank corners (orb_structures.h:29-33). */
	enum ScoreType
	{
		HARRIS_SCORE = 0,
		FAST_SCORE = 1
	};

#ifdef __cplusplus
	static_assert(sizeof(OrbPoint) == 28, "OrbPoint must stay 28 bytes");
	static_assert(offsetof(OrbPoint, x) == 0 && offsetof(OrbPoint, y) == 4 &&
		offsetof(OrbPoint, octave) == 8 && offsetof(OrbPoint, score) == 12 &&
		offsetof(OrbPoint, angle) == 16 && offsetof(OrbPoint, match) == 20 &&
		offsetof(OrbPoint, distance) == 24, "OrbPoint field offsets");
}
#endif
