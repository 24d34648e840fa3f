/*
 * orb::Orbor -- the reference's public C++ API (orb.h:11-76) on top of the
 * C ABI of orb_capi.h.  Same signatures, ownership and print-and-exit error
 * behaviour; one addition, setScaleFactor(), selects the bilinear pyramid
 * (default 2.0 keeps the reference's octaves).
 */
#pragma once

#include "orb_structures.h"
#include "orb_hip_utils.h"
#include "orb_capi.h"

namespace orb
{

	class Orbor
	{
	public:
		Orbor();
		~Orbor();

		/* Initialization (orb.h:18-19, orb.cpp:32-54) */
		void init(int _noctaves = 5, int _edge_threshold = 31, int _wta_k = 2, ScoreType _score_type = ScoreType::HARRIS_SCORE,
			int _patch_size = 31, int _fast_threshold = 20, int _retain_topn = -1, int _max_pts = 1000);

		/* Pyramid scale factor; call before init().  2.0 (default) is the
		   reference's nearest-neighbour octaves, a value in (1,2) the
		   bilinear levels.  No counterpart in the reference. */
		void setScaleFactor(double scale_factor);

		/* Detect keypoints and compute descriptors (orb.h:22, orb.cpp:57-101).
		   image: device pointer, row pitch iAlignUp(whp0.x, 128).
		   *desc_addr is freed and reallocated with num_pts * 32 bytes (the
		   caller frees it with hipFree); with desc_addr == NULL only the
		   orientation is computed. */
		void detectAndCompute(unsigned char* image, OrbData& result, int3 whp0, void** desc_addr = NULL, const bool compute_desc = true);

		/* Match two results (orb.h:25, orb.cpp:104-113) */
		void match(OrbData& result1, OrbData& result2, unsigned char* desc1, unsigned char* desc2);

		/* Initialize data (orb.h:28, orb.cpp:116-127) */
		void initOrbData(OrbData& data, const int max_pts, const bool host, const bool dev);

		/* Free data (orb.h:31, orb.cpp:130-142) */
		void freeOrbData(OrbData& data);

	private:
		orb_ctx* ctx = nullptr;
		orb_params prm{};
		double scale_factor = 2.0;
		int* d_count = nullptr;
		unsigned char* d_desc_work = nullptr;
		int desc_work_cap = 0;

		void check(int rc, const char* what);
	};

}
