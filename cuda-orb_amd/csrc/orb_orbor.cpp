/*
 * orb::Orbor over the C ABI: the reference's class (orb.cpp:14-201) with its
 * ownership rules and exit-on-error behaviour, sequencing the asynchronous
 * library calls and doing the host copies the reference does.
 */
#include "orb.h"

#include <cstdio>
#include <cstdlib>

namespace orb
{

	Orbor::Orbor()
	{
	}

	Orbor::~Orbor()
	{
		if (d_count) (void)hipFree(d_count);
		if (d_desc_work) (void)hipFree(d_desc_work);
		if (ctx) orb_ctx_destroy(ctx);
	}

	void Orbor::check(int rc, const char* what)
	{
		if (rc != ORB_OK)
		{
			fprintf(stderr, "ORB error in %s: %s\n", what, orb_last_error(ctx));
			exit(-1);
		}
	}

	void Orbor::setScaleFactor(double s)
	{
		scale_factor = s;
	}

	void Orbor::init(int _noctaves, int _edge_threshold, int _wta_k, ScoreType _score_type,
		int _patch_size, int _fast_threshold, int _retain_topn, int _max_pts)
	{
		if (!ctx)
		{
			int dev = 0;
			CHECK(hipGetDevice(&dev));
			check(orb_ctx_create(&ctx, dev), "orb_ctx_create");
			CHECK(hipMalloc((void**)&d_count, sizeof(int)));
		}
		prm.noctaves = _noctaves;
		prm.edge_threshold = _edge_threshold;
		prm.wta_k = _wta_k;
		prm.score_type = _score_type;
		prm.patch_size = _patch_size;
		prm.fast_threshold = _fast_threshold;
		prm.retain_topn = _retain_topn;
		prm.max_pts = _max_pts;
		prm.scale_factor = scale_factor;
		check(orb_init(ctx, &prm), "init");
		if (desc_work_cap < _max_pts)
		{
			if (d_desc_work) CHECK(hipFree(d_desc_work));
			CHECK(hipMalloc((void**)&d_desc_work, (size_t)_max_pts * 32));
			desc_work_cap = _max_pts;
		}
	}

	void Orbor::detectAndCompute(unsigned char* image, OrbData& result, int3 whp0, void** desc_addr, const bool compute_desc)
	{
		if (!ctx)
		{
			fprintf(stderr, "ORB error: detectAndCompute before init\n");
			exit(-1);
		}
		/* the image pitch is iAlignUp(width, 128) (orb.cpp:162, orbd.cu:1584),
		   which the reference assumes without reading whp0.z; a nonzero whp0.z
		   that disagrees would make it read the image wrongly, so it is an error */
		const int pitch = iAlignUp(whp0.x, 128);
		if (whp0.z != 0 && whp0.z != pitch)
		{
			fprintf(stderr, "ORB error: detectAndCompute: whp0.z = %d, but the image pitch must be iAlignUp(%d, 128) = %d\n",
				whp0.z, whp0.x, pitch);
			exit(-1);
		}
		const int want_desc = compute_desc && desc_addr != NULL;
		check(orb_detect_and_compute(ctx, image, whp0.x, whp0.y, pitch, result.d_data, want_desc ? d_desc_work : NULL,
			d_count, compute_desc ? 1 : 0), "detectAndCompute");
		int n = 0;
		check(orb_sync(ctx), "detectAndCompute");
		CHECK(hipMemcpy(&n, d_count, sizeof(int), hipMemcpyDeviceToHost));
		result.num_pts = n < prm.max_pts ? n : prm.max_pts;

		if (want_desc && result.num_pts > 0)
		{
			/* orb.cpp:84-90: the descriptor buffer is reallocated per call */
			unsigned char* desc = (unsigned char*)(*desc_addr);
			if (desc)
			{
				CHECK(hipFree(desc));
			}
			CHECK(hipMalloc(desc_addr, (size_t)result.num_pts * 32));
			CHECK(hipMemcpy(*desc_addr, d_desc_work, (size_t)result.num_pts * 32, hipMemcpyDeviceToDevice));
		}

		/* orb.cpp:95-100: x, y, octave, score, angle to the host */
		if (result.h_data != NULL && result.num_pts > 0)
		{
			CHECK(hipMemcpy2D(&result.h_data[0].x, sizeof(OrbPoint), &result.d_data[0].x, sizeof(OrbPoint),
				3 * sizeof(int) + 2 * sizeof(float), result.num_pts, hipMemcpyDeviceToHost));
		}
	}

	void Orbor::match(OrbData& result1, OrbData& result2, unsigned char* desc1, unsigned char* desc2)
	{
		if (result1.num_pts <= 0) return; /* the reference launches an empty grid (orbd.cu:1789) */
		check(orb_match(ctx, result1.d_data, desc1, result1.num_pts, desc2, result2.num_pts), "match");
		check(orb_sync(ctx), "match");
		if (result1.h_data)
		{
			CHECK(hipMemcpy2D(&result1.h_data[0].match, sizeof(OrbPoint), &result1.d_data[0].match, sizeof(OrbPoint),
				2 * sizeof(int), result1.num_pts, hipMemcpyDeviceToHost));
		}
	}

	void Orbor::initOrbData(OrbData& data, const int max_pts, const bool host, const bool dev)
	{
		data.num_pts = 0;
		const size_t size = sizeof(OrbPoint) * max_pts;
		data.h_data = host ? (OrbPoint*)malloc(size) : NULL;
		data.d_data = NULL;
		if (dev)
		{
			CHECK(hipMalloc((void**)&data.d_data, size));
		}
	}

	void Orbor::freeOrbData(OrbData& data)
	{
		if (data.d_data != NULL)
		{
			CHECK(hipFree(data.d_data));
		}
		if (data.h_data != NULL)
		{
			free(data.h_data);
		}
		data.num_pts = 0;
	}

}
