/*
 * Demo and timing program over the drop-in C++ API (include/orb.h): the
 * counterpart of the reference's cudaOrbDemo2 (main.cpp:141-271).  Reads two
 * PGM (P5) images, or makes a seeded synthetic pair, runs the reference's
 * protocol -- N x detectAndCompute of both images, then N x match, timed with
 * GpuTimer -- prints the same report, and writes the last results to
 * <outdir>/{points1,points2}.bin (OrbPoint arrays) and {desc1,desc2}.bin.
 *
 *   orb_demo <device> <img1.pgm | synth:W:H:SEED> <img2.pgm | warp> <outdir>
 *            [--scale S] [--levels N] [--wta K] [--max-pts N] [--repeats N]
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "orb.h"

namespace
{

bool read_pgm(const char* path, std::vector<unsigned char>& img, int& w, int& h)
{
	FILE* f = fopen(path, "rb");
	if (!f) return false;
	char magic[3] = {0};
	int maxval = 0;
	if (fscanf(f, "%2s %d %d %d", magic, &w, &h, &maxval) != 4 || strcmp(magic, "P5") != 0 || maxval != 255)
	{
		fclose(f);
		return false;
	}
	fgetc(f);
	img.resize((size_t)w * h);
	const bool ok = fread(img.data(), 1, img.size(), f) == img.size();
	fclose(f);
	return ok;
}

void write_file(const std::string& path, const void* data, size_t bytes)
{
	FILE* f = fopen(path.c_str(), "wb");
	if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(-1); }
	fwrite(data, 1, bytes, f);
	fclose(f);
}

} // namespace

int main(int argc, char** argv)
{
	if (argc < 5)
	{
		fprintf(stderr, "usage: %s <device> <img1.pgm|synth:W:H:SEED> <img2.pgm|warp> <outdir> "
			"[--scale S] [--levels N] [--wta K] [--max-pts N] [--repeats N]\n", argv[0]);
		return 1;
	}
	/* the reference's demo parameters (main.cpp:176-197) unless overridden */
	double scale = 2.0;
	int noctaves = 5, wta_k = 4, max_pts = 10000, repeats = 100;
	for (int i = 5; i + 1 < argc; i += 2)
	{
		if (!strcmp(argv[i], "--scale")) scale = atof(argv[i + 1]);
		else if (!strcmp(argv[i], "--levels")) noctaves = atoi(argv[i + 1]);
		else if (!strcmp(argv[i], "--wta")) wta_k = atoi(argv[i + 1]);
		else if (!strcmp(argv[i], "--max-pts")) max_pts = atoi(argv[i + 1]);
		else if (!strcmp(argv[i], "--repeats")) repeats = atoi(argv[i + 1]);
	}
	printf("===== Registration by MI355X ORB =====\n");
	if (!initDevice(atoi(argv[1]))) return 1;

	int w1 = 0, h1 = 0, w2 = 0, h2 = 0;
	std::vector<unsigned char> host1, host2;
	if (!strncmp(argv[2], "synth:", 6))
	{
		int sw = 0, sh = 0;
		unsigned long long seed = 0;
		if (sscanf(argv[2] + 6, "%d:%d:%llu", &sw, &sh, &seed) != 3) { fprintf(stderr, "bad synth spec\n"); return 1; }
		w1 = w2 = sw;
		h1 = h2 = sh;
		host1.resize((size_t)sw * sh);
		host2.resize((size_t)sw * sh);
		const int nshapes = (int)(40.0 * sw * sh / 307200.0);
		orb_synth_image(seed, sw, sh, sw, nshapes < 1 ? 1 : nshapes, host1.data());
		orb_synth_warp(seed, host1.data(), sw, sh, sw, host2.data(), NULL);
	}
	else if (!read_pgm(argv[2], host1, w1, h1) || !read_pgm(argv[3], host2, w2, h2))
	{
		fprintf(stderr, "cannot read %s / %s (binary PGM, maxval 255)\n", argv[2], argv[3]);
		return 1;
	}
	printf("Image size = (%d,%d)\n", w1, h1);

	/* upload with row pitch iAlignUp(w,128), as main.cpp:200-216 does */
	GpuTimer timer(0);
	int3 whp1, whp2;
	whp1.x = w1; whp1.y = h1; whp1.z = iAlignUp(w1, 128);
	whp2.x = w2; whp2.y = h2; whp2.z = iAlignUp(w2, 128);
	unsigned char *img1 = NULL, *img2 = NULL;
	CHECK(hipMalloc((void**)&img1, (size_t)whp1.y * whp1.z));
	CHECK(hipMalloc((void**)&img2, (size_t)whp2.y * whp2.z));
	CHECK(hipMemset(img1, 0, (size_t)whp1.y * whp1.z));
	CHECK(hipMemset(img2, 0, (size_t)whp2.y * whp2.z));
	CHECK(hipMemcpy2D(img1, whp1.z, host1.data(), w1, w1, h1, hipMemcpyHostToDevice));
	CHECK(hipMemcpy2D(img2, whp2.z, host2.data(), w2, w2, h2, hipMemcpyHostToDevice));
	const float t0 = timer.read();

	orb::Orbor detector;
	detector.setScaleFactor(scale);
	detector.init(noctaves, 31, wta_k, orb::HARRIS_SCORE, 31, 20, 0, max_pts);
	orb::OrbData d1, d2;
	detector.initOrbData(d1, max_pts, true, true);
	detector.initOrbData(d2, max_pts, true, true);
	unsigned char* desc1 = NULL;
	unsigned char* desc2 = NULL;

	const float t1 = timer.read();
	for (int i = 0; i < repeats; i++)
	{
		detector.detectAndCompute(img1, d1, whp1, (void**)&desc1, true);
		detector.detectAndCompute(img2, d2, whp2, (void**)&desc2, true);
	}
	const float t2 = timer.read();
	for (int i = 0; i < repeats; i++)
	{
		detector.match(d1, d2, desc1, desc2);
	}
	const float t3 = timer.read();

	printf("Number of features1: %d\nNumber of features2: %d\n", d1.num_pts, d2.num_pts);
	printf("Time for allocating image memory:  %g\n", t0);
	printf("Time for allocating point memory:  %g\n", t1 - t0);
	printf("Time of detection and computation: %g\n", (t2 - t1) / repeats);
	printf("Time of matching ORB keypoints:   %g\n", (t3 - t2) / repeats);

	/* full OrbPoint records (device) and descriptors for checking */
	const std::string out = argv[4];
	std::vector<orb::OrbPoint> p1(d1.num_pts > 0 ? d1.num_pts : 1), p2(d2.num_pts > 0 ? d2.num_pts : 1);
	std::vector<unsigned char> q1((size_t)(d1.num_pts > 0 ? d1.num_pts : 1) * 32), q2((size_t)(d2.num_pts > 0 ? d2.num_pts : 1) * 32);
	if (d1.num_pts > 0)
	{
		CHECK(hipMemcpy(p1.data(), d1.d_data, sizeof(orb::OrbPoint) * d1.num_pts, hipMemcpyDeviceToHost));
		CHECK(hipMemcpy(q1.data(), desc1, (size_t)d1.num_pts * 32, hipMemcpyDeviceToHost));
	}
	if (d2.num_pts > 0)
	{
		CHECK(hipMemcpy(p2.data(), d2.d_data, sizeof(orb::OrbPoint) * d2.num_pts, hipMemcpyDeviceToHost));
		CHECK(hipMemcpy(q2.data(), desc2, (size_t)d2.num_pts * 32, hipMemcpyDeviceToHost));
	}
	write_file(out + "/points1.bin", p1.data(), sizeof(orb::OrbPoint) * d1.num_pts);
	write_file(out + "/points2.bin", p2.data(), sizeof(orb::OrbPoint) * d2.num_pts);
	write_file(out + "/desc1.bin", q1.data(), (size_t)d1.num_pts * 32);
	write_file(out + "/desc2.bin", q2.data(), (size_t)d2.num_pts * 32);
	/* the host copies detectAndCompute / match made must agree with the device records */
	for (int i = 0; i < d1.num_pts; i++)
	{
		if (d1.h_data[i].x != p1[i].x || d1.h_data[i].y != p1[i].y || d1.h_data[i].match != p1[i].match)
		{
			fprintf(stderr, "host copy mismatch at %d\n", i);
			return 1;
		}
	}

	detector.freeOrbData(d1);
	detector.freeOrbData(d2);
	if (desc1) CHECK(hipFree(desc1));
	if (desc2) CHECK(hipFree(desc2));
	CHECK(hipFree(img1));
	CHECK(hipFree(img2));
	return 0;
}
