/*
 * Demo and timing program over the drop-in C++ API (include/orb.h): the
 * counterpart of the reference's cudaOrbDemo2 (main.cpp:141-271).  Reads two
 * PGM (P5) images, or makes a seeded synthetic pair, runs the reference's
 * protocol -- N x detectAndCompute of both images, then N x match, timed with
 * GpuTimer -- prints the same report, and writes the last results to
 * <outdir>/{points1,points2}.bin (OrbPoint arrays) and {desc1,desc2}.bin.
 * With --draw 1 it also writes the pictures the reference's demo draws with
 * OpenCV (drawKeypoints, drawMatches: main.cpp:27-99) as binary PPM files:
 * keypoints1.ppm, keypoints2.ppm (a circle per keypoint) and matches.ppm (the
 * two images side by side, a line per accepted match).
 *
 *   orb_demo <device> <img1.pgm | synth:W:H:SEED> <img2.pgm | warp> <outdir>
 *            [--scale S] [--levels N] [--wta K] [--max-pts N] [--repeats N] [--draw 0|1]
 */
#include <cmath>
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

/* An RGB picture for the keypoint and match drawings. */
struct Rgb {
	int w = 0, h = 0;
	std::vector<unsigned char> px;
	Rgb(int w_, int h_) : w(w_), h(h_), px((size_t)w_ * h_ * 3, 0) {}
	void gray(const unsigned char* img, int iw, int ih, int x0, int y0)
	{
		for (int y = 0; y < ih; y++)
			for (int x = 0; x < iw; x++)
				for (int c = 0; c < 3; c++) px[((size_t)(y + y0) * w + x + x0) * 3 + c] = img[(size_t)y * iw + x];
	}
	void dot(int x, int y, const unsigned char* col)
	{
		if (x < 0 || y < 0 || x >= w || y >= h) return;
		for (int c = 0; c < 3; c++) px[((size_t)y * w + x) * 3 + c] = col[c];
	}
	/* midpoint circle outline */
	void circle(int cx, int cy, int r, const unsigned char* col)
	{
		for (int x = r, y = 0, e = 1 - r; x >= y; y++) {
			const int pts[8][2] = {{x, y}, {y, x}, {-y, x}, {-x, y}, {-x, -y}, {-y, -x}, {y, -x}, {x, -y}};
			for (auto& q : pts) dot(cx + q[0], cy + q[1], col);
			if (e < 0) e += 2 * y + 3;
			else { e += 2 * (y - x) + 5; x--; }
		}
	}
	/* Bresenham segment */
	void line(int x0, int y0, int x1, int y1, const unsigned char* col)
	{
		const int dx = abs(x1 - x0), dy = -abs(y1 - y0), sx = x0 < x1 ? 1 : -1, sy = y0 < y1 ? 1 : -1;
		for (int e = dx + dy;;) {
			dot(x0, y0, col);
			if (x0 == x1 && y0 == y1) break;
			const int e2 = 2 * e;
			if (e2 >= dy) { e += dy; x0 += sx; }
			if (e2 <= dx) { e += dx; y0 += sy; }
		}
	}
	void save(const std::string& path) const
	{
		FILE* f = fopen(path.c_str(), "wb");
		if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(-1); }
		fprintf(f, "P6\n%d %d\n255\n", w, h);
		fwrite(px.data(), 1, px.size(), f);
		fclose(f);
	}
};

/* a bright colour per item from a fixed sequence (the reference picks rand() colours) */
void color_of(unsigned i, unsigned char* col)
{
	unsigned v = i * 2654435761u + 0x9e3779b9u;
	for (int c = 0; c < 3; c++, v = v * 1103515245u + 12345u) col[c] = (unsigned char)(64 + (v >> 24) % 192);
}

/* main.cpp:27-38: a circle of radius max(1, min(5, log10(score))) per keypoint */
void draw_keypoints(const std::vector<unsigned char>& img, int w, int h, const orb::OrbData& d, const std::string& path)
{
	Rgb pic(w, h);
	pic.gray(img.data(), w, h, 0, 0);
	for (int i = 0; i < d.num_pts; i++) {
		const orb::OrbPoint& p = d.h_data[i];
		const float lg = p.score > 0.f ? log10f(p.score) : 1.f;
		const int r = lg < 1.f ? 1 : (lg > 5.f ? 5 : (int)lg);
		unsigned char col[3];
		color_of((unsigned)i, col);
		pic.circle(p.x, p.y, r, col);
	}
	pic.save(path);
}

/* main.cpp:41-99 (horizontal layout): a line per accepted match */
void draw_matches(const std::vector<unsigned char>& img1, int w1, int h1, const std::vector<unsigned char>& img2,
                  int w2, int h2, const orb::OrbData& a, const orb::OrbData& b, const std::string& path)
{
	Rgb pic(w1 + w2, h1 > h2 ? h1 : h2);
	pic.gray(img1.data(), w1, h1, 0, 0);
	pic.gray(img2.data(), w2, h2, w1, 0);
	for (int i = 0; i < a.num_pts; i++) {
		const int k = a.h_data[i].match;
		if (k < 0 || k >= b.num_pts) continue;
		unsigned char col[3];
		color_of((unsigned)i, col);
		pic.line(a.h_data[i].x, a.h_data[i].y, b.h_data[k].x + w1, b.h_data[k].y, col);
	}
	pic.save(path);
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
	int noctaves = 5, wta_k = 4, max_pts = 10000, repeats = 100, draw = 0;
	for (int i = 5; i + 1 < argc; i += 2)
	{
		if (!strcmp(argv[i], "--scale")) scale = atof(argv[i + 1]);
		else if (!strcmp(argv[i], "--levels")) noctaves = atoi(argv[i + 1]);
		else if (!strcmp(argv[i], "--wta")) wta_k = atoi(argv[i + 1]);
		else if (!strcmp(argv[i], "--max-pts")) max_pts = atoi(argv[i + 1]);
		else if (!strcmp(argv[i], "--repeats")) repeats = atoi(argv[i + 1]);
		else if (!strcmp(argv[i], "--draw")) draw = atoi(argv[i + 1]);
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

	if (draw)
	{
		draw_keypoints(host1, w1, h1, d1, out + "/keypoints1.ppm");
		draw_keypoints(host2, w2, h2, d2, out + "/keypoints2.ppm");
		draw_matches(host1, w1, h1, host2, w2, h2, d1, d2, out + "/matches.ppm");
	}

	detector.freeOrbData(d1);
	detector.freeOrbData(d2);
	if (desc1) CHECK(hipFree(desc1));
	if (desc2) CHECK(hipFree(desc2));
	CHECK(hipFree(img1));
	CHECK(hipFree(img2));
	return 0;
}
