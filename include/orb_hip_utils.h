/*
 * HIP-native counterparts of the helpers users of the reference's orb.h
 * reach through cuda_utils.h (main.cpp:170,200,202,209-216): CHECK /
 * CheckMsg (print + exit, cuda_utils.h:10-37), initDevice (:41-67),
 * GpuTimer (:81-108) and iAlignUp (:170-173).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#define H_PI 1.5707963267948966f

#define CHECK(err) __check(err, __FILE__, __LINE__)
#define CheckMsg(msg) __checkMsg(msg, __FILE__, __LINE__)

/* Print a HIP runtime error and exit (cuda_utils.h:18-25). */
inline void __check(hipError_t err, const char* file, const int line)
{
	if (err != hipSuccess)
	{
		fprintf(stderr, "CHECK() Runtime API error in file <%s>, line %i : %s.\n", file, line, hipGetErrorString(err));
		exit(-1);
	}
}

/* Print the last HIP error with a message and exit (cuda_utils.h:29-37). */
inline void __checkMsg(const char* msg, const char* file, const int line)
{
	hipError_t err = hipGetLastError();
	if (err != hipSuccess)
	{
		fprintf(stderr, "CheckMsg() HIP error: %s in file <%s>, line %i : %s.\n", msg, file, line, hipGetErrorString(err));
		exit(-1);
	}
}

/* Select device `dev` (clamped to the available range) and print its name
   (cuda_utils.h:41-67). */
inline bool initDevice(int dev)
{
	int device_count = 0;
	CHECK(hipGetDeviceCount(&device_count));
	if (device_count == 0)
	{
		fprintf(stderr, "HIP error: no devices.\n");
		return false;
	}
	dev = std::max<int>(0, std::min<int>(dev, device_count - 1));
	hipDeviceProp_t prop;
	CHECK(hipGetDeviceProperties(&prop, dev));
	CHECK(hipSetDevice(dev));
	int driver_version = 0, runtime_version = 0;
	CHECK(hipDriverGetVersion(&driver_version));
	CHECK(hipRuntimeGetVersion(&runtime_version));
	fprintf(stderr, "Using Device %d: %s (%s), driver %d, runtime %d\n", dev, prop.name, prop.gcnArchName,
		driver_version, runtime_version);
	return true;
}

/* Event timer on a stream, milliseconds since construction (cuda_utils.h:81-108). */
class GpuTimer
{
public:
	GpuTimer(hipStream_t stream_ = 0) : stream(stream_)
	{
		(void)hipEventCreate(&start);
		(void)hipEventCreate(&stop);
		(void)hipEventRecord(start, stream);
	}

	~GpuTimer()
	{
		(void)hipEventDestroy(start);
		(void)hipEventDestroy(stop);
	}

	float read()
	{
		(void)hipEventRecord(stop, stream);
		(void)hipEventSynchronize(stop);
		float time = 0.f;
		(void)hipEventElapsedTime(&time, start, stop);
		return time;
	}

private:
	hipEvent_t start, stop;
	hipStream_t stream;
};

/* Round a up to a multiple of b (cuda_utils.h:170-173). */
inline int iAlignUp(const int a, const int b)
{
	return (a % b != 0) ? (a - a % b + b) : a;
}
