/*
 * Host helpers that users of the reference's orb.h reach through its utility
 * header (cuda_utils.h:10-108,170-173; used by main.cpp:170,200,202,209-216):
 * CHECK / CheckMsg, initDevice, GpuTimer and iAlignUp.  The names and call
 * signatures are the reference's, so its demo code compiles unchanged against
 * HIP; the implementations are this library's.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define H_PI 1.5707963267948966f

namespace orb_util
{

/* Report a failed HIP call with the source position and terminate, the error
   behaviour the reference's API promises its callers. */
[[noreturn]] inline void die(const char* what, hipError_t err, const char* file, int line)
{
    std::fprintf(stderr, "orb: %s failed at %s:%d: %s (%s)\n", what, file, line, hipGetErrorName(err),
                 hipGetErrorString(err));
    std::exit(-1);
}

inline void check_call(hipError_t err, const char* expr, const char* file, int line)
{
    if (err == hipSuccess) return;
    die(expr, err, file, line);
}

inline void check_last(const char* msg, const char* file, int line)
{
    const hipError_t err = hipGetLastError();
    if (err == hipSuccess) return;
    die(msg, err, file, line);
}

} // namespace orb_util

/* CHECK(call): exit with a message when a HIP runtime call fails. */
#define CHECK(err) orb_util::check_call((err), #err, __FILE__, __LINE__)
/* CheckMsg(msg): exit with `msg` when a launch or call left an error pending. */
#define CheckMsg(msg) orb_util::check_last((msg), __FILE__, __LINE__)

/* Make device `dev` current, clamped into [0, device count), and print which
   device that is; false when no HIP device exists. */
inline bool initDevice(int dev)
{
    int count = 0;
    CHECK(hipGetDeviceCount(&count));
    if (count < 1) {
        std::fprintf(stderr, "orb: no HIP device found\n");
        return false;
    }
    const int chosen = dev < 0 ? 0 : (dev >= count ? count - 1 : dev);
    CHECK(hipSetDevice(chosen));
    hipDeviceProp_t props;
    CHECK(hipGetDeviceProperties(&props, chosen));
    int rt = 0;
    CHECK(hipRuntimeGetVersion(&rt));
    std::fprintf(stderr, "orb: device %d of %d, %s [%s], %d CUs, HIP runtime %d\n", chosen, count, props.name,
                 props.gcnArchName, props.multiProcessorCount, rt);
    return true;
}

/* Milliseconds of stream work since construction: read() marks the end on the
   stream, waits for it and returns the time between the two marks. */
class GpuTimer
{
public:
    explicit GpuTimer(hipStream_t stream_ = 0) : stream(stream_)
    {
        CHECK(hipEventCreate(&marks[0]));
        CHECK(hipEventCreate(&marks[1]));
        CHECK(hipEventRecord(marks[0], stream));
    }
    GpuTimer(const GpuTimer&) = delete;
    GpuTimer& operator=(const GpuTimer&) = delete;
    ~GpuTimer()
    {
        (void)hipEventDestroy(marks[1]);
        (void)hipEventDestroy(marks[0]);
    }

    float read()
    {
        CHECK(hipEventRecord(marks[1], stream));
        CHECK(hipEventSynchronize(marks[1]));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, marks[0], marks[1]));
        return ms;
    }

private:
    hipEvent_t marks[2] = {nullptr, nullptr};
    hipStream_t stream;
};

/* The smallest multiple of b at or above a (for a >= 0; a negative a moves
   toward zero, as the reference's helper does). */
inline int iAlignUp(const int a, const int b)
{
    const int r = a % b;
    return r == 0 ? a : a + (b - r);
}
