/*
 * Seeded synthetic test images for the ORB path (host code, no GPU).
 *
 * Image A: a uniform background, `nshapes` filled rotated rectangles or
 * ellipses with intensity U{0..255} (later shapes on top), plus 3-octave value
 * noise (lattice spacing 32/16/8 px, weights 4/7, 2/7, 1/7 of amplitude 30),
 * rounded half to even and clipped to uint8.
 * Image B: A resampled (bilinear, border clamped) through the inverse of a
 * similarity H = T(c + t) R(theta) S(s) T(-c), theta ~ U[-15, 15] deg,
 * s ~ U[0.9, 1.1], t ~ U[-20, 20]^2, c the image centre, plus N(0, 3^2) noise
 * (Box-Muller), rounded half to even and clipped.
 * All draws come from SplitMix64 in double arithmetic, so the bytes are the
 * same on every x86-64 host.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "orb_capi.h"

namespace
{

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal()
    {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

uint8_t to_u8(double v)
{
    v = std::nearbyint(v);
    if (v < 0) v = 0;
    if (v > 255) v = 255;
    return (uint8_t)v;
}

} // namespace

extern "C" int orb_synth_image(uint64_t seed, int w, int h, int pitch, int nshapes, uint8_t* out)
{
    if (w <= 0 || h <= 0 || pitch < w || nshapes < 0 || !out) return ORB_E_ARG;
    Rng rng(seed * 0x2545F4914F6CDD1Dull + 1);
    std::vector<double> img((size_t)w * h, rng.uniform(0.0, 255.0));
    const double ms = (double)(w < h ? w : h);
    for (int i = 0; i < nshapes; i++) {
        const bool ellipse = rng.uniform() < 0.5;
        const double cx = rng.uniform(0.0, w), cy = rng.uniform(0.0, h);
        const double a = rng.uniform(0.03, 0.2) * ms, b = rng.uniform(0.03, 0.2) * ms;
        const double th = rng.uniform(0.0, 3.141592653589793);
        const double val = (double)(rng.next() % 256);
        const double c = std::cos(th), s = std::sin(th);
        const double rad = a > b ? a : b;
        const int x0 = (int)std::floor(cx - rad) - 1, x1 = (int)std::ceil(cx + rad) + 1;
        const int y0 = (int)std::floor(cy - rad) - 1, y1 = (int)std::ceil(cy + rad) + 1;
        for (int y = y0 < 0 ? 0 : y0; y <= y1 && y < h; y++) {
            for (int x = x0 < 0 ? 0 : x0; x <= x1 && x < w; x++) {
                const double dx = x - cx, dy = y - cy;
                const double u = dx * c + dy * s, v = -dx * s + dy * c;
                const bool in = ellipse ? (u * u) / (a * a) + (v * v) / (b * b) <= 1.0
                                        : std::fabs(u) <= a && std::fabs(v) <= b;
                if (in) img[(size_t)y * w + x] = val;
            }
        }
    }
    const int spacing[3] = {32, 16, 8};
    const double weight[3] = {4.0 / 7.0, 2.0 / 7.0, 1.0 / 7.0};
    for (int o = 0; o < 3; o++) {
        const int sp = spacing[o];
        const int gw = w / sp + 2, gh = h / sp + 2;
        std::vector<double> lat((size_t)gw * gh);
        for (auto& v : lat) v = rng.uniform(-1.0, 1.0);
        const double amp = 30.0 * weight[o];
        for (int y = 0; y < h; y++) {
            const int gy = y / sp;
            const double fy = (double)(y - gy * sp) / sp;
            for (int x = 0; x < w; x++) {
                const int gx = x / sp;
                const double fx = (double)(x - gx * sp) / sp;
                const double v00 = lat[(size_t)gy * gw + gx], v01 = lat[(size_t)gy * gw + gx + 1];
                const double v10 = lat[(size_t)(gy + 1) * gw + gx], v11 = lat[(size_t)(gy + 1) * gw + gx + 1];
                const double v = (1 - fy) * ((1 - fx) * v00 + fx * v01) + fy * ((1 - fx) * v10 + fx * v11);
                img[(size_t)y * w + x] += amp * v;
            }
        }
    }
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) out[(size_t)y * pitch + x] = to_u8(img[(size_t)y * w + x]);
        if (pitch > w) memset(out + (size_t)y * pitch + w, 0, (size_t)(pitch - w));
    }
    return ORB_OK;
}

extern "C" int orb_synth_warp(uint64_t seed, const uint8_t* src, int w, int h, int pitch, uint8_t* out, double* Hout)
{
    if (w <= 0 || h <= 0 || pitch < w || !src || !out) return ORB_E_ARG;
    Rng rng(seed * 0x9E3779B97F4A7C15ull + 7);
    const double th = rng.uniform(-15.0, 15.0) * 3.141592653589793 / 180.0;
    const double sc = rng.uniform(0.9, 1.1);
    const double tx = rng.uniform(-20.0, 20.0), ty = rng.uniform(-20.0, 20.0);
    const double cx = 0.5 * (w - 1), cy = 0.5 * (h - 1);
    const double c = std::cos(th) * sc, s = std::sin(th) * sc;
    /* H: p_B = R S (p_A - c) + c + t */
    const double H[9] = {c, -s, cx + tx - c * cx + s * cy, s, c, cy + ty - s * cx - c * cy, 0, 0, 1};
    if (Hout) memcpy(Hout, H, sizeof(H));
    /* inverse similarity: p_A = (R S)^-1 (p_B - c - t) + c */
    const double det = c * c + s * s;
    const double ic = c / det, is = -s / det;
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const double bx = x - cx - tx, by = y - cy - ty;
            double sx = ic * bx - is * by + cx, sy = is * bx + ic * by + cy;
            if (sx < 0) sx = 0;
            if (sy < 0) sy = 0;
            if (sx > w - 1) sx = w - 1;
            if (sy > h - 1) sy = h - 1;
            const int x0 = (int)std::floor(sx), y0 = (int)std::floor(sy);
            const int x1 = x0 + 1 < w ? x0 + 1 : w - 1, y1 = y0 + 1 < h ? y0 + 1 : h - 1;
            const double fx = sx - x0, fy = sy - y0;
            const double v = (1 - fy) * ((1 - fx) * src[(size_t)y0 * pitch + x0] + fx * src[(size_t)y0 * pitch + x1]) +
                             fy * ((1 - fx) * src[(size_t)y1 * pitch + x0] + fx * src[(size_t)y1 * pitch + x1]);
            out[(size_t)y * pitch + x] = to_u8(v + 3.0 * rng.normal());
        }
        if (pitch > w) memset(out + (size_t)y * pitch + w, 0, (size_t)(pitch - w));
    }
    return ORB_OK;
}
