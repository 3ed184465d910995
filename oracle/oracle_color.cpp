// Oracle for the optimize stage's colour operations: Lab statistics and
// Reinhard colour transfer (GainColorOps.cpp:28-157).
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
// calls this file; only the test suite may use it, and only as the checker.
//
// A scalar restatement in the reference's arithmetic: binary32 operations in
// source order, uncontracted (-ffp-contract=off), the C library's powf and
// cbrtf, binary64 sums in pixel order.  Parity pins
// (tests/test_oracle_color.py): equality with the compiled reference sources
// (oracle/_ref) where they build; closed forms (black and white pixels, the
// sRGB curve's two branches, neutral greys with a = b = 0 up to rounding,
// the identity transfer) and the round trip RGB -> Lab -> RGB.

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace {

// GainColorOps.cpp:30-38
float SrgbLin(float c) { return (c <= 0.04045f) ? (c / 12.92f) : std::pow((c + 0.055f) / 1.055f, 2.4f); }
float LinSrgb(float c) { return (c <= 0.0031308f) ? (c * 12.92f) : (1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f); }

// GainColorOps.cpp:40-52
constexpr float kD = 6.0f / 29.0f;
float Fwd(float t) { return (t > kD * kD * kD) ? std::cbrt(t) : (t / (3.0f * kD * kD) + 4.0f / 29.0f); }
float Inv(float t) { return (t > kD) ? (t * t * t) : (3.0f * kD * kD * (t - 4.0f / 29.0f)); }

// GainColorOps.cpp:58-75
void ToLab(const uint16_t* p, float* lab) {
    const float r = SrgbLin(p[0] / 65535.0f), g = SrgbLin(p[1] / 65535.0f), b = SrgbLin(p[2] / 65535.0f);
    const float X = 0.4124564f * r + 0.3575761f * g + 0.1804375f * b;
    const float Y = 0.2126729f * r + 0.7151522f * g + 0.0721750f * b;
    const float Z = 0.0193339f * r + 0.1191920f * g + 0.9503041f * b;
    const float fx = Fwd(X / 0.95047f), fy = Fwd(Y / 1.0f), fz = Fwd(Z / 1.08883f);
    lab[0] = 116.0f * fy - 16.0f;
    lab[1] = 500.0f * (fx - fy);
    lab[2] = 200.0f * (fy - fz);
}

uint16_t Sat(float c) {
    const float v = c * 65535.0f;
    const float lo = (v < 0.0f) ? 0.0f : v;
    return static_cast<uint16_t>((65535.0f < lo) ? 65535.0f : lo);
}

// GainColorOps.cpp:77-102
void ToRgb(const float* lab, uint16_t* p) {
    const float fy = (lab[0] + 16.0f) / 116.0f, fx = fy + lab[1] / 500.0f, fz = fy - lab[2] / 200.0f;
    const float X = 0.95047f * Inv(fx), Y = 1.0f * Inv(fy), Z = 1.08883f * Inv(fz);
    const float rl = 3.2404542f * X - 1.5371385f * Y - 0.4985314f * Z;
    const float gl = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
    const float bl = 0.0556434f * X - 0.2040259f * Y + 1.0572252f * Z;
    p[0] = Sat(LinSrgb(rl < 0.0f ? 0.0f : rl));
    p[1] = Sat(LinSrgb(gl < 0.0f ? 0.0f : gl));
    p[2] = Sat(LinSrgb(bl < 0.0f ? 0.0f : bl));
}

}  // namespace

extern "C" {

void orc_rgb_to_lab(const uint16_t* rgb, long long n, float* lab) {
    for (long long i = 0; i < n; ++i) ToLab(rgb + 3 * i, lab + 3 * i);
}

void orc_lab_to_rgb(const float* lab, long long n, uint16_t* rgb) {
    for (long long i = 0; i < n; ++i) ToRgb(lab + 3 * i, rgb + 3 * i);
}

// GainColorOps.cpp:103-130
void orc_lab_stats(const uint16_t* rgb, int w, int h, double* mean, double* sd) {
    const long long n = static_cast<long long>(w) * h;
    double s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
    for (long long i = 0; i < n; ++i) {
        float lab[3];
        ToLab(rgb + 3 * i, lab);
        for (int c = 0; c < 3; ++c) {
            const double v = lab[c];
            s[c] += v;
            q[c] += v * v;
        }
    }
    for (int c = 0; c < 3; ++c) {
        const double m = s[c] / n, var = q[c] / n - m * m;
        mean[c] = m;
        sd[c] = std::sqrt(var > 0.0 ? var : 0.0);
    }
}

// GainColorOps.cpp:132-157
void orc_reinhard(uint16_t* rgb, int w, int h, const double* sm, const double* ss, const double* rm, const double* rs) {
    const long long n = static_cast<long long>(w) * h;
    for (long long i = 0; i < n; ++i) {
        float lab[3];
        ToLab(rgb + 3 * i, lab);
        for (int c = 0; c < 3; ++c) {
            const double st = (ss[c] > 1e-6) ? ss[c] : 1.0;
            lab[c] = static_cast<float>((static_cast<double>(lab[c]) - sm[c]) * (rs[c] / st) + rm[c]);
        }
        ToRgb(lab, rgb + 3 * i);
    }
}

}  // extern "C"
