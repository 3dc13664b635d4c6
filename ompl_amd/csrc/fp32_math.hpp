// Project-owned FP32 transcendentals for the device.  Written from published
// methods, evaluated without FMA contraction (the library is compiled with
// -ffp-contract=off), so a CPU evaluation of the same expressions in the same
// order gives the same bits:
//   sincos: Cody-Waite three-part reduction by pi/2 (leading 12 bits of pi/2,
//           next 12 bits, FP32 remainder) + Taylor polynomials of degree 9 (sin)
//           and 10 (cos) on [-pi/4, pi/4], Horner form.  Valid for |x| < 6000.
//   acos:   Abramowitz & Stegun 4.4.46 (Hastings), |err| <= 2e-8 on [0, 1]:
//           acos(x) = sqrt(1 - x) * (a0 + a1 x + ... + a7 x^7).
#pragma once

#include <hip/hip_runtime.h>

namespace ompl_amd {

struct Fp32Consts {
    static constexpr float inv_pio2 = 0.636619772367581343f;
    static constexpr float pio2_1 = 1.5703125f;
    static constexpr float pio2_2 = 4.837512969970703125e-4f;
    static constexpr float pio2_3 = 7.549790126404332e-08f;
    static constexpr float s1 = -1.66666666666666667e-1f;
    static constexpr float s2 = 8.33333333333333333e-3f;
    static constexpr float s3 = -1.98412698412698413e-4f;
    static constexpr float s4 = 2.75573192239858907e-6f;
    static constexpr float c1 = 4.16666666666666667e-2f;
    static constexpr float c2 = -1.38888888888888889e-3f;
    static constexpr float c3 = 2.48015873015873016e-5f;
    static constexpr float c4 = -2.75573192239858907e-7f;
};

__device__ __forceinline__ void sincos_pl(float x, float &s, float &c)
{
    using K = Fp32Consts;
    float kf = __builtin_rintf(x * K::inv_pio2);
    int k = (int)kf;
    float r = x - kf * K::pio2_1;
    r = r - kf * K::pio2_2;
    r = r - kf * K::pio2_3;
    float z = r * r;
    float ps = r + (r * z) * (K::s1 + z * (K::s2 + z * (K::s3 + z * K::s4)));
    float pc = 1.0f + z * (-0.5f + z * (K::c1 + z * (K::c2 + z * (K::c3 + z * K::c4))));
    // quadrant selection without divergence
    int qd = k & 3;
    float a = (qd & 1) ? pc : ps;  // |sin|
    float b = (qd & 1) ? ps : pc;  // |cos|
    s = (qd & 2) ? -a : a;
    c = (qd == 1 || qd == 2) ? -b : b;
}

__device__ __forceinline__ float acos_as(float x)
{
    float p = -0.0012624911f;
    p = p * x + 0.0066700901f;
    p = p * x + -0.0170881256f;
    p = p * x + 0.0308918810f;
    p = p * x + -0.0501743046f;
    p = p * x + 0.0889789874f;
    p = p * x + -0.2145988016f;
    p = p * x + 1.5707963050f;
    return __builtin_sqrtf(1.0f - x) * p;
}

} // namespace ompl_amd
