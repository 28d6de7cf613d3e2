// Photon sources (include/mlgpu.h, mlgpu_source), shared by the HIP kernel
// and host builds of it.  Draws come from Philox4x32-10 counter blocks
// (block, 1, id_lo, id_hi) with key (seed_lo, seed_hi):
//   block 0      direction: two 53-bit uniforms (words 0-1, 2-3)
//   block 1      wavelength: one 53-bit uniform (words 0-1)
//   block 2 + k  disc try k (k < 64): x from words 0-1, y from words 2-3
// Only +, -, *, / and sqrt are used, all correctly rounded in IEEE double
// on both sides and never contracted (the library is built with
// -ffp-contract=off), so every photon is reproducible bit for bit.
#pragma once

#include <cstdint>
#include <cstring>

#include "propagate_core.hpp"

namespace mlgpu {

constexpr uint32_t kSourceStream = 1u;
constexpr int kDiscTries = 64;

MLGPU_HD inline void source_block(uint64_t seed, uint64_t id, uint32_t block, uint32_t w[4]) {
    philox4x32_10(block, kSourceStream, static_cast<uint32_t>(id), static_cast<uint32_t>(id >> 32),
                  static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), w[0], w[1], w[2], w[3]);
}

// uniform on [0, 1) with 53 random bits
MLGPU_HD inline double u53(uint32_t a, uint32_t b) {
    return (static_cast<double>(a >> 5) * 67108864.0 + static_cast<double>(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Natural logarithm of a normal x > 0: x = m 2^e with m in [sqrt(1/2),
// sqrt(2)), log m = 2 atanh(z), z = (m - 1) / (m + 1), |z| <= 0.1716, and
// atanh(z) / z = sum z^2k / (2k + 1) to k = 12 (error < 1e-17).
MLGPU_HD inline double log_series(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = static_cast<int>((b >> 52) & 0x7ffu) - 1022;  // x = m 2^e, m in [0.5, 1)
    b = (b & 0x800fffffffffffffull) | (1022ull << 52);
    double m;
    memcpy(&m, &b, 8);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    const double z = (m - 1.0) / (m + 1.0);
    const double s = z * z;
    double p = 1.0 / 25.0;
    p = p * s + 1.0 / 23.0;
    p = p * s + 1.0 / 21.0;
    p = p * s + 1.0 / 19.0;
    p = p * s + 1.0 / 17.0;
    p = p * s + 1.0 / 15.0;
    p = p * s + 1.0 / 13.0;
    p = p * s + 1.0 / 11.0;
    p = p * s + 1.0 / 9.0;
    p = p * s + 1.0 / 7.0;
    p = p * s + 1.0 / 5.0;
    p = p * s + 1.0 / 3.0;
    p = p * s + 1.0;
    const double ln2_hi = 6.93147180559945286227e-01, ln2_lo = 2.31904681384629955842e-17;
    const double de = static_cast<double>(e);
    return de * ln2_hi + (de * ln2_lo + 2.0 * z * p);
}

// sin and cos of t in [0, pi/2] by their Taylor series to t^25 (error < 1e-19)
MLGPU_HD inline void sincos_series(double t, double& sn, double& cs) {
    const double t2 = t * t;
    // sin t = t (1 - t^2/(2.3) (1 - t^2/(4.5) (1 - ...))), cos likewise
    double ps = 1.0, pc = 1.0;
    for (int k = 12; k >= 1; k--) {
        const double a = static_cast<double>(2 * k) * static_cast<double>(2 * k + 1);
        const double c = static_cast<double>(2 * k - 1) * static_cast<double>(2 * k);
        ps = 1.0 - t2 / a * ps;
        pc = 1.0 - t2 / c * pc;
    }
    sn = t * ps;
    cs = pc;
}

// sin and cos of 2 pi u for u in [0, 1)
MLGPU_HD inline void sincos_turn(double u, double& sn, double& cs) {
    const double u4 = u * 4.0;
    const int q = static_cast<int>(u4);  // quadrant
    const double f = u4 - static_cast<double>(q);
    double s, c;
    sincos_series(f * 1.57079632679489661923, s, c);
    if (q == 0) { sn = s; cs = c; }
    else if (q == 1) { sn = c; cs = -s; }
    else if (q == 2) { sn = -s; cs = -c; }
    else { sn = -c; cs = s; }
}

// Photon id of the source before the pose: the support (x, y) on the xy
// plane, the direction d about +z, the wavelength
MLGPU_HD inline void source_local(const mlgpu_source& src, const uint64_t seed, const uint64_t id, double& x,
                                  double& y, D3& d, double& lambda) {
    uint32_t w[4];
    // direction about +z
    source_block(seed, id, 0u, w);
    const double ua = u53(w[0], w[1]), ub = u53(w[2], w[3]);
    switch (src.direction) {
    case MLGPU_DIRECTION_TILT: {
        double st, ct, sp, cp;
        sincos_series(src.angle * ua, st, ct);
        sincos_turn(ub, sp, cp);
        d = D3{st * cp, st * sp, ct};
        break;
    }
    case MLGPU_DIRECTION_GAUSSIAN: {
        // Box-Muller: radius from 1 - ua in (0, 1], angle 2 pi ub
        const double r = sqrt(-2.0 * log_series(1.0 - ua));
        double sp, cp;
        sincos_turn(ub, sp, cp);
        const double tx = src.angle * (r * cp), ty = src.angle * (r * sp);
        const double nrm = sqrt(tx * tx + ty * ty + 1.0);
        d = D3{tx / nrm, ty / nrm, 1.0 / nrm};
        break;
    }
    case MLGPU_DIRECTION_CAP: {
        // ZenithDistancePicker (merlict/random/SpherePointPicker.cpp:8-20) with
        // cos(zenith) uniform on [cos(angle), 1]
        double sa, ca;
        sincos_series(src.angle, sa, ca);
        const double cz = 1.0 - ua * (1.0 - ca);
        const double sz = sqrt(1.0 - cz * cz);
        double sp, cp;
        sincos_turn(ub, sp, cp);
        d = D3{sz * cp, sz * sp, cz};
        break;
    }
    case MLGPU_DIRECTION_RING: {
        // zenith angle + angle_width z, z standard normal (Box-Muller), folded
        // at 0 (the validation bounds it below pi/2); azimuth from the
        // wavelength block's second pair
        const double r = sqrt(-2.0 * log_series(1.0 - ua));
        double sb, cb;
        sincos_turn(ub, sb, cb);
        double zen = src.angle + src.angle_width * (r * cb);
        if (zen < 0.0) zen = -zen;
        if (zen > 1.57079632679489661923) zen = 1.57079632679489661923;
        double st, ct;
        sincos_series(zen, st, ct);
        source_block(seed, id, 1u, w);
        double sp, cp;
        sincos_turn(u53(w[2], w[3]), sp, cp);
        d = D3{st * cp, st * sp, ct};
        break;
    }
    default: d = D3{0.0, 0.0, 1.0};
    }
    // wavelength: pdf ~ lambda^-2 means 1 / lambda uniform
    source_block(seed, id, 1u, w);
    if (src.lambda_min == src.lambda_max) {
        lambda = src.lambda_min;
    } else if (src.spectrum == MLGPU_SPECTRUM_UNIFORM) {
        lambda = src.lambda_min + u53(w[0], w[1]) * (src.lambda_max - src.lambda_min);
    } else {
        const double a = 1.0 / src.lambda_min, b = 1.0 / src.lambda_max;
        lambda = 1.0 / (a - u53(w[0], w[1]) * (a - b));
    }
    // support: Generator::get_point_on_xy_disc_within_radius, with a
    // deterministic cap of kDiscTries (all failing: p < 1e-42) at the centre
    x = 0.0;
    y = 0.0;
    if (src.support == MLGPU_SUPPORT_DISC) {
        const double r = src.radius, r2 = r * r;
        for (int k = 0; k < kDiscTries; k++) {
            source_block(seed, id, 2u + static_cast<uint32_t>(k), w);
            const double tx = r * (u53(w[0], w[1]) * 2.0 - 1.0);
            const double ty = r * (u53(w[2], w[3]) * 2.0 - 1.0);
            if (!(tx * tx + ty * ty > r2)) {
                x = tx;
                y = ty;
                break;
            }
        }
    }
}

// Photon id of the source (support s, direction d, wavelength): the local
// photon started start_distance back along d, then posed
MLGPU_HD inline void source_photon(const mlgpu_source& src, const uint64_t seed, const uint64_t id, D3& s, D3& d,
                                   double& lambda) {
    double x, y;
    source_local(src, seed, id, x, y, d, lambda);
    const double g = src.start_distance;
    const D3 p{x - d.x * g, y - d.y * g, 0.0 - d.z * g};
    const double* R = src.rot;
    s = D3{p.x * R[0] + p.y * R[1] + p.z * R[2] + src.pos[0], p.x * R[3] + p.y * R[4] + p.z * R[5] + src.pos[1],
           p.x * R[6] + p.y * R[7] + p.z * R[8] + src.pos[2]};
    d = D3{d.x * R[0] + d.y * R[1] + d.z * R[2], d.x * R[3] + d.y * R[4] + d.z * R[5],
           d.x * R[6] + d.y * R[7] + d.z * R[8]};
}

}  // namespace mlgpu
