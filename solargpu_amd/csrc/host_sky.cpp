/* Host-side sky tables: PSA sun positions, sun-visited sky bins, the 2592 ray
 * directions and the per-hour constants of the solar kernel.  FP64 stays in
 * the sun position only; everything else is FP32 in the reference's
 * expression order.  Built with -ffp-contract=off. */
#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "sgpu_internal.h"

namespace sgpu {

void fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Failure(code, buf);
}

void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e == hipSuccess) return;
    int code = (e == hipErrorOutOfMemory) ? SGPU_ERR_OOM : SGPU_ERR_HIP;
    fail(code, "%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
}

/* PSA algorithm (Blanco-Muriel et al. 2001), as used by the reference's
 * SunPositionCalculator (sun_position.h:24-40 date walk, :58-107 formulae).
 * Year 2026, local standard time hour -> UT via the fixed zone offset. */
static void psa_position(float lon, float lat, int tz, int month, int day, int hour, float& az, float& alt) {
    const int year = 2026;
    const double pi = 3.14159265358979323846, rad = pi / 180.0;
    const double ut = hour - tz + 0 / 60.0;
    const long m14 = (month - 14) / 12;
    const long jdn = (1461 * (year + 4800 + m14)) / 4 + (367 * (month - 2 - 12 * m14)) / 12 -
                     (3 * ((year + 4900 + m14) / 100)) / 4 + day - 32075;
    const double n = (double)jdn - 0.5 + ut / 24.0 - 2451545.0;   // days since J2000.0

    const double omega = 2.1429 - 0.0010394594 * n;
    const double L = 4.8950630 + 0.017202791698 * n;
    const double g = 6.2400600 + 0.0172019699 * n;
    const double lambda = L + 0.03341607 * std::sin(g) + 0.00034894 * std::sin(2 * g) - 0.0001134 -
                          0.0000203 * std::sin(omega);
    const double eps = 0.4090928 - 6.2140e-9 * n + 0.0000396 * std::cos(omega);

    const double sl = std::sin(lambda);
    double ra = std::atan2(std::cos(eps) * sl, std::cos(lambda));
    if (ra < 0.0) ra += 2.0 * pi;
    const double dec = std::asin(std::sin(eps) * sl);

    const double gmst = 6.6974243242 + 0.0657098283 * n + ut;
    const double lmst = (gmst * 15.0 + lon) * rad;
    const double w = lmst - ra;
    const double phi = lat * rad;
    const double cphi = std::cos(phi), sphi = std::sin(phi), cw = std::cos(w);
    double zen = std::acos(cphi * cw * std::cos(dec) + std::sin(dec) * sphi);
    double azi = std::atan2(-std::sin(w), std::tan(dec) * cphi - sphi * cw);
    if (azi < 0.0) azi += 2.0 * pi;
    azi /= rad;
    zen = (zen + (6371.01 / 149597890.0) * std::sin(zen)) / rad;   // parallax
    alt = (float)(90.0 - zen);
    az = (float)azi;
}

void sun_positions(float lon, float lat, int tz, float* az_alt) {
    static const int kDaysIn[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    int k = 0;
    for (int month = 1; month <= 12; ++month)
        for (int day = 1; day <= kDaysIn[month - 1]; ++day)
            for (int hour = 0; hour < 24; ++hour, ++k)
                psa_position(lon, lat, tz, month, day, hour, az_alt[2 * k], az_alt[2 * k + 1]);
}

/* sky_bins.cpp:9-52: nudge, mark sun-visited 5x5 deg upper bins, dense-number them
 * azimuth-major, index every hour's sun bin. */
SkyTables sky_bins(const float* az_alt) {
    SkyTables t;
    t.nudged.resize(2 * kHours);
    t.full_to_dense.assign(kAzBins * kElBins, -1);
    t.sun_dense.assign(kHours, -1);
    std::vector<int> bin_of(kHours, -1);
    std::vector<char> used(kAzBins * kElBins, 0);
    for (int k = 0; k < kHours; ++k) {
        const float az = az_alt[2 * k] - 1e-8f, alt = az_alt[2 * k + 1] - 1e-8f;
        t.nudged[2 * k] = az;
        t.nudged[2 * k + 1] = alt;
        if (alt < 1.0f) continue;
        const unsigned bi = (unsigned)(az / kBinDeg), bj = (unsigned)(alt / kBinDeg);
        if (bi < (unsigned)kAzBins && bj < (unsigned)kElBins) {
            bin_of[k] = (int)(bi * kElBins + bj);
            used[bin_of[k]] = 1;
        }
    }
    for (int b = 0; b < kAzBins * kElBins; ++b)
        if (used[b]) t.full_to_dense[b] = t.active++;
    for (int k = 0; k < kHours; ++k)
        if (bin_of[k] >= 0) t.sun_dense[k] = t.full_to_dense[bin_of[k]];
    t.svf_bytes = (int)std::ceil((float)t.active / 8.0f);
    return t;
}

/* viewshed_stage.cuh:64-80: azimuth-major; per azimuth 18 upward then 18 downward
 * elevations at bin centres; host cosf/sinf, then normalised. */
void directions(const SkyTables& sky, DirConst* out) {
    int d = 0;
    for (unsigned i = 0; i < (unsigned)kAzBins; ++i) {
        const float az = ((i + 0.5f) * kBinDeg) * kDegToRad;
        for (unsigned j = 0; j < 2u * kElBins; ++j, ++d) {
            const bool upper = j < (unsigned)kElBins;
            const float e = ((upper ? j : j - kElBins) + 0.5f) * kBinDeg;
            float cx = cosf(e * kDegToRad) * sinf(az);
            float cy = -cosf(e * kDegToRad) * cosf(az);
            float cz = upper ? sinf(e * kDegToRad) : -sinf(e * kDegToRad);
            const float len = sqrtf(cx * cx + cy * cy + cz * cz);
            if (len > 0.0f) { cx /= len; cy /= len; cz /= len; }
            DirConst& c = out[d];
            c.cx = cx; c.cy = cy; c.cz = cz;
            c.ax = fmaxf(fabsf(cx), 1e-8f);
            c.ay = fmaxf(fabsf(cy), 1e-8f);
            c.rx = 1.0f / c.ax;
            c.ry = 1.0f / c.ay;
            const int dense = upper ? sky.full_to_dense[i * kElBins + j] : -1;
            c.info = (upper ? 1 : 0) | ((dense + 1) << 8);
        }
    }
}

/* Perez (1990) coefficient table, F[row][eps bin] (device_globals.cu:50-57). */
static const float kPerez[6][8] = {
    {-0.0083117f, 0.1299457f, 0.3296958f, 0.5682053f, 0.8730280f, 1.1326077f, 1.0601591f, 0.6777470f},
    {0.5877285f, 0.6825954f, 0.4868735f, 0.1874525f, -0.3920403f, -1.2367284f, -1.5999137f, -0.3272588f},
    {-0.0620636f, -0.1513752f, -0.2210958f, -0.2951290f, -0.3616149f, -0.4118494f, -0.3589221f, -0.2504286f},
    {-0.0596012f, -0.0189325f, 0.0554140f, 0.1088631f, 0.2255647f, 0.2877813f, 0.2642124f, 0.1561313f},
    {0.0721249f, 0.0659650f, -0.0639588f, -0.1519229f, -0.4620442f, -0.8230357f, -1.1272340f, -1.3765031f},
    {-0.0220216f, -0.0288748f, -0.0260542f, -0.0139754f, 0.0012448f, 0.0558651f, 0.1310694f, 0.2506212f},
};

/* The hour-only part of Perez() and calc_kernel_solar (solar_kernel.cu:8-50, 62-72),
 * evaluated once per sunlit hour in the reference's FP32 operation order. */
std::vector<HourConst> hour_constants(const SkyTables& sky, const float* dni, const float* dhi) {
    std::vector<HourConst> hc;
    const float pi_f = (float)3.14159265358979323846;
    for (int day = 0; day < 365; ++day) {
        for (int hour = 0; hour < 24; ++hour) {
            const int idx = day * 24 + hour;
            const float az = sky.nudged[2 * idx], alt = sky.nudged[2 * idx + 1];
            if (alt < 1.0f) continue;   // gpu_irradiance_uncompressed.cu:117
            HourConst c{};
            const float dir = dni[idx], dif = dhi[idx];
            c.cos_z = cosf((90.0f - alt) * kDegToRad);
            const float za = acosf(c.cos_z);
            c.sin_z = sinf(za);
            const float azk = az * kDegToRad;
            c.cos_az = cosf(azk);
            c.sin_az = sinf(azk);
            c.dni_eff = fmaxf(dir, 1e-3f);
            c.dhi = dif;
            const float z_deg = za * (180.0f / pi_f);
            const float am = 1.0f / (c.cos_z + 0.50572f * powf(96.07995f - z_deg, -1.6364f));
            c.inv_b = 1.0f / fmaxf(0.08715574f, c.cos_z);
            float eps = ((dif + dir / cosf(za)) / dif + 1.041f * (za * za * za)) / (1.0f + 1.041f * za * za * za);
            const float B = 2 * pi_f * (day / 365.0f);
            const float e0 = 1367 * (1.00011f + 0.034221f * cosf(B) + 0.00128f * sinf(B) + 0.000719f * cosf(2 * B) +
                                     0.000077f * sinf(2 * B));
            const float delta = (am * dif) / e0;
            eps = fmaxf(1.0f, eps);
            static const float kEdges[7] = {1.065f, 1.230f, 1.500f, 1.950f, 2.800f, 4.500f, 6.200f};
            int bin = 7;
            for (int q = 0; q < 7; ++q)
                if (eps < kEdges[q]) { bin = q; break; }
            c.f1 = fmaxf(0.0f, kPerez[0][bin] + kPerez[1][bin] * delta + kPerez[2][bin] * za);
            c.f2 = kPerez[3][bin] + kPerez[4][bin] * delta + kPerez[5][bin] * za;
            c.dense = sky.sun_dense[idx];
            c.index = idx;
            c.eps_bin = bin;
            hc.push_back(c);
        }
    }
    return hc;
}

}  // namespace sgpu
