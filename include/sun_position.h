/* sun_position.h -- hourly sun azimuth/altitude for a non-leap year (2026)
 * by the PSA algorithm, FP64 internally, FP32 out.  Same type and class
 * interface as the SolarGPU reference's include/sun_position.h:15,19-40
 * (SunSample, SunPositionCalculator::calc), implemented by libsolargpu
 * (sgpu_sun_positions). */
#pragma once

#include <tuple>
#include <vector>

#include "sgpu.h"

namespace solar_gpu {

using SunSample = std::tuple<float, float>;  // (azimuth_deg, altitude_deg)

class SunPositionCalculator {
public:
    int timezone = 0;

    SunPositionCalculator(float lon, float lat, int tz = 0) : timezone(tz), lon_(lon), lat_(lat) {}

    /* Appends 8760 hourly samples (day-major, local standard time hours 0..23). */
    void calc(std::vector<SunSample>& out) {
        std::vector<float> buf(2 * SGPU_HOURS_PER_YEAR);
        sgpu_sun_positions(lon_, lat_, timezone, buf.data());
        out.reserve(out.size() + SGPU_HOURS_PER_YEAR);
        for (int k = 0; k < SGPU_HOURS_PER_YEAR; ++k) out.emplace_back(buf[2 * k], buf[2 * k + 1]);
    }

    static bool is_leap_year(unsigned year) {
        return (year % 400 == 0) || (year % 4 == 0 && year % 100 != 0);
    }

private:
    float lon_ = 0.0f, lat_ = 0.0f;
};

}  // namespace solar_gpu
