/* Seeded synthetic inputs: an urban DSM (fBm terrain + box buildings +
 * Gaussian trees) and a TMY-like DNI/DHI year (clear-sky model scaled by a
 * 3-state cloud Markov chain).  Deterministic for a given seed on any host:
 * integer PRNG (splitmix64) and IEEE double arithmetic only. */
#include <algorithm>
#include <cmath>
#include <vector>

#include "sgpu_internal.h"

namespace sgpu {
namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
};

double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

/* Value-noise fBm on a lattice; 5 octaves, base wavelength `base` cells. */
std::vector<double> fbm(Rng& rng, int w, int h, double base) {
    std::vector<double> acc((size_t)w * h, 0.0);
    double amp = 1.0, wl = base;
    for (int oct = 0; oct < 5; ++oct, amp *= 0.5, wl *= 0.5) {
        if (wl < 2.0) break;
        const int lw = (int)std::ceil(w / wl) + 2, lh = (int)std::ceil(h / wl) + 2;
        std::vector<double> lat((size_t)lw * lh);
        for (double& v : lat) v = rng.uni(-1.0, 1.0);
        for (int y = 0; y < h; ++y) {
            const double fy = y / wl;
            const int y0 = (int)fy;
            const double ty = smooth(fy - y0);
            for (int x = 0; x < w; ++x) {
                const double fx = x / wl;
                const int x0 = (int)fx;
                const double tx = smooth(fx - x0);
                const double a = lat[(size_t)y0 * lw + x0], b = lat[(size_t)y0 * lw + x0 + 1];
                const double c = lat[(size_t)(y0 + 1) * lw + x0], d = lat[(size_t)(y0 + 1) * lw + x0 + 1];
                acc[(size_t)y * w + x] += amp * ((a + (b - a) * tx) + ((c + (d - c) * tx) - (a + (b - a) * tx)) * ty);
            }
        }
    }
    return acc;
}

}  // namespace

/* Ground 200-220 m; buildings until `density` of the cells are covered (footprint
 * sides U(8,30) m, height U(5,40) m above the ground at the footprint centre);
 * `trees` Gaussian crowns of 4-12 m on open ground. */
void synth_dsm(uint64_t seed, int w, int h, float cell_m, float density, int trees, float* out) {
    Rng rng(seed);
    std::vector<double> g = fbm(rng, w, h, 256.0 / cell_m);
    double lo = g[0], hi = g[0];
    for (double v : g) { lo = std::min(lo, v); hi = std::max(hi, v); }
    const double span = (hi > lo) ? hi - lo : 1.0;
    std::vector<double> z((size_t)w * h);
    for (size_t i = 0; i < z.size(); ++i) z[i] = 200.0 + 20.0 * (g[i] - lo) / span;
    const std::vector<double> ground = z;
    std::vector<char> bldg((size_t)w * h, 0);
    const int64_t target = (int64_t)((double)density * w * h);
    int64_t covered = 0;
    const int64_t max_tries = 64 + (int64_t)w * h;
    for (int64_t tries = 0; covered < target && tries < max_tries; ++tries) {
        const int sx = std::max(1, (int)std::lround(rng.uni(8.0, 30.0) / cell_m));
        const int sy = std::max(1, (int)std::lround(rng.uni(8.0, 30.0) / cell_m));
        const int x0 = (int)(rng.uni() * w) - sx / 2, y0 = (int)(rng.uni() * h) - sy / 2;
        const double height = rng.uni(5.0, 40.0);
        const int cx = std::min(std::max(x0 + sx / 2, 0), w - 1), cy = std::min(std::max(y0 + sy / 2, 0), h - 1);
        const double roof = ground[(size_t)cy * w + cx] + height;
        for (int y = std::max(y0, 0); y < std::min(y0 + sy, h); ++y)
            for (int x = std::max(x0, 0); x < std::min(x0 + sx, w); ++x) {
                size_t i = (size_t)y * w + x;
                if (!bldg[i]) { bldg[i] = 1; ++covered; z[i] = roof; }
                else z[i] = std::max(z[i], roof);
            }
    }
    for (int t = 0; t < trees; ++t) {
        const double tx = rng.uni() * w, ty = rng.uni() * h;
        const double th = rng.uni(4.0, 12.0), sig = rng.uni(1.5, 3.5) / cell_m;
        const int r = (int)std::ceil(3.0 * sig);
        for (int y = std::max((int)ty - r, 0); y <= std::min((int)ty + r, h - 1); ++y)
            for (int x = std::max((int)tx - r, 0); x <= std::min((int)tx + r, w - 1); ++x) {
                size_t i = (size_t)y * w + x;
                if (bldg[i]) continue;
                const double d2 = (x - tx) * (x - tx) + (y - ty) * (y - ty);
                z[i] += th * std::exp(-d2 / (2.0 * sig * sig));
            }
    }
    for (size_t i = 0; i < z.size(); ++i) out[i] = (float)z[i];
}

/* Clear sky: Kasten-Young air mass, Meinel beam transmittance, a diffuse share of
 * extraterrestrial-on-horizontal; clouds: 3-state hourly Markov chain
 * (clear / broken / overcast).  DHI >= 1 W/m^2 whenever the sun is >= 1 deg up. */
void synth_weather(uint64_t seed, const float* az_alt, float* dni, float* dhi) {
    Rng rng(seed ^ 0x5DEECE66Dull);
    static const double P[3][3] = {{0.86, 0.11, 0.03}, {0.18, 0.64, 0.18}, {0.04, 0.14, 0.82}};
    int state = 0;
    for (int k = 0; k < kHours; ++k) {
        const double u = rng.uni();
        state = (u < P[state][0]) ? 0 : (u < P[state][0] + P[state][1]) ? 1 : 2;
        const double alt = az_alt[2 * k + 1];
        if (alt - 1e-8 < 1.0) { dni[k] = 0.0f; dhi[k] = 0.0f; continue; }
        const double pi = 3.14159265358979323846;
        const double cz = std::sin(alt * pi / 180.0);
        const double am = 1.0 / (cz + 0.50572 * std::pow(6.07995 + alt, -1.6364));
        const double beam = 1361.0 * std::pow(0.7, std::pow(am, 0.678));
        const double diff = std::max(1.0, 1361.0 * cz * (0.02 + 0.06 * std::sqrt(am) / (1.0 + 0.1 * am)));
        double fb, fd;
        if (state == 0) { fb = rng.uni(0.92, 1.0); fd = rng.uni(0.95, 1.1); }
        else if (state == 1) { fb = rng.uni(0.25, 0.7); fd = rng.uni(1.4, 2.2); }
        else { fb = rng.uni(0.0, 0.08); fd = rng.uni(0.9, 1.5); }
        dni[k] = (float)(beam * fb);
        dhi[k] = (float)std::max(1.0, diff * fd);
    }
}

}  // namespace sgpu
