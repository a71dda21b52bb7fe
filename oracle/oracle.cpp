/* SolarGPU CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restatement of the SolarGPU reference method, written from its source as a
 * specification.  Citations are to the reference repository's files.
 * Build: g++ -O2 -std=c++17 -ffp-contract=off (no fast-math), see oracle/Makefile.
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

namespace {

constexpr float kDegToRadF = 0.0174532925f;          // pipeline_math_constants.h:15
constexpr float kAlbedoF = 0.2f;                     // pipeline_constants.h:9
constexpr double kPiD = 3.14159265358979323846;

/* Perez coefficient table F[6][8] (device_globals.cu:50-57; Perez et al. 1990). */
const float kPerezF[6][8] = {
    {-0.0083117f, 0.1299457f, 0.3296958f, 0.5682053f, 0.8730280f, 1.1326077f, 1.0601591f, 0.6777470f},
    {0.5877285f, 0.6825954f, 0.4868735f, 0.1874525f, -0.3920403f, -1.2367284f, -1.5999137f, -0.3272588f},
    {-0.0620636f, -0.1513752f, -0.2210958f, -0.2951290f, -0.3616149f, -0.4118494f, -0.3589221f, -0.2504286f},
    {-0.0596012f, -0.0189325f, 0.0554140f, 0.1088631f, 0.2255647f, 0.2877813f, 0.2642124f, 0.1561313f},
    {0.0721249f, 0.0659650f, -0.0639588f, -0.1519229f, -0.4620442f, -0.8230357f, -1.1272340f, -1.3765031f},
    {-0.0220216f, -0.0288748f, -0.0260542f, -0.0139754f, 0.0012448f, 0.0558651f, 0.1310694f, 0.2506212f},
};

int resolve_threads(int nthreads) {
    if (nthreads > 0) return nthreads;
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? (int)hc : 1;
}

/* Dynamic block scheduler over [0, n) in blocks of `block`. */
template <typename F>
void parallel_blocks(int64_t n, int64_t block, int nthreads, F&& fn) {
    nthreads = resolve_threads(nthreads);
    int64_t nblocks = (n + block - 1) / block;
    if (nthreads <= 1 || nblocks <= 1) {
        for (int64_t b = 0; b < nblocks; ++b) fn(b * block, std::min(n, (b + 1) * block), 0);
        return;
    }
    std::atomic<int64_t> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) {
        pool.emplace_back([&, t]() {
            for (;;) {
                int64_t b = next.fetch_add(1);
                if (b >= nblocks) break;
                fn(b * block, std::min(n, (b + 1) * block), t);
            }
        });
    }
    for (auto& th : pool) th.join();
}

/* ---------------- sun position: sun_position.h:24-40 (calc) and :58-107 (psa_calc) ---------- */

void month_day_2026(int day_of_year, int* month, int* mday) {
    /* sun_position.h:28-33 normalises tm_mday = dayOfYear with mktime in non-leap 2026. */
    static const int kLen[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    int m = 0, d = day_of_year;
    while (d > kLen[m]) { d -= kLen[m]; ++m; }
    *month = m + 1;
    *mday = d;
}

void psa(int year, float lon, float lat, int tz, int month, int day, int hour, float* az, float* alt) {
    const double kRad = kPiD / 180.0;
    const double kTwoPi = 2.0 * kPiD;
    double ut = hour - tz + 0 / 60.0;
    long a1 = (month - 14) / 12;
    long a2 = (1461 * (year + 4800 + a1)) / 4 + (367 * (month - 2 - 12 * a1)) / 12 -
              (3 * ((year + 4900 + a1) / 100)) / 4 + day - 32075;
    double n = (double)a2 - 0.5 + ut / 24.0 - 2451545.0;
    double omega = 2.1429 - 0.0010394594 * n;
    double mean_lon = 4.8950630 + 0.017202791698 * n;
    double anomaly = 6.2400600 + 0.0172019699 * n;
    double ecl_lon = mean_lon + 0.03341607 * std::sin(anomaly) + 0.00034894 * std::sin(2 * anomaly) -
                     0.0001134 - 0.0000203 * std::sin(omega);
    double obliq = 0.4090928 - 6.2140e-9 * n + 0.0000396 * std::cos(omega);
    double sin_el = std::sin(ecl_lon);
    double ra = std::atan2(std::cos(obliq) * sin_el, std::cos(ecl_lon));
    if (ra < 0.0) ra += kTwoPi;
    double decl = std::asin(std::sin(obliq) * sin_el);
    double gmst = 6.6974243242 + 0.0657098283 * n + ut;
    double lmst = (gmst * 15.0 + lon) * kRad;
    double ha = lmst - ra;
    double lat_r = lat * kRad;
    double cl = std::cos(lat_r), sl = std::sin(lat_r), ch = std::cos(ha);
    double zen = std::acos(cl * ch * std::cos(decl) + std::sin(decl) * sl);
    double azi = std::atan2(-std::sin(ha), std::tan(decl) * cl - sl * ch);
    if (azi < 0.0) azi += kTwoPi;
    azi /= kRad;
    double parallax = (6371.01 / 149597890.0) * std::sin(zen);
    zen = (zen + parallax) / kRad;
    *alt = (float)(90.0 - zen);
    *az = (float)azi;
}

/* ---------------- DDA: viewshed_kernels.cu:15-58 ---------------- */

struct DirConst {
    float cx, cy, cz;
    float ax, ay;     // fmaxf(fabsf(C), 1e-8f)
    float rx, ry;     // 1.0f / a
    int sx, sy;
    int upper;
};

DirConst make_dir(const float* c, int upper) {
    DirConst d;
    d.cx = c[0]; d.cy = c[1]; d.cz = c[2];
    d.ax = std::fmax(std::fabs(d.cx), 1e-8f);
    d.ay = std::fmax(std::fabs(d.cy), 1e-8f);
    d.rx = 1.0f / d.ax;
    d.ry = 1.0f / d.ay;
    d.sx = (d.cx >= 0.0f) ? 1 : -1;
    d.sy = (d.cy >= 0.0f) ? 1 : -1;
    d.upper = upper;
    return d;
}

/* Returns cell id + 1 of the first hit, 0 if none (viewshed_kernels.cu:27-58). */
inline int trace(const float* grid, int w, int h, float maxz, float gxf, float gyf, float base_z,
                 const DirConst& d, int* steps_out) {
    const float ox = gxf + d.cx * 2.0f;
    const float oy = gyf + d.cy * 2.0f;
    const float oz = base_z + d.cz * 2.0f;
    int ix = (int)std::floor(ox);
    int iy = (int)std::floor(oy);
    float tmx = ((d.cx >= 0.0f) ? ((float)ix + 1.0f - ox) : (ox - (float)ix)) / d.ax;
    float tmy = ((d.cy >= 0.0f) ? ((float)iy + 1.0f - oy) : (oy - (float)iy)) / d.ay;
    float t;
    int steps = 0;
    int result = 0;
    for (;;) {
        ++steps;
        if (tmx < tmy) { t = tmx; ix += d.sx; tmx += d.rx; }
        else { t = tmy; iy += d.sy; tmy += d.ry; }
        if ((unsigned)ix >= (unsigned)w || (unsigned)iy >= (unsigned)h) break;
        float zr = oz + d.cz * t;
        if (d.upper && zr > maxz) break;
        if (!d.upper && zr < 0.0f) break;
        if (grid[ix + iy * w] - zr > 0.01f) { result = ix + iy * w + 1; break; }
    }
    if (steps_out) *steps_out = steps;
    return result;
}

/* Small open-addressing set with generation stamps (first-occurrence dedupe). */
struct IdSet {
    std::vector<int32_t> key;
    std::vector<uint32_t> gen;
    uint32_t cur = 0;
    explicit IdSet(int cap_pow2) : key(cap_pow2), gen(cap_pow2, 0) {}
    void clear() { ++cur; }
    bool insert(int32_t id) {  // true if newly inserted
        uint32_t mask = (uint32_t)key.size() - 1;
        uint32_t hsh = ((uint32_t)id * 2654435761u) & mask;
        for (;;) {
            if (gen[hsh] != cur) { gen[hsh] = cur; key[hsh] = id; return true; }
            if (key[hsh] == id) return false;
            hsh = (hsh + 1) & mask;
        }
    }
};

/* ---------------- solar: solar_kernel.cu:8-77 ---------------- */

int eps_bin(float eps) {  // solar_kernel.cu:33-40
    int i = 7;
    if (eps < 1.065f) i = 0;
    else if (eps < 1.230f) i = 1;
    else if (eps < 1.500f) i = 2;
    else if (eps < 1.950f) i = 3;
    else if (eps < 2.800f) i = 4;
    else if (eps < 4.500f) i = 5;
    else if (eps < 6.200f) i = 6;
    return i;
}

float perez_f32(float cos_z, float cos_aoi, float dif, float dir, float slope, int n) {
    float z_deg = acosf(cos_z) * (180.0f / (float)kPiD);
    float am = 1.0f / (cos_z + 0.50572f * powf(96.07995f - z_deg, -1.6364f));
    float a = fmaxf(0.0f, cos_aoi);
    float b = fmaxf(0.08715574f, cos_z);
    float a_div_b = 0;
    if (b > 0) a_div_b = a / b;
    float za = acosf(cos_z);
    float eps = ((dif + dir / cosf(za)) / dif + 1.041f * (za * za * za)) / (1.0f + 1.041f * za * za * za);
    float bb = 2 * (float)kPiD * (n / 365.0f);
    float e0 = 1367 * (1.00011f + 0.034221f * cosf(bb) + 0.00128f * sinf(bb) + 0.000719f * cosf(2 * bb) +
                       0.000077f * sinf(2 * bb));
    float delta = (am * dif) / e0;
    eps = fmaxf(1.0f, eps);
    int i = eps_bin(eps);
    float f1 = fmaxf(0.0f, kPerezF[0][i] + kPerezF[1][i] * delta + kPerezF[2][i] * za);
    float f2 = kPerezF[3][i] + kPerezF[4][i] * delta + kPerezF[5][i] * za;
    return dif * ((1 - f1) * ((1 + cosf(slope)) / 2) + f1 * a_div_b + f2 * sinf(slope));
}

/* The FP32 epsilon bin of an hour: a discrete decision shared with the FP64 oracle. */
int hour_eps_bin(float alt, float dif, float dir) {
    float eps;
    return orc_hour_eps(alt, dif, dir, &eps);
}

inline bool svf_bit(const uint8_t* svf, int64_t v, int di, int64_t n_total) {
    return (svf[v + (int64_t)(di / 8) * n_total] >> (di % 8)) & 1;  // voxel_types.h:28-44
}

float solar_voxel(float az, float alt, int di, float dni, float dhi, int day, float slope, float aspect,
                  bool blocked) {
    float cos_z = cosf((90.0f - alt) * kDegToRadF);
    float cos_aoi = cos_z * cosf(slope) + sinf(acosf(cos_z)) * sinf(slope) * cosf(az * kDegToRadF - aspect);
    float I = 0.0f;
    if (cos_aoi > 0.0f && di >= 0 && !blocked) I += fmaxf(0.0f, fmaxf(dni, 1e-3f) * cos_aoi);
    I += fmaxf(0.0f, perez_f32(cos_z, cos_aoi, dhi, dni, slope, day));
    return I;
}

/* ---------------- reflection: reflection_kernels.cu:18-65 ---------------- */

struct Receiver { int gx, gy; float z, nx, ny, nz; };

Receiver receiver(int64_t t, int w, int n_top, const int* pp, const float* grid, const float* wall_z,
                  const float* slopes, const float* aspects) {
    Receiver r;
    int top = pp[t];
    r.gx = top % w;
    r.gy = top / w;
    r.z = (t < n_top) ? grid[r.gx + r.gy * w] : wall_z[t - n_top];
    float s = slopes[t], a = aspects[t];
    float ss = sinf(s);
    r.nx = ss * sinf(a);
    r.ny = -ss * cosf(a);
    r.nz = cosf(s);
    return r;
}

inline void add_reflection(int ngid, const Receiver& r, const float* grid, int w, int h, const float* hourly,
                           float& R) {
    if (ngid < 0 || ngid >= w * h) return;
    int ngx = ngid % w, ngy = ngid / w;
    float dx = (float)(ngx - r.gx), dy = (float)(ngy - r.gy), dz = grid[ngid] - r.z;
    float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < 1e-4f) return;
    float d = sqrtf(d2);
    float cr = (r.nx * dx + r.ny * dy + r.nz * dz) / d;
    if (cr <= 0.0f) return;
    float cs = fmaxf(fabsf(-dz) / d, 0.1f);
    R += kAlbedoF * hourly[ngy * w + ngx] * cr * cs / (3.14159265f * d2);
}

/* ---------------- ba-RLE: barle_compression.cpp:12-187 ---------------- */

struct Run { uint8_t w; int start; int len; };

uint8_t width_class(uint32_t x) {  // barle_compression.cpp:12-18
    if (x <= 15) return 0;
    if (x <= 255) return 1;
    if (x <= 65535) return 2;
    if (x <= 16777215) return 3;
    return 4;
}

std::vector<Run> runs_of(const std::vector<uint32_t>& dl) {  // :90-105
    std::vector<Run> out;
    size_t i = 0;
    while (i < dl.size()) {
        uint8_t wc = width_class(dl[i]);
        size_t j = i + 1;
        while (j < dl.size() && width_class(dl[j]) == wc) ++j;
        out.push_back({wc, (int)i, (int)(j - i)});
        i = j;
    }
    return out;
}

std::vector<Run> optimise(const std::vector<Run>& in) {  // :112-152
    if (in.size() < 2) return in;
    std::vector<Run> cur = in;
    for (;;) {
        bool merged_any = false;
        std::vector<Run> nxt;
        size_t i = 0;
        while (i + 1 < cur.size()) {
            if (cur[i + 1].len == 1 && cur[i + 1].w + 1 == cur[i].w) {
                Run m = cur[i];
                m.len += 1;
                nxt.push_back(m);
                merged_any = true;
                i += 2;
            } else {
                nxt.push_back(cur[i]);
                i += 1;
            }
        }
        if (i < cur.size()) nxt.push_back(cur[i]);
        if (nxt.size() == 1) return nxt;
        if (!merged_any) break;
        cur.swap(nxt);
    }
    std::vector<Run> out;
    for (const Run& r : cur) {
        if (!out.empty() && out.back().w == r.w) out.back().len += r.len;
        else out.push_back(r);
    }
    return out;
}

void encode_list(std::vector<uint32_t>& ids, std::vector<uint8_t>& out) {
    out.clear();
    if (ids.empty()) return;
    std::sort(ids.begin(), ids.end(), [](uint32_t a, uint32_t b) { return a > b; });
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    if (!ids.empty() && ids.back() == 0) ids.pop_back();
    if (ids.empty()) return;
    std::vector<uint32_t> dl(ids.size());
    dl[0] = ids[0];
    for (size_t i = 1; i < ids.size(); ++i) dl[i] = ids[i - 1] - ids[i];
    std::vector<Run> opt = optimise(runs_of(dl));
    std::vector<Run> runs;  // split > 255 (:29-44)
    for (const Run& r : opt) {
        int left = r.len, s = r.start;
        while (left > 255) { runs.push_back({r.w, s, 255}); s += 255; left -= 255; }
        if (left > 0) runs.push_back({r.w, s, left});
    }
    for (size_t b = 0; b < runs.size(); b += 8) {
        size_t cnt = std::min<size_t>(8, runs.size() - b);
        uint32_t bits = 0;
        for (size_t j = 0; j < 8; ++j) {
            uint32_t code = (j < cnt) ? runs[b + j].w : 7u;
            bits |= (code & 7u) << ((7 - j) * 3);
        }
        out.push_back((uint8_t)(bits >> 16));
        out.push_back((uint8_t)(bits >> 8));
        out.push_back((uint8_t)bits);
        for (size_t j = 0; j < cnt; ++j) out.push_back((uint8_t)runs[b + j].len);
        for (size_t j = 0; j < cnt; ++j) {
            const Run& r = runs[b + j];
            if (r.w == 0) {
                for (int k = 0; k < r.len; k += 2) {
                    uint32_t hi = dl[r.start + k];
                    uint32_t lo = (k + 1 < r.len) ? dl[r.start + k + 1] : 0u;
                    out.push_back((uint8_t)((hi << 4) | (lo & 0x0F)));
                }
            } else {
                int nb = r.w;  // 8/16/24/32-bit widths are 1..4 big-endian bytes
                for (int k = 0; k < r.len; ++k) {
                    uint32_t x = dl[r.start + k];
                    for (int q = nb - 1; q >= 0; --q) out.push_back((uint8_t)(x >> (8 * q)));
                }
            }
        }
    }
}

/* Decoder walk (reflection_kernels.cu:132-180); calls emit(val) for every decoded value. */
template <typename Emit>
bool decode_stream(const uint8_t* p, int64_t size, Emit&& emit) {
    int64_t i = 0;
    uint32_t prev = 0;
    bool first = true;
    while (i < size) {
        if (i + 3 > size) return false;
        uint32_t bits = ((uint32_t)p[i] << 16) | ((uint32_t)p[i + 1] << 8) | p[i + 2];
        i += 3;
        int wid[8], nc = 0;
        for (int j = 0; j < 8; ++j) {
            int c = (bits >> (21 - 3 * j)) & 7;
            if (c == 7) break;
            wid[nc++] = c;
        }
        if (i + nc > size) return false;
        int len[8];
        for (int j = 0; j < nc; ++j) len[j] = p[i++];
        for (int j = 0; j < nc; ++j) {
            if (wid[j] == 0) {
                for (int k = 0; k < len[j];) {
                    if (i >= size) return false;
                    uint8_t b = p[i++];
                    uint32_t hi = (b >> 4) & 0x0F;
                    uint32_t v = first ? hi : prev - hi;
                    first = false;
                    prev = v;
                    emit(v);
                    ++k;
                    if (k < len[j]) { v = prev - (b & 0x0Fu); prev = v; emit(v); ++k; }
                }
            } else {
                for (int k = 0; k < len[j]; ++k) {
                    if (i + wid[j] > size) return false;
                    uint32_t d = 0;
                    for (int q = 0; q < wid[j]; ++q) d = (d << 8) | p[i++];
                    uint32_t v = first ? d : prev - d;
                    first = false;
                    prev = v;
                    emit(v);
                }
            }
        }
    }
    return true;
}

/* ---------------- FP64 per-voxel hour terms ---------------- */

struct HourD {
    int idx, day, di, eps_bin;
    double cos_z, sin_z, az, za, am, e0, dni, dhi;
};

HourD hour_d(const float* naa, const int* sdi, const float* dni, const float* dhi, int idx) {
    HourD h;
    h.idx = idx;
    h.day = idx / 24;
    h.di = sdi[idx];
    float az = naa[2 * idx], alt = naa[2 * idx + 1];
    h.eps_bin = hour_eps_bin(alt, dhi[idx], dni[idx]);
    const double rad = kPiD / 180.0;
    h.cos_z = std::cos((90.0 - (double)alt) * rad);
    h.za = std::acos(h.cos_z);
    h.sin_z = std::sin(h.za);
    h.az = (double)az * rad;
    double z_deg = h.za / rad;
    h.am = 1.0 / (h.cos_z + 0.50572 * std::pow(96.07995 - z_deg, -1.6364));
    double B = 2.0 * kPiD * (h.day / 365.0);
    h.e0 = 1367.0 * (1.00011 + 0.034221 * std::cos(B) + 0.00128 * std::sin(B) + 0.000719 * std::cos(2 * B) +
                     0.000077 * std::sin(2 * B));
    h.dni = dni[idx];
    h.dhi = dhi[idx];
    return h;
}

double hourly_d(const HourD& h, double slope, double aspect, bool blocked) {
    double cs = std::cos(slope), ss = std::sin(slope);
    double cos_aoi = h.cos_z * cs + h.sin_z * ss * std::cos(h.az - aspect);
    double I = 0.0;
    if (cos_aoi > 0.0 && h.di >= 0 && !blocked) I += std::max(0.0, std::max(h.dni, (double)1e-3f) * cos_aoi);
    double a = std::max(0.0, cos_aoi);
    double b = std::max((double)0.08715574f, h.cos_z);
    double delta = h.am * h.dhi / h.e0;
    int i = h.eps_bin;
    double f1 = std::max(0.0, (double)kPerezF[0][i] + (double)kPerezF[1][i] * delta + (double)kPerezF[2][i] * h.za);
    double f2 = (double)kPerezF[3][i] + (double)kPerezF[4][i] * delta + (double)kPerezF[5][i] * h.za;
    double dif = h.dhi * ((1.0 - f1) * ((1.0 + cs) / 2.0) + f1 * (a / b) + f2 * ss);
    I += std::max(0.0, dif);
    return I;
}

/* Hour-invariant FP64 single-bounce weight of neighbour cell ngid for receiver t. */
double weight_d(int64_t t, int ngid, int w, int h, int n_top, const int* pp, const float* grid,
                const float* wall_z, const float* slopes, const float* aspects) {
    if (ngid < 0 || ngid >= w * h) return 0.0;
    int top = pp[t];
    int gx = top % w, gy = top / w;
    double z = (t < n_top) ? grid[gx + gy * w] : wall_z[t - n_top];
    double s = slopes[t], a = aspects[t];
    double nx = std::sin(s) * std::sin(a), ny = -std::sin(s) * std::cos(a), nz = std::cos(s);
    int ngx = ngid % w, ngy = ngid / w;
    double dx = ngx - gx, dy = ngy - gy, dz = (double)grid[ngid] - z;
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < 1e-4) return 0.0;
    double d = std::sqrt(d2);
    double cr = (nx * dx + ny * dy + nz * dz) / d;
    if (cr <= 0.0) return 0.0;
    double cs = std::max(std::fabs(dz) / d, 0.1);
    return 0.2 * cr * cs / (kPiD * d2);
}

std::vector<int> sunlit(const float* naa) {
    std::vector<int> hs;
    for (int day = 0; day < 365; ++day)
        for (int hour = 0; hour < 24; ++hour) {
            int idx = day * 24 + hour;
            if (naa[2 * idx + 1] < 1.0f) continue;  // gpu_irradiance_uncompressed.cu:117
            hs.push_back(idx);
        }
    return hs;
}

}  // namespace

extern "C" {

void orc_sun_positions(float lon, float lat, int tz, float* out) {
    for (int doy = 1; doy <= 365; ++doy) {
        int mon, md;
        month_day_2026(doy, &mon, &md);
        for (int hr = 0; hr < 24; ++hr) {
            float az, alt;
            psa(2026, lon, lat, tz, mon, md, hr, &az, &alt);
            size_t k = (size_t)(doy - 1) * 24 + hr;
            out[2 * k] = az;
            out[2 * k + 1] = alt;
        }
    }
}

int orc_sky_bins(const float* az_alt, float h_res, float v_res, float* nudged, int* f2d, int* sdi, int* hsize_o,
                 int* vsize_o, int* hgs_o) {
    unsigned hs = (unsigned)std::ceil(360.0f / h_res);
    unsigned vs = (unsigned)std::ceil(90.0f / v_res);
    std::vector<char> used(hs * vs, 0);
    for (int k = 0; k < ORC_HOURS; ++k) {
        float az = az_alt[2 * k] - 1e-8f, alt = az_alt[2 * k + 1] - 1e-8f;
        nudged[2 * k] = az;
        nudged[2 * k + 1] = alt;
        if (alt < 1.0f) continue;
        unsigned bi = (unsigned)(az / h_res), bj = (unsigned)(alt / v_res);
        if (bi < hs && bj < vs) used[bi * vs + bj] = 1;
    }
    int active = 0;
    for (unsigned i = 0; i < hs * vs; ++i) f2d[i] = used[i] ? active++ : -1;
    for (int k = 0; k < ORC_HOURS; ++k) {
        float az = nudged[2 * k], alt = nudged[2 * k + 1];
        if (alt < 1.0f) { sdi[k] = -1; continue; }
        unsigned bi = (unsigned)(az / h_res), bj = (unsigned)(alt / v_res);
        sdi[k] = (bi < hs && bj < vs) ? f2d[bi * vs + bj] : -1;
    }
    *hsize_o = (int)hs;
    *vsize_o = (int)vs;
    *hgs_o = (int)std::ceil((float)active / 8.0f);
    return active;
}

int orc_directions(int hsize, int vsize, float hres, float vres, const int* f2d, float* cdir, int* up,
                   int* dense) {
    int d = 0;
    for (int i = 0; i < hsize; ++i) {
        for (int j = 0; j < 2 * vsize; ++j, ++d) {
            float cx, cy, cz;
            float az = ((unsigned)i + 0.5f) * hres;
            if (j < vsize) {
                float e = ((unsigned)j + 0.5f) * vres;
                cx = cosf(e * kDegToRadF) * sinf(az * kDegToRadF);
                cy = -cosf(e * kDegToRadF) * cosf(az * kDegToRadF);
                cz = sinf(e * kDegToRadF);
                up[d] = 1;
                dense[d] = f2d[i * vsize + j];
            } else {
                float e = ((unsigned)j - (unsigned)vsize + 0.5f) * vres;
                cx = cosf(e * kDegToRadF) * sinf(az * kDegToRadF);
                cy = -cosf(e * kDegToRadF) * cosf(az * kDegToRadF);
                cz = -sinf(e * kDegToRadF);
                up[d] = 0;
                dense[d] = -1;
            }
            float len = sqrtf(cx * cx + cy * cy + cz * cz);
            if (len > 0.0f) { cx /= len; cy /= len; cz /= len; }
            cdir[3 * d] = cx;
            cdir[3 * d + 1] = cy;
            cdir[3 * d + 2] = cz;
        }
    }
    return d;
}

int64_t orc_detect_walls(const float* H, int w, int h, float voxel, int* parent, float* z, float* aspect,
                         int64_t cap) {
    const float kOut = -1e30f;
    auto at = [&](int x, int y) -> float {
        return (x >= 0 && x < w && y >= 0 && y < h) ? H[(size_t)y * w + x] : kOut;
    };
    int64_t n = 0;
    for (int gy = 0; gy < h; ++gy) {
        for (int gx = 0; gx < w; ++gx) {
            float z0 = H[(size_t)gy * w + gx];
            float hE = at(gx + 1, gy), hW = at(gx - 1, gy), hN = at(gx, gy - 1), hS = at(gx, gy + 1);
            float cE = (hE <= kOut * 0.5f) ? z0 : hE;
            float cW = (hW <= kOut * 0.5f) ? z0 : hW;
            float cN = (hN <= kOut * 0.5f) ? z0 : hN;
            float cS = (hS <= kOut * 0.5f) ? z0 : hS;
            float mn = std::min(std::min(cE, cW), std::min(cN, cS));
            float drop = z0 - mn;
            if (drop <= voxel) continue;
            int nv = (int)(drop / voxel);
            for (int v = 0; v < nv; ++v) {
                float zz = z0 - (v + 0.5f) * voxel;
                float oE = std::max(0.0f, zz - hE), oW = std::max(0.0f, zz - hW);
                float oN = std::max(0.0f, zz - hN), oS = std::max(0.0f, zz - hS);
                float asp = atan2f(oE - oW, oN - oS);
                if (asp < 0.0f) asp += 2.0f * (float)kPiD;
                if (n < cap) { parent[n] = gy * w + gx; z[n] = zz; aspect[n] = asp; }
                ++n;
            }
        }
    }
    return n;
}

void orc_top_slope_aspect(const float* H, int w, int h, double sx, double sy, float* slope, float* aspect) {
    for (int gy = 0; gy < h; ++gy) {
        for (int gx = 0; gx < w; ++gx) {
            int l = std::max(gx - 1, 0), r = std::min(gx + 1, w - 1);
            int u = std::max(gy - 1, 0), d = std::min(gy + 1, h - 1);
            float zl = H[(size_t)gy * w + l], zr = H[(size_t)gy * w + r];
            float zu = H[(size_t)u * w + gx], zd = H[(size_t)d * w + gx];
            float dx = (zr - zl) / ((r - l) * (float)sx);
            float dy = (zd - zu) / ((d - u) * (float)sy);
            size_t idx = (size_t)gy * w + gx;
            slope[idx] = atan2f(sqrtf(dx * dx + dy * dy), 1.0f);
            float a = atan2f(-dx, dy);
            if (a < 0.0f) a += 2.0f * (float)kPiD;
            aspect[idx] = a;
        }
    }
}

int64_t orc_check_setup_division(const float* cdir, int ndir, int gmax) {
    int64_t bad = 0;
    for (int k = 0; k < ndir; ++k) {
        const DirConst d = make_dir(cdir + 3 * k, 1);
        for (int axis = 0; axis < 2; ++axis) {
            const float c = axis ? d.cy : d.cx, a = axis ? d.ay : d.ax, r = axis ? d.ry : d.rx;
            for (int g = 0; g < gmax; ++g) {
                const float o = (float)g + c * 2.0f;   // the set-up of trace() above
                const int i = (int)std::floor(o);
                const float num = (c >= 0.0f) ? ((float)i + 1.0f - o) : (o - (float)i);
                const float q_ieee = num / a;
                const float q0 = num * r;
                const float q = std::fma(std::fma(-q0, a, num), r, q0);
                uint32_t x, y;
                std::memcpy(&x, &q_ieee, 4);
                std::memcpy(&y, &q, 4);
                bad += (x != y);
            }
        }
    }
    return bad;
}

int orc_trace_ray(const float* grid, int w, int h, float maxz, float gxf, float gyf, float base_z,
                  const float* cdir3, int is_upper, int* steps) {
    DirConst d = make_dir(cdir3, is_upper);
    return trace(grid, w, h, maxz, gxf, gyf, base_z, d, steps);
}

void orc_trace_raw(const float* grid, int w, int h, int n_top, const int* pp, const float* wall_z, float maxz,
                   const float* cdir, const int* up, int ndir, int v0, int count, int nthreads, int32_t* raw) {
    std::vector<DirConst> dirs(ndir);
    for (int d = 0; d < ndir; ++d) dirs[d] = make_dir(cdir + 3 * d, up[d]);
    parallel_blocks(count, 256, nthreads, [&](int64_t a, int64_t b, int) {
        for (int64_t k = a; k < b; ++k) {
            int64_t t = v0 + k;
            int top = pp[t];
            int gx = top % w, gy = top / w;
            float bz = (t < n_top) ? grid[gx + gy * w] : wall_z[t - n_top];
            for (int d = 0; d < ndir; ++d)
                raw[k * ndir + d] = trace(grid, w, h, maxz, (float)gx, (float)gy, bz, dirs[d], nullptr);
        }
    });
}

}  // extern "C"

namespace {

/* Viewshed of n voxels (voxel_of(i) for i < n): SVF bits at svf[i + byte * n], CSR lists by
 * position.  The loop of run_viewshed_out_of_core + the CSR consumer (viewshed_stage.cuh:56-121,
 * gpu_irradiance_uncompressed.cu:79-95). */
template <typename VoxelOf>
int64_t viewshed_impl(VoxelOf voxel_of, int64_t n, const float* grid, int w, int h, int n_top, const int* pp,
                      const float* wall_z, float maxz, const float* cdir, const int* up, const int* dense, int ndir,
                      int hgs, int nthreads, uint8_t* svf, int64_t* off, int32_t* ids, int64_t cap,
                      int64_t* total_steps) {
    std::vector<DirConst> dirs(ndir);
    for (int d = 0; d < ndir; ++d) dirs[d] = make_dir(cdir + 3 * d, up[d]);
    if (svf) std::memset(svf, 0, (size_t)hgs * n);
    const int64_t kBlock = 512;
    int64_t nblocks = (n + kBlock - 1) / kBlock;
    std::vector<std::vector<int32_t>> block_ids(nblocks);
    std::vector<int32_t> counts(n);
    int nt = resolve_threads(nthreads);
    std::vector<int64_t> steps_per_thread(nt, 0);
    parallel_blocks(n, kBlock, nt, [&](int64_t a, int64_t b, int tid) {
        IdSet set(8192);
        std::vector<int32_t>& out = block_ids[a / kBlock];
        int64_t st = 0;
        for (int64_t i = a; i < b; ++i) {
            const int64_t t = voxel_of(i);
            int top = pp[t];
            int gx = top % w, gy = top / w;
            float bz = (t < n_top) ? grid[gx + gy * w] : wall_z[t - n_top];
            set.clear();
            int c = 0;
            for (int d = 0; d < ndir; ++d) {
                int s = 0;
                int hit = trace(grid, w, h, maxz, (float)gx, (float)gy, bz, dirs[d], &s);
                st += s;
                if (!hit) continue;
                if (svf && dense[d] >= 0) svf[i + (int64_t)(dense[d] / 8) * n] |= (uint8_t)(1u << (dense[d] % 8));
                if (set.insert(hit)) { out.push_back(hit); ++c; }
            }
            counts[i] = c;
        }
        steps_per_thread[tid] += st;
    });
    int64_t nnz = 0;
    for (int64_t i = 0; i < n; ++i) { off[i] = nnz; nnz += counts[i]; }
    off[n] = nnz;
    if (total_steps) {
        int64_t s = 0;
        for (int64_t x : steps_per_thread) s += x;
        *total_steps = s;
    }
    if (nnz > cap) return -nnz;
    int64_t pos = 0;
    for (auto& v : block_ids) {
        std::memcpy(ids + pos, v.data(), v.size() * sizeof(int32_t));
        pos += (int64_t)v.size();
    }
    return nnz;
}

}  // namespace

extern "C" {

int64_t orc_viewshed(const float* grid, int w, int h, int n_total, int n_top, const int* pp, const float* wall_z,
                     float maxz, const float* cdir, const int* up, const int* dense, int ndir, int hgs,
                     int nthreads, uint8_t* svf, int64_t* off, int32_t* ids, int64_t cap, int64_t* total_steps) {
    return viewshed_impl([](int64_t i) { return i; }, n_total, grid, w, h, n_top, pp, wall_z, maxz, cdir, up, dense,
                         ndir, hgs, nthreads, svf, off, ids, cap, total_steps);
}

int64_t orc_viewshed_subset(const float* grid, int w, int h, int n_top, const int* pp, const float* wall_z,
                            float maxz, const float* cdir, const int* up, const int* dense, int ndir, int hgs,
                            const int32_t* voxels, int64_t n, int nthreads, uint8_t* svf, int64_t* off,
                            int32_t* ids, int64_t cap, int64_t* total_steps) {
    return viewshed_impl([voxels](int64_t i) { return (int64_t)voxels[i]; }, n, grid, w, h, n_top, pp, wall_z, maxz,
                         cdir, up, dense, ndir, hgs, nthreads, svf, off, ids, cap, total_steps);
}

int64_t orc_barle_encode_voxel(const int32_t* ids, int n, uint8_t* out, int64_t cap) {
    std::vector<uint32_t> v;
    for (int k = 0; k < n; ++k) {
        if (ids[k] == 0) break;  // barle_compression.cpp:160-163
        v.push_back((uint32_t)ids[k]);
    }
    std::vector<uint8_t> enc;
    encode_list(v, enc);
    if ((int64_t)enc.size() > cap) return -(int64_t)enc.size();
    if (!enc.empty()) std::memcpy(out, enc.data(), enc.size());
    return (int64_t)enc.size();
}

int64_t orc_barle_encode_all(const int64_t* off, const int32_t* ids, int n_total, int nthreads, int64_t* off_out,
                             uint8_t* bytes, int64_t cap) {
    const int64_t kBlock = 1024;
    int64_t nblocks = ((int64_t)n_total + kBlock - 1) / kBlock;
    std::vector<std::vector<uint8_t>> blk(nblocks);
    std::vector<int64_t> sizes(n_total);
    parallel_blocks(n_total, kBlock, nthreads, [&](int64_t a, int64_t b, int) {
        std::vector<uint32_t> v;
        std::vector<uint8_t> enc;
        std::vector<uint8_t>& dst = blk[a / kBlock];
        for (int64_t t = a; t < b; ++t) {
            v.clear();
            for (int64_t k = off[t]; k < off[t + 1]; ++k) {
                if (ids[k] == 0) break;
                v.push_back((uint32_t)ids[k]);
            }
            encode_list(v, enc);
            sizes[t] = (int64_t)enc.size();
            dst.insert(dst.end(), enc.begin(), enc.end());
        }
    });
    int64_t total = 0;
    for (int64_t t = 0; t < n_total; ++t) { off_out[t] = total; total += sizes[t]; }
    off_out[n_total] = total;
    if (total > cap) return -total;
    int64_t pos = 0;
    for (auto& b : blk) {
        if (!b.empty()) std::memcpy(bytes + pos, b.data(), b.size());
        pos += (int64_t)b.size();
    }
    return total;
}

int orc_barle_decode(const uint8_t* bytes, int64_t size, uint32_t* vals, int cap) {
    int n = 0;
    bool ok = decode_stream(bytes, size, [&](uint32_t v) {
        if (n < cap) vals[n] = v;
        ++n;
    });
    return ok ? n : -1;
}

int orc_hour_eps(float alt, float dif, float dir, float* eps_out) {
    const float cos_z = cosf((90.0f - alt) * kDegToRadF);
    const float za = acosf(cos_z);
    float eps = ((dif + dir / cosf(za)) / dif + 1.041f * (za * za * za)) / (1.0f + 1.041f * za * za * za);
    eps = fmaxf(1.0f, eps);
    *eps_out = eps;
    return eps_bin(eps);
}

int orc_sunlit_hours(const float* naa, int* hours_out) {
    std::vector<int> hs = sunlit(naa);
    for (size_t i = 0; i < hs.size(); ++i) hours_out[i] = hs[i];
    return (int)hs.size();
}

void orc_solar_hour(const float* naa, const int* sdi, const float* dni, const float* dhi, int idx, int n,
                    const float* slopes, const float* aspects, const uint8_t* svf, int n_total, int nthreads,
                    float* hourly) {
    float az = naa[2 * idx], alt = naa[2 * idx + 1];
    int di = sdi[idx];
    int day = idx / 24;
    parallel_blocks(n, 4096, nthreads, [&](int64_t a, int64_t b, int) {
        for (int64_t t = a; t < b; ++t) {
            bool blk = (di >= 0) ? svf_bit(svf, t, di, n_total) : false;
            hourly[t] = solar_voxel(az, alt, di, dni[idx], dhi[idx], day, slopes[t], aspects[t], blk);
        }
    });
}

void orc_reflect_hour_csr(const float* grid, int w, int h, int n_total, int n_top, const int* pp,
                          const float* wall_z, const float* slopes, const float* aspects, const int64_t* off,
                          const int32_t* ids, const float* hourly, int nthreads, float* refl) {
    parallel_blocks(n_total, 2048, nthreads, [&](int64_t a, int64_t b, int) {
        for (int64_t t = a; t < b; ++t) {
            Receiver r = receiver(t, w, n_top, pp, grid, wall_z, slopes, aspects);
            float R = 0.0f;
            for (int64_t k = off[t]; k < off[t + 1]; ++k) add_reflection(ids[k] - 1, r, grid, w, h, hourly, R);
            refl[t] = R;
        }
    });
}

void orc_reflect_hour_barle(const float* grid, int w, int h, int n_total, int n_top, const int* pp,
                            const float* wall_z, const float* slopes, const float* aspects, const int64_t* off,
                            const uint8_t* bytes, const float* hourly, int nthreads, float* refl) {
    parallel_blocks(n_total, 2048, nthreads, [&](int64_t a, int64_t b, int) {
        for (int64_t t = a; t < b; ++t) {
            Receiver r = receiver(t, w, n_top, pp, grid, wall_z, slopes, aspects);
            float R = 0.0f;
            decode_stream(bytes + off[t], off[t + 1] - off[t], [&](uint32_t v) {
                if (v > 0) add_reflection((int)v - 1, r, grid, w, h, hourly, R);
            });
            refl[t] = R;
        }
    });
}

int orc_annual_fp32(const float* grid, int w, int h, int n_total, int n_top, const int* pp, const float* wall_z,
                    const float* slopes, const float* aspects, const float* naa, const int* sdi, const float* dni,
                    const float* dhi, const uint8_t* svf, const int64_t* off, const int32_t* ids,
                    const uint8_t* bytes, int mode, int stride, int kahan, int nthreads, float* out) {
    std::vector<int> hs = sunlit(naa);
    std::vector<float> hourly(n_total), refl(n_total), comp(n_total, 0.0f);
    std::fill(out, out + n_total, 0.0f);
    if (stride < 1) stride = 1;
    int done = 0;
    for (size_t q = 0; q < hs.size(); q += stride) {
        int idx = hs[q];
        orc_solar_hour(naa, sdi, dni, dhi, idx, n_total, slopes, aspects, svf, n_total, nthreads, hourly.data());
        if (mode == 0)
            orc_reflect_hour_csr(grid, w, h, n_total, n_top, pp, wall_z, slopes, aspects, off, ids, hourly.data(),
                                 nthreads, refl.data());
        else
            orc_reflect_hour_barle(grid, w, h, n_total, n_top, pp, wall_z, slopes, aspects, off, bytes,
                                   hourly.data(), nthreads, refl.data());
        for (int64_t t = 0; t < n_total; ++t) {  // solar_kernel.cu:79-85
            float add = refl[t] + hourly[t];
            if (kahan) {
                float y = add - comp[t];
                float s = out[t] + y;
                comp[t] = (s - out[t]) - y;
                out[t] = s;
            } else {
                out[t] += add;
            }
        }
        ++done;
    }
    return done;
}

}  // extern "C"

namespace {

/* The reference's FP32 hour loop (gpu_irradiance_compressed.cu:119-143 with the CSR walk of
 * reflection_kernels.cu:70-89) for `count` voxels voxel_of(k), with their lists by position:
 * per sampled hour, the solar term of every top cell (the neighbours' values), then each
 * voxel's solar + reflection + Kahan accumulate.  Times the two parts separately. */
template <typename VoxelOf>
int annual_fp32_sample(VoxelOf voxel_of, int64_t count, const float* grid, int w, int h, int n_total, int n_top,
                       const int* pp, const float* wall_z, const float* slopes, const float* aspects,
                       const float* naa, const int* sdi, const float* dni, const float* dhi, const uint8_t* svf,
                       const int64_t* off, const int32_t* ids, int stride, int nthreads, float* out,
                       double* t_top_s, double* t_range_s) {
    std::vector<int> hs = sunlit(naa);
    std::vector<float> hourly(n_top), comp(count, 0.0f);
    double tt = 0.0, tr = 0.0;
    std::fill(out, out + count, 0.0f);
    if (stride < 1) stride = 1;
    int done = 0;
    for (size_t q = 0; q < hs.size(); q += stride) {
        const int idx = hs[q];
        auto t0 = std::chrono::steady_clock::now();
        orc_solar_hour(naa, sdi, dni, dhi, idx, n_top, slopes, aspects, svf, n_total, nthreads, hourly.data());
        auto t1 = std::chrono::steady_clock::now();
        const float az = naa[2 * idx], alt = naa[2 * idx + 1];
        const int di = sdi[idx], day = idx / 24;
        parallel_blocks(count, 1024, nthreads, [&](int64_t a, int64_t b, int) {
            for (int64_t k = a; k < b; ++k) {
                const int64_t t = voxel_of(k);
                bool blk = (di >= 0) ? svf_bit(svf, t, di, n_total) : false;
                float hv = solar_voxel(az, alt, di, dni[idx], dhi[idx], day, slopes[t], aspects[t], blk);
                Receiver r = receiver(t, w, n_top, pp, grid, wall_z, slopes, aspects);
                float R = 0.0f;
                for (int64_t e = off[k]; e < off[k + 1]; ++e) add_reflection(ids[e] - 1, r, grid, w, h, hourly.data(), R);
                float add = R + hv;
                float y = add - comp[k];
                float s = out[k] + y;
                comp[k] = (s - out[k]) - y;
                out[k] = s;
            }
        });
        auto t2 = std::chrono::steady_clock::now();
        tt += std::chrono::duration<double>(t1 - t0).count();
        tr += std::chrono::duration<double>(t2 - t1).count();
        ++done;
    }
    if (t_top_s) *t_top_s = tt;
    if (t_range_s) *t_range_s = tr;
    return done;
}

}  // namespace

extern "C" {

int orc_annual_fp32_range(const float* grid, int w, int h, int n_total, int n_top, const int* pp,
                          const float* wall_z, const float* slopes, const float* aspects, const float* naa,
                          const int* sdi, const float* dni, const float* dhi, const uint8_t* svf, int64_t v0,
                          int64_t count, const int64_t* off, const int32_t* ids, int stride, int nthreads,
                          float* out, double* t_top_s, double* t_range_s) {
    return annual_fp32_sample([v0](int64_t k) { return v0 + k; }, count, grid, w, h, n_total, n_top, pp, wall_z,
                              slopes, aspects, naa, sdi, dni, dhi, svf, off, ids, stride, nthreads, out, t_top_s,
                              t_range_s);
}

int orc_annual_fp32_subset(const float* grid, int w, int h, int n_total, int n_top, const int* pp,
                           const float* wall_z, const float* slopes, const float* aspects, const float* naa,
                           const int* sdi, const float* dni, const float* dhi, const uint8_t* svf,
                           const int32_t* voxels, int64_t count, const int64_t* off, const int32_t* ids, int stride,
                           int nthreads, float* out, double* t_top_s, double* t_range_s) {
    return annual_fp32_sample([voxels](int64_t k) { return (int64_t)voxels[k]; }, count, grid, w, h, n_total, n_top,
                              pp, wall_z, slopes, aspects, naa, sdi, dni, dhi, svf, off, ids, stride, nthreads, out,
                              t_top_s, t_range_s);
}

}  // extern "C"

namespace {

/* FP64 annual totals by single-bounce linearity (the Q24 identity, checked against the per-hour
 * FP64 sum): each output voxel's own annual direct + diffuse plus the sum over its neighbours of
 * the hour-invariant weight times the neighbour cell's annual direct + diffuse.  lists(i, t)
 * gives output i's (voxel t's) neighbour ids as a [begin, end) pair of pointers. */
template <typename Lists>
void annual_fp64_impl(const float* grid, int w, int h, int n_total, int n_top, const int* pp, const float* wall_z,
                      const float* slopes, const float* aspects, const float* naa, const int* sdi, const float* dni,
                      const float* dhi, const uint8_t* svf, const int32_t* voxels, int64_t n_out, int nthreads,
                      Lists lists, double* out) {
    std::vector<int> hs = sunlit(naa);
    std::vector<HourD> hd;
    for (int idx : hs) hd.push_back(hour_d(naa, sdi, dni, dhi, idx));
    auto voxel_of = [&](int64_t i) -> int64_t { return voxels ? voxels[i] : i; };
    auto annual_direct = [&](int64_t t) -> double {
        double s = 0.0;
        for (const HourD& x : hd) {
            bool blk = (x.di >= 0) ? svf_bit(svf, t, x.di, n_total) : false;
            s += hourly_d(x, slopes[t], aspects[t], blk);
        }
        return s;
    };
    /* Annual hourly sums of every top cell some selected voxel sees. */
    std::vector<char> need((size_t)n_top, 0);
    for (int64_t i = 0; i < n_out; ++i) {
        auto l = lists(i, voxel_of(i));
        for (const int32_t* p = l.first; p < l.second; ++p) {
            int c = *p - 1;
            if (c >= 0 && c < n_top) need[c] = 1;
        }
    }
    std::vector<int32_t> cells;
    for (int c = 0; c < n_top; ++c) if (need[c]) cells.push_back(c);
    std::vector<double> A((size_t)n_top, 0.0);
    parallel_blocks((int64_t)cells.size(), 64, nthreads, [&](int64_t a, int64_t b, int) {
        for (int64_t q = a; q < b; ++q) A[cells[q]] = annual_direct(cells[q]);
    });
    parallel_blocks(n_out, 64, nthreads, [&](int64_t a, int64_t b, int) {
        for (int64_t i = a; i < b; ++i) {
            int64_t t = voxel_of(i);
            double s = (t < n_top && need[t]) ? A[t] : annual_direct(t);
            double r = 0.0;
            auto l = lists(i, t);
            for (const int32_t* p = l.first; p < l.second; ++p) {
                int c = *p - 1;
                double wgt = weight_d(t, c, w, h, n_top, pp, grid, wall_z, slopes, aspects);
                if (wgt != 0.0) r += wgt * A[c];
            }
            out[i] = s + r;
        }
    });
}

}  // namespace

extern "C" {

void orc_annual_fp64(const float* grid, int w, int h, int n_total, int n_top, const int* pp, const float* wall_z,
                     const float* slopes, const float* aspects, const float* naa, const int* sdi, const float* dni,
                     const float* dhi, const uint8_t* svf, const int64_t* off, const int32_t* ids,
                     const int32_t* voxels, int n_sel, int nthreads, double* out) {
    const int64_t n_out = voxels ? n_sel : n_total;
    annual_fp64_impl(grid, w, h, n_total, n_top, pp, wall_z, slopes, aspects, naa, sdi, dni, dhi, svf, voxels, n_out,
                     nthreads,
                     [&](int64_t, int64_t t) { return std::make_pair(ids + off[t], ids + off[t + 1]); }, out);
}

void orc_annual_fp64_subset(const float* grid, int w, int h, int n_total, int n_top, const int* pp,
                            const float* wall_z, const float* slopes, const float* aspects, const float* naa,
                            const int* sdi, const float* dni, const float* dhi, const uint8_t* svf,
                            const int32_t* voxels, int64_t n_sel, const int64_t* off_sub, const int32_t* ids_sub,
                            int nthreads, double* out) {
    annual_fp64_impl(grid, w, h, n_total, n_top, pp, wall_z, slopes, aspects, naa, sdi, dni, dhi, svf, voxels, n_sel,
                     nthreads,
                     [&](int64_t i, int64_t) { return std::make_pair(ids_sub + off_sub[i], ids_sub + off_sub[i + 1]); },
                     out);
}

void orc_annual_fp64_direct(const float* grid, int w, int h, int n_total, int n_top, const int* pp,
                            const float* wall_z, const float* slopes, const float* aspects, const float* naa,
                            const int* sdi, const float* dni, const float* dhi, const uint8_t* svf,
                            const int64_t* off, const int32_t* ids, double* out) {
    std::vector<int> hs = sunlit(naa);
    std::vector<double> hourly(n_total);
    std::fill(out, out + n_total, 0.0);
    for (int idx : hs) {
        HourD x = hour_d(naa, sdi, dni, dhi, idx);
        for (int64_t t = 0; t < n_total; ++t) {
            bool blk = (x.di >= 0) ? svf_bit(svf, t, x.di, n_total) : false;
            hourly[t] = hourly_d(x, slopes[t], aspects[t], blk);
        }
        for (int64_t t = 0; t < n_total; ++t) {
            double R = 0.0;
            for (int64_t k = off[t]; k < off[t + 1]; ++k) {
                int c = ids[k] - 1;
                if (c < 0 || c >= w * h) continue;
                R += weight_d(t, c, w, h, n_top, pp, grid, wall_z, slopes, aspects) * hourly[c];
            }
            out[t] += hourly[t] + R;
        }
    }
}

}  // extern "C"
