/* Persisted neighbour lists (no reference counterpart: the reference recomputes the viewshed on
 * every run).  sgpu_save_lists writes a context's SVF bitmap and its device-resident CSR and/or
 * ba-RLE lists (int64 offsets) to a file; sgpu_load_lists reads them into a context whose scene
 * and sky match, skipping stage 1 on reruns and weather sweeps.
 *
 * File (little-endian): ListFileHeader, then the SVF bytes (svf_bytes x num_total, word-major),
 * CSR offsets (num_total + 1 int64) and ids (nnz int32) when present, ba-RLE offsets (num_total
 * + 1 int64) and bytes when present.  The header carries the scene's and the sky's checksums
 * (a position-weighted 64-bit sum of the words of the heightmap, parents and wall heights; of
 * the dense sun-bin map; the scene's also takes the height limit), so lists are only loaded
 * into the scene and sky they were made for.
 * Sections stream through a pinned host buffer in 64 MB pieces.  A load checks the file's
 * length against its header and every loaded section on the device (offsets monotone and
 * ending at the section size, ids within the grid, ba-RLE streams walking block by block to
 * their ends) before the lists are used: a damaged file is refused with SGPU_ERR_FORMAT. */
#include <cstdio>
#include <cstring>
#include <memory>

#include "sgpu_ctx.h"

namespace sgpu {
namespace {

constexpr char kMagic[8] = {'S', 'G', 'P', 'U', 'L', 'S', 'T', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kPiece = (size_t)64 << 20;

struct ListFileHeader {
    char magic[8];
    uint32_t version;
    uint32_t formats;          // SGPU_FMT_CSR | SGPU_FMT_BARLE sections present
    int32_t width, height;
    int64_t num_total, num_top;
    uint64_t scene_sum, sky_sum;
    int32_t svf_bytes, reserved;
    int64_t nnz;               // CSR ids (0 without the CSR section)
    int64_t barle_bytes;       // ba-RLE bytes (0 without the ba-RLE section)
};
static_assert(sizeof(ListFileHeader) == 80, "header layout");

/* sum over i of word(x[i]) * (2 i + 1 + salt), mod 2^64 (order-sensitive, associative). */
__global__ __launch_bounds__(256) void k_checksum(const uint32_t* x, int64_t n, uint64_t salt,
                                                  unsigned long long* out) {
    uint64_t acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        acc += (uint64_t)x[i] * (2ull * (uint64_t)i + 1ull + salt);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += (uint64_t)__shfl_xor((unsigned long long)acc, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)acc);
}

uint64_t device_checksum(hipStream_t s, const void* p, int64_t words, uint64_t salt, unsigned long long* tmp) {
    SGPU_HIP(hipMemsetAsync(tmp, 0, sizeof(unsigned long long), s));
    if (words > 0) {
        const int64_t blocks = std::min<int64_t>((words + 255) / 256, 4096);
        hipLaunchKernelGGL(k_checksum, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<const uint32_t*>(p),
                           words, salt, tmp);
        SGPU_HIP(hipGetLastError());
    }
    unsigned long long v = 0;
    SGPU_HIP(hipMemcpyAsync(&v, tmp, sizeof v, hipMemcpyDeviceToHost, s));
    SGPU_HIP(hipStreamSynchronize(s));
    return v;
}

uint64_t scene_checksum(sgpu_ctx* c) {
    unsigned long long* tmp = dalloc<unsigned long long>(1);
    uint64_t sum = 0;
    try {
        sum = device_checksum(c->stream, c->grid, c->n_top, 1, tmp);
        sum = sum * 0x9E3779B97F4A7C15ull + device_checksum(c->stream, c->pix_parent, c->n_total, 2, tmp);
        sum = sum * 0x9E3779B97F4A7C15ull + device_checksum(c->stream, c->wall_z, c->n_total - c->n_top, 3, tmp);
        uint32_t zbits = 0;   // the height limit rays leave through
        std::memcpy(&zbits, &c->maxz, sizeof zbits);
        sum = sum * 0x9E3779B97F4A7C15ull + zbits;
    } catch (...) {
        dfree(tmp);
        throw;
    }
    dfree(tmp);
    return sum;
}

uint64_t sky_checksum(const sgpu_ctx* c) {
    uint64_t sum = 0;
    for (size_t i = 0; i < c->sky.full_to_dense.size(); ++i)
        sum += (uint64_t)(uint32_t)c->sky.full_to_dense[i] * (2ull * i + 1ull);
    return sum;
}

/* Checks of loaded sections, on the device so that 8192^2-sized sections need no host copy.
 * Each records the first bad index in *bad (left at ~0 when the section is sound). */

/* offsets[0..n]: off[0] == 0, 0 <= off[i+1] - off[i] <= max_len, off[n] == total. */
__global__ __launch_bounds__(256) void k_check_offsets(const int64_t* __restrict__ off, int64_t n, int64_t total,
                                                       int64_t max_len, unsigned long long* bad) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i <= n; i += (int64_t)gridDim.x * 256) {
        bool b;
        if (i == 0) {
            b = off[0] != 0;
        } else {
            const int64_t len = off[i] - off[i - 1];
            b = len < 0 || len > max_len || (i == n && off[n] != total);
        }
        if (b) atomicMin(bad, (unsigned long long)i);
    }
}

/* CSR ids are 1-based top-cell ids: 1..cells. */
__global__ __launch_bounds__(256) void k_check_ids(const int32_t* __restrict__ ids, int64_t nnz, int32_t cells,
                                                   unsigned long long* bad) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * 256) {
        const int32_t id = ids[k];
        if (id <= 0 || id > cells) atomicMin(bad, (unsigned long long)k);
    }
}

/* Every ba-RLE stream (offsets already checked) walks block by block to exactly its end:
 * 3 header bytes of width codes (0..4, 7 = pad), a length byte per run, then the payload --
 * the structure the reflection pass's decoder relies on to stay inside a stream.  Every byte
 * read is inside [off[v], off[v+1]). */
__global__ __launch_bounds__(256) void k_check_barle(const int64_t* __restrict__ off,
                                                     const uint8_t* __restrict__ bytes, int64_t n,
                                                     unsigned long long* bad) {
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (int64_t)gridDim.x * 256) {
        const int64_t b = off[v + 1];
        int64_t i = off[v];
        bool ok = true;
        while (ok && i < b) {
            if (i + 3 > b) { ok = false; break; }
            const unsigned bits = ((unsigned)bytes[i] << 16) | ((unsigned)bytes[i + 1] << 8) | bytes[i + 2];
            i += 3;
            unsigned w[8];
            int nc = 0;
            for (int j = 0; j < 8; ++j) {
                const unsigned cd = (bits >> (21 - 3 * j)) & 7u;
                if (cd == 7u) break;
                if (cd > 4u) ok = false;
                w[nc++] = cd;
            }
            if (!ok || i + nc > b) { ok = false; break; }
            int64_t payload = 0;
            for (int j = 0; j < nc; ++j) {
                const int64_t len = bytes[i + j];
                payload += (w[j] == 0) ? (len + 1) / 2 : len * (int64_t)w[j];
            }
            i += nc + payload;
            if (i > b) ok = false;
        }
        if (!ok) atomicMin(bad, (unsigned long long)v);
    }
}

/* Runs one check kernel and fails with SGPU_ERR_FORMAT naming the section and index. */
template <typename Launch>
void check_section(sgpu_ctx* c, unsigned long long* flag, const char* what, Launch&& launch) {
    SGPU_HIP(hipMemsetAsync(flag, 0xFF, sizeof(unsigned long long), c->stream));
    launch();
    SGPU_HIP(hipGetLastError());
    unsigned long long first = 0;
    SGPU_HIP(hipMemcpyAsync(&first, flag, sizeof first, hipMemcpyDeviceToHost, c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    if (first != ~0ull) fail(SGPU_ERR_FORMAT, "list file: %s damaged (first bad entry %llu)", what, first);
}

unsigned check_blocks(int64_t items) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, 8192)); }

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

struct Pinned {
    void* p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

/* device -> file, or file -> device, through the pinned buffer */
void write_section(sgpu_ctx* c, FILE* f, const void* dev, size_t bytes, void* buf) {
    for (size_t o = 0; o < bytes; o += kPiece) {
        const size_t n = std::min(kPiece, bytes - o);
        SGPU_HIP(hipMemcpyAsync(buf, static_cast<const uint8_t*>(dev) + o, n, hipMemcpyDeviceToHost, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        if (std::fwrite(buf, 1, n, f) != n) fail(SGPU_ERR_IO, "list file: write failed");
    }
}

void read_section(sgpu_ctx* c, FILE* f, void* dev, size_t bytes, void* buf) {
    for (size_t o = 0; o < bytes; o += kPiece) {
        const size_t n = std::min(kPiece, bytes - o);
        if (std::fread(buf, 1, n, f) != n) fail(SGPU_ERR_FORMAT, "list file: truncated");
        SGPU_HIP(hipMemcpyAsync(static_cast<uint8_t*>(dev) + o, buf, n, hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
    }
}

}  // namespace
}  // namespace sgpu

using namespace sgpu;

int sgpu_save_lists(sgpu_ctx* c, const char* path, unsigned formats) {
    return guarded([&] {
        need(c && path, SGPU_ERR_ARG, "null argument");
        need((formats & ~(unsigned)(SGPU_FMT_CSR | SGPU_FMT_BARLE)) == 0 && formats != 0, SGPU_ERR_ARG,
             "bad formats");
        need(c->scene && c->sky_set && c->svf_ok, SGPU_ERR_STATE, "no viewshed to save");
        if (formats & SGPU_FMT_CSR)
            need(c->nnz >= 0 && c->csr_off, SGPU_ERR_STATE, "no device-resident CSR lists");
        if (formats & SGPU_FMT_BARLE)
            need(c->br_size >= 0 && c->br_off, SGPU_ERR_STATE, "no device-resident ba-RLE lists");
        c->set_device();
        ListFileHeader hd{};
        std::memcpy(hd.magic, kMagic, sizeof kMagic);
        hd.version = kVersion;
        hd.formats = formats;
        hd.width = c->W;
        hd.height = c->H;
        hd.num_total = c->n_total;
        hd.num_top = c->n_top;
        hd.scene_sum = scene_checksum(c);
        hd.sky_sum = sky_checksum(c);
        hd.svf_bytes = c->sky.svf_bytes;
        hd.nnz = (formats & SGPU_FMT_CSR) ? c->nnz : 0;
        hd.barle_bytes = (formats & SGPU_FMT_BARLE) ? c->br_size : 0;
        File out;
        out.f = std::fopen(path, "wb");
        need(out.f != nullptr, SGPU_ERR_IO, "list file: cannot open for writing");
        if (std::fwrite(&hd, sizeof hd, 1, out.f) != 1) fail(SGPU_ERR_IO, "list file: write failed");
        Pinned buf;
        SGPU_HIP(hipHostMalloc(&buf.p, kPiece, hipHostMallocDefault));
        const size_t offs = (size_t)(c->n_total + 1) * sizeof(int64_t);
        write_section(c, out.f, c->svf, (size_t)c->sky.svf_bytes * c->n_total, buf.p);
        if (formats & SGPU_FMT_CSR) {
            write_section(c, out.f, c->csr_off, offs, buf.p);
            write_section(c, out.f, c->csr_ids, (size_t)c->nnz * sizeof(int32_t), buf.p);
        }
        if (formats & SGPU_FMT_BARLE) {
            write_section(c, out.f, c->br_off, offs, buf.p);
            write_section(c, out.f, c->br_bytes, (size_t)c->br_size, buf.p);
        }
        if (std::fclose(out.f) != 0) {
            out.f = nullptr;
            fail(SGPU_ERR_IO, "list file: write failed");
        }
        out.f = nullptr;
    });
}

int sgpu_load_lists(sgpu_ctx* c, const char* path) {
    return guarded([&] {
        need(c && path, SGPU_ERR_ARG, "null argument");
        need(c->scene && c->sky_set, SGPU_ERR_STATE, "scene/sky not set");
        c->set_device();
        File in;
        in.f = std::fopen(path, "rb");
        need(in.f != nullptr, SGPU_ERR_IO, "list file: cannot open");
        ListFileHeader hd{};
        if (std::fread(&hd, sizeof hd, 1, in.f) != 1) fail(SGPU_ERR_FORMAT, "list file: truncated header");
        need(std::memcmp(hd.magic, kMagic, sizeof kMagic) == 0 && hd.version == kVersion, SGPU_ERR_FORMAT,
             "not a neighbour-list file of this version");
        need(hd.width == c->W && hd.height == c->H && hd.num_total == c->n_total && hd.num_top == c->n_top,
             SGPU_ERR_ARG, "list file is for another scene (dimensions)");
        need(hd.scene_sum == scene_checksum(c), SGPU_ERR_ARG, "list file is for another scene (checksum)");
        need(hd.sky_sum == sky_checksum(c) && hd.svf_bytes == c->sky.svf_bytes, SGPU_ERR_ARG,
             "list file is for another sky (sun bins)");
        need((hd.formats & ~(unsigned)(SGPU_FMT_CSR | SGPU_FMT_BARLE)) == 0 && hd.formats != 0 && hd.nnz >= 0 &&
                 hd.barle_bytes >= 0 && hd.nnz <= (int64_t)kDirs * c->n_total &&
                 hd.barle_bytes <= ((int64_t)1 << 20) * c->n_total,
             SGPU_ERR_FORMAT, "list file: bad header");
        if ((hd.formats & SGPU_FMT_CSR) == 0) need(hd.nnz == 0, SGPU_ERR_FORMAT, "list file: bad header");
        if ((hd.formats & SGPU_FMT_BARLE) == 0) need(hd.barle_bytes == 0, SGPU_ERR_FORMAT, "list file: bad header");
        const size_t offs = (size_t)(c->n_total + 1) * sizeof(int64_t);
        {   /* the sections the header announces must be exactly the file's length */
            uint64_t want = sizeof hd + (uint64_t)c->sky.svf_bytes * c->n_total;
            if (hd.formats & SGPU_FMT_CSR) want += offs + (uint64_t)hd.nnz * sizeof(int32_t);
            if (hd.formats & SGPU_FMT_BARLE) want += offs + (uint64_t)hd.barle_bytes;
            need(std::fseek(in.f, 0, SEEK_END) == 0, SGPU_ERR_IO, "list file: cannot seek");
            const long long size = std::ftell(in.f);
            if (size < 0 || (uint64_t)size != want)
                fail(SGPU_ERR_FORMAT, "list file: %lld bytes, its header announces %llu", size,
                     (unsigned long long)want);
            need(std::fseek(in.f, (long)sizeof hd, SEEK_SET) == 0, SGPU_ERR_IO, "list file: cannot seek");
        }
        c->drop_products();
        Pinned buf;
        SGPU_HIP(hipHostMalloc(&buf.p, kPiece, hipHostMallocDefault));
        unsigned long long* flag = nullptr;
        try {
            flag = dalloc<unsigned long long>(1);
            c->svf = dalloc<uint8_t>((size_t)std::max(c->sky.svf_bytes, 1) * c->n_total);
            read_section(c, in.f, c->svf, (size_t)c->sky.svf_bytes * c->n_total, buf.p);
            const int64_t N = c->n_total;
            if (hd.formats & SGPU_FMT_CSR) {
                c->csr_off = dalloc<int64_t>(N + 1);
                c->csr_ids = dalloc<int32_t>((size_t)std::max<int64_t>(hd.nnz, 1));
                read_section(c, in.f, c->csr_off, offs, buf.p);
                read_section(c, in.f, c->csr_ids, (size_t)hd.nnz * sizeof(int32_t), buf.p);
                const int64_t* o = c->csr_off;
                const int32_t* ids = c->csr_ids;
                check_section(c, flag, "CSR offsets", [&] {
                    hipLaunchKernelGGL(k_check_offsets, dim3(check_blocks(N + 1)), dim3(256), 0, c->stream, o, N,
                                       hd.nnz, (int64_t)kDirs, flag);
                });
                check_section(c, flag, "CSR ids", [&] {
                    hipLaunchKernelGGL(k_check_ids, dim3(check_blocks(hd.nnz)), dim3(256), 0, c->stream, ids,
                                       hd.nnz, (int32_t)c->n_top, flag);
                });
                c->nnz = hd.nnz;
            }
            if (hd.formats & SGPU_FMT_BARLE) {
                c->br_off = dalloc<int64_t>(N + 1);
                c->br_bytes = dalloc<uint8_t>((size_t)std::max<int64_t>(hd.barle_bytes, 1));
                read_section(c, in.f, c->br_off, offs, buf.p);
                read_section(c, in.f, c->br_bytes, (size_t)hd.barle_bytes, buf.p);
                const int64_t* o = c->br_off;
                const uint8_t* bytes = c->br_bytes;
                check_section(c, flag, "ba-RLE offsets", [&] {
                    hipLaunchKernelGGL(k_check_offsets, dim3(check_blocks(N + 1)), dim3(256), 0, c->stream, o, N,
                                       hd.barle_bytes, (int64_t)1 << 20, flag);
                });
                check_section(c, flag, "ba-RLE streams", [&] {
                    hipLaunchKernelGGL(k_check_barle, dim3(check_blocks(N)), dim3(256), 0, c->stream, o, bytes, N,
                                       flag);
                });
                c->br_size = hd.barle_bytes;
            }
        } catch (...) {
            dfree(flag);
            c->drop_products();
            throw;
        }
        dfree(flag);
        c->svf_ok = true;
        c->info.nnz = hd.nnz;
        c->info.barle_bytes = hd.barle_bytes;
    });
}
