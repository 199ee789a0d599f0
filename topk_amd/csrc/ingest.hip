// Data ingest: the reference's text loader (File<T>, input/File.h) with the
// text parsed on the GPU, its writer, and a binary column file that loads
// straight into the column store.
//
// Text format (input/File.h:86-104 line_specifier, :264-295 read_scanf_t):
// one object per line, d values separated by ',', each value read as
// fscanf's "%f" reads it (leading white space skipped; decimal or exponent
// notation; inf, nan and hex floats), stored column-major x[m*n + i].
//
// Device pipeline over the text in HBM:
//   csv_count_kernel   newlines per 64 KiB tile (16-byte loads)
//   csv_scan_kernel    exclusive prefix over tiles (one workgroup)
//   csv_index_kernel   byte offset of every newline, in order
//   csv_rows_kernel    non-blank lines per 256-line group; scanned again
//   csv_parse_kernel   one thread per line over the workgroup's lines staged
//                      in LDS: fields to binary32, written column-major
//                      (coalesced across a wave's lines)
//   csv_patch_kernel   rows the host parsed (below) written into place
//
// Exactness: a field with at most 19 significant digits w and decimal
// exponent q, |q| <= 22, w < 2^53, is w*10^q or w/10^-q in one correctly
// rounded binary64 operation on exact operands.  Rounding that result to
// binary32 equals rounding the decimal directly unless an inexact binary64
// result lands exactly on a binary32 midpoint (every midpoint is a binary64
// value, and binary64 rounding is monotone, so the decimal and its binary64
// image lie on the same side of every other midpoint; an exact result is the
// decimal itself).  Such midpoints, subnormal or overflowing results, longer
// digit strings, inf / nan / hex and anything malformed go to the host,
// which parses the line with strtof, the conversion fscanf's "%f" uses.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace topk {

constexpr uint32_t CSV_BS = 256;
constexpr uint32_t CSV_TILE = 65536;             // text bytes per workgroup in the newline passes
constexpr uint32_t CSV_LDS_WORDS = 8192;        // text staged in LDS per parse workgroup (32 KiB)
constexpr uint32_t CSV_SCAN_BS = 1024;
constexpr uint64_t STAGE_BYTES = 64ull << 20;    // pinned staging buffer (two alternate)
constexpr uint32_t IO_THREADS = 16;

// Status of a runtime call; with TOPK_DEBUG set in the environment a
// failure also names the call on stderr.
static int ingest_status(hipError_t e, const char *what, int line) {
    if (e == hipSuccess) return TOPK_OK;
    if (getenv("TOPK_DEBUG")) fprintf(stderr, "topk ingest.hip:%d %s: %s\n", line, what, hipGetErrorString(e));
    return hip_status(e);
}
#define IK_ST(call) ingest_status((call), #call, __LINE__)
#define IK_HIP(call)                   \
    do {                               \
        const int r_ = IK_ST(call);    \
        if (r_) return r_;             \
    } while (0)

// A line the device parser leaves to the host.
struct CsvSlow {
    uint64_t start;  // byte offset of the line
    uint32_t len;    // bytes up to (not including) its newline
    uint32_t row;    // object index it fills
};

__constant__ double c_pow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                   1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// 0x80 in every byte of w that equals '\n', 0 elsewhere (no carries cross bytes).
__device__ __forceinline__ uint32_t nl_mask(uint32_t w) {
    const uint32_t x = w ^ 0x0A0A0A0Au;
    const uint32_t t = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return ~(t | x | 0x7F7F7F7Fu);
}

__host__ __device__ __forceinline__ bool csv_ws(uint32_t c) {
    return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

// The text is zero-padded to whole tiles, so no tile read needs a bound.
__global__ __launch_bounds__(CSV_BS) void csv_count_kernel(const uint4 *text, uint32_t *tile_nl) {
    const uint4 *t = text + (uint64_t)blockIdx.x * (CSV_TILE / 16);
    uint32_t c = 0;
#pragma unroll 4
    for (uint32_t j = threadIdx.x; j < CSV_TILE / 16; j += CSV_BS) {
        const uint4 v = t[j];
        c += __popc(nl_mask(v.x)) + __popc(nl_mask(v.y)) + __popc(nl_mask(v.z)) + __popc(nl_mask(v.w));
    }
    __shared__ uint32_t ws[CSV_BS / 64];
    c = wave_sum_u32(c);
    if (lane_id() == 0) ws[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < CSV_BS / 64; w++) s += ws[w];
        tile_nl[blockIdx.x] = s;
    }
}

// Exclusive prefix of cnt[0..m) into base[0..m], base[m] = total.  One
// workgroup; each thread sums a contiguous run.
__global__ __launch_bounds__(CSV_SCAN_BS) void csv_scan_kernel(const uint32_t *cnt, uint64_t m, uint64_t *base) {
    __shared__ uint64_t part[CSV_SCAN_BS];
    const uint64_t per = (m + CSV_SCAN_BS - 1) / CSV_SCAN_BS;
    const uint64_t b = threadIdx.x * per, e = min(b + per, m);
    uint64_t s = 0;
    for (uint64_t i = b; i < e; i++) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (uint32_t t = 0; t < CSV_SCAN_BS; t++) {
            const uint64_t v = part[t];
            part[t] = acc;
            acc += v;
        }
        base[m] = acc;
    }
    __syncthreads();
    uint64_t acc = part[threadIdx.x];
    for (uint64_t i = b; i < e; i++) {
        base[i] = acc;
        acc += cnt[i];
    }
}

// Writes the byte offset of every newline, in file order.  The tile is
// read in coalesced rounds of CSV_BS 16-byte words; a block scan per round
// orders the round's newlines (word order is file order).
__global__ __launch_bounds__(CSV_BS) void csv_index_kernel(const uint4 *text, const uint64_t *tile_base,
                                                           uint64_t *ends) {
    const uint4 *t = text + (uint64_t)blockIdx.x * (CSV_TILE / 16);
    __shared__ uint32_t wsum[CSV_BS / 64];
    __shared__ uint32_t total;
    uint64_t o = tile_base[blockIdx.x];
    for (uint32_t r = 0; r < CSV_TILE / 16 / CSV_BS; r++) {
        const uint32_t wi = r * CSV_BS + threadIdx.x;
        const uint4 v = t[wi];
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
        uint32_t mk[4], c = 0;
#pragma unroll
        for (int h = 0; h < 4; h++) {
            mk[h] = nl_mask(w4[h]);
            c += __popc(mk[h]);
        }
        const uint32_t ex = block_excl_scan<CSV_BS>(c, wsum, &total);
        uint64_t q = o + ex;
        const uint64_t byte0 = (uint64_t)blockIdx.x * CSV_TILE + (uint64_t)wi * 16;
#pragma unroll
        for (int h = 0; h < 4; h++) {
            uint32_t m = mk[h];
            while (m) {
                ends[q++] = byte0 + h * 4 + ((uint32_t)__builtin_ctz(m) >> 3);
                m &= m - 1;
            }
        }
        o += total;
    }
}

// Reads the text a byte at a time through aligned 32-bit loads, from HBM
// or from a copy of words [base, ...) in LDS.
struct ByteReader {
    const uint32_t *w;
    uint64_t cur = ~0ull;
    uint32_t word = 0;
    __device__ __forceinline__ uint32_t operator()(uint64_t p) {
        const uint64_t wi = p >> 2;
        if (wi != cur) {
            cur = wi;
            word = w[wi];
        }
        return (word >> ((p & 3) * 8)) & 0xFFu;
    }
};

struct LdsByteReader {
    const uint8_t *lds;  // LDS bytes holding text bytes [base, ...)
    uint64_t base;
    __device__ __forceinline__ uint32_t operator()(uint64_t p) const { return lds[(uint32_t)(p - base)]; }
};

// Line j spans [start, end): end is its newline, or the text's end for the
// line after the last newline.
__device__ __forceinline__ void csv_line(const uint64_t *ends, uint64_t nl, uint64_t size, uint64_t j,
                                         uint64_t &start, uint64_t &end) {
    start = j ? ends[j - 1] + 1 : 0;
    end = j < nl ? ends[j] : size;
}

template <class Reader>
__device__ __forceinline__ bool csv_nonblank(Reader &rd, uint64_t start, uint64_t end) {
    for (uint64_t p = start; p < end; p++)
        if (!csv_ws(rd(p))) return true;
    return false;
}

// Non-blank lines per group of CSV_BS lines (fscanf's "%f" skips blank lines).
__global__ __launch_bounds__(CSV_BS) void csv_rows_kernel(const uint32_t *text, const uint64_t *ends, uint64_t nl,
                                                          uint64_t size, uint32_t *grp_rows) {
    const uint64_t j = (uint64_t)blockIdx.x * CSV_BS + threadIdx.x;
    bool f = false;
    if (j <= nl) {
        uint64_t s, e;
        csv_line(ends, nl, size, j, s, e);
        ByteReader rd{text};
        f = csv_nonblank(rd, s, e);
    }
    const uint32_t c = __syncthreads_count(f);
    if (threadIdx.x == 0) grp_rows[blockIdx.x] = c;
}

// Converts one field (see the exactness note at the top); false sends the
// line to the host.
__device__ __forceinline__ bool csv_finish(uint64_t w, int q, bool neg, bool digits, float &out) {
    if (!digits) return false;
    if (w == 0) {
        out = neg ? -0.0f : 0.0f;
        return true;
    }
    if (w >= (1ull << 53) || q < -22 || q > 22) return false;
    const double wd = (double)w, p10 = c_pow10[q < 0 ? -q : q];
    const double x = q < 0 ? wd / p10 : wd * p10;
    if (!(x >= 0x1p-126) || x > 0x1.fffffep127) return false;  // binary32 subnormal or overflow
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    if ((b & 0x1FFFFFFFull) == 0x10000000ull) {
        // On a binary32 midpoint: fine if the binary64 operation was exact
        // (the residual, exactly representable, is zero), else the host decides.
        const double r = q < 0 ? fma(x, p10, -wd) : fma(wd, p10, -x);
        if (r != 0.0) return false;
    }
    const float f = (float)x;
    out = neg ? -f : f;
    return true;
}

// Parses line [start, end) into row `row`; false if the host must.  One
// straight pass per field (white space, sign, integer digits, fraction,
// exponent), so the lanes of a wave walk their lines in step.
template <class Reader>
__device__ __forceinline__ bool csv_parse_line(Reader &rd, uint64_t start, uint64_t end, uint32_t d, float *out,
                                               uint64_t cs, uint64_t rs, uint64_t row) {
    uint64_t p = start;
    auto at = [&](uint64_t i) -> uint32_t { return i < end ? rd(i) : (uint32_t)'\n'; };
    uint32_t c = at(p);
    for (uint32_t m = 0; m < d; m++) {
        while (csv_ws(c)) c = at(++p);
        bool neg = false;
        if (c == '+' || c == '-') {
            neg = c == '-';
            c = at(++p);
        }
        uint64_t w = 0;
        int q = 0, nd = 0;
        bool digits = false;
        for (uint32_t dig = c - '0'; dig < 10; dig = c - '0') {  // integer digits
            digits = true;
            if (w || dig) {
                if (nd < 19) {
                    w = w * 10 + dig;
                    nd++;
                } else {
                    q++;  // past 19 digits w >= 2^53: only the scale matters
                }
            }
            c = at(++p);
        }
        if (c == '.') {
            c = at(++p);
            for (uint32_t dig = c - '0'; dig < 10; dig = c - '0') {  // fraction digits
                digits = true;
                if (w || dig) {
                    if (nd < 19) {
                        w = w * 10 + dig;
                        nd++;
                        q--;
                    }
                } else {
                    q--;  // leading zeros scale only
                }
                c = at(++p);
            }
        }
        if (!digits) return false;  // empty field, inf / nan, or garbage
        if (c == 'e' || c == 'E') {
            c = at(++p);
            bool eneg = false;
            if (c == '+' || c == '-') {
                eneg = c == '-';
                c = at(++p);
            }
            if (c - '0' >= 10u) return false;
            int e = 0;
            for (uint32_t dig = c - '0'; dig < 10; dig = c - '0') {
                e = e < 100000 ? e * 10 + (int)dig : e;
                c = at(++p);
            }
            q += eneg ? -e : e;
        }
        float v;
        if (!csv_finish(w, q, neg, digits, v)) return false;
        out[m * cs + row * rs] = v;
        if (m + 1 < d) {
            if (c != ',') return false;  // also white space before the separator
            c = at(++p);
        }
    }
    while (csv_ws(c)) c = at(++p);
    return p >= end;
}

// One thread per line; row = rank of the line among the non-blank lines.
// Value m of row i goes to out[m*cs + i*rs].  A workgroup's lines are
// staged in LDS (coalesced 32-bit loads) when they fit in CSV_LDS_WORDS.
__global__ __launch_bounds__(CSV_BS) void csv_parse_kernel(const uint32_t *text, const uint64_t *ends, uint64_t nl,
                                                           uint64_t size, const uint64_t *row_base, uint64_t nwant,
                                                           uint32_t d, float *out, uint64_t cs, uint64_t rs,
                                                           unsigned long long *slow_cnt, CsvSlow *slow) {
    __shared__ uint32_t lds[CSV_LDS_WORDS];
    __shared__ uint32_t wsum[CSV_BS / 64];
    __shared__ uint32_t total;
    const uint64_t j0 = (uint64_t)blockIdx.x * CSV_BS;
    const uint64_t j = j0 + threadIdx.x;
    const uint64_t jl = min(j0 + CSV_BS - 1, nl);  // the group's last line
    uint64_t gs, ge;
    {
        uint64_t s0, e0, s1, e1;
        csv_line(ends, nl, size, j0, s0, e0);
        csv_line(ends, nl, size, jl, s1, e1);
        gs = s0;
        ge = e1;
    }
    const uint64_t wb = gs >> 2, we = (ge + 3) >> 2;
    const bool staged = we - wb <= CSV_LDS_WORDS;
    if (staged)
        for (uint64_t w = wb + threadIdx.x; w < we; w += CSV_BS) lds[(uint32_t)(w - wb)] = text[w];
    __syncthreads();
    uint64_t start = 0, end = 0;
    bool f = false;
    ByteReader grd{text};
    LdsByteReader lrd{reinterpret_cast<const uint8_t *>(lds), wb * 4};
    if (j <= nl) {
        csv_line(ends, nl, size, j, start, end);
        f = staged ? csv_nonblank(lrd, start, end) : csv_nonblank(grd, start, end);
    }
    const uint32_t rank = block_excl_scan<CSV_BS>(f ? 1u : 0u, wsum, &total);
    const uint64_t row = row_base[blockIdx.x] + rank;
    if (!f || row >= nwant) return;
    const bool ok = staged ? csv_parse_line(lrd, start, end, d, out, cs, rs, row)
                           : csv_parse_line(grd, start, end, d, out, cs, rs, row);
    if (!ok) {
        const unsigned long long s = atomicAdd(slow_cnt, 1ull);
        slow[s] = CsvSlow{start, (uint32_t)(end - start), (uint32_t)row};
    }
}

// Writes the host-parsed rows: vals[s*d + m] is value m of row rows[s].
__global__ void csv_patch_kernel(const uint32_t *rows, const float *vals, uint32_t cnt, uint32_t d, float *out,
                                 uint64_t cs, uint64_t rs) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)cnt * d) return;
    const uint32_t s = (uint32_t)(t / d), m = (uint32_t)(t % d);
    out[m * cs + (uint64_t)rows[s] * rs] = vals[t];
}

// ---------------------------------------------------------------- host side

// Parses one line the way fscanf "%f,%f,...,%f" reads a row: strtof per
// field (it skips leading white space), a ',' right after each value but the
// last, then only white space.  false if the line is malformed.
static bool parse_line_host(const char *p, size_t len, uint32_t d, float *out) {
    std::string s(p, len);  // NUL-terminated copy for strtof
    const char *c = s.c_str();
    for (uint32_t m = 0; m < d; m++) {
        if (m > 0) {
            if (*c != ',') return false;
            c++;
        }
        char *q = nullptr;
        const float v = strtof(c, &q);
        if (q == c) return false;
        out[m] = v;
        c = q;
    }
    while (*c && csv_ws((unsigned char)*c)) c++;
    return *c == 0;
}

struct MappedFile {
    int fd = -1;
    const char *map = nullptr;
    uint64_t size = 0;
    int open(const char *path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return TOPK_ERR_IO;
        struct stat st;
        if (fstat(fd, &st) != 0) return TOPK_ERR_IO;
        size = (uint64_t)st.st_size;
        if (size) {
            void *m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m == MAP_FAILED) return TOPK_ERR_IO;
            map = static_cast<const char *>(m);
        }
        return TOPK_OK;
    }
    ~MappedFile() {
        if (map) munmap(const_cast<char *>(map), size);
        if (fd >= 0) ::close(fd);
    }
};

// Reads [off, off+bytes) of fd with IO_THREADS preads.
static bool pread_all(int fd, char *dst, uint64_t off, uint64_t bytes) {
    std::atomic<bool> ok{true};
    const uint64_t per = (bytes + IO_THREADS - 1) / IO_THREADS;
    auto work = [&](uint64_t b, uint64_t e) {
        while (b < e) {
            const ssize_t r = ::pread(fd, dst + b, e - b, (off_t)(off + b));
            if (r <= 0) {
                if (r < 0 && errno == EINTR) continue;
                ok = false;
                return;
            }
            b += (uint64_t)r;
        }
    };
    if (bytes < (4ull << 20)) {
        work(0, bytes);
        return ok;
    }
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < IO_THREADS; t++) {
        const uint64_t b = t * per, e = std::min(bytes, b + per);
        if (b < e) th.emplace_back(work, b, e);
    }
    for (auto &t : th) t.join();
    return ok;
}

// Streams file bytes to device memory through two pinned buffers: the
// read of one chunk overlaps the copy of the previous one.
struct Uploader {
    char *stage[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    int init(hipStream_t s) {
        st = s;
        for (int i = 0; i < 2; i++) {
            IK_HIP(hipHostMalloc(&stage[i], STAGE_BYTES, hipHostMallocDefault));
            IK_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
        }
        return TOPK_OK;
    }
    // Appends [off, off+bytes) of fd to dst (device).
    int copy(int fd, uint64_t off, uint64_t bytes, char *dst, int &slot) {
        for (uint64_t b = 0; b < bytes; b += STAGE_BYTES) {
            const uint64_t c = std::min<uint64_t>(STAGE_BYTES, bytes - b);
            IK_HIP(hipEventSynchronize(done[slot]));
            if (!pread_all(fd, stage[slot], off + b, c)) return TOPK_ERR_IO;
            IK_HIP(hipMemcpyAsync(dst + b, stage[slot], c, hipMemcpyHostToDevice, st));
            IK_HIP(hipEventRecord(done[slot], st));
            slot ^= 1;
        }
        return TOPK_OK;
    }
    ~Uploader() {
        if (st) hipStreamSynchronize(st);
        for (int i = 0; i < 2; i++) {
            if (stage[i]) hipHostFree(stage[i]);
            if (done[i]) hipEventDestroy(done[i]);
        }
    }
};

static double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Fields of the first non-blank line (commas + 1, as File::count,
// input/File.h:114-119); 0 if the text holds none.
static uint32_t first_line_fields(const char *t, uint64_t size) {
    uint64_t p = 0;
    while (p < size) {
        uint64_t e = p;
        bool blank = true;
        uint32_t commas = 0;
        while (e < size && t[e] != '\n') {
            if (!csv_ws((unsigned char)t[e])) blank = false;
            if (t[e] == ',') commas++;
            e++;
        }
        if (!blank) return commas + 1;
        p = e + 1;
    }
    return 0;
}

// The text of one file in HBM with its line index.
struct CsvText {
    MappedFile f;
    hipStream_t st = nullptr;
    char *text = nullptr;       // size rounded up to whole tiles, zero padded
    uint64_t *ends = nullptr;   // newline offsets
    uint64_t *row_base = nullptr;  // exclusive prefix of non-blank lines per group
    uint64_t ntiles = 0, nl = 0, groups = 0, rows = 0;
    uint32_t fields = 0;
    double ms_read = 0, ms_index = 0;

    ~CsvText() {
        if (st) hipStreamSynchronize(st);
        hipFree(text);
        hipFree(ends);
        hipFree(row_base);
    }

    int load(const char *path, hipStream_t s) {
        st = s;
        int rc = f.open(path);
        if (rc) return rc;
        fields = first_line_fields(f.map, f.size);
        if (!fields) return TOPK_ERR_FORMAT;
        if (fields > TOPK_MAX_DIMS) return TOPK_ERR_UNSUPPORTED;
        auto t0 = std::chrono::steady_clock::now();
        ntiles = (f.size + CSV_TILE - 1) / CSV_TILE;
        const uint64_t cap = ntiles * CSV_TILE;
        IK_HIP(hipMalloc(&text, cap));
        IK_HIP(hipMemsetAsync(text + f.size, 0, cap - f.size, st));
        {
            Uploader up;
            rc = up.init(st);
            int slot = 0;
            if (!rc) rc = up.copy(f.fd, 0, f.size, text, slot);
            if (rc) return rc;
        }  // drains the stream
        ms_read = ms_since(t0);
        t0 = std::chrono::steady_clock::now();
        uint32_t *tile_nl = nullptr;
        uint64_t *tile_base = nullptr;
        IK_HIP(hipMalloc(&tile_nl, sizeof(uint32_t) * ntiles));
        IK_HIP(hipMalloc(&tile_base, sizeof(uint64_t) * (ntiles + 1)));
        hipLaunchKernelGGL(csv_count_kernel, dim3((uint32_t)ntiles), dim3(CSV_BS), 0, st, (const uint4 *)text,
                           tile_nl);
        hipLaunchKernelGGL(csv_scan_kernel, dim3(1), dim3(CSV_SCAN_BS), 0, st, tile_nl, ntiles, tile_base);
        rc = IK_ST(hipGetLastError());
        if (!rc) rc = IK_ST(hipMemcpyAsync(&nl, tile_base + ntiles, sizeof(nl), hipMemcpyDeviceToHost, st));
        if (!rc) rc = IK_ST(hipStreamSynchronize(st));
        if (!rc && hipMalloc(&ends, sizeof(uint64_t) * (nl + 1)) != hipSuccess) rc = TOPK_ERR_NOMEM;
        if (!rc) {
            hipLaunchKernelGGL(csv_index_kernel, dim3((uint32_t)ntiles), dim3(CSV_BS), 0, st, (const uint4 *)text,
                               tile_base, ends);
            rc = IK_ST(hipGetLastError());
        }
        if (!rc) rc = IK_ST(hipStreamSynchronize(st));
        hipFree(tile_nl);
        hipFree(tile_base);
        if (rc) return rc;
        // lines: nl terminated ones plus the text after the last newline
        groups = (nl + 1 + CSV_BS - 1) / CSV_BS;
        uint32_t *grp = nullptr;
        IK_HIP(hipMalloc(&grp, sizeof(uint32_t) * groups));
        IK_HIP(hipMalloc(&row_base, sizeof(uint64_t) * (groups + 1)));
        hipLaunchKernelGGL(csv_rows_kernel, dim3((uint32_t)groups), dim3(CSV_BS), 0, st, (const uint32_t *)text,
                           ends, nl, f.size, grp);
        hipLaunchKernelGGL(csv_scan_kernel, dim3(1), dim3(CSV_SCAN_BS), 0, st, grp, groups, row_base);
        rc = IK_ST(hipGetLastError());
        if (!rc) rc = IK_ST(hipMemcpyAsync(&rows, row_base + groups, sizeof(rows), hipMemcpyDeviceToHost, st));
        if (!rc) rc = IK_ST(hipStreamSynchronize(st));
        hipFree(grp);
        ms_index = ms_since(t0);
        return rc;
    }

    // Parses the first n rows (d fields each) into out[m*cs + i*rs] (device).
    int parse(uint64_t n, uint32_t d, float *out, uint64_t cs, uint64_t rs, uint64_t &slow_lines, double &ms) {
        auto t0 = std::chrono::steady_clock::now();
        unsigned long long *cnt = nullptr;
        CsvSlow *slow = nullptr;
        const uint64_t cap = std::min<uint64_t>(n, nl + 1);
        IK_HIP(hipMalloc(&cnt, sizeof(unsigned long long)));
        int rc = IK_ST(hipMalloc(&slow, sizeof(CsvSlow) * cap));
        if (!rc) rc = IK_ST(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), st));
        if (!rc) {
            hipLaunchKernelGGL(csv_parse_kernel, dim3((uint32_t)groups), dim3(CSV_BS), 0, st,
                               (const uint32_t *)text, ends, nl, f.size, row_base, n, d, out, cs, rs, cnt, slow);
            rc = IK_ST(hipGetLastError());
        }
        unsigned long long hs = 0;
        if (!rc) rc = IK_ST(hipMemcpyAsync(&hs, cnt, sizeof(hs), hipMemcpyDeviceToHost, st));
        if (!rc) rc = IK_ST(hipStreamSynchronize(st));
        slow_lines = hs;
        if (!rc && hs) rc = patch(slow, hs, d, out, cs, rs);
        hipFree(cnt);
        hipFree(slow);
        ms = ms_since(t0);
        return rc;
    }

    // Parses the lines the device left, on host threads, and writes them.
    int patch(const CsvSlow *d_slow, uint64_t cnt, uint32_t d, float *out, uint64_t cs, uint64_t rs) {
        std::vector<CsvSlow> sl(cnt);
        IK_HIP(hipMemcpy(sl.data(), d_slow, sizeof(CsvSlow) * cnt, hipMemcpyDeviceToHost));
        std::vector<float> vals(cnt * d);
        std::vector<uint32_t> rows_h(cnt);
        std::atomic<bool> ok{true};
        auto work = [&](uint64_t b, uint64_t e) {
            for (uint64_t s = b; s < e && ok; s++) {
                rows_h[s] = sl[s].row;
                if (!parse_line_host(f.map + sl[s].start, sl[s].len, d, &vals[s * d])) ok = false;
            }
        };
        const uint32_t nt = cnt < 4096 ? 1u : 16u;
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < nt; t++) {
            const uint64_t b = cnt * t / nt, e = cnt * (t + 1) / nt;
            th.emplace_back(work, b, e);
        }
        for (auto &t : th) t.join();
        if (!ok) return TOPK_ERR_FORMAT;
        float *dv = nullptr;
        uint32_t *dr = nullptr;
        IK_HIP(hipMalloc(&dv, sizeof(float) * vals.size()));
        int rc = IK_ST(hipMalloc(&dr, sizeof(uint32_t) * cnt));
        if (!rc) rc = IK_ST(hipMemcpyAsync(dv, vals.data(), sizeof(float) * vals.size(), hipMemcpyHostToDevice, st));
        if (!rc) rc = IK_ST(hipMemcpyAsync(dr, rows_h.data(), sizeof(uint32_t) * cnt, hipMemcpyHostToDevice, st));
        if (!rc) {
            const uint64_t tot = cnt * d;
            hipLaunchKernelGGL(csv_patch_kernel, dim3((uint32_t)((tot + 255) / 256)), dim3(256), 0, st, dr, dv,
                               (uint32_t)cnt, d, out, cs, rs);
            rc = IK_ST(hipGetLastError());
        }
        if (!rc) rc = IK_ST(hipStreamSynchronize(st));
        hipFree(dv);
        hipFree(dr);
        return rc;
    }
};

// Binary column file: a 4096-byte header, then d columns of n binary32
// values (little-endian), column m at data_offset + m*n*4.
constexpr char COL_MAGIC[8] = {'T', 'O', 'P', 'K', 'C', 'O', 'L', '1'};
constexpr uint64_t COL_HEADER = 4096;
struct ColHeader {
    char magic[8];
    uint32_t version;
    uint32_t d;
    uint64_t n;
    uint64_t data_offset;
};

static int read_col_header(int fd, ColHeader &h, uint64_t fsize) {
    if (::pread(fd, &h, sizeof(h), 0) != (ssize_t)sizeof(h)) return TOPK_ERR_FORMAT;
    if (memcmp(h.magic, COL_MAGIC, 8) != 0 || h.version != 1 || h.d < 1 || h.d > TOPK_MAX_DIMS || h.n < 1)
        return TOPK_ERR_FORMAT;
    // (no 64-bit wrap for a corrupt header: n is checked against what the file can hold)
    if (h.data_offset < sizeof(h) || h.data_offset > fsize || h.n > (fsize - h.data_offset) / (4ull * h.d))
        return TOPK_ERR_FORMAT;
    return TOPK_OK;
}

static void fill_report(topk_ingest_report *rep, int format, uint64_t bytes, uint64_t lines, uint64_t rows,
                        uint32_t fields, uint64_t slow, double ms_read, double ms_parse, double ms_total) {
    if (!rep) return;
    rep->format = format;
    rep->fields = fields;
    rep->bytes = bytes;
    rep->lines = lines;
    rep->rows = rows;
    rep->slow_lines = slow;
    rep->ms_read = ms_read;
    rep->ms_parse = ms_parse;
    rep->ms_total = ms_total;
}

// Text to device rows: checks the requested shape against the file.
static int csv_shape(const CsvText &t, uint64_t &n, uint32_t &d) {
    if (d == 0) d = t.fields;
    if (d != t.fields) return TOPK_ERR_FORMAT;
    if (n == 0) n = t.rows;
    if (n == 0 || n > t.rows) return TOPK_ERR_FORMAT;
    if (n >= 0xFFFFFFFFull) return TOPK_ERR_UNSUPPORTED;  // as topk_ctx_create: ids are 32-bit
    return TOPK_OK;
}

}  // namespace topk

using namespace topk;

static int check_device_ingest() {
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) return TOPK_ERR_NO_DEVICE;
    return TOPK_OK;
}

// A file descriptor closed on scope exit.
struct FileFd {
    int fd = -1;
    ~FileFd() {
        if (fd >= 0) ::close(fd);
    }
};

// col = the file starts with the column-file magic (else it is read as text).
static int sniff_format(const char *path, bool &col) {
    FileFd f;
    f.fd = ::open(path, O_RDONLY);
    if (f.fd < 0) return TOPK_ERR_IO;
    char m[8] = {0};
    col = ::pread(f.fd, m, sizeof(m), 0) == (ssize_t)sizeof(m) && memcmp(m, COL_MAGIC, 8) == 0;
    return TOPK_OK;
}

// Column file -> the store of ctx (first n objects, first d attributes).
static int load_colfile(topk_ctx *ctx, const char *path, uint64_t n, uint32_t d, topk_ingest_report *rep,
                        std::chrono::steady_clock::time_point t0) {
    FileFd f;
    f.fd = ::open(path, O_RDONLY);
    if (f.fd < 0) return TOPK_ERR_IO;
    struct stat st;
    if (fstat(f.fd, &st) != 0) return TOPK_ERR_IO;
    ColHeader h;
    int rc = read_col_header(f.fd, h, (uint64_t)st.st_size);
    if (rc) return rc;
    if (d == 0) d = h.d;
    if (n == 0) n = h.n;
    if (d > h.d || n > h.n || n >= 0xFFFFFFFFull) return TOPK_ERR_FORMAT;
    ctx->n = n;
    ctx->d = d;
    rc = ctx_alloc(ctx);
    if (!rc) {
        Uploader up;
        rc = up.init(ctx->stream);
        int slot = 0;
        for (uint32_t m = 0; m < d && !rc; m++)
            rc = up.copy(f.fd, h.data_offset + (uint64_t)m * h.n * 4, n * 4,
                         reinterpret_cast<char *>(ctx->cols + (uint64_t)m * ctx->ld), slot);
    }  // drains the copies
    const double ms_read = ms_since(t0);
    if (!rc) rc = check_finite(ctx);
    if (!rc) fill_report(rep, 1, (uint64_t)st.st_size, 0, n, d, 0, ms_read, 0, ms_since(t0));
    return rc;
}

// Text file -> uploaded, indexed and parsed into the store of ctx.
static int load_text(topk_ctx *ctx, const char *path, uint64_t n, uint32_t d, topk_ingest_report *rep,
                     std::chrono::steady_clock::time_point t0) {
    hipStream_t ls;
    int rc = IK_ST(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking));
    if (rc) return rc;
    {
        CsvText t;  // released (its stream drained) before the stream is destroyed
        rc = t.load(path, ls);
        if (!rc) rc = csv_shape(t, n, d);
        if (!rc) {
            ctx->n = n;
            ctx->d = d;
            rc = ctx_alloc(ctx);
        }
        uint64_t slow = 0;
        double ms_parse = 0;
        if (!rc) rc = IK_ST(hipStreamSynchronize(ctx->stream));  // the store is zeroed first
        if (!rc) rc = t.parse(n, d, ctx->cols, ctx->ld, 1, slow, ms_parse);
        if (!rc) rc = check_finite(ctx);
        if (!rc)
            fill_report(rep, 0, t.f.size, t.nl + (t.f.size && t.f.map[t.f.size - 1] != '\n' ? 1 : 0), n, d, slow,
                        t.ms_read, t.ms_index + ms_parse, ms_since(t0));
    }
    (void)hipStreamDestroy(ls);
    return rc;
}

extern "C" {

int topk_csv_count(const char *path, uint64_t *rows, uint32_t *fields) {
    if (!path || !rows || !fields) return TOPK_ERR_INVALID;
    MappedFile f;
    int rc = f.open(path);
    if (rc) return rc;
    const uint32_t fl = f.size ? first_line_fields(f.map, f.size) : 0;
    uint64_t r = 0;
    uint64_t p = 0;
    while (p < f.size) {
        const char *nlp = static_cast<const char *>(memchr(f.map + p, '\n', f.size - p));
        const uint64_t e = nlp ? (uint64_t)(nlp - f.map) : f.size;
        for (uint64_t i = p; i < e; i++)
            if (!csv_ws((unsigned char)f.map[i])) {
                r++;
                break;
            }
        p = e + 1;
    }
    *rows = r;
    *fields = fl;
    return TOPK_OK;
}

int topk_csv_read(const char *path, uint64_t n, uint32_t d, uint32_t flags, float *out, topk_ingest_report *rep) {
    if (!path || !out || n >= 0xFFFFFFFFull || d > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
    int rc = check_device_ingest();
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st;
    IK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    float *dout = nullptr;
    uint64_t slow = 0;
    double ms_parse = 0;
    CsvText *t = new CsvText();
    rc = t->load(path, st);
    if (!rc) rc = csv_shape(*t, n, d);
    if (!rc && hipMalloc(&dout, sizeof(float) * n * d) != hipSuccess) rc = TOPK_ERR_NOMEM;
    if (!rc) {
        const bool rowmajor = flags & TOPK_ROWMAJOR;
        rc = t->parse(n, d, dout, rowmajor ? 1 : n, rowmajor ? d : 1, slow, ms_parse);
    }
    if (!rc) rc = IK_ST(hipMemcpyAsync(out, dout, sizeof(float) * n * d, hipMemcpyDeviceToHost, st));
    if (!rc) rc = IK_ST(hipStreamSynchronize(st));
    if (!rc)
        fill_report(rep, 0, t->f.size, t->nl + (t->f.size && t->f.map[t->f.size - 1] != '\n' ? 1 : 0), n, d, slow,
                    t->ms_read, t->ms_index + ms_parse, ms_since(t0));
    delete t;
    hipFree(dout);
    hipStreamDestroy(st);
    return rc;
}

int topk_csv_write(const char *path, const float *data, uint64_t n, uint32_t d, uint32_t flags) {
    if (!path || !data || n < 1 || d < 1 || d > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
    const bool rowmajor = flags & TOPK_ROWMAJOR;
    const uint32_t nt = n < 65536 ? 1u : 16u;
    std::vector<std::string> part(nt);
    auto work = [&](uint32_t t) {
        const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
        std::string &s = part[t];
        s.reserve((e - b) * d * 10);
        char buf[64];
        for (uint64_t i = b; i < e; i++) {
            for (uint32_t m = 0; m < d; m++) {
                const float v = rowmajor ? data[i * d + m] : data[(uint64_t)m * n + i];
                const int l = snprintf(buf, sizeof(buf), m + 1 < d ? "%f," : "%f\n", (double)v);
                if (l > 0 && l < (int)sizeof(buf)) {
                    s.append(buf, (size_t)l);
                } else {  // very large magnitudes: let the string grow
                    std::vector<char> big((size_t)l + 1);
                    snprintf(big.data(), big.size(), m + 1 < d ? "%f," : "%f\n", (double)v);
                    s.append(big.data(), (size_t)l);
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < nt; t++) th.emplace_back(work, t);
    for (auto &t : th) t.join();
    FILE *f = fopen(path, "wb");
    if (!f) return TOPK_ERR_IO;
    bool ok = true;
    for (auto &s : part) ok = ok && fwrite(s.data(), 1, s.size(), f) == s.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? TOPK_OK : TOPK_ERR_IO;
}

int topk_colfile_write(const char *path, const float *data, uint64_t n, uint32_t d, uint32_t flags) {
    if (!path || !data || n < 1 || d < 1 || d > TOPK_MAX_DIMS) return TOPK_ERR_INVALID;
    FILE *f = fopen(path, "wb");
    if (!f) return TOPK_ERR_IO;
    std::vector<char> hdr(COL_HEADER, 0);
    ColHeader h;
    memcpy(h.magic, COL_MAGIC, 8);
    h.version = 1;
    h.d = d;
    h.n = n;
    h.data_offset = COL_HEADER;
    memcpy(hdr.data(), &h, sizeof(h));
    bool ok = fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    if (flags & TOPK_ROWMAJOR) {
        std::vector<float> col(n);
        for (uint32_t m = 0; m < d && ok; m++) {
            for (uint64_t i = 0; i < n; i++) col[i] = data[i * d + m];
            ok = fwrite(col.data(), sizeof(float), n, f) == n;
        }
    } else {
        ok = ok && fwrite(data, sizeof(float), n * d, f) == n * d;
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? TOPK_OK : TOPK_ERR_IO;
}

int topk_colfile_info(const char *path, uint64_t *n, uint32_t *d) {
    if (!path || !n || !d) return TOPK_ERR_INVALID;
    FileFd f;
    f.fd = ::open(path, O_RDONLY);
    if (f.fd < 0) return TOPK_ERR_IO;
    struct stat st;
    if (fstat(f.fd, &st) != 0) return TOPK_ERR_IO;
    ColHeader h;
    const int rc = read_col_header(f.fd, h, (uint64_t)st.st_size);
    if (rc) return rc;
    *n = h.n;
    *d = h.d;
    return TOPK_OK;
}

int topk_ctx_create_file(topk_ctx **out, const char *path, uint64_t n, uint32_t d, uint32_t flags,
                         topk_ingest_report *rep) {
    if (!out || !path || n >= 0xFFFFFFFFull || d > TOPK_MAX_DIMS || flags) return TOPK_ERR_INVALID;
    *out = nullptr;
    int rc = check_device_ingest();
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    bool col = false;
    rc = sniff_format(path, col);
    if (rc) return rc;
    topk_ctx *ctx = new topk_ctx();
    hipGetDevice(&ctx->device);
    rc = col ? load_colfile(ctx, path, n, d, rep, t0) : load_text(ctx, path, n, d, rep, t0);
    if (rc) {
        if (ctx->stream) topk_ctx_destroy(ctx);
        else delete ctx;
        return rc;
    }
    *out = ctx;
    return TOPK_OK;
}

}  // extern "C"
