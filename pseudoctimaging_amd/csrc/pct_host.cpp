// Host-side part of libpct.so: pre/post-processing around the fusion hot
// path, the MatrixIO text format, the configuration file of the reference
// program, and seeded input generators.  None of this computes the fusion.
#include <omp.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "pct_internal.h"

namespace pct {

uint64_t splitmix64(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void parallel_copy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kSerial = (size_t)4 << 20;
    if (bytes < kSerial) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const int nt = std::max(1, omp_get_max_threads());
    const size_t part = ((bytes + nt - 1) / nt + 4095) & ~(size_t)4095;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; ++t) {
        const size_t o = (size_t)t * part;
        if (o < bytes)
            std::memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o,
                        std::min(part, bytes - o));
    }
}

namespace {

// Counter-based stream: value number `i` of stream `key`.  Independent of
// evaluation order, so generation can be parallel and stays bit-identical.
inline uint64_t hash_u64(uint64_t key, uint64_t i) {
    uint64_t s = key ^ (i * 0xD1B54A32D192ED03ULL);
    return splitmix64(s);
}

inline double uniform01(uint64_t key, uint64_t i) {  // [0, 1)
    return (double)(hash_u64(key, i) >> 11) * (1.0 / 9007199254740992.0);
}

inline double normal01(uint64_t key, uint64_t i) {  // Box-Muller, one variate per index
    const double u1 = ((double)(hash_u64(key, 2 * i) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(hash_u64(key, 2 * i + 1) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

inline uint64_t derive_key(uint64_t seed, uint64_t a, uint64_t b) {
    uint64_t s = seed ^ (0xA0761D6478BD642FULL * (a + 1)) ^ (0xE7037ED1A0B428DBULL * (b + 1));
    splitmix64(s);
    return splitmix64(s);
}

inline int mirror(int i, int n) {  // symmetric reflection without edge repeat
    if (n == 1) return 0;
    const int period = 2 * n - 2;
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - i;
}

// Separable Gaussian blur, sigma 3 voxels, taps to 4 sigma, in place.
void gaussian_blur(std::vector<double>& f, int nx, int ny, int nz) {
    const double sigma = 3.0;
    const int R = 12;
    double taps[2 * R + 1];
    double sum = 0;
    for (int k = -R; k <= R; ++k) sum += taps[k + R] = std::exp(-(double)(k * k) / (2 * sigma * sigma));
    for (double& t : taps) t /= sum;
    std::vector<double> tmp(f.size());
    const long long pl = (long long)nx * ny;
    // y
#pragma omp parallel for schedule(static)
    for (long long row = 0; row < (long long)nz * nx; ++row)
        for (int y = 0; y < ny; ++y) {
            double acc = 0;
            for (int k = -R; k <= R; ++k) acc += taps[k + R] * f[row * ny + mirror(y + k, ny)];
            tmp[row * ny + y] = acc;
        }
    // x
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                double acc = 0;
                for (int k = -R; k <= R; ++k)
                    acc += taps[k + R] * tmp[z * pl + (long long)mirror(x + k, nx) * ny + y];
                f[z * pl + (long long)x * ny + y] = acc;
            }
    // z
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (long long xy = 0; xy < pl; ++xy) {
            double acc = 0;
            for (int k = -R; k <= R; ++k) acc += taps[k + R] * f[mirror(z + k, nz) * pl + xy];
            tmp[z * pl + xy] = acc;
        }
    f.swap(tmp);
}

// Tissue classes 0..3 from blurred noise thresholded at its quartiles.
std::vector<unsigned char> tissue_classes(uint64_t key, int nx, int ny, int nz) {
    const long long n = (long long)nx * ny * nz;
    std::vector<double> f(n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) f[i] = normal01(key, (uint64_t)i);
    gaussian_blur(f, nx, ny, nz);
    std::vector<double> s(f);
    auto kth = [&](long long k) {
        std::nth_element(s.begin(), s.begin() + k, s.end());
        return s[k];
    };
    const double q25 = kth(n / 4), q50 = kth(n / 2), q75 = kth((3 * n) / 4);
    std::vector<unsigned char> c(n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i)
        c[i] = (unsigned char)((f[i] > q25) + (f[i] > q50) + (f[i] > q75));
    return c;
}

const double kMriMean[4] = {0.0, 0.8, 0.5, 0.2};
const double kCtMean[4] = {-1000.0, -100.0, 40.0, 1000.0};

void fill_intensities(const std::vector<unsigned char>& cls, uint64_t key, long long n, float* mri,
                      float* ct) {
    const uint64_t kmri = derive_key(key, 101, 0), kct = derive_key(key, 102, 0);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        const int c = cls[i];
        if (mri) mri[i] = (float)(kMriMean[c] + 0.03 * normal01(kmri, (uint64_t)i));
        if (ct) ct[i] = (float)(kCtMean[c] + 20.0 * normal01(kct, (uint64_t)i));
    }
}

}  // namespace
}  // namespace pct

using namespace pct;

extern "C" const char* pct_strerror(int status) {
    switch (status) {
        case PCT_OK: return "ok";
        case PCT_ERR_INVALID: return "invalid argument";
        case PCT_ERR_NO_DEVICE: return "no gfx950 HIP device available";
        case PCT_ERR_HIP: return "HIP runtime error";
        case PCT_ERR_NO_MEMORY: return "out of memory";
        case PCT_ERR_IO: return "file could not be opened or written";
        case PCT_ERR_FORMAT: return "malformed file";
        case PCT_ERR_DIMS: return "dimensions do not match";
        case PCT_ERR_UNSUPPORTED: return "no kernel configuration fits the request";
        case PCT_ERR_TOO_LARGE: return "request exceeds a documented maximum";
        default: return "unknown status";
    }
}

extern "C" int pct_abi_version(void) { return PCT_ABI_VERSION; }

extern "C" double pct_weight_divisor(double sigma, double beta) {
    // cpp:1601 - (2 * 27 * beta*pow(sigma, 2)), evaluated left to right
    return (2 * 27 * beta * std::pow(sigma, 2));
}

extern "C" void pct_normalize_labels(float* label, size_t n) {
    if (!label) return;
    for (size_t i = 0; i < n; ++i) label[i] = 255.0f * ((label[i] + 1024.0f) / 4095.0f);
}

// ---- post-processing ------------------------------------------------------

namespace {
// Median of the non-NaN values (cpp:1020-1043): stable sort, middle value or
// float64 mean of the two middle values.  NaN when every value is NaN.
double median27(const float* patch) {
    float v[27];
    int cnt = 0;
    for (int j = 0; j < 27; ++j)
        if (!std::isnan(patch[j])) v[cnt++] = patch[j];
    if (cnt == 0) return NAN;
    std::stable_sort(v, v + cnt);
    if (cnt % 2 == 0) return ((double)v[cnt / 2 - 1] + (double)v[cnt / 2]) / 2;
    return (double)v[cnt / 2];
}
}  // namespace

extern "C" int pct_regularize(int nh, int nx, int ny, int nz, float* arr) {
    if (!arr || nh < 1 || nx < 1 || ny < 1 || nz < 1) return PCT_ERR_INVALID;
    const int hp = 1, hw = (nh - 1) / 2, lo = hw + hp;
    const long long pl = (long long)nx * ny;
    // A volume without an interior NaN leaves the sweep after one pass that
    // changes nothing; find that out in parallel first.
    int any_nan = 0;
#pragma omp parallel for schedule(static) reduction(| : any_nan)
    for (int z = lo; z < nz - lo; ++z)
        for (int x = lo; x < nx - lo; ++x) {
            const float* row = arr + z * pl + (long long)x * ny;
            for (int y = lo; y < ny - lo; ++y) any_nan |= std::isnan(row[y]) ? 1 : 0;
        }
    if (!any_nan) return 1;
    int passes = 0;
    bool again = true;
    while (again) {
        again = false;
        long long filled = 0;
        ++passes;
        for (int z = lo; z < nz - lo; ++z)
            for (int x = lo; x < nx - lo; ++x)
                for (int y = lo; y < ny - lo; ++y) {
                    const long long id = z * pl + (long long)x * ny + y;
                    if (!std::isnan(arr[id])) continue;
                    float patch[27];
                    int c = 0;
                    for (int r = -1; r <= 1; ++r)
                        for (int s = -1; s <= 1; ++s)
                            for (int t = -1; t <= 1; ++t)
                                patch[c++] = arr[id + r * pl + (long long)s * ny + t];
                    arr[id] = (float)median27(patch);
                    if (std::isnan(arr[id]))
                        again = true;
                    else
                        ++filled;
                }
        if (again && filled == 0) break;
    }
    return passes;
}

// Both are elementwise, so the parallel loops give the sequential result.
extern "C" void pct_round(float* arr, size_t n) {
    if (!arr) return;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) arr[i] = std::round(arr[i]);
}

extern "C" void pct_to_hu(float* arr, size_t n) {
    if (!arr) return;
    const float max_labels = 4095.0f, min_labels = 1024.0f;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i)
        arr[i] = std::round(max_labels * (arr[i] / 255.0f)) - min_labels;
}

extern "C" int pct_postprocess(int nh, int nx, int ny, int nz, float* arr) {
    int rc = pct_regularize(nh, nx, ny, nz, arr);
    if (rc < 0) return rc;
    // RoundMatrix then the HU rescale, in one pass with the same operations
    const long long n = (long long)nx * ny * nz;
    const float max_labels = 4095.0f, min_labels = 1024.0f;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i)
        arr[i] = std::round(max_labels * (std::round(arr[i]) / 255.0f)) - min_labels;
    return PCT_OK;
}

// ---- MatrixIO --------------------------------------------------------------

extern "C" void pct_free(void* p) { std::free(p); }

namespace {

bool is_blank(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// The whole file in memory, with a terminating NUL.
int slurp(const char* path, std::vector<char>* buf) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return PCT_ERR_IO;
    struct stat st;
    std::vector<char>& b = *buf;
    b.clear();
    if (::fstat(fileno(f), &st) == 0 && S_ISREG(st.st_mode)) {
        b.resize((size_t)st.st_size);
        const size_t got = std::fread(b.data(), 1, b.size(), f);
        b.resize(got);
    }
    char tmp[1 << 16];   // anything beyond the size at open (or a non-regular file)
    size_t got;
    while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) b.insert(b.end(), tmp, tmp + got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) return PCT_ERR_IO;
    b.push_back('\0');
    return PCT_OK;
}

// One value token [p, e) as the reference's `iss >> float` reads it
// (correctly rounded decimal; an optional leading '+').
bool parse_float(const char* p, const char* e, float* v) {
    if (p < e && *p == '+') {
        ++p;
        if (p < e && (*p == '+' || *p == '-')) return false;
    }
    const auto r = std::from_chars(p, e, *v);
    return r.ec == std::errc() && r.ptr == e;
}

// Parses exactly `size` blank-separated values of [b, e) into `out`.  The
// text is cut into chunks at blanks; every chunk counts its tokens, a prefix
// sum gives its first index, and the chunks parse in parallel.
int parse_values(const char* b, const char* e, float* out, long long size) {
    const long long len = e - b;
    const int chunks = (int)std::max<long long>(1, std::min<long long>(256, len >> 16));
    std::vector<const char*> cut(chunks + 1);
    cut[0] = b;
    cut[chunks] = e;
    for (int c = 1; c < chunks; ++c) {
        const char* q = b + len * c / chunks;
        while (q < e && !is_blank(*q)) ++q;   // never split a token
        cut[c] = std::max(q, cut[c - 1]);
    }
    std::vector<long long> count(chunks + 1, 0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < chunks; ++c) {
        long long k = 0;
        bool in_tok = false;
        for (const char* q = cut[c]; q < cut[c + 1]; ++q) {
            const bool blank = is_blank(*q);
            if (!blank && !in_tok) ++k;
            in_tok = !blank;
        }
        count[c + 1] = k;
    }
    for (int c = 0; c < chunks; ++c) count[c + 1] += count[c];
    if (count[chunks] != size) return PCT_ERR_FORMAT;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int c = 0; c < chunks; ++c) {
        long long idx = count[c];
        const char* q = cut[c];
        const char* end = cut[c + 1];
        while (q < end) {
            while (q < end && is_blank(*q)) ++q;
            if (q == end) break;
            const char* t = q;
            while (q < end && !is_blank(*q)) ++q;
            if (!parse_float(t, q, &out[idx++])) bad = 1;
        }
    }
    return bad ? PCT_ERR_FORMAT : PCT_OK;
}

// Binary sidecar "<path>.pctbin": a header naming the text file it was made
// from (byte size and modification time) followed by the values as raw
// little-endian float32.  A sidecar that does not match its text file is
// ignored.
struct SidecarHeader {
    char magic[8];
    long long num, rows, cols, depth;
    long long text_bytes;
    long long text_mtime_ns;
};
constexpr char kSidecarMagic[8] = {'P', 'C', 'T', 'M', 'I', 'O', '1', '\0'};

std::string sidecar_path(const char* path) { return std::string(path) + ".pctbin"; }

bool text_stamp(const char* path, long long* bytes, long long* mtime_ns) {
    struct stat st;
    if (::stat(path, &st) != 0) return false;
    *bytes = (long long)st.st_size;
    *mtime_ns = (long long)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
    return true;
}

// Reads the sidecar of `path` if it is there and current; false otherwise.
bool read_sidecar(const char* path, float** data, int dims4[4]) {
    long long bytes = 0, mtime = 0;
    if (!text_stamp(path, &bytes, &mtime)) return false;
    FILE* f = std::fopen(sidecar_path(path).c_str(), "rb");
    if (!f) return false;
    SidecarHeader h;
    bool ok = std::fread(&h, sizeof(h), 1, f) == 1 &&
              std::memcmp(h.magic, kSidecarMagic, sizeof(kSidecarMagic)) == 0 &&
              h.text_bytes == bytes && h.text_mtime_ns == mtime;
    long long per_volume = 0, size = 0;
    ok = ok && checked_volume_elements(h.num, h.rows, h.cols, h.depth, &per_volume, &size);
    float* v = ok ? static_cast<float*>(std::malloc(sizeof(float) * (size_t)size)) : nullptr;
    ok = ok && v && std::fread(v, sizeof(float), (size_t)size, f) == (size_t)size &&
         std::fgetc(f) == EOF;
    std::fclose(f);
    if (!ok) {
        std::free(v);
        return false;
    }
    dims4[0] = (int)h.rows;
    dims4[1] = (int)h.cols;
    dims4[2] = (int)h.depth;
    dims4[3] = (int)h.num;
    *data = v;
    return true;
}

int read_text(const char* path, float** data, int dims4[4]) {
    std::vector<char> buf;
    int rc = slurp(path, &buf);
    if (rc) return rc;
    const char* p = buf.data();
    const char* e = buf.data() + buf.size() - 1;   // before the NUL
    // four header lines, each read with atoi semantics (cpp:256-265)
    long long hdr[4];
    for (int i = 0; i < 4; ++i) {
        if (p >= e) return PCT_ERR_FORMAT;
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(e - p)));
        const char* line_end = nl ? nl : e;
        hdr[i] = std::atoll(std::string(p, line_end).c_str());
        p = nl ? nl + 1 : e;
    }
    const long long num = hdr[0], rows = hdr[1], cols = hdr[2], depth = hdr[3];
    if (num <= 0 || rows <= 0 || cols <= 0 || depth <= 0) return PCT_ERR_FORMAT;
    // every product is checked before the allocation: a wrapped count would
    // give a short buffer that later loops overrun
    long long per_volume = 0, size = 0;
    if (!checked_volume_elements(num, rows, cols, depth, &per_volume, &size))
        return PCT_ERR_TOO_LARGE;
    float* v = static_cast<float*>(std::malloc(sizeof(float) * (size_t)size));
    if (!v) return PCT_ERR_NO_MEMORY;
    // values separated by blanks/newlines, in memory order (cpp:273-285)
    rc = parse_values(p, e, v, size);
    if (rc) {
        std::free(v);
        return rc;
    }
    dims4[0] = (int)rows;
    dims4[1] = (int)cols;
    dims4[2] = (int)depth;
    dims4[3] = (int)num;
    *data = v;
    return PCT_OK;
}

}  // namespace

extern "C" int pct_matrixio_read(const char* path, float** data, int dims4[4]) {
    if (!path || !data || !dims4) return PCT_ERR_INVALID;
    *data = nullptr;
    if (read_sidecar(path, data, dims4)) return PCT_OK;
    return read_text(path, data, dims4);
}

extern "C" int pct_matrixio_write_sidecar(const char* path) {
    if (!path) return PCT_ERR_INVALID;
    long long bytes = 0, mtime = 0;
    if (!text_stamp(path, &bytes, &mtime)) return PCT_ERR_IO;
    float* v = nullptr;
    int dims4[4];
    int rc = read_text(path, &v, dims4);
    if (rc) return rc;
    SidecarHeader h;
    std::memcpy(h.magic, kSidecarMagic, sizeof(kSidecarMagic));
    h.num = dims4[3];
    h.rows = dims4[0];
    h.cols = dims4[1];
    h.depth = dims4[2];
    h.text_bytes = bytes;
    h.text_mtime_ns = mtime;
    const size_t size = (size_t)h.num * h.rows * h.cols * h.depth;
    // written under a temporary name and renamed, so a reader never sees a
    // partial sidecar
    const std::string final_path = sidecar_path(path), tmp_path = final_path + ".tmp";
    FILE* f = std::fopen(tmp_path.c_str(), "wb");
    if (!f) {
        std::free(v);
        return PCT_ERR_IO;
    }
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(v, sizeof(float), size, f) == size;
    ok = (std::fclose(f) == 0) && ok;
    std::free(v);
    if (!ok || std::rename(tmp_path.c_str(), final_path.c_str()) != 0) {
        std::remove(tmp_path.c_str());
        return PCT_ERR_IO;
    }
    return PCT_OK;
}

extern "C" int pct_matrixio_write(const char* path, const float* data, int num, int rows, int cols,
                                  int depth) {
    if (!path || (!data && (long long)num * rows * cols * depth > 0)) return PCT_ERR_INVALID;
    if (num < 0 || rows < 0 || cols < 0 || depth < 0) return PCT_ERR_INVALID;
    FILE* f = std::fopen(path, "w");
    if (!f) return PCT_ERR_IO;
    const long long size = (long long)num * rows * cols * depth;
    std::fprintf(f, "%d\n%d\n%d\n%d\n", num, rows, cols, depth);
    // formatted in parallel blocks, written in order
    constexpr long long kBlock = 1 << 16;
    const long long nblocks = (size + kBlock - 1) / kBlock;
    bool ok = true;
    for (long long b0 = 0; b0 < nblocks && ok; b0 += 256) {
        const long long b1 = std::min(nblocks, b0 + 256);
        std::vector<std::string> text((size_t)(b1 - b0));
#pragma omp parallel for schedule(dynamic)
        for (long long b = b0; b < b1; ++b) {
            std::string& t = text[(size_t)(b - b0)];
            t.reserve((size_t)kBlock * 14);
            char one[64];
            const long long e = std::min(size, (b + 1) * kBlock);
            for (long long i = b * kBlock; i < e; ++i) {
                // std::ostream << float with the default precision 6 (cpp:307)
                const int k = std::snprintf(one, sizeof(one), "%g ", (double)data[i]);
                t.append(one, (size_t)k);
            }
        }
        for (const auto& t : text)
            if (std::fwrite(t.data(), 1, t.size(), f) != t.size()) ok = false;
    }
    ok = ok && std::ferror(f) == 0;
    return (std::fclose(f) == 0 && ok) ? PCT_OK : PCT_ERR_IO;
}

// ---- configuration file and generators ----------------------------------

extern "C" void pct_volume_set_free(pct_volume_set* vs) {
    if (!vs) return;
    std::free(vs->img);
    std::free(vs->atlas);
    std::free(vs->label);
    vs->img = vs->atlas = vs->label = nullptr;
}

extern "C" int pct_synth_uniform(int nx, int ny, int nz, int n_atlas, uint64_t seed, float* img,
                                 float* atlas, float* label) {
    if (nx < 1 || ny < 1 || nz < 1 || n_atlas < 1 || !img || !atlas || !label)
        return PCT_ERR_INVALID;
    const long long n = (long long)nx * ny * nz;
    const long long na = n * n_atlas;
    const uint64_t ka = derive_key(seed, 1, 0), kl = derive_key(seed, 2, 0),
                   ki = derive_key(seed, 3, 0);
    // cpp:961-973: 10*U values for atlas and label, then the image
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < na; ++i) {
        atlas[i] = 10.0f * (float)uniform01(ka, (uint64_t)i);
        label[i] = 10.0f * (float)uniform01(kl, (uint64_t)i);
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) img[i] = 10.0f * (float)uniform01(ki, (uint64_t)i);
    return PCT_OK;
}

extern "C" int pct_synth_tissue(int nx, int ny, int nz, int n_atlas, int max_shift,
                                int independent_anatomy, uint64_t seed, float* target_mri,
                                float* atlas_mri, float* atlas_ct) {
    if (nx < 1 || ny < 1 || nz < 1 || n_atlas < 0 || max_shift < 0) return PCT_ERR_INVALID;
    if (!target_mri || (n_atlas > 0 && (!atlas_mri || !atlas_ct))) return PCT_ERR_INVALID;
    const long long n = (long long)nx * ny * nz;
    const long long pl = (long long)nx * ny;
    std::vector<unsigned char> cls = tissue_classes(derive_key(seed, 0, 0), nx, ny, nz);
    fill_intensities(cls, derive_key(seed, 0, 1), n, target_mri, nullptr);
    for (int a = 0; a < n_atlas; ++a) {
        std::vector<unsigned char> acls;
        if (independent_anatomy) {
            acls = tissue_classes(derive_key(seed, (uint64_t)a + 1, 0), nx, ny, nz);
        } else {
            const uint64_t ks = derive_key(seed, (uint64_t)a + 1, 2);
            const int span = 2 * max_shift + 1;
            const int sz = (int)(hash_u64(ks, 0) % (uint64_t)span) - max_shift;
            const int sx = (int)(hash_u64(ks, 1) % (uint64_t)span) - max_shift;
            const int sy = (int)(hash_u64(ks, 2) % (uint64_t)span) - max_shift;
            acls.resize(n);
#pragma omp parallel for schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int x = 0; x < nx; ++x)
                    for (int y = 0; y < ny; ++y) {
                        const int zz = std::min(std::max(z + sz, 0), nz - 1);
                        const int xx = std::min(std::max(x + sx, 0), nx - 1);
                        const int yy = std::min(std::max(y + sy, 0), ny - 1);
                        acls[z * pl + (long long)x * ny + y] = cls[zz * pl + (long long)xx * ny + yy];
                    }
        }
        fill_intensities(acls, derive_key(seed, (uint64_t)a + 1, 1), n, atlas_mri + a * n,
                         atlas_ct + a * n);
    }
    return PCT_OK;
}

extern "C" int pct_read_config(const char* path, uint64_t seed, pct_volume_set* out) {
    if (!path || !out) return PCT_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    // the MATLAB build takes a .mat file as its data argument (cpp:1513-1519)
    const size_t len = std::strlen(path);
    if (len > 4 && std::strcmp(path + len - 4, ".mat") == 0) return pct_mat_read(path, out);
    std::ifstream in(path);
    if (!in.is_open()) return PCT_ERR_IO;
    // cpp:922-958: a line-indexed state machine
    std::string line, atlas_path, label_path, img_path;
    int state = 0, random_data = 0, num = 0, nx = 0, ny = 0, nz = 0;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        switch (state) {
            case 0: random_data = std::atoi(line.c_str()); state++; break;
            case 1: if (random_data) num = std::atoi(line.c_str()); else atlas_path = line; state++; break;
            case 2: if (random_data) nx = std::atoi(line.c_str()); else label_path = line; state++; break;
            case 3: if (random_data) ny = std::atoi(line.c_str()); else img_path = line; state++; break;
            case 4: if (random_data) { nz = std::atoi(line.c_str()); state++; } break;
            default: break;
        }
    }
    if (random_data) {
        if (state < 5 || num < 1 || nx < 1 || ny < 1 || nz < 1) return PCT_ERR_FORMAT;
        const size_t n = (size_t)nx * ny * nz;
        out->nx = nx; out->ny = ny; out->nz = nz; out->n_atlas = num;
        out->img = static_cast<float*>(std::malloc(n * 4));
        out->atlas = static_cast<float*>(std::malloc(n * num * 4));
        out->label = static_cast<float*>(std::malloc(n * num * 4));
        if (!out->img || !out->atlas || !out->label) {
            pct_volume_set_free(out);
            return PCT_ERR_NO_MEMORY;
        }
        return pct_synth_uniform(nx, ny, nz, num, seed, out->img, out->atlas, out->label);
    }
    if (state < 4) return PCT_ERR_FORMAT;
    int da[4], dl[4], di[4];
    float *a = nullptr, *l = nullptr, *im = nullptr;
    int rc = pct_matrixio_read(atlas_path.c_str(), &a, da);
    if (!rc) rc = pct_matrixio_read(label_path.c_str(), &l, dl);
    if (!rc) rc = pct_matrixio_read(img_path.c_str(), &im, di);
    // cpp:991-1007: atlas and label dims equal, image dims equal in x, y, z
    if (!rc && !(da[0] == dl[0] && da[1] == dl[1] && da[2] == dl[2] && da[3] == dl[3] &&
                 da[0] == di[0] && da[1] == di[1] && da[2] == di[2]))
        rc = PCT_ERR_DIMS;
    if (rc) {
        std::free(a); std::free(l); std::free(im);
        return rc;
    }
    out->nx = dl[0]; out->ny = dl[1]; out->nz = dl[2]; out->n_atlas = dl[3];
    out->img = im; out->atlas = a; out->label = l;
    return PCT_OK;
}
