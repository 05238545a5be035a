// hegj_dataset.cpp -- dataset generator and loader of libhegjoin.
//
// hegj_dataset_generate_expo follows datasets/dataset_gen_fixed_len_expo_dist.cpp:57-84
// (std::mt19937 + std::exponential_distribution<double>, redraw outside
// [0, 1], point-major order), so a seed reproduces the reference generator's
// values exactly.  hegj_dataset_load reads what importNDDataset
// (import_dataset.cpp:10-40) reads -- numbers separated by commas and/or
// whitespace, count = values / dim -- parsing chunks of the file in parallel;
// a *.bin file is raw row-major doubles (the format importNDDatasetBinary,
// import_dataset.cpp:63-112, intended).
//
// hegj_dataset_load_ex adds two options:
//   HEGJ_LOAD_F32: each number is parsed as `ss >> float` parses it (strtof,
//     correctly rounded to float), for a DTYPE=float build (params.h:24);
//   HEGJ_LOAD_BIN_CACHE: a parsed text file is written beside itself as a
//     binary cache, <path>.d<dim>[f].hegjbin = a 56-byte header (magic,
//     version, dim, precision, point count, size and mtime of the text file)
//     + the row-major doubles; a later load whose header still matches the
//     text file reads the cache instead of parsing.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "hegj_internal.h"

using namespace hegj;

namespace {

constexpr char kCacheMagic[8] = {'H', 'E', 'G', 'J', 'B', 'I', 'N', '1'};

struct CacheHeader {
    char magic[8];
    uint32_t version;
    uint32_t dim;
    uint32_t f32;        // values were parsed as float
    uint32_t pad;
    uint64_t n;          // points
    uint64_t src_size;   // bytes of the text file
    int64_t src_mtime_s;
    int64_t src_mtime_ns;
};
static_assert(sizeof(CacheHeader) == 56, "cache header layout");

std::string cache_path(const char *path, uint32_t dim, bool f32) {
    return std::string(path) + ".d" + std::to_string(dim) + (f32 ? "f" : "") + ".hegjbin";
}

bool read_file(const char *path, std::string &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.assign((size_t)std::max(sz, 0L), '\0');
    bool ok = sz <= 0 || std::fread(&buf[0], 1, (size_t)sz, f) == (size_t)sz;
    std::fclose(f);
    return ok;
}

// Reads the cache when its header matches the text file; false otherwise.
bool load_cache(const std::string &cpath, const struct stat &src, uint32_t dim, bool f32,
                std::vector<double> &vals) {
    FILE *f = std::fopen(cpath.c_str(), "rb");
    if (!f) return false;
    CacheHeader h;
    bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kCacheMagic, 8) == 0 &&
              h.version == 1 && h.dim == dim && h.f32 == (f32 ? 1u : 0u) &&
              h.src_size == (uint64_t)src.st_size && h.src_mtime_s == (int64_t)src.st_mtim.tv_sec &&
              h.src_mtime_ns == (int64_t)src.st_mtim.tv_nsec && h.n <= UINT32_MAX;
    if (ok) {
        vals.resize(h.n * dim);
        ok = vals.empty() || std::fread(vals.data(), sizeof(double), vals.size(), f) == vals.size();
    }
    std::fclose(f);
    return ok;
}

// Writes the cache through a temporary file renamed into place, so a reader
// never sees a partial cache.  Failure to write is not an error of the load.
void store_cache(const std::string &cpath, const struct stat &src, uint32_t dim, bool f32,
                 const double *vals, uint64_t npts) {
    CacheHeader h{};
    std::memcpy(h.magic, kCacheMagic, 8);
    h.version = 1;
    h.dim = dim;
    h.f32 = f32 ? 1u : 0u;
    h.n = npts;
    h.src_size = (uint64_t)src.st_size;
    h.src_mtime_s = (int64_t)src.st_mtim.tv_sec;
    h.src_mtime_ns = (int64_t)src.st_mtim.tv_nsec;
    const std::string tmp = cpath + ".tmp" + std::to_string((long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 &&
              (npts == 0 || std::fwrite(vals, sizeof(double), npts * dim, f) == npts * dim);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), cpath.c_str()) != 0) std::remove(tmp.c_str());
}

// Parses comma / whitespace separated numbers in parallel chunks (cut at
// line ends).  Like `ss >> v` in importNDDataset, parsing stops at the first
// token that is not a number.
void parse_text(const std::string &buf, bool f32, std::vector<double> &vals) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<size_t> cut(nt + 1, buf.size());
    cut[0] = 0;
    for (unsigned t = 1; t < nt; t++) {
        size_t p = std::max(cut[t - 1], buf.size() * t / nt);
        while (p < buf.size() && buf[p] != '\n') p++;
        cut[t] = p;
    }
    std::vector<std::vector<double>> parts(nt);
    std::vector<char> stop(nt, 0);
    std::vector<std::thread> th;
    auto sep = [](char c) { return c == ',' || c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            const char *p = buf.data() + cut[t];
            const char *e = buf.data() + cut[t + 1];
            std::string tok;
            while (p < e) {
                while (p < e && sep(*p)) p++;
                if (p >= e) break;
                const char *s = p;
                while (p < e && !sep(*p)) p++;
                tok.assign(s, p);
                char *endp;
                // float: rounded once, from the decimal string (not via double)
                const double v = f32 ? (double)std::strtof(tok.c_str(), &endp) : std::strtod(tok.c_str(), &endp);
                if (endp == tok.c_str()) {  // operator>> stops at the first bad token
                    stop[t] = 1;
                    break;
                }
                parts[t].push_back(v);
            }
        });
    for (auto &x : th) x.join();
    for (unsigned t = 0; t < nt; t++) {
        vals.insert(vals.end(), parts[t].begin(), parts[t].end());
        if (stop[t]) break;
    }
}

}  // namespace

extern "C" {

hegj_status hegj_dataset_generate_expo(uint32_t seed, uint32_t n, uint32_t dim, double lambda,
                                       double *out) {
    if (!out || dim == 0 || !(lambda > 0)) return fail(HEGJ_ERR_INVALID_ARG, "bad generator arguments");
    std::mt19937 gen(seed);
    std::exponential_distribution<double> dis(lambda);
    const unsigned int length = 1;
    const uint64_t total = (uint64_t)n * dim;
    for (uint64_t i = 0; i < total; i++) {
        double val;
        do {
            val = dis(gen) * length;
        } while (val < 0 || val > 1);
        out[i] = val;
    }
    return HEGJ_OK;
}

hegj_status hegj_dataset_load_ex(const char *path, uint32_t dim, uint32_t flags, double **out,
                                 uint32_t *n_out) {
    if (!path || !out || !n_out || dim == 0) return fail(HEGJ_ERR_INVALID_ARG, "bad loader arguments");
    if (flags & ~(uint32_t)(HEGJ_LOAD_BIN_CACHE | HEGJ_LOAD_F32))
        return fail(HEGJ_ERR_INVALID_ARG, "unknown loader flags");
    const bool f32 = (flags & HEGJ_LOAD_F32) != 0;
    std::vector<double> vals;
    const size_t plen = std::strlen(path);
    const bool bin = plen > 4 && std::strcmp(path + plen - 4, ".bin") == 0;
    struct stat src {};
    const bool use_cache = !bin && (flags & HEGJ_LOAD_BIN_CACHE) && stat(path, &src) == 0;
    const std::string cpath = use_cache ? cache_path(path, dim, f32) : std::string();
    if (!(use_cache && load_cache(cpath, src, dim, f32, vals))) {
        vals.clear();
        std::string buf;
        if (!read_file(path, buf)) return fail(HEGJ_ERR_IO, std::string("cannot read ") + path);
        if (bin) {
            vals.resize(buf.size() / sizeof(double));
            std::memcpy(vals.data(), buf.data(), vals.size() * sizeof(double));
            if (f32)
                for (double &v : vals) v = (double)(float)v;
        } else {
            parse_text(buf, f32, vals);
        }
        if (use_cache) store_cache(cpath, src, dim, f32, vals.data(), vals.size() / dim);
    }
    uint64_t npts = vals.size() / dim;
    if (npts > UINT32_MAX) return fail(HEGJ_ERR_UNSUPPORTED, "more than 2^32-1 points");
    double *o = (double *)std::malloc(sizeof(double) * (npts * dim + 1));
    if (!o) return fail(HEGJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::memcpy(o, vals.data(), sizeof(double) * npts * dim);
    *out = o;
    *n_out = (uint32_t)npts;
    return HEGJ_OK;
}

hegj_status hegj_dataset_load(const char *path, uint32_t dim, double **out, uint32_t *n_out) {
    return hegj_dataset_load_ex(path, dim, 0, out, n_out);
}

void hegj_dataset_free(double *p) { std::free(p); }

}  // extern "C"
