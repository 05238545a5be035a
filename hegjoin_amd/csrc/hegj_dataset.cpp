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
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "hegj_internal.h"

using namespace hegj;

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

hegj_status hegj_dataset_load(const char *path, uint32_t dim, double **out, uint32_t *n_out) {
    if (!path || !out || !n_out || dim == 0) return fail(HEGJ_ERR_INVALID_ARG, "bad loader arguments");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(HEGJ_ERR_IO, std::string("cannot open ") + path);
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::string buf((size_t)std::max(sz, 0L), '\0');
    bool ok = sz <= 0 || std::fread(&buf[0], 1, (size_t)sz, f) == (size_t)sz;
    std::fclose(f);
    if (!ok) return fail(HEGJ_ERR_IO, std::string("cannot read ") + path);
    std::vector<double> vals;
    size_t plen = std::strlen(path);
    if (plen > 4 && std::strcmp(path + plen - 4, ".bin") == 0) {
        vals.resize(buf.size() / sizeof(double));
        std::memcpy(vals.data(), buf.data(), vals.size() * sizeof(double));
    } else {
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
                    double v = std::strtod(tok.c_str(), &endp);
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
    uint64_t npts = vals.size() / dim;
    if (npts > UINT32_MAX) return fail(HEGJ_ERR_UNSUPPORTED, "more than 2^32-1 points");
    double *o = (double *)std::malloc(sizeof(double) * (npts * dim + 1));
    if (!o) return fail(HEGJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::memcpy(o, vals.data(), sizeof(double) * npts * dim);
    *out = o;
    *n_out = (uint32_t)npts;
    return HEGJ_OK;
}

void hegj_dataset_free(double *p) { std::free(p); }

}  // extern "C"
