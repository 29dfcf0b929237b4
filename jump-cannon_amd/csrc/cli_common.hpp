// Shared pieces of the one-shot analytics CLIs (graph-pagerank, graph-bfs,
// graph-cc), which mirror the reference's crates/graph-compute/src/bin/*.rs over
// the C ABI: argument errors, reading graph-compute's binary CSR file
// ([u32 n_nodes][u32 n_edges][u32 x (n+1) offsets][u32 x m neighbors],
// sim.rs:53-128) and writing the per-node JSON array to stdout or --out.
#pragma once
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/jc.h"

namespace jc_cli {

inline const char *&prog() {
    static const char *p = "graph-cli";
    return p;
}

// Message to stderr, then exit: 2 for usage errors, 1 for load / GPU errors.
[[noreturn]] inline void fail(const std::string &msg, int code = 1) {
    std::fprintf(stderr, "%s: %s\n", prog(), msg.c_str());
    std::exit(code);
}

inline std::vector<uint8_t> read_file(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) fail("loading CSR from " + path + ": " + std::strerror(errno));
    std::vector<uint8_t> bytes;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) fail("loading CSR from " + path + ": read error");
    return bytes;
}

struct Csr {
    uint32_t n = 0, m = 0;
    std::vector<uint32_t> offsets, neighbors;
};

inline Csr load_csr(const std::string &path) {
    const std::vector<uint8_t> bytes = read_file(path);
    Csr g;
    if (jc_csr_bin_header(bytes.data(), bytes.size(), &g.n, &g.m) != JC_OK)
        fail("loading CSR from " + path + ": " + jc_last_error());
    g.offsets.resize((size_t)g.n + 1);
    g.neighbors.resize(g.m ? g.m : 1);
    if (jc_csr_from_bin(bytes.data(), bytes.size(), g.offsets.data(), g.neighbors.data()) != JC_OK)
        fail("loading CSR from " + path + ": " + jc_last_error());
    return g;
}

// Shortest decimal that reads back as the same f32 (null for a non-finite
// value, as serde_json writes it).
inline std::string f32_text(float v) {
    if (!std::isfinite(v)) return "null";
    char s[64];
    for (int prec = 1; prec <= 9; prec++) {
        std::snprintf(s, sizeof s, "%.*g", prec, (double)v);
        if (std::strtof(s, nullptr) == v) break;
    }
    return s;
}

inline std::string json_array(const std::vector<float> &v) {
    std::string s = "[";
    for (size_t i = 0; i < v.size(); i++) {
        if (i) s += ',';
        s += f32_text(v[i]);
    }
    return s + "]";
}

inline std::string json_array(const std::vector<uint32_t> &v) {
    std::string s = "[";
    char num[16];
    for (size_t i = 0; i < v.size(); i++) {
        if (i) s += ',';
        std::snprintf(num, sizeof num, "%u", v[i]);
        s += num;
    }
    return s + "]";
}

// JSON to `out` (reporting "wrote <count> <what> to <out>" on stderr) or, when
// out is empty, to stdout followed by a newline.
inline void emit(const std::string &json, const std::string &out, size_t count, const char *what) {
    if (!out.empty()) {
        FILE *f = std::fopen(out.c_str(), "wb");
        if (!f || std::fwrite(json.data(), 1, json.size(), f) != json.size())
            fail("writing " + out + ": " + std::strerror(errno));
        std::fclose(f);
        std::fprintf(stderr, "%s: wrote %zu %s to %s\n", prog(), count, what, out.c_str());
    } else {
        std::fwrite(json.data(), 1, json.size(), stdout);
        std::fputc('\n', stdout);
    }
}

inline uint32_t parse_u32(const std::string &flag, const std::string &v) {
    char *end = nullptr;
    errno = 0;
    const unsigned long long x = std::strtoull(v.c_str(), &end, 10);
    if (errno || end == v.c_str() || *end || v.empty() || v[0] == '-' || x > 0xFFFFFFFFull)
        fail("parsing " + flag + ": invalid u32 '" + v + "'", 2);
    return (uint32_t)x;
}

}  // namespace jc_cli
