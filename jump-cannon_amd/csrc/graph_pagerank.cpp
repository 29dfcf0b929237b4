// graph-pagerank: one-shot PageRank over a CSR file on the GPU, through the C ABI.
//
// Mirrors the reference's `graph-pagerank` binary
// (crates/graph-compute/src/bin/pagerank.rs): reads a CSR in graph-compute's
// binary format ([u32 n_nodes][u32 n_edges][u32 x (n+1) offsets][u32 x m
// neighbors], sim.rs:53-128), runs `iters` power iterations with the given
// damping (no tolerance: the reference runs a fixed count), and writes a JSON
// array of per-node f32 ranks in CSR node order to stdout or --out.  The chosen
// backend and the timing go to stderr.
//
// Usage:
//     graph-pagerank <csr.bin> [--damping 0.85] [--iters 100] [--out ranks.json]
//
// The reference falls back to its CPU implementation when no GPU adapter is
// present or the GPU path rejects the graph (dangling nodes).  This build has
// no CPU path: those cases end with the error and a non-zero exit status, and
// --cpu is rejected.
#include <chrono>
#include <cmath>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/jc.h"

namespace {

const char *kUsage = "graph-pagerank <csr.bin> [--damping 0.85] [--iters 100] [--out ranks.json]";

struct Args {
    std::string csr_path;
    float damping = 0.85f;
    uint32_t iters = 100;
    std::string out;
};

[[noreturn]] void fail(const std::string &msg, int code = 1) {
    std::fprintf(stderr, "graph-pagerank: %s\n", msg.c_str());
    std::exit(code);
}

Args parse_args(int argc, char **argv) {
    Args a;
    for (int i = 1; i < argc; i++) {
        const std::string arg = argv[i];
        auto value = [&](const char *flag) -> std::string {
            if (i + 1 >= argc) fail(std::string(flag) + " needs a value", 2);
            return argv[++i];
        };
        if (arg == "--damping") {
            const std::string v = value("--damping");
            char *end = nullptr;
            errno = 0;
            a.damping = std::strtof(v.c_str(), &end);
            if (errno || end == v.c_str() || *end) fail("parsing --damping: invalid float '" + v + "'", 2);
        } else if (arg == "--iters") {
            const std::string v = value("--iters");
            char *end = nullptr;
            errno = 0;
            const unsigned long long it = std::strtoull(v.c_str(), &end, 10);
            if (errno || end == v.c_str() || *end || v[0] == '-' || it > 0xFFFFFFFFull)
                fail("parsing --iters: invalid u32 '" + v + "'", 2);
            a.iters = (uint32_t)it;
        } else if (arg == "--out") {
            a.out = value("--out");
        } else if (arg == "--cpu") {
            fail("--cpu: this build has no CPU PageRank; it runs on the GPU only", 2);
        } else if (arg == "-h" || arg == "--help") {
            std::fprintf(stderr, "%s\n", kUsage);
            std::exit(0);
        } else if (a.csr_path.empty() && !arg.empty() && arg[0] != '-') {
            a.csr_path = arg;
        } else {
            fail("unexpected argument: " + arg, 2);
        }
    }
    if (a.csr_path.empty()) fail("missing required <csr.bin> path", 2);
    return a;
}

std::vector<uint8_t> read_file(const std::string &path) {
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

// Shortest decimal that reads back as the same f32 (null for a non-finite
// value, as serde_json writes it).
std::string f32_text(float v) {
    if (!std::isfinite(v)) return "null";
    char s[64];
    for (int prec = 1; prec <= 9; prec++) {
        std::snprintf(s, sizeof s, "%.*g", prec, (double)v);
        if (std::strtof(s, nullptr) == v) break;
    }
    return s;
}

}  // namespace

int main(int argc, char **argv) {
    const Args args = parse_args(argc, argv);
    const std::vector<uint8_t> bytes = read_file(args.csr_path);
    uint32_t n = 0, m = 0;
    if (jc_csr_bin_header(bytes.data(), bytes.size(), &n, &m) != JC_OK)
        fail("loading CSR from " + args.csr_path + ": " + jc_last_error());
    std::vector<uint32_t> offsets((size_t)n + 1), neighbors(m ? m : 1);
    if (jc_csr_from_bin(bytes.data(), bytes.size(), offsets.data(), neighbors.data()) != JC_OK)
        fail("loading CSR from " + args.csr_path + ": " + jc_last_error());
    std::fprintf(stderr, "graph-pagerank: %u nodes, %u edges, damping=%g, iters=%u\n", n, m,
                 (double)args.damping, args.iters);

    const auto t0 = std::chrono::steady_clock::now();
    std::vector<float> ranks(n);
    uint32_t iters_done = 0;
    double resid = 0.0;
    const int rc = jc_pagerank(n, offsets.data(), neighbors.data(), m, args.damping, args.iters,
                               0.0, ranks.data(), &iters_done, &resid);
    if (rc != JC_OK) fail(std::string("GPU PageRank failed: ") + jc_last_error());
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "graph-pagerank: backend=gpu (hip), elapsed=%.3fms\n", ms);

    std::string json = "[";
    for (uint32_t i = 0; i < n; i++) {
        if (i) json += ',';
        json += f32_text(ranks[i]);
    }
    json += "]";
    if (!args.out.empty()) {
        FILE *f = std::fopen(args.out.c_str(), "wb");
        if (!f || std::fwrite(json.data(), 1, json.size(), f) != json.size())
            fail("writing " + args.out + ": " + std::strerror(errno));
        std::fclose(f);
        std::fprintf(stderr, "graph-pagerank: wrote %u ranks to %s\n", n, args.out.c_str());
    } else {
        json += "\n";
        std::fwrite(json.data(), 1, json.size(), stdout);
    }
    return 0;
}
