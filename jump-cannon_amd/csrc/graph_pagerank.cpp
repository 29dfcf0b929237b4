// graph-pagerank: one-shot PageRank over a CSR file on the GPU, through the C ABI.
//
// Mirrors the reference's `graph-pagerank` binary
// (crates/graph-compute/src/bin/pagerank.rs): reads a CSR in graph-compute's
// binary format, runs `iters` power iterations with the given damping (no
// tolerance: the reference runs a fixed count), and writes a JSON array of
// per-node f32 ranks in CSR node order to stdout or --out.  The chosen backend
// and the timing go to stderr.
//
// Usage:
//     graph-pagerank <csr.bin> [--damping 0.85] [--iters 100] [--out ranks.json]
//
// The reference falls back to its CPU implementation when no GPU adapter is
// present or the GPU path rejects the graph (dangling nodes).  This build has
// no CPU path: those cases end with the error and a non-zero exit status, and
// --cpu is rejected.
#include <chrono>

#include "cli_common.hpp"

using namespace jc_cli;

namespace {

const char *kUsage = "graph-pagerank <csr.bin> [--damping 0.85] [--iters 100] [--out ranks.json]";

struct Args {
    std::string csr_path;
    float damping = 0.85f;
    uint32_t iters = 100;
    std::string out;
};

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
            a.iters = parse_u32("--iters", value("--iters"));
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

}  // namespace

int main(int argc, char **argv) {
    prog() = "graph-pagerank";
    const Args args = parse_args(argc, argv);
    const Csr g = load_csr(args.csr_path);
    std::fprintf(stderr, "graph-pagerank: %u nodes, %u edges, damping=%g, iters=%u\n", g.n, g.m,
                 (double)args.damping, args.iters);

    const auto t0 = std::chrono::steady_clock::now();
    std::vector<float> ranks(g.n);
    uint32_t iters_done = 0;
    double resid = 0.0;
    const int rc = jc_pagerank(g.n, g.offsets.data(), g.neighbors.data(), g.m, args.damping, args.iters,
                               0.0, ranks.data(), &iters_done, &resid);
    if (rc != JC_OK) fail(std::string("GPU PageRank failed: ") + jc_last_error());
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "graph-pagerank: backend=gpu (hip), elapsed=%.3fms\n", ms);
    emit(json_array(ranks), args.out, ranks.size(), "ranks");
    return 0;
}
