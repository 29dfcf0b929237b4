// graph-bfs: one-shot single-source BFS over a CSR file on the GPU, through the
// C ABI.
//
// Mirrors the reference's `graph-bfs` binary (crates/graph-compute/src/bin/bfs.rs):
// reads a CSR in graph-compute's binary format and writes a JSON array of
// per-node hop distances from --source, in CSR node order; unreachable nodes
// are 4294967295 (u32::MAX, bfs.rs:20-22).  The backend and the timing go to
// stderr.  A source outside the graph is an error, as in the reference.
//
// Usage:
//     graph-bfs <csr.bin> --source <u32> [--out dists.json]
//
// This build has no CPU path: --cpu is rejected, and a missing GPU ends with the
// error and a non-zero exit status instead of the reference's CPU fallback.
#include <chrono>

#include "cli_common.hpp"

using namespace jc_cli;

namespace {

const char *kUsage = "graph-bfs <csr.bin> --source <u32> [--out dists.json]";

struct Args {
    std::string csr_path;
    bool has_source = false;
    uint32_t source = 0;
    std::string out;
};

Args parse_args(int argc, char **argv) {
    Args a;
    for (int i = 1; i < argc; i++) {
        const std::string arg = argv[i];
        auto value = [&](const char *flag, const char *what) -> std::string {
            if (i + 1 >= argc) fail(std::string(flag) + " needs a " + what, 2);
            return argv[++i];
        };
        if (arg == "--source") {
            a.source = parse_u32("--source", value("--source", "value"));
            a.has_source = true;
        } else if (arg == "--out") {
            a.out = value("--out", "path");
        } else if (arg == "--cpu") {
            fail("--cpu: this build has no CPU BFS; it runs on the GPU only", 2);
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
    prog() = "graph-bfs";
    const Args args = parse_args(argc, argv);
    if (!args.has_source) fail("missing required --source <u32>", 2);
    const Csr g = load_csr(args.csr_path);
    if (args.source >= g.n)
        fail("--source " + std::to_string(args.source) + " out of range (graph has " + std::to_string(g.n) +
             " nodes)");
    std::fprintf(stderr, "graph-bfs: %u nodes, %u edges, source=%u\n", g.n, g.m, args.source);

    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> dist(g.n);
    const int rc = jc_bfs(g.n, g.offsets.data(), g.neighbors.data(), g.m, args.source, dist.data());
    if (rc != JC_OK) fail(std::string("GPU BFS failed: ") + jc_last_error());
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "graph-bfs: backend=gpu (hip), elapsed=%.3fms\n", ms);
    emit(json_array(dist), args.out, dist.size(), "distances");
    return 0;
}
