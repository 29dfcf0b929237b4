// graph-cc: one-shot connected components over a CSR file on the GPU, through
// the C ABI.
//
// Mirrors the reference's `graph-cc` binary (crates/graph-compute/src/bin/cc.rs):
// reads a CSR in graph-compute's binary format and writes a JSON array of
// per-node component labels (the smallest node index of the component), in CSR
// node order.  The backend and the timing go to stderr.
//
// Usage:
//     graph-cc <csr.bin> [--out labels.json]
//
// This build has no CPU path: --cpu is rejected, and a missing GPU ends with the
// error and a non-zero exit status instead of the reference's CPU fallback.
#include <chrono>

#include "cli_common.hpp"

using namespace jc_cli;

namespace {

const char *kUsage = "graph-cc <csr.bin> [--out labels.json]";

struct Args {
    std::string csr_path;
    std::string out;
};

Args parse_args(int argc, char **argv) {
    Args a;
    for (int i = 1; i < argc; i++) {
        const std::string arg = argv[i];
        if (arg == "--out") {
            if (i + 1 >= argc) fail("--out needs a path", 2);
            a.out = argv[++i];
        } else if (arg == "--cpu") {
            fail("--cpu: this build has no CPU connected components; it runs on the GPU only", 2);
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
    prog() = "graph-cc";
    const Args args = parse_args(argc, argv);
    const Csr g = load_csr(args.csr_path);
    std::fprintf(stderr, "graph-cc: %u nodes, %u edges\n", g.n, g.m);

    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> labels(g.n);
    const int rc = jc_connected_components(g.n, g.offsets.data(), g.neighbors.data(), g.m, labels.data());
    if (rc != JC_OK) fail(std::string("GPU connected components failed: ") + jc_last_error());
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stderr, "graph-cc: backend=gpu (hip), elapsed=%.3fms\n", ms);
    emit(json_array(labels), args.out, labels.size(), "labels");
    return 0;
}
