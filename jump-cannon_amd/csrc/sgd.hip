// SGD sparse stress engine (placeholder until the HIP kernels land).
#include "engine.hpp"

namespace jc {
LayoutEngine *make_sgd(const jc_sgd_settings &) { return nullptr; }
}  // namespace jc
