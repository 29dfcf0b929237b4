#include "engine.hpp"
extern "C" int jc_pagerank(uint32_t, const uint32_t *, const uint32_t *, uint64_t, float, uint32_t,
                           double, float *, uint32_t *, double *) {
    jc::set_error("not built");
    return JC_E_STATE;
}
