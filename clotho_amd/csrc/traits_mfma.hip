// traits_mfma.hip — placeholder until the matrix-core contraction lands.
#include <hip/hip_runtime.h>

#include "traits.hpp"

namespace clotho {

size_t mfma_effect_elems(uint32_t sites, uint32_t traits, int terms) {
    return (size_t)sites * traits * terms;
}
void mfma_pack_effects(const double*, uint32_t, uint32_t, int, uint16_t*) {}
size_t mfma_partial_elems(uint32_t h, uint32_t, uint32_t traits) { return (size_t)h * traits; }
hipError_t launch_traits_mfma(const uint64_t*, uint32_t, uint32_t, const uint16_t*, uint32_t,
                              uint32_t, int, float*, double*, hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace clotho
