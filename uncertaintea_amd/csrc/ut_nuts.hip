// ut_nuts.hip -- launcher of the persistent NUTS kernel (k_nuts, ut_chain.h) for
// every device model and chain layout.
#include "ut_chain.h"

namespace ut {

hipError_t launch_nuts(const DevModelDesc& d, const ChainState& st, const SamplerParams& sp,
                       int it_begin, int it_end, hipStream_t stream) {
  return dispatch_chain(d, [&](auto m) {
    constexpr int K = decltype(m)::K;
    constexpr int NW = WavesOf<decltype(m)>::value;
    const size_t ck = (size_t)NW * nuts_checkpoint_bytes(K, sp.max_depth);
    const size_t lds = model_lds<float>(m, ck, (const void*)k_nuts<decltype(m)>);
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)k_nuts<decltype(m)>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    hipLaunchKernelGGL((k_nuts<decltype(m)>), dim3(sp.C), dim3(WAVE * NW), lds, stream, m, st, sp,
                       it_begin, it_end);
    return hipGetLastError();
  });
}

}  // namespace ut
