// ut_chain.hip -- launchers of the per-chain density, HMC and step-search kernels
// (ut_chain.h) for every device model and chain layout.
#include "ut_chain.h"

namespace ut {

hipError_t launch_density_f32(const DevModelDesc& d, const float* params, float* values,
                              float* grads, int C, hipStream_t stream) {
  return dispatch_chain(d, [&](auto m) {
    const size_t lds = model_lds<float>(m, 0, (const void*)k_density<decltype(m), float>);
    hipLaunchKernelGGL((k_density<decltype(m), float>), dim3(C), dim3(WAVE * WavesOf<decltype(m)>::value),
                       lds, stream, m, params, values, grads, d.P);
    return hipGetLastError();
  });
}

hipError_t launch_density_f64(const DevModelDesc& d, const double* params, double* values,
                              double* grads, int C, hipStream_t stream) {
  return dispatch_chain(d, [&](auto m) {
    const size_t lds = model_lds<double>(m, 0, (const void*)k_density<decltype(m), double>);
    hipLaunchKernelGGL((k_density<decltype(m), double>), dim3(C), dim3(WAVE * WavesOf<decltype(m)>::value),
                       lds, stream, m, params, values, grads, d.P);
    return hipGetLastError();
  });
}

hipError_t launch_hmc(const DevModelDesc& d, const ChainState& st, const SamplerParams& sp,
                      int it_begin, int it_end, hipStream_t stream) {
  return dispatch_chain(d, [&](auto m) {
    const size_t lds = model_lds<float>(m, 0, (const void*)k_hmc<decltype(m)>);
    hipLaunchKernelGGL((k_hmc<decltype(m)>), dim3(sp.C), dim3(WAVE * WavesOf<decltype(m)>::value), lds,
                       stream, m, st, sp, it_begin, it_end);
    return hipGetLastError();
  });
}

hipError_t launch_step_search(const DevModelDesc& d, const ChainState& st, int C,
                              uint32_t itc, int do_search, double target, hipStream_t stream) {
  return dispatch_chain(d, [&](auto m) {
    const size_t lds = model_lds<float>(m, 0, (const void*)k_step_search<decltype(m)>);
    hipLaunchKernelGGL((k_step_search<decltype(m)>), dim3(C), dim3(WAVE * WavesOf<decltype(m)>::value), lds,
                       stream, m, st, d.P, itc, do_search, target);
    return hipGetLastError();
  });
}

}  // namespace ut
