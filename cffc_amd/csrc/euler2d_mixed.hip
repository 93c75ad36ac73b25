// euler2d_mixed.hip -- the fused stage kernel with FP32 reconstruction,
// limiter and Roe flux (FT = float) over FP64 state, update, time step and
// norms.  Same source as the FP64 instantiation (euler2d_stage.hpp); this
// translation unit is compiled with floating-point contraction enabled.
#include <hip/hip_runtime.h>

#include "euler2d_kernels.hpp"
#include "euler2d_stage.hpp"

using namespace cffc;

extern "C" hipError_t cffc_launch_stage_mixed(const StageParams *p, int ti, int tj, int fused, hipStream_t st) {
#define L(TI, TJ)                                                                                          \
    if (ti == TI && tj == TJ) {                                                                            \
        if (fused) hipLaunchKernelGGL((stage_kernel<float, TI, TJ, true>), dim3(p->ntiles), dim3((TileThreads<TI, TJ>::value)), 0, st, *p); \
        else hipLaunchKernelGGL((stage_kernel<float, TI, TJ, false>), dim3(p->ntiles), dim3((TileThreads<TI, TJ>::value)), 0, st, *p); \
        return hipGetLastError();                                                                          \
    }
    CFFC_TILE_LIST(L)
#undef L
    return hipErrorInvalidValue;
}

extern "C" hipError_t cffc_launch_update_tile_mixed(const StageParams *p, int ti, int tj, hipStream_t st) {
    hipLaunchKernelGGL((update_tile_kernel<float>), dim3(p->tile_list ? p->nlist : p->ntiles), dim3(UPD_NT), 0, st, *p, ti,
                       tj);
    return hipGetLastError();
}
