// psi / local-energy kernels, float bulk precision.
#include "qmc_launch.h"

namespace qmc {

template <int NM, int RAD>
static void go(const KArgs& a, int mode, hipStream_t s) {
    const dim3 g = grid_for(a.t, a.nw);
    if (mode == QMC_EVAL_ELOC)
        hipLaunchKernelGGL((wf_kernel<float, NM, RAD, true>), g, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((wf_kernel<float, NM, RAD, false>), g, dim3(kBlock), 0, s, a);
}

template <int RAD>
static int bucket(const KArgs& a, int mode, hipStream_t s) {
    switch (nm_bucket(a.t)) {
        case 2: go<2, RAD>(a, mode, s); return 0;
        case 4: go<4, RAD>(a, mode, s); return 0;
        case 8: go<8, RAD>(a, mode, s); return 0;
        case 16: go<16, RAD>(a, mode, s); return 0;
        case 24: go<24, RAD>(a, mode, s); return 0;
        default: return -1;
    }
}

int launch_wf_f32(const KArgs& a, int mode, hipStream_t s) {
    return a.t.radial == QMC_RADIAL_STO ? bucket<QMC_RADIAL_STO>(a, mode, s) : bucket<QMC_RADIAL_GTO_PURE>(a, mode, s);
}

}  // namespace qmc
