// Persistent Metropolis sweep kernels, float bulk precision.
#include "qmc_launch.h"

namespace qmc {

template <int NM, int RAD>
static void go(const SArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((sweep_kernel<float, NM, RAD>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
}

template <int RAD>
static int bucket(const SArgs& a, hipStream_t s) {
    switch (nm_bucket(a.t)) {
        case 2: go<2, RAD>(a, s); return 0;
        case 4: go<4, RAD>(a, s); return 0;
        case 8: go<8, RAD>(a, s); return 0;
        case 16: go<16, RAD>(a, s); return 0;
        case 24: go<24, RAD>(a, s); return 0;
        default: return -1;
    }
}

int launch_sweep_f32(const SArgs& a, hipStream_t s) {
    return a.t.radial == QMC_RADIAL_STO ? bucket<QMC_RADIAL_STO>(a, s) : bucket<QMC_RADIAL_GTO_PURE>(a, s);
}

}  // namespace qmc
