// Adjoint-row and AO-layout kernels (FP64).
#include "qmc_launch.h"

namespace qmc {

template <int NM, int RAD>
static void go(const GArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((grad_kernel<NM, RAD>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
}

template <int RAD>
static int bucket(const GArgs& a, hipStream_t s) {
    switch (nm_bucket(a.t)) {
        case 2: go<2, RAD>(a, s); return 0;
        case 4: go<4, RAD>(a, s); return 0;
        case 8: go<8, RAD>(a, s); return 0;
        case 16: go<16, RAD>(a, s); return 0;
        case 24: go<24, RAD>(a, s); return 0;
        default: return -1;
    }
}

int launch_grad(const GArgs& a, hipStream_t s) {
    return a.t.radial == QMC_RADIAL_STO ? bucket<QMC_RADIAL_STO>(a, s) : bucket<QMC_RADIAL_GTO_PURE>(a, s);
}

int launch_ao(const AArgs& a, bool deriv, hipStream_t s) {
    const dim3 g = grid_for(a.t, a.nw);
    if (a.t.radial == QMC_RADIAL_STO) {
        if (deriv)
            hipLaunchKernelGGL((ao_kernel<QMC_RADIAL_STO, true>), g, dim3(kBlock), 0, s, a);
        else
            hipLaunchKernelGGL((ao_kernel<QMC_RADIAL_STO, false>), g, dim3(kBlock), 0, s, a);
    } else {
        if (deriv)
            hipLaunchKernelGGL((ao_kernel<QMC_RADIAL_GTO_PURE, true>), g, dim3(kBlock), 0, s, a);
        else
            hipLaunchKernelGGL((ao_kernel<QMC_RADIAL_GTO_PURE, false>), g, dim3(kBlock), 0, s, a);
    }
    return 0;
}

template <int NM, int RAD>
static void go_e(const EArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((eloc_grad_kernel<NM, RAD>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
}

template <int RAD>
static int bucket_e(const EArgs& a, hipStream_t s) {
    switch (nm_bucket(a.t)) {
        case 2: go_e<2, RAD>(a, s); return 0;
        case 4: go_e<4, RAD>(a, s); return 0;
        case 8: go_e<8, RAD>(a, s); return 0;
        case 16: go_e<16, RAD>(a, s); return 0;
        case 24: go_e<24, RAD>(a, s); return 0;
        default: return -1;
    }
}

int launch_eloc_grad(const EArgs& a, hipStream_t s) {
    return a.t.radial == QMC_RADIAL_STO ? bucket_e<QMC_RADIAL_STO>(a, s) : bucket_e<QMC_RADIAL_GTO_PURE>(a, s);
}

}  // namespace qmc
