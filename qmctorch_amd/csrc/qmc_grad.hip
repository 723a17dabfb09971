// Adjoint-row and AO-layout kernels (FP64).
#include "qmc_adjoint.h"
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

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

template <>
int launch_grad<QMC_RAD>(const GArgs& a, hipStream_t s) {
    return bucket<QMC_RAD>(a, s);
}

template <>
int launch_ao<QMC_RAD>(const AArgs& a, bool deriv, hipStream_t s) {
    const dim3 g = grid_for(a.t, a.nw);
    if (deriv)
        hipLaunchKernelGGL((ao_kernel<QMC_RAD, true>), g, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((ao_kernel<QMC_RAD, false>), g, dim3(kBlock), 0, s, a);
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

template <>
int launch_eloc_grad<QMC_RAD>(const EArgs& a, hipStream_t s) {
    return bucket_e<QMC_RAD>(a, s);
}

template <>
int launch_grad_reduce<QMC_RAD>(const GRedArgs& a, const RedOut& o, hipStream_t s) {
    const dim3 grid = grid_for_waves(a.r.nslot);
    switch (nm_bucket(a.t)) {
        case 2: hipLaunchKernelGGL((grad_reduce_kernel<2, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 4: hipLaunchKernelGGL((grad_reduce_kernel<4, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 8: hipLaunchKernelGGL((grad_reduce_kernel<8, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 16: hipLaunchKernelGGL((grad_reduce_kernel<16, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 24: hipLaunchKernelGGL((grad_reduce_kernel<24, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        default: return -1;
    }
    red_slot_sum(a.t, a.r, o, s);
    return 0;
}

template <>
int launch_eloc_reduce<QMC_RAD>(const ERedArgs& a, const RedOut& o, hipStream_t s) {
    const dim3 grid = grid_for_waves(a.r.nslot);
    switch (nm_bucket(a.t)) {
        case 2: hipLaunchKernelGGL((eloc_grad_reduce_kernel<2, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 4: hipLaunchKernelGGL((eloc_grad_reduce_kernel<4, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 8: hipLaunchKernelGGL((eloc_grad_reduce_kernel<8, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 16: hipLaunchKernelGGL((eloc_grad_reduce_kernel<16, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        case 24: hipLaunchKernelGGL((eloc_grad_reduce_kernel<24, QMC_RAD>), grid, dim3(kBlock), 0, s, a); break;
        default: return -1;
    }
    red_slot_sum(a.t, a.r, o, s);
    return 0;
}

}  // namespace qmc
