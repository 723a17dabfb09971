// Drift and Hamiltonian Monte Carlo for plans with more than QMC_MAXCONF
// unique configurations per spin (the excitation path's Gamma rows), both
// precisions, register buckets up to 16.
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

namespace qmc {

template <int NM, int RAD>
static void go_hmc(const HArgs& a, int prec, hipStream_t s) {
    const size_t lds = exc_lds_bytes(a.t, NM);
    const dim3 g = exc_grid(a.t, a.nw);
    if (prec == QMC_PREC_FP64)
        hipLaunchKernelGGL((hmc_kernel<double, NM, RAD, 1, AoGeneric, true>), g, dim3(64), lds, s, a);
    else
        hipLaunchKernelGGL((hmc_kernel<float, NM, RAD, 1, AoGeneric, true>), g, dim3(64), lds, s, a);
}

template <>
int launch_hmc_exc<QMC_RAD>(const HArgs& a, int prec, hipStream_t s) {
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)                     \
    case K:                             \
        go_hmc<K, QMC_RAD>(a, prec, s); \
        return 0;
        QMC_EXC_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

}  // namespace qmc
