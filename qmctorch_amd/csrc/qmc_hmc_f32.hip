// Hamiltonian Monte Carlo and drift kernels, float bulk precision.
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif
#ifndef QMC_HMC_MC
#error "compile with -DQMC_HMC_MC=1 (single configuration) or -DQMC_HMC_MC=QMC_MAXCONF"
#endif

namespace qmc {

template <int NM, int RAD>
static void go(const HArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((hmc_kernel<float, NM, RAD, QMC_HMC_MC>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
}

template <int RAD>
static int bucket(const HArgs& a, hipStream_t s) {
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)       \
    case K:               \
        go<K, RAD>(a, s); \
        return 0;
        QMC_FINE_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

template <>
int launch_hmc_f32<QMC_RAD, QMC_HMC_MC>(const HArgs& a, hipStream_t s) {
    return bucket<QMC_RAD>(a, s);
}

}  // namespace qmc
