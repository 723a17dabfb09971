// Persistent Metropolis sweep kernels, double bulk precision.
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

namespace qmc {

template <int NM, int RAD>
static void go(const SArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((sweep_kernel<double, NM, RAD>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
}

template <int RAD>
static int bucket(const SArgs& a, hipStream_t s) {
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)           \
    case K:                   \
        go<K, RAD>(a, s); \
        return 0;
        QMC_FINE_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

template <>
int launch_sweep_f64<QMC_RAD>(const SArgs& a, hipStream_t s) {
    return bucket<QMC_RAD>(a, s);
}

}  // namespace qmc
