// Batched persistent Metropolis sweeps (many systems per launch), float bulk precision.
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

namespace qmc {

template <int NM, int RAD>
static void go(const SJob* jobs, int njobs, int nwaves, hipStream_t s) {
    hipLaunchKernelGGL((sweep_batch_kernel<float, NM, RAD>), grid_for_waves(nwaves), dim3(kBlock), 0, s, jobs, njobs,
                       nwaves);
}

template <>
int launch_sweep_batch_f32<QMC_RAD>(const SJob* jobs, int njobs, int nwaves, int bucket, hipStream_t s) {
    switch (bucket) {
#define QMC_CASE(K)                  \
    case K:                          \
        go<K, QMC_RAD>(jobs, njobs, nwaves, s); \
        return 0;
        QMC_FINE_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

}  // namespace qmc
