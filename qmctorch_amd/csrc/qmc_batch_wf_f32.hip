// Batched psi evaluation (many systems per launch), float bulk precision (local energies: FP64 units).
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

namespace qmc {

template <int NM, int RAD>
static void go(const WJob* jobs, int njobs, int nwaves, int mode, hipStream_t s) {
    // psi only: local energies are evaluated by the FP64 kernels in every precision
    (void)mode;
    hipLaunchKernelGGL((wf_batch_kernel<float, NM, RAD, false>), grid_for_waves(nwaves), dim3(kBlock), 0, s, jobs,
                       njobs, nwaves);
}

template <>
int launch_wf_batch_f32<QMC_RAD>(const WJob* jobs, int njobs, int nwaves, int bucket, int mode, hipStream_t s) {
    switch (bucket) {
#define QMC_CASE(K)                  \
    case K:                          \
        go<K, QMC_RAD>(jobs, njobs, nwaves, mode, s); \
        return 0;
        QMC_FINE_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

}  // namespace qmc
