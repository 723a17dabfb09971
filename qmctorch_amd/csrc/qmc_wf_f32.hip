// psi kernels, float bulk precision (local energies: qmc_wf_f64.hip).
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

namespace qmc {

template <int NM, int RAD>
static void go(const KArgs& a, int mode, hipStream_t s) {
    // psi only: local energies are evaluated by the FP64 kernels in every precision
    (void)mode;
    hipLaunchKernelGGL((wf_kernel<float, NM, RAD, false>), grid_for(a.t, a.nw), dim3(kBlock), 0, s, a);
}

template <int RAD>
static int bucket(const KArgs& a, int mode, hipStream_t s) {
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)           \
    case K:                   \
        go<K, RAD>(a, mode, s); \
        return 0;
        QMC_FINE_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

template <>
int launch_wf_f32<QMC_RAD>(const KArgs& a, int mode, hipStream_t s) {
    return bucket<QMC_RAD>(a, mode, s);
}

}  // namespace qmc
