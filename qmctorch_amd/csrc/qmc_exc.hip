// Excitation (trace-trick) kernels for plans with more than QMC_MAXCONF unique
// configurations per spin: psi / local energy and the Metropolis sweep, both
// precisions, register buckets up to 16.
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

namespace qmc {

// dynamic LDS of one wavefront: 2 (G + 1) groups of M, X and Mk rows
static size_t exc_lds_bytes(const Tables& t, int nm) {
    const int G = 64 / t.nelec;
    return (size_t)2 * (G + 1) * 3 * t.nsig_max * nm * sizeof(double);
}

static dim3 exc_grid(const Tables& t, int64_t nw) {
    const int G = 64 / t.nelec;
    return dim3((unsigned)((nw + G - 1) / G));
}

template <int NM, int RAD>
static void go_wf(const KArgs& a, int prec, int mode, hipStream_t s) {
    const size_t lds = exc_lds_bytes(a.t, NM);
    const dim3 g = exc_grid(a.t, a.nw);
    if (prec == QMC_PREC_FP64) {
        if (mode == QMC_EVAL_ELOC)
            hipLaunchKernelGGL((wf_kernel_exc<double, NM, RAD, true>), g, dim3(64), lds, s, a);
        else
            hipLaunchKernelGGL((wf_kernel_exc<double, NM, RAD, false>), g, dim3(64), lds, s, a);
    } else {
        if (mode == QMC_EVAL_ELOC)
            hipLaunchKernelGGL((wf_kernel_exc<float, NM, RAD, true>), g, dim3(64), lds, s, a);
        else
            hipLaunchKernelGGL((wf_kernel_exc<float, NM, RAD, false>), g, dim3(64), lds, s, a);
    }
}

template <int NM, int RAD>
static void go_sweep(const SArgs& a, int prec, hipStream_t s) {
    const size_t lds = exc_lds_bytes(a.t, NM);
    const dim3 g = exc_grid(a.t, a.nw);
    if (prec == QMC_PREC_FP64)
        hipLaunchKernelGGL((sweep_kernel_exc<double, NM, RAD>), g, dim3(64), lds, s, a);
    else
        hipLaunchKernelGGL((sweep_kernel_exc<float, NM, RAD>), g, dim3(64), lds, s, a);
}

#define QMC_EXC_BUCKETS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(12) X(16)

template <>
int launch_wf_exc<QMC_RAD>(const KArgs& a, int prec, int mode, hipStream_t s) {
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)                          \
    case K:                                  \
        go_wf<K, QMC_RAD>(a, prec, mode, s); \
        return 0;
        QMC_EXC_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

template <>
int launch_sweep_exc<QMC_RAD>(const SArgs& a, int prec, hipStream_t s) {
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)                       \
    case K:                               \
        go_sweep<K, QMC_RAD>(a, prec, s); \
        return 0;
        QMC_EXC_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

}  // namespace qmc
