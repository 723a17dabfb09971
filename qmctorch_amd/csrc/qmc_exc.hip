// Excitation (trace-trick) kernels for plans with more than QMC_MAXCONF unique
// configurations per spin: psi / local energy and the Metropolis sweep (single
// plans and batched jobs), both precisions, register buckets up to 16.
#include "qmc_adjoint.h"
#include "qmc_launch.h"

#ifndef QMC_RAD
#error "compile with -DQMC_RAD=0 (gto_pure) or -DQMC_RAD=1 (sto)"
#endif

namespace qmc {

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

template <int NM, int RAD>
static void go_sweep_batch(const SJob* jobs, int njobs, int nwaves, size_t lds, int prec, hipStream_t s) {
    const dim3 g((unsigned)nwaves);
    if (prec == QMC_PREC_FP64)
        hipLaunchKernelGGL((sweep_batch_kernel_exc<double, NM, RAD>), g, dim3(64), lds, s, jobs, njobs, nwaves);
    else
        hipLaunchKernelGGL((sweep_batch_kernel_exc<float, NM, RAD>), g, dim3(64), lds, s, jobs, njobs, nwaves);
}

template <>
int launch_sweep_batch_exc<QMC_RAD>(const SJob* jobs, int njobs, int nwaves, int bucket, size_t lds, int prec,
                                    hipStream_t s) {
    switch (bucket) {
#define QMC_CASE(K)                                                    \
    case K:                                                            \
        go_sweep_batch<K, QMC_RAD>(jobs, njobs, nwaves, lds, prec, s); \
        return 0;
        QMC_EXC_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

template <int NM, int RAD>
static void go_wf_batch(const WJob* jobs, int njobs, int nwaves, size_t lds, int prec, int mode, hipStream_t s) {
    const dim3 g((unsigned)nwaves);
    if (prec == QMC_PREC_FP64) {
        if (mode == QMC_EVAL_ELOC)
            hipLaunchKernelGGL((wf_batch_kernel_exc<double, NM, RAD, true>), g, dim3(64), lds, s, jobs, njobs, nwaves);
        else
            hipLaunchKernelGGL((wf_batch_kernel_exc<double, NM, RAD, false>), g, dim3(64), lds, s, jobs, njobs,
                               nwaves);
    } else {
        if (mode == QMC_EVAL_ELOC)
            hipLaunchKernelGGL((wf_batch_kernel_exc<float, NM, RAD, true>), g, dim3(64), lds, s, jobs, njobs, nwaves);
        else
            hipLaunchKernelGGL((wf_batch_kernel_exc<float, NM, RAD, false>), g, dim3(64), lds, s, jobs, njobs, nwaves);
    }
}

template <>
int launch_wf_batch_exc<QMC_RAD>(const WJob* jobs, int njobs, int nwaves, int bucket, size_t lds, int prec, int mode,
                                 hipStream_t s) {
    switch (bucket) {
#define QMC_CASE(K)                                                       \
    case K:                                                               \
        go_wf_batch<K, QMC_RAD>(jobs, njobs, nwaves, lds, prec, mode, s); \
        return 0;
        QMC_EXC_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
}

template <int NM, int RAD>
static void go_grad_reduce(const GRedArgs& a, hipStream_t s) {
    const size_t lds = exc_lds_bytes(a.t, NM);
    hipLaunchKernelGGL((grad_reduce_kernel_exc<NM, RAD>), dim3((unsigned)a.r.nslot), dim3(64), lds, s, a);
}

template <>
int launch_grad_reduce_exc<QMC_RAD>(const GRedArgs& a, const RedOut& o, hipStream_t s) {
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)                          \
    case K:                                  \
        go_grad_reduce<K, QMC_RAD>(a, s);    \
        break;
        QMC_EXC_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
    red_slot_sum(a.t, a.r, o, s);
    return 0;
}

// static LDS of a kernel plus dynamic bytes within the device's per-block limit
template <typename K>
static bool lds_fits(K kernel, size_t dynamic) {
    hipFuncAttributes fa;
    int dev = 0, limit = 0;
    if (hipFuncGetAttributes(&fa, (const void*)kernel) != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return dynamic <= kExcEnergyAdjointLds;
    return fa.sharedSizeBytes + dynamic <= (size_t)limit;
}

template <int NM, int RAD>
static int go_eloc_reduce(const ERedArgs& a, hipStream_t s) {
    const size_t lds = 2 * exc_lds_bytes(a.t, NM);  // six row blocks per group
    if (lds > kExcEnergyAdjointLds || !lds_fits(eloc_grad_reduce_kernel_exc<NM, RAD>, lds)) return -2;
    hipLaunchKernelGGL((eloc_grad_reduce_kernel_exc<NM, RAD>), dim3((unsigned)a.r.nslot), dim3(64), lds, s, a);
    return 0;
}

template <>
int launch_eloc_reduce_exc<QMC_RAD>(const ERedArgs& a, const RedOut& o, hipStream_t s) {
    int rc = -1;
    switch (nm_bucket_fine(a.t)) {
#define QMC_CASE(K)                             \
    case K:                                     \
        rc = go_eloc_reduce<K, QMC_RAD>(a, s);  \
        break;
        QMC_EXC_BUCKETS(QMC_CASE)
#undef QMC_CASE
        default: return -1;
    }
    if (rc) return rc;
    red_slot_sum(a.t, a.r, o, s);
    return 0;
}

}  // namespace qmc
