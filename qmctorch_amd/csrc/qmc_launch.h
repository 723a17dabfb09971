// Host launchers of the kernel families, one translation unit each so the
// library builds in parallel.  Every launcher returns 0, or -1 when the
// orbital count has no compiled bucket.
#pragma once
#include "qmc_kernels.h"

namespace qmc {

// register-array bucket for max(nmo, nup)
inline int nm_bucket(const Tables& t) {
    const int need = t.nmo > t.nup ? t.nmo : t.nup;
    if (need <= 2) return 2;
    if (need <= 4) return 4;
    if (need <= 8) return 8;
    if (need <= 16) return 16;
    if (need <= 24) return 24;
    return -1;
}

// register-array bound of the psi / local-energy / sweep kernels: exact up to
// 8 occupied orbitals per spin, so the lane elimination and the MO rows carry
// no idle columns, then 12, 16, 24
inline int nm_bucket_fine(const Tables& t) {
    int need = t.nmo > t.nup ? t.nmo : t.nup;
    need = need > t.ndown ? need : t.ndown;
    if (need < 1) need = 1;
    if (need <= 8) return need;
    if (need <= 12) return 12;
    if (need <= 16) return 16;
    if (need <= 24) return 24;
    return -1;
}
#define QMC_FINE_BUCKETS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(12) X(16) X(24)

inline dim3 grid_for(const Tables& t, int64_t nw) {
    const int G = 64 / t.nelec;
    const int64_t waves = (nw + G - 1) / G;
    return dim3((unsigned)((waves + (kBlock / 64) - 1) / (kBlock / 64)));
}

// Each family is compiled once per radial type (-DQMC_RAD=0 gto_pure, 1 sto);
// the specialisations live in different translation units.
template <int RAD> int launch_wf_f32(const KArgs& a, int mode, hipStream_t s);
template <int RAD> int launch_wf_f64(const KArgs& a, int mode, hipStream_t s);
template <int RAD> int launch_sweep_f32(const SArgs& a, hipStream_t s);
template <int RAD> int launch_sweep_f64(const SArgs& a, hipStream_t s);
template <int RAD> int launch_grad(const GArgs& a, hipStream_t s);
template <int RAD> int launch_ao(const AArgs& a, bool deriv, hipStream_t s);
template <int RAD> int launch_eloc_grad(const EArgs& a, hipStream_t s);
// fused adjoints (qmc_adjoint.h): the reduction kernel plus the slot sum
struct GRedArgs;
struct ERedArgs;
struct RedOut;
template <int RAD> int launch_grad_reduce(const GRedArgs& a, const RedOut& o, hipStream_t s);
template <int RAD> int launch_eloc_reduce(const ERedArgs& a, const RedOut& o, hipStream_t s);
// the log|psi| adjoint, the drift and HMC of plans with exc_mode (register
// buckets up to 16)
template <int RAD> int launch_grad_reduce_exc(const GRedArgs& a, const RedOut& o, hipStream_t s);
template <int RAD> int launch_hmc_exc(const HArgs& a, int precision, hipStream_t s);
// the local-energy adjoint of plans with exc_mode; -2 when its LDS rows
// (2 (G + 1) groups of 6 x nsig_max x bucket doubles) exceed kExcEnergyAdjointLds
template <int RAD> int launch_eloc_reduce_exc(const ERedArgs& a, const RedOut& o, hipStream_t s);
constexpr size_t kExcEnergyAdjointLds = 64 * 1024;
// excitation (trace-trick) kernels, any precision; register buckets up to 16
template <int RAD> int launch_wf_exc(const KArgs& a, int precision, int mode, hipStream_t s);
template <int RAD> int launch_sweep_exc(const SArgs& a, int precision, hipStream_t s);
// batched excitation-plan launches: jobs (device) of one register bucket, one
// wavefront per block, lds = the largest job's exc_lds_bytes
template <int RAD>
int launch_sweep_batch_exc(const SJob* jobs, int njobs, int nwaves, int bucket, size_t lds, int precision,
                           hipStream_t s);
template <int RAD>
int launch_wf_batch_exc(const WJob* jobs, int njobs, int nwaves, int bucket, size_t lds, int precision, int mode,
                        hipStream_t s);
// HMC / drift: MC = 1 for single-configuration plans, QMC_MAXCONF otherwise
template <int RAD, int MC> int launch_hmc_f32(const HArgs& a, hipStream_t s);
template <int RAD, int MC> int launch_hmc_f64(const HArgs& a, hipStream_t s);
// batched launches: jobs (device) sorted into one register bucket, nwaves in total
template <int RAD> int launch_sweep_batch_f32(const SJob* jobs, int njobs, int nwaves, int bucket, hipStream_t s);
template <int RAD> int launch_sweep_batch_f64(const SJob* jobs, int njobs, int nwaves, int bucket, hipStream_t s);
template <int RAD>
int launch_wf_batch_f32(const WJob* jobs, int njobs, int nwaves, int bucket, int mode, hipStream_t s);
template <int RAD>
int launch_wf_batch_f64(const WJob* jobs, int njobs, int nwaves, int bucket, int mode, hipStream_t s);

inline dim3 grid_for_waves(int nwaves) { return dim3((unsigned)((nwaves + (kBlock / 64) - 1) / (kBlock / 64))); }

// Excitation-path launches: one wavefront per block, its dynamic LDS holding
// 2 (G + 1) groups (walker, spin) of three nsig_max x nm row blocks.
inline size_t exc_lds_bytes(const Tables& t, int nm) {
    const int G = 64 / t.nelec;
    return (size_t)2 * (G + 1) * 3 * t.nsig_max * nm * sizeof(double);
}
inline dim3 exc_grid(const Tables& t, int64_t nw) {
    const int G = 64 / t.nelec;
    return dim3((unsigned)((nw + G - 1) / G));
}
#define QMC_EXC_BUCKETS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(12) X(16)

// Kernels built for one basis layout (generated units, specialize.py).  Each
// launcher returns 0, or -1 when the layout has no instance for that
// precision / mode (the caller then uses the generic kernels).
struct SpecEntry {
    uint64_t signature;  // layout_signature of the layout
    int nm;              // register bucket the kernels were built for
    int max_rows;        // lanes of real walkers per wavefront (G Ne) the sweeps' matrix tiles cover
    const char* name;
    int (*sweep)(const SArgs& a, int precision, hipStream_t s);
    int (*sweep_batch)(const SJob* jobs, int njobs, int nwaves, int precision, hipStream_t s);
    int (*wf)(const KArgs& a, int precision, int mode, hipStream_t s);  // FP64 local energy
    int (*wf_batch)(const WJob* jobs, int njobs, int nwaves, int precision, int mode, hipStream_t s);
    int (*hmc)(const HArgs& a, int precision, hipStream_t s);  // mixed HMC / drift
};

// the generated layout with this signature and bucket whose kernels cover
// `rows` lanes of real walkers per wavefront, or nullptr
const SpecEntry* find_spec(uint64_t signature, int nm, int rows);

#define QMC_BY_RAD(fn, rad, ...) ((rad) == QMC_RADIAL_STO ? fn<QMC_RADIAL_STO>(__VA_ARGS__) : fn<QMC_RADIAL_GTO_PURE>(__VA_ARGS__))

}  // namespace qmc
