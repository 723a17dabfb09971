// libqmcmi355x: qmc_logpsi_grad without a caller workspace.
//
// The partial-sum slots of the fused log|psi| adjoint live in a file-scope
// device array (one copy per GPU, held until the process ends) instead of a
// workspace argument, so callers that cannot allocate -- code under HIP graph
// capture, tight serving loops -- can run the adjoint.  The array's address is
// resolved once per device with hipGetSymbolAddress; host code never passes the
// symbol itself.  Every call on a GPU shares that GPU's array: calls that
// overlap in time on one GPU race (order them on one stream).  The kernels
// zero every slot they use at the start of each call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "qmc_adjoint.h"
#include "qmc_plan.h"

using namespace qmc;

namespace {

constexpr size_t kScratchDoubles = (size_t)QMC_NOALLOC_SCRATCH_BYTES / sizeof(double);

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (resolved on the first call per device).
int scratch_address(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(QMC_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_error(QMC_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(QMC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = (double*)g_scratch_ptr[dev];
    return QMC_OK;
}

}  // namespace

extern "C" {

int qmc_logpsi_grad_noalloc(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                            const double* weights, QmcGradOut* out, qmc_stream_t stream) {
    if (!plan || !params) return set_error(QMC_ERR_ARG, "null plan or params");
    if (nw < 0) return set_error(QMC_ERR_ARG, "negative walker count");
    if (!params->bas_exp || !params->atom_coords || !params->mo_w || !params->ci || !params->grp_exp ||
        !params->grp_expf)
        return set_error(QMC_ERR_ARG, "missing parameter pointer");
    if (plan->t.use_ee && !params->w_ee) return set_error(QMC_ERR_ARG, "e-e Jastrow weight missing");
    if (plan->t.use_en && !params->w_en) return set_error(QMC_ERR_ARG, "e-n Jastrow weight missing");
    if (!out || !out->grad_w || !out->grad_ci || !out->grad_jas)
        return set_error(QMC_ERR_ARG, "null gradient output");
    if (nw == 0) return qmc_logpsi_grad(plan, params, pos, 0, weights, out, nullptr, 0, stream);
    if (!pos || !weights) return set_error(QMC_ERR_ARG, "null pos or weights");
    double* ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    const RedSlots r = red_layout(plan->t, nw, out->grad_exp != nullptr, ws, QMC_NOALLOC_SCRATCH_BYTES, weights);
    if (r.nslot < 1)
        return set_error(QMC_ERR_UNSUPPORTED,
                         "one partial-sum slot of this system exceeds the static scratch of "
                         "QMC_NOALLOC_SCRATCH_BYTES; use qmc_logpsi_grad with a workspace");
    GRedArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.r = r;
    RedOut o{out->grad_w, out->grad_ci, out->grad_jas, out->grad_exp};
    hipStream_t s = (hipStream_t)stream;
    const int miss = a.t.exc_mode ? QMC_BY_RAD(launch_grad_reduce_exc, a.t.radial, a, o, s)
                                  : QMC_BY_RAD(launch_grad_reduce, a.t.radial, a, o, s);
    if (miss)
        return set_error(QMC_ERR_UNSUPPORTED, a.t.exc_mode ? "more than 16 MO columns per electron (excitation path)"
                                                           : "more than 24 occupied orbitals per spin");
    QMC_CHECK_HIP(hipGetLastError());
    return QMC_OK;
}

}  // extern "C"
