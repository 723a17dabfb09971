// Plan object and error reporting shared by the C-ABI translation units of
// libqmcmi355x (qmc_abi.hip, qmc_api.hip, qmc_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "qmc_launch.h"

struct QmcPlan {
    qmc::Tables t;
    const qmc::SpecEntry* spec = nullptr;  // kernels generated for this basis layout, if any
    bool use_spec = true;
    int nexp_grp = 0;  // exponents in group order (length of QmcParams.grp_exp)
    std::vector<void*> allocs;
    ~QmcPlan() {
        for (void* p : allocs) (void)hipFree(p);
    }
    template <typename X>
    const X* upload(const X* host, size_t n) {
        if (n == 0 || host == nullptr) return nullptr;
        void* d = nullptr;
        if (hipMalloc(&d, n * sizeof(X)) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        allocs.push_back(d);
        if (hipMemcpy(d, host, n * sizeof(X), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy failed");
        return (const X*)d;
    }
};

namespace qmc {
// Records msg as the calling thread's last error (qmc_last_error) and returns code.
int set_error(int code, const std::string& msg);
}  // namespace qmc

#define QMC_CHECK_HIP(call)                                                                              \
    do {                                                                                                 \
        hipError_t _e = (call);                                                                          \
        if (_e != hipSuccess)                                                                            \
            return qmc::set_error(QMC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e));       \
    } while (0)
