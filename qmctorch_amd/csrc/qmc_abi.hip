// libqmcmi355x: C ABI of the QMCTorch VMC hot path on MI355X (gfx950).
// Plan lifetime, argument checks and dispatch to the kernel families
// (qmc_wf_*.hip, qmc_sweep_*.hip, qmc_grad.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "qmc_adjoint.h"
#include "qmc_plan.h"

using namespace qmc;

namespace {
thread_local std::string g_last_error;
}  // namespace

int qmc::set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

namespace {
int fail(int code, const std::string& msg) { return qmc::set_error(code, msg); }

#define QMC_HIP(call)                                                                                    \
    do {                                                                                                 \
        hipError_t _e = (call);                                                                          \
        if (_e != hipSuccess) return fail(QMC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

const char* kNoBucket = "more than 24 occupied orbitals per spin";
const char* kNoExcBucket = "the excitation path supports up to 16 MO columns per electron";

__global__ void philox_kernel(const uint32_t* in, uint32_t* out, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* c = in + 6 * i;
    const u4 r = philox4x32_10(u4{c[0], c[1], c[2], c[3]}, c[4], c[5]);
    out[4 * i + 0] = r.x;
    out[4 * i + 1] = r.y;
    out[4 * i + 2] = r.z;
    out[4 * i + 3] = r.w;
}
}  // namespace


namespace {

// FNV-1a over the layout fields a generated AO row bakes in (specialize.py
// computes the same value over the same fields, 32-bit little-endian words).
struct Fnv {
    uint64_t h = 0xCBF29CE484222325ull;
    void word(int32_t v) {
        const uint32_t u = (uint32_t)v;
        for (int b = 0; b < 4; ++b) {
            h ^= (u >> (8 * b)) & 0xFFu;
            h *= 0x100000001B3ull;
        }
    }
    void words(const int32_t* v, int n) {
        for (int i = 0; i < n; ++i) word(v ? v[i] : 0);
    }
};

uint64_t layout_signature(const QmcSystemDesc* d) {
    Fnv f;
    f.word(d->radial_type);
    f.word(d->natom);
    f.word(d->ngroup);
    f.words(d->grp_atom, d->ngroup);
    f.words(d->grp_nprim, d->ngroup);
    f.words(d->grp_prim0, d->ngroup);
    f.words(d->grp_ao0, d->ngroup);
    f.words(d->grp_nao, d->ngroup);
    f.words(d->grp_tight, d->ngroup);
    f.word(d->nentry);
    f.words(d->gao_mu, d->nentry);
    f.words(d->gao_k, d->nentry);
    // coefficients and r powers (literals of a generated row), as 32-bit words
    for (const double* arr : {d->gao_K, d->gao_n}) {
        const size_t n = (size_t)d->nentry * QMC_MAXP;
        for (size_t k = 0; k < n; ++k) {
            uint64_t u = 0;
            if (arr) std::memcpy(&u, arr + k, sizeof(u));
            f.word((int32_t)(uint32_t)(u & 0xFFFFFFFFu));
            f.word((int32_t)(uint32_t)(u >> 32));
        }
    }
    return f.h;
}

}  // namespace

namespace {
// Range check of every count and index table of a system description: the
// kernels index registers and device tables with these values unchecked.
// Returns the reason of the first violation, or nullptr.
const char* invalid_tables(const QmcSystemDesc* d) {
    if (d->natom < 1 || d->nao < 1 || d->nprim_ref < 1 || d->ngroup < 1 || d->nentry < 1 || d->nmo < 1 ||
        d->nci < 1)
        return "natom, nao, nprim_ref, ngroup, nentry, nmo and nci must be positive";
    if (d->nmo < d->ndown) return "nmo smaller than ndown";
    if (!d->grp_atom || !d->grp_nprim || !d->grp_prim0 || !d->grp_prim_src || !d->grp_ao0 || !d->grp_nao ||
        !d->gao_mu || !d->gao_k || !d->gao_K || !d->gao_n || !d->gao_src || !d->atom_z || !d->jastrow_atoms ||
        !d->conf_up || !d->ci_up || !d->ci_dn || (d->ndown > 0 && !d->conf_dn))
        return "missing table pointer";
    std::vector<int> entry_group(d->nentry, -1);
    int next_entry = 0;  // the groups' entries follow each other in group order
    for (int g = 0; g < d->ngroup; ++g) {
        if (d->grp_ao0[g] != next_entry) return "AO entries of the groups must follow in group order";
        next_entry += std::max(d->grp_nao[g], 0);
        if (d->grp_atom[g] < 0 || d->grp_atom[g] >= d->natom) return "grp_atom out of range";
        const int np = d->grp_nprim[g];
        if (np < 1 || np > QMC_MAXP) return "grp_nprim out of range";
        if (d->grp_prim0[g] < 0) return "negative grp_prim0";
        for (int q = 0; q < np; ++q) {
            const int32_t s = d->grp_prim_src[d->grp_prim0[g] + q];
            if (s < 0 || s >= d->nprim_ref) return "grp_prim_src out of range";
        }
        if (d->grp_ao0[g] < 0 || d->grp_nao[g] < 1 || (int64_t)d->grp_ao0[g] + d->grp_nao[g] > d->nentry)
            return "AO entries of a group out of range";
        for (int i = d->grp_ao0[g]; i < d->grp_ao0[g] + d->grp_nao[g]; ++i) {
            if (entry_group[i] >= 0) return "an AO entry belongs to two groups";
            entry_group[i] = g;
        }
    }
    for (int i = 0; i < d->nentry; ++i) {
        if (entry_group[i] < 0) return "an AO entry belongs to no group";
        if (d->gao_mu[i] < 0 || d->gao_mu[i] >= d->nao) return "gao_mu out of range";
        const int k = d->gao_k[i];
        if (k < 0 || (k & 0xF) > 3 || ((k >> 4) & 0xF) > 3 || ((k >> 8) & 0xF) > 3 || (k >> 12) != 0)
            return "gao_k: Cartesian powers must be in [0, 3]";
        for (int q = 0; q < QMC_MAXP; ++q) {
            const int32_t s = d->gao_src[(size_t)i * QMC_MAXP + q];
            if (s < 0 || s >= d->nprim_ref) return "gao_src out of range";
        }
    }
    for (int j = 0; j < d->nconf_up * d->nup; ++j)
        if (d->conf_up[j] < 0 || d->conf_up[j] >= d->nmo) return "conf_up column out of range";
    if (d->ndown > 0)
        for (int j = 0; j < d->nconf_dn * d->ndown; ++j)
            if (d->conf_dn[j] < 0 || d->conf_dn[j] >= d->nmo) return "conf_dn column out of range";
    for (int n = 0; n < d->nci; ++n) {
        if (d->ci_up[n] < 0 || d->ci_up[n] >= d->nconf_up) return "ci_up out of range";
        if (d->ci_dn[n] < 0 || d->ci_dn[n] >= d->nconf_dn) return "ci_dn out of range";
    }
    if (d->grp_tight)
        for (int g = 0; g < d->ngroup; ++g)
            if (d->grp_tight[g] != 0 && d->grp_tight[g] != 1) return "grp_tight must be 0 or 1";
    return nullptr;
}

// Excitation table of nconf configurations of n columns each (row c of conf
// lists its MO columns in the order of the determinant): per configuration
// {level, h1, h2, p1, p2, sign} with the holes h (ground-state columns 0..n-1
// missing, ascending), the particles p (columns >= n, ascending, stored as
// p - n) and the sign of the listed order relative to the ground-state order
// with h_i replaced in place by p_i.  False when a configuration is more than
// a double excitation of the ground state or repeats a column.
bool excitation_table(const int32_t* conf, int nconf, int n, std::vector<int32_t>& out) {
    out.assign((size_t)nconf * 6, 0);
    for (int c = 0; c < nconf; ++c) {
        const int32_t* cc = conf + (size_t)c * n;
        int holes[2] = {0, 0}, parts[2] = {0, 0}, nh = 0, np = 0;
        for (int g = 0; g < n; ++g) {
            bool in = false;
            for (int j = 0; j < n; ++j) in = in || cc[j] == g;
            if (!in) {
                if (nh == 2) return false;
                holes[nh++] = g;
            }
        }
        for (int j = 0; j < n; ++j) {
            if (cc[j] < 0) return false;
            if (cc[j] >= n) {
                if (np == 2) return false;
                parts[np++] = cc[j];
            }
        }
        if (nh != np) return false;
        if (np == 2 && parts[0] > parts[1]) std::swap(parts[0], parts[1]);
        if (np == 2 && parts[0] == parts[1]) return false;
        std::vector<int> inplace(n);
        for (int g = 0; g < n; ++g) inplace[g] = g;
        for (int i = 0; i < nh; ++i) inplace[holes[i]] = parts[i];
        std::vector<int> perm(n, -1);
        for (int j = 0; j < n; ++j)
            for (int g = 0; g < n; ++g)
                if (inplace[g] == cc[j]) perm[j] = g;
        int inv = 0;
        for (int i = 0; i < n; ++i) {
            if (perm[i] < 0) return false;
            for (int j = i + 1; j < n; ++j) inv += perm[i] > perm[j];
        }
        int32_t* o = out.data() + (size_t)c * 6;
        o[0] = nh;
        o[1] = holes[0];
        o[2] = holes[1];
        o[3] = nh > 0 ? parts[0] - n : 0;
        o[4] = nh > 1 ? parts[1] - n : 0;
        o[5] = (inv & 1) ? -1 : 1;
    }
    return true;
}

// Entry points without the trace-trick path refuse plans that need it.
int refuse_excitations(const QmcPlan* plan) {
    if (plan && plan->t.exc_mode)
        return fail(QMC_ERR_UNSUPPORTED, "more than 8 unique configurations per spin: this entry point has no "
                                         "excitation (trace-trick) path (psi, pdf, local energy, the Metropolis "
                                         "sweep, the drift, HMC, the log|psi| and local-energy adjoints and the "
                                         "batched launches have)");
    return QMC_OK;
}

int launch_hmc(const HArgs& a, int precision, hipStream_t s) {
    const bool f64 = precision == QMC_PREC_FP64, sto = a.t.radial == QMC_RADIAL_STO;
    if (a.t.exc_mode) return QMC_BY_RAD(launch_hmc_exc, a.t.radial, a, precision, s);
    if (a.t.nconf_max <= 1) {
        if (f64) return sto ? launch_hmc_f64<QMC_RADIAL_STO, 1>(a, s) : launch_hmc_f64<QMC_RADIAL_GTO_PURE, 1>(a, s);
        return sto ? launch_hmc_f32<QMC_RADIAL_STO, 1>(a, s) : launch_hmc_f32<QMC_RADIAL_GTO_PURE, 1>(a, s);
    }
    if (f64)
        return sto ? launch_hmc_f64<QMC_RADIAL_STO, QMC_MAXCONF>(a, s)
                   : launch_hmc_f64<QMC_RADIAL_GTO_PURE, QMC_MAXCONF>(a, s);
    return sto ? launch_hmc_f32<QMC_RADIAL_STO, QMC_MAXCONF>(a, s)
               : launch_hmc_f32<QMC_RADIAL_GTO_PURE, QMC_MAXCONF>(a, s);
}
}  // namespace

extern "C" {

const char* qmc_last_error(void) { return g_last_error.c_str(); }

const char* qmc_version(void) { return "libqmcmi355x 0.1 (gfx950)"; }

int qmc_plan_create(const QmcSystemDesc* d, QmcPlan** out) {
    if (!d || !out) return fail(QMC_ERR_ARG, "null argument");
    if (d->nelec < 1 || d->nelec > 64) return fail(QMC_ERR_UNSUPPORTED, "nelec must be in [1, 64]");
    if (d->nup + d->ndown != d->nelec || d->ndown > d->nup) return fail(QMC_ERR_ARG, "inconsistent spin counts");
    if (d->radial_type != QMC_RADIAL_GTO_PURE && d->radial_type != QMC_RADIAL_STO)
        return fail(QMC_ERR_UNSUPPORTED, "radial type must be gto_pure or sto");
    if (d->nconf_up < 1 || d->nconf_up > QMC_MAXEXC || d->nconf_dn < 1 || d->nconf_dn > QMC_MAXEXC)
        return fail(QMC_ERR_UNSUPPORTED, "unique configurations per spin must be in [1, 1024]");
    // every index table is range-checked before anything reaches the device
    if (const char* why = invalid_tables(d)) return fail(QMC_ERR_ARG, why);
    // beyond QMC_MAXCONF per spin: single and double excitations of the ground
    // state through the trace-trick kernels
    const bool exc = d->nconf_up > QMC_MAXCONF || d->nconf_dn > QMC_MAXCONF;
    std::vector<int32_t> exc_up, exc_dn;
    if (exc) {
        if (!excitation_table(d->conf_up, d->nconf_up, d->nup, exc_up) ||
            !excitation_table(d->conf_dn, d->ndown > 0 ? d->nconf_dn : 0, d->ndown, exc_dn))
            return fail(QMC_ERR_UNSUPPORTED, "more than 8 unique configurations per spin are supported as single and "
                                             "double excitations of the ground state only");
        if (exc_dn.empty()) exc_dn.assign(6, 0);
        if (d->nelec < 2) return fail(QMC_ERR_UNSUPPORTED, "the excitation path needs at least two electrons");
        if (std::max(d->nmo, std::max(d->nup, d->ndown)) > 16)
            return fail(QMC_ERR_UNSUPPORTED, "the excitation path supports up to 16 MO columns per electron");
    }
    if (d->nmo < d->nup) return fail(QMC_ERR_ARG, "nmo smaller than nup");
    for (int g = 0; g < d->ngroup; ++g)
        if (d->grp_nprim[g] > QMC_MAXP || d->grp_nprim[g] < 1) return fail(QMC_ERR_UNSUPPORTED, "group too large");
    QmcPlan* p = nullptr;
    try {
        p = new QmcPlan();
        Tables& t = p->t;
        std::memset(&t, 0, sizeof(t));
        t.nelec = d->nelec;
        t.nup = d->nup;
        t.ndown = d->ndown;
        t.natom = d->natom;
        t.nao = d->nao;
        t.nprim_ref = d->nprim_ref;
        t.radial = d->radial_type;
        t.ngroup = d->ngroup;
        t.nentry = d->nentry;
        t.nmo = d->nmo;
        t.nci = d->nci;
        t.nconf_up = d->nconf_up;
        t.nconf_dn = d->nconf_dn;
        t.nconf_max = d->nconf_up > d->nconf_dn ? d->nconf_up : d->nconf_dn;
        t.nsig_max = d->nup;
        t.use_ee = d->use_jastrow_ee;
        t.use_en = d->use_jastrow_en;
        t.vnn = d->vnn;
        // identity configuration: a single configuration per spin with columns 0..n-1
        bool ident = d->nconf_up == 1 && d->nconf_dn == 1 && d->nci == 1;
        for (int j = 0; ident && j < d->nup; ++j) ident = d->conf_up[j] == j;
        for (int j = 0; ident && j < d->ndown; ++j) ident = d->conf_dn[j] == j;
        t.identity_conf = ident ? 1 : 0;
        int nprim_tot = 0;
        for (int g = 0; g < d->ngroup; ++g) nprim_tot = std::max(nprim_tot, d->grp_prim0[g] + d->grp_nprim[g]);
        t.grp_atom = (decltype(t.grp_atom))p->upload(d->grp_atom, d->ngroup);
        t.grp_nprim = (decltype(t.grp_nprim))p->upload(d->grp_nprim, d->ngroup);
        t.grp_prim0 = (decltype(t.grp_prim0))p->upload(d->grp_prim0, d->ngroup);
        t.grp_prim_src = (decltype(t.grp_prim_src))p->upload(d->grp_prim_src, nprim_tot);
        p->nexp_grp = nprim_tot;
        t.grp_ao0 = (decltype(t.grp_ao0))p->upload(d->grp_ao0, d->ngroup);
        t.grp_nao = (decltype(t.grp_nao))p->upload(d->grp_nao, d->ngroup);
        t.grp_tight = (decltype(t.grp_tight))p->upload(d->grp_tight, d->grp_tight ? d->ngroup : 0);
        t.gao_mu = (decltype(t.gao_mu))p->upload(d->gao_mu, d->nentry);
        t.gao_k = (decltype(t.gao_k))p->upload(d->gao_k, d->nentry);
        t.gao_K = (decltype(t.gao_K))p->upload(d->gao_K, (size_t)d->nentry * QMC_MAXP);
        {
            std::vector<float> kf((size_t)d->nentry * QMC_MAXP);
            for (size_t i = 0; i < kf.size(); ++i) kf[i] = (float)d->gao_K[i];
            t.gao_Kf = (decltype(t.gao_Kf))p->upload(kf.data(), kf.size());
        }
        t.gao_n = (decltype(t.gao_n))p->upload(d->gao_n, (size_t)d->nentry * QMC_MAXP);
        t.gao_src = (decltype(t.gao_src))p->upload(d->gao_src, (size_t)d->nentry * QMC_MAXP);
        t.atom_z = (decltype(t.atom_z))p->upload(d->atom_z, d->natom);
        t.jas_atoms = (decltype(t.jas_atoms))p->upload(d->jastrow_atoms, (size_t)d->natom * 3);
        t.conf_up = (decltype(t.conf_up))p->upload(d->conf_up, (size_t)d->nconf_up * d->nup);
        t.conf_dn = (decltype(t.conf_dn))p->upload(d->conf_dn, (size_t)d->nconf_dn * (d->ndown > 0 ? d->ndown : 1));
        t.ci_up = (decltype(t.ci_up))p->upload(d->ci_up, d->nci);
        t.ci_dn = (decltype(t.ci_dn))p->upload(d->ci_dn, d->nci);
        if (!t.conf_dn) t.conf_dn = t.conf_up;
        {
            // inverse maps for the adjoint slot sum: entries per AO, and the
            // real exponent terms (a group's own primitives) per bas_exp index
            std::vector<int32_t> ao_ptr(d->nao + 1, 0), ao_ent(d->nentry);
            for (int i = 0; i < d->nentry; ++i) ++ao_ptr[d->gao_mu[i] + 1];
            for (int mu = 0; mu < d->nao; ++mu) ao_ptr[mu + 1] += ao_ptr[mu];
            {
                std::vector<int32_t> fill(ao_ptr.begin(), ao_ptr.end() - 1);
                for (int i = 0; i < d->nentry; ++i) ao_ent[fill[d->gao_mu[i]]++] = i;
            }
            std::vector<int32_t> exp_ptr(d->nprim_ref + 1, 0), exp_col;
            std::vector<std::pair<int32_t, int32_t>> terms;  // (bas_exp index, column)
            for (int g = 0; g < d->ngroup; ++g)
                for (int i = d->grp_ao0[g]; i < d->grp_ao0[g] + d->grp_nao[g]; ++i)
                    for (int q = 0; q < d->grp_nprim[g]; ++q)
                        terms.emplace_back(d->gao_src[(size_t)i * QMC_MAXP + q], i * QMC_MAXP + q);
            std::stable_sort(terms.begin(), terms.end(),
                             [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
                                 return a.first < b.first;
                             });
            for (const auto& tc : terms) {
                ++exp_ptr[tc.first + 1];
                exp_col.push_back(tc.second);
            }
            for (int k = 0; k < d->nprim_ref; ++k) exp_ptr[k + 1] += exp_ptr[k];
            t.ao_ptr = (decltype(t.ao_ptr))p->upload(ao_ptr.data(), ao_ptr.size());
            t.ao_ent = (decltype(t.ao_ent))p->upload(ao_ent.data(), ao_ent.size());
            t.exp_ptr = (decltype(t.exp_ptr))p->upload(exp_ptr.data(), exp_ptr.size());
            t.exp_col = (decltype(t.exp_col))p->upload(exp_col.data(), exp_col.size());
        }
        t.exc_mode = exc ? 1 : 0;
        if (exc) {
            t.exc_up = (decltype(t.exc_up))p->upload(exc_up.data(), exc_up.size());
            t.exc_dn = (decltype(t.exc_dn))p->upload(exc_dn.data(), exc_dn.size());
        }
        // generated kernels have no excitation path
        p->spec = exc ? nullptr : find_spec(layout_signature(d), nm_bucket_fine(t), (64 / t.nelec) * t.nelec);
    } catch (const std::exception& ex) {
        delete p;
        return fail(QMC_ERR_HIP, ex.what());
    }
    *out = p;
    return QMC_OK;
}

int qmc_plan_destroy(QmcPlan* plan) {
    delete plan;
    return QMC_OK;
}

int qmc_plan_set_specialized(QmcPlan* plan, int32_t enable) {
    if (!plan) return fail(QMC_ERR_ARG, "null plan");
    plan->use_spec = enable != 0;
    return QMC_OK;
}

int qmc_plan_bucket(const QmcPlan* plan) { return plan ? nm_bucket_fine(plan->t) : -1; }

const char* qmc_plan_specialization(const QmcPlan* plan) {
    return (plan && plan->spec && plan->use_spec) ? plan->spec->name : nullptr;
}

}  // extern "C"

namespace {

int check_common(const QmcPlan* plan, const QmcParams* prm, int64_t nw) {
    if (!plan || !prm) return fail(QMC_ERR_ARG, "null plan or params");
    if (nw < 0) return fail(QMC_ERR_ARG, "negative walker count");
    if (!prm->bas_exp || !prm->atom_coords || !prm->mo_w || !prm->ci || !prm->grp_exp || !prm->grp_expf)
        return fail(QMC_ERR_ARG, "missing parameter pointer");
    if (plan->t.use_ee && !prm->w_ee) return fail(QMC_ERR_ARG, "e-e Jastrow weight missing");
    if (plan->t.use_en && !prm->w_en) return fail(QMC_ERR_ARG, "e-n Jastrow weight missing");
    return QMC_OK;
}

}  // namespace

extern "C" {

int qmc_wf_eval(const QmcPlan* plan, const QmcParams* params, int32_t precision, int32_t mode, const double* pos,
                int64_t nw, QmcWfOut* out, qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if (!pos || !out) return fail(QMC_ERR_ARG, "null pos/out");
    if (!params->entry_w) return fail(QMC_ERR_ARG, "missing parameter pointer entry_w");
    if (nw == 0) return QMC_OK;
    if (mode == QMC_EVAL_ELOC) precision = QMC_PREC_FP64;  // local energies are FP64 in every precision
    KArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.out = *out;
    hipStream_t s = (hipStream_t)stream;
    if (plan->t.exc_mode) {
        rc = QMC_BY_RAD(launch_wf_exc, a.t.radial, a, precision, mode, s);
        if (rc) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    if (plan->spec && plan->use_spec && plan->spec->wf(a, precision, mode, s) == 0) {
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    rc = precision == QMC_PREC_FP64 ? QMC_BY_RAD(launch_wf_f64, a.t.radial, a, mode, s)
                                    : QMC_BY_RAD(launch_wf_f32, a.t.radial, a, mode, s);
    if (rc) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_metropolis_sweep(const QmcPlan* plan, const QmcParams* params, int32_t precision, QmcWalkers* wk, int64_t nw,
                         const QmcSweepDesc* sw, qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if (!wk || !sw || !wk->pos || !wk->logabs || !wk->sign) return fail(QMC_ERR_ARG, "null walker state");
    if (sw->keep && sw->keep_every < 1) return fail(QMC_ERR_ARG, "keep_every must be >= 1");
    if (!params->entry_w) return fail(QMC_ERR_ARG, "missing parameter pointer entry_w");
    if (nw == 0 || sw->nstep <= 0) return QMC_OK;
    SArgs a;
    a.t = plan->t;
    a.p = *params;
    a.wk = *wk;
    a.sw = *sw;
    a.nw = nw;
    hipStream_t s = (hipStream_t)stream;
    if (plan->t.exc_mode) {
        rc = QMC_BY_RAD(launch_sweep_exc, a.t.radial, a, precision, s);
        if (rc) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    if (plan->spec && plan->use_spec && plan->spec->sweep(a, precision, s) == 0) {
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    rc = precision == QMC_PREC_FP64 ? QMC_BY_RAD(launch_sweep_f64, a.t.radial, a, s)
                                    : QMC_BY_RAD(launch_sweep_f32, a.t.radial, a, s);
    if (rc) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_hmc_sweep(const QmcPlan* plan, const QmcParams* params, int32_t precision, QmcWalkers* wk, int64_t nw,
                  const QmcHmcDesc* hm, qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if (!wk || !hm || !wk->pos || !wk->logabs || !wk->sign) return fail(QMC_ERR_ARG, "null walker state");
    if (hm->traj_length < 1) return fail(QMC_ERR_ARG, "traj_length must be >= 1");
    if (hm->keep && hm->keep_every < 1) return fail(QMC_ERR_ARG, "keep_every must be >= 1");
    if (!params->entry_w) return fail(QMC_ERR_ARG, "missing parameter pointer entry_w");
    if (nw == 0 || hm->nstep <= 0) return QMC_OK;
    HArgs a;
    a.t = plan->t;
    a.p = *params;
    a.wk = *wk;
    a.hm = *hm;
    a.nw = nw;
    a.drift = nullptr;
    a.write_pos = 1;
    hipStream_t s = (hipStream_t)stream;
    if (plan->spec && plan->use_spec && plan->spec->hmc(a, precision, s) == 0) {
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    rc = launch_hmc(a, precision, s);
    if (rc) return fail(QMC_ERR_UNSUPPORTED, a.t.exc_mode ? kNoExcBucket : kNoBucket);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_wf_drift(const QmcPlan* plan, const QmcParams* params, int32_t precision, const double* pos, int64_t nw,
                 double* drift, double* logabs, double* sign, qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if (!pos || !drift) return fail(QMC_ERR_ARG, "null pos/drift");
    if (!params->entry_w) return fail(QMC_ERR_ARG, "missing parameter pointer entry_w");
    if (nw == 0) return QMC_OK;
    HArgs a;
    a.t = plan->t;
    a.p = *params;
    a.wk = QmcWalkers{const_cast<double*>(pos), logabs, sign, nullptr, nullptr};  // positions are only read
    a.hm = QmcHmcDesc{};
    a.hm.traj_length = 1;
    a.nw = nw;
    a.drift = drift;
    a.write_pos = 0;
    hipStream_t s = (hipStream_t)stream;
    if (plan->spec && plan->use_spec && plan->spec->hmc(a, precision, s) == 0) {
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    rc = launch_hmc(a, precision, s);
    if (rc) return fail(QMC_ERR_UNSUPPORTED, a.t.exc_mode ? kNoExcBucket : kNoBucket);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_ao_eval(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw, double* ao,
                double* dao, double* d2ao, qmc_stream_t stream) {
    if (!plan || !params || !params->bas_exp || !params->atom_coords || !params->grp_exp || !params->grp_expf)
        return fail(QMC_ERR_ARG, "null argument");
    if (!pos || !ao) return fail(QMC_ERR_ARG, "null pos/ao");
    if (nw == 0) return QMC_OK;
    AArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.ao = ao;
    a.dao = dao;
    a.d2ao = d2ao;
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)nw * plan->t.nelec * plan->t.nao;
    // the kernel accumulates (AO entries split over groups), so clear first
    fill_zero(ao, n, s);
    if (dao) fill_zero(dao, 3 * n, s);
    if (d2ao) fill_zero(d2ao, n, s);
    QMC_BY_RAD(launch_ao, a.t.radial, a, dao || d2ao, s);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_logpsi_grad_rows(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                         double* ao_rows, double* gamma_rows, double* dci, double* djas, double* dexp_rows,
                         qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if ((rc = refuse_excitations(plan))) return rc;
    if (nw == 0) return QMC_OK;  // an empty batch: its tensors may have no storage
    if (!pos || !ao_rows || !gamma_rows || !dci || !djas) return fail(QMC_ERR_ARG, "null output");
    GArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.ao_rows = ao_rows;
    a.gamma_rows = gamma_rows;
    a.dci = dci;
    a.djas = djas;
    a.dexp_rows = dexp_rows;
    hipStream_t s = (hipStream_t)stream;
    fill_zero(ao_rows, (size_t)nw * plan->t.nelec * plan->t.nao, s);
    if (dexp_rows) fill_zero(dexp_rows, (size_t)nw * plan->t.nelec * plan->t.nprim_ref, s);
    if (QMC_BY_RAD(launch_grad, a.t.radial, a, s)) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

size_t qmc_adjoint_workspace_size(const QmcPlan* plan, int64_t nw, int32_t with_exp) {
    if (!plan || nw < 0) return 0;
    const RedSlots r = red_layout(plan->t, nw, with_exp != 0, nullptr, (size_t)-1, nullptr);
    return (size_t)(r.nslot > 0 ? r.nslot : 1) * r.stride * sizeof(double);
}

int qmc_logpsi_grad(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                    const double* weights, QmcGradOut* out, void* workspace, size_t workspace_bytes,
                    qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if (!out || !out->grad_w || !out->grad_ci || !out->grad_jas) return fail(QMC_ERR_ARG, "null gradient output");
    const RedSlots r = red_layout(plan->t, nw, out->grad_exp != nullptr, (double*)workspace, workspace_bytes,
                                  weights);
    if (nw > 0 && (!pos || !weights)) return fail(QMC_ERR_ARG, "null pos or weights");
    if (nw > 0 && (!workspace || r.nslot < 1))
        return fail(QMC_ERR_ARG, "workspace smaller than one partial-sum slot (qmc_adjoint_workspace_size)");
    hipStream_t s = (hipStream_t)stream;
    RedOut o{out->grad_w, out->grad_ci, out->grad_jas, out->grad_exp};
    if (nw == 0) {  // empty batch: zero gradients
        fill_zero(o.grad_w, (size_t)plan->t.nao * plan->t.nmo, s);
        fill_zero(o.grad_ci, plan->t.nci, s);
        fill_zero(o.grad_jas, 2, s);
        if (o.grad_exp) fill_zero(o.grad_exp, plan->t.nprim_ref, s);
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    GRedArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.r = r;
    const int miss = a.t.exc_mode ? QMC_BY_RAD(launch_grad_reduce_exc, a.t.radial, a, o, s)
                                  : QMC_BY_RAD(launch_grad_reduce, a.t.radial, a, o, s);
    if (miss) return fail(QMC_ERR_UNSUPPORTED, a.t.exc_mode ? kNoExcBucket : kNoBucket);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_eloc_grad(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                  const double* weights, double* eloc, QmcGradOut* out, void* workspace, size_t workspace_bytes,
                  qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if (!out || !out->grad_w || !out->grad_ci || !out->grad_jas) return fail(QMC_ERR_ARG, "null gradient output");
    if (out->grad_exp) return fail(QMC_ERR_UNSUPPORTED, "d E_L / d bas_exp is not provided");
    const RedSlots r = red_layout(plan->t, nw, false, (double*)workspace, workspace_bytes, weights);
    if (nw > 0 && (!pos || !weights)) return fail(QMC_ERR_ARG, "null pos or weights");
    if (nw > 0 && (!workspace || r.nslot < 1))
        return fail(QMC_ERR_ARG, "workspace smaller than one partial-sum slot (qmc_adjoint_workspace_size)");
    hipStream_t s = (hipStream_t)stream;
    RedOut o{out->grad_w, out->grad_ci, out->grad_jas, nullptr};
    if (nw == 0) {
        fill_zero(o.grad_w, (size_t)plan->t.nao * plan->t.nmo, s);
        fill_zero(o.grad_ci, plan->t.nci, s);
        fill_zero(o.grad_jas, 2, s);
        QMC_HIP(hipGetLastError());
        return QMC_OK;
    }
    ERedArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.eloc = eloc;
    a.r = r;
    if (a.t.exc_mode) {
        const int miss = QMC_BY_RAD(launch_eloc_reduce_exc, a.t.radial, a, o, s);
        if (miss == -2)
            return fail(QMC_ERR_UNSUPPORTED, "the local-energy adjoint of this excitation plan needs more LDS per "
                                             "wavefront than a block can hold (64 KiB of row blocks at most)");
        if (miss) return fail(QMC_ERR_UNSUPPORTED, kNoExcBucket);
    } else if (QMC_BY_RAD(launch_eloc_reduce, a.t.radial, a, o, s)) {
        return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
    }
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_philox4x32_10(const uint32_t* in, uint32_t* out, int64_t n, qmc_stream_t stream) {
    if (!in || !out || n < 0) return fail(QMC_ERR_ARG, "bad philox arguments");
    if (n == 0) return QMC_OK;
    hipLaunchKernelGGL(philox_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, in, out, n);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_eloc_grad_rows(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw, double* eloc,
                       double* ao_rows, double* b_rows, double* ga_rows, double* gb_rows, double* dci,
                       double* djas, double* drift, qmc_stream_t stream) {
    int rc = check_common(plan, params, nw);
    if (rc) return rc;
    if ((rc = refuse_excitations(plan))) return rc;
    if ((ao_rows == nullptr) != (b_rows == nullptr)) return fail(QMC_ERR_ARG, "ao_rows and b_rows go together");
    if (nw == 0) return QMC_OK;  // an empty batch: its tensors may have no storage
    if (!pos) return fail(QMC_ERR_ARG, "null pos");
    EArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.eloc = eloc;
    a.ao_rows = ao_rows;
    a.b_rows = b_rows;
    a.ga_rows = ga_rows;
    a.gb_rows = gb_rows;
    a.dci = dci;
    a.djas = djas;
    a.drift = drift;
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)nw * plan->t.nelec * plan->t.nao;
    fill_zero(ao_rows, n, s);
    fill_zero(b_rows, n, s);
    if (QMC_BY_RAD(launch_eloc_grad, a.t.radial, a, s)) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
    QMC_HIP(hipGetLastError());
    return QMC_OK;
}

}  // extern "C"

// ========================================================================= batch
struct QmcBatch {
    static constexpr int kSlots = 8;
    int max_jobs = 0;
    size_t slot_bytes = 0;
    char* host = nullptr;  // pinned staging, kSlots * slot_bytes
    char* dev = nullptr;   // device job tables, kSlots * slot_bytes
    hipEvent_t done[kSlots] = {};
    bool pending[kSlots] = {};
    int next = 0;
    ~QmcBatch() {
        for (int i = 0; i < kSlots; ++i) {
            if (pending[i]) (void)hipEventSynchronize(done[i]);
            if (done[i]) (void)hipEventDestroy(done[i]);
        }
        if (host) (void)hipHostFree(host);
        if (dev) (void)hipFree(dev);
    }
};

namespace {

// Relative per-wave cost of a system (lane work: AO groups x MO columns,
// pair terms, elimination), used to dispatch the costliest waves first.
double wave_cost(const Tables& t) {
    return (double)t.nentry * (t.nmo + 8) + 12.0 * t.ngroup + 6.0 * t.nelec * (t.nelec + t.natom) +
           (double)t.nsig_max * t.nsig_max * t.nconf_max;
}

// Takes the next ring slot, waiting (rarely) for the launches that last read it.
int take_slot(QmcBatch* b, int& slot) {
    slot = b->next;
    b->next = (b->next + 1) % QmcBatch::kSlots;
    if (b->pending[slot]) {
        QMC_HIP(hipEventSynchronize(b->done[slot]));
        b->pending[slot] = false;
    }
    return QMC_OK;
}

struct Group {
    int rad, bucket;
    bool exc;  // excitation plans: their own kernels, one wavefront per block
    const SpecEntry* spec;
    std::vector<int> idx;
    size_t lds;  // exc: the largest exc_lds_bytes of the group's jobs
};

// Groups the jobs by (radial type, register bucket, excitation path, generated
// layout), costliest systems first inside a group.  Returns an error for jobs
// without a bucket.
template <typename JobT>
int group_jobs(const JobT* jobs, int njobs, std::vector<Group>& groups) {
    for (int j = 0; j < njobs; ++j) {
        const Tables& t = jobs[j].plan->t;
        const int bk = nm_bucket_fine(t);
        if (bk < 0) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
        const bool exc = t.exc_mode != 0;
        const SpecEntry* sp = (!exc && jobs[j].plan->use_spec) ? jobs[j].plan->spec : nullptr;
        Group* g = nullptr;
        for (auto& x : groups)
            if (x.rad == t.radial && x.bucket == bk && x.exc == exc && x.spec == sp) g = &x;
        if (!g) {
            groups.push_back(Group{t.radial, bk, exc, sp, {}, 0});
            g = &groups.back();
        }
        g->idx.push_back(j);
        if (exc) g->lds = std::max(g->lds, exc_lds_bytes(t, bk));
    }
    for (auto& g : groups)
        std::stable_sort(g.idx.begin(), g.idx.end(),
                         [&](int a, int b) { return wave_cost(jobs[a].plan->t) > wave_cost(jobs[b].plan->t); });
    return QMC_OK;
}

}  // namespace

extern "C" {

int qmc_batch_create(int32_t max_jobs, QmcBatch** out) {
    if (!out || max_jobs <= 0) return fail(QMC_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    QmcBatch* b = new QmcBatch();
    b->max_jobs = max_jobs;
    const size_t per = std::max(sizeof(SJob), sizeof(WJob));
    b->slot_bytes = ((per * (size_t)max_jobs + 255) / 256) * 256;
    const size_t total = b->slot_bytes * QmcBatch::kSlots;
    if (hipHostMalloc((void**)&b->host, total, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&b->dev, total) != hipSuccess) {
        delete b;
        return fail(QMC_ERR_HIP, "batch table allocation failed");
    }
    for (int i = 0; i < QmcBatch::kSlots; ++i) {
        if (hipEventCreateWithFlags(&b->done[i], hipEventDisableTiming) != hipSuccess) {
            delete b;
            return fail(QMC_ERR_HIP, "hipEventCreate failed");
        }
    }
    *out = b;
    return QMC_OK;
}

int qmc_batch_destroy(QmcBatch* b) {
    delete b;
    return QMC_OK;
}

int qmc_metropolis_sweep_batch(QmcBatch* b, int32_t precision, const QmcSweepJob* jobs, int32_t njobs,
                               qmc_stream_t stream) {
    if (!b || (njobs > 0 && !jobs)) return fail(QMC_ERR_ARG, "null batch or jobs");
    if (njobs < 0 || njobs > b->max_jobs) return fail(QMC_ERR_ARG, "job count outside [0, max_jobs]");
    std::vector<QmcSweepJob> live;
    for (int j = 0; j < njobs; ++j) {
        const QmcSweepJob& x = jobs[j];
        int rc = check_common(x.plan, &x.params, x.nw);
        if (rc) return rc;
        if (!x.walkers.pos || !x.walkers.logabs || !x.walkers.sign) return fail(QMC_ERR_ARG, "null walker state");
        if (!x.params.entry_w) return fail(QMC_ERR_ARG, "missing parameter pointer entry_w");
        if (x.sweep.keep && x.sweep.keep_every < 1) return fail(QMC_ERR_ARG, "keep_every must be >= 1");
        if (x.nw > 0 && x.sweep.nstep > 0) live.push_back(x);
    }
    if (live.empty()) return QMC_OK;
    std::vector<Group> groups;
    int rc = group_jobs(live.data(), (int)live.size(), groups);
    if (rc) return rc;
    int slot;
    rc = take_slot(b, slot);
    if (rc) return rc;
    SJob* host = (SJob*)(b->host + slot * b->slot_bytes);
    SJob* dev = (SJob*)(b->dev + slot * b->slot_bytes);
    struct Launch {
        int rad, bucket, first, n, nwaves;
        bool exc;
        size_t lds;
        const SpecEntry* spec;
    };
    std::vector<Launch> launches;
    int k = 0;
    for (const auto& g : groups) {
        int64_t waves = 0;
        const int first = k;
        for (int j : g.idx) {
            const QmcSweepJob& x = live[j];
            SJob& e = host[k++];
            e.a.t = x.plan->t;
            e.a.p = x.params;
            e.a.wk = x.walkers;
            e.a.sw = x.sweep;
            e.a.nw = x.nw;
            e.wave0 = waves;
            const int G = 64 / x.plan->t.nelec;
            waves += (x.nw + G - 1) / G;
        }
        if (waves > (int64_t)1 << 30) return fail(QMC_ERR_ARG, "batch too large");
        launches.push_back(Launch{g.rad, g.bucket, first, k - first, (int)waves, g.exc, g.lds, g.spec});
    }
    hipStream_t s = (hipStream_t)stream;
    QMC_HIP(hipMemcpyAsync(dev, host, (size_t)k * sizeof(SJob), hipMemcpyHostToDevice, s));
    for (const auto& L : launches) {
        if (L.exc) {
            if (QMC_BY_RAD(launch_sweep_batch_exc, L.rad, dev + L.first, L.n, L.nwaves, L.bucket, L.lds, precision, s))
                return fail(QMC_ERR_UNSUPPORTED, kNoExcBucket);
            continue;
        }
        if (L.spec && L.spec->sweep_batch(dev + L.first, L.n, L.nwaves, precision, s) == 0) continue;
        const int r = precision == QMC_PREC_FP64
                          ? QMC_BY_RAD(launch_sweep_batch_f64, L.rad, dev + L.first, L.n, L.nwaves, L.bucket, s)
                          : QMC_BY_RAD(launch_sweep_batch_f32, L.rad, dev + L.first, L.n, L.nwaves, L.bucket, s);
        if (r) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
    }
    QMC_HIP(hipGetLastError());
    QMC_HIP(hipEventRecord(b->done[slot], s));
    b->pending[slot] = true;
    return QMC_OK;
}

int qmc_wf_eval_batch(QmcBatch* b, int32_t precision, int32_t mode, const QmcEvalJob* jobs, int32_t njobs,
                      qmc_stream_t stream) {
    if (!b || (njobs > 0 && !jobs)) return fail(QMC_ERR_ARG, "null batch or jobs");
    if (njobs < 0 || njobs > b->max_jobs) return fail(QMC_ERR_ARG, "job count outside [0, max_jobs]");
    if (mode == QMC_EVAL_ELOC) precision = QMC_PREC_FP64;  // local energies are FP64 in every precision
    std::vector<QmcEvalJob> live;
    for (int j = 0; j < njobs; ++j) {
        const QmcEvalJob& x = jobs[j];
        int rc = check_common(x.plan, &x.params, x.nw);
        if (rc) return rc;
        if (!x.pos) return fail(QMC_ERR_ARG, "null pos");
        if (!x.params.entry_w) return fail(QMC_ERR_ARG, "missing parameter pointer entry_w");
        if (x.nw > 0) live.push_back(x);
    }
    if (live.empty()) return QMC_OK;
    std::vector<Group> groups;
    int rc = group_jobs(live.data(), (int)live.size(), groups);
    if (rc) return rc;
    int slot;
    rc = take_slot(b, slot);
    if (rc) return rc;
    WJob* host = (WJob*)(b->host + slot * b->slot_bytes);
    WJob* dev = (WJob*)(b->dev + slot * b->slot_bytes);
    struct Launch {
        int rad, bucket, first, n, nwaves;
        bool exc;
        size_t lds;
        const SpecEntry* spec;
    };
    std::vector<Launch> launches;
    int k = 0;
    for (const auto& g : groups) {
        int64_t waves = 0;
        const int first = k;
        for (int j : g.idx) {
            const QmcEvalJob& x = live[j];
            WJob& e = host[k++];
            e.a.t = x.plan->t;
            e.a.p = x.params;
            e.a.pos = x.pos;
            e.a.nw = x.nw;
            e.a.out = x.out;
            e.wave0 = waves;
            const int G = 64 / x.plan->t.nelec;
            waves += (x.nw + G - 1) / G;
        }
        if (waves > (int64_t)1 << 30) return fail(QMC_ERR_ARG, "batch too large");
        launches.push_back(Launch{g.rad, g.bucket, first, k - first, (int)waves, g.exc, g.lds, g.spec});
    }
    hipStream_t s = (hipStream_t)stream;
    QMC_HIP(hipMemcpyAsync(dev, host, (size_t)k * sizeof(WJob), hipMemcpyHostToDevice, s));
    for (const auto& L : launches) {
        if (L.exc) {
            if (QMC_BY_RAD(launch_wf_batch_exc, L.rad, dev + L.first, L.n, L.nwaves, L.bucket, L.lds, precision, mode,
                           s))
                return fail(QMC_ERR_UNSUPPORTED, kNoExcBucket);
            continue;
        }
        if (L.spec && L.spec->wf_batch(dev + L.first, L.n, L.nwaves, precision, mode, s) == 0) continue;
        const int r =
            precision == QMC_PREC_FP64
                ? QMC_BY_RAD(launch_wf_batch_f64, L.rad, dev + L.first, L.n, L.nwaves, L.bucket, mode, s)
                : QMC_BY_RAD(launch_wf_batch_f32, L.rad, dev + L.first, L.n, L.nwaves, L.bucket, mode, s);
        if (r) return fail(QMC_ERR_UNSUPPORTED, kNoBucket);
    }
    QMC_HIP(hipGetLastError());
    QMC_HIP(hipEventRecord(b->done[slot], s));
    b->pending[slot] = true;
    return QMC_OK;
}

}  // extern "C"
