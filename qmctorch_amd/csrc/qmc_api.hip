// libqmcmi355x: kernels behind the component-level entry points of the
// reference's module API (parameter views, Jastrow factors, Slater pooling,
// atomic orbitals with full second derivatives, walker initialisation).  The
// fused wavefunction / sweep kernels of qmc_kernels.h are the hot path; these
// serve direct calls of the individual modules with the reference's shapes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "qmc_plan.h"

using namespace qmc;

namespace {

int fail(int code, const char* msg) { return set_error(code, msg); }

// ----------------------------------------------------------------- parameters
// W = mo_scf * mo_modifier (first nmo columns), its rows in AO-entry order
// padded to the plan's register bucket, and the exponents gathered in group
// order, in FP64 and FP32 (molecular_orbitals.py:65-95, atomic_orbitals.py:53-55).
struct GatherArgs {
    Tables t;
    const double* mo_scf;
    const double* mo_mod;
    int ld, bucket, nexp;
    const double* bas_exp;
    QmcParamViews v;
};

__global__ __launch_bounds__(kBlock) void params_gather_kernel(GatherArgs a) {
    const Tables& t = a.t;
    const int nmo = t.nmo;
    const int n_w = t.nao * nmo, n_e = t.nentry * a.bucket;
    const int n = max(max(n_w, n_e), a.nexp);
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        if (idx < n_w) {
            const int mu = idx / nmo, m = idx - mu * nmo;
            const size_t k = (size_t)mu * a.ld + m;
            const double w = a.mo_scf[k] * a.mo_mod[k];
            if (a.v.mo_w64) a.v.mo_w64[idx] = w;
            if (a.v.mo_w32) a.v.mo_w32[idx] = (float)w;
        }
        if (idx < n_e) {
            const int i = idx / a.bucket, m = idx - i * a.bucket;
            double w = 0.0;
            if (m < nmo) {
                const size_t k = (size_t)t.gao_mu[i] * a.ld + m;
                w = a.mo_scf[k] * a.mo_mod[k];
            }
            if (a.v.entry_w64) a.v.entry_w64[idx] = w;
            if (a.v.entry_w32) a.v.entry_w32[idx] = (float)w;
        }
        if (idx < a.nexp) {
            const double e = a.bas_exp[t.grp_prim_src[idx]];
            if (a.v.grp_exp) a.v.grp_exp[idx] = e;
            if (a.v.grp_expf) a.v.grp_expf[idx] = (float)e;
        }
    }
}

// ------------------------------------------------------------------- Jastrow
// J, dJ/dx_{e,d} and lap_e J of the selected Pade factors, and the potentials
// (jastrow_factor_electron_electron.py:124-260, jastrow_factor_electron_nuclei.py,
// combine_jastrow.py:33-195; wf_base.py:49-95).  The product of factors is
// evaluated in log space: lnJ = sum of the terms, dJ = J grad lnJ,
// lap J = J (lap lnJ + |grad lnJ|^2), which equals the reference's product rule.
struct JasArgs {
    Tables t;  // use_ee / use_en select the factors
    QmcParams p;
    const double* pos;
    int64_t nw;
    QmcJastrowOut out;
};

__global__ __launch_bounds__(kBlock) void jastrow_eval_kernel(JasArgs a) {
    const Tables& t = a.t;
    const WalkerLane L = lane_layout(t, a.nw);
    if ((int64_t)L.wave * L.G >= a.nw) return;
    const double* p = a.pos + (size_t)L.wclamp * 3 * t.nelec + 3 * L.e;
    double kee, ken, gx, gy, gz, lap, vee, ven;
    jastrow_and_potential<true>(t, a.p, L, p[0], p[1], p[2], kee, ken, gx, gy, gz, lap, vee, ven);
    const double lnj = seg_sum(0.5 * kee + ken, L.wbase, t.nelec, t.nelec);
    const double svee = seg_sum(vee, L.wbase, t.nelec, t.nelec);
    const double sven = seg_sum(ven, L.wbase, t.nelec, t.nelec);
    const double J = exp(lnj);
    if (!L.valid) return;
    const size_t w = (size_t)L.w;
    if (a.out.grad) {
        double* g = a.out.grad + w * 3 * t.nelec + L.e;
        g[0] = J * gx;
        g[t.nelec] = J * gy;
        g[2 * t.nelec] = J * gz;
    }
    if (a.out.lap) a.out.lap[w * t.nelec + L.e] = J * (lap + gx * gx + gy * gy + gz * gz);
    if (L.e == 0) {
        if (a.out.value) a.out.value[w] = J;
        if (a.out.v_ee) a.out.v_ee[w] = svee;
        if (a.out.v_en) a.out.v_en[w] = sven;
    }
}

// ------------------------------------------------------------ Slater pooling
// Determinants and Jacobi traces of the configurations of every CI term from
// MO matrices given by the caller (slater_pooling.py:64-111 det_explicit,
// 262-409 operator / operator_explicit / operator_ground_state): for CI term n
// and spin s, A_s = mo[rows of s][conf_s(n)], B_s = bop[rows of s][conf_s(n)];
//   det[b][n]  = det A_up det A_dn
//   up[b][n]   = tr(A_up^-1 B_up)   (tr((A^-1 B)^2) with op_squared)
//   dn[b][n]   = the same for spin down.
// Batch b = op * nw + w reads mo of walker w and bop of (op, w) (the reference's
// multi-operator layout [nop][nw][nelec][nmo]).
struct PoolArgs {
    Tables t;
    const double* mo;
    const double* bop;
    int64_t nw, nb;
    int ld, op_squared;
    double* det;
    double* up;
    double* dn;
};

template <int NM>
__global__ __launch_bounds__(kBlock) void pool_kernel(PoolArgs a) {
    const Tables& t = a.t;
    const WalkerLane L = lane_layout(t, a.nb);
    if ((int64_t)L.wave * L.G >= a.nb) return;
    const int64_t b = L.wclamp;
    const int64_t w = b % a.nw;
    const double* mrow = a.mo + ((size_t)w * t.nelec + L.e) * a.ld;
    const double* brow = a.bop ? a.bop + ((size_t)b * t.nelec + L.e) * a.ld : nullptr;
    double* rows = wave_rowbuf<NM>();
    const int n = L.n_sig;
    for (int ci = 0; ci < t.nci; ++ci) {
        const int c = L.spin == 0 ? t.ci_up[ci] : t.ci_dn[ci];
        const QMC_CONST int* conf = L.spin == 0 ? t.conf_up + c * t.nup : t.conf_dn + c * t.ndown;
        double aa[NM], bb[NM];
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            const int col = j < n ? conf[j] : 0;
            aa[j] = j < n ? mrow[col] : 0.0;
            bb[j] = (j < n && brow) ? brow[col] : 0.0;
        }
        double lg, sg, tp;
        int col;
        lane_gauss_jordan<NM, true>(aa, bb, n, t.nsig_max, L.myrow, L.gbase, L.gmask, lg, sg, tp, col);
        double tr = seg_sum(tp, L.gbase, n, t.nsig_max);
        if (a.op_squared) {
            // tr(X^2) = sum_kj X_kj X_jk: the lane that pivoted column k holds
            // row k of X = A^-1 B; rows go through the wave's LDS row buffer
            wave_lds_sync();
            if (L.lane_real && col >= 0) {
#pragma unroll
                for (int j = 0; j < NM; ++j) rows[(L.gbase + col) * NM + j] = bb[j];
            }
            wave_lds_sync();
            double part = 0.0;
            if (col >= 0)
                for (int j = 0; j < n; ++j) part += bb[j] * rows[(L.gbase + j) * NM + col];
            tr = seg_sum(part, L.gbase, n, t.nsig_max);
        }
        // every lane fetches its walker's up (lane wbase) and down values
        const int up_lane = L.wbase, dn_lane = L.wbase + (t.ndown > 0 ? t.nup : 0);
        const double lu = __shfl(lg, up_lane), su = __shfl(sg, up_lane), tu = __shfl(tr, up_lane);
        const double ldn = __shfl(lg, dn_lane), sdn = __shfl(sg, dn_lane), tdn = __shfl(tr, dn_lane);
        if (L.valid && L.e == 0) {
            const size_t o = (size_t)b * t.nci + ci;
            const double l = lu + (t.ndown > 0 ? ldn : 0.0);
            const double s = su * (t.ndown > 0 ? sdn : 1.0);
            if (a.det && b < a.nw) a.det[o] = s * exp(l);  // det is [nw][nci]
            if (a.up) a.up[o] = tu;
            if (a.dn) a.dn[o] = t.ndown > 0 ? tdn : 0.0;
        }
    }
}

// ---------------------------------------------------- AO second derivatives
// Values, gradients, Laplacians, Hessian diagonals and mixed second
// derivatives (xy, xz, yz) of the AOs of one electron in the reference layouts
// (atomic_orbitals.py:131-218, 330-575).  Radial part of an AO entry, with P1,
// P2 its reduced radial derivatives:
//   d_a R = x_a P1,   d_a d_b R = x_a x_b P2 + delta_ab P1,
//   GTO  R = sum K e^{-a r^2}:    P1 = -2 sum K a e,  P2 = 4 sum K a^2 e
//   STO  R = sum K r^n e^{-a r}:  with q = n/r^2 - a/r, P1 = sum K R_p q,
//                                 P2 = sum K R_p (q^2 - 2n/r^4 + a/r^3),
// and AO = Y R: d_a d_b AO = Y_ab R + (Y_a x_b + Y_b x_a) P1 + Y (x_a x_b P2 + delta_ab P1).
struct AoFullArgs {
    Tables t;
    QmcParams p;
    const double* pos;
    int64_t nw;
    int nelec;  // electrons per walker in pos (1 for the one-electron update)
    QmcAoOut out;
};

__device__ __forceinline__ double pw(double x, int k) { return k <= 0 ? 1.0 : (k == 1 ? x : (k == 2 ? x * x : x * x * x)); }
__device__ __forceinline__ double d1p(double x, int k) { return k <= 0 ? 0.0 : k * pw(x, k - 1); }
__device__ __forceinline__ double d2p(double x, int k) { return k <= 1 ? 0.0 : k * (k - 1) * pw(x, k - 2); }

template <int RAD>
__global__ __launch_bounds__(kBlock) void ao_full_kernel(AoFullArgs a) {
    const Tables& t = a.t;
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // (walker, electron)
    if (row >= a.nw * a.nelec) return;
    const double* pe = a.pos + row * 3;
    const size_t base = (size_t)row * t.nao;
    const QmcAoOut& o = a.out;
    for (int g = 0; g < t.ngroup; ++g) {
        const int at = t.grp_atom[g];
        const double x[3] = {pe[0] - ldc(a.p.atom_coords, 3 * at), pe[1] - ldc(a.p.atom_coords, 3 * at + 1),
                             pe[2] - ldc(a.p.atom_coords, 3 * at + 2)};
        const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        const double r = sqrt(r2), ir = 1.0 / r;
        const int np = t.grp_nprim[g], p0 = t.grp_prim0[g];
        for (int i = t.grp_ao0[g]; i < t.grp_ao0[g] + t.grp_nao[g]; ++i) {
            const int mu = t.gao_mu[i], kk = t.gao_k[i];
            const int k[3] = {kk & 15, (kk >> 4) & 15, (kk >> 8) & 15};
            double R = 0.0, P1 = 0.0, P2 = 0.0;
            for (int q = 0; q < np; ++q) {
                const double al = ldc(a.p.grp_exp, p0 + q);
                const double K = t.gao_K[(size_t)i * QMC_MAXP + q];
                if (RAD == QMC_RADIAL_GTO_PURE) {
                    const double e = K * exp(-al * r2);
                    R += e;
                    P1 += -2.0 * al * e;
                    P2 += 4.0 * al * al * e;
                } else {
                    const double nn = t.gao_n[(size_t)i * QMC_MAXP + q];
                    const double e = K * pw(r, (int)nn) * exp(-al * r);
                    const double qq = nn * ir * ir - al * ir;
                    R += e;
                    P1 += e * qq;
                    P2 += e * (qq * qq - 2.0 * nn * ir * ir * ir * ir + al * ir * ir * ir);
                }
            }
            const double px = pw(x[0], k[0]), py = pw(x[1], k[1]), pz = pw(x[2], k[2]);
            const double Y = px * py * pz;
            const double Yd[3] = {d1p(x[0], k[0]) * py * pz, px * d1p(x[1], k[1]) * pz, px * py * d1p(x[2], k[2])};
            const double Ydd[3] = {d2p(x[0], k[0]) * py * pz, px * d2p(x[1], k[1]) * pz, px * py * d2p(x[2], k[2])};
            const double Ym[3] = {d1p(x[0], k[0]) * d1p(x[1], k[1]) * pz, d1p(x[0], k[0]) * py * d1p(x[2], k[2]),
                                  px * d1p(x[1], k[1]) * d1p(x[2], k[2])};
            double grad[3], hd[3];
            for (int d = 0; d < 3; ++d) {
                grad[d] = Yd[d] * R + Y * x[d] * P1;
                hd[d] = Ydd[d] * R + 2.0 * Yd[d] * x[d] * P1 + Y * (x[d] * x[d] * P2 + P1);
            }
            const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
            const size_t e = base + mu;
            if (o.ao) o.ao[e] += Y * R;
            if (o.grad)
                for (int d = 0; d < 3; ++d) o.grad[e * 3 + d] += grad[d];
            if (o.lap) o.lap[e] += hd[0] + hd[1] + hd[2];
            if (o.hess_diag)
                for (int d = 0; d < 3; ++d) o.hess_diag[e * 3 + d] += hd[d];
            if (o.hess_mixed)
                for (int m = 0; m < 3; ++m) {
                    const int u = pa[m], v = pb[m];
                    o.hess_mixed[e * 3 + m] +=
                        Ym[m] * R + (Yd[u] * x[v] + Yd[v] * x[u]) * P1 + Y * x[u] * x[v] * P2;
                }
        }
    }
}

__global__ __launch_bounds__(kBlock) void zero_kernel(double* p, size_t n) {
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

void zero(double* p, size_t n, hipStream_t s) {
    if (!p || n == 0) return;
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, n);
}

// ----------------------------------------------------------- walker init
// One thread per walker; Philox4x32-10 keyed by the seed, counter (walker id,
// electron, draw, stream tag).  Modes follow the keys of Molecule.domain()
// (sampler/walkers.py:8-150): uniform box, multivariate normal per electron,
// near the centre, or around the atoms with shell-dependent widths.
struct InitArgs {
    QmcWalkerInit d;
    double* pos;
    int64_t nw;
    int nelec;
};

__device__ __forceinline__ u4 init_rng(const QmcWalkerInit& d, int64_t w, uint32_t e, uint32_t draw, uint32_t tag) {
    const uint64_t wid = (uint64_t)w + d.walker0;
    return philox4x32_10(u4{(uint32_t)wid, (uint32_t)(wid >> 32), e | (draw << 16), tag}, (uint32_t)d.seed,
                         (uint32_t)(d.seed >> 32));
}

__global__ __launch_bounds__(kBlock) void walker_init_kernel(InitArgs a) {
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (w >= a.nw) return;
    const QmcWalkerInit& d = a.d;
    double* out = a.pos + (size_t)w * 3 * a.nelec;
    if (d.mode == QMC_INIT_ATOMIC) {
        // electron -> atom assignment: the list (atom a repeated atom_nelec[a]
        // times) in a uniformly random order (Fisher-Yates), then electron i at
        // atom[i] + s N(0, I) with s from its rank k on that atom
        int at[64];
        int cnt[QMC_INIT_MAXATOM];
        int m = 0;
        for (int A = 0; A < d.natom; ++A) {
            cnt[A] = 0;
            for (int j = 0; j < d.atom_nelec[A] && m < a.nelec; ++j) at[m++] = A;
        }
        for (int i = a.nelec - 1; i > 0; --i) {
            const u4 r = init_rng(d, w, (uint32_t)i, 0, 0x51A7u);
            const int j = (int)(((uint64_t)r.x * (uint32_t)(i + 1)) >> 32);
            const int tmp = at[i];
            at[i] = at[j];
            at[j] = tmp;
        }
        for (int e = 0; e < a.nelec; ++e) {
            const int A = at[e];
            const int k = cnt[A]++;
            double s = d.atom_width[A * 3 + (k == 0 ? 0 : (k < 5 ? 1 : 2))];
            double z0, z1, z2;
            gauss3_f64(init_rng(d, w, (uint32_t)e, 1, 0xA70Du), z0, z1, z2);
            out[3 * e + 0] = d.atom_xyz[3 * A + 0] + s * z0;
            out[3 * e + 1] = d.atom_xyz[3 * A + 1] + s * z1;
            out[3 * e + 2] = d.atom_xyz[3 * A + 2] + s * z2;
        }
        return;
    }
    for (int e = 0; e < a.nelec; ++e) {
        const u4 r = init_rng(d, w, (uint32_t)e, 2, 0xD0Au);
        double v[3];
        if (d.mode == QMC_INIT_NORMAL) {
            double z[3];
            gauss3_f64(r, z[0], z[1], z[2]);
            for (int i = 0; i < 3; ++i)
                v[i] = d.mean[i] + d.chol[3 * i + 0] * z[0] + d.chol[3 * i + 1] * z[1] + d.chol[3 * i + 2] * z[2];
        } else {
            const double u[3] = {u01_closed0(r.x), u01_closed0(r.y), u01_closed0(r.z)};
            for (int i = 0; i < 3; ++i) v[i] = d.lo + (d.hi - d.lo) * u[i];
        }
        out[3 * e + 0] = v[0];
        out[3 * e + 1] = v[1];
        out[3 * e + 2] = v[2];
    }
}

int pool_bucket(const Tables& t) {
    const int need = std::max(std::max(t.nup, t.ndown), 1);
    if (need <= 2) return 2;
    if (need <= 4) return 4;
    if (need <= 8) return 8;
    if (need <= 16) return 16;
    if (need <= 24) return 24;
    return -1;
}

}  // namespace

extern "C" {

int qmc_params_gather(const QmcPlan* plan, const double* mo_scf, const double* mo_modifier, int32_t ld,
                      const double* bas_exp, QmcParamViews* views, qmc_stream_t stream) {
    if (!plan || !views) return fail(QMC_ERR_ARG, "null plan or views");
    if (!mo_scf || !mo_modifier || !bas_exp) return fail(QMC_ERR_ARG, "null parameter");
    if (ld < plan->t.nmo) return fail(QMC_ERR_ARG, "ld smaller than the plan's MO count");
    const int bucket = nm_bucket_fine(plan->t);
    if (bucket < 0) return fail(QMC_ERR_UNSUPPORTED, "more than 24 occupied orbitals per spin");
    GatherArgs a;
    a.t = plan->t;
    a.mo_scf = mo_scf;
    a.mo_mod = mo_modifier;
    a.ld = ld;
    a.bucket = bucket;
    a.nexp = plan->nexp_grp;
    a.bas_exp = bas_exp;
    a.v = *views;
    const int n = std::max(std::max(plan->t.nao * plan->t.nmo, plan->t.nentry * bucket), plan->nexp_grp);
    const int blocks = std::max(1, std::min((n + kBlock - 1) / kBlock, 1024));
    hipLaunchKernelGGL(params_gather_kernel, dim3(blocks), dim3(kBlock), 0, (hipStream_t)stream, a);
    QMC_CHECK_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_plan_sizes(const QmcPlan* plan, QmcPlanSizes* out) {
    if (!plan || !out) return fail(QMC_ERR_ARG, "null plan or output");
    out->nexp_grp = plan->nexp_grp;
    out->nentry = plan->t.nentry;
    out->bucket = nm_bucket_fine(plan->t);
    out->nao = plan->t.nao;
    out->nmo = plan->t.nmo;
    return QMC_OK;
}

int qmc_jastrow_eval(const QmcPlan* plan, const QmcParams* params, int32_t terms, const double* pos, int64_t nw,
                     QmcJastrowOut* out, qmc_stream_t stream) {
    if (!plan || !params || !out) return fail(QMC_ERR_ARG, "null plan, params or output");
    if (nw < 0) return fail(QMC_ERR_ARG, "negative walker count");
    if (terms & ~3) return fail(QMC_ERR_ARG, "terms: bit 0 e-e, bit 1 e-n");
    const bool ee = (terms & 1) != 0, en = (terms & 2) != 0;
    if ((ee && !params->w_ee) || (en && !params->w_en)) return fail(QMC_ERR_ARG, "Jastrow weight missing");
    if ((out->v_en || en) && !params->atom_coords) return fail(QMC_ERR_ARG, "missing atom coordinates");
    if (en && !plan->t.jas_atoms) return fail(QMC_ERR_ARG, "plan without e-n Jastrow centres");
    if (nw == 0) return QMC_OK;
    if (!pos) return fail(QMC_ERR_ARG, "null pos");
    JasArgs a;
    a.t = plan->t;
    a.t.use_ee = ee ? 1 : 0;
    a.t.use_en = en ? 1 : 0;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.out = *out;
    hipLaunchKernelGGL(jastrow_eval_kernel, grid_for(a.t, nw), dim3(kBlock), 0, (hipStream_t)stream, a);
    QMC_CHECK_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_slater_pool(const QmcPlan* plan, const double* mo, int32_t ld, const double* bop, int32_t nop,
                    int32_t op_squared, int64_t nw, double* det, double* op_up, double* op_dn,
                    qmc_stream_t stream) {
    if (!plan) return fail(QMC_ERR_ARG, "null plan");
    if (nw < 0 || nop < 1) return fail(QMC_ERR_ARG, "negative walker count or nop < 1");
    if (!bop && (nop != 1 || op_up || op_dn)) return fail(QMC_ERR_ARG, "operator outputs need bop");
    if (ld < plan->t.nmo) return fail(QMC_ERR_ARG, "ld smaller than the highest configured MO + 1");
    if (nw == 0) return QMC_OK;
    if (!mo) return fail(QMC_ERR_ARG, "null mo");
    PoolArgs a;
    a.t = plan->t;
    a.mo = mo;
    a.bop = bop;
    a.nw = nw;
    a.nb = nw * nop;
    a.ld = ld;
    a.op_squared = op_squared ? 1 : 0;
    a.det = det;
    a.up = op_up;
    a.dn = op_dn;
    const dim3 grid = grid_for(a.t, a.nb);
    hipStream_t s = (hipStream_t)stream;
    switch (pool_bucket(a.t)) {
        case 2: hipLaunchKernelGGL(pool_kernel<2>, grid, dim3(kBlock), 0, s, a); break;
        case 4: hipLaunchKernelGGL(pool_kernel<4>, grid, dim3(kBlock), 0, s, a); break;
        case 8: hipLaunchKernelGGL(pool_kernel<8>, grid, dim3(kBlock), 0, s, a); break;
        case 16: hipLaunchKernelGGL(pool_kernel<16>, grid, dim3(kBlock), 0, s, a); break;
        case 24: hipLaunchKernelGGL(pool_kernel<24>, grid, dim3(kBlock), 0, s, a); break;
        default: return fail(QMC_ERR_UNSUPPORTED, "more than 24 electrons per spin");
    }
    QMC_CHECK_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_ao_eval_full(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw, int32_t nelec,
                     QmcAoOut* out, qmc_stream_t stream) {
    if (!plan || !params || !out) return fail(QMC_ERR_ARG, "null plan, params or output");
    if (!params->atom_coords || !params->grp_exp) return fail(QMC_ERR_ARG, "missing parameter pointer");
    if (nw < 0 || nelec < 1) return fail(QMC_ERR_ARG, "negative walker count or nelec < 1");
    if (nw == 0) return QMC_OK;
    if (!pos) return fail(QMC_ERR_ARG, "null pos");
    hipStream_t s = (hipStream_t)stream;
    const size_t nrow = (size_t)nw * nelec * plan->t.nao;
    zero(out->ao, nrow, s);
    zero(out->grad, nrow * 3, s);
    zero(out->lap, nrow, s);
    zero(out->hess_diag, nrow * 3, s);
    zero(out->hess_mixed, nrow * 3, s);
    AoFullArgs a;
    a.t = plan->t;
    a.p = *params;
    a.pos = pos;
    a.nw = nw;
    a.nelec = nelec;
    a.out = *out;
    const int64_t rows = nw * nelec;
    const dim3 grid((unsigned)((rows + kBlock - 1) / kBlock));
    if (a.t.radial == QMC_RADIAL_STO)
        hipLaunchKernelGGL(ao_full_kernel<QMC_RADIAL_STO>, grid, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL(ao_full_kernel<QMC_RADIAL_GTO_PURE>, grid, dim3(kBlock), 0, s, a);
    QMC_CHECK_HIP(hipGetLastError());
    return QMC_OK;
}

int qmc_walkers_init(const QmcWalkerInit* init, int32_t nelec, int64_t nw, double* pos, qmc_stream_t stream) {
    if (!init) return fail(QMC_ERR_ARG, "null init");
    if (nw < 0 || nelec < 1 || nelec > 64) return fail(QMC_ERR_ARG, "nelec must be in [1, 64], nw >= 0");
    if (init->mode < QMC_INIT_CENTER || init->mode > QMC_INIT_ATOMIC) return fail(QMC_ERR_ARG, "unknown mode");
    if (init->mode == QMC_INIT_ATOMIC) {
        if (init->natom < 1 || init->natom > QMC_INIT_MAXATOM) return fail(QMC_ERR_ARG, "natom out of range");
        int tot = 0;
        for (int A = 0; A < init->natom; ++A) {
            if (init->atom_nelec[A] < 0) return fail(QMC_ERR_ARG, "negative electron count");
            tot += init->atom_nelec[A];
        }
        if (tot != nelec) return fail(QMC_ERR_ARG, "atom_nelec does not sum to nelec");
    }
    if (nw == 0) return QMC_OK;
    if (!pos) return fail(QMC_ERR_ARG, "null pos");
    InitArgs a;
    a.d = *init;
    a.pos = pos;
    a.nw = nw;
    a.nelec = nelec;
    hipLaunchKernelGGL(walker_init_kernel, dim3((unsigned)((nw + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, a);
    QMC_CHECK_HIP(hipGetLastError());
    return QMC_OK;
}

}  // extern "C"
