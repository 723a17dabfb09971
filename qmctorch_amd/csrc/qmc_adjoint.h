// Adjoint kernels with the parameter-gradient reduction fused in (FP64).
//
// For walker weights g_w the wavefunction optimisation needs
//   dL/dW[mu,m]   = sum_w g_w sum_i AO_wi[mu] Gamma_wi[m]          (log|psi|)
//                 = sum_w g_w sum_i AO_wi[mu] GA_wi[m] + b_wi[mu] GB_wi[m]   (E_L)
//   dL/dc_n, dL/dw_ee, dL/dw_en, dL/dbas_exp   (sums of per-walker values)
// (solver.py:352-431).  The row kernels of qmc_kernels.h write AO and Gamma
// rows to memory and leave the contraction to a GEMM; here every wavefront
// evaluates its walkers' rows in registers, recomputes the AO row once more
// after the Slater inverse, and accumulates its share of the sums into a
// partial-sum slot of a workspace; one small kernel adds the slots.
//
// Ownership (no atomics, deterministic for a given slot count): output o of a
// slot is read, accumulated and written by one fixed lane of the slot's
// wavefront -- for W entries (i, m) lane (i mod 8) + 8 (m mod 8), for the
// other outputs lane o mod 64 -- in every iteration of the wavefront's loop
// over the walkers it serves.
#pragma once
#include "qmc_kernels.h"

namespace qmc {

constexpr int kRedCols = 8;  // AO columns staged in LDS per flush (8 row groups of 8 lanes)
constexpr int kMaxSlots = 4096;  // partial-sum slots (wavefronts) of one launch at most

struct RedSlots {
    double* ws;         // [nslot][stride]
    int stride;         // doubles per slot
    int nslot;          // wavefronts of the launch (one slot each)
    int off_ci, off_jas, off_exp;  // off_exp < 0: no exponent terms
    int64_t nwaves;     // wavefronts needed to cover the walkers
    const double* gw;   // [nw] per-walker weights
};

__device__ __forceinline__ int red_owner_w(int i, int m) { return (i & 7) + 8 * (m & 7); }
__device__ __forceinline__ int red_owner(int o) { return o & 63; }

// Sum of v over the 64 lanes, in every lane (xor butterfly, fixed order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Zero the outputs of this wavefront's slot owned by this lane.
__device__ __forceinline__ void red_zero_slot(const RedSlots& r, int nentry, int nmo, int slot, int lane) {
    double* s = r.ws + (size_t)slot * r.stride;
    for (int i = 0; i < nentry; ++i)
        for (int m = 0; m < nmo; ++m)
            if (red_owner_w(i, m) == lane) s[i * nmo + m] = 0.0;
    for (int o = nentry * nmo; o < r.stride; ++o)
        if (red_owner(o) == lane) s[o] = 0.0;
}

// Accumulates sum over lanes r of x[r][i] * G[r][m] (+ y[r][i] * H[r][m]) for
// the AO entries i of a chunk staged in LDS (column i mod 8 of the wave's
// stage rows, 9 doubles per lane), every m < nmo.  Lane l combines rows
// 8 q .. 8 q + 7 (q = l / 8) of column l mod 8, the eight partial sums meet
// by three xor shuffles, and the owner lane adds the result to the slot.
template <int NM, bool TWO>
__device__ __forceinline__ void red_flush_w(double* slot, int i0, int ncol, int nmo, const double* xs,
                                            const double* ys, const double (&G)[NM], const double (&H)[NM],
                                            int lane) {
    const int c = lane & 7, q = lane >> 3;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        if (m < nmo) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int r = 8 * q + k;
                acc += xs[r * (kRedCols + 1) + c] * __shfl(G[m], r);
                if (TWO) acc += ys[r * (kRedCols + 1) + c] * __shfl(H[m], r);
            }
            acc += __shfl_xor(acc, 8);
            acc += __shfl_xor(acc, 16);
            acc += __shfl_xor(acc, 32);
            if (c < ncol && red_owner_w(i0 + c, m) == lane) slot[(i0 + c) * nmo + m] += acc;
        }
    }
}

// Staging of AO entries for red_flush_w during an AO row evaluation: the sink
// stores this lane's value of entry i in column i - i0 and flushes a full chunk.
template <int NM, bool TWO>
struct RedStage {
    double* xs;  // this wave's [64][9] stage rows (values)
    double* ys;  // second stage (the kinetic b rows) when TWO
    double* slot;
    int lane, nmo, i0, ncol;
    const double (*G)[NM];
    const double (*H)[NM];
    __device__ __forceinline__ void put(int i, double x, double y) {
        if (i - i0 >= kRedCols) flush();
        if (ncol == 0) i0 = i - (i & (kRedCols - 1));  // chunks start at multiples of 8
        const int c = i - i0;
        xs[lane * (kRedCols + 1) + c] += x;  // entries of one AO split over groups accumulate
        if (TWO) ys[lane * (kRedCols + 1) + c] += y;
        ncol = c + 1;
    }
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int c = 0; c < kRedCols; ++c) {
            xs[lane * (kRedCols + 1) + c] = 0.0;
            if (TWO) ys[lane * (kRedCols + 1) + c] = 0.0;
        }
    }
    __device__ __forceinline__ void flush() {
        if (ncol == 0) return;
        wave_lds_sync();
        red_flush_w<NM, TWO>(slot, i0, ncol, nmo, xs, ys, *G, *H, lane);
        wave_lds_sync();
        clear();
        wave_lds_sync();
        i0 += kRedCols;
        ncol = 0;
    }
};

// CI combination of a walker from the per-lane unique-configuration values
// (the lanes' spins, shared through the walker's first up and down lanes):
// lmax, S = sum_n c_n s_n exp(l_n - lmax), and per term D_n / S.
struct CiTerms {
    double lu[QMC_MAXCONF], su[QMC_MAXCONF], tu[QMC_MAXCONF];
    double lo[QMC_MAXCONF], so[QMC_MAXCONF], to[QMC_MAXCONF];
    double lmax, S;
    __device__ __forceinline__ void term(const Tables& t, int n, double& l, double& sg, double& tr) const {
        l = 0.0;
        sg = 1.0;
        tr = 0.0;
#pragma unroll
        for (int c = 0; c < QMC_MAXCONF; ++c) {
            if (c == t.ci_up[n]) {
                l += lu[c];
                sg *= su[c];
                tr += tu[c];
            }
            if (c == t.ci_dn[n]) {
                l += lo[c];
                sg *= so[c];
                tr += to[c];
            }
        }
    }
};

__device__ __forceinline__ void ci_terms(const Tables& t, const QmcParams& prm, const WalkerLane& L,
                                         const double (&ld)[QMC_MAXCONF], const double (&sd)[QMC_MAXCONF],
                                         const double (&td)[QMC_MAXCONF], CiTerms& ct) {
    const int up_lane = L.wbase, dn_lane = L.wbase + (t.ndown > 0 ? t.nup : 0);
#pragma unroll
    for (int c = 0; c < QMC_MAXCONF; ++c) {
        ct.lu[c] = __shfl(ld[c], up_lane);
        ct.su[c] = __shfl(sd[c], up_lane);
        ct.tu[c] = __shfl(td[c], up_lane);
        const double l2 = __shfl(ld[c], dn_lane), s2 = __shfl(sd[c], dn_lane), t2 = __shfl(td[c], dn_lane);
        ct.lo[c] = t.ndown > 0 ? l2 : 0.0;
        ct.so[c] = t.ndown > 0 ? s2 : 1.0;
        ct.to[c] = t.ndown > 0 ? t2 : 0.0;
    }
    double lmax = -INFINITY;
    for (int n = 0; n < t.nci; ++n) {
        double l, sg, tr;
        ct.term(t, n, l, sg, tr);
        lmax = fmax(lmax, l);
    }
    if (!(lmax > -INFINITY)) lmax = 0.0;
    double S = 0.0;
    for (int n = 0; n < t.nci; ++n) {
        double l, sg, tr;
        ct.term(t, n, l, sg, tr);
        S += ldc(prm.ci, n) * sg * exp(l - lmax);
    }
    ct.lmax = lmax;
    ct.S = S;
}

// ------------------------------------------------------------------ log|psi|
struct GRedArgs {
    Tables t;
    QmcParams p;
    const double* pos;
    int64_t nw;
    RedSlots r;
};

// Adds g * d log|psi| / d theta of the wavefront's walkers to its slot
// (grad_kernel's quantities: Gamma rows, D_n / S, the Pade weight
// derivatives and the exponent terms, contracted instead of written).
// EXC: plans with exc_mode (Gamma from exc_gamma; xlds its LDS rows).
template <int NM, int RAD, bool EXC = false>
__device__ __forceinline__ void logpsi_adjoint_reduce(const GRedArgs& a, const WalkerLane& L, double* slot,
                                                      double* stage, double* xlds = nullptr) {
    const Tables& t = a.t;
    const double* p = a.pos + (size_t)L.wclamp * 3 * t.nelec + 3 * L.e;
    const double x = p[0], y = p[1], z = p[2];
    // lanes without a walker of this batch (past nw, or the spare lanes of a
    // wavefront) run throw-away arithmetic that may be singular: every value
    // they contribute is selected to zero, never multiplied by a zero weight
    const bool v = L.valid;
    const double gw = v ? a.r.gw[L.wclamp] : 0.0;
    const double* W = (const double*)a.p.mo_w;
    double mo[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) mo[m] = 0.0;
    {
        auto sink = [&](int, int mu, double val, double, double, double, double) {
            const QMC_CONST double* wr = (const QMC_CONST double*)W + (size_t)mu * t.nmo;
#pragma unroll
            for (int m = 0; m < NM; ++m)
                if (m < t.nmo) mo[m] += val * wr[m];
        };
        ao_row<double, RAD, false>(t, a.p, x, y, z, sink);
    }
    double gam[NM];
    if constexpr (EXC) {
        double lg, sg, S;
        exc_gamma<NM>(t, a.p, L, mo, xlds, gam, lg, sg, S);
        // d log|psi| / d c_n = R_n^up R_n^dn / S
        const int blk = t.nsig_max * NM;
        const double* Mu = xlds + (size_t)(L.wl * 2) * 3 * blk;
        const double* Md = Mu + 3 * blk;
        for (int n = 0; n < t.nci; ++n) {
            double Ru, Rd = 1.0, q;
            exc_ratio<NM, false>(t.exc_up + 6 * t.ci_up[n], Mu, nullptr, t.nup, 0.0, Ru, q);
            if (t.ndown > 0) exc_ratio<NM, false>(t.exc_dn + 6 * t.ci_dn[n], Md, nullptr, t.ndown, 0.0, Rd, q);
            const double sum = wave_sum((v && L.e == 0) ? gw * (Ru * Rd / S) : 0.0);
            const int o = a.r.off_ci + n;
            if (red_owner(o) == L.lane) slot[o] += sum;
        }
    } else {
        double ld[QMC_MAXCONF], sd[QMC_MAXCONF], td[QMC_MAXCONF], om[QMC_MAXCONF];
#pragma unroll
        for (int c = 0; c < QMC_MAXCONF; ++c) {
            ld[c] = 0.0;
            sd[c] = 1.0;
            td[c] = 0.0;
            om[c] = 0.0;
        }
#pragma unroll
        for (int m = 0; m < NM; ++m) gam[m] = 0.0;
        conf_dets<NM, false>(t, L, mo, ld, sd, om, gam);
        CiTerms ct;
        ci_terms(t, a.p, L, ld, sd, td, ct);
        for (int n = 0; n < t.nci; ++n) {
            double l, sg, tr;
            ct.term(t, n, l, sg, tr);
            const double dn_s = sg * exp(l - ct.lmax) / ct.S;  // d log|psi| / d c_n
            const int mine = L.spin == 0 ? t.ci_up[n] : t.ci_dn[n];
#pragma unroll
            for (int c = 0; c < QMC_MAXCONF; ++c)
                if (c == mine) om[c] += ldc(a.p.ci, n) * dn_s;
            const double sum = wave_sum((v && L.e == 0) ? gw * dn_s : 0.0);
            const int o = a.r.off_ci + n;
            if (red_owner(o) == L.lane) slot[o] += sum;
        }
        conf_dets<NM, true>(t, L, mo, ld, sd, om, gam);
    }
    // W: stage the recomputed AO row, contract with g * Gamma
    double gG[NM], unused[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        gG[m] = v ? gw * gam[m] : 0.0;
        unused[m] = 0.0;
    }
    {
        RedStage<NM, false> st{stage, nullptr, slot, L.lane, t.nmo, 0, 0, &gG, &unused};
        st.clear();
        wave_lds_sync();
        auto sink = [&](int i, int, double val, double, double, double, double) { st.put(i, v ? val : 0.0, 0.0); };
        ao_row<double, RAD, false>(t, a.p, x, y, z, sink);
        st.flush();
    }
    // exponents: per term slot (entry i, primitive q) the walker sum of
    // g G_i d AO_i / d a_q (G_i = sum_m Gamma[m] W[mu_i][m]), as exp_adjoint
    if (a.r.off_exp >= 0) {
        for (int g = 0; g < t.ngroup; ++g) {
            const int at = t.grp_atom[g];
            const double dx = x - ldc(a.p.atom_coords, 3 * at), dy = y - ldc(a.p.atom_coords, 3 * at + 1),
                         dz = z - ldc(a.p.atom_coords, 3 * at + 2);
            const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
            const int np = t.grp_nprim[g], p0 = t.grp_prim0[g];
            for (int i = t.grp_ao0[g]; i < t.grp_ao0[g] + t.grp_nao[g]; ++i) {
                const int mu = t.gao_mu[i];
                const QMC_CONST double* wr = (const QMC_CONST double*)W + (size_t)mu * t.nmo;
                double G = 0.0;
#pragma unroll
                for (int m = 0; m < NM; ++m)
                    if (m < t.nmo) G += gam[m] * wr[m];
                const int kk = t.gao_k[i];
                double Y, d0, d1, d2, d3;
                cart_poly<double, false>(dx, dy, dz, kk & 15, (kk >> 4) & 15, (kk >> 8) & 15, Y, d0, d1, d2, d3);
                for (int q = 0; q < np; ++q) {
                    const double al = ldc(a.p.bas_exp, t.grp_prim_src[p0 + q]);
                    const double K = t.gao_K[(size_t)i * QMC_MAXP + q];
                    double term;
                    if (RAD == QMC_RADIAL_GTO_PURE) {
                        term = -r2 * K * exp(-al * r2) * Y;
                    } else {
                        const int nn = (int)t.gao_n[(size_t)i * QMC_MAXP + q];
                        term = -r * K * ipow(r, nn) * exp(-al * r) * Y;
                    }
                    const double sum = wave_sum(v ? gw * G * term : 0.0);
                    const int o = a.r.off_exp + i * QMC_MAXP + q;
                    if (red_owner(o) == L.lane) slot[o] += sum;
                }
            }
        }
    }
    // Pade weights: d/dw [w0 r/(1+w r)] = -w0 r^2/(1+w r)^2
    double dee = 0.0, den_ = 0.0;
    const double wee = (t.use_ee && a.p.w_ee) ? ldc(a.p.w_ee, 0) : 0.0;
    const double wen = (t.use_en && a.p.w_en) ? ldc(a.p.w_en, 0) : 0.0;
    for (int j = 0; j < t.nelec; ++j) {
        const double xj = __shfl(x, L.wbase + j), yj = __shfl(y, L.wbase + j), zj = __shfl(z, L.wbase + j);
        if (t.use_ee && j != L.e) {
            const double rr = sqrt((x - xj) * (x - xj) + (y - yj) * (y - yj) + (z - zj) * (z - zj));
            const double w0 = ((j < t.nup) == (L.e < t.nup)) ? 0.25 : 0.5;
            const double qq = 1.0 / (1.0 + wee * rr);
            dee -= 0.5 * w0 * rr * rr * qq * qq;
        }
    }
    if (t.use_en) {
        for (int A = 0; A < t.natom; ++A) {
            const double dx = x - t.jas_atoms[3 * A], dy = y - t.jas_atoms[3 * A + 1], dz = z - t.jas_atoms[3 * A + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double qq = 1.0 / (1.0 + wen * sqrt(r2));
            den_ -= r2 * qq * qq;
        }
    }
    const double sums[2] = {wave_sum(v ? gw * dee : 0.0), wave_sum(v ? gw * den_ : 0.0)};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int o = a.r.off_jas + k;
        if (red_owner(o) == L.lane) slot[o] += sums[k];
    }
}

// Stage rows of the wavefronts of a block: NS x [64][9] doubles per wave.
template <int NS>
__device__ __forceinline__ double* red_stage_rows() {
    __shared__ double st[kBlockWaves * NS * 64 * (kRedCols + 1)];
    return st + (threadIdx.x >> 6) * NS * 64 * (kRedCols + 1);
}

template <int NM, int RAD>
__global__ __launch_bounds__(kBlock) void grad_reduce_kernel(GRedArgs a) {
    const int slot_id = global_wave();
    if (slot_id >= a.r.nslot) return;
    const int lane = threadIdx.x & 63;
    double* slot = a.r.ws + (size_t)slot_id * a.r.stride;
    red_zero_slot(a.r, a.t.nentry, a.t.nmo, slot_id, lane);
    double* stage = red_stage_rows<1>();
    for (int64_t vw = slot_id; vw < a.r.nwaves; vw += a.r.nslot) {
        const WalkerLane L = lane_layout_w(a.t, a.nw, (int)vw);
        logpsi_adjoint_reduce<NM, RAD>(a, L, slot, stage);
    }
}

// Plans with exc_mode: one wavefront per block (its exc_gamma rows in the
// dynamic LDS, 2 (G + 1) groups of 3 x nsig_max x NM doubles), one slot each.
template <int NM, int RAD>
__global__ __launch_bounds__(64) void grad_reduce_kernel_exc(GRedArgs a) {
    const int slot_id = blockIdx.x;
    if (slot_id >= a.r.nslot) return;
    const int lane = threadIdx.x & 63;
    double* slot = a.r.ws + (size_t)slot_id * a.r.stride;
    red_zero_slot(a.r, a.t.nentry, a.t.nmo, slot_id, lane);
    __shared__ double stage[64 * (kRedCols + 1)];
    for (int64_t vw = slot_id; vw < a.r.nwaves; vw += a.r.nslot) {
        const WalkerLane L = lane_layout_w(a.t, a.nw, (int)vw);
        logpsi_adjoint_reduce<NM, RAD, true>(a, L, slot, stage, (double*)qmc_dyn_lds);
    }
}

// ------------------------------------------------------------------ E_L
struct ERedArgs {
    Tables t;
    QmcParams p;
    const double* pos;
    int64_t nw;
    double* eloc;  // [nw] or null
    RedSlots r;
};

// Adds g * d E_L / d theta of the wavefront's walkers to its slot
// (eloc_grad_kernel's quantities, contracted instead of written).
// EXC: plans with exc_mode (exc_eloc_gamma's rows in xlds).
template <int NM, int RAD, bool EXC = false>
__device__ __forceinline__ void eloc_adjoint_reduce(const ERedArgs& a, const WalkerLane& L, double* slot,
                                                    double* stage, double* xlds = nullptr) {
    const Tables& t = a.t;
    const double* p = a.pos + (size_t)L.wclamp * 3 * t.nelec + 3 * L.e;
    const double x = p[0], y = p[1], z = p[2];
    const bool v = L.valid;
    const double gw = v ? a.r.gw[L.wclamp] : 0.0;
    const double* W = (const double*)a.p.mo_w;

    double g[3], l, dge[3], dle, dgn[3], dln;
    jastrow_w_derivs(t, a.p, L, x, y, z, g, l, dge, dle, dgn, dln);
    const double l2 = l + g[0] * g[0] + g[1] * g[1] + g[2] * g[2];

    double mo[NM], gx[NM], gy[NM], gz[NM], lp[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) mo[m] = gx[m] = gy[m] = gz[m] = lp[m] = 0.0;
    {
        auto sink = [&](int, int mu, double ao, double ax, double ay, double az, double al) {
            const QMC_CONST double* wr = (const QMC_CONST double*)W + (size_t)mu * t.nmo;
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                if (m < t.nmo) {
                    const double w = wr[m];
                    mo[m] += ao * w;
                    gx[m] += ax * w;
                    gy[m] += ay * w;
                    gz[m] += az * w;
                    lp[m] += al * w;
                }
            }
        };
        ao_row<double, RAD, true>(t, a.p, x, y, z, sink);
    }
    double bk[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m)
        bk[m] = -0.5 * (lp[m] + 2.0 * (g[0] * gx[m] + g[1] * gy[m] + g[2] * gz[m]) + l2 * mo[m]);

    double GA[NM], GB[NM];
    double s[3] = {0.0, 0.0, 0.0};
    double T;
    if constexpr (EXC) {
        T = exc_eloc_gamma<NM>(t, a.p, L, mo, bk, xlds, GA, GB, [&](int nn, double val) {
            const double sum = wave_sum((v && L.e == 0) ? gw * val : 0.0);
            const int o = a.r.off_ci + nn;
            if (red_owner(o) == L.lane) slot[o] += sum;
        });
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            if (m < t.nmo) {
                s[0] += GB[m] * gx[m];
                s[1] += GB[m] * gy[m];
                s[2] += GB[m] * gz[m];
            }
        }
    } else {
        double ld[QMC_MAXCONF], sd[QMC_MAXCONF], td[QMC_MAXCONF];
#pragma unroll
        for (int c = 0; c < QMC_MAXCONF; ++c) {
            ld[c] = 0.0;
            sd[c] = 1.0;
            td[c] = 0.0;
        }
#pragma unroll 1
        for (int c = 0; c < t.nconf_max; ++c) {
            double aa[NM], bb[NM];
            gather_conf<NM>(t, L, c, mo, aa);
            gather_conf<NM>(t, L, c, bk, bb);
            double lg, sg, tp;
            int col;
            lane_gj<NM>(L, t.nsig_max, aa, bb, lg, sg, tp, col);
            const double tr = seg_sum(tp, L.gbase, L.n_sig, t.nsig_max);
#pragma unroll
            for (int c2 = 0; c2 < QMC_MAXCONF; ++c2)
                if (c2 == c) {
                    ld[c2] = lg;
                    sd[c2] = sg;
                    td[c2] = tr;
                }
        }
        CiTerms ct;
        ci_terms(t, a.p, L, ld, sd, td, ct);
        double ST = 0.0;
        for (int n = 0; n < t.nci; ++n) {
            double lg, sg, tr;
            ct.term(t, n, lg, sg, tr);
            ST += ldc(a.p.ci, n) * sg * exp(lg - ct.lmax) * tr;
        }
        T = ST / ct.S;
        double rho[QMC_MAXCONF], om[QMC_MAXCONF];
#pragma unroll
        for (int c = 0; c < QMC_MAXCONF; ++c) rho[c] = om[c] = 0.0;
        for (int n = 0; n < t.nci; ++n) {
            double lg, sg, tr;
            ct.term(t, n, lg, sg, tr);
            const double dns = sg * exp(lg - ct.lmax) / ct.S;
            const double wn = ldc(a.p.ci, n) * dns;
            const int mine = L.spin == 0 ? t.ci_up[n] : t.ci_dn[n];
#pragma unroll
            for (int c = 0; c < QMC_MAXCONF; ++c)
                if (c == mine) {
                    rho[c] += wn * (tr - T);
                    om[c] += wn;
                }
            const double sum = wave_sum((v && L.e == 0) ? gw * dns * (tr - T) : 0.0);
            const int o = a.r.off_ci + n;
            if (red_owner(o) == L.lane) slot[o] += sum;
        }
#pragma unroll
        for (int m = 0; m < NM; ++m) GA[m] = GB[m] = 0.0;
#pragma unroll 1
        for (int c = 0; c < t.nconf_max; ++c) {
            double aa[NM], xb[NM], ib[NM];
            gather_conf<NM>(t, L, c, mo, aa);
            gather_conf<NM>(t, L, c, bk, xb);
            double lg, sg, tp;
            int col, col2;
            lane_gj<NM>(L, t.nsig_max, aa, xb, lg, sg, tp, col);
            gather_conf<NM>(t, L, c, mo, aa);
#pragma unroll
            for (int j = 0; j < NM; ++j) ib[j] = (j == L.myrow) ? 1.0 : 0.0;
            lane_gj<NM>(L, t.nsig_max, aa, ib, lg, sg, tp, col2);
            double inv[NM], zz[NM];
            inverse_columns<NM>(L, t.nsig_max, col, xb, ib, inv, zz);
            double rc = 0.0, oc = 0.0;
#pragma unroll
            for (int c2 = 0; c2 < QMC_MAXCONF; ++c2)
                if (c2 == c) {
                    rc = rho[c2];
                    oc = om[c2];
                }
            const QMC_CONST int* conf = L.spin == 0 ? t.conf_up + c * t.nup : t.conf_dn + c * t.ndown;
            const int nconf = L.spin == 0 ? t.nconf_up : t.nconf_dn;
            if (c < nconf) {
#pragma unroll
                for (int k = 0; k < NM; ++k) {
                    if (k < L.n_sig) {
                        const int mcol = t.identity_conf ? k : conf[k];
#pragma unroll
                        for (int m = 0; m < NM; ++m) {
                            if (m == mcol) {
                                GA[m] += rc * inv[k] - oc * zz[k];
                                GB[m] += oc * inv[k];
                                s[0] += oc * inv[k] * gx[m];
                                s[1] += oc * inv[k] * gy[m];
                                s[2] += oc * inv[k] * gz[m];
                            }
                        }
                    }
                }
            }
        }
    }
    // W: stage the recomputed AO and kinetic rows, contract with g GA and g GB
    {
        double gA[NM], gB[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            gA[m] = v ? gw * GA[m] : 0.0;
            gB[m] = v ? gw * GB[m] : 0.0;
        }
        RedStage<NM, true> st{stage, stage + 64 * (kRedCols + 1), slot, L.lane, t.nmo, 0, 0, &gA, &gB};
        st.clear();
        wave_lds_sync();
        auto sink = [&](int i, int, double ao, double ax, double ay, double az, double al) {
            st.put(i, v ? ao : 0.0, v ? -0.5 * (al + 2.0 * (g[0] * ax + g[1] * ay + g[2] * az) + l2 * ao) : 0.0);
        };
        ao_row<double, RAD, true>(t, a.p, x, y, z, sink);
        st.flush();
    }
    // Pade weights
    double dte = 0.0, dtn = 0.0;
    for (int d = 0; d < 3; ++d) {
        dte -= dge[d] * (s[d] + g[d]);
        dtn -= dgn[d] * (s[d] + g[d]);
    }
    dte -= 0.5 * dle;
    dtn -= 0.5 * dln;
    const double sdte = seg_sum(dte, L.wbase, t.nelec, t.nelec);
    const double sdtn = seg_sum(dtn, L.wbase, t.nelec, t.nelec);
    const double sums[2] = {wave_sum((v && L.e == 0) ? gw * sdte : 0.0), wave_sum((v && L.e == 0) ? gw * sdtn : 0.0)};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int o = a.r.off_jas + k;
        if (red_owner(o) == L.lane) slot[o] += sums[k];
    }
    if (a.eloc) {
        double kee, ken, d0, d1, d2, d3, vee, ven;
        jastrow_and_potential<true>(t, a.p, L, x, y, z, kee, ken, d0, d1, d2, d3, vee, ven);
        const double ep = seg_sum(vee + ven, L.wbase, t.nelec, t.nelec) + t.vnn;
        if (v && L.e == 0) a.eloc[L.w] = T + ep;
    }
}

template <int NM, int RAD>
__global__ __launch_bounds__(kBlock) void eloc_grad_reduce_kernel(ERedArgs a) {
    const int slot_id = global_wave();
    if (slot_id >= a.r.nslot) return;
    const int lane = threadIdx.x & 63;
    double* slot = a.r.ws + (size_t)slot_id * a.r.stride;
    red_zero_slot(a.r, a.t.nentry, a.t.nmo, slot_id, lane);
    double* stage = red_stage_rows<2>();
    for (int64_t vw = slot_id; vw < a.r.nwaves; vw += a.r.nslot) {
        const WalkerLane L = lane_layout_w(a.t, a.nw, (int)vw);
        eloc_adjoint_reduce<NM, RAD>(a, L, slot, stage);
    }
}

// Plans with exc_mode: one wavefront per block, exc_eloc_gamma's rows in the
// dynamic LDS (2 (G + 1) groups of 6 x nsig_max x NM doubles), one slot each.
template <int NM, int RAD>
__global__ __launch_bounds__(64) void eloc_grad_reduce_kernel_exc(ERedArgs a) {
    const int slot_id = blockIdx.x;
    if (slot_id >= a.r.nslot) return;
    const int lane = threadIdx.x & 63;
    double* slot = a.r.ws + (size_t)slot_id * a.r.stride;
    red_zero_slot(a.r, a.t.nentry, a.t.nmo, slot_id, lane);
    __shared__ double stage[2 * 64 * (kRedCols + 1)];
    for (int64_t vw = slot_id; vw < a.r.nwaves; vw += a.r.nslot) {
        const WalkerLane L = lane_layout_w(a.t, a.nw, (int)vw);
        eloc_adjoint_reduce<NM, RAD, true>(a, L, slot, stage, (double*)qmc_dyn_lds);
    }
}

// ------------------------------------------------------------ slot reduction
// Two launches.  slot_fold_kernel: slot p (p < nfold) receives, column by
// column, the sum of the slots p, p + nfold, p + 2 nfold, ...  Block (x, p)
// owns columns [64 x, 64 x + 64) of exactly those slots, so its in-place
// write races with no other block; its four wavefronts take every fourth slot
// of the partition and their sums are combined in a fixed order.
// slot_reduce_kernel, one wavefront per output, lane s reading folded slot s:
// dW[mu][m] over the AO entries of mu, the other outputs directly, exponent
// terms (entry i, primitive q) to bas_exp index gao_src[i][q] (the plan's
// inverse maps list both), the 64 lanes summed by a fixed xor butterfly.
struct RedOut {
    double* grad_w;    // [nao][nmo]
    double* grad_ci;   // [nci]
    double* grad_jas;  // [2]
    double* grad_exp;  // [nprim_ref] or null (its exponent terms must be in the slots)
};

constexpr int kFoldBlock = 256;  // four wavefronts per block of slot_fold_kernel

inline dim3 grid_for_waves_of(int nwaves) { return dim3((unsigned)((nwaves + kBlockWaves - 1) / kBlockWaves)); }
constexpr int kMaxFold = 64;     // folded slots left for slot_reduce_kernel (one per lane)
static_assert(kMaxFold <= 64, "slot_reduce_kernel reads one folded slot per lane");

inline int red_nfold(const RedSlots& r) { return r.nslot < kMaxFold ? r.nslot : kMaxFold; }

static __global__ __launch_bounds__(kFoldBlock) void slot_fold_kernel(RedSlots r, int nfold) {
    __shared__ double part[kFoldBlock / 64][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + lane;
    const int64_t p = blockIdx.y;
    double acc = 0.0;
    if (col < r.stride) {
#pragma unroll 4
        for (int64_t sl = p + (int64_t)nfold * w; sl < r.nslot; sl += (int64_t)nfold * (kFoldBlock / 64))
            acc += r.ws[sl * r.stride + col];
    }
    part[w][lane] = acc;
    __syncthreads();
    if (w == 0 && col < r.stride) {
        double sum = part[0][lane];
#pragma unroll
        for (int k = 1; k < kFoldBlock / 64; ++k) sum += part[k][lane];
        r.ws[p * r.stride + col] = sum;
    }
}

static __global__ __launch_bounds__(kBlock) void slot_reduce_kernel(Tables t, RedSlots r, RedOut o, int nfold) {
    const int n_w = t.nao * t.nmo;
    const int n_all = n_w + t.nci + 2 + (o.grad_exp ? t.nprim_ref : 0);
    const int idx = global_wave();  // one wavefront per output, lane s reads folded slot s
    if (idx >= n_all) return;
    const int lane = threadIdx.x & 63;
    const bool mine = lane < nfold;
    const double* ws = r.ws + (size_t)(mine ? lane : 0) * r.stride;
    double acc = 0.0;
    if (idx < n_w) {
        const int mu = idx / t.nmo, m = idx - mu * t.nmo;
        for (int e = t.ao_ptr[mu]; e < t.ao_ptr[mu + 1]; ++e) acc += mine ? ws[t.ao_ent[e] * t.nmo + m] : 0.0;
    } else if (idx < n_w + t.nci) {
        acc = mine ? ws[r.off_ci + idx - n_w] : 0.0;
    } else if (idx < n_w + t.nci + 2) {
        acc = mine ? ws[r.off_jas + idx - n_w - t.nci] : 0.0;
    } else {
        const int src = idx - n_w - t.nci - 2;  // bas_exp index
        for (int c = t.exp_ptr[src]; c < t.exp_ptr[src + 1]; ++c) acc += mine ? ws[r.off_exp + t.exp_col[c]] : 0.0;
    }
    acc = wave_sum(acc);
    if (lane == 0) {
        if (idx < n_w)
            o.grad_w[idx] = acc;
        else if (idx < n_w + t.nci)
            o.grad_ci[idx - n_w] = acc;
        else if (idx < n_w + t.nci + 2)
            o.grad_jas[idx - n_w - t.nci] = acc;
        else
            o.grad_exp[idx - n_w - t.nci - 2] = acc;
    }
}

// The slot sum on stream s: the fold (when there are more slots than
// kMaxFold) and the per-output reduction.
inline void red_slot_sum(const Tables& t, const RedSlots& r, const RedOut& o, hipStream_t s) {
    const int nfold = red_nfold(r);
    if (r.nslot > nfold)
        hipLaunchKernelGGL(slot_fold_kernel, dim3((unsigned)((r.stride + 63) / 64), (unsigned)nfold), dim3(kFoldBlock),
                           0, s, r, nfold);
    const int n = t.nao * t.nmo + t.nci + 2 + (o.grad_exp ? t.nprim_ref : 0);
    hipLaunchKernelGGL(slot_reduce_kernel, grid_for_waves_of(n), dim3(kBlock), 0, s, t, r, o, nfold);
}

// Layout of a slot and the slot count for nw walkers within ws_bytes.
inline RedSlots red_layout(const Tables& t, int64_t nw, bool with_exp, double* ws, size_t ws_bytes,
                           const double* gw) {
    RedSlots r;
    r.ws = ws;
    r.gw = gw;
    r.off_ci = t.nentry * t.nmo;
    r.off_jas = r.off_ci + t.nci;
    r.off_exp = with_exp ? r.off_jas + 2 : -1;
    r.stride = r.off_jas + 2 + (with_exp ? t.nentry * QMC_MAXP : 0);
    const int G = 64 / t.nelec;
    r.nwaves = (nw + G - 1) / G;
    const int64_t fit = (int64_t)(ws_bytes / ((size_t)r.stride * sizeof(double)));
    int64_t ns = r.nwaves < kMaxSlots ? r.nwaves : kMaxSlots;
    r.nslot = (int)(ns < fit ? ns : fit);
    return r;
}

}  // namespace qmc
