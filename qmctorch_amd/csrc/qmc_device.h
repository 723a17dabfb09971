// Device building blocks of the VMC kernels (gfx950, wave64).
//
// Lane layout used by every kernel: a wavefront holds G = 64 / nelec walkers,
// lane l = wl * nelec + e serves electron e of walker wl.  Basis, MO and CI
// tables are identical for all lanes of a launch (scalar-cache loads); each
// lane keeps its electron's MO row in registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qmc_mi355x.h"

namespace qmc {

// The tables and parameters of a launch are never written while its kernels
// run: device code reads them through the constant address space, so that
// wave-uniform reads become scalar loads whether the pointer arrived as a
// kernel argument or from a job table in device memory.
#if defined(__HIP_DEVICE_COMPILE__)
#define QMC_CONST __attribute__((address_space(4)))
#define QMC_LDS __attribute__((address_space(3)))
#else
#define QMC_CONST
#define QMC_LDS
#endif

template <typename X>
__device__ __forceinline__ X ldc(const X* p, size_t i) {
    return ((const __attribute__((address_space(4))) X*)p)[i];
}

struct Tables {
    int nelec, nup, ndown, natom, nao, nprim_ref, radial;
    int ngroup, nentry, nmo, nci, nconf_up, nconf_dn, nconf_max, nsig_max;
    int use_ee, use_en, identity_conf;
    int exc_mode;  // configurations as single/double excitations (trace-trick kernels)
    double vnn;
    const QMC_CONST int* grp_atom;
    const QMC_CONST int* grp_nprim;
    const QMC_CONST int* grp_prim0;
    const QMC_CONST int* grp_prim_src;
    const QMC_CONST int* grp_ao0;
    const QMC_CONST int* grp_nao;
    const QMC_CONST int* grp_tight;
    const QMC_CONST int* gao_mu;
    const QMC_CONST int* gao_k;
    const QMC_CONST double* gao_K;
    const QMC_CONST float* gao_Kf;
    const QMC_CONST double* gao_n;
    const QMC_CONST int* gao_src;
    const QMC_CONST double* atom_z;
    const QMC_CONST double* jas_atoms;
    const QMC_CONST int* conf_up;
    const QMC_CONST int* conf_dn;
    const QMC_CONST int* ci_up;
    const QMC_CONST int* ci_dn;
    const QMC_CONST int* exc_up;  // [nconf_up][6] {level, h1, h2, p1, p2, sign}, or null
    const QMC_CONST int* exc_dn;
    // inverse maps of the adjoint slot sum (CSR): the AO entries of AO mu are
    // ao_ent[ao_ptr[mu] .. ao_ptr[mu + 1]), the exponent-term columns
    // (entry i, primitive q) -> i * QMC_MAXP + q of bas_exp index k are
    // exp_col[exp_ptr[k] .. exp_ptr[k + 1]), both ascending
    const QMC_CONST int* ao_ptr;
    const QMC_CONST int* ao_ent;
    const QMC_CONST int* exp_ptr;
    const QMC_CONST int* exp_col;
};

// ------------------------------------------------------------------ math
// exp(x) in FP32 as the hardware 2^y of y = x log2 e: relative error about
// (|x| + 1) 2^-23, of the order of the rounding already in the argument
// (x = -a r^2 in FP32 carries |x| 2^-24); a compensated y (x log2 e split into
// its rounded product and the fma residual) halves the first term for three
// more instructions per exponential, which the sweeps do not need.
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269502162933349609375f); }
__device__ __forceinline__ double fexp(double x) { return exp(x); }

// 1/x: the hardware reciprocal (1 ulp) in FP32, a division in FP64
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }

// log(m) for m in [0.5, 1) (a frexp mantissa): with m' = m or 2m in
// [1/sqrt 2, sqrt 2), z = (m' - 1)/(m' + 1), |z| < 0.172,
// log m' = 2 atanh z = 2 (z + z^3/3 + ... + z^21/21) (truncation < 1e-19),
// log m = log m' - [m' = 2m] log 2
// A wave-uniform constant formed where it is used (scalar moves) rather than
// hoisted by the compiler over an enclosing loop into registers it then spills.
__device__ __forceinline__ double sconst(double v) {
    asm volatile("" : "+s"(v));
    return v;
}

__device__ __forceinline__ double log_mantissa(double m) {
    const bool up = m < 0.70710678118654752;
    const double mm = up ? 2.0 * m : m;
    const double z = (mm - 1.0) / (mm + 1.0);
    const double s = z * z;
    double p = sconst(1.0 / 21.0);
    p = fma(p, s, sconst(1.0 / 19.0));
    p = fma(p, s, sconst(1.0 / 17.0));
    p = fma(p, s, sconst(1.0 / 15.0));
    p = fma(p, s, sconst(1.0 / 13.0));
    p = fma(p, s, sconst(1.0 / 11.0));
    p = fma(p, s, sconst(1.0 / 9.0));
    p = fma(p, s, sconst(1.0 / 7.0));
    p = fma(p, s, sconst(1.0 / 5.0));
    p = fma(p, s, sconst(1.0 / 3.0));
    p = fma(p, s, 1.0);
    return 2.0 * z * p - (up ? sconst(0.6931471805599453) : 0.0);
}
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }  // hardware, 1 ulp
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }

template <typename T>
__device__ __forceinline__ T ipow(T x, int k) {
    // k in 0..3 (Cartesian exponents / STO r powers of the supported bases)
    T x2 = x * x;
    return k == 0 ? T(1) : (k == 1 ? x : (k == 2 ? x2 : x2 * x));
}

// Cartesian polynomial x^kx y^ky z^kz with gradient and Laplacian
// (spherical_harmonics.py:102-199, Cartesian branch).
template <typename T, bool DERIV>
__device__ __forceinline__ void cart_poly(T x, T y, T z, int kx, int ky, int kz, T& Y, T& Yx, T& Yy, T& Yz,
                                          T& Ylap) {
    T px = ipow(x, kx), py = ipow(y, ky), pz = ipow(z, kz);
    Y = px * py * pz;
    if (DERIV) {
        T dx = kx > 0 ? T(kx) * ipow(x, kx - 1) : T(0);
        T dy = ky > 0 ? T(ky) * ipow(y, ky - 1) : T(0);
        T dz = kz > 0 ? T(kz) * ipow(z, kz - 1) : T(0);
        Yx = dx * py * pz;
        Yy = px * dy * pz;
        Yz = px * py * dz;
        T d2x = kx > 1 ? T(kx * (kx - 1)) * ipow(x, kx - 2) : T(0);
        T d2y = ky > 1 ? T(ky * (ky - 1)) * ipow(y, ky - 2) : T(0);
        T d2z = kz > 1 ? T(kz * (kz - 1)) * ipow(z, kz - 2) : T(0);
        Ylap = d2x * py * pz + px * d2y * pz + px * py * d2z;
    }
}

// ------------------------------------------------------------ AO -> MO row
// Evaluates the AOs of one electron group by group and contracts them into
// the MO row on the fly (the AO row never leaves registers).
//
//   GTO (radial_functions.py:241-320): with S_k = sum_p K_p a_p^k exp(-a_p r^2)
//     AO = Y S0,  grad AO = S0 grad Y - 2 S1 Y r,
//     lap AO = Y (4 r^2 S2 - (6 + 4L) S1) + S0 lap Y          (L = kx+ky+kz)
//   STO (radial_functions.py:6-123): R_p = r^n exp(-a r), q = n/r^2 - a/r,
//     AO = Y sum K R,  grad AO = (sum K R) grad Y + Y (sum K R q) r,
//     lap AO = Y sum K R (n(n+1)/r^2 - 2a(n+1)/r + a^2 + 2Lq) + (sum K R) lap Y
//
// sink(i, mu, ao, gx, gy, gz, lap) is called once per AO entry i (AO mu); entries of one
// AO that were split across groups are delivered separately (all consumers
// are linear in the AO).  The primitive count of a group selects a fully
// unrolled instance (no predicated slots).
template <typename T>
__device__ __forceinline__ const QMC_CONST T* coeff_table(const Tables& t);
template <>
__device__ __forceinline__ const QMC_CONST float* coeff_table<float>(const Tables& t) {
    return t.gao_Kf;
}
template <>
__device__ __forceinline__ const QMC_CONST double* coeff_table<double>(const Tables& t) {
    return t.gao_K;
}

// exponents gathered in group-primitive order, in the evaluation precision
template <typename T>
__device__ __forceinline__ const T* group_exponents(const QmcParams& prm);
template <>
__device__ __forceinline__ const float* group_exponents<float>(const QmcParams& prm) {
    return prm.grp_expf;
}
template <>
__device__ __forceinline__ const double* group_exponents<double>(const QmcParams& prm) {
    return prm.grp_exp;
}

// exp(-a r^2) (or exp(-a r)) of the NP primitives of a group whose exponents
// start at p0 of the gathered exponent list
template <typename T, int RAD, int NP>
__device__ __forceinline__ void group_prims_from(const T* gexp, int p0, T r2, T r, T (&E)[NP], T (&A)[NP]) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const T a = ldc(gexp, p0 + p);
        A[p] = a;
        E[p] = (RAD == QMC_RADIAL_GTO_PURE) ? fexp(-a * r2) : fexp(-a * r);
    }
}

template <typename T, int RAD, int NP>
__device__ __forceinline__ void group_prims(const QmcParams& prm, int p0, T r2, T r, T (&E)[NP], T (&A)[NP]) {
    group_prims_from<T, RAD, NP>(group_exponents<T>(prm), p0, r2, r, E, A);
}

// AO entry i (AO mu, Cartesian powers kx, ky, kz) of a group from its
// primitive exponentials; hands value, gradient and Laplacian to the sink.
// kcoef(p) is the contraction coefficient K_p of primitive p, rpow(p) its STO
// r power (unused for Gaussians): read from the plan tables, or written into a
// generated row as literals.
template <typename T, int RAD, bool DERIV, int NP, typename KF, typename NF, typename Sink>
__device__ __forceinline__ void ao_entry_with(KF kcoef, NF rpow, int i, int mu, int kx, int ky, int kz,
                                              const T (&E)[NP], const T (&A)[NP], T dx, T dy, T dz, T r2, T r,
                                              T ir, Sink& sink) {
    T Y, Yx = 0, Yy = 0, Yz = 0, Yl = 0;
    cart_poly<T, DERIV>(dx, dy, dz, kx, ky, kz, Y, Yx, Yy, Yz, Yl);
    T s0 = 0, s1 = 0, s2 = 0;
    if (RAD == QMC_RADIAL_GTO_PURE) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const T ke = kcoef(p) * E[p];
            s0 += ke;
            if (DERIV) {
                s1 += ke * A[p];
                s2 += ke * A[p] * A[p];
            }
        }
        const T ao = Y * s0;
        if (DERIV) {
            const T L = T(kx + ky + kz);
            const T c = T(-2) * s1 * Y;
            sink(i, mu, ao, s0 * Yx + c * dx, s0 * Yy + c * dy, s0 * Yz + c * dz,
                 Y * (T(4) * r2 * s2 - (T(6) + T(4) * L) * s1) + s0 * Yl);
        } else {
            sink(i, mu, ao, T(0), T(0), T(0), T(0));
        }
    } else {
        const T L = T(kx + ky + kz);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const double np_ = rpow(p);
            const T n = (T)np_;
            const int ni = (int)np_;
            const T kr = kcoef(p) * ipow(r, ni) * E[p];
            s0 += kr;
            if (DERIV) {
                const T a = A[p];
                const T q = n * ir * ir - a * ir;
                s1 += kr * q;
                s2 += kr * (n * (n + T(1)) * ir * ir - T(2) * a * (n + T(1)) * ir + a * a + T(2) * L * q);
            }
        }
        const T ao = Y * s0;
        if (DERIV) {
            const T c = Y * s1;
            sink(i, mu, ao, s0 * Yx + c * dx, s0 * Yy + c * dy, s0 * Yz + c * dz, Y * s2 + s0 * Yl);
        } else {
            sink(i, mu, ao, T(0), T(0), T(0), T(0));
        }
    }
}

template <typename T, int RAD, bool DERIV, int NP, typename Sink>
__device__ __forceinline__ void ao_entry(const Tables& t, int i, int mu, int kx, int ky, int kz, const T (&E)[NP],
                                         const T (&A)[NP], T dx, T dy, T dz, T r2, T r, T ir, Sink& sink) {
    const QMC_CONST T* Kp = coeff_table<T>(t) + (size_t)i * QMC_MAXP;
    const QMC_CONST double* Np = t.gao_n + (size_t)i * QMC_MAXP;
    ao_entry_with<T, RAD, DERIV, NP>([&](int p) { return Kp[p]; }, [&](int p) { return Np[p]; }, i, mu, kx, ky, kz,
                                     E, A, dx, dy, dz, r2, r, ir, sink);
}

template <typename T, int RAD, bool DERIV, int NP, typename Sink>
__device__ __forceinline__ void ao_group(const Tables& t, int g, const QmcParams& prm, T dx, T dy, T dz, T r2, T r,
                                         T ir, Sink& sink) {
    T E[NP], A[NP];
    group_prims<T, RAD, NP>(prm, t.grp_prim0[g], r2, r, E, A);
    const int i0 = t.grp_ao0[g], i1 = i0 + t.grp_nao[g];
    for (int i = i0; i < i1; ++i) {
        const int kk = t.gao_k[i];
        ao_entry<T, RAD, DERIV, NP>(t, i, t.gao_mu[i], kk & 15, (kk >> 4) & 15, (kk >> 8) & 15, E, A, dx, dy, dz, r2,
                                    r, ir, sink);
    }
}

// TIGHT: groups flagged in grp_tight are evaluated in FP64 whatever T is (the
// mixed-precision local energy: tight core primitives cancel in lap AO near the
// nucleus, which FP32 cannot resolve).
template <typename T, int RAD, bool DERIV, bool TIGHT = false, typename Sink>
__device__ __forceinline__ void ao_row(const Tables& t, const QmcParams& prm, double ex, double ey, double ez,
                                       Sink&& sink) {
    const double* atom_xyz = prm.atom_coords;
    int cur_atom = -1;
    double dxd = 0, dyd = 0, dzd = 0, r2d = 0, rd = 0, ird = 0;
    T dx = 0, dy = 0, dz = 0, r2 = 0, r = 0, ir = 0;
    for (int g = 0; g < t.ngroup; ++g) {
        const int at = t.grp_atom[g];
        if (at != cur_atom) {
            cur_atom = at;
            dxd = ex - ldc(atom_xyz, 3 * at + 0);
            dyd = ey - ldc(atom_xyz, 3 * at + 1);
            dzd = ez - ldc(atom_xyz, 3 * at + 2);
            dx = (T)dxd;
            dy = (T)dyd;
            dz = (T)dzd;
            r2 = dx * dx + dy * dy + dz * dz;
            if (RAD == QMC_RADIAL_STO) {
                r = fsqrt(r2);
                ir = T(1) / r;
            }
            if (TIGHT) {
                r2d = dxd * dxd + dyd * dyd + dzd * dzd;
                if (RAD == QMC_RADIAL_STO) {
                    rd = sqrt(r2d);
                    ird = 1.0 / rd;
                }
            }
        }
        const bool tight = TIGHT && t.grp_tight != nullptr && t.grp_tight[g] != 0;
        switch (t.grp_nprim[g]) {
#define QMC_AO_CASE(K)                                                                               \
    case K:                                                                                          \
        if (tight)                                                                                   \
            ao_group<double, RAD, DERIV, K>(t, g, prm, dxd, dyd, dzd, r2d, rd, ird, sink);           \
        else                                                                                         \
            ao_group<T, RAD, DERIV, K>(t, g, prm, dx, dy, dz, r2, r, ir, sink);                      \
        break;
            QMC_AO_CASE(1) QMC_AO_CASE(2) QMC_AO_CASE(3) QMC_AO_CASE(4) QMC_AO_CASE(5) QMC_AO_CASE(6)
            QMC_AO_CASE(7) QMC_AO_CASE(8) QMC_AO_CASE(9) QMC_AO_CASE(10) QMC_AO_CASE(11) QMC_AO_CASE(12)
#undef QMC_AO_CASE
            default: break;
        }
    }
}

// AO evaluation policies of the kernels: AoGeneric walks the group tables
// above; a generated policy (qmc_spec_*.hip) walks one basis layout with every
// group, entry, primitive offset and Cartesian power a compile-time constant,
// built from the same atom_geo / group_prims / ao_entry pieces in the same
// order, so both give the same values.
struct AoGeneric {
    static constexpr int kEntries = 0;    // AO entries of the row: known at run time only
    static constexpr int kRowBlocks = 4;  // 16-lane row blocks of the MFMA contraction: all
    template <typename T, int RAD, bool DERIV, bool TIGHT, typename Sink>
    static __device__ __forceinline__ void row(const Tables& t, const QmcParams& prm, double ex, double ey, double ez,
                                               Sink&& sink) {
        ao_row<T, RAD, DERIV, TIGHT>(t, prm, ex, ey, ez, sink);
    }
};

template <typename T>
struct AtomGeo {
    T dx, dy, dz, r2, r, ir;
};

template <typename T, int RAD>
__device__ __forceinline__ AtomGeo<T> atom_geo(double dxd, double dyd, double dzd) {
    AtomGeo<T> g;
    g.dx = (T)dxd;
    g.dy = (T)dyd;
    g.dz = (T)dzd;
    g.r2 = g.dx * g.dx + g.dy * g.dy + g.dz * g.dz;
    g.r = g.ir = T(0);
    if (RAD == QMC_RADIAL_STO) {
        g.r = fsqrt(g.r2);
        g.ir = T(1) / g.r;
    }
    return g;
}

// ------------------------------------------------------- lane-row elimination
// Gauss-Jordan elimination with partial pivoting on the n x n matrix whose
// row r lives in lane gbase + r (r < n), in FP64.  nmax (wave-uniform) is the
// largest n over the groups of the wave; lanes whose group is smaller simply
// idle through the extra columns.  On exit every lane of the group holds
// log|det| and sign(det); with WITH_B the augmented block b (n columns) is
// reduced to A^-1 B and the lane that pivoted column k returns in trace_part
// its entry (A^-1 B)_kk, otherwise 0.  Pivot choice: largest |a_rk| among the
// rows not used yet, first one on ties (LAPACK getrf's rule).
template <int NM, bool WITH_B>
__device__ __forceinline__ void lane_gauss_jordan(double (&a)[NM], double (&b)[NM], int n, int nmax, int myrow,
                                                  int gbase, uint64_t gmask, double& logdet, double& sgn,
                                                  double& trace_part, int& my_col) {
    bool done = false;
    my_col = -1;
    logdet = 0.0;
    sgn = 1.0;
    int inversions = 0;
    bool singular = false;
    double mant = 1.0;  // |det| = mant * 2^expo, renormalised after every pivot
    int expo = 0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        if (k < nmax) {
            const bool colk = k < n;
            const double cand = (!done && colk && myrow < n) ? fabs(a[k]) : -1.0;
            // all candidates of the group are fetched first (independent
            // shuffles, one wait), then scanned in row order
            double c[NM];
#pragma unroll
            for (int s = 0; s < NM; ++s) c[s] = __shfl(cand, gbase + (s < n ? s : 0));
            double best = -1.0;
            int br = 0;
#pragma unroll
            for (int s = 0; s < NM; ++s) {
                if (s < n && c[s] > best) {
                    best = c[s];
                    br = s;
                }
            }
            const int plane = gbase + br;
            // rows not yet used with a smaller index than the pivot row
            const uint64_t bal = __ballot(!done && myrow < br);
            if (colk) inversions += __popcll(bal & gmask);
            const double piv = __shfl(a[k], plane);
            const bool is_piv = colk && (myrow == br);
            if (colk) {
                singular = singular || (piv == 0.0);
                int ex;
                mant = frexp(mant * fabs(piv), &ex);
                expo += ex;
                sgn *= (piv < 0.0) ? -1.0 : 1.0;
            }
            const double ip = (piv != 0.0) ? 1.0 / piv : 0.0;
            const double f = (colk && !is_piv && (WITH_B || !done)) ? a[k] * ip : 0.0;
            // pivot row broadcast: columns past nmax hold zeros and stay zero,
            // so every column is updated without a per-column branch
            double pa[NM];
#pragma unroll
            for (int j = k + 1; j < NM; ++j) pa[j] = __shfl(a[j], plane);
#pragma unroll
            // (rows already used are never read again by the determinant-only
            // elimination, so only the eliminated rows are written there)
            for (int j = k + 1; j < NM; ++j) a[j] = (WITH_B && is_piv) ? pa[j] * ip : a[j] - f * pa[j];
            if (WITH_B) {
                double pb[NM];
#pragma unroll
                for (int j = 0; j < NM; ++j) pb[j] = __shfl(b[j], plane);
#pragma unroll
                for (int j = 0; j < NM; ++j) b[j] = is_piv ? pb[j] * ip : b[j] - f * pb[j];
            }
            if (is_piv) {
                done = true;
                my_col = k;
            }
        }
    }
    if (inversions & 1) sgn = -sgn;
    logdet = singular ? -INFINITY : log_mantissa(mant) + 0.6931471805599453 * (double)expo;
    trace_part = 0.0;
    if (WITH_B) {
#pragma unroll
        for (int j = 0; j < NM; ++j)
            if (j == my_col) trace_part = b[j];
    }
}

// ------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11.  Round constants and Weyl key increments from that paper.
struct u4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32 x 32 -> 64-bit product per multiplier (v_mad_u64_u32) instead
        // of separate high and low multiplies
        const uint64_t p0 = (uint64_t)M0 * (uint64_t)c.x, p1 = (uint64_t)M1 * (uint64_t)c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// (0,1] and [0,1) uniforms from 32 random bits
__device__ __forceinline__ double u01_open0(uint32_t v) { return ((double)v + 1.0) * 2.3283064365386963e-10; }
__device__ __forceinline__ double u01_closed0(uint32_t v) { return (double)v * 2.3283064365386963e-10; }

// Three standard normals from one Philox block, Box-Muller in FP64:
// (x, y) -> sqrt(-2 ln u) (cos 2 pi v, sin 2 pi v), (z, w) -> sqrt(-2 ln u) cos 2 pi v.
__device__ __forceinline__ void gauss3_f64(const u4 r, double& a, double& b, double& c) {
    const double r1 = sqrt(-2.0 * log(u01_open0(r.x)));
    const double r2 = sqrt(-2.0 * log(u01_open0(r.z)));
    double s1, c1;
    sincospi(2.0 * u01_closed0(r.y), &s1, &c1);
    a = r1 * c1;
    b = r1 * s1;
    c = r2 * cospi(2.0 * u01_closed0(r.w));
}

// ------------------------------------------------------------ wave-local LDS
// Orders one wavefront's LDS accesses across this point: waits for the
// outstanding ones and keeps the compiler from moving accesses over it.  The
// LDS serves a wavefront's requests in order, so stores issued before the call
// are seen by loads of any lane of the same wavefront issued after it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

constexpr int kBlock = 256;  // threads per block of the wave-per-walker-group kernels
constexpr int kBlockWaves = kBlock / 64;

// Pivot row (group-relative) of column k for the eliminations of the large
// buckets: the candidate with the largest |v| (the first one on ties), by an
// LDS atomic maximum into a slot of the group instead of ceil(log2 n)
// dependent shuffle rounds.  Each lane's key is the high word of |v|'s FP64 bit
// pattern (exponent and top 20 mantissa bits, monotonic in |v|) with its five
// low bits replaced by 31 - row; non-candidates do not take part.  Magnitudes
// that differ beyond their top 15 mantissa bits are ordered exactly; closer
// ones count as tied and go to the first row (partial pivoting only needs a
// large pivot, not the exact maximum).  No candidate (a column past this
// group's n): row 0.
// The slot of a group is its first lane's index; columns alternate between two
// slot sets, and row 0 of each group clears the other set's slot for column
// k + 1 while column k's maxima are formed.  The LDS runs one wavefront's
// instructions in order, so that clear follows the reads of column k - 1 and
// precedes the maxima of column k + 1.  Lanes outside every walker (real =
// false, whose group may share its first lane index with a real group) only
// read.  seg_pivot_init clears set 0 before column 0.  The wait after the
// maxima also completes the caller's LDS stores issued before the call.
__device__ __forceinline__ uint32_t* seg_pivot_slots() {
    __shared__ uint32_t slots[kBlockWaves * 128];
    return slots + (threadIdx.x >> 6) * 128;
}

// Reciprocal of each group's published pivot (lane_det_lds), at the group's
// first lane index.
__device__ __forceinline__ double* det_recips() {
    __shared__ double r[kBlockWaves * 64];
    return r + (threadIdx.x >> 6) * 64;
}

__device__ __forceinline__ void seg_pivot_init(int myrow, int gbase, bool real) {
    uint32_t* S = seg_pivot_slots();
    wave_lds_sync();  // earlier eliminations' reads of the slots are done
    if (real && myrow == 0) S[gbase] = 0u;
}

// seg_pivot_post issues column k's maximum (and the next column's clear);
// seg_pivot_get waits for it, and for every LDS store the caller issued in
// between, and returns the pivot row.  seg_pivot does both.
template <int NM>
__device__ __forceinline__ void seg_pivot_post(int k, bool cand, double v, int myrow, int gbase, bool real) {
    static_assert(NM <= 32, "five bits of row index");
    uint32_t* S = seg_pivot_slots();
    const uint32_t hi = (uint32_t)((uint64_t)__double_as_longlong(fabs(v)) >> 32);
    const uint32_t key = (hi & ~31u) | (uint32_t)(31 - myrow);
    if (real && cand)
        __hip_atomic_fetch_max(S + (k & 1) * 64 + gbase, key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (real && myrow == 0) S[((k + 1) & 1) * 64 + gbase] = 0u;
}

__device__ __forceinline__ int seg_pivot_get(int k, int gbase) {
    wave_lds_sync();
    const uint32_t m = seg_pivot_slots()[(k & 1) * 64 + gbase];
    return m ? 31 - (int)(m & 31u) : 0;
}

template <int NM>
__device__ __forceinline__ int seg_pivot(int k, bool cand, double v, int myrow, int gbase, bool real) {
    seg_pivot_post<NM>(k, cand, v, myrow, gbase, real);
    return seg_pivot_get(k, gbase);
}

// ------------------------------------------------- determinant-only elimination
// log|det| and sign of the n x n matrix whose row r lives in lane gbase + r,
// with the row arithmetic of lane_gauss_jordan<NM, false> but fewer cross-lane
// moves per column:
//  * the pivot (largest |a_rk| among the rows not used yet, first one on ties,
//    compared through seg_pivot's 32-bit keys): up to 8 columns from the n
//    keys gathered into every lane by shuffles; for the larger buckets by
//    seg_pivot's LDS atomic maximum;
//  * the pivot row through LDS, read back as broadcast loads (two columns per
//    load) instead of one shuffle pair per column, with the pivot itself and
//    its reciprocal (det_recips()), which every candidate forms while the
//    pivot is searched (the division leaves the column's critical path): for
//    the larger buckets the pivot lane stores its remaining columns at
//    rowbuf[0..NM) once the pivot is known; up to 8 columns every candidate
//    row r stores them at rowbuf[r NM ..] during the search (n NM doubles).
// Stores happen only when may_store.
// Rows already used are never read again (f = 0 for them), so only the rows
// still to be eliminated are written.
template <int NM>
__device__ __forceinline__ void lane_det_lds(double (&a)[NM], int n, int nmax, int myrow, int gbase, int spin,
                                             uint64_t gmask, bool may_store, double* rowbuf, double& logdet,
                                             double& sgn) {
    if constexpr (NM > 8) seg_pivot_init(myrow, gbase, may_store);
    double* const recip = det_recips() + gbase;  // the group's pivot reciprocal
    bool done = false;
    sgn = 1.0;
    int inversions = 0;
    bool singular = false;
    double mant = 1.0;  // |det| = mant * 2^expo, renormalised after every pivot
    int expo = 0;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        if (k < nmax) {
            const bool colk = k < n;
            if constexpr (NM > 8) {
                // the reciprocal of every candidate's pivot is formed while the
                // search runs; the pivot lane then publishes its remaining
                // columns and that reciprocal (rowbuf[0..NM) and the group's
                // det_recips() entry), which the group reads back at once
                const bool cand = !done && colk && myrow < n;
                seg_pivot_post<NM>(k, cand, a[k], myrow, gbase, may_store);
                double rk = (a[k] != 0.0) ? 1.0 / a[k] : 0.0;
                asm volatile("" : "+v"(rk));  // formed here, not sunk into the pivot lane's store
                const int br = seg_pivot_get(k, gbase);
                const uint64_t bal = __ballot(!done && myrow < br);
                if (colk) inversions += __popcll(bal & gmask);
                const bool is_piv = colk && (myrow == br);
                wave_lds_sync();  // the previous column's reads are done
                if (is_piv && may_store) {
#pragma unroll
                    for (int j = k; j < NM; ++j) rowbuf[j] = a[j];
                    *recip = rk;
                }
                wave_lds_sync();
                // past this group's n nobody stores and the row is stale: colk
                // is false there, so neither the determinant nor a uses it
                const double piv = rowbuf[k];
                const double ip = *recip;
                {
                    const double pv = colk ? piv : 1.0;
                    singular = singular || (pv == 0.0);
                    int ex;
                    mant = frexp(mant * fabs(pv), &ex);
                    expo += ex;
                    sgn *= (pv < 0.0) ? -1.0 : 1.0;
                }
                const double f = (colk && !is_piv && !done) ? a[k] * ip : 0.0;
#pragma unroll
                for (int j = k + 1; j < NM; ++j) a[j] = a[j] - f * rowbuf[j];
                if (is_piv) done = true;
            } else {
                // keys as in seg_pivot, gathered from the group's n lanes (reads
                // past the last row repeat it, which leaves the maximum alone);
                // meanwhile every candidate row r publishes its remaining
                // columns at rowbuf[r NM + k ..] and its pivot's reciprocal at
                // recip[r], and the group then reads the chosen row back (slots
                // of rows already used are not written again, so rows read in
                // earlier columns stay intact)
                const bool cand = !done && colk && myrow < n;
                const uint32_t hi = (uint32_t)((uint64_t)__double_as_longlong(fabs(a[k])) >> 32);
                const uint32_t key = cand ? ((hi & ~31u) | (uint32_t)(31 - myrow)) : 0u;
                if (cand && may_store) {
                    double* const slot = rowbuf + myrow * NM;
#pragma unroll
                    for (int j = k; j < NM; ++j) slot[j] = a[j];
                    recip[myrow] = (a[k] != 0.0) ? 1.0 / a[k] : 0.0;
                }
                uint32_t m = 0;
#pragma unroll
                for (int s = 0; s < NM; ++s) m = max(m, (uint32_t)__shfl((int)key, gbase + (s < n ? s : n - 1)));
                const int br = m ? 31 - (int)(m & 31u) : 0;
                const uint64_t bal = __ballot(!done && myrow < br);
                if (colk) inversions += __popcll(bal & gmask);
                const bool is_piv = colk && (myrow == br);
                wave_lds_sync();  // the candidates' stores are done
                const double* const prow = rowbuf + br * NM;
                const double piv = prow[k];
                const double ip = recip[br];
                {
                    const double pv = colk ? piv : 1.0;
                    singular = singular || (pv == 0.0);
                    int ex;
                    mant = frexp(mant * fabs(pv), &ex);
                    expo += ex;
                    sgn *= (pv < 0.0) ? -1.0 : 1.0;
                }
                const double f = (colk && !is_piv && !done) ? a[k] * ip : 0.0;
#pragma unroll
                for (int j = k + 1; j < NM; ++j) a[j] = a[j] - f * prow[j];
                if (is_piv) done = true;
            }
        }
    }
    if (inversions & 1) sgn = -sgn;
    logdet = singular ? -INFINITY : log_mantissa(mant) + 0.6931471805599453 * (double)expo;
}

// ------------------------------------------------ augmented elimination (LDS)
// lane_gauss_jordan<NM, true> with the pivot row passed through LDS and, above
// 8 columns, the pivot found by seg_pivot: same arithmetic, same outputs up to
// the choice between pivots equal to 15 mantissa bits, for matrices a whose
// columns n..NM-1 are zero on entry (gathered configurations).
// rowbuf is the group's n * NM doubles (the lane_det_lds buffer).  Up to
// NM = 16 the pivot lane stores its b[0..NM) at rowbuf[0..NM) and its a[j],
// k < j < n, at rowbuf[NM - 1 + j] (at most NM + n - 1 doubles, so it fits
// for every n <= NM); readers take the a-part only for j < n (the columns
// past n stay zero, as in the shuffle version, where the pivot row's entries
// there are zero).  For the larger buckets, where unmasked loads of whole b
// rows raise the register pressure, b must also be zero past n: the row is
// b[0..n) at rowbuf[0..n) and a[j] at rowbuf[n - 1 + j], both read for j < n
// only (at most 2n - 1 doubles).  The loads of the masked columns may read up
// to 2 NM - 2 doubles past rowbuf, which the buffer's tail padding keeps in
// bounds.
template <int NM>
__device__ __forceinline__ void lane_gj_lds(double (&a)[NM], double (&b)[NM], int n, int nmax, int myrow, int gbase,
                                            int spin, uint64_t gmask, bool may_store, double* rowbuf, double& logdet,
                                            double& sgn, double& trace_part, int& my_col) {
    if constexpr (NM > 8) seg_pivot_init(myrow, gbase, may_store);
    bool done = false;
    my_col = -1;
    sgn = 1.0;
    int inversions = 0;
    bool singular = false;
    double mant = 1.0;  // |det| = mant * 2^expo, renormalised after every pivot
    int expo = 0;
    constexpr bool kFullB = NM <= 16;
    double* const ra = rowbuf + ((kFullB ? NM : n) - 1);  // a-part: ra[j] for k < j < n
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        if (k < nmax) {
            const bool colk = k < n;
            int br;
            if constexpr (NM <= 8) {
                const double cand = (!done && colk && myrow < n) ? fabs(a[k]) : -1.0;
                double c[NM];
#pragma unroll
                for (int s = 0; s < NM; ++s) c[s] = __shfl(cand, gbase + (s < n ? s : 0));
                double bb = -1.0;
                br = 0;
#pragma unroll
                for (int s = 0; s < NM; ++s) {
                    if (s < n && c[s] > bb) {
                        bb = c[s];
                        br = s;
                    }
                }
            } else {
                br = seg_pivot<NM>(k, !done && colk && myrow < n, a[k], myrow, gbase, may_store);
            }
            const int plane = gbase + br;
            const uint64_t bal = __ballot(!done && myrow < br);
            if (colk) inversions += __popcll(bal & gmask);
            const double piv = __shfl(a[k], plane);
            const bool is_piv = colk && (myrow == br);
            if (colk) {
                singular = singular || (piv == 0.0);
                int ex;
                mant = frexp(mant * fabs(piv), &ex);
                expo += ex;
                sgn *= (piv < 0.0) ? -1.0 : 1.0;
            }
            const double ip = (piv != 0.0) ? 1.0 / piv : 0.0;
            const double f = (colk && !is_piv) ? a[k] * ip : 0.0;
            wave_lds_sync();  // the previous column's reads are done
            if (is_piv && may_store) {
#pragma unroll
                for (int j = 0; j < NM; ++j)
                    if (kFullB || j < n) rowbuf[j] = b[j];
#pragma unroll
                for (int j = k + 1; j < NM; ++j)
                    if (j < n) ra[j] = a[j];
            }
            wave_lds_sync();
#pragma unroll
            for (int j = k + 1; j < NM; ++j) {
                const double pa = (j < n) ? ra[j] : 0.0;
                a[j] = is_piv ? a[j] * ip : a[j] - f * pa;
            }
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                const double pb = (kFullB || j < n) ? rowbuf[j] : 0.0;
                b[j] = is_piv ? b[j] * ip : b[j] - f * pb;
            }
            if (is_piv) {
                done = true;
                my_col = k;
            }
        }
    }
    if (inversions & 1) sgn = -sgn;
    logdet = singular ? -INFINITY : log_mantissa(mant) + 0.6931471805599453 * (double)expo;
    trace_part = 0.0;
#pragma unroll
    for (int j = 0; j < NM; ++j)
        if (j == my_col) trace_part = b[j];
}

}  // namespace qmc
