/*
 * qmc_mi355x.h -- C ABI of libqmcmi355x.so, the MI355X (gfx950) kernels of the
 * QMCTorch VMC hot path.
 *
 * The reference (QMCTorch 0.4.0) is pure Python; its "operator API" for this
 * path is the Python module surface listed next to each entry point.  The
 * Python package qmctorch_amd keeps that surface and calls these functions
 * through ctypes; every pointer argument is a device pointer (torch caching
 * allocator), sizes are element counts, every function returns 0 on success
 * or a nonzero status whose text is available from qmc_last_error().  No entry
 * point synchronises the stream it is given.
 */
#ifndef QMC_MI355X_H
#define QMC_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* qmc_stream_t; /* hipStream_t of the caller (0 = default stream) */

enum { QMC_OK = 0, QMC_ERR_ARG = 1, QMC_ERR_HIP = 2, QMC_ERR_UNSUPPORTED = 3 };

enum { QMC_RADIAL_GTO_PURE = 0, QMC_RADIAL_STO = 1 };

/* MIXED: psi, pdf, sweeps and drift with FP32 AO/MO/Jastrow and FP64 Slater
 * elimination, acceptance and walker state.  Local energies (QMC_EVAL_ELOC) are
 * evaluated in FP64 whatever the precision argument says. */
enum { QMC_PREC_FP64 = 0, QMC_PREC_MIXED = 1 };

/* Maximum number of primitives sharing one exponent group, and of unique
 * configurations per spin handled by the kernels. */
#define QMC_MAXP 12
#define QMC_MAXCONF 8
/* Plans with more unique configurations per spin (up to QMC_MAXEXC) must be
 * single and double excitations of the ground state (columns 0..n-1); psi, pdf,
 * the local energy, the Metropolis sweep, the drift, HMC, the adjoints
 * (qmc_logpsi_grad, qmc_logpsi_grad_noalloc, qmc_eloc_grad) and the batched
 * launches then use the excitation (trace-trick) kernels; the row kernels
 * return QMC_ERR_UNSUPPORTED. */
#define QMC_MAXEXC 1024

/*
 * Static description of one Slater-Jastrow wavefunction (host pointers; the
 * plan copies them to the device).  Built by qmctorch_amd.wavefunction.plan
 * from the same basis namespace the reference consumes
 * (qmctorch/scf/calculator/pyscf.py:104-252, molecule.py:305-350):
 *
 *  exponent groups  -- primitives of one atom that share an exponent list;
 *                      group g has grp_nprim[g] exponents taken from bas_exp at
 *                      indices grp_prim_src[grp_prim0[g] + p] and grp_nao[g]
 *                      AO entries starting at grp_ao0[g];
 *  AO entries       -- gao_mu (AO index), gao_k (kx | ky<<4 | kz<<8) and
 *                      QMC_MAXP coefficients gao_K (norm_cst * bas_coeffs of
 *                      atomic_orbitals.py:236-249,654-669) and r powers gao_n
 *                      (STO only);
 *  configurations   -- unique spin-up/down orbital lists and the CI map of
 *                      slater_pooling.py:96-111 / orbital_projector.py:29-94.
 */
typedef struct QmcSystemDesc {
    int32_t nelec, nup, ndown, natom;
    int32_t nao;
    int32_t nprim_ref;   /* length of bas_exp (reference primitive count Np) */
    int32_t radial_type; /* QMC_RADIAL_* */
    int32_t ngroup;
    const int32_t* grp_atom;     /* [ngroup] */
    const int32_t* grp_nprim;    /* [ngroup] <= QMC_MAXP */
    const int32_t* grp_prim0;    /* [ngroup] offset into grp_prim_src */
    const int32_t* grp_prim_src; /* [sum nprim] index into bas_exp */
    const int32_t* grp_ao0;      /* [ngroup] offset into AO entries */
    const int32_t* grp_nao;      /* [ngroup] */
    const int32_t* grp_tight;    /* [ngroup] 1: evaluated in FP64 by the mixed local energy, or NULL */
    int32_t nentry;              /* number of AO entries */
    const int32_t* gao_mu;       /* [nentry] */
    const int32_t* gao_k;        /* [nentry] */
    const double* gao_K;         /* [nentry * QMC_MAXP] */
    const double* gao_n;         /* [nentry * QMC_MAXP] (STO r powers, else 0) */
    const int32_t* gao_src;      /* [nentry * QMC_MAXP] index into bas_exp of each term */
    const double* atom_z;        /* [natom] nuclear charges */
    const double* jastrow_atoms; /* [natom*3] e-n Jastrow centres (mol.atom_coords) */
    int32_t nmo;                 /* MO columns used (highest occupied index + 1) */
    int32_t nci;
    int32_t nconf_up, nconf_dn;  /* unique configurations per spin (<= QMC_MAXEXC) */
    const int32_t* conf_up;      /* [nconf_up * nup] MO columns */
    const int32_t* conf_dn;      /* [nconf_dn * ndown] */
    const int32_t* ci_up;        /* [nci] unique up configuration of CI term n */
    const int32_t* ci_dn;        /* [nci] */
    int32_t use_jastrow_ee, use_jastrow_en;
    double vnn;                  /* nuclear repulsion (wf_base.py:97-116) */
} QmcSystemDesc;

typedef struct QmcPlan QmcPlan;

/*
 * Trainable parameters, read on every call (device pointers, float64), so
 * optimizer updates made in place are seen without re-planning:
 *   bas_exp     ao.bas_exp            [nprim_ref]   (atomic_orbitals.py:53-55)
 *   atom_coords ao.atom_coords        [natom*3]     (atomic_orbitals.py:33-37)
 *   mo_w        mo_scf * mo_modifier  [nao*nmo]     row-major, first nmo columns
 *               (molecular_orbitals.py:79-95); float32 for MIXED psi and sweeps,
 *               float64 otherwise (FP64 runs and every local-energy evaluation)
 *   ci          fc.weight             [nci]         (slater_jastrow.py:179-191)
 *   w_ee, w_en  Pade weights          [1] or NULL   (pade_jastrow_kernel.py:28)
 * and the same values gathered into the order the kernels read them, so that
 * no table lookup sits between a kernel and a parameter value:
 *   grp_exp     bas_exp[grp_prim_src[j]]  [sum of grp_nprim]  float64
 *   grp_expf    the same in float32
 *   entry_w     mo_w row of AO entry i (mo_w[gao_mu[i], :]) padded with zeros to
 *               B = qmc_plan_bucket(plan) columns  [nentry*B], type of mo_w
 */
typedef struct QmcParams {
    const double* bas_exp;
    const double* atom_coords;
    const void* mo_w;
    const double* ci;
    const double* w_ee;
    const double* w_en;
    const double* grp_exp;
    const float* grp_expf;
    const void* entry_w;
} QmcParams;

/* Outputs of qmc_wf_eval; any pointer may be NULL.  All float64, [nw]. */
typedef struct QmcWfOut {
    double* psi;     /* J * fc(dets)                 slater_jastrow.py:243-286 */
    double* logabs;  /* log|psi| */
    double* sign;    /* sign(psi) */
    double* kinetic; /* kinetic_energy_jacobi        slater_jastrow.py:312-344 */
    double* eloc;    /* local_energy                 wf_base.py:184-215 */
    double* epot;    /* V_en + V_ee + V_nn           wf_base.py:49-116 */
    double* v_ee;    /* sum_{i<j} 1/r_ij             WaveFunction.electronic_potential, wf_base.py:49-70 */
    double* v_en;    /* -sum_{i,A} Z_A/r_iA          WaveFunction.nuclear_potential, wf_base.py:72-95 */
} QmcWfOut;

enum { QMC_EVAL_PSI = 0, QMC_EVAL_ELOC = 1 };

/* Walker state of a Metropolis chain (device pointers). */
typedef struct QmcWalkers {
    double* pos;      /* [nw * 3*nelec], updated in place */
    double* logabs;   /* [nw] log|psi| of pos, updated in place */
    double* sign;     /* [nw] */
    int32_t* naccept; /* [nw] accepted moves, incremented */
    const int64_t* walker_id; /* [nw] global walker ids keying the RNG, or NULL (= index) */
} QmcWalkers;

/*
 * Metropolis sweep (sampler/metropolis.py:85-298): all-electron Gaussian moves
 * with covariance sigma2 * I per electron (metropolis.py:206-212: sigma2 =
 * step/(2 sqrt(2 ln 2))), acceptance min(1, psi'^2/psi^2) >= U.  Steps
 * step0 .. step0+nstep-1 of the chain are performed; random numbers come from
 * Philox4x32-10 keyed by seed and counted by (walker id, step, electron), one
 * block per electron and step (the acceptance uniform is made of the bits of
 * electron 0's block that its proposal leaves unused), so a walker's
 * trajectory does not depend on how walkers are batched.  Positions
 * of steps s with s >= keep_from and (s - keep_from) % keep_every == 0 are
 * written to keep[(s - keep_from)/keep_every][walker] when keep != NULL.
 */
typedef struct QmcSweepDesc {
    int64_t step0;
    int32_t nstep;
    int32_t init_logpsi; /* 1: evaluate logabs/sign of pos first */
    double sigma2;
    uint64_t seed;
    int64_t keep_from;
    int32_t keep_every;
    int32_t uniform_move; /* 1: uniform cube proposal step*(2U-1) (metropolis.py:263-269) */
    double step_size;
    double* keep;        /* [nkeep * nw * 3*nelec] or NULL */
    int32_t one_electron; /* 1: each step moves one electron per walker, chosen uniformly
                             (move type "one-elec", metropolis.py:228-250); 0: all electrons */
    int32_t reserved;
} QmcSweepDesc;

/* Plan lifetime.  Replaces SlaterJastrow.__init__ (slater_jastrow.py:31-117)
 * bookkeeping: device copies of the basis, configuration and Jastrow tables. */
int qmc_plan_create(const QmcSystemDesc* desc, QmcPlan** plan);
int qmc_plan_destroy(QmcPlan* plan);

/* Kernels generated for one basis layout (a row with every group, entry and
 * Cartesian power a compile-time constant) are built into the library for the
 * benchmark molecules; a plan whose layout and register bucket match one uses
 * them for the mixed-precision sweep and the FP64 local energy.  No
 * counterpart in the reference (the layout there is a set of index tensors).
 * qmc_plan_set_specialized(plan, 0) makes the plan use the table-driven
 * kernels only (for comparisons); qmc_plan_specialization returns the
 * layout's name, or NULL when the plan uses the table-driven kernels. */
int qmc_plan_set_specialized(QmcPlan* plan, int32_t enable);

/* Register bucket of a plan: the column count of its kernels' MO rows
 * (>= nmo, nup, ndown), i.e. the row stride of QmcParams.entry_w. */
int qmc_plan_bucket(const QmcPlan* plan);
const char* qmc_plan_specialization(const QmcPlan* plan);

/* psi / log|psi| / kinetic / local energy of nw walkers.
 * Replaces SlaterJastrow.forward (slater_jastrow.py:243-286), WaveFunction.pdf
 * (wf_base.py:241-247), SlaterJastrow.kinetic_energy_jacobi
 * (slater_jastrow.py:312-344) and WaveFunction.local_energy (wf_base.py:184-215). */
int qmc_wf_eval(const QmcPlan* plan, const QmcParams* params, int32_t precision, int32_t mode,
                const double* pos, int64_t nw, QmcWfOut* out, qmc_stream_t stream);

/* Atomic orbitals and their derivatives at nw walkers, reference layouts:
 * ao [nw][nelec][nao]; dao [nw][nelec][nao][3]; d2ao [nw][nelec][nao] (float64).
 * Replaces AtomicOrbitals.forward(pos, derivative=[0] or [0,1,2], sum_grad=False)
 * (atomic_orbitals.py:131-218, 578-609).  dao/d2ao may be NULL. */
int qmc_ao_eval(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                double* ao, double* dao, double* d2ao, qmc_stream_t stream);

/* Metropolis sweeps.  Replaces the step loop of Metropolis.__call__
 * (metropolis.py:134-165) for move={'type':'all-elec'}. */
int qmc_metropolis_sweep(const QmcPlan* plan, const QmcParams* params, int32_t precision,
                         QmcWalkers* walkers, int64_t nw, const QmcSweepDesc* sweep, qmc_stream_t stream);

/*
 * Hamiltonian Monte Carlo (sampler/hamiltonian.py:80-201) on U = -ln psi^2:
 * per step, momenta p ~ N(0, 1) per coordinate, E0 = U(q) + |p|^2/2, a
 * leapfrog trajectory of traj_length position updates of step_size
 * (half momentum steps at both ends), E1 = U(q') + |p'|^2/2, and the move is
 * kept iff exp(E0 - E1) >= u, u ~ U[0,1) (a NaN energy rejects the move).
 * grad U = -2 grad ln|psi| is evaluated on the device (no autograd).  Random
 * numbers: Philox4x32-10 keyed by seed, counter (walker id, step, electron |
 * 2^30) for the momenta (Box-Muller in FP64), (walker id, step, 2^31) for u.
 * log|psi| of the walkers is evaluated at the start of every call (the stored
 * walker logabs is not read); logabs/sign are written back.  Positions of the
 * kept steps (keep_from / keep_every as in QmcSweepDesc) go to keep.
 */
typedef struct QmcHmcDesc {
    int64_t step0;
    int32_t nstep;
    int32_t traj_length; /* L >= 1 */
    double step_size;    /* epsilon */
    uint64_t seed;
    int64_t keep_from;
    int32_t keep_every;
    int32_t reserved;
    double* keep; /* [nkeep * nw * 3*nelec] or NULL */
} QmcHmcDesc;

int qmc_hmc_sweep(const QmcPlan* plan, const QmcParams* params, int32_t precision, QmcWalkers* walkers, int64_t nw,
                  const QmcHmcDesc* hmc, qmc_stream_t stream);

/* grad_i ln|psi| of every electron, drift [nw][nelec][3], with log|psi| and
 * sign (either may be NULL), in the precision of the sampling kernels.
 * Replaces SlaterJastrow.gradients_jacobi (slater_jastrow.py:346-447), which
 * returns psi * drift (pdf=False) or 2 psi^2 * drift (pdf=True). */
int qmc_wf_drift(const QmcPlan* plan, const QmcParams* params, int32_t precision, const double* pos, int64_t nw,
                 double* drift, double* logabs, double* sign, qmc_stream_t stream);

/* Adjoint of log|psi| with respect to the parameters, as per-walker rows
 * (AO values are recomputed here, not stored by the forward pass):
 *   ao_rows    [nw*nelec][nao]  AO values of every electron (reference AO layout)
 *   gamma_rows [nw*nelec][nmo]  d log|psi| / d MO[i,m] = sum_c omega_c (A_c^-1)[k(m), i]
 *   dci        [nw][nci]        d log|psi| / d c_n = D_n / sum_k c_k D_k
 *   djas       [nw][2]          d log|psi| / d (w_ee, w_en)
 *   dexp_rows  [nw*nelec][nprim_ref] d log|psi| / d bas_exp per electron, or NULL
 *              (norm_cst is a constant of the AO layer, atomic_orbitals.py:90-94)
 * so that for walker weights g: dL/dW = ao_rows^T (g * gamma_rows) (one GEMM),
 * dL/dc = g^T dci, dL/dw = g^T djas, dL/dexp = g_rows^T dexp_rows.
 * FP64 throughout; mo_w must be float64.
 * Replaces autograd through SlaterJastrow.forward in Solver.evaluate_grad_manual
 * (solver.py:372-431). */
int qmc_logpsi_grad_rows(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                         double* ao_rows, double* gamma_rows, double* dci, double* djas, double* dexp_rows,
                         qmc_stream_t stream);

/* Adjoint of the local energy with respect to the parameters, as per-walker
 * rows (FP64; AO values and derivatives are recomputed):
 *   eloc    [nw]              local energy (or NULL)
 *   ao_rows [nw*nelec][nao]   AO values          } dE_L/dW = ao_rows^T (g*ga_rows)
 *   b_rows  [nw*nelec][nao]   AO-level kinetic   }         + b_rows^T (g*gb_rows)
 *                             rows -1/2(lap AO + 2 grad lnJ . grad AO + (lap J/J) AO)
 *   ga_rows, gb_rows [nw*nelec][nmo]
 *   dci     [nw][nci]         dE_L/dc_n
 *   djas    [nw][2]           dE_L/d(w_ee, w_en)
 *   drift   [nw][nelec][3]    grad_i psi / psi (slater_jastrow.py:346-447) or NULL
 * ao_rows and b_rows are both given or both NULL.  Replaces autograd through
 * WaveFunction.local_energy in Solver.evaluate_grad_auto (solver.py:352-370,
 * loss.py:49-79) and SlaterJastrow.gradients_jacobi. */
int qmc_eloc_grad_rows(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw, double* eloc,
                       double* ao_rows, double* b_rows, double* ga_rows, double* gb_rows, double* dci,
                       double* djas, double* drift, qmc_stream_t stream);

/*
 * Parameter gradients of a weighted sum over walkers with the reduction fused
 * into the adjoint kernels (FP64; mo_w must be float64): for per-walker
 * weights g,
 *   qmc_logpsi_grad  L = sum_w g_w log|psi(x_w)|    (g_w = dL/dpsi_w * psi_w)
 *   qmc_eloc_grad    L = sum_w g_w E_L(x_w)         (eloc [nw] written if non-NULL)
 * and the gradients dL/dW [nao][nmo] (W = mo_scf * mo_modifier, first nmo
 * columns), dL/dc [nci], dL/d(w_ee, w_en) [2] and, for log|psi| only,
 * dL/dbas_exp [nprim_ref] (grad_exp, or NULL to skip).  AO values are
 * recomputed in the kernel; every wavefront adds its walkers' contributions to
 * a partial-sum slot of the workspace and one more launch adds the slots, so
 * no per-walker row reaches memory.  The workspace holds up to 4096 slots of
 * (nentry nmo + nci + 2 [+ 12 nentry]) doubles; qmc_adjoint_workspace_size
 * gives the size that uses the most slots, any size of at least one slot
 * works (fewer slots, each serving more walkers).  The slot count fixes the
 * order of the sums: equal workspace sizes give bit-identical results.
 * Replace autograd through SlaterJastrow.forward in Solver.evaluate_grad_manual
 * (solver.py:372-431) and through WaveFunction.local_energy in
 * Solver.evaluate_grad_auto (solver.py:352-370, loss.py:49-79).
 */
typedef struct QmcGradOut {
    double* grad_w;   /* [nao * nmo] */
    double* grad_ci;  /* [nci] */
    double* grad_jas; /* [2] */
    double* grad_exp; /* [nprim_ref] or NULL */
} QmcGradOut;

size_t qmc_adjoint_workspace_size(const QmcPlan* plan, int64_t nw, int32_t with_exp);
int qmc_logpsi_grad(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                    const double* weights, QmcGradOut* out, void* workspace, size_t workspace_bytes,
                    qmc_stream_t stream);
int qmc_eloc_grad(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                  const double* weights, double* eloc, QmcGradOut* out, void* workspace, size_t workspace_bytes,
                  qmc_stream_t stream);

/*
 * qmc_logpsi_grad without a workspace argument: the partial-sum slots live in
 * a static device array of QMC_NOALLOC_SCRATCH_BYTES per GPU (allocated when
 * the library is loaded, held until the process ends), so the call allocates
 * nothing -- for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  No counterpart in
 * the reference.  Returns QMC_ERR_UNSUPPORTED when one slot of the system
 * exceeds the array.  Every call on one GPU shares that GPU's array: calls
 * that overlap in time on one GPU race (order them on one stream).  Results
 * equal qmc_logpsi_grad's with a workspace of QMC_NOALLOC_SCRATCH_BYTES bit
 * for bit (same slot count, same summation order).
 */
#define QMC_NOALLOC_SCRATCH_BYTES (64u << 20)
int qmc_logpsi_grad_noalloc(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw,
                            const double* weights, QmcGradOut* out, qmc_stream_t stream);

/*
 * Batched launches: many systems (plans), each with its own parameters,
 * walkers, walker count and step count, in one launch per register bucket
 * (occupied orbitals per spin <= 2, 4, 8, 16, 24) and radial type.  Waves of
 * the costliest systems are dispatched first.  A walker's arithmetic does not
 * depend on the other jobs of a batch, so every result equals that of a batch
 * holding its job alone, bit for bit.  No counterpart in the reference, whose
 * driver handles one wavefunction per Solver (solver_base.py:316-387); the
 * callers are request loops that run uneven systems side by side.
 *
 * A QmcBatch owns a ring of job tables (pinned host staging + device copy) for
 * up to max_jobs jobs per call; a call reuses a slot only after the launches
 * that last read it have completed.  One QmcBatch serves one host thread.
 */
typedef struct QmcBatch QmcBatch;

typedef struct QmcSweepJob {
    const QmcPlan* plan;
    QmcParams params;
    QmcWalkers walkers;
    int64_t nw;
    QmcSweepDesc sweep;
} QmcSweepJob;

typedef struct QmcEvalJob {
    const QmcPlan* plan;
    QmcParams params;
    const double* pos;
    int64_t nw;
    QmcWfOut out;
} QmcEvalJob;

int qmc_batch_create(int32_t max_jobs, QmcBatch** batch);
int qmc_batch_destroy(QmcBatch* batch);

/* qmc_metropolis_sweep for every job of the list (jobs with nw == 0 or
 * nstep <= 0 are skipped). */
int qmc_metropolis_sweep_batch(QmcBatch* batch, int32_t precision, const QmcSweepJob* jobs, int32_t njobs,
                               qmc_stream_t stream);

/* qmc_wf_eval for every job of the list. */
int qmc_wf_eval_batch(QmcBatch* batch, int32_t precision, int32_t mode, const QmcEvalJob* jobs, int32_t njobs,
                      qmc_stream_t stream);

/* Philox4x32-10 of the sweep kernel applied to n (counter[4], key[2]) tuples
 * (device pointers: in [n*6] uint32, out [n*4] uint32).  No counterpart in the
 * reference (whose proposals come from torch's host generator); exposed so the
 * device generator can be checked against the published known-answer vectors. */
int qmc_philox4x32_10(const uint32_t* in, uint32_t* out, int64_t n, qmc_stream_t stream);

/*
 * Parameter views gathered on the device in one launch (the layouts of
 * QmcParams): W = mo_scf * mo_modifier over the plan's first nmo columns
 * (molecular_orbitals.py:65-95) in FP64 and FP32, its rows in AO-entry order
 * padded to the register bucket, and bas_exp gathered in group order
 * (atomic_orbitals.py:53-55).  mo_scf and mo_modifier are [nao][ld] float64;
 * any output pointer may be NULL.  Sizes from qmc_plan_sizes.  Replaces the
 * per-call torch ops that rebuilt these views (no reference counterpart: the
 * reference recomputes W inside MolecularOrbitals.forward).
 */
typedef struct QmcParamViews {
    double* mo_w64;    /* [nao * nmo] */
    float* mo_w32;     /* [nao * nmo] */
    double* entry_w64; /* [nentry * bucket] */
    float* entry_w32;  /* [nentry * bucket] */
    double* grp_exp;   /* [nexp_grp] */
    float* grp_expf;   /* [nexp_grp] */
} QmcParamViews;

typedef struct QmcPlanSizes {
    int32_t nexp_grp; /* exponents in group order (length of grp_exp) */
    int32_t nentry;
    int32_t bucket;   /* qmc_plan_bucket */
    int32_t nao;
    int32_t nmo;
} QmcPlanSizes;

int qmc_plan_sizes(const QmcPlan* plan, QmcPlanSizes* out);
int qmc_params_gather(const QmcPlan* plan, const double* mo_scf, const double* mo_modifier, int32_t ld,
                      const double* bas_exp, QmcParamViews* views, qmc_stream_t stream);

/*
 * Pade Jastrow factors and potentials at nw walkers (float64; any output may be
 * NULL).  terms selects the factors: bit 0 the e-e factor, bit 1 the e-n factor
 * (their product when both; the plan must hold the e-n centres for bit 1):
 *   value [nw]            J                              (Nbatch, 1) of forward(pos)
 *   grad  [nw][3][nelec]  dJ/dx_{e,d}                    forward(pos, 1, sum_grad=False)
 *   lap   [nw][nelec]     sum_d d^2 J / dx_{e,d}^2       forward(pos, 2)
 *   v_ee  [nw]            sum_{i<j} 1/r_ij               wf_base.py:49-70
 *   v_en  [nw]            -sum_{i,A} Z_A / r_iA          wf_base.py:72-95
 * Replaces JastrowFactorElectronElectron.forward
 * (jastrow_factor_electron_electron.py:124-260), JastrowFactorElectronNuclei.forward
 * (jastrow_factor_electron_nuclei.py:56-161), CombineJastrow.forward
 * (combine_jastrow.py:33-195) and WaveFunction.electronic_potential /
 * nuclear_potential (wf_base.py:49-95).
 */
typedef struct QmcJastrowOut {
    double* value;
    double* grad;
    double* lap;
    double* v_ee;
    double* v_en;
} QmcJastrowOut;

int qmc_jastrow_eval(const QmcPlan* plan, const QmcParams* params, int32_t terms, const double* pos, int64_t nw,
                     QmcJastrowOut* out, qmc_stream_t stream);

/*
 * Slater determinants and operator traces from caller-given MO matrices
 * mo [nw][nelec][ld] (float64), for every CI term n of the plan's
 * configurations: det [nw][nci] = det A_up det A_dn; op_up / op_dn [nop][nw][nci]
 * = tr(A_s^-1 B_s) (tr((A_s^-1 B_s)^2) with op_squared), A_s and B_s the
 * configuration's columns of mo and bop [nop][nw][nelec][ld] on the rows of
 * spin s.  Any output may be NULL.  Replaces SlaterPooling.forward
 * (slater_pooling.py:64-111) and SlaterPooling.operator (slater_pooling.py:262-409).
 */
int qmc_slater_pool(const QmcPlan* plan, const double* mo, int32_t ld, const double* bop, int32_t nop,
                    int32_t op_squared, int64_t nw, double* det, double* op_up, double* op_dn,
                    qmc_stream_t stream);

/*
 * Atomic orbitals with all second derivatives at nw walkers of nelec electrons
 * (pos [nw][nelec][3]; nelec = 1 for one-electron updates), reference layouts,
 * float64, any output NULL:
 *   ao [nw][nelec][nao], grad [..][nao][3], lap [..][nao],
 *   hess_diag [..][nao][3] (d2/dx2, d2/dy2, d2/dz2), hess_mixed [..][nao][3] (xy, xz, yz).
 * Replaces AtomicOrbitals.forward with derivative 0 / 1 / 2 / 3 / [0,1,2],
 * sum_grad, sum_hess and one_elec (atomic_orbitals.py:131-218, 330-575).
 */
typedef struct QmcAoOut {
    double* ao;
    double* grad;
    double* lap;
    double* hess_diag;
    double* hess_mixed;
} QmcAoOut;

int qmc_ao_eval_full(const QmcPlan* plan, const QmcParams* params, const double* pos, int64_t nw, int32_t nelec,
                     QmcAoOut* out, qmc_stream_t stream);

/*
 * Initial walker positions pos [nw][nelec][3] drawn on the device
 * (Philox4x32-10 keyed by seed, counted by walker walker0 + w, electron and draw),
 * for the domains of Molecule.domain (molecule.py:176-214) that
 * Walkers.initialize accepts (sampler/walkers.py:8-150):
 *   CENTER   uniform in [lo, hi) per coordinate (lo = -hi = -1e-3 by convention)
 *   UNIFORM  uniform in [lo, hi) per coordinate
 *   NORMAL   mean + L z per electron, z ~ N(0, I), L the lower Cholesky factor
 *            of the covariance
 *   ATOMIC   electron e at atom a_e + s N(0, I): the assignment a_e is the list
 *            (atom A repeated atom_nelec[A] times) in a uniformly random order,
 *            and s = atom_width[3A + r], r = 0 for the first electron of its atom
 *            in electron order, 1 for the second to fifth, 2 from the sixth on.
 */
enum { QMC_INIT_CENTER = 0, QMC_INIT_UNIFORM = 1, QMC_INIT_NORMAL = 2, QMC_INIT_ATOMIC = 3 };
#define QMC_INIT_MAXATOM 32

typedef struct QmcWalkerInit {
    int32_t mode;
    int32_t natom;
    uint64_t seed;
    int64_t walker0;
    double lo, hi;
    double mean[3];
    double chol[9];
    double atom_xyz[3 * QMC_INIT_MAXATOM];
    double atom_width[3 * QMC_INIT_MAXATOM];
    int32_t atom_nelec[QMC_INIT_MAXATOM];
} QmcWalkerInit;

int qmc_walkers_init(const QmcWalkerInit* init, int32_t nelec, int64_t nw, double* pos, qmc_stream_t stream);

/* Text of the last error on the calling thread. */
const char* qmc_last_error(void);

/* Library version string. */
const char* qmc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* QMC_MI355X_H */
