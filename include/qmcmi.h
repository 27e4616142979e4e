/* qmcmi: MI355X-native batched walker-move path of real-space VMC.
 *
 * C ABI (plain pointers and sizes, no exceptions across the boundary).  Every
 * entry point returns 0 on success and a nonzero status otherwise; the message
 * is available from qmi_last_error(ctx).  All calls on one context must come
 * from one host thread (as one crowd is driven by one thread in the
 * reference); contexts are independent.
 *
 * Each entry point states the reference interface it replaces (QMCPACK source
 * tree, file:line).  Walker data lives in HBM for the lifetime of the context;
 * host arrays passed in or out are copied.
 */
#ifndef QMCMI_H
#define QMCMI_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct qmi_ctx qmi_ctx;

enum
{
  QMI_OK            = 0,
  QMI_ERR_ARG       = 1, /* invalid argument or call sequence */
  QMI_ERR_DEVICE    = 2, /* HIP runtime / device error */
  QMI_ERR_NUMERIC   = 3, /* NaN ratio/gradient, zero ratio on accept, singular matrix */
  QMI_ERR_UNSUPPORTED = 4
};

/* Precision of the orbital table and of the delayed-update engine (the reference's full and mixed
 * precision builds, CMakeLists.txt:83-111).  In the MIXED modes the determinant state (inverse,
 * orbital gradients/Laplacians, delayed rows) is stored in FP32 and the flush runs on FP32
 * matrix cores; in the TABLE modes only the spline table is narrowed.  The LU re-inversion and the
 * log determinant stay FP64 (DiracMatrix.h:204-220), as do positions, Jastrows and every ratio and
 * gradient dot product. */
enum
{
  QMI_PREC_FP64       = 0, /* FP64 table and determinant engine */
  QMI_PREC_MIXED_FP32 = 1, /* FP32 table, FP32 determinant engine */
  QMI_PREC_MIXED_BF16 = 2, /* BF16 table storage (FP32 and wider arithmetic), FP32 determinant engine */
  QMI_PREC_TABLE_FP32 = 3, /* FP32 table, FP64 determinant engine */
  QMI_PREC_TABLE_BF16 = 4  /* BF16 table storage, FP64 determinant engine */
};

/* System description.  Orthorhombic cell (L[d] <= 0: open along d); ions;
 * two spin groups (electrons [0, n_up) up, [n_up, N) down); one real periodic
 * cubic B-spline orbital table given as its unique periodic coefficients
 * u[Mx][My][Mz][norb] (orbital fastest); J1 (e-ion, one ion species) and J2
 * (uu = dd, ud = du) B-spline functors with NumParams coefficients each.
 * Replaces the objects built by the reference's XML builders:
 * SplineR2R (QMCWaveFunctions/BsplineFactory/SplineR2R.h:35-233),
 * BsplineFunctor (QMCWaveFunctions/Jastrow/BsplineFunctor.h:104-133),
 * DiracDeterminantBatched with delay_rank (Fermion/SlaterDetBuilder.cpp:393-412). */
typedef struct
{
  double L[3];
  int n_up, n_dn;
  int n_ion;
  const double* ions; /* [n_ion][3] */
  int M[3];
  int norb;
  const double* coefs; /* [Mx][My][Mz][norb] */
  int j1_np;
  const double* j1_params;
  double j1_rcut, j1_cusp;
  int j2_np;
  const double* j2_uu;
  const double* j2_ud;
  double j2_rcut, j2_cusp_uu, j2_cusp_ud;
  int delay_rank; /* 1: Sherman-Morrison (Fahy) update; k > 1: rank-k delayed Woodbury */
  /* Complex orbitals (SplineC2C, QMCWaveFunctions/BsplineFactory/SplineC2C.cpp:200-279):
   * psi_j(r) = e^{-i k_j.r} u_j(r); coefs then hold [Mx][My][Mz][norb][2] (real and imaginary
   * parts of u_j interleaved) and kpoints the Cartesian k_j.  With complex orbitals every
   * value-typed array crossing this interface (ratios, gradients, G, L, orbital values,
   * inverse rows) holds (re, im) pairs, and logpsi holds (log|psi|, phase) per walker. */
  int complex_orbitals;
  const double* kpoints; /* [norb][3] */
} qmi_system_desc;

/* ---- context and data ---------------------------------------------------------------------- */
/* device: HIP device ordinal.  Creates one stream; all work of the context runs on it. */
int qmi_ctx_create(qmi_ctx** ctx, int device, int precision);
int qmi_ctx_destroy(qmi_ctx* ctx);
const char* qmi_last_error(const qmi_ctx* ctx);

/* Upload the system: spline table (wrapped to the padded [M+3]^3 layout on the device, as
 * einspline/bspline_create.c:1027-1034), ions, functors.  Replaces SplineR2R construction and
 * SplineSetReader, J1/J2 RadialJastrowBuilder functors. */
int qmi_system_create(qmi_ctx* ctx, const qmi_system_desc* desc);

/* Load nw walkers' electron positions R[nw][N][3] and allocate per-walker state.
 * walkers_per_crowd: walkers sharing one RNG stream and one delay counter
 * (QMCDrivers/Crowd.h:37-135; DelayedUpdateBatched.h:183). */
int qmi_walkers_load(qmi_ctx* ctx, int nw, const double* R, int walkers_per_crowd);
/* Copy positions back: R[nw][N][3]. */
int qmi_walkers_get(qmi_ctx* ctx, double* R);

/* ---- multi-walker wave-function API (one call = all walkers of the context) ----------------- */
/* TrialWaveFunction::mw_evaluateLog (TrialWaveFunction.cpp:161-213) including
 * DiracDeterminantBatched::mw_recompute / mw_invertPsiM (DiracDeterminantBatched.cpp:117-146,
 * 1121-1198), J1/J2 recompute.  Outputs (nullable): logpsi[nw], G[nw][N][3], L[nw][N]. */
int qmi_mw_evaluate_log(qmi_ctx* ctx, double* logpsi, double* G, double* L);

/* TrialWaveFunction::mw_evalGrad (TrialWaveFunction.cpp:568-593) ->
 * DiracDeterminantBatched::mw_evalGrad (DiracDeterminantBatched.cpp:183-211),
 * DelayedUpdateBatched::mw_evalGrad (DelayedUpdateBatched.h:353-400).  grads[nw][3]. */
int qmi_mw_eval_grad(qmi_ctx* ctx, int iat, double* grads);

/* ParticleSet::mw_makeMove + mw_computeNewPosDistTables (ParticleSet.cpp:397-495),
 * DistanceTable::mw_move (DistanceTable.h:153-161).  displ[nw][3]; proposed position
 * R[iat] + displ (not wrapped into the cell). */
int qmi_mw_make_move(qmi_ctx* ctx, int iat, const double* displ);

/* TrialWaveFunction::mw_calcRatioGrad (TrialWaveFunction.cpp:685-741) ->
 * DiracDeterminantBatched::mw_ratioGrad (DiracDeterminantBatched.cpp:307-354) with
 * SPOSet::mw_evaluateVGLandDetRatioGrads (SPOSet.h:346-352), J1/J2 ratioGrad.
 * ratios[nw]; grads[nw][3] (total gradient of log psi at the proposed position). */
int qmi_mw_ratio_grad(qmi_ctx* ctx, int iat, double* ratios, double* grads);

/* ---- per-component form ----------------------------------------------------------------------
 * A context built with only some of the components (qmi_system_desc without coefs: Jastrows only;
 * without j1_params / j2 arrays: no J1 / J2) is one WaveFunctionComponent of the reference, or a
 * subset of them; its qmi_mw_eval_grad is the component's mw_evalGrad (WaveFunctionComponent.h:
 * 234-237, grad_now assigned), qmi_mw_evaluate_log returns the component's log value, G and L, and
 * qmi_mw_accept_reject / qmi_mw_complete_updates are its mw_accept_rejectMove / mw_completeUpdates.
 * Several such contexts loaded with the same walkers and driven by the same calls (make_move with
 * the same displacements, the same acceptance masks) keep identical positions, as the components of
 * one TrialWaveFunction share one ParticleSet.
 *
 * WaveFunctionComponent::mw_ratioGrad (WaveFunctionComponent.h:305-309): ratios[nw] = the ratio of
 * this context's components, grad_new[nw][3] += their gradient at the proposed position (the
 * accumulation of DiracDeterminantBatched.cpp:352 and TwoBodyJastrow.cpp:576). */
int qmi_mw_component_ratio_grad(qmi_ctx* ctx, int iat, double* ratios, double* grad_new);

/* SPOSet::mw_evaluateVGLandDetRatioGrads (SPOSet.h:346-352; SplineR2R.cpp:414-580) with the caller's
 * inverse rows (invRow_ptr_list): orbital value, gradient and Laplacian of the first n orbitals at
 * pos[nw][3] into phi_vgl[5][nw][n] (the reference's phi_vgl_v layout; nullable), the determinant
 * ratio ratios[nw] = sum_j inv_rows[w][j] phi_j and grads[nw][3] = sum_j inv_rows[w][j] dphi_j / ratio. */
int qmi_mw_spo_vgl_det_ratio_grads(qmi_ctx* ctx, const double* pos, int n, const double* inv_rows, double* phi_vgl,
                                   double* ratios, double* grads);

/* Stateful distance tables over the walkers of a context (DistanceTable, Particle/DistanceTable.h:
 * 104-183), dense SoA semantics: kind 0 = electron pairs (SoaDistanceTableAA.h:45-127; row i holds
 * electrons j < i, the rest is not maintained), kind 1 = electrons to ions (SoaDistanceTableAB.h:
 * 41-86).  Minimum-image distances, displacements source - target.  The table is sized for the
 * context's walkers at creation; it reads the context's current positions. */
typedef struct qmi_dtable qmi_dtable;
int qmi_dt_create(qmi_ctx* ctx, int kind, qmi_dtable** dt);
int qmi_dt_destroy(qmi_dtable* dt);
/* DistanceTable::mw_evaluate (DistanceTable.h:104-121): the full table from the current positions */
int qmi_dt_mw_evaluate(qmi_dtable* dt);
/* DistanceTable::mw_move (DistanceTable.h:153-161; SoaDistanceTableAA.h:57-75, SoaDistanceTableAB.h:
 * 60-71): proposal rows from rnew[nw][3] for particle iat; prepare_old = 1 also computes the rows of
 * its current position (AA: old rows with the largest double at iat; AB: written into row iat). */
int qmi_dt_mw_move(qmi_dtable* dt, int iat, const double* rnew, int prepare_old);
/* partial = 0: DistanceTable::update (SoaDistanceTableAA.h:80-98, row [0, iat) and the column below
 * it from the proposal rows) for the walkers with accepted[w] != 0; partial = 1:
 * DistanceTable::mw_updatePartial (DistanceTable.h:166-183, SoaDistanceTableAA.h:100-120): row
 * [0, iat) from the proposal where accepted[w], else from the old rows (AB: from the proposal where
 * accepted, else unchanged) */
int qmi_dt_mw_update(qmi_dtable* dt, int iat, const int* accepted, int partial);
/* copies (nullable): r[nw][N][ns], dr[nw][N][3][ns], temp_r[nw][ns], temp_dr[nw][3][ns], old_r[nw][ns],
 * old_dr[nw][3][ns] (AA only); ns = N (AA) or the ion count (AB) */
int qmi_dt_get(qmi_dtable* dt, double* r, double* dr, double* temp_r, double* temp_dr, double* old_r, double* old_dr);

/* TrialWaveFunction::mw_accept_rejectMove (TrialWaveFunction.cpp:790-821) ->
 * DelayedUpdateBatched::mw_accept_rejectRow (DelayedUpdateBatched.h:542-670) and
 * ParticleSet::mw_accept_rejectMove forward mode (ParticleSet.cpp:717-758).
 * accepted[nw] (0/1).  safe_to_delay = 0 flushes the delayed updates immediately. */
int qmi_mw_accept_reject(qmi_ctx* ctx, int iat, const int* accepted, int safe_to_delay);

/* TrialWaveFunction::mw_completeUpdates -> DiracDeterminantBatched::mw_completeUpdates
 * (DiracDeterminantBatched.cpp:544-591): flush pending delayed rows of both determinants. */
int qmi_mw_complete_updates(qmi_ctx* ctx);

/* TrialWaveFunction::mw_evaluateGL (TrialWaveFunction.cpp:869-918); fromscratch = 1 recomputes
 * the determinants (re-inversion in FP64).  Outputs (nullable): logpsi[nw], G, L, and the kinetic
 * local energy -1/2 (sum L + sum G.G) per walker (BareKineticEnergy.cpp:100-118) in ekin[nw]. */
int qmi_mw_evaluate_gl(qmi_ctx* ctx, int fromscratch, double* logpsi, double* G, double* L, double* ekin);

/* SPOSet::mw_evaluateVGL (SPOSet.h:313-318): orbital value, gradient, Laplacian of the first n
 * orbitals at one position per walker.  pos[nw][3]; out[nw][5][n] (v, gx, gy, gz, lapl). */
int qmi_mw_spo_vgl(qmi_ctx* ctx, const double* pos, int n, double* out);

/* SPOSet::mw_evaluateDetRatios (SPOSet.h:257-261): value-only orbitals at nvp virtual positions
 * per walker dotted with a given inverse row.  pos[nw][nvp][3], inv_row[nw][n],
 * ratios[nw][nvp]. */
int qmi_mw_det_ratios(qmi_ctx* ctx, int nvp, const double* pos, int n, const double* inv_row, double* ratios);

/* DiracMatrixInverter::mw_invert_transpose (Fermion/DiracMatrixInverter.hpp:41-44; CUDA form
 * DiracMatrixInverterCUDA.hpp:305-360 with getrf/getri and computeLogDet_kernel,
 * detail/CUDA/cuBLAS_LU.cu:62-108; per matrix DiracMatrix::invert_transpose, DiracMatrix.h:189-220):
 * nmat matrices a[nmat][n][n] (row-major; (re, im) pairs when complex_values) ->
 * inv_a[nmat][n][n] = (a^-1)^T and log_values[nmat][2] = (log|det a|, phase) (nullable).  The batched
 * LU with partial pivoting of the determinant recompute, on the context's device and stream (no system
 * needed).  A singular matrix returns QMI_ERR_NUMERIC. */
int qmi_mw_invert_transpose(qmi_ctx* ctx, int complex_values, int n, int nmat, const double* a, double* inv_a,
                            double* log_values);
/* No counterpart in the reference.  qmi_mw_invert_transpose for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): a,
 * inv_a, log_values (nullable) and info are DEVICE arrays (a[nmat][n][n], inv_a[nmat][n][n] with
 * (re, im) pairs when complex_values, log_values[nmat][2], info[nmat] = 0 or the 1-based column of the
 * first zero pivot).  The LU factors, the solve work array and the pivots live in a fixed device
 * scratch of QMI_NOALLOC_SCRATCH_BYTES held by the library on each GPU it runs on; a call needing
 * more (16 n^2 nmat bytes real, 32 n^2 nmat complex, plus 8 n nmat) returns QMI_ERR_ARG.  The work is
 * enqueued on the context's stream without synchronisation or allocation; singular matrices are
 * reported only through info.  Calls that overlap on one GPU (any contexts) race on the scratch. */
#define QMI_NOALLOC_SCRATCH_BYTES (64u << 20)
int qmi_mw_invert_transpose_noalloc(qmi_ctx* ctx, int complex_values, int n, int nmat, const double* a,
                                    double* inv_a, double* log_values, int* info);
/* DelayedUpdateBatched::mw_transferAinv_D2H (DelayedUpdateBatched.h:813-824, guard_no_delay) /
 * DiracDeterminantBatched::get_psiMinv (DiracDeterminantBatched.h:276): the inverse Slater matrix of
 * spin group g of every walker, ainv[nw][n][n] (row i = inverse row of electron i of the group; FP32
 * storage widened), after the delayed updates are completed. */
int qmi_mw_get_inverse(qmi_ctx* ctx, int g, double* ainv);
/* DistanceTable::mw_move (DistanceTable.h:153-161) with SoaDistanceTableAA/AB rows
 * (SoaDistanceTableAA.h:58-75, SoaDistanceTableAB.h:59-78): distances and displacements
 * (source - target, minimum image) from pos[nw][3] to all electrons and all ions.
 * ee_r[nw][N], ee_dr[nw][N][3], ei_r[nw][n_ion], ei_dr[nw][n_ion][3] (nullable). */
/* TrialWaveFunction::mw_evaluateRatios (TrialWaveFunction.cpp:1079-1110) over one
 * VirtualParticleSet per walker (Particle/VirtualParticleSet.cpp; the virtual moves of non-local
 * pseudopotentials, NonLocalECPComponent.cpp:208-228): ratios[w][v] = psi with electron iat at
 * pos[w][v] over psi (determinant: stored inverse row, DiracDeterminantBatched.cpp:812-846;
 * J1OrbitalSoA.h:295-299; TwoBodyJastrow.cpp:166-171).  Needs no pending delayed update in iat's
 * spin group (call qmi_mw_complete_updates first). */
int qmi_mw_evaluate_ratios(qmi_ctx* ctx, int iat, int nvp, const double* pos, double* ratios);
/* Periodic Coulomb energy with the Ewald breakup (CoulombPBCAA / CoulombPBCAB with
 * Particle/LongRange/EwaldHandler3D; QMCHamiltonians/CoulombPBCAA.cpp:568-840,
 * CoulombPBCAB.cpp:373-453): electrons carry charge -1, the ions ion_charge; dim_cutoff = k_c r_c
 * (the reference's LR_dim_cutoff, 15 by default).  Periodic orthorhombic cells; tables built once
 * per system. */
int qmi_coulomb_setup(qmi_ctx* ctx, double ion_charge, double dim_cutoff);
/* StructFact::mw_updateAllPart (Particle/LongRange/StructFact.cpp:63-135, computeRhok 138-205) for
 * the electrons of every walker: rhok[nw][2][nk][2] = for spin group g (up, down: the electron
 * species) the sums over its electrons of cos(k . r) and sin(k . r) at the wave vectors
 * kvecs[nk][3] (Cartesian; e.g. a KContainer's list).  At most 1024 electrons. */
int qmi_mw_structure_factor(qmi_ctx* ctx, int nk, const double* kvecs, double* rhok);
/* Cell vectors of the orbitals' lattice, R[3][3] row-major (rows are the vectors; the
 * reference's CrystalLattice R, Lattice/CrystalLattice.h, with G = R^-1 used by SplineR2R /
 * SplineC2C to reduce positions and to map orbital gradients and Hessians,
 * SplineR2R.cpp:338-374).  Replaces the diagonal cell of the system description.  A cell with
 * off-diagonal vectors serves the orbital paths only (qmi_mw_spo_*, determinant ratios and
 * moves without Jastrow factors): with Jastrow functors in the system it returns
 * QMI_ERR_UNSUPPORTED, and qmi_coulomb_setup / qmi_nlpp_setup / qmi_dt_create refuse it afterwards
 * (their minimum images are orthorhombic).  Any lattice is set before those set-ups (after them:
 * QMI_ERR_ARG for a diagonal cell, QMI_ERR_UNSUPPORTED for a skewed one). */
int qmi_system_set_lattice(qmi_ctx* ctx, const double* R);
/* Long-range handlers of the periodic Coulomb sum (the lattice input's LR_handler,
 * Particle/ParticleIO/LatticeIO.cpp:123-143; LRCoulombSingleton::getHandler,
 * Particle/LongRange/LRCoulombSingleton.cpp:58-97) */
#define QMI_LR_EWALD 0       /* "ewald": EwaldHandler3D */
#define QMI_LR_OPT_BREAKUP 1 /* "opt_breakup", the reference's default: LRHandlerTemp<CoulombFunctor,
                              * LPQHIBasis>, the optimized breakup fitted on 15 knots */
/* qmi_coulomb_setup with the long-range handler chosen (QMI_LR_*); qmi_coulomb_setup is
 * qmi_coulomb_setup_handler(..., QMI_LR_EWALD). */
int qmi_coulomb_setup_handler(qmi_ctx* ctx, double ion_charge, double dim_cutoff, int lr_handler);
/* The optimized breakup of an orthorhombic cell with sides L[3] (host computation, no device
 * work and no context): the 45 basis coefficients (LRHandlerTemp::coefs after InitBreakup,
 * LRHandlerTemp.h:195-227, 3 per knot: value, slope, curvature) into coefs[ncoefs >= 45] and
 * info[4] = {r_c, v_l(r = 0) (evaluateLR_r0), v_s(k = 0) (evaluateSR_k0), chi^2 of the fit}. */
int qmi_lr_breakup(const double* L, double dim_cutoff, double* coefs, int ncoefs, double* info);
/* Non-local pseudopotential energy of one ion species: NonLocalECPotential::evaluate
 * (QMCHamiltonians/NonLocalECPotential.cpp:200-228) with NonLocalECPComponent::evaluateOne
 * (NonLocalECPComponent.cpp:109-166, 826-888).  Channel c has angular momentum l[c] and radial
 * potential v_c(r) given as a cubic spline on the linear grid r_i = i dr, i < ngrid: values
 * v[c][ngrid] and second derivatives v2[c][ngrid] (OneDimCubicSpline data); electron-ion pairs
 * with r < rcut (<= grid end) contribute.  The quadrature rule (Quadrature3D): nknot unit vectors
 * knots[nknot][3] and weights[nknot].  Real orbitals. */
int qmi_nlpp_setup(qmi_ctx* ctx, int nchannel, const int* l, int ngrid, double dr, const double* v, const double* v2,
                   double rcut, int nknot, const double* knots, const double* weights);
/* v[nw]: non-local energy of every walker, the grid rotated by rot[w][3][3] (row-major, as
 * rotateQuadratureGrid with generateRandomRotationMatrix); after mw_completeUpdates */
int qmi_mw_evaluate_nlpp(qmi_ctx* ctx, const double* rot, double* v);

/* DMC driver: DMCBatched::advanceWalkers / runDMCStep (QMCDrivers/DMC/DMCBatched.cpp:67-414),
 * SFNBranch (QMCDrivers/SFNBranch.{h,cpp}: branch weight, trial / reference energy, effective
 * time step, branching cutoff) and WalkerControl::branch with a fluctuating population
 * (QMCDrivers/DMC/WalkerControl.cpp:150-260).  Needs one walker per crowd and
 * qmi_coulomb_setup; the local energy is kinetic + V_ee + V_eI + V_II. */
typedef struct qmi_dmc_params
{
  double tau;             /* time step */
  double feedback;        /* population feedback (DMC input "feedback", 1.0) */
  double sigma_bound;     /* SFNBranch sigmaBound (10) */
  int warmup_steps;       /* SFNBranch warmupSteps (200) */
  int refenergy_history;  /* 0: limited_history (default), 1: unlimited_history */
  int cutoff_scheme;      /* branching_cutoff_scheme: 0 classic, 1 DRV, 2 ZSGMA, 3 YL */
  int capacity;           /* walker slots; a population beyond it is an error */
  int target_walkers;     /* 0: the population at qmi_dmc_init */
  int recompute_period;   /* full re-inversion every this many steps (0: never) */
  uint32_t seed;          /* walker slot w draws from seed + w, walker control from seed - 1 */
  int nonlocal_moves;     /* DMC input "nonlocalmoves": 0 locality approximation, 1 T-moves v0
                           * (NonLocalECPotential::makeNonLocalMovesPbyP, NonLocalECPotential.cpp:
                           * 546-567; NonLocalTOperator::selectMove, NonLocalTOperator.cpp:79-106;
                           * needs qmi_nlpp_setup) */
} qmi_dmc_params;
#define QMI_DMC_STATS 11
/* re-evaluates every walker (log psi, local energy), sets E_T = E_ref = mean energy */
int qmi_dmc_init(qmi_ctx* ctx, const qmi_dmc_params* params);
/* stats[steps][QMI_DMC_STATS] per step: weighted mean energy and variance before branching,
 * total weight, walkers before and after branching, E_trial and E_ref after the update,
 * effective time step, accepted and rejected moves, accepted non-local moves */
int qmi_dmc_run(qmi_ctx* ctx, int steps, double* stats);
/* current population (walkers in slots 0 .. n-1) */
int qmi_num_walkers(qmi_ctx* ctx);
/* weights[n], local energies[n] of the current population (nullable outputs) */
int qmi_dmc_walker_data(qmi_ctx* ctx, double* weights, double* energies);

/* ---- restart files ---------------------------------------------------------------------------
 * A checkpoint holds the walkers and the random streams (HDFWalkerOutput::dump,
 * Particle/HDFWalkerOutput.cpp:81-160: state_0/{block, number_of_walkers, walker_partition,
 * walkers [nw][N][3], walker_weights}; RandomNumberControl::write, Utilities/RandomNumberControl.cpp:
 * 194-209 with StdRandom::save, Utilities/StdRandom.cpp: one std::mt19937 state per stream) and, for
 * a DMC context, the branch state (BranchIO<SFNBranch>::write, QMCDrivers/BranchIO.cpp:109-160) with
 * the walker-control stream and the walkers' local energies.  The datasets keep those names in a
 * flat little-endian container ("QMICKPT1", u32 version 1, u32 dataset count; per dataset u32 name
 * length, name, u32 type 1 f64 / 2 i64 / 3 u32, u32 rank, u64 dims, row-major payload). */
/* writes the checkpoint into buf (capacity bytes); *bytes receives its size (buf NULL: size only).
 * Needs no pending delayed updates. */
int qmi_checkpoint_save(qmi_ctx* ctx, void* buf, size_t capacity, size_t* bytes);
/* restart (HDFWalkerInput_0_4::read_hdf5, Particle/HDFWalkerInput_0_4.cpp:126-190;
 * RandomNumberControl::read, :183-192; BranchIO::read, QMCDrivers/BranchIO.cpp:166-208): reloads the
 * walkers and re-evaluates them from scratch (QMCDriverNew::initialLogEvaluation), restores the
 * streams and the step count, and for a DMC checkpoint the branch state (qmi_dmc_run continues;
 * qmi_coulomb_setup and, if the checkpoint has the pseudopotential term, qmi_nlpp_setup first).
 * qmi_vmc_run continues the restored streams when given the checkpoint's seed.  The checkpoint is
 * validated before the context changes. */
int qmi_checkpoint_load(qmi_ctx* ctx, const void* buf, size_t bytes);
/* number of k vectors of the long-range sum (0 before qmi_coulomb_setup) */
int qmi_coulomb_nk(qmi_ctx* ctx);
/* v[nw][3] = {electron-electron, electron-ion, ion-ion} energy of every walker */
int qmi_mw_evaluate_potential(qmi_ctx* ctx, double* v);
int qmi_mw_distance_rows(qmi_ctx* ctx, const double* pos, double* ee_r, double* ee_dr, double* ei_r,
                         double* ei_dr);

/* ---- device-resident driver ---------------------------------------------------------------- */
/* VMCBatched::advanceWalkers<POS> (QMCDrivers/VMC/VMCBatched.cpp:57-228) for every crowd of the
 * context, `steps` times: each step runs sub_steps sweeps with drift-diffusion moves
 * (DriftModifierUNR, a = 1), Metropolis-Hastings on the device with the reference's RNG order
 * (one std::mt19937 per crowd seeded seed + crowd index, boost-style uniforms, Box-Muller
 * Gaussians, uniforms drawn only for |ratio|^2 >= eps), then evaluateGL (fromscratch on steps
 * where (step_count + 1) % recompute_period == 0) and the kinetic local energy.
 * energies[steps][nw] (nullable); trace[steps][sub_steps][N][nw][8] = {ratio, accepted,
 * proposed position[3], gradient at the current position[3]} (nullable: parity only). */
int qmi_vmc_run(qmi_ctx* ctx, int steps, int sub_steps, double tau, int recompute_period, uint32_t seed,
                double* energies, double* trace);
/* accepted / rejected move counts since the context was created: stats[2] */
int qmi_vmc_stats(qmi_ctx* ctx, double* stats);

/* ---- learned per-electron proposal ----------------------------------------------------------
 * No counterpart in the reference: an MLP proposal inside the same Metropolis-Hastings step, for
 * callers that want fewer correlated samples per move.  Features of the moving electron (68): the
 * 4 nearest ions (dx, dy, dz, |d|), the 16 nearest other electrons (dx, dy, dz), the drift and the
 * timestep.  Network 68 -> 128 -> 128 -> 128 -> 4 (SiLU; BF16 weights, FP32 accumulation on the
 * matrix cores); outputs (mu[3], s) give r' = r + drift + mu + sqrt(tau) e^s xi, and the acceptance
 * carries the asymmetric-proposal correction q(r|r')/q(r'|r) with the reverse network evaluated at
 * r' (exact sampling of |psi|^2 for any weights).  A zero output layer reproduces plain
 * drift-diffusion.  Training (optional, on the device): Adam on -mean of
 * min(1, acceptance) |r' - r|^2 over the moves of the run.  Needs one walker per crowd. */
#define QMI_MLP_NPARAM 42372 /* W1[128][68], b1, W2[128][128], b2, W3[128][128], b3, W4[4][128], b4 */
/* set = 1: upload params[QMI_MLP_NPARAM] (FP32); set = 0: download them */
int qmi_mlp_params(qmi_ctx* ctx, int set, float* params);
/* forward pass of the current network on X[B][68] -> out[B][4] */
int qmi_mlp_forward(qmi_ctx* ctx, int B, const float* X, float* out);
/* gradient of sum_b dout[b] . out[b] with respect to the parameters -> grad[QMI_MLP_NPARAM] */
int qmi_mlp_gradient(qmi_ctx* ctx, int B, const float* X, const float* dout, float* grad);
/* learned: use the learned proposal in qmi_vmc_run; train: update the network with Adam (rate lr)
 * after every move; init_seed != 0: fresh network (uniform fan-in weights, zero output layer) */
int qmi_proposal_config(qmi_ctx* ctx, int learned, int train, double lr, uint32_t init_seed);
/* since the last call: out[4] = {mean objective per move, mean acceptance probability, moves,
 * training steps so far} */
int qmi_proposal_stats(qmi_ctx* ctx, double* out);
/* features of electron iat at pos[nw][3] with drift[nw][3] in each walker -> out[nw][68] */
int qmi_mw_proposal_features(qmi_ctx* ctx, int iat, const double* pos, const double* drift, double tau,
                             float* out);
/* qmi_vmc_run with, for the learned proposal, ltrace[steps][sub_steps][N][nw][15] = {mu[3], s,
 * mu'[3], s', log q_fwd, log q_rev, acceptance probability, objective, drift at r'[3]} (nullable) */
int qmi_vmc_run_ex(qmi_ctx* ctx, int steps, int sub_steps, double tau, int recompute_period, uint32_t seed,
                   double* energies, double* trace, double* ltrace);

/* ---- uneven batching: tuning knobs, graph replay, batches of contexts ------------------------
 * No counterpart in the reference (its batched driver runs crowds of one system).  A batch
 * advances several contexts (walker populations of different systems: electron counts, cells,
 * tables, real or complex orbitals) together: one sub-step of every context is captured into one
 * HIP graph, each context on its own stream (concurrent branches) or chained, and replayed every
 * sub-step.  Results of each context are those of qmi_vmc_run on it alone. */
/* delay_rank (<= 0: unchanged; else 1..64, at most the determinant size), move_kernel (0: one
 * launch per move, 1: one launch per delay window, real orbitals with one walker per crowd),
 * flush_variant (0: matrix-core flush for n <= 512, 1: three LDS-tiled passes).  Between sweeps. */
int qmi_tune_set(qmi_ctx* ctx, int delay_rank, int move_kernel, int flush_variant);
int qmi_tune_get(qmi_ctx* ctx, int* out3);
/* Walker partitions of the device-resident sweep (no reference counterpart): partitions > 0 sets P
 * (1..8); the fused single-move path then runs the moves of P walker ranges on P streams forked from
 * the context's stream (at least 64 walkers per range), so that launches of different ranges overlap
 * on the compute units.  Results are unchanged.  *current (nullable) receives P.  Partitioned
 * sub-steps are launched eagerly, not from a captured graph (replayed, the partitions' chains did
 * not overlap). */
int qmi_tune_partitions(qmi_ctx* ctx, int partitions, int* current);
/* on = 1: qmi_vmc_run replays each sub-step from a captured graph (runs without trace, with one
 * walker partition) */
int qmi_graph_enable(qmi_ctx* ctx, int on);
typedef struct qmi_batch qmi_batch;
int qmi_batch_create(qmi_batch** batch, qmi_ctx* const* ctxs, int nctx);
int qmi_batch_destroy(qmi_batch* batch);
const char* qmi_batch_last_error(const qmi_batch* batch);
/* seeds[nctx] (one std::mt19937 per crowd seeded seed + crowd, as qmi_vmc_run);
 * energies[nctx] -> [steps][nw of that context] kinetic local energies (entries nullable) */
int qmi_batch_vmc_run(qmi_batch* batch, int steps, int sub_steps, double tau, int recompute_period,
                      const uint32_t* seeds, double* const* energies);
/* Autotuner: per context, delay rank, move-kernel form and flush form by timing the context's
 * sub-step alone (its own graph, median of `trials`), then concurrent vs chained branches of the
 * batch.  choice[nctx][3] = {delay rank, move_kernel, flush_variant}; ms[nctx + 2] = best sub-step
 * time of each context, then the chained and the concurrent batch (nullable).  The chains advance
 * while tuning. */
int qmi_batch_autotune(qmi_batch* batch, double tau, int trials, int* choice, double* ms);
int qmi_batch_set_concurrent(qmi_batch* batch, int concurrent);
int qmi_batch_get_concurrent(qmi_batch* batch, int* concurrent);

/* ---- measurement ---------------------------------------------------------------------------- */
/* Per-kernel device time accumulated with HIP events on the context stream (enabled by
 * qmi_timing_enable(ctx, 1)).  names: '\n'-separated kernel names written to buf (size buflen);
 * ms[i], calls[i] for up to maxk kernels; returns the number of kernels in *nk. */
int qmi_timing_enable(qmi_ctx* ctx, int on);
int qmi_timing_get(qmi_ctx* ctx, char* buf, int buflen, double* ms, double* calls, int maxk, int* nk);
int qmi_timing_reset(qmi_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
