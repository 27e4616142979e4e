/* CPU FP64 oracle for the batched walker-move path of real-space VMC.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it. The product
 * library (qmcpack_amd/) never links, loads or calls anything here.
 *
 * Every routine is a from-scratch restatement of the reference algorithm
 * (QMCPACK, see the file:line citations in qmc_oracle.cpp), written with the
 * reference's operation order where that is cheap, so that GPU results can be
 * compared at 1e-10 relative.
 */
#ifndef QMC_ORACLE_H
#define QMC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Synthetic system: orthorhombic periodic cell (L[d] <= 0 means open along d),
 * ions, two spin groups, one real B-spline orbital table, J1 and J2 B-spline
 * functors.  Mirrors the fields of qmi_system_desc in include/qmcmi.h. */
typedef struct
{
  double L[3];
  int n_up, n_dn;
  int n_ion;
  const double* ions; /* [n_ion][3] */
  int M[3];           /* periodic grid points per dimension */
  int norb;           /* orbitals in the table (>= max(n_up, n_dn)); 0 = no determinant */
  const double* coefs; /* unique coefficients [Mx][My][Mz][norb] */
  int j1_np;           /* 0 = no J1 */
  const double* j1_params;
  double j1_rcut, j1_cusp;
  int j2_np; /* 0 = no J2 */
  const double* j2_uu;
  const double* j2_ud;
  double j2_rcut, j2_cusp_uu, j2_cusp_ud;
  int delay_rank; /* k: 1 = Sherman-Morrison, >1 delayed Woodbury */
  int complex_orbitals;   /* 1: C2C orbitals, coefs [Mx][My][Mz][norb][2] (re, im) */
  const double* kpoints;  /* [norb][3] Cartesian k of each orbital (complex orbitals) */
} orc_system_desc;

/* RNG: mt19937 + boost-style uniform + Box-Muller */
int orc_rng_uniform(uint32_t seed, int n, double* out);
int orc_rng_gauss(uint32_t seed, int n, double* out);

/* spline pieces */
void orc_spline_prefactors(double tx, double* a, double* da, double* d2a);
double orc_symtrace(const double* h6, const double* gg6);
/* raw 3D VGH in unit-cube coordinates: out[10][norb] = v, gx,gy,gz, hxx,hxy,hxz,hyy,hyz,hzz */
int orc_spline_vgh_raw(const int* M, int norb, const double* unique_coefs, const double* ru, double* out);
/* SplineR2R::evaluateVGL at Cartesian r for an orthorhombic cell: out[5][n] = v, gx,gy,gz, lapl */
int orc_spo_vgl(const double* L, const int* M, int norb, const double* unique_coefs, const double* r, int n,
                double* out);

/* radial functor value, du/dr, d2u/dr2 */
int orc_functor_vgl(int np, const double* params, double rcut, double cusp, double r, double* out3);

/* minimum-image distances from pos to each src point: r[n], dr[n][3] = src - pos */
int orc_min_image(const double* L, const double* src, int nsrc, const double* pos, double* r, double* dr);

/* DiracMatrix::invert_transpose: ainv = inverse of transpose(a), logdet = log det(a) (complex) */
int orc_invert_transpose(int n, const double* a, double* ainv, double* logdet2);

/* batched determinant + delayed-update engine, nw walkers in lock-step */
void* orc_det_create(int nw, int n, int k);
void orc_det_destroy(void* h);
int orc_det_recompute(void* h, int iw, const double* psiM, const double* dpsiM, const double* d2psiM);
int orc_det_eval_grad(void* h, int row, double* grads);
int orc_det_ratio_grad(void* h, int row, const double* phi_vgl, double* ratios, double* grads);
int orc_det_accept_reject(void* h, int row, const int* accepted, int safe_to_delay);
int orc_det_complete(void* h);
int orc_det_get(void* h, int iw, double* ainv, double* logdet2);

/* wavefunction + VMC driver over a walker population split in crowds */
void* orc_vmc_create(const orc_system_desc* d, int nw, const double* R, int walkers_per_crowd, uint32_t seed);
void orc_vmc_destroy(void* h);
/* load the crowds' std::mt19937 states (StdRandom::load): words[nstreams][state_size] */
int orc_vmc_set_rng(void* h, int nstreams, int state_size, const uint32_t* words);
/* initial log evaluation (recompute everything); logpsi[nw], G[nw][N][3], L[nw][N] (nullable) */
int orc_vmc_evaluate_log(void* h, double* logpsi, double* G, double* L);
/* single-walker protocol used by the known-answer tests */
int orc_wf_ratio_grad(void* h, int iw, int iat, const double* newpos, double* ratio, double* grad);
int orc_wf_accept_reject(void* h, int iw, int iat, int accepted);
int orc_wf_eval_grad(void* h, int iw, int iat, double* grad);
int orc_wf_evaluate_gl(void* h, int iw, int fromscratch, double* logpsi, double* G, double* L);
/* VMC steps.  energies[steps][nw] (kinetic local energy), trace (nullable) per
 * sub-step, electron and walker: {ratio, accepted, proposed position[3], gradient at the
 * current position[3]} */
int orc_vmc_run(void* h, int steps, int sub_steps, double tau, int recompute_period, int nthreads,
                double* energies, double* trace);
/* Replay: as orc_vmc_run, but each move's proposed position and accept decision are taken from
 * `forced` (a trace of another implementation); everything else is computed here and written to
 * `trace` (own proposed position in fields 2..4, own decision in field 1). */
int orc_vmc_replay(void* h, int steps, int sub_steps, double tau, int recompute_period, int nthreads,
                   double* energies, double* trace, const double* forced);
/* replay with the long double determinant shadow: exact[step][sub][iat][iw][8], field 0 the ratio,
 * 5-7 the drift gradient of every traced move */
int orc_vmc_replay_exact(void* h, int steps, int sub_steps, double tau, int recompute_period, int nthreads,
                         double* energies, double* trace, const double* forced, double* exact);
/* Driver known answers of the reference's tests: a generator returning `value` on every call
 * (FakeRandom), test-only wave-function components (mode 1: ConstantOrbital with ratio fake_ratio,
 * mode 2: LinearOrbital), one sweep with the VMC or DMC acceptance, the drift, the tau parameters
 * and the Box-Muller deviates of a constant generator. */
int orc_vmc_fake_rng(void* h, double value);
int orc_vmc_test_orbital(void* h, int mode, double fake_ratio);
int orc_vmc_sweep(void* h, double tau, int dmc, double* energies);
int orc_get_drift(double tau, const double* g, double* drift);
int orc_tau_params(double tau, double inv_mass, double* out);
int orc_rng_gauss_const(double value, int n, double* out);
/* R[nw][N][3], logpsi[nw], G, L, stats[2] = {accepted, rejected} (nullable each) */
int orc_vmc_get(void* h, double* R, double* logpsi, double* G, double* L, double* stats);

/* DMC (DMCBatched, SFNBranch, WalkerControl with a fluctuating population) on the walkers of an
 * orc_vmc_create handle with one walker per crowd.  Mirrors qmi_dmc_params. */
typedef struct
{
  double tau, feedback, sigma_bound;
  int warmup_steps, refenergy_history, cutoff_scheme, capacity, target_walkers, recompute_period;
  uint32_t seed;
  int nonlocal_moves; /* 0 locality approximation, 1 T-moves v0 */
} orc_dmc_params;
int orc_dmc_init(void* h, double ion_charge, double dim_cutoff, const orc_dmc_params* p);
/* stats[steps][11]: as qmi_dmc_run */
int orc_dmc_run(void* h, int steps, double* stats);
int orc_num_walkers(void* h);
/* pseudopotential in the DMC local energy (before orc_dmc_init; real orbitals) */
int orc_dmc_set_nlpp(void* h, int nch, const int* l, int ng, double dr, const double* v, const double* v2, double rcut,
                     int nknot, const double* knots, const double* weights);
/* generateRotationMatrix(u1, u2, u3), row-major */
int orc_rotation_matrix(double u1, double u2, double u3, double* out9);
/* non-local pseudopotential energy out[nw] (see qmi_nlpp_setup / qmi_mw_evaluate_nlpp) */
int orc_nlpp_energy(void* h, int nch, const int* l, int ng, double dr, const double* v, const double* v2, double rcut,
                    int nknot, const double* knots, const double* weights, const double* rot, double* out);
int orc_dmc_walker_data(void* h, double* weights, double* energies);
/* NonLocalTOperator::selectMove (alpha = gamma = 0): selected candidate index or -1 */
int orc_select_move(double prob, int n, const double* weights, double tau);

#ifdef __cplusplus
}
#endif
#endif
