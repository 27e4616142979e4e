// Kernel launch interface shared by the kernel translation units and the host library.
#pragma once
#include "qmi_device.hpp"

namespace qmi
{
// Calls f(ST{}, DT{}) with the spline storage type ST and determinant storage type DT of a
// precision mode.
template<typename F>
hipError_t dispatch_prec(int prec, F&& f)
{
  switch (prec)
  {
  case PREC_FP64: return f(double{}, double{});
  case PREC_FP32: return f(float{}, float{});
  case PREC_BF16: return f(bf16s{}, float{});
  case PREC_T32: return f(float{}, double{});
  case PREC_TBF16: return f(bf16s{}, double{});
  default: return hipErrorInvalidValue;
  }
}

// arguments of the per-walker move kernels
struct MoveArgs
{
  int iat0, iat1; // moves [iat0, iat1) (one move except for the fused segment)
  int g;          // spin group of the moves
  int d0;         // delay count at iat0
  int nw, wpc;    // walkers, walkers per crowd
  // walkers [w0, w0 + wn) of this launch (wn = 0: all nw); only the fused move kernels (k_move_one,
  // k_sweep_segment, k_move_one_c) take a partition, the others run with w0 = 0
  int w0, wn;
  int prepare;    // API ratio kernel: prepare the inverse row first
  double tau;
  const double* gauss; // [N][nw][3] standard normals of the sub-step
  const double* unif;  // [ncrowd][N*wpc] uniforms of the sub-step (drawn lazily in order)
  int* ubase;          // [ncrowd][N+1] uniforms consumed before each move
  const int* host_acc; // [nw] accept decisions (API accept)
  double* out;         // API outputs
  double* trace;       // [N][nw][TRACE_FIELDS] (nullable)
  // DMC acceptance (DMCBatched.cpp:190-232): moves across a node of a real wave function are
  // rejected without a draw, the uniform is compared only when prob_eff >= eps, and each walker
  // accumulates tau |xi|^2 of its proposed and accepted moves in rr[nw][2] (one walker per crowd)
  int dmc;
  double* rr;
  // NaNguard (NaNguard.cpp:18-31) on the device: nanflag[w] = 1 + the first electron of walker w
  // whose ratio, or whose gradient, is NaN (0: none); read by the host once per step
  int* nanflag;
};

// the check of TrialWaveFunction::mw_evalGrad / mw_calcRatioGrad (TrialWaveFunction.cpp:592, 736-739):
// squared norms of the ratio, of the gradient at the proposed position (checked only for a non-zero
// ratio) and of the gradient at the current position
__device__ inline void nan_guard(const MoveArgs& a, int w, int iat, double ratio2, double gnew2, double gnow2)
{
  if (a.nanflag && (isnan(ratio2) || (ratio2 != 0.0 && isnan(gnew2)) || isnan(gnow2)) && a.nanflag[w] == 0)
    a.nanflag[w] = iat + 1;
}

enum MoveKernel
{
  MK_SEGMENT = 0,
  MK_PROPOSE,
  MK_ACCEPT_CROWD,
  MK_API_GRAD,
  MK_API_RATIO,
  MK_API_ACCEPT,
  MK_MOVE_ONE
};

// coefs: the spline table in the storage type of s.prec (double, float or bf16s)
hipError_t launch_move(int which, const SysC& s, const State& st, const void* coefs, const MoveArgs& a,
                       hipStream_t stream);
hipError_t launch_make_move(const SysC& s, const State& st, int iat, int nw, const double* displ, hipStream_t stream);

// initialisation / recompute / GL
hipError_t launch_wrap_coefs(const SysC& s, const double* unique, int norb, void* table, hipStream_t stream);
// orbital rows of determinant g: values transposed into the FP64 LU workspace lu ([nw][n][ld]),
// gradients/Laplacians into dpsi
hipError_t launch_spo_rows(const SysC& s, const State& st, const void* coefs, int g, int nw, double* lu,
                           hipStream_t stream);
// Batched LU with partial pivoting of nmat matrices A(i, c) = a[m][c * ld + i] (FP64 real, or complex
// (re, im) pairs): factors in place, ipiv[m][n] (1-based), perm[m][k] = row of the unit of P e_k,
// info[m] = first zero pivot (1-based) or 0
hipError_t launch_lu_factor(bool complex_values, int n, int ld, int nmat, double* a, int* piv, int* perm, int* info,
                            hipStream_t stream);
// log det of the factored matrices: out[m * ostride + {0, 1}] = {log|det|, phase}
hipError_t launch_lu_logdet(bool complex_values, int n, int ld, int nmat, const double* a, const int* piv, double* out,
                            int ostride, hipStream_t stream);
// inverse from the factors, written transposed (out[m][k * ld + i] = (A^-1)(i, k)) in the storage type
// out_kind; work: FP64 [nmat][n][ld] (complex: pairs)
enum LuOut
{
  LU_OUT_F64 = 0,
  LU_OUT_F32,
  LU_OUT_C128
};
hipError_t launch_lu_inverse(int out_kind, int n, int ld, int nmat, const double* lu, const int* perm, double* work,
                             void* out, hipStream_t stream);
// allocation-free inverter: the address of this device's scratch (resolved once per device, with that
// device current), and the transposing copy of row-major matrices into the factorisation's layout
constexpr int QMI_MAX_DEVICES = 64;
hipError_t lu_noalloc_scratch(int device, void** out);
hipError_t launch_lu_transpose_in(int n, int nmat, int vw, const double* a, double* t, hipStream_t stream);
hipError_t launch_jastrow_recompute(const SysC& s, const State& st, int nw, hipStream_t stream);
// per-walker outputs: out3[nw][3] = {ekin, logpsi, phase}; G[nw][N][3], L[nw][N] (nullable)
hipError_t launch_eval_gl(const SysC& s, const State& st, int nw, double* out3, double* G, double* L,
                          hipStream_t stream);

// delayed-update flush of determinant g with d delayed rows (Ainv -= (Ainv U^T - W) Binv V).
// variant 0: one matrix-core pass for n <= 256, three LDS-tiled passes above; 1: always the three passes
hipError_t launch_flush(const SysC& s, const State& st, int g, int d, int nw, double* scratchT, hipStream_t stream,
                        int variant = 0);
// p[0, n) = v (a kernel, so that it can be captured in a graph)
hipError_t launch_fill_i32(int* p, size_t n, int v, hipStream_t stream);
// non-local moves: acc[w] = chosen[w] && ratio[w] != 0 (ratio[nw] as written by the ratio kernel),
// count[0] += accepted, count[1] += chosen moves with a NaN ratio
hipError_t launch_tmove_accept_mask(int nw, const double* ratio, const int* chosen, int* acc, int* count,
                                    hipStream_t stream);

// API helpers
hipError_t launch_api_spo_vgl(const SysC& s, const void* coefs, int nw, int n, const double* pos, double* out,
                              hipStream_t stream);
hipError_t launch_api_det_ratios(const SysC& s, const void* coefs, int nw, int nvp, int n, const double* pos,
                                 const double* inv, double* out, hipStream_t stream);
// SPOSet::mw_evaluateVGLandDetRatioGrads: phi[5][nw][n] (nullable), rg[nw][4] = ratio, gradient / ratio
hipError_t launch_api_vgl_ratio_grads(const SysC& s, const void* coefs, int nw, int n, const double* pos,
                                      const double* inv, double* phi, double* rg, hipStream_t stream);
// stateful distance tables (AA: electron pairs, lower triangle; AB: electron-ion)
enum DTableKind
{
  DT_AA = 0,
  DT_AB = 1
};
struct DTableArrays
{
  int kind, ns;          // ns sources per row: N (AA) or the ion count (AB)
  double* r;             // [nw][N][ns]
  double* dr;            // [nw][N][3][ns]
  double* temp_r;        // [nw][ns]
  double* temp_dr;       // [nw][3][ns]
  double* old_r;         // [nw][ns] (AA)
  double* old_dr;        // [nw][3][ns] (AA)
};
hipError_t launch_dt_evaluate(const SysC& s, const State& st, const DTableArrays& t, int nw, hipStream_t stream);
hipError_t launch_dt_move(const SysC& s, const State& st, const DTableArrays& t, int nw, int iat, const double* rnew,
                          int prepare_old, hipStream_t stream);
hipError_t launch_dt_update(const SysC& s, const DTableArrays& t, int nw, int iat, const int* acc, int partial,
                            hipStream_t stream);
// complex (C2C) orbitals: complex FP64 determinant engine; value arrays hold (re, im) pairs
hipError_t launch_move_c(int which, const SysC& s, const State& st, const void* coefs, const MoveArgs& a,
                         hipStream_t stream);
hipError_t launch_spo_rows_c(const SysC& s, const State& st, const void* coefs, int g, int nw, double* lu,
                             hipStream_t stream);
hipError_t launch_eval_gl_c(const SysC& s, const State& st, int nw, double* out3, double* G, double* L,
                            hipStream_t stream);
hipError_t launch_flush_c(const SysC& s, const State& st, int g, int d, int nw, double* scratchT, hipStream_t stream);
hipError_t launch_api_spo_vgl_c(const SysC& s, const void* coefs, const double* kpts, int nw, int n,
                                const double* pos, double* out, hipStream_t stream);
hipError_t launch_api_det_ratios_c(const SysC& s, const void* coefs, const double* kpts, int nw, int nvp, int n,
                                   const double* pos, const double* inv, double* out, hipStream_t stream);
hipError_t launch_api_vgl_ratio_grads_c(const SysC& s, const void* coefs, const double* kpts, int nw, int n,
                                        const double* pos, const double* inv, double* phi, double* rg,
                                        hipStream_t stream);
// Non-local pseudopotential of one ion species (host-built tables, device pointers)
constexpr int NLPP_MAX_CH = 4, NLPP_MAX_L = 4, NLPP_MAX_IONS = 2048, NLPP_MAX_KNOTS = 64;
struct NlppC
{
  int nch, lmax;          // channels, largest l among them
  int l[NLPP_MAX_CH];     // angular momentum of each channel
  int ng;                 // radial grid points on [0, (ng - 1) dr]
  double dr, drinv, rcut; // grid spacing, cutoff radius (<= grid end)
  const double* vpoly;    // [nch][ng - 1][4] v_l(r) = a + t (b + t (c + t d)) per interval
  int nknot;
  const double* knots;    // [nknot][4] unit vector, weight
  int kc;                 // quadrature points evaluated together by the kernel (6 or 12)
};
// Non-local moves (NonLocalECPotential::makeNonLocalMovesPbyP v0 with NonLocalTOperator::selectMove):
// every quadrature point of every electron-ion pair is a candidate move r_e -> r_e + deltaV_j with
// weight W_j = knot partial product x ratio, scanned with W plus (W > 0) or W minus (W <= 0).
// mode 0: energy only; 1: energy and, per electron, the scan weight of its points (tsum[w][N],
// -1 for an electron without pairs); 2: re-evaluate electron sel_iel[w] of each walker (grid of
// one electron) and pick the last point whose scan value (from sel_base[w]) is <= sel_target[w]:
// sel_out[w][4] = {found, deltaV}.
struct TmoveArgs
{
  int mode;
  double plus, minus;
  double* tsum;
  const int* sel_iel;
  const double* sel_base;
  const double* sel_target;
  double* sel_out;
};
// Pair-list form of modes 0 and 1: the electron-ion pairs of all walkers within the cutoff,
// counted and sorted by ion on the device, then evaluated by a persistent grid in which the
// workgroups of one XCD take consecutive pairs, so that the pairs of one ion run together there and
// their quadrature points gather the table around that ion from the XCD's L2 and the MALL.  Slot
// (w N + e) ke + k holds the k-th ion (in ion order) within the cutoff of electron e of walker w;
// ke bounds the ions within the cutoff of any one point (1 + the most ions within 2 r_cut of one
// ion, from the fixed ion positions).
struct NlppPairs
{
  int ke;
  int grid;      // workgroups of the persistent pair kernel (a multiple of 8)
  int* icnt;     // [nion] pairs of each ion, then the scatter cursors
  int* total;    // [1] pairs of all walkers
  int* ecnt;     // [nw N] ions within the cutoff of each electron (> ke: reported as a NaN energy)
  int* pion;     // [nw N ke] ion of each slot
  int* list;     // [nw N ke] slots sorted by ion
  double* wres;  // [nw N ke] pair sums (mode 0)
  double* wpts;  // [nw N ke][nknot] point weights (mode 1)
};
// out[nw] = non-local energy of each walker; rot[nw][9] row-major rotation of the grid; part[nw][N]
// the per-electron sums.  pairs: the pair-list form for mode 0 (nullptr: one workgroup per electron)
hipError_t launch_nlpp(const SysC& s, const State& st, const void* coefs, const NlppC& c, int nw, const double* rot,
                       double* part, double* out, const TmoveArgs& tm, hipStream_t stream,
                       const NlppPairs* pairs = nullptr);

// per-walker slices of the walker state (base of slot 0 and bytes per slot, a multiple of 16)
constexpr int WA_MAX = 8;
struct WalkerArrays
{
  char* base[WA_MAX];
  size_t bytes[WA_MAX];
  int n;
};
// slot dst <- slot src for each pair {src, dst} of pairs[npairs][2] (distinct sources and
// destinations within one launch)
hipError_t launch_copy_walkers(const WalkerArrays& t, const int* pairs, int npairs, hipStream_t stream);

// Ewald tables of the periodic Coulomb energy (built on the host, device pointers)
constexpr int COUL_MAX_NG = 1024;
constexpr int COUL_MAX_N  = 1024; // electrons staged in LDS by the structure-factor kernel
struct CoulombC
{
  const double* kvec; // [3][nk] one k of every pair (k, -k) with |k| <= k_c, k != 0
  const double* fk;   // [nk] 2 F_k (the pair's two terms are equal for real charges)
  const double* rhoi; // [2][nk] rho_k of the ions (charges included): real, imaginary parts
  const double* poly; // [ng - 1][4] r v_sr(r) = a + t (b + t (c + t d)) on [i dr, (i + 1) dr]
  int nk, ng;
  double dr, drinv, rmax, zion;
  double const_ee, const_ei, vii; // constant terms of V_ee and V_eI, the ion-ion energy
};
// out[nw][3] = {V_ee, V_eI, V_II}
hipError_t launch_coulomb(const SysC& s, const State& st, const CoulombC& c, int nw, double* out, hipStream_t stream);
// rhok[nw][2][nk][2]: structure factor of each spin group at wave vectors kv[nk][3]
hipError_t launch_rhok(const SysC& s, const State& st, int nw, int nk, const double* kv, double* out, hipStream_t stream);

// TrialWaveFunction::mw_evaluateRatios for virtual moves of electron iat: pos [nw][nvp][3] ->
// out [nw][nvp] (complex pairs for complex orbitals); no delayed update pending for iat's group
hipError_t launch_api_vp_ratios(const SysC& s, const State& st, const void* coefs, int nw, int iat, int nvp,
                                const double* pos, double* out, hipStream_t stream);
hipError_t launch_api_vp_ratios_c(const SysC& s, const State& st, const void* coefs, int nw, int iat, int nvp,
                                  const double* pos, double* out, hipStream_t stream);
hipError_t launch_api_distances(const SysC& s, const State& st, int nw, const double* pos, double* ee_r,
                                double* ee_dr, double* ei_r, double* ei_dr, hipStream_t stream);

// ---- learned proposal (no counterpart in the reference) -----------------------------------------
constexpr int MLP_IN = 68, MLP_HID = 128, MLP_OUT = 4, MLP_K1 = 96; // K1: input width padded for MFMA
constexpr int N_ION_NB = 4, N_EL_NB = 16;                          // feature neighbour counts
// flat FP32 parameter vector: W1 [128][68], b1, W2 [128][128], b2, W3 [128][128], b3, W4 [4][128], b4
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_HID * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_HID * MLP_HID;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_HID;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_HID * MLP_HID;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_HID;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_OUT * MLP_HID;
constexpr int MLP_NPARAM = MLP_OFF_B4 + MLP_OUT;

struct MlpParams
{
  const float* master;              // FP32 parameters (flat layout above)
  uint16_t *w1b, *w2b, *w3b, *w4b;  // BF16 copies: W1 [128][96], W2/W3 [128][128], W4 [16][128]
};

// per-walker state of the learned move between its kernels
enum LearnField
{
  LS_R = 0,      // 3 position of the moving electron
  LS_V = 3,      // 3 drift at r
  LS_GNOW = 6,   // 3 total gradient at r
  LS_RP = 9,     // 3 proposed position
  LS_XI = 12,    // 3 standard normal of the move
  LS_LQF = 15,   // log q_fwd
  LS_VR = 16,    // 3 drift at r'
  LS_GNEW = 19,  // 3 total gradient at r'
  LS_SIZE = 24
};
// mu[3], s, mu'[3], s', log q_fwd, log q_rev, acceptance probability, objective f, drift at r'[3]
constexpr int LTRACE_FIELDS = 15;

struct LearnArgs
{
  float* X;        // [2 nw][68] features at r (rows [0, nw)) and at r' (rows [nw, 2 nw))
  float* out;      // [2 nw][4] network outputs for those rows
  double* ls;      // [nw][LS_SIZE]
  float* dout;     // [2 nw][4] gradient of the training loss with respect to out (nullable)
  double* fsum;    // [nw][2] running sums of the objective f and of the acceptance probability
  double* ltrace;  // [N][nw][LTRACE_FIELDS] (nullable)
};

hipError_t launch_mlp_pack(const MlpParams& P, hipStream_t stream);
// X [B][68] -> out [B][4]; z (nullable): pre-activations, layer l of sample b at
// z[(l * zstride + zoff + b) * 128]
hipError_t launch_mlp_forward(const MlpParams& P, int B, const float* X, float* out, float* z, int zstride, int zoff,
                              hipStream_t stream);
// gradient of sum_b dOut[b] . out[b] into grad[MLP_NPARAM]; z from a forward with zstride = B, zoff = 0;
// dz_a, dz_b: [B][128] scratch
hipError_t launch_mlp_backward(const MlpParams& P, int B, const float* X, const float* z, const float* dOut,
                               float* grad, float* dz_a, float* dz_b, hipStream_t stream);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, int t, hipStream_t stream);

enum LearnKernel
{
  LK_PRE = 0,  // inverse row, gradient, drift and features at r
  LK_PROP,     // proposal from the network output, ratio / gradient at r', features at r'
  LK_ACCEPT,   // asymmetric Metropolis-Hastings, objective gradient, accept / reject
  LK_FEATURES, // features only (API test hook; position in LS_R, drift in LS_V)
  LK_FUSED     // the whole move with a frozen network in one kernel (network evaluated per walker)
};
hipError_t launch_learn(int which, const SysC& s, const State& st, const void* coefs, const MoveArgs& a,
                        const LearnArgs& la, hipStream_t stream, const MlpParams* P = nullptr);
} // namespace qmi
