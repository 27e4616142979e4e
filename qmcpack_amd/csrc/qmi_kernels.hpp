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
  int prepare;    // API ratio kernel: prepare the inverse row first
  double tau;
  const double* gauss; // [N][nw][3] standard normals of the sub-step
  const double* unif;  // [ncrowd][N*wpc] uniforms of the sub-step (drawn lazily in order)
  int* ubase;          // [ncrowd][N+1] uniforms consumed before each move
  const int* host_acc; // [nw] accept decisions (API accept)
  double* out;         // API outputs
  double* trace;       // [N][nw][TRACE_FIELDS] (nullable)
};

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
// orbital rows of determinant g: values transposed into the FP64 LU workspace lu ([nw][n][ld]; Ainv
// itself in FP64 contexts), gradients/Laplacians into dpsi
hipError_t launch_spo_rows(const SysC& s, const State& st, const void* coefs, int g, int nw, double* lu,
                           hipStream_t stream);
hipError_t launch_logdet(const SysC& s, const State& st, int g, int nw, const int* ipiv, const double* lu,
                         hipStream_t stream);
// FP64 inverse in lu -> Ainv in the determinant storage type (mixed precisions)
hipError_t launch_store_inverse(const SysC& s, const State& st, int g, int nw, const double* lu, hipStream_t stream);
hipError_t launch_jastrow_recompute(const SysC& s, const State& st, int nw, hipStream_t stream);
// per-walker outputs: out3[nw][3] = {ekin, logpsi, phase}; G[nw][N][3], L[nw][N] (nullable)
hipError_t launch_eval_gl(const SysC& s, const State& st, int nw, double* out3, double* G, double* L,
                          hipStream_t stream);

// delayed-update flush of determinant g with d delayed rows (Ainv -= (Ainv U^T - W) Binv V)
hipError_t launch_flush(const SysC& s, const State& st, int g, int d, int nw, double* scratchT, hipStream_t stream);

// API helpers
hipError_t launch_api_spo_vgl(const SysC& s, const void* coefs, int nw, int n, const double* pos, double* out,
                              hipStream_t stream);
hipError_t launch_api_det_ratios(const SysC& s, const void* coefs, int nw, int nvp, int n, const double* pos,
                                 const double* inv, double* out, hipStream_t stream);
hipError_t launch_api_distances(const SysC& s, const State& st, int nw, const double* pos, double* ee_r,
                                double* ee_dr, double* ei_r, double* ei_dr, hipStream_t stream);
} // namespace qmi
