// Host side of the qmcmi C ABI: context, device memory, the multi-walker API and the
// device-resident VMC driver.  All device work of a context runs on one HIP stream.
#include "../../include/qmcmi.h"
#include "qmi_breakup.hpp"
#include "qmi_checkpoint.hpp"
#include "qmi_kernels.hpp"
#include "qmi_rng.hpp"

#include <algorithm>
#include <array>
#include <deque>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

using namespace qmi;

namespace
{
struct QmiError : std::runtime_error
{
  int code;
  QmiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                                      \
  do                                                                                                        \
  {                                                                                                         \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess)                                                                                   \
      throw QmiError(QMI_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));                    \
  } while (0)

template<typename T>
struct DevBuf
{
  T* p     = nullptr;
  size_t n = 0;
  void alloc(size_t count)
  {
    free();
    if (count == 0)
      return;
    HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void free()
  {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
};

template<typename T>
struct PinnedBuf
{
  T* p     = nullptr;
  size_t n = 0;
  void alloc(size_t count)
  {
    if (count <= n)
      return;
    free();
    HIPCHECK(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void free()
  {
    if (p)
      (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  ~PinnedBuf() { free(); }
};

// pool of timing events
struct Timing
{
  bool on = false;
  struct Rec
  {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> free_ev;
  std::vector<std::string> names;
  std::map<std::string, int> ids;
  std::vector<double> ms, calls;

  hipEvent_t get()
  {
    if (!free_ev.empty())
    {
      hipEvent_t e = free_ev.back();
      free_ev.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  int id(const char* name)
  {
    auto it = ids.find(name);
    if (it != ids.end())
      return it->second;
    const int i = int(names.size());
    names.push_back(name);
    ids[name] = i;
    ms.push_back(0);
    calls.push_back(0);
    return i;
  }
  void resolve()
  {
    for (auto& r : pending)
    {
      float t = 0;
      HIPCHECK(hipEventSynchronize(r.b));
      HIPCHECK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.id] += t;
      calls[r.id] += 1;
      free_ev.push_back(r.a);
      free_ev.push_back(r.b);
    }
    pending.clear();
  }
  ~Timing()
  {
    for (auto e : free_ev)
      (void)hipEventDestroy(e);
    for (auto& r : pending)
    {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
  }
};

} // namespace

struct qmi_ctx
{
  int device     = 0;
  int precision  = QMI_PREC_FP64;
  hipStream_t stream = nullptr;
  std::string err;
  bool have_system = false;
  bool skew_cell   = false; // qmi_system_set_lattice gave a cell with off-diagonal vectors
  bool cplx        = false; // complex (C2C) orbitals
  SysC sys{};
  DevBuf<double> coefs, ions, kpts;
  // walkers
  int nw = 0, wpc = 1, ncrowd = 0;
  int cap = 0; // walker slots allocated (>= nw)
  DevBuf<double> R, jas, logdet, stats, invRow, phi, mv, flushT, out, gl3, Gbuf, Lbuf, hostin;
  // determinant engine, in elements of the determinant storage type (double or float)
  DevBuf<double> Ainv[2], dpsi[2], U[2], V[2], Binv[2];
  // recompute: FP64 LU factors and the solve's work array ([nw][n][ld] values), unit rows of P e_k
  DevBuf<double> lu, luw;
  DevBuf<int> dlist[2], ipiv, lperm, info, ubase, acc;
  DevBuf<int> nanflag; // [cap] device NaN guard: 1 + the first electron with a NaN ratio or gradient
  PinnedBuf<int> h_nan;
  DevBuf<double> gauss, unif, trace;
  State st{};
  int dcount[2] = {0, 0};
  int invrow_id = -1;
  // fused move kernel: 1 = one launch per move (k_move_one), 0 = one launch per delay window
  // (k_sweep_segment).  QMI_MOVE_KERNEL=segment selects the latter.
  int move_one = 1;
  int flush_variant = 0; // 0: matrix-core flush for n <= 256; 1: three LDS-tiled passes
  // sub-step replayed from a captured HIP graph (no trace, no learned proposal)
  bool graph_on              = false;
  hipGraphExec_t sweep_exec  = nullptr;
  hipGraphExec_t split_exec[2] = {nullptr, nullptr};
  double graph_tau           = 0;
  hipEvent_t ev_done         = nullptr;
  PinnedBuf<int> h_ubase;
  // generation of the device buffers and knobs baked into captured graphs: every change that
  // invalidates a graph (reallocation, tuning, new walkers) increments it
  unsigned long long graph_gen = 0;
  hipEvent_t ev_ubase = nullptr; // recorded when a sub-step's uniform counts have reached the host
  // Walker partitions of the fused sweep: the moves of walkers [p nw / P, (p + 1) nw / P) run on
  // pstream[p], forked from and joined to the context stream, so that the single-move launches of
  // different partitions overlap on the compute units instead of all walkers passing through the
  // same latency-bound phase at once.  Results are those of one partition (each walker's arithmetic
  // is unchanged).
  int npart = 1;
  std::vector<hipStream_t> pstream;
  std::vector<hipEvent_t> pev_moves, pev_done;
  hipEvent_t pev_fork = nullptr;
  long long step_count = 0;
  // RNG: one stream per crowd
  std::vector<std::mt19937> rng;
  uint32_t rng_seed   = 0;
  bool rng_seeded     = false;
  PinnedBuf<double> h_gauss, h_unif;
  Timing timing;
  // learned proposal (no counterpart in the reference)
  bool learned = false, train = false, mlp_ready = false;
  // frozen network: the whole learned move in one kernel (QMI_LEARN_FUSED=0 selects the five-launch
  // form used for training)
  bool fused_learn = true;
  float lr     = 1e-3f;
  int adam_t   = 0;
  DevBuf<float> mlp_master, mlp_grad, adam_m, adam_v, lX, lout, ldout, lz, ldz_a, ldz_b;
  DevBuf<uint16_t> w1b, w2b, w3b, w4b;
  DevBuf<double> lstate, lfsum, ltrace_d;
  double learn_moves = 0;
  bool ltrace_on     = false;

  // -------------------------------------------------------------------------------------------
  struct Scoped
  {
    qmi_ctx* c;
    int id;
    hipEvent_t a;
    Scoped(qmi_ctx* ctx, const char* name) : c(ctx), id(-1), a(nullptr)
    {
      if (c->timing.on)
      {
        id = c->timing.id(name);
        a  = c->timing.get();
        HIPCHECK(hipEventRecord(a, c->stream));
      }
    }
    ~Scoped()
    {
      if (c->timing.on && a)
      {
        hipEvent_t b = c->timing.get();
        (void)hipEventRecord(b, c->stream);
        c->timing.pending.push_back({id, a, b});
      }
    }
  };

  void sync()
  {
    HIPCHECK(hipStreamSynchronize(stream));
    timing.resolve();
  }

  // an executable graph is destroyed only after every launch of it has completed: its launches run on
  // this context's stream (alone, or as a branch forked from a batch's origin stream, which joins it)
  void invalidate_graph()
  {
    if (sweep_exec || split_exec[0] || split_exec[1])
      (void)hipStreamSynchronize(stream);
    if (sweep_exec)
      (void)hipGraphExecDestroy(sweep_exec);
    sweep_exec = nullptr;
    for (auto& x : split_exec)
    {
      if (x)
        (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
    graph_gen++;
  }

  // distance tables over this context's walkers: detached (their context pointer cleared) when the
  // context is destroyed first, so that destroying them later touches no freed context
  std::vector<qmi_dtable*> dtables;
  void detach_dtables();

  void free_partitions()
  {
    for (auto s_ : pstream)
    {
      (void)hipStreamSynchronize(s_);
      (void)hipStreamDestroy(s_);
    }
    for (auto e : pev_moves)
      (void)hipEventDestroy(e);
    for (auto e : pev_done)
      (void)hipEventDestroy(e);
    if (pev_fork)
      (void)hipEventDestroy(pev_fork);
    pstream.clear();
    pev_moves.clear();
    pev_done.clear();
    pev_fork = nullptr;
  }

  void ensure_partitions(int P)
  {
    if ((int)pstream.size() >= P)
      return;
    if (!pev_fork)
      HIPCHECK(hipEventCreateWithFlags(&pev_fork, hipEventDisableTiming));
    while ((int)pstream.size() < P)
    {
      hipStream_t s_ = nullptr;
      hipEvent_t e1 = nullptr, e2 = nullptr;
      HIPCHECK(hipStreamCreate(&s_));
      pstream.push_back(s_);
      HIPCHECK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
      pev_moves.push_back(e1);
      HIPCHECK(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
      pev_done.push_back(e2);
    }
  }

  // the partitions of the next fused sub-step: 1 unless one walker per crowd runs the single-move
  // kernel without the learned proposal and every partition keeps at least 64 walkers
  int partitions() const
  {
    if (npart <= 1 || wpc != 1 || learned || !(move_one || cplx))
      return 1;
    return std::max(1, std::min(npart, nw / 64));
  }

  // walkers [w0, ...) as a population of their own: every per-walker array advanced by w0 walkers
  State state_from(int w0) const
  {
    State t  = st;
    auto adv = [&](auto*& p, size_t bytes) {
      if (p)
        p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(reinterpret_cast<char*>(p) + bytes * w0);
    };
    const size_t es = det_esize(), ld = sys.ld, k = sys.k, Np = sys.Npad, v = vw();
    adv(t.R, 3 * Np * sizeof(double));
    adv(t.jas, J_ROWS * Np * sizeof(double));
    adv(t.logdet, 4 * sizeof(double));
    adv(t.stats, 2 * sizeof(double));
    adv(t.invRow, ld * v * sizeof(double));
    adv(t.phi, 5 * ld * v * sizeof(double));
    adv(t.mv, MV_SIZE * sizeof(double));
    for (int g = 0; g < 2; g++)
    {
      const size_t n = sys.n[g];
      adv(t.Ainv[g], n * ld * es);
      adv(t.dpsi[g], n * 4 * ld * es);
      adv(t.U[g], ld * k * es);
      adv(t.V[g], ld * k * es);
      adv(t.Binv[g], k * k * es);
      adv(t.dlist[g], k * sizeof(int));
    }
    return t;
  }

  // flush of spin group g at delay d for walkers [w0, w0 + wn) on stream s_
  void flush_part(int g, int d, int w0, int wn, hipStream_t s_)
  {
    if (sys.k == 1 || d == 0)
      return;
    const State t    = state_from(w0);
    double* scratchT = flushT.p ? flushT.p + (size_t)w0 * sys.ld * sys.k * vw() : nullptr;
    HIPCHECK(cplx ? launch_flush_c(sys, t, g, d, wn, scratchT, s_)
                  : launch_flush(sys, t, g, d, wn, scratchT, s_, flush_variant));
  }

  ~qmi_ctx()
  {
    free_partitions();
    invalidate_graph();
    if (ev_done)
      (void)hipEventDestroy(ev_done);
    if (ev_ubase)
      (void)hipEventDestroy(ev_ubase);
  }

  int group(int iat) const { return iat < sys.nup ? 0 : 1; }

  // bytes per element of the spline table and of the determinant engine
  size_t table_esize() const
  {
    switch (precision)
    {
    case QMI_PREC_FP64: return 8;
    case QMI_PREC_MIXED_FP32:
    case QMI_PREC_TABLE_FP32: return 4;
    default: return 2;
    }
  }
  bool det_fp32() const { return precision == QMI_PREC_MIXED_FP32 || precision == QMI_PREC_MIXED_BF16; }
  size_t det_esize() const { return cplx ? 16 : det_fp32() ? 4 : 8; }
  // doubles per value crossing the interface (ratios, gradients, G, L, orbitals)
  int vw() const { return cplx ? 2 : 1; }

  hipError_t move_kernel(int which, const MoveArgs& a)
  {
    return cplx ? launch_move_c(which, sys, st, coefs.p, a, stream) : launch_move(which, sys, st, coefs.p, a, stream);
  }
  // device buffer (in doubles) holding count elements of the determinant type
  size_t det_words(size_t count) const { return (count * det_esize() + 7) / 8; }

  void require_walkers() const
  {
    if (!have_system || nw == 0)
      throw QmiError(QMI_ERR_ARG, "system and walkers must be loaded first");
  }

  // ---------------------------------------------------------------------------------------------
  // Cell vectors R (rows) of the orbitals' lattice: reduced coordinates u = r R^-1, orbital
  // gradients G g and Laplacians sum_ij H_ij (G^T G)_ij with G = R^-1 (CrystalLattice::toUnit,
  // SplineR2R.cpp:338-374 assign_vgl).  A cell with off-diagonal vectors serves the orbitals only:
  // the Jastrow, Coulomb, pseudopotential and distance-table paths take minimum images in an
  // orthorhombic cell and refuse it.
  void set_lattice(const double* R)
  {
    if (!have_system)
      throw QmiError(QMI_ERR_ARG, "system not created");
    if (!R)
      throw QmiError(QMI_ERR_ARG, "null lattice");
    const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
        R[2] * (R[3] * R[7] - R[4] * R[6]);
    if (!(std::fabs(det) > 0) || !std::isfinite(det))
      throw QmiError(QMI_ERR_ARG, "singular lattice");
    bool diag = true;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        if (i != j && R[i * 3 + j] != 0.0)
          diag = false;
    if (!diag && (sys.has_j1 || sys.has_j2))
      throw QmiError(QMI_ERR_UNSUPPORTED, "Jastrow distances are built for orthorhombic cells");
    if (coul_ready || nlpp_ready || !dtables.empty())
      throw QmiError(diag ? QMI_ERR_ARG : QMI_ERR_UNSUPPORTED,
                     "set the lattice before the Coulomb, pseudopotential and distance-table set-up");
    if (diag && !(R[0] > 0 && R[4] > 0 && R[8] > 0))
      throw QmiError(QMI_ERR_ARG, "diagonal lattice needs positive sides");
    // G = R^-1 by cofactors
    double G[9];
    G[0] = (R[4] * R[8] - R[5] * R[7]) / det;
    G[1] = (R[2] * R[7] - R[1] * R[8]) / det;
    G[2] = (R[1] * R[5] - R[2] * R[4]) / det;
    G[3] = (R[5] * R[6] - R[3] * R[8]) / det;
    G[4] = (R[0] * R[8] - R[2] * R[6]) / det;
    G[5] = (R[2] * R[3] - R[0] * R[5]) / det;
    G[6] = (R[3] * R[7] - R[4] * R[6]) / det;
    G[7] = (R[1] * R[6] - R[0] * R[7]) / det;
    G[8] = (R[0] * R[4] - R[1] * R[3]) / det;
    double GtG[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        GtG[i * 3 + j] = G[0 * 3 + i] * G[0 * 3 + j] + G[1 * 3 + i] * G[1 * 3 + j] + G[2 * 3 + i] * G[2 * 3 + j];
    std::copy(G, G + 9, sys.G);
    const double symGG[6] = {GtG[0], GtG[1] + GtG[3], GtG[2] + GtG[6], GtG[4], GtG[5] + GtG[7], GtG[8]};
    std::copy(symGG, symGG + 6, sys.symGG);
    if (diag)
      for (int i = 0; i < 3; i++)
      {
        sys.periodic[i] = 1;
        sys.L[i]        = R[i * 4];
        sys.Linv[i]     = 1.0 / R[i * 4];
      }
    skew_cell = !diag;
    invalidate_graph();
  }

  void system_create(const qmi_system_desc* d)
  {
    if (!d)
      throw QmiError(QMI_ERR_ARG, "null system description");
    skew_cell = false;
    SysC s{};
    s.nup  = d->n_up;
    s.ndn  = d->n_dn;
    s.N    = s.nup + s.ndn;
    s.nion = d->n_ion;
    if (s.nup < 0 || s.ndn < 0 || s.N <= 0 || s.N > 1024)
      throw QmiError(QMI_ERR_ARG, "electron counts must satisfy 0 < N <= 1024");
    s.first[0] = 0;
    s.first[1] = s.nup;
    s.n[0]     = s.nup;
    s.n[1]     = s.ndn;
    s.Npad     = (s.N + 31) / 32 * 32;
    const int nmax = std::max(s.nup, s.ndn);
    if (nmax > MAX_N)
      throw QmiError(QMI_ERR_UNSUPPORTED, "more than 512 electrons per spin");
    s.ld = std::max(8, (nmax + 7) / 8 * 8);
    s.k  = std::max(1, d->delay_rank);
    if (s.k > MAX_K || (s.k > 1 && s.k > nmax))
      throw QmiError(QMI_ERR_ARG, "delay rank must be in [1, 64] and not exceed the determinant size");
    for (int i = 0; i < 3; i++)
    {
      s.periodic[i] = d->L[i] > 0;
      s.L[i]        = s.periodic[i] ? d->L[i] : 1.0;
      s.Linv[i]     = 1.0 / s.L[i];
    }
    std::fill(s.G, s.G + 9, 0.0);
    double GGt[9] = {0};
    for (int i = 0; i < 3; i++)
    {
      s.G[i * 4]  = s.Linv[i];
      GGt[i * 4] = s.Linv[i] * s.Linv[i];
    }
    const double symGG[6] = {GGt[0], GGt[1] + GGt[3], GGt[2] + GGt[6], GGt[4], GGt[5] + GGt[7], GGt[8]};
    std::copy(symGG, symGG + 6, s.symGG);
    s.prec    = precision;
    s.has_det = d->norb > 0 && d->coefs != nullptr;
    if (!s.has_det)
      s.k = 1; // the delay rank is the determinants' (no delayed-update buffers without them)
    const bool zc = s.has_det && d->complex_orbitals;
    if (zc && precision != QMI_PREC_FP64)
      throw QmiError(QMI_ERR_UNSUPPORTED, "complex orbitals need the FP64 precision mode");
    if (zc && !d->kpoints)
      throw QmiError(QMI_ERR_ARG, "complex orbitals need k-points");
    s.cplx = zc ? 1 : 0;
    // spline columns: norb, or 2 norb (real and imaginary parts adjacent) for complex orbitals
    const int ncol = zc ? 2 * d->norb : d->norb;
    if (s.has_det)
    {
      if (d->norb < nmax)
        throw QmiError(QMI_ERR_ARG, "norb smaller than the electrons per spin");
      for (int i = 0; i < 3; i++)
      {
        s.M[i] = d->M[i];
        if (s.M[i] < 4)
          throw QmiError(QMI_ERR_ARG, "spline grid needs at least 4 points per dimension");
        const double delta = (1.0 - 0.0) / (double)(s.M[i]);
        s.dinv[i]          = 1.0 / delta;
      }
      s.npad = (ncol + 7) / 8 * 8;
      s.zs   = s.npad;
      s.ys   = (long long)(s.M[2] + 3) * s.npad;
      s.xs   = (long long)(s.M[1] + 3) * s.ys;
    }
    auto functor = [](FunctorC& f, int np, const double* P, double rcut, double cusp) {
      if (np < 2 || np + 4 > MAX_JCOEF)
        throw QmiError(QMI_ERR_ARG, "functor needs 2 <= NumParams <= 64");
      std::memset(&f, 0, sizeof(f));
      const int numCoefs = np + 4, numKnots = numCoefs - 2;
      const double DeltaR = rcut / (double)(numKnots - 1);
      f.DeltaRInv         = 1.0 / DeltaR;
      f.rcut              = rcut;
      f.np                = np;
      f.max_index         = numCoefs - 4;
      f.c[1]              = P[0];
      f.c[2]              = P[1];
      f.c[0]              = P[1] - 2.0 * DeltaR * cusp;
      for (int i = 2; i < np; i++)
        f.c[i + 1] = P[i];
    };
    s.has_j1 = d->j1_np > 0 && d->n_ion > 0 && d->j1_params;
    if (s.has_j1)
      functor(s.j1, d->j1_np, d->j1_params, d->j1_rcut, d->j1_cusp);
    s.has_j2 = d->j2_np > 0 && d->j2_uu && d->j2_ud;
    if (s.has_j2)
    {
      functor(s.j2[0], d->j2_np, d->j2_uu, d->j2_rcut, d->j2_cusp_uu);
      functor(s.j2[1], d->j2_np, d->j2_ud, d->j2_rcut, d->j2_cusp_ud);
    }
    // device copies
    ions.alloc(std::max(3 * s.nion, 3));
    if (s.nion > 0)
    {
      std::vector<double> t(3 * s.nion);
      for (int j = 0; j < s.nion; j++)
        for (int q = 0; q < 3; q++)
          t[q * s.nion + j] = d->ions[3 * j + q];
      HIPCHECK(hipMemcpy(ions.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (s.has_det)
    {
      const size_t nuniq = (size_t)s.M[0] * s.M[1] * s.M[2] * ncol;
      const size_t ntab  = (size_t)(s.M[0] + 3) * (s.M[1] + 3) * (s.M[2] + 3) * s.npad;
      DevBuf<double> tmp;
      tmp.alloc(nuniq);
      HIPCHECK(hipMemcpy(tmp.p, d->coefs, nuniq * sizeof(double), hipMemcpyHostToDevice));
      coefs.alloc((ntab * table_esize() + 7) / 8);
      HIPCHECK(launch_wrap_coefs(s, tmp.p, ncol, coefs.p, stream));
      HIPCHECK(hipStreamSynchronize(stream));
    }
    s.kld = std::max(s.ld, (std::max(d->norb, 1) + 7) / 8 * 8);
    kpts.alloc((size_t)3 * s.kld);
    {
      std::vector<double> t((size_t)3 * s.kld, 0.0);
      if (zc)
        for (int j = 0; j < d->norb; j++)
          for (int q = 0; q < 3; q++)
            t[(size_t)q * s.kld + j] = d->kpoints[3 * j + q];
      HIPCHECK(hipMemcpy(kpts.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    cplx        = zc;
    sys         = s;
    have_system = true;
    nw          = 0;
    invalidate_graph();
    coul_ready = false;
    nlpp_ready = false;
  }

  // ---------------------------------------------------------------------------------------------
  // capacity: walker slots to allocate (>= nw; DMC populations grow into them)
  void walkers_load(int nw_in, const double* Rin, int wpc_in, int capacity = 0)
  {
    if (!have_system)
      throw QmiError(QMI_ERR_ARG, "system not created");
    if (nw_in <= 0 || !Rin)
      throw QmiError(QMI_ERR_ARG, "need at least one walker");
    const SysC& s = sys;
    nw            = nw_in;
    cap           = std::max(nw, capacity);
    wpc           = std::max(1, std::min(wpc_in, nw));
    ncrowd        = (nw + wpc - 1) / wpc;
    dmc.on        = false;
    const int Np = s.Npad, ld = s.ld, k = s.k;
    R.alloc((size_t)cap * 3 * Np);
    {
      std::vector<double> t((size_t)cap * 3 * Np, 0.0);
      for (int w = 0; w < nw; w++)
        for (int i = 0; i < s.N; i++)
          for (int q = 0; q < 3; q++)
            t[((size_t)w * 3 + q) * Np + i] = Rin[((size_t)w * s.N + i) * 3 + q];
      HIPCHECK(hipMemcpy(R.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    jas.alloc((size_t)cap * J_ROWS * Np);
    HIPCHECK(hipMemset(jas.p, 0, jas.n * sizeof(double)));
    logdet.alloc((size_t)cap * 4);
    HIPCHECK(hipMemset(logdet.p, 0, logdet.n * sizeof(double)));
    stats.alloc((size_t)cap * 2);
    HIPCHECK(hipMemset(stats.p, 0, stats.n * sizeof(double)));
    nanflag.alloc(cap);
    HIPCHECK(hipMemset(nanflag.p, 0, nanflag.n * sizeof(int)));
    h_nan.alloc(cap);
    mv.alloc((size_t)cap * MV_SIZE);
    HIPCHECK(hipMemset(mv.p, 0, mv.n * sizeof(double)));
    out.alloc((size_t)cap * 8 + 16);
    gl3.alloc((size_t)cap * 3);
    hostin.alloc((size_t)cap * 3 + 16);
    acc.alloc(cap);
    ubase.alloc((size_t)((cap + wpc - 1) / wpc) * (s.N + 1));
    lstate.alloc((size_t)cap * LS_SIZE);
    lfsum.alloc((size_t)cap * 2);
    HIPCHECK(hipMemset(lfsum.p, 0, lfsum.n * sizeof(double)));
    learn_moves = 0;
    if (s.has_det)
    {
      invRow.alloc((size_t)cap * ld * vw());
      phi.alloc((size_t)cap * 5 * ld * vw());
      for (int g = 0; g < 2; g++)
      {
        const int n = s.n[g];
        Ainv[g].alloc(det_words((size_t)cap * std::max(n, 1) * ld));
        dpsi[g].alloc(det_words((size_t)cap * std::max(n, 1) * 4 * ld));
        U[g].alloc(det_words((size_t)cap * ld * k));
        V[g].alloc(det_words((size_t)cap * ld * k));
        Binv[g].alloc(det_words((size_t)cap * k * k));
        dlist[g].alloc((size_t)cap * k);
        HIPCHECK(hipMemset(U[g].p, 0, U[g].n * sizeof(double)));
        HIPCHECK(hipMemset(V[g].p, 0, V[g].n * sizeof(double)));
        HIPCHECK(hipMemset(Binv[g].p, 0, Binv[g].n * sizeof(double)));
      }
      lu.alloc((size_t)cap * std::max(s.nup, s.ndn) * ld * vw());
      luw.alloc((size_t)cap * std::max(s.nup, s.ndn) * ld * vw());
      ipiv.alloc((size_t)cap * std::max(s.nup, s.ndn));
      lperm.alloc((size_t)cap * std::max(s.nup, s.ndn));
      info.alloc((size_t)2 * cap);
      if (k > 1)
        flushT.alloc((size_t)cap * ld * k * vw()); // T (n x k) or BV_T (ld x k) per walker
    }
    st        = State{};
    st.R      = R.p;
    st.jas    = jas.p;
    st.logdet = logdet.p;
    st.stats  = stats.p;
    st.invRow = invRow.p;
    st.phi    = phi.p;
    st.mv     = mv.p;
    st.ions   = ions.p;
    st.kpts   = kpts.p;
    for (int g = 0; g < 2; g++)
    {
      st.Ainv[g]  = Ainv[g].p;
      st.dpsi[g]  = dpsi[g].p;
      st.U[g]     = U[g].p;
      st.V[g]     = V[g].p;
      st.Binv[g]  = Binv[g].p;
      st.dlist[g] = dlist[g].p;
    }
    dcount[0] = dcount[1] = 0;
    invrow_id             = -1;
    step_count            = 0;
    rng_seeded            = false;
    invalidate_graph();
  }

  // delayed-update rank of the engine (between sub-steps: no pending delayed rows); the delayed
  // block buffers are reallocated and cleared
  void set_delay_rank(int k)
  {
    require_walkers();
    const int nmax = std::max(sys.nup, sys.ndn);
    if (k < 1 || k > MAX_K || (k > 1 && k > nmax))
      throw QmiError(QMI_ERR_ARG, "delay rank must be in [1, 64] and not exceed the determinant size");
    if (dcount[0] || dcount[1])
      throw QmiError(QMI_ERR_ARG, "delay rank change with pending delayed updates");
    if (k == sys.k || !sys.has_det)
      return;
    sys.k = k;
    if (sys.has_det)
    {
      const int ld = sys.ld;
      for (int g = 0; g < 2; g++)
      {
        U[g].alloc(det_words((size_t)cap * ld * k));
        V[g].alloc(det_words((size_t)cap * ld * k));
        Binv[g].alloc(det_words((size_t)cap * k * k));
        dlist[g].alloc((size_t)cap * k);
        HIPCHECK(hipMemset(U[g].p, 0, U[g].n * sizeof(double)));
        HIPCHECK(hipMemset(V[g].p, 0, V[g].n * sizeof(double)));
        HIPCHECK(hipMemset(Binv[g].p, 0, Binv[g].n * sizeof(double)));
        st.U[g]     = U[g].p;
        st.V[g]     = V[g].p;
        st.Binv[g]  = Binv[g].p;
        st.dlist[g] = dlist[g].p;
      }
      if (k > 1)
        flushT.alloc((size_t)cap * ld * k * vw());
    }
    invrow_id = -1;
    invalidate_graph();
  }

  // ---------------------------------------------------------------------------------------------
  // determinant recompute: orbital rows, batched LU, log det, inverse (FP64 in every precision;
  // the FP32 engine stores the inverse rounded to FP32, DiracMatrix.h:204-220)
  void recompute_dets()
  {
    const SysC& s = sys;
    if (!s.has_det)
      return;
    if (dcount[0] || dcount[1])
      throw QmiError(QMI_ERR_ARG, "recompute with pending delayed updates");
    const int out_kind = cplx ? LU_OUT_C128 : det_fp32() ? LU_OUT_F32 : LU_OUT_F64;
    for (int g = 0; g < 2; g++)
    {
      const int n = s.n[g];
      if (n == 0)
        continue;
      {
        Scoped t(this, "spo_rows");
        HIPCHECK(cplx ? launch_spo_rows_c(s, st, coefs.p, g, nw, lu.p, stream)
                      : launch_spo_rows(s, st, coefs.p, g, nw, lu.p, stream));
      }
      {
        Scoped t(this, "lu_factor");
        HIPCHECK(launch_lu_factor(cplx, n, s.ld, nw, lu.p, ipiv.p, lperm.p, info.p + (size_t)g * cap, stream));
      }
      {
        Scoped t(this, "logdet");
        HIPCHECK(launch_lu_logdet(cplx, n, s.ld, nw, lu.p, ipiv.p, logdet.p + 2 * g, 4, stream));
      }
      {
        Scoped t(this, "lu_inverse");
        HIPCHECK(launch_lu_inverse(out_kind, n, s.ld, nw, lu.p, lperm.p, luw.p, Ainv[g].p, stream));
      }
    }
    check_info("LU factorisation: singular Slater matrix");
    invrow_id = -1;
  }

  // info of both determinants ([2][cap]): the first zero pivot of any walker is an error
  void check_info(const char* what)
  {
    std::vector<int> h((size_t)2 * cap);
    HIPCHECK(hipMemcpyAsync(h.data(), info.p, h.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
    sync();
    for (int g = 0; g < 2; g++)
      for (int w = 0; w < nw; w++)
        if (sys.n[g] > 0 && h[(size_t)g * cap + w] != 0)
          throw QmiError(QMI_ERR_NUMERIC, std::string(what) + " (walker " + std::to_string(w) + ", determinant " +
                                              std::to_string(g) + ")");
  }

  // mw_evaluateGL in two halves: the kernel and the copies (per-walker {ekin, log psi, phase} to
  // pinned memory) on the stream, then, after a synchronisation, the outputs
  PinnedBuf<double> h_gl3;
  void eval_gl_enqueue()
  {
    h_gl3.alloc((size_t)cap * 3);
    {
      Scoped t(this, "eval_gl");
      HIPCHECK(cplx ? launch_eval_gl_c(sys, st, nw, gl3.p, nullptr, nullptr, stream)
                    : launch_eval_gl(sys, st, nw, gl3.p, nullptr, nullptr, stream));
    }
    HIPCHECK(hipMemcpyAsync(h_gl3.p, gl3.p, (size_t)nw * 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    nan_enqueue();
  }

  // the device NaN guard of the moves since the last check, read once per step (NaNguard.cpp:18-31,
  // TrialWaveFunction.cpp:592, 736-739): copied to pinned memory on the stream, checked after a
  // synchronisation; a flagged walker clears the flags and raises QMI_ERR_NUMERIC
  void nan_enqueue()
  {
    HIPCHECK(hipMemcpyAsync(h_nan.p, nanflag.p, (size_t)nw * sizeof(int), hipMemcpyDeviceToHost, stream));
  }
  void nan_check()
  {
    for (int w = 0; w < nw; w++)
      if (h_nan.p[w] != 0)
      {
        const int iat = h_nan.p[w] - 1;
        HIPCHECK(hipMemsetAsync(nanflag.p, 0, nanflag.n * sizeof(int), stream));
        HIPCHECK(hipStreamSynchronize(stream));
        throw QmiError(QMI_ERR_NUMERIC, "NaN ratio or gradient of walker " + std::to_string(w) + " at particle " +
                                            std::to_string(iat) + " (TWF::mw_calcRatioGrad / mw_evalGrad)");
      }
  }

  void eval_gl_finish(double* ekin)
  {
    nan_check();
    for (int w = 0; w < nw; w++)
    {
      const double* h = h_gl3.p + (size_t)w * 3;
      if (ekin)
        ekin[w] = h[0];
      if (!std::isfinite(h[0]) || !std::isfinite(h[1]))
        throw QmiError(QMI_ERR_NUMERIC, "non-finite log psi or local energy (walker " + std::to_string(w) + ")");
    }
  }

  void eval_gl(double* logpsi, double* G, double* L, double* ekin)
  {
    const SysC& s = sys;
    double* dG    = nullptr;
    double* dL    = nullptr;
    const size_t nG = (size_t)nw * s.N * 3 * vw(), nL = (size_t)nw * s.N * vw();
    if (G)
    {
      Gbuf.alloc(nG);
      dG = Gbuf.p;
    }
    if (L)
    {
      Lbuf.alloc(nL);
      dL = Lbuf.p;
    }
    {
      Scoped t(this, "eval_gl");
      HIPCHECK(cplx ? launch_eval_gl_c(s, st, nw, gl3.p, dG, dL, stream) : launch_eval_gl(s, st, nw, gl3.p, dG, dL, stream));
    }
    std::vector<double> h((size_t)nw * 3);
    HIPCHECK(hipMemcpyAsync(h.data(), gl3.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (G)
      HIPCHECK(hipMemcpyAsync(G, dG, nG * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (L)
      HIPCHECK(hipMemcpyAsync(L, dL, nL * sizeof(double), hipMemcpyDeviceToHost, stream));
    nan_enqueue();
    sync();
    nan_check();
    for (int w = 0; w < nw; w++)
    {
      if (ekin)
        ekin[w] = h[w * 3 + 0];
      if (logpsi && cplx)
      {
        logpsi[2 * w]     = h[w * 3 + 1];
        logpsi[2 * w + 1] = h[w * 3 + 2];
      }
      else if (logpsi)
        logpsi[w] = h[w * 3 + 1];
      if (!std::isfinite(h[w * 3 + 0]) || !std::isfinite(h[w * 3 + 1]))
        throw QmiError(QMI_ERR_NUMERIC, "non-finite log psi or local energy (walker " + std::to_string(w) + ")");
    }
  }

  void flush(int g)
  {
    if (sys.k == 1 || dcount[g] == 0)
      return;
    {
      Scoped t(this, "flush");
      HIPCHECK(cplx ? launch_flush_c(sys, st, g, dcount[g], nw, flushT.p, stream)
                    : launch_flush(sys, st, g, dcount[g], nw, flushT.p, stream, flush_variant));
    }
    dcount[g] = 0;
    invrow_id = -1;
  }

  MoveArgs move_args(int iat)
  {
    MoveArgs a{};
    a.iat0 = iat;
    a.iat1 = iat + 1;
    a.g    = group(iat);
    a.d0   = sys.k > 1 ? dcount[a.g] : 0;
    a.nw   = nw;
    a.wpc  = wpc;
    a.nanflag = nanflag.p;
    return a;
  }

  // ---------------------------------------------------------------------------------------------
  // Periodic Coulomb tables (Ewald breakup, Particle/LongRange/EwaldHandler3D.{h,cpp} as used by
  // CoulombPBCAA / CoulombPBCAB): r_c = min L / 2, k_c = dim_cutoff / r_c, sigma = sqrt(k_c / 2 r_c);
  // k vectors 2 pi n / L with |k|^2 <= k_c^2 grouped in shells (KContainer.cpp:54-230); the
  // clamped cubic spline of r erfc(sigma r) / r on 1001 points (LRCoulombSingleton.cpp:137-158).
  bool coul_ready = false;
  CoulombC coul{};
  DevBuf<double> coul_k, coul_fk, coul_rhoi, coul_poly, coul_out;

  static void clamped_spline(const std::vector<double>& x, const std::vector<double>& y, double yp0, double ypn,
                             std::vector<double>& y2)
  {
    const int n = (int)x.size();
    std::vector<double> c(n), d(n);
    const double h0 = x[1] - x[0];
    c[0]            = 0.5;
    d[0]            = 3.0 * ((y[1] - y[0]) / h0 - yp0) / h0;
    for (int i = 1; i < n - 1; i++)
    {
      const double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
      const double den = 2.0 * (hl + hr) - hl * c[i - 1];
      c[i]             = hr / den;
      d[i]             = (6.0 * ((y[i + 1] - y[i]) / hr - (y[i] - y[i - 1]) / hl) - hl * d[i - 1]) / den;
    }
    const double hn = x[n - 1] - x[n - 2];
    y2[n - 1]       = (6.0 * (ypn - (y[n - 1] - y[n - 2]) / hn) - hn * d[n - 2]) / (2.0 * hn - hn * c[n - 2]);
    for (int i = n - 2; i >= 0; i--)
      y2[i] = d[i] - c[i] * y2[i + 1];
  }

  // handler QMI_LR_EWALD: erfc / Gaussian split (EwaldHandler3D); QMI_LR_OPT_BREAKUP: the fitted
  // LPQHI split (LRHandlerTemp<CoulombFunctor, LPQHIBasis>, built by lr_breakup). Both give the
  // device the same tables: F_k per k vector, r v_sr(r) on the spline grid, the constants.
  void coulomb_setup(double zion, double dim_cutoff, int handler = QMI_LR_EWALD)
  {
    if (!have_system)
      throw QmiError(QMI_ERR_ARG, "system not created");
    const SysC& s = sys;
    if (!s.periodic[0] || !s.periodic[1] || !s.periodic[2] || skew_cell)
      throw QmiError(QMI_ERR_UNSUPPORTED, "the periodic Coulomb energy needs a periodic orthorhombic cell");
    if (!(dim_cutoff > 0))
      throw QmiError(QMI_ERR_ARG, "dim_cutoff must be positive");
    if (handler != QMI_LR_EWALD && handler != QMI_LR_OPT_BREAKUP)
      throw QmiError(QMI_ERR_ARG, "unknown long-range handler");
    const bool opt   = handler == QMI_LR_OPT_BREAKUP;
    const double* L  = s.L;
    const double rc  = 0.5 * std::min(L[0], std::min(L[1], L[2]));
    const double kc  = dim_cutoff / rc;
    const double sig = std::sqrt(kc / (2.0 * rc));
    const double vol = L[0] * L[1] * L[2];
    LrBreakup brk;
    if (opt)
      brk = lr_breakup(L, dim_cutoff);
    // |k|^2 bound of the long-range sum (the fitted split keeps at least the first shell)
    const double kc2 = opt ? brk.kc2_lr : kc * kc;
    struct K
    {
      double x, y, z, k2;
      int i, j, k;
    };
    std::vector<K> ks;
    int mmax[3];
    for (int d = 0; d < 3; d++)
      mmax[d] = (int)std::floor(L[d] * std::sqrt(kc2) / (2 * M_PI)) + 1;
    for (int i = -mmax[0]; i <= mmax[0]; i++)
      for (int j = -mmax[1]; j <= mmax[1]; j++)
        for (int k = -mmax[2]; k <= mmax[2]; k++)
        {
          if (i == 0 && j == 0 && k == 0)
            continue;
          const double x = 2 * M_PI * i / L[0], y = 2 * M_PI * j / L[1], z = 2 * M_PI * k / L[2];
          const double k2 = x * x + y * y + z * z;
          if (k2 <= kc2)
            ks.push_back({x, y, z, k2, i, j, k});
        }
    std::stable_sort(ks.begin(), ks.end(),
                     [](const K& a, const K& b) { return std::llround(a.k2 * 1e8) < std::llround(b.k2 * 1e8); });
    // F_k by shell over the full list (the reference's grouping), then one k of each pair (k, -k):
    // for real charges rho_{-k} = conj(rho_k), so the pair contributes twice the kept term
    std::vector<double> fall(ks.size());
    {
      const double kgauss = 1.0 / (4 * sig * sig), knorm = 4 * M_PI / vol;
      long long key = -1;
      double fshell = 0;
      for (size_t k = 0; k < ks.size(); k++)
      {
        const long long kk = std::llround(ks[k].k2 * 1e8);
        if (kk != key)
        {
          key    = kk;
          fshell = opt ? brk.fk(std::sqrt(ks[k].k2)) : knorm * std::exp(-ks[k].k2 * kgauss) / ks[k].k2;
        }
        fall[k] = fshell;
      }
    }
    std::vector<K> kh;
    std::vector<double> fh;
    for (size_t k = 0; k < ks.size(); k++)
      if (ks[k].i > 0 || (ks[k].i == 0 && (ks[k].j > 0 || (ks[k].j == 0 && ks[k].k > 0))))
      {
        kh.push_back(ks[k]);
        fh.push_back(2.0 * fall[k]);
      }
    const int nk = (int)kh.size();
    std::vector<double> kv(3 * (size_t)nk), fk(nk), rhoi(2 * (size_t)nk, 0.0);
    std::vector<double> ions_h(3 * (size_t)s.nion);
    if (s.nion)
      HIPCHECK(hipMemcpy(ions_h.data(), ions.p, ions_h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < nk; k++)
    {
      kv[k]                  = kh[k].x;
      kv[nk + k]             = kh[k].y;
      kv[2 * (size_t)nk + k] = kh[k].z;
      fk[k]                  = fh[k];
      for (int a = 0; a < s.nion; a++)
      {
        const double ph = kh[k].x * ions_h[a] + kh[k].y * ions_h[s.nion + a] + kh[k].z * ions_h[2 * s.nion + a];
        rhoi[k] += zion * std::cos(ph);
        rhoi[nk + k] += zion * std::sin(ph);
      }
    }
    const int ng = 1001;
    std::vector<double> gx(ng), gy(ng), gy2(ng);
    const double dr = rc / (ng - 1);
    for (int i = 0; i < ng; i++)
      gx[i] = i * dr;
    for (int i = 1; i < ng - 1; i++)
      gy[i] = opt ? brk.rvsr(gx[i]) : gx[i] * (std::erfc(sig * gx[i]) / gx[i]);
    gy[0]      = 2.0 * gy[1] - gy[2];
    gy[ng - 1] = 0.0;
    clamped_spline(gx, gy, (gy[1] - gy[0]) / dr, 0.0, gy2);
    auto vsr = [&](double r) {
      if (r >= gx[ng - 1])
        return 0.0;
      const int loc     = (int)(r / dr);
      const double dist = r - gx[loc];
      const double cl = dist / dr, cr = (dr - dist) / dr, h6 = dr / 6.0;
      const double q1 = cr * (cr * cr - 1.0) * h6 * dr, q2 = cl * (cl * cl - 1.0) * h6 * dr;
      return (cr * gy[loc] + cl * gy[loc + 1] + q1 * gy2[loc] + q2 * gy2[loc + 1]) / r;
    };
    const double vlr0 = opt ? brk.vlr_r0() : 2.0 * sig / std::sqrt(M_PI);
    const double vsk0 = opt ? brk.vsr_k0() : M_PI / (sig * sig * vol);
    const int N = s.N, ni = s.nion;
    // ion-ion energy (CoulombPBCAA on the ions)
    double vii = 0;
    if (ni > 0)
    {
      double sr = 0, lr = 0;
      for (int a = 1; a < ni; a++)
        for (int b = 0; b < a; b++)
        {
          double d2 = 0;
          for (int q = 0; q < 3; q++)
          {
            const double u  = (ions_h[(size_t)q * ni + b] - ions_h[(size_t)q * ni + a]) * s.Linv[q];
            const double dd = s.L[q] * (u - std::round(u));
            d2 += dd * dd;
          }
          sr += zion * zion * vsr(std::sqrt(d2));
        }
      for (int k = 0; k < nk; k++)
        lr += fk[k] * (rhoi[k] * rhoi[k] + rhoi[nk + k] * rhoi[nk + k]);
      vii = sr + 0.5 * lr + ni * (-0.5 * zion * zion * vlr0 - 0.5 * zion * (zion * ni) * vsk0);
    }
    // r v_sr on interval i as a cubic in t = (r - i dr) / dr: the spline's
    // (1 - t) y_i + t y_{i+1} + h ((1 - t)^3 - (1 - t)) y2_i + h (t^3 - t) y2_{i+1}, h = dr^2 / 6
    std::vector<double> poly(4 * (size_t)(ng - 1));
    const double hh = dr * dr / 6.0;
    for (int i = 0; i < ng - 1; i++)
    {
      poly[4 * i + 0] = gy[i];
      poly[4 * i + 1] = (gy[i + 1] - gy[i]) - hh * (2.0 * gy2[i] + gy2[i + 1]);
      poly[4 * i + 2] = 3.0 * hh * gy2[i];
      poly[4 * i + 3] = hh * (gy2[i + 1] - gy2[i]);
    }
    coul_k.alloc(kv.size());
    coul_fk.alloc(nk);
    coul_rhoi.alloc(rhoi.size());
    coul_poly.alloc(poly.size());
    HIPCHECK(hipMemcpy(coul_k.p, kv.data(), kv.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(coul_fk.p, fk.data(), fk.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(coul_rhoi.p, rhoi.data(), rhoi.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(coul_poly.p, poly.data(), poly.size() * sizeof(double), hipMemcpyHostToDevice));
    coul          = CoulombC{};
    coul.kvec     = coul_k.p;
    coul.fk       = coul_fk.p;
    coul.rhoi     = coul_rhoi.p;
    coul.poly     = coul_poly.p;
    coul.nk       = nk;
    coul.ng       = ng;
    coul.dr       = dr;
    coul.drinv    = 1.0 / dr;
    coul.rmax     = gx[ng - 1];
    coul.zion     = zion;
    coul.const_ee = N * (-0.5 * vlr0) - 0.5 * (double)N * (double)N * vsk0;
    coul.const_ei = (double)N * zion * ni * vsk0;
    coul.vii      = vii;
    coul_ready    = true;
  }

  // ---------------------------------------------------------------------------------------------
  // Non-local pseudopotential (one species): radial channels as cubic splines on a linear grid
  // (values and second derivatives, the OneDimCubicSpline form), the quadrature grid and weights
  bool nlpp_ready = false;
  NlppC nlpp{};
  DevBuf<double> nlpp_poly, nlpp_knots, nlpp_rot, nlpp_part, nlpp_out;
  // pair-list form of the energy (QMI_NLPP_FORM=electron at set-up selects one workgroup per
  // electron instead); ke from the fixed ion positions
  bool nlpp_pairs_on = true;
  int nlpp_ke        = 1;
  int nlpp_grid      = 8;
  DevBuf<int> np_icnt, np_total, np_ecnt, np_pion, np_list;
  DevBuf<double> np_wres, np_wpts;
  // non-local moves: per-electron scan weights of the last evaluation, selection in / out
  DevBuf<double> tm_tsum, tm_base, tm_target, tm_out, tm_disp;
  DevBuf<int> tm_iel, tm_chosen, tm_acc, tm_count;
  std::vector<double> h_tsum;

  void nlpp_setup(int nch, const int* ls, int ng, double dr, const double* y, const double* y2, double rcut, int nknot,
                  const double* knots, const double* weights)
  {
    if (!have_system)
      throw QmiError(QMI_ERR_ARG, "system not created");
    if (skew_cell)
      throw QmiError(QMI_ERR_UNSUPPORTED, "the pseudopotential energy needs an orthorhombic cell");
    if (nch < 1 || nch > NLPP_MAX_CH || !ls || ng < 2 || !(dr > 0) || !y || !y2 || !(rcut > 0) ||
        rcut > (ng - 1) * dr || nknot < 1 || nknot > NLPP_MAX_KNOTS || !knots || !weights)
      throw QmiError(QMI_ERR_ARG, "invalid pseudopotential tables (1..4 channels, r_cut inside the grid, 1..64 points)");
    if (sys.nion > NLPP_MAX_IONS)
      throw QmiError(QMI_ERR_UNSUPPORTED, "more than 2048 ions");
    NlppC c{};
    c.nch  = nch;
    c.lmax = 0;
    for (int i = 0; i < nch; i++)
    {
      if (ls[i] < 0 || ls[i] > NLPP_MAX_L)
        throw QmiError(QMI_ERR_ARG, "channel l must be in [0, 4]");
      c.l[i]  = ls[i];
      c.lmax  = std::max(c.lmax, ls[i]);
    }
    c.ng    = ng;
    c.dr    = dr;
    c.drinv = 1.0 / dr;
    c.rcut  = rcut;
    c.nknot = nknot;
    {
      // quadrature points per pass of the kernel: QMI_NLPP_POINTS = 6 (default) or 12
      const char* e = std::getenv("QMI_NLPP_POINTS");
      c.kc          = (e && std::atoi(e) == 12) ? 12 : 6;
    }
    // the spline's (1 - t) y_i + t y_{i+1} + h ((1 - t)^3 - (1 - t)) y2_i + h (t^3 - t) y2_{i+1},
    // h = dr^2 / 6, as a cubic in t per interval
    std::vector<double> poly((size_t)nch * (ng - 1) * 4);
    const double hh = dr * dr / 6.0;
    for (int ch = 0; ch < nch; ch++)
      for (int i = 0; i < ng - 1; i++)
      {
        const double* yy  = y + (size_t)ch * ng;
        const double* yy2 = y2 + (size_t)ch * ng;
        double* q         = poly.data() + ((size_t)ch * (ng - 1) + i) * 4;
        q[0]              = yy[i];
        q[1]              = (yy[i + 1] - yy[i]) - hh * (2.0 * yy2[i] + yy2[i + 1]);
        q[2]              = 3.0 * hh * yy2[i];
        q[3]              = hh * (yy2[i + 1] - yy2[i]);
      }
    std::vector<double> kw((size_t)nknot * 4);
    for (int j = 0; j < nknot; j++)
    {
      for (int a = 0; a < 3; a++)
        kw[4 * j + a] = knots[3 * j + a];
      kw[4 * j + 3] = weights[j];
    }
    nlpp_poly.alloc(poly.size());
    nlpp_knots.alloc(kw.size());
    HIPCHECK(hipMemcpy(nlpp_poly.p, poly.data(), poly.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(nlpp_knots.p, kw.data(), kw.size() * sizeof(double), hipMemcpyHostToDevice));
    c.vpoly    = nlpp_poly.p;
    c.knots    = nlpp_knots.p;
    nlpp       = c;
    {
      const char* e = std::getenv("QMI_NLPP_FORM");
      nlpp_pairs_on = !(e && std::string(e) == "electron");
    }
    // Ions within r_cut of one point lie within 2 r_cut of each other (triangle inequality of the
    // minimum-image distance), so 1 + the most ions within 2 r_cut of any one ion bounds the pairs of
    // an electron.  Same distance formula as the device; a relative margin of 1e-9 covers rounding.
    {
      const int ni = sys.nion;
      std::vector<double> ih(3 * (size_t)ni);
      if (ni)
        HIPCHECK(hipMemcpy(ih.data(), ions.p, ih.size() * sizeof(double), hipMemcpyDeviceToHost));
      const double lim = 2.0 * rcut * (1.0 + 1e-9);
      int most         = 0;
      for (int a = 0; a < ni; a++)
      {
        int cnt = 0;
        for (int b = 0; b < ni; b++)
        {
          if (b == a)
            continue;
          double d2 = 0.0;
          for (int q = 0; q < 3; q++)
          {
            double d = ih[(size_t)q * ni + b] - ih[(size_t)q * ni + a];
            if (sys.periodic[q])
            {
              const double u = d * sys.Linv[q];
              d              = sys.L[q] * (u - std::round(u));
            }
            d2 += d * d;
          }
          cnt += std::sqrt(d2) < lim ? 1 : 0;
        }
        most = std::max(most, cnt);
      }
      nlpp_ke = std::max(1, std::min(ni, 1 + most));
      int ncu = 0;
      HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
      // all workgroups resident: 3 per compute unit at 6 points per pass (the register bound), 1 at 12
      nlpp_grid = std::max(8, (ncu * (c.kc == 6 ? 4 : 1) + 7) / 8 * 8);
    }
    nlpp_ready = true;
  }

  // with_txy: also the per-electron scan weights of the non-local moves (into h_tsum)
  void nlpp_evaluate(const double* rot, double* v, bool with_txy = false)
  {
    require_walkers();
    if (!nlpp_ready)
      throw QmiError(QMI_ERR_ARG, "call qmi_nlpp_setup first");
    if (cplx)
      throw QmiError(QMI_ERR_UNSUPPORTED, "the pseudopotential energy is built for real orbitals");
    if (dcount[0] || dcount[1])
      throw QmiError(QMI_ERR_ARG, "pseudopotential energy with pending delayed updates (call completeUpdates)");
    if (!rot || !v)
      throw QmiError(QMI_ERR_ARG, "null rotation or output");
    if (nlpp_rot.n < (size_t)cap * 9)
      nlpp_rot.alloc((size_t)cap * 9);
    if (nlpp_part.n < (size_t)cap * sys.N)
      nlpp_part.alloc((size_t)cap * sys.N);
    if (nlpp_out.n < (size_t)cap)
      nlpp_out.alloc(cap);
    HIPCHECK(hipMemcpyAsync(nlpp_rot.p, rot, (size_t)nw * 9 * sizeof(double), hipMemcpyHostToDevice, stream));
    TmoveArgs tm{};
    if (with_txy)
    {
      if (tm_tsum.n < (size_t)cap * sys.N)
        tm_tsum.alloc((size_t)cap * sys.N);
      tm.mode  = 1;
      tm.plus  = tmove_plus();
      tm.minus = tmove_minus();
      tm.tsum  = tm_tsum.p;
    }
    NlppPairs pl{};
    const bool pairs = nlpp_pairs_on;
    if (pairs)
    {
      const size_t nslot = (size_t)cap * sys.N * nlpp_ke;
      if (nslot > (size_t)INT32_MAX)
        throw QmiError(QMI_ERR_UNSUPPORTED, "too many electron-ion pair slots for the pair list");
      if (np_icnt.n < (size_t)std::max(1, sys.nion))
        np_icnt.alloc(std::max(1, sys.nion));
      if (!np_total.n)
        np_total.alloc(1);
      if (np_ecnt.n < (size_t)cap * sys.N)
        np_ecnt.alloc((size_t)cap * sys.N);
      if (np_pion.n < nslot)
      {
        np_pion.alloc(nslot);
        np_list.alloc(nslot);
        np_wres.alloc(nslot);
      }
      if (with_txy && np_wpts.n < nslot * nlpp.nknot)
        np_wpts.alloc(nslot * nlpp.nknot);
      pl = NlppPairs{nlpp_ke,   nlpp_grid, np_icnt.p, np_total.p, np_ecnt.p,
                     np_pion.p, np_list.p, np_wres.p, with_txy ? np_wpts.p : nullptr};
    }
    {
      Scoped t(this, "nlpp");
      HIPCHECK(launch_nlpp(sys, st, coefs.p, nlpp, nw, nlpp_rot.p, nlpp_part.p, nlpp_out.p, tm, stream,
                           pairs ? &pl : nullptr));
    }
    HIPCHECK(hipMemcpyAsync(v, nlpp_out.p, (size_t)nw * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (with_txy)
    {
      h_tsum.resize((size_t)nw * sys.N);
      HIPCHECK(hipMemcpyAsync(h_tsum.data(), tm_tsum.p, h_tsum.size() * sizeof(double), hipMemcpyDeviceToHost,
                              stream));
    }
    sync();
  }

  // NonLocalTOperator plusFactor = tau gamma, minusFactor = -tau (1 - alpha (1 + gamma)) with the
  // defaults alpha = gamma = 0 (NonLocalTOperator.cpp:62-70)
  double tmove_plus() const { return 0.0; }
  double tmove_minus() const { return -dmc.p.tau; }

  // T-moves v0 (NonLocalECPotential::makeNonLocalMovesPbyP, NonLocalECPotential.cpp:546-567): each
  // walker draws one uniform u (after its grid rotation) and NonLocalTOperator::selectMove
  // (NonLocalTOperator.cpp:79-106) picks the candidate with the last scan value <= u x total, the
  // scan starting at 1 (no move) over the electron-major list of the last evaluation; the chosen
  // electron is moved by deltaV and the move accepted when the ratio is non-zero.  The scan is
  // summed per electron on the device, across electrons here; the selected electron's points are
  // re-evaluated to find the point.  The moves go through the multi-walker protocol kernels one
  // electron at a time (walkers that do not move that electron take a zero displacement and are
  // rejected).
  int dmc_tmoves(const std::vector<double>& u)
  {
    const int N = sys.N;
    std::vector<int> sel(nw, -1);
    std::vector<double> base(nw, 0.0), target(nw, 0.0);
    bool any = false;
    for (int w = 0; w < nw; w++)
    {
      const double* T = h_tsum.data() + (size_t)w * N;
      double total    = 1.0;
      for (int e = 0; e < N; e++)
        if (T[e] >= 0.0)
          total += T[e];
      const double tg = u[w] * total;
      double acc      = 1.0;
      for (int e = 0; e < N; e++)
      {
        if (T[e] < 0.0)
          continue;
        if (acc <= tg)
        {
          sel[w]  = e;
          base[w] = acc;
        }
        acc += T[e];
      }
      target[w] = tg;
      any |= sel[w] >= 0;
    }
    if (!any)
      return 0;
    if (tm_iel.n < (size_t)cap)
    {
      tm_iel.alloc(cap);
      tm_base.alloc(cap);
      tm_target.alloc(cap);
      tm_out.alloc((size_t)cap * 4);
    }
    HIPCHECK(hipMemcpyAsync(tm_iel.p, sel.data(), (size_t)nw * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(tm_base.p, base.data(), (size_t)nw * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(tm_target.p, target.data(), (size_t)nw * sizeof(double), hipMemcpyHostToDevice, stream));
    TmoveArgs tm{};
    tm.mode       = 2;
    tm.plus       = tmove_plus();
    tm.minus      = tmove_minus();
    tm.sel_iel    = tm_iel.p;
    tm.sel_base   = tm_base.p;
    tm.sel_target = tm_target.p;
    tm.sel_out    = tm_out.p;
    {
      Scoped t(this, "tmove_select");
      HIPCHECK(launch_nlpp(sys, st, coefs.p, nlpp, nw, nlpp_rot.p, nlpp_part.p, nlpp_out.p, tm, stream));
    }
    std::vector<double> so((size_t)nw * 4);
    HIPCHECK(hipMemcpyAsync(so.data(), tm_out.p, so.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    // the moves of each electron that some walker moves, enqueued without host round trips:
    // displacements (zero for the walkers that do not move it), ratio, the acceptance mask on the
    // device (moved and a non-zero ratio), accept; the delayed block flushed when full
    std::vector<int> elist;
    std::vector<double> disp;
    std::vector<int> chosen;
    for (int e = 0; e < N; e++)
    {
      bool here = false;
      for (int w = 0; w < nw && !here; w++)
        here = sel[w] == e && so[(size_t)w * 4] != 0.0;
      if (!here)
        continue;
      elist.push_back(e);
      for (int w = 0; w < nw; w++)
      {
        const bool m = sel[w] == e && so[(size_t)w * 4] != 0.0;
        for (int a = 0; a < 3; a++)
          disp.push_back(m ? so[(size_t)w * 4 + 1 + a] : 0.0);
        chosen.push_back(m ? 1 : 0);
      }
    }
    if (elist.empty())
      return 0;
    if (tm_disp.n < disp.size())
      tm_disp.alloc(disp.size());
    if (tm_chosen.n < chosen.size())
      tm_chosen.alloc(chosen.size());
    if (tm_acc.n < (size_t)cap)
      tm_acc.alloc(cap);
    if (tm_count.n < 2)
      tm_count.alloc(2);
    HIPCHECK(hipMemcpyAsync(tm_disp.p, disp.data(), disp.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(tm_chosen.p, chosen.data(), chosen.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHECK(launch_fill_i32(tm_count.p, 2, 0, stream));
    for (size_t i = 0; i < elist.size(); i++)
    {
      const int e = elist[i], g = group(e);
      HIPCHECK(launch_make_move(sys, st, e, nw, tm_disp.p + i * (size_t)nw * 3, stream));
      MoveArgs a = move_args(e);
      a.out      = out.p;
      a.prepare  = (sys.k == 1 || invrow_id != e) ? 1 : 0;
      {
        Scoped t(this, "api_ratio_grad");
        HIPCHECK(move_kernel(MK_API_RATIO, a));
      }
      HIPCHECK(launch_tmove_accept_mask(nw, out.p, tm_chosen.p + i * (size_t)nw, tm_acc.p, tm_count.p, stream));
      MoveArgs b = move_args(e);
      b.host_acc = tm_acc.p;
      {
        Scoped t(this, "api_accept");
        HIPCHECK(move_kernel(MK_API_ACCEPT, b));
      }
      invrow_id = -1;
      if (sys.k > 1 && ++dcount[g] == sys.k)
        flush(g);
    }
    flush(0);
    flush(1);
    int counts[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(counts, tm_count.p, sizeof(counts), hipMemcpyDeviceToHost, stream));
    sync();
    if (counts[1])
      throw QmiError(QMI_ERR_NUMERIC, "NaN ratio in a non-local move");
    return counts[0];
  }

  // ---------------------------------------------------------------------------------------------
  // DMC: DMCBatched::advanceWalkers / runDMCStep (QMCDrivers/DMC/DMCBatched.cpp:67-414) with the
  // branch engine SFNBranch (SFNBranch.h:161-208, SFNBranch.cpp:93-245, 284-317) and the
  // fluctuating-population walker control (WalkerControl.cpp:150-260, MCPopulation.cpp:34-66).
  // One walker per crowd: each walker slot owns an RNG stream (seed + slot, created when the slot
  // is first used); walker control draws from its own stream (seed - 1).  A branched copy takes
  // the parent's whole state (positions, inverses, orbital derivatives, Jastrow accumulators,
  // log det) where the reference reloads and recomputes it: the same values to rounding.
  struct DmcState
  {
    bool on = false, nlpp = false;
    qmi_dmc_params p{};
    long long iter = 0;
    std::vector<double> weight, eloc; // per slot
    std::mt19937 wc_rng;
    double tau_eff = 0, etrial = 0, eref = 0, sigma2 = 0, enow = 0;
    double cutoff = 0, branch_max = 0, filter = 0;
    double r2acc = 0, r2prop = 0; // accumulator sums (SFNBranch R2Accepted / R2Proposed)
    int warmup_todo = 0, etrial_todo = 1, target = 0;
    bool main_stage = false;
    std::deque<std::array<double, 3>> hist; // limited history: weight, energy, variance
    size_t hist_limit = 1;
    double ehist = 0, vhist = 0, nhist = 0; // unlimited history sums
  } dmc;
  DevBuf<double> dmc_rr;
  DevBuf<int> dmc_pairs;
  PinnedBuf<int> h_pairs;
  PinnedBuf<double> h_rr, h_pot;

  // random rotation of the quadrature grid from three uniforms (RandomRotationMatrix.h:24-32 with
  // Numerics/RotationMatrix3D.h: phi = 2 pi u1, psi = 2 pi u2, cos theta = 1 - 2 u3), row-major
  static void rotation_from_uniforms(double u1, double u2, double u3, double* m)
  {
    const double phi = 2.0 * M_PI * u1, psi = 2.0 * M_PI * u2, cth = 1.0 - 2 * u3;
    const double sph = std::sin(phi), cph = std::cos(phi);
    const double sth = std::sqrt(1.0 - cth * cth);
    const double sps = std::sin(psi), cps = std::cos(psi);
    m[0] = cph * cth * cps - sph * sps;
    m[1] = sph * cth * cps + cph * sps;
    m[2] = -sth * cps;
    m[3] = -cph * cth * sps - sph * cps;
    m[4] = -sph * cth * sps + cph * cps;
    m[5] = sth * sps;
    m[6] = cph * sth;
    m[7] = sph * sth;
    m[8] = cth;
  }

  // SFNBranch::setBranchCutoff (SFNBranch.cpp:284-317)
  void dmc_set_cutoff(double variance, double target_sigma, double max_sigma, int nelec)
  {
    const double tau = dmc.p.tau;
    switch (dmc.p.cutoff_scheme)
    {
    case 1: dmc.cutoff = 2.0 / std::sqrt(tau); break;
    case 2:
      if (nelec > 0)
        dmc.cutoff = 0.2 * std::sqrt(nelec / tau);
      break;
    case 3: dmc.cutoff = std::sqrt(variance) * std::min(target_sigma, std::sqrt(1.0 / tau)); break;
    default: dmc.cutoff = std::min(std::max(variance * target_sigma, max_sigma), 2.5 / tau); break;
    }
    dmc.branch_max = dmc.cutoff * 1.5;
    dmc.filter     = 1.0 / (dmc.branch_max - dmc.cutoff);
  }

  // SFNBranch::branchWeight (SFNBranch.h:199-208)
  double dmc_branch_weight(double enew, double eold) const
  {
    double taueff  = dmc.tau_eff * 0.5;
    const double x = std::max(dmc.eref - enew, dmc.eref - eold);
    if (x > dmc.branch_max)
      taueff = 0.0;
    else if (x > dmc.cutoff)
      taueff *= (1.0 - (x - dmc.cutoff) * dmc.filter);
    return std::exp(taueff * (dmc.etrial * 2.0 - enew - eold));
  }

  // local energies of the walkers: kinetic (evaluateGL) + V_ee + V_eI + V_II (+ V_NL when a
  // pseudopotential is set: each walker's grid rotated by three draws from its slot stream, after
  // the sweep's draws, as NonLocalECPotential draws from the crowd's generator), summed in that order
  void dmc_local_energies(std::vector<double>& e, bool with_rr)
  {
    std::vector<double> vnl;
    if (dmc.nlpp)
    {
      std::vector<double> rot((size_t)nw * 9);
      for (int w = 0; w < nw; w++)
      {
        const double u1 = rng_uniform01(rng[w]);
        const double u2 = rng_uniform01(rng[w]);
        const double u3 = rng_uniform01(rng[w]);
        rotation_from_uniforms(u1, u2, u3, rot.data() + (size_t)w * 9);
      }
      vnl.resize(nw);
      nlpp_evaluate(rot.data(), vnl.data(), with_rr && dmc.p.nonlocal_moves);
    }
    eval_gl_enqueue();
    if (coul_out.n < (size_t)cap * 3)
      coul_out.alloc((size_t)cap * 3);
    h_pot.alloc((size_t)cap * 3);
    {
      Scoped t(this, "coulomb");
      HIPCHECK(launch_coulomb(sys, st, coul, nw, coul_out.p, stream));
    }
    HIPCHECK(hipMemcpyAsync(h_pot.p, coul_out.p, (size_t)nw * 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (with_rr)
      HIPCHECK(hipMemcpyAsync(h_rr.p, dmc_rr.p, (size_t)nw * 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    e.resize(nw);
    eval_gl_finish(e.data());
    for (int w = 0; w < nw; w++)
    {
      const double* v = h_pot.p + (size_t)w * 3;
      e[w]            = e[w] + v[0] + v[1] + v[2];
      if (dmc.nlpp)
        e[w] += vnl[w];
    }
  }

  void dmc_init(const qmi_dmc_params& p)
  {
    require_walkers();
    if (wpc != 1)
      throw QmiError(QMI_ERR_UNSUPPORTED, "DMC needs one walker per crowd (one RNG stream per walker)");
    if (learned)
      throw QmiError(QMI_ERR_UNSUPPORTED, "DMC with the learned proposal is not built");
    if (!coul_ready)
      throw QmiError(QMI_ERR_ARG, "DMC needs the Coulomb energy: call qmi_coulomb_setup first");
    if (!(p.tau > 0) || !(p.feedback > 0) || p.warmup_steps < 0 || p.capacity < nw || p.recompute_period < 0 ||
        p.cutoff_scheme < 0 || p.cutoff_scheme > 3 || p.target_walkers < 0)
      throw QmiError(QMI_ERR_ARG, "invalid DMC parameters (tau, feedback > 0; capacity >= walkers)");
    if (dcount[0] || dcount[1])
      throw QmiError(QMI_ERR_ARG, "DMC start with pending delayed updates");
    if (cap < p.capacity)
    {
      // re-allocate the walker slots for the capacity, keeping the walkers' positions
      const int N = sys.N, Np = sys.Npad, n0 = nw;
      std::vector<double> t((size_t)nw * 3 * Np), Rh((size_t)nw * N * 3);
      HIPCHECK(hipMemcpy(t.data(), R.p, t.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int w = 0; w < nw; w++)
        for (int i = 0; i < N; i++)
          for (int q = 0; q < 3; q++)
            Rh[((size_t)w * N + i) * 3 + q] = t[((size_t)w * 3 + q) * Np + i];
      walkers_load(n0, Rh.data(), 1, p.capacity);
    }
    invalidate_graph();
    // initialLogEvaluation: recompute the wave function, local energies
    {
      Scoped t(this, "jastrow_recompute");
      HIPCHECK(launch_jastrow_recompute(sys, st, nw, stream));
    }
    recompute_dets();
    dmc_rr.alloc((size_t)cap * 2);
    dmc_pairs.alloc((size_t)cap * 2);
    h_pairs.alloc((size_t)cap * 2);
    h_rr.alloc((size_t)cap * 2);
    if (nlpp_ready && cplx)
      throw QmiError(QMI_ERR_UNSUPPORTED, "the pseudopotential energy is built for real orbitals");
    if (p.nonlocal_moves != 0 && (p.nonlocal_moves != 1 || !nlpp_ready))
      throw QmiError(QMI_ERR_ARG, "non-local moves (v0) need the pseudopotential: call qmi_nlpp_setup first");
    dmc      = DmcState{};
    dmc.p    = p;
    dmc.nlpp = nlpp_ready;
    rng.clear();
    for (int c = 0; c < nw; c++)
      rng.emplace_back(p.seed + uint32_t(c));
    rng_seed   = p.seed;
    rng_seeded = true;
    std::vector<double> e;
    dmc_local_energies(e, false);
    dmc.wc_rng.seed(p.seed - 1u);
    dmc.weight.assign(cap, 1.0);
    dmc.eloc.assign(cap, 0.0);
    std::copy(e.begin(), e.end(), dmc.eloc.begin());
    // MCPopulation::measureGlobalEnergyVariance + SFNBranch::initParam
    double ws = 0, we = 0, wee = 0;
    for (int w = 0; w < nw; w++)
    {
      ws += dmc.weight[w];
      we += dmc.weight[w] * e[w];
      wee += dmc.weight[w] * e[w] * e[w];
    }
    const double ene = we / ws, var = wee / ws - ene * ene;
    dmc.r2acc   = 1.0e-10;
    dmc.r2prop  = 1.0e-10;
    dmc.etrial  = ene;
    dmc.eref    = ene;
    dmc.sigma2  = var;
    dmc.tau_eff = p.tau * dmc.r2acc / dmc.r2prop;
    dmc_set_cutoff(var, p.sigma_bound, 50, sys.N);
    dmc.target      = p.target_walkers > 0 ? p.target_walkers : nw;
    dmc.warmup_todo = p.warmup_steps;
    dmc.main_stage  = p.warmup_steps == 0;
    dmc.etrial_todo = 1;
    dmc.hist_limit  = (size_t)std::max(1, (int)(1.0 / (p.feedback * p.tau)));
    dmc.iter        = 0;
    dmc.on          = true;
  }

  // one DMC step; out[QMI_DMC_STATS]
  void dmc_step(double* out)
  {
    const SysC& s = sys;
    const int N   = s.N;
    const double tau = dmc.p.tau;
    const int nw0 = nw;
    // advanceWalkers: one sweep with the DMC acceptance
    substep_buffers(false);
    // the sweep in two launches around the second part of the host's numbers (see vmc_run)
    split_n = N >= 2 ? std::min(N - 1, std::max(1, N / 8)) : 0;
    if (split_n > 0)
      rng_generate_part(1);
    else
      rng_generate();
    HIPCHECK(hipMemsetAsync(dmc_rr.p, 0, (size_t)nw * 2 * sizeof(double), stream));
    HIPCHECK(hipMemsetAsync(stats.p, 0, (size_t)nw * 2 * sizeof(double), stream));
    if (split_n > 0)
    {
      enqueue_substep(tau, false, false, 1);
      rng_generate_part(2);
      enqueue_substep(tau, false, true, 2);
    }
    else
      enqueue_substep(tau, false, true);
    HIPCHECK(hipEventSynchronize(ev_ubase));
    finish_substep();
    const bool recompute = dmc.p.recompute_period > 0 && ((step_count + 1) % dmc.p.recompute_period == 0);
    if (recompute)
      recompute_dets();
    std::vector<double> hs((size_t)nw * 2);
    HIPCHECK(hipMemcpyAsync(hs.data(), stats.p, hs.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    std::vector<double> enew;
    dmc_local_energies(enew, true);
    // non-local moves after the local energy (DMCBatched.cpp:298-325): one uniform per walker from
    // its stream, after the grid rotation's three
    int ntmove = 0;
    if (dmc.p.nonlocal_moves)
    {
      std::vector<double> u(nw);
      for (int w = 0; w < nw; w++)
        u[w] = rng_uniform01(rng[w]);
      ntmove = dmc_tmoves(u);
    }
    step_count++;
    double nacc = 0, nrej = 0;
    for (int w = 0; w < nw; w++)
    {
      nacc += hs[2 * w];
      nrej += hs[2 * w + 1];
    }
    // branch weights (DMCBatched.cpp:262-272)
    for (int w = 0; w < nw; w++)
    {
      dmc.weight[w] *= dmc_branch_weight(enew[w], dmc.eloc[w]);
      dmc.eloc[w] = enew[w];
    }
    // WalkerControl::branch: multiplicities, ensemble properties (computeCurData, writeDMCdat)
    const bool no_branch = dmc.iter == 0;
    std::vector<int> mult(nw);
    for (int w = 0; w < nw; w++)
      mult[w] = no_branch ? 1 : static_cast<int>(dmc.weight[w] + rng_uniform01(dmc.wc_rng));
    double esum = 0, e2sum = 0, wsum = 0, r2a = 0, r2p = 0;
    long long total = 0;
    for (int w = 0; w < nw; w++)
    {
      total += mult[w];
      r2a += h_rr.p[2 * w + 1];
      r2p += h_rr.p[2 * w];
      const double e = dmc.eloc[w], wg = dmc.weight[w];
      esum += wg * e;
      e2sum += wg * e * e;
      wsum += wg;
    }
    const double winv = 1.0 / wsum;
    const double eavg = esum * winv;
    const double evar = e2sum * winv - eavg * eavg;
    if (total < 1)
      throw QmiError(QMI_ERR_NUMERIC, "the DMC population died out");
    if (total > cap)
      throw QmiError(QMI_ERR_NUMERIC, "DMC population " + std::to_string(total) + " exceeds the capacity " +
                                          std::to_string(cap));
    dmc_branch_slots(mult);
    if (!no_branch)
      for (int w = 0; w < nw; w++)
        dmc.weight[w] = 1.0;
    // SFNBranch::updateParamAfterPopControl
    const double logN = std::log(static_cast<double>(dmc.target));
    dmc.enow          = eavg;
    dmc.r2acc += r2a;
    dmc.r2prop += r2p;
    dmc.tau_eff = tau * dmc.r2acc / dmc.r2prop;
    if (dmc.main_stage)
    {
      if (dmc.p.refenergy_history == 1)
      {
        dmc.ehist += eavg;
        dmc.vhist += evar;
        dmc.nhist += 1.0;
        dmc.eref = dmc.ehist / dmc.nhist;
      }
      else
      {
        if (dmc.hist.size() == dmc.hist_limit)
          dmc.hist.pop_front();
        dmc.hist.push_back({wsum, eavg, evar});
        double sw = 0, se = 0;
        for (auto& h : dmc.hist)
        {
          sw += h[0];
          se += h[1] * h[0];
        }
        dmc.eref = se / sw;
      }
      if (--dmc.etrial_todo == 0)
      {
        dmc.etrial      = dmc.eref + dmc.p.feedback * (logN - std::log(wsum));
        dmc.etrial_todo = 1;
      }
    }
    else
    {
      dmc.eref   = eavg;
      dmc.etrial = eavg + (logN - std::log(wsum)) / tau;
      if (--dmc.warmup_todo == 0)
      {
        dmc.sigma2 = evar;
        dmc_set_cutoff(dmc.sigma2, dmc.p.sigma_bound, 10, N);
        dmc.main_stage = true;
      }
    }
    dmc.iter++;
    if (out)
    {
      out[0] = eavg;
      out[1] = evar;
      out[2] = wsum;
      out[3] = nw0;
      out[4] = nw;
      out[5] = dmc.etrial;
      out[6] = dmc.eref;
      out[7] = dmc.tau_eff;
      out[8] = nacc;
      out[9] = nrej;
      out[10] = ntmove;
    }
  }

  // Kill walkers of multiplicity 0 and copy those of multiplicity m > 1 (m - 1 copies): copies go
  // to the dead slots in ascending order, then past the end; remaining holes are filled from the
  // tail.  Two launches: copies (sources live, destinations dead or new), then tail moves.
  // a VMC section after DMC (as consecutive driver sections of one input) samples |psi|^2 again: the
  // DMC acceptance (node rejection, rr sums) is switched off and graphs captured with it are dropped
  void end_dmc()
  {
    if (!dmc.on)
      return;
    dmc.on = false;
    invalidate_graph();
  }

  void dmc_branch_slots(const std::vector<int>& mult)
  {
    std::vector<int> holes;
    for (int w = 0; w < nw; w++)
      if (mult[w] == 0)
        holes.push_back(w);
    std::vector<int> pairs;
    size_t hi = 0;
    int next  = nw;
    for (int w = 0; w < nw; w++)
      for (int c = 1; c < mult[w]; c++)
      {
        const int dst = hi < holes.size() ? holes[hi++] : next++;
        pairs.push_back(w);
        pairs.push_back(dst);
        dmc.weight[dst] = dmc.weight[w];
        dmc.eloc[dst]   = dmc.eloc[w];
      }
    const int n1 = next;
    std::vector<char> live(n1, 1);
    for (size_t h = hi; h < holes.size(); h++)
      live[holes[h]] = 0;
    const size_t ncopy = pairs.size() / 2;
    int lo = 0, top = n1 - 1;
    for (;;)
    {
      while (lo < n1 && live[lo])
        lo++;
      while (top >= 0 && !live[top])
        top--;
      if (lo >= top)
        break;
      pairs.push_back(top);
      pairs.push_back(lo);
      dmc.weight[lo] = dmc.weight[top];
      dmc.eloc[lo]   = dmc.eloc[top];
      live[lo]       = 1;
      live[top]      = 0;
    }
    const int nw_new = top + 1;
    if (!pairs.empty())
    {
      std::copy(pairs.begin(), pairs.end(), h_pairs.p);
      HIPCHECK(hipMemcpyAsync(dmc_pairs.p, h_pairs.p, pairs.size() * sizeof(int), hipMemcpyHostToDevice, stream));
      WalkerArrays t{};
      auto add = [&](void* base, size_t bytes) {
        t.base[t.n]  = static_cast<char*>(base);
        t.bytes[t.n] = bytes;
        t.n++;
      };
      const int Np = sys.Npad;
      add(R.p, (size_t)3 * Np * sizeof(double));
      add(jas.p, (size_t)J_ROWS * Np * sizeof(double));
      add(logdet.p, 4 * sizeof(double));
      if (sys.has_det)
        for (int g = 0; g < 2; g++)
        {
          const size_t n = std::max(sys.n[g], 1);
          add(Ainv[g].p, det_words(n * sys.ld) * sizeof(double));
          add(dpsi[g].p, det_words(n * 4 * sys.ld) * sizeof(double));
        }
      {
        Scoped tm(this, "copy_walkers");
        HIPCHECK(launch_copy_walkers(t, dmc_pairs.p, (int)ncopy, stream));
        HIPCHECK(launch_copy_walkers(t, dmc_pairs.p + 2 * ncopy, (int)(pairs.size() / 2 - ncopy), stream));
      }
      // pinned pairs are reused by the next step: wait for the upload
      sync();
    }
    if (nw_new != nw)
      invalidate_graph(); // captured sweeps are bound to the walker count
    nw     = nw_new;
    ncrowd = nw;
    while ((int)rng.size() < nw)
      rng.emplace_back(rng_seed + uint32_t(rng.size()));
  }

  // ---------------------------------------------------------------------------------------------
  // Restart files.  Walkers (positions [nw][N][3], weights) as HDFWalkerOutput::write_configuration
  // (Particle/HDFWalkerOutput.cpp:110-160), the random streams as RandomNumberControl::write_rank_0
  // (Utilities/RandomNumberControl.cpp:382-) with StdRandom::save, and for a DMC context the branch
  // state (BranchIO::write, QMCDrivers/BranchIO.cpp:109-160: parameters, energy history, R2
  // accumulators) with the walker-control stream and the walkers' local energies.
  enum
  {
    CK_V_TAU,
    CK_V_TAUEFF,
    CK_V_ETRIAL,
    CK_V_EREF,
    CK_V_BRANCHMAX,
    CK_V_CUTOFF,
    CK_V_FILTER,
    CK_V_SIGMA2,
    CK_V_ENOW,
    CK_V_R2ACC,
    CK_V_R2PROP,
    CK_V_FEEDBACK,
    CK_V_SIGMABOUND,
    CK_V_EHIST,
    CK_V_VHIST,
    CK_V_NHIST,
    CK_NV
  };
  enum
  {
    CK_I_WARMUP,
    CK_I_COUNTER,
    CK_I_TARGET,
    CK_I_WARMUP_TODO,
    CK_I_ETRIAL_TODO,
    CK_I_MAIN_STAGE,
    CK_I_HIST_LIMIT,
    CK_I_CUTOFF_SCHEME,
    CK_I_REFHIST,
    CK_I_RECOMPUTE,
    CK_I_CAPACITY,
    CK_I_TARGET_IN,
    CK_I_SEED,
    CK_I_NLPP,
    CK_I_TMOVE,
    CK_NI
  };

  std::vector<char> checkpoint_save()
  {
    require_walkers();
    if (dcount[0] || dcount[1])
      throw QmiError(QMI_ERR_ARG, "checkpoint with pending delayed updates: call qmi_mw_complete_updates first");
    const int N = sys.N, Np = sys.Npad;
    std::vector<double> t((size_t)nw * 3 * Np), Rh((size_t)nw * N * 3);
    HIPCHECK(hipMemcpyAsync(t.data(), R.p, t.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    for (int w = 0; w < nw; w++)
      for (int i = 0; i < N; i++)
        for (int q = 0; q < 3; q++)
          Rh[((size_t)w * N + i) * 3 + q] = t[((size_t)w * 3 + q) * Np + i];
    CkptWriter wr;
    wr.add_i64("state_0/block", {step_count});
    wr.add_i64("state_0/number_of_walkers", {nw});
    wr.add_i64("state_0/walker_partition", {0, nw});
    wr.add("state_0/walkers", CKPT_F64, {uint64_t(nw), uint64_t(N), 3}, Rh.data());
    std::vector<double> wts(nw, 1.0);
    if (dmc.on)
      std::copy(dmc.weight.begin(), dmc.weight.begin() + nw, wts.begin());
    wr.add_f64("state_0/walker_weights", wts);
    wr.add_i64("crowd_size", {wpc});
    const size_t ss = mt_state_size(), ns = rng_seeded ? rng.size() : 0;
    std::vector<uint32_t> words;
    words.reserve(ns * ss);
    for (size_t c = 0; c < ns; c++)
    {
      const std::vector<uint32_t> v = mt_save(rng[c]);
      words.insert(words.end(), v.begin(), v.end());
    }
    wr.add_i64("nprocs_nthreads_statesize", {1, (int64_t)ns, (int64_t)ss});
    wr.add("random/std::mt19937", CKPT_U32, {uint64_t(ns), uint64_t(ss)}, words.data());
    wr.add_i64("random_seed", {rng_seeded ? 1 : 0, (int64_t)rng_seed});
    if (dmc.on)
    {
      std::vector<double> v(CK_NV);
      v[CK_V_TAU]        = dmc.p.tau;
      v[CK_V_TAUEFF]     = dmc.tau_eff;
      v[CK_V_ETRIAL]     = dmc.etrial;
      v[CK_V_EREF]       = dmc.eref;
      v[CK_V_BRANCHMAX]  = dmc.branch_max;
      v[CK_V_CUTOFF]     = dmc.cutoff;
      v[CK_V_FILTER]     = dmc.filter;
      v[CK_V_SIGMA2]     = dmc.sigma2;
      v[CK_V_ENOW]       = dmc.enow;
      v[CK_V_R2ACC]      = dmc.r2acc;
      v[CK_V_R2PROP]     = dmc.r2prop;
      v[CK_V_FEEDBACK]   = dmc.p.feedback;
      v[CK_V_SIGMABOUND] = dmc.p.sigma_bound;
      v[CK_V_EHIST]      = dmc.ehist;
      v[CK_V_VHIST]      = dmc.vhist;
      v[CK_V_NHIST]      = dmc.nhist;
      wr.add_f64("branch/vparam", v);
      std::vector<int64_t> iv(CK_NI);
      iv[CK_I_WARMUP]        = dmc.p.warmup_steps;
      iv[CK_I_COUNTER]       = dmc.iter;
      iv[CK_I_TARGET]        = dmc.target;
      iv[CK_I_WARMUP_TODO]   = dmc.warmup_todo;
      iv[CK_I_ETRIAL_TODO]   = dmc.etrial_todo;
      iv[CK_I_MAIN_STAGE]    = dmc.main_stage ? 1 : 0;
      iv[CK_I_HIST_LIMIT]    = (int64_t)dmc.hist_limit;
      iv[CK_I_CUTOFF_SCHEME] = dmc.p.cutoff_scheme;
      iv[CK_I_REFHIST]       = dmc.p.refenergy_history;
      iv[CK_I_RECOMPUTE]     = dmc.p.recompute_period;
      iv[CK_I_CAPACITY]      = dmc.p.capacity;
      iv[CK_I_TARGET_IN]     = dmc.p.target_walkers;
      iv[CK_I_SEED]          = dmc.p.seed;
      iv[CK_I_NLPP]          = dmc.nlpp ? 1 : 0;
      iv[CK_I_TMOVE]         = dmc.p.nonlocal_moves;
      wr.add_i64("branch/iparam", iv);
      std::vector<double> h;
      for (auto& e : dmc.hist)
        h.insert(h.end(), e.begin(), e.end());
      wr.add("branch/history", CKPT_F64, {uint64_t(dmc.hist.size()), 3}, h.data());
      const std::vector<uint32_t> wc = mt_save(dmc.wc_rng);
      wr.add("random_master/std::mt19937", CKPT_U32, {1, uint64_t(wc.size())}, wc.data());
      wr.add_f64("dmc/local_energies", std::vector<double>(dmc.eloc.begin(), dmc.eloc.begin() + nw));
    }
    return wr.finish();
  }

  // Restart: the walkers are reloaded and re-evaluated from scratch (QMCDriverNew::
  // initialLogEvaluation, QMCDriverNew.cpp:258-), the streams, the step count and the DMC state
  // restored.  Everything is validated before the context changes.
  void checkpoint_load(const void* buf, size_t bytes)
  {
    if (!have_system)
      throw QmiError(QMI_ERR_ARG, "system not created");
    const int N = sys.N;
    try
    {
      CkptReader rd(buf, bytes);
      const int64_t nwc = rd.i64("state_0/number_of_walkers", 1)[0];
      if (nwc < 1 || nwc > (1 << 24))
        throw QmiError(QMI_ERR_ARG, "checkpoint walker count out of range");
      const CkptDataset& W = rd.need("state_0/walkers", CKPT_F64);
      if (W.dims.size() != 3 || W.dims[1] != (uint64_t)N || W.dims[2] != 3)
        throw QmiError(QMI_ERR_ARG, "checkpoint walkers do not have " + std::to_string(N) + " electrons");
      if (W.dims[0] != (uint64_t)nwc)
        throw QmiError(QMI_ERR_ARG, "checkpoint walkers and number_of_walkers disagree");
      std::vector<double> Rh(W.size());
      std::memcpy(Rh.data(), W.data, Rh.size() * sizeof(double));
      const std::vector<double> wts = rd.f64("state_0/walker_weights", nwc);
      const int64_t wpc_in          = rd.i64("crowd_size", 1)[0];
      const int64_t block           = rd.i64("state_0/block", 1)[0];
      const std::vector<int64_t> sh = rd.i64("nprocs_nthreads_statesize", 3);
      const std::vector<int64_t> sd = rd.i64("random_seed", 2);
      const size_t ss               = mt_state_size();
      if (sh[0] != 1 || sh[1] < 0 || sh[2] != (int64_t)ss)
        throw QmiError(QMI_ERR_ARG, "checkpoint random streams of another shape (processes, streams, state size)");
      const CkptDataset& RS = rd.need("random/std::mt19937", CKPT_U32, sh[1] * (int64_t)ss);
      std::vector<std::mt19937> streams(sh[1]);
      for (int64_t c = 0; c < sh[1]; c++)
        mt_load(streams[c], reinterpret_cast<const uint32_t*>(RS.data) + c * ss, ss);
      if (wpc_in < 1 || block < 0)
        throw QmiError(QMI_ERR_ARG, "checkpoint crowd size or block out of range");
      const int64_t ncrowd_in = (nwc + std::min<int64_t>(wpc_in, nwc) - 1) / std::min<int64_t>(wpc_in, nwc);
      const bool is_dmc = rd.find("branch/vparam") != nullptr;
      if (!is_dmc && sd[0] && sh[1] != ncrowd_in)
        throw QmiError(QMI_ERR_ARG, "checkpoint has one random stream per crowd: stream count does not match");
      DmcState d{};
      if (is_dmc)
      {
        const std::vector<double> v   = rd.f64("branch/vparam", CK_NV);
        const std::vector<int64_t> iv = rd.i64("branch/iparam", CK_NI);
        const CkptDataset& H          = rd.need("branch/history", CKPT_F64);
        const CkptDataset& WC         = rd.need("random_master/std::mt19937", CKPT_U32, (int64_t)ss);
        const std::vector<double> el  = rd.f64("dmc/local_energies", nwc);
        if (wpc_in != 1 || !sd[0] || sh[1] < nwc)
          throw QmiError(QMI_ERR_ARG, "DMC checkpoint needs one seeded random stream per walker slot");
        if (!coul_ready)
          throw QmiError(QMI_ERR_ARG, "DMC checkpoint: call qmi_coulomb_setup first");
        if (iv[CK_I_NLPP] && !nlpp_ready)
          throw QmiError(QMI_ERR_ARG, "DMC checkpoint with a pseudopotential term: call qmi_nlpp_setup first");
        if (iv[CK_I_NLPP] && cplx)
          throw QmiError(QMI_ERR_UNSUPPORTED, "the pseudopotential energy is built for real orbitals");
        if (learned)
          throw QmiError(QMI_ERR_UNSUPPORTED, "DMC with the learned proposal is not built");
        if (iv[CK_I_CAPACITY] < nwc || iv[CK_I_HIST_LIMIT] < 1 || H.dims.size() != 2 || H.dims[1] != 3 ||
            H.dims[0] > (uint64_t)iv[CK_I_HIST_LIMIT] || iv[CK_I_CUTOFF_SCHEME] < 0 || iv[CK_I_CUTOFF_SCHEME] > 3 ||
            !(v[CK_V_TAU] > 0) || iv[CK_I_TARGET] < 1 || iv[CK_I_TMOVE] < 0 || iv[CK_I_TMOVE] > 1 ||
            (iv[CK_I_TMOVE] && !iv[CK_I_NLPP]))
          throw QmiError(QMI_ERR_ARG, "DMC checkpoint branch state out of range");
        d.p.tau               = v[CK_V_TAU];
        d.p.feedback          = v[CK_V_FEEDBACK];
        d.p.sigma_bound       = v[CK_V_SIGMABOUND];
        d.p.warmup_steps      = (int)iv[CK_I_WARMUP];
        d.p.refenergy_history = (int)iv[CK_I_REFHIST];
        d.p.cutoff_scheme     = (int)iv[CK_I_CUTOFF_SCHEME];
        d.p.capacity          = (int)iv[CK_I_CAPACITY];
        d.p.target_walkers    = (int)iv[CK_I_TARGET_IN];
        d.p.recompute_period  = (int)iv[CK_I_RECOMPUTE];
        d.p.seed              = (uint32_t)iv[CK_I_SEED];
        d.p.nonlocal_moves    = (int)iv[CK_I_TMOVE];
        d.nlpp                = iv[CK_I_NLPP] != 0;
        d.iter                = iv[CK_I_COUNTER];
        d.tau_eff             = v[CK_V_TAUEFF];
        d.etrial              = v[CK_V_ETRIAL];
        d.eref                = v[CK_V_EREF];
        d.branch_max          = v[CK_V_BRANCHMAX];
        d.cutoff              = v[CK_V_CUTOFF];
        d.filter              = v[CK_V_FILTER];
        d.sigma2              = v[CK_V_SIGMA2];
        d.enow                = v[CK_V_ENOW];
        d.r2acc               = v[CK_V_R2ACC];
        d.r2prop              = v[CK_V_R2PROP];
        d.ehist               = v[CK_V_EHIST];
        d.vhist               = v[CK_V_VHIST];
        d.nhist               = v[CK_V_NHIST];
        d.target              = (int)iv[CK_I_TARGET];
        d.warmup_todo         = (int)iv[CK_I_WARMUP_TODO];
        d.etrial_todo         = (int)iv[CK_I_ETRIAL_TODO];
        d.main_stage          = iv[CK_I_MAIN_STAGE] != 0;
        d.hist_limit          = (size_t)iv[CK_I_HIST_LIMIT];
        const double* h       = reinterpret_cast<const double*>(H.data);
        for (uint64_t r = 0; r < H.dims[0]; r++)
          d.hist.push_back({h[3 * r], h[3 * r + 1], h[3 * r + 2]});
        mt_load(d.wc_rng, reinterpret_cast<const uint32_t*>(WC.data), ss);
        d.weight.assign(iv[CK_I_CAPACITY], 1.0);
        d.eloc.assign(iv[CK_I_CAPACITY], 0.0);
        std::copy(wts.begin(), wts.end(), d.weight.begin());
        std::copy(el.begin(), el.end(), d.eloc.begin());
        d.on = true;
      }
      // the context changes from here on
      walkers_load((int)nwc, Rh.data(), (int)wpc_in, is_dmc ? d.p.capacity : 0);
      {
        Scoped t(this, "jastrow_recompute");
        HIPCHECK(launch_jastrow_recompute(sys, st, nw, stream));
      }
      recompute_dets();
      sync();
      step_count = block;
      rng        = std::move(streams);
      rng_seed   = (uint32_t)sd[1];
      rng_seeded = sd[0] != 0;
      if (is_dmc)
      {
        dmc_rr.alloc((size_t)cap * 2);
        dmc_pairs.alloc((size_t)cap * 2);
        h_pairs.alloc((size_t)cap * 2);
        h_rr.alloc((size_t)cap * 2);
        dmc = std::move(d);
      }
    }
    catch (const QmiError&)
    {
      throw;
    }
    catch (const std::runtime_error& e)
    {
      throw QmiError(QMI_ERR_ARG, e.what());
    }
  }

  // ---------------------------------------------------------------------------------------------
  // RNG for one sub-step: Gaussians [iat][iw][3] per crowd (the reference layout of
  // walker_deltas, VMCBatched.cpp:109) scattered into the global [iat][w][3] array, then the crowd's
  // possible uniforms (N * crowd size), drawn speculatively and rewound after the sub-step.
  std::vector<std::mt19937> rng_saved;
  void rng_generate()
  {
    const int N = sys.N;
    h_gauss.alloc((size_t)N * nw * 3);
    h_unif.alloc((size_t)ncrowd * N * wpc);
    rng_saved.resize(ncrowd);
    auto work = [&](int c0, int c1) {
      std::vector<double> buf;
      for (int c = c0; c < c1; c++)
      {
        const int w0 = c * wpc, cw = std::min(nw, w0 + wpc) - w0;
        buf.resize((size_t)N * cw * 3);
        rng_fill_gauss(buf.data(), unsigned(buf.size()), rng[c]);
        for (int iat = 0; iat < N; iat++)
          std::memcpy(h_gauss.p + ((size_t)iat * nw + w0) * 3, buf.data() + (size_t)iat * cw * 3,
                      (size_t)cw * 3 * sizeof(double));
        rng_saved[c] = rng[c];
        rng_fill_uniform(h_unif.p + (size_t)c * N * wpc, (size_t)N * cw, rng[c]);
      }
    };
    const int nth = std::max(1, std::min<int>({ncrowd, 16, (int)std::thread::hardware_concurrency()}));
    if (nth == 1)
      work(0, ncrowd);
    else
    {
      std::vector<std::thread> th;
      const int per = (ncrowd + nth - 1) / nth;
      for (int i = 0; i < nth; i++)
        th.emplace_back(work, i * per, std::min(ncrowd, (i + 1) * per));
      for (auto& x : th)
        x.join();
    }
  }

  // run f(c0, c1) over the crowds on up to 16 host threads
  template<typename F>
  void over_crowds(F&& work)
  {
    const int nth = std::max(1, std::min<int>({ncrowd, 16, (int)std::thread::hardware_concurrency()}));
    if (nth == 1)
      work(0, ncrowd);
    else
    {
      std::vector<std::thread> th;
      const int per = (ncrowd + nth - 1) / nth;
      for (int i = 0; i < nth; i++)
        th.emplace_back(work, i * per, std::min(ncrowd, (i + 1) * per));
      for (auto& x : th)
        x.join();
    }
  }

  // The same numbers in two parts (one walker per crowd), so that the Box-Muller work of the
  // second half of the electrons overlaps the device's first half of the sweep.  Part 1: the
  // uniforms (drawn from a copy of the stream advanced past the sub-step's Gaussians, as the
  // sequential order puts them after all Gaussians) and the Gaussians of electrons [0, split_n);
  // part 2: the Gaussians of electrons [split_n, N).  The Gaussians are drawn in pairs, part 1
  // ending on an even count, so both parts together equal one sequential call.
  int split_n = 0; // electrons of part 1 (0: sub-steps are not split)
  int split_g1 = 0; // Gaussians drawn in part 1 (even)
  void rng_generate_part(int part)
  {
    const int N = sys.N, n = 3 * N;
    h_gauss.alloc((size_t)N * nw * 3);
    h_unif.alloc((size_t)ncrowd * N * wpc);
    rng_saved.resize(ncrowd);
    split_g1       = std::min(n, (3 * split_n + 1) / 2 * 2);
    const int g0   = part == 1 ? 0 : split_g1;
    const int g1   = part == 1 ? split_g1 : n;
    auto work = [&](int c0, int c1) {
      std::vector<double> buf;
      for (int c = c0; c < c1; c++)
      {
        if (part == 1)
        {
          std::mt19937 e = rng[c];
          e.discard(2ull * (unsigned long long)((n + 1) / 2)); // the draws of all n Gaussians
          rng_saved[c] = e;
          rng_fill_uniform(h_unif.p + (size_t)c * N, (size_t)N, e);
        }
        buf.resize((size_t)(g1 - g0));
        if (g1 > g0)
          rng_fill_gauss(buf.data(), unsigned(g1 - g0), rng[c]);
        for (int j = g0; j < g1; j++)
          h_gauss.p[((size_t)(j / 3) * nw + c) * 3 + j % 3] = buf[j - g0];
      }
    };
    over_crowds(work);
  }

  void rng_rewind(const std::vector<int>& consumed)
  {
    for (int c = 0; c < ncrowd; c++)
    {
      rng[c] = rng_saved[c];
      rng[c].discard((unsigned long long)consumed[c]);
    }
  }

  // ---------------------------------------------------------------------------------------------
  // learned proposal
  MlpParams mlp() const { return MlpParams{mlp_master.p, w1b.p, w2b.p, w3b.p, w4b.p}; }

  void mlp_alloc()
  {
    if (mlp_master.p)
      return;
    mlp_master.alloc(MLP_NPARAM);
    mlp_grad.alloc(MLP_NPARAM);
    adam_m.alloc(MLP_NPARAM);
    adam_v.alloc(MLP_NPARAM);
    w1b.alloc((size_t)MLP_HID * MLP_K1);
    w2b.alloc((size_t)MLP_HID * MLP_HID);
    w3b.alloc((size_t)MLP_HID * MLP_HID);
    w4b.alloc((size_t)16 * MLP_HID);
  }

  void mlp_set(const float* hparams)
  {
    mlp_alloc();
    HIPCHECK(hipMemcpyAsync(mlp_master.p, hparams, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, stream));
    HIPCHECK(launch_mlp_pack(mlp(), stream));
    HIPCHECK(hipMemsetAsync(adam_m.p, 0, MLP_NPARAM * sizeof(float), stream));
    HIPCHECK(hipMemsetAsync(adam_v.p, 0, MLP_NPARAM * sizeof(float), stream));
    sync();
    adam_t    = 0;
    mlp_ready = true;
  }

  // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases, zero output layer: the initial
  // proposal is plain drift-diffusion
  void mlp_init(uint32_t seed)
  {
    std::vector<float> P(MLP_NPARAM, 0.0f);
    std::mt19937 e(seed);
    auto fill = [&](int off, int count, int fan_in) {
      const double lim = 1.0 / std::sqrt((double)fan_in);
      for (int i = 0; i < count; i++)
        P[off + i] = (float)((rng_uniform01(e) * 2.0 - 1.0) * lim);
    };
    fill(MLP_OFF_W1, MLP_HID * MLP_IN, MLP_IN);
    fill(MLP_OFF_B1, MLP_HID, MLP_IN);
    fill(MLP_OFF_W2, MLP_HID * MLP_HID, MLP_HID);
    fill(MLP_OFF_B2, MLP_HID, MLP_HID);
    fill(MLP_OFF_W3, MLP_HID * MLP_HID, MLP_HID);
    fill(MLP_OFF_B3, MLP_HID, MLP_HID);
    mlp_set(P.data());
  }

  void learn_alloc(int B)
  {
    mlp_alloc();
    if (lX.n < (size_t)B * MLP_IN)
    {
      lX.alloc((size_t)B * MLP_IN);
      lout.alloc((size_t)B * MLP_OUT);
      ldout.alloc((size_t)B * MLP_OUT);
      lz.alloc((size_t)3 * B * MLP_HID);
      ldz_a.alloc((size_t)B * MLP_HID);
      ldz_b.alloc((size_t)B * MLP_HID);
    }
  }

  LearnArgs learn_args(bool with_train)
  {
    LearnArgs la{};
    la.X    = lX.p;
    la.out  = lout.p;
    la.ls   = lstate.p;
    la.dout = with_train ? ldout.p : nullptr;
    la.fsum   = lfsum.p;
    la.ltrace = ltrace_on ? ltrace_d.p : nullptr;
    return la;
  }

  // one learned move of electron iat for every walker (a: the sub-step's move arguments)
  void learned_move(MoveArgs& a)
  {
    const SysC& s = sys;
    LearnArgs la  = learn_args(train);
    const MlpParams P = mlp();
    if (!train && fused_learn)
    {
      {
        Scoped t(this, "learn_move");
        HIPCHECK(launch_learn(LK_FUSED, s, st, coefs.p, a, la, stream, &P));
      }
      learn_moves += 1;
      return;
    }
    {
      Scoped t(this, "learn_pre");
      HIPCHECK(launch_learn(LK_PRE, s, st, coefs.p, a, la, stream));
    }
    {
      Scoped t(this, "mlp_forward");
      HIPCHECK(launch_mlp_forward(P, nw, lX.p, lout.p, train ? lz.p : nullptr, 2 * nw, 0, stream));
    }
    {
      Scoped t(this, "learn_prop");
      HIPCHECK(launch_learn(LK_PROP, s, st, coefs.p, a, la, stream));
    }
    {
      Scoped t(this, "mlp_forward");
      HIPCHECK(launch_mlp_forward(P, nw, lX.p + (size_t)nw * MLP_IN, lout.p + (size_t)nw * MLP_OUT,
                                  train ? lz.p : nullptr, 2 * nw, nw, stream));
    }
    {
      Scoped t(this, "learn_accept");
      HIPCHECK(launch_learn(LK_ACCEPT, s, st, coefs.p, a, la, stream));
    }
    if (train)
    {
      Scoped t(this, "mlp_train");
      HIPCHECK(launch_mlp_backward(P, 2 * nw, lX.p, lz.p, ldout.p, mlp_grad.p, ldz_a.p, ldz_b.p, stream));
      HIPCHECK(launch_adam(mlp_master.p, mlp_grad.p, adam_m.p, adam_v.p, MLP_NPARAM, lr, ++adam_t, stream));
      HIPCHECK(launch_mlp_pack(P, stream));
    }
    learn_moves += 1;
  }

  // device buffers and pinned host buffers of one sub-step (allocated before any capture)
  void substep_buffers(bool with_trace)
  {
    const int N = sys.N;
    // sized for the walker slots (a DMC population changes nw between sub-steps); reallocation
    // only on growth, which invalidates captured graphs through walkers_load
    if (gauss.n < (size_t)N * cap * 3)
      gauss.alloc((size_t)N * cap * 3);
    const size_t nunif = (size_t)((cap + wpc - 1) / wpc) * wpc * N; // whole crowds
    if (unif.n < nunif)
      unif.alloc(nunif);
    if (with_trace && trace.n < (size_t)N * cap * TRACE_FIELDS)
      trace.alloc((size_t)N * cap * TRACE_FIELDS);
    h_gauss.alloc((size_t)N * cap * 3);
    h_unif.alloc(nunif);
    h_ubase.alloc(ubase.n);
    if (!ev_ubase)
      HIPCHECK(hipEventCreateWithFlags(&ev_ubase, hipEventDisableTiming));
  }

  void seed_rng(uint32_t seed)
  {
    if (!rng_seeded || seed != rng_seed)
    {
      rng.clear();
      for (int c = 0; c < ncrowd; c++)
        rng.emplace_back(seed + uint32_t(c));
      rng_seed   = seed;
      rng_seeded = true;
    }
  }

  // One sub-step on the context's stream: RNG upload (from the pinned buffers filled by
  // rng_generate), every move of both spin groups with their flushes, and the per-crowd uniform
  // counts back to pinned memory.  Contains no memset and no pageable copy, so it can be captured.
  // part 0: the whole sub-step; 1: its moves of electrons [0, split_n); 2: the rest and the closing
  // flushes (two launches around the second part of the host's random numbers)
  void enqueue_substep(double tau, bool with_trace, bool mark_counts = false, int part = 0)
  {
    const SysC& s = sys;
    const int N   = s.N;
    const int e0 = part == 2 ? split_n : 0, e1 = part == 1 ? split_n : N;
    if (part != 2)
    {
      HIPCHECK(hipMemcpyAsync(unif.p, h_unif.p, (size_t)ncrowd * N * wpc * sizeof(double), hipMemcpyHostToDevice,
                              stream));
      HIPCHECK(launch_fill_i32(ubase.p, ubase.n, 0, stream));
    }
    HIPCHECK(hipMemcpyAsync(gauss.p + (size_t)e0 * nw * 3, h_gauss.p + (size_t)e0 * nw * 3,
                            (size_t)(e1 - e0) * nw * 3 * sizeof(double), hipMemcpyHostToDevice, stream));
    MoveArgs a{};
    a.nw    = nw;
    a.wpc   = wpc;
    a.tau   = tau;
    a.gauss = gauss.p;
    a.unif  = unif.p;
    a.ubase = ubase.p;
    a.trace = with_trace ? trace.p : nullptr;
    a.dmc   = dmc.on ? 1 : 0;
    a.rr    = dmc.on ? dmc_rr.p : nullptr;
    a.nanflag = nanflag.p;
    const int P = partitions();
    if (P > 1)
    {
      enqueue_partitioned(a, P, e0, e1, part, mark_counts);
      return;
    }
    for (int g = 0; g < 2; g++)
    {
      int iat        = std::max(s.first[g], e0);
      const int last = std::min(s.first[g] + s.n[g], e1);
      while (iat < last)
      {
        a.g  = g;
        a.d0 = s.k > 1 ? dcount[g] : 0;
        if (learned)
        {
          a.iat0 = iat;
          a.iat1 = iat + 1;
          learned_move(a);
          iat += 1;
          if (s.k > 1)
            dcount[g] += 1;
        }
        else if (wpc == 1 && (move_one || cplx))
        {
          a.iat0 = iat;
          a.iat1 = iat + 1;
          {
            Scoped t(this, "move_one");
            HIPCHECK(move_kernel(MK_MOVE_ONE, a));
          }
          iat += 1;
          if (s.k > 1)
            dcount[g] += 1;
        }
        else if (wpc == 1)
        {
          const int seg = s.k > 1 ? std::min(s.k - dcount[g], last - iat) : last - iat;
          a.iat0        = iat;
          a.iat1        = iat + seg;
          {
            Scoped t(this, "sweep_segment");
            HIPCHECK(launch_move(MK_SEGMENT, s, st, coefs.p, a, stream));
          }
          iat += seg;
          if (s.k > 1)
            dcount[g] += seg;
        }
        else
        {
          a.iat0 = iat;
          a.iat1 = iat + 1;
          {
            Scoped t(this, "propose");
            HIPCHECK(move_kernel(MK_PROPOSE, a));
          }
          {
            Scoped t(this, "accept_crowd");
            HIPCHECK(move_kernel(MK_ACCEPT_CROWD, a));
          }
          iat += 1;
          if (s.k > 1)
            dcount[g] += 1;
        }
        if (s.k > 1 && dcount[g] == s.k)
          flush(g);
      }
    }
    if (part == 1)
      return;
    // the uniform counts are final after the last move: copy them before the closing flushes, so
    // that the host can rewind and draw the next sub-step's numbers while the flushes run
    HIPCHECK(hipMemcpyAsync(h_ubase.p, ubase.p, ubase.n * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (mark_counts)
      HIPCHECK(hipEventRecord(ev_ubase, stream));
    flush(0);
    flush(1);
  }

  // moves [e0, e1) of the sub-step (and, unless part == 1, the closing flushes) for P walker
  // partitions, each on its own stream forked from the context stream; the uniform counts reach the
  // host once every partition's moves are done, before the closing flushes
  void enqueue_partitioned(const MoveArgs& a, int P, int e0, int e1, int part, bool mark_counts)
  {
    const SysC& s = sys;
    ensure_partitions(P);
    HIPCHECK(hipEventRecord(pev_fork, stream));
    int dc[2] = {dcount[0], dcount[1]};
    for (int p = 0; p < P; p++)
    {
      hipStream_t ps = pstream[p];
      const int w0 = (int)((long long)p * nw / P), w1 = (int)((long long)(p + 1) * nw / P);
      HIPCHECK(hipStreamWaitEvent(ps, pev_fork, 0));
      MoveArgs b = a;
      b.w0       = w0;
      b.wn       = w1 - w0;
      dc[0]      = dcount[0];
      dc[1]      = dcount[1];
      for (int g = 0; g < 2; g++)
      {
        const int last = std::min(s.first[g] + s.n[g], e1);
        for (int iat = std::max(s.first[g], e0); iat < last;)
        {
          // one launch per move, or with the per-window kernel (real orbitals) one per delay window
          const bool window = !move_one && !cplx;
          const int seg     = !window ? 1 : s.k > 1 ? std::min(s.k - dc[g], last - iat) : last - iat;
          b.g               = g;
          b.d0              = s.k > 1 ? dc[g] : 0;
          b.iat0            = iat;
          b.iat1            = iat + seg;
          HIPCHECK(cplx ? launch_move_c(MK_MOVE_ONE, s, st, coefs.p, b, ps)
                        : launch_move(window ? MK_SEGMENT : MK_MOVE_ONE, s, st, coefs.p, b, ps));
          iat += seg;
          if (s.k > 1 && (dc[g] += seg) == s.k)
          {
            flush_part(g, dc[g], w0, w1 - w0, ps);
            dc[g] = 0;
          }
        }
      }
      HIPCHECK(hipEventRecord(pev_moves[p], ps));
    }
    dcount[0] = dc[0];
    dcount[1] = dc[1];
    invrow_id = -1;
    if (part != 1)
    {
      for (int p = 0; p < P; p++)
        HIPCHECK(hipStreamWaitEvent(stream, pev_moves[p], 0));
      HIPCHECK(hipMemcpyAsync(h_ubase.p, ubase.p, ubase.n * sizeof(int), hipMemcpyDeviceToHost, stream));
      if (mark_counts)
        HIPCHECK(hipEventRecord(ev_ubase, stream));
      for (int p = 0; p < P; p++)
      {
        const int w0 = (int)((long long)p * nw / P), w1 = (int)((long long)(p + 1) * nw / P);
        flush_part(0, dcount[0], w0, w1 - w0, pstream[p]);
        flush_part(1, dcount[1], w0, w1 - w0, pstream[p]);
        HIPCHECK(hipEventRecord(pev_done[p], pstream[p]));
      }
      dcount[0] = dcount[1] = 0;
    }
    for (int p = 0; p < P; p++)
      HIPCHECK(hipStreamWaitEvent(stream, part != 1 ? pev_done[p] : pev_moves[p], 0));
  }

  // after the sub-step's work has completed: rewind each crowd's generator to the uniforms used
  void finish_substep()
  {
    const int N = sys.N;
    std::vector<int> consumed(ncrowd);
    for (int c = 0; c < ncrowd; c++)
      consumed[c] = h_ubase.p[(size_t)c * (N + 1) + N];
    rng_rewind(consumed);
  }

  // capture enqueue_substep(tau) into sweep_exec (kept until the system, walkers, tau or a tuning
  // knob change)
  void ensure_graph(double tau)
  {
    if (sweep_exec && graph_tau == tau)
      return;
    invalidate_graph();
    hipGraph_t graph = nullptr;
    capture(stream, [&] { enqueue_substep(tau, false); }, &graph);
    const hipError_t e = hipGraphInstantiate(&sweep_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(e);
    graph_tau = tau;
  }

  // the two graphs of a split sub-step
  void ensure_graph_split(double tau)
  {
    if (split_exec[0] && graph_tau == tau)
      return;
    invalidate_graph();
    for (int part = 1; part <= 2; part++)
    {
      hipGraph_t graph = nullptr;
      capture(stream, [&] { enqueue_substep(tau, false, false, part); }, &graph);
      const hipError_t e = hipGraphInstantiate(&split_exec[part - 1], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHECK(e);
    }
    graph_tau = tau;
  }

  // record f()'s work on stream s (and the streams it joins) into *graph
  template<typename F>
  void capture(hipStream_t s, F&& f, hipGraph_t* graph)
  {
    const bool t_on = timing.on;
    timing.on       = false;
    HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try
    {
      f();
    }
    catch (...)
    {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g)
        (void)hipGraphDestroy(g);
      timing.on = t_on;
      throw;
    }
    timing.on = t_on;
    HIPCHECK(hipStreamEndCapture(s, graph));
  }

  void vmc_run(int steps, int sub_steps, double tau, int recompute_period, uint32_t seed, double* energies,
               double* htrace, double* hltrace = nullptr)
  {
    require_walkers();
    const SysC& s = sys;
    const int N   = s.N;
    if (steps < 0 || sub_steps < 1 || !(tau > 0))
      throw QmiError(QMI_ERR_ARG, "invalid steps / sub_steps / tau");
    end_dmc();
    seed_rng(seed);
    substep_buffers(htrace != nullptr);
    if (learned && cplx)
      throw QmiError(QMI_ERR_UNSUPPORTED, "the learned proposal needs real orbitals");
    if (learned)
    {
      if (wpc != 1)
        throw QmiError(QMI_ERR_UNSUPPORTED, "the learned proposal needs one walker per crowd (one RNG stream each)");
      if (!s.has_det || s.nion > 1280)
        throw QmiError(QMI_ERR_UNSUPPORTED, "the learned proposal needs orbitals and at most 1280 ions");
      if (!mlp_ready)
        mlp_init(12345u);
      learn_alloc(2 * nw);
      if (hltrace)
        ltrace_d.alloc((size_t)N * nw * LTRACE_FIELDS);
      ltrace_on = hltrace != nullptr;
    }
    std::vector<double> ek(nw);
    // a partitioned sub-step runs eagerly: replayed from a captured graph its two chains took as long
    // as one after the other (measured at 1024 x 512: 56 ms per sweep from a graph, 43 ms eager)
    const bool use_graph = graph_on && !htrace && !learned && partitions() == 1;
    // one walker per crowd: split each sub-step after its first N/8 electrons, so that the Gaussians
    // of the rest are drawn while the device runs the first moves (the first part's share hides
    // behind the previous step's evaluateGL)
    split_n          = (wpc == 1 && N >= 2) ? std::min(N - 1, std::max(1, N / 8)) : 0;
    const bool split = split_n > 0;
    auto gen1        = [&] {
      if (split)
        rng_generate_part(1);
      else
        rng_generate();
    };
    // The host draws a sub-step's random numbers while the device finishes the previous one (its
    // closing flushes, evaluateGL): the draws depend only on the uniforms already consumed, whose
    // counts arrive before those kernels run.
    bool rng_ready = false;
    for (int step = 0; step < steps; step++)
    {
      for (int sub = 0; sub < sub_steps; sub++)
      {
        if (!rng_ready)
          gen1();
        rng_ready = false;
        if (split)
        {
          if (use_graph)
          {
            ensure_graph_split(tau);
            HIPCHECK(hipGraphLaunch(split_exec[0], stream));
          }
          else
            enqueue_substep(tau, htrace != nullptr, false, 1);
          rng_generate_part(2);
          if (use_graph)
            HIPCHECK(hipGraphLaunch(split_exec[1], stream));
          else
            enqueue_substep(tau, htrace != nullptr, true, 2);
        }
        else if (use_graph)
        {
          ensure_graph(tau);
          HIPCHECK(hipGraphLaunch(sweep_exec, stream));
        }
        else
          enqueue_substep(tau, htrace != nullptr, true);
        if (htrace)
          HIPCHECK(hipMemcpyAsync(htrace + ((size_t)step * sub_steps + sub) * N * nw * TRACE_FIELDS, trace.p,
                                  (size_t)N * nw * TRACE_FIELDS * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (learned && hltrace)
          HIPCHECK(hipMemcpyAsync(hltrace + ((size_t)step * sub_steps + sub) * N * nw * LTRACE_FIELDS, ltrace_d.p,
                                  (size_t)N * nw * LTRACE_FIELDS * sizeof(double), hipMemcpyDeviceToHost, stream));
        // a graph replay is waited for as a whole: an event recorded during capture does not
        // signal the replay here (measured: the host read the counts before they were written)
        if (use_graph || htrace || hltrace)
          sync();
        else
          HIPCHECK(hipEventSynchronize(ev_ubase));
        finish_substep();
        if (sub + 1 < sub_steps)
        {
          gen1();
          rng_ready = true;
        }
      }
      const bool recompute = recompute_period > 0 && ((step_count + 1) % recompute_period == 0);
      if (recompute)
        recompute_dets();
      eval_gl_enqueue();
      if (step + 1 < steps)
      {
        gen1();
        rng_ready = true;
      }
      sync();
      eval_gl_finish(ek.data());
      if (energies)
        std::copy(ek.begin(), ek.end(), energies + (size_t)step * nw);
      step_count++;
    }
  }
};

//--------------------------------------------------------------------------------------------------
// A stateful distance table over the walkers of one context (DistanceTable, Particle/DistanceTable.h):
// the electron positions are the context's; the table, the proposal rows and (AA) the rows of the
// current position live in device memory sized for the context's walkers at creation.
struct qmi_dtable
{
  qmi_ctx* c = nullptr; // null once the context is destroyed
  int device = 0;
  int kind = DT_AA, nw = 0, N = 0, ns = 0;
  DevBuf<double> r, dr, temp_r, temp_dr, old_r, old_dr, rnew;
  DevBuf<int> acc;

  DTableArrays arrays()
  {
    DTableArrays t;
    t.kind    = kind;
    t.ns      = ns;
    t.r       = r.p;
    t.dr      = dr.p;
    t.temp_r  = temp_r.p;
    t.temp_dr = temp_dr.p;
    t.old_r   = old_r.p;
    t.old_dr  = old_dr.p;
    return t;
  }
  void check() const
  {
    if (c->nw != nw || c->sys.N != N || c->device != device)
      throw QmiError(QMI_ERR_ARG, "the context's walkers changed since the distance table was created");
  }
};

void qmi_ctx::detach_dtables()
{
  for (qmi_dtable* t : dtables)
    t->c = nullptr;
  dtables.clear();
}

//--------------------------------------------------------------------------------------------------
// Several contexts (walker populations of different systems: electron counts, cells, tables, real
// or complex orbitals) advanced together: one sub-step of every context is captured into one HIP
// graph, each context on its own stream (concurrent branches) or chained one after the other.
struct qmi_batch
{
  std::vector<qmi_ctx*> ctxs;
  int device = 0;
  hipStream_t origin = nullptr;
  hipEvent_t ev_fork = nullptr;
  hipGraphExec_t exec = nullptr;
  double exec_tau     = 0;
  bool concurrent     = true;
  std::vector<unsigned long long> exec_gen; // graph_gen of every context when exec was captured
  std::string err;

  // the two graphs of a split batched sub-step (all contexts' first parts, then their second parts)
  hipGraphExec_t split[2] = {nullptr, nullptr};
  double split_tau        = 0;
  std::vector<unsigned long long> split_gen;

  void invalidate()
  {
    if (exec || split[0] || split[1])
      (void)hipStreamSynchronize(origin);
    if (exec)
      (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    for (auto& x : split)
    {
      if (x)
        (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
  }

  bool splittable() const
  {
    for (auto* c : ctxs)
      if (c->wpc != 1 || c->sys.N < 2)
        return false;
    return true;
  }

  void prepare_split(double tau)
  {
    for (auto* c : ctxs)
    {
      c->require_walkers();
      if (c->learned)
        throw QmiError(QMI_ERR_UNSUPPORTED, "batched sweeps run the drift-diffusion proposal");
      c->substep_buffers(false);
      if (!c->ev_done)
        HIPCHECK(hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming));
      c->split_n = std::min(c->sys.N - 1, std::max(1, c->sys.N / 8));
    }
    bool stale = !split[0] || split_tau != tau || split_gen.size() != ctxs.size();
    for (size_t i = 0; !stale && i < ctxs.size(); i++)
      stale = split_gen[i] != ctxs[i]->graph_gen;
    if (!stale)
      return;
    invalidate();
    for (int part = 1; part <= 2; part++)
    {
      hipGraph_t graph = nullptr;
      ctxs[0]->capture(origin, [&] {
        HIPCHECK(hipEventRecord(ev_fork, origin));
        hipEvent_t prev = ev_fork;
        for (auto* c : ctxs)
        {
          HIPCHECK(hipStreamWaitEvent(c->stream, concurrent ? ev_fork : prev, 0));
          c->enqueue_substep(tau, false, false, part);
          HIPCHECK(hipEventRecord(c->ev_done, c->stream));
          prev = c->ev_done;
        }
        for (auto* c : ctxs)
          HIPCHECK(hipStreamWaitEvent(origin, c->ev_done, 0));
      }, &graph);
      const hipError_t e = hipGraphInstantiate(&split[part - 1], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHECK(e);
    }
    split_tau = tau;
    split_gen.clear();
    for (auto* c : ctxs)
      split_gen.push_back(c->graph_gen);
  }

  // one batched sub-step in two graphs around the second part of every context's numbers
  void substep_split(double tau, bool numbers_ready)
  {
    prepare_split(tau);
    if (!numbers_ready)
      for (auto* c : ctxs)
        c->rng_generate_part(1);
    HIPCHECK(hipGraphLaunch(split[0], origin));
    for (auto* c : ctxs)
      c->rng_generate_part(2);
    HIPCHECK(hipGraphLaunch(split[1], origin));
    HIPCHECK(hipStreamSynchronize(origin));
    for (auto* c : ctxs)
      c->finish_substep();
  }

  ~qmi_batch()
  {
    invalidate();
    if (ev_fork)
      (void)hipEventDestroy(ev_fork);
    if (origin)
      (void)hipStreamDestroy(origin);
  }

  void prepare(double tau)
  {
    for (auto* c : ctxs)
    {
      c->require_walkers();
      if (c->learned)
        throw QmiError(QMI_ERR_UNSUPPORTED, "batched sweeps run the drift-diffusion proposal");
      c->substep_buffers(false);
      if (!c->ev_done)
        HIPCHECK(hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming));
    }
    // the batch graph is rebuilt when any context changed buffers or knobs since it was captured
    bool stale = !exec || exec_tau != tau || exec_gen.size() != ctxs.size();
    for (size_t i = 0; !stale && i < ctxs.size(); i++)
      stale = exec_gen[i] != ctxs[i]->graph_gen;
    if (!stale)
      return;
    invalidate();
    hipGraph_t graph = nullptr;
    ctxs[0]->capture(origin, [&] {
      HIPCHECK(hipEventRecord(ev_fork, origin));
      hipEvent_t prev = ev_fork;
      for (auto* c : ctxs)
      {
        HIPCHECK(hipStreamWaitEvent(c->stream, concurrent ? ev_fork : prev, 0));
        c->enqueue_substep(tau, false);
        HIPCHECK(hipEventRecord(c->ev_done, c->stream));
        prev = c->ev_done;
      }
      for (auto* c : ctxs)
        HIPCHECK(hipStreamWaitEvent(origin, c->ev_done, 0));
    }, &graph);
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(e);
    exec_tau = tau;
    exec_gen.clear();
    for (auto* c : ctxs)
      exec_gen.push_back(c->graph_gen);
  }

  // one sub-step of every context: host RNG, one graph launch, rewind
  void substep(double tau, bool numbers_ready = false)
  {
    prepare(tau);
    if (!numbers_ready)
      for (auto* c : ctxs)
        c->rng_generate();
    HIPCHECK(hipGraphLaunch(exec, origin));
    HIPCHECK(hipStreamSynchronize(origin));
    for (auto* c : ctxs)
      c->finish_substep();
  }

  void run(int steps, int sub_steps, double tau, int recompute_period, const uint32_t* seeds,
           double* const* energies)
  {
    if (steps < 0 || sub_steps < 1 || !(tau > 0))
      throw QmiError(QMI_ERR_ARG, "invalid steps / sub_steps / tau");
    for (size_t i = 0; i < ctxs.size(); i++)
    {
      ctxs[i]->require_walkers();
      ctxs[i]->end_dmc();
      ctxs[i]->seed_rng(seeds ? seeds[i] : ctxs[i]->rng_seed);
    }
    std::vector<std::vector<double>> ek(ctxs.size());
    const bool two = splittable();
    auto gen       = [&] {
      for (auto* c : ctxs)
      {
        if (two)
        {
          c->split_n = std::min(c->sys.N - 1, std::max(1, c->sys.N / 8));
          c->rng_generate_part(1);
        }
        else
          c->rng_generate();
      }
    };
    bool ready = false; // the next sub-step's numbers (or their first part) are drawn
    for (int step = 0; step < steps; step++)
    {
      for (int sub = 0; sub < sub_steps; sub++)
      {
        if (two)
          substep_split(tau, ready);
        else
          substep(tau, ready);
        ready = false;
        if (sub + 1 < sub_steps)
        {
          gen();
          ready = true;
        }
      }
      for (size_t i = 0; i < ctxs.size(); i++)
      {
        qmi_ctx* c           = ctxs[i];
        const bool recompute = recompute_period > 0 && ((c->step_count + 1) % recompute_period == 0);
        if (recompute)
          c->recompute_dets();
      }
      // evaluateGL of every context on its own stream; the host draws the next step's numbers
      // meanwhile
      for (auto* c : ctxs)
        c->eval_gl_enqueue();
      if (step + 1 < steps)
      {
        gen();
        ready = true;
      }
      for (size_t i = 0; i < ctxs.size(); i++)
      {
        qmi_ctx* c = ctxs[i];
        c->sync();
        ek[i].resize(c->nw);
        c->eval_gl_finish(ek[i].data());
        if (energies && energies[i])
          std::copy(ek[i].begin(), ek[i].end(), energies[i] + (size_t)step * c->nw);
        c->step_count++;
      }
    }
  }

  // wall time of one batched sub-step (ms), median of `trials` after one warm-up
  double time_substep(double tau, int trials)
  {
    substep(tau);
    std::vector<double> t;
    for (int i = 0; i < trials; i++)
    {
      const auto a = std::chrono::steady_clock::now();
      substep(tau);
      t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
  }

  // wall time of one sub-step of context c alone (its own captured graph), median of `trials`
  double time_ctx(qmi_ctx* c, double tau, int trials)
  {
    auto one = [&] {
      c->ensure_graph(tau);
      c->rng_generate();
      HIPCHECK(hipGraphLaunch(c->sweep_exec, c->stream));
      c->sync();
      c->finish_substep();
    };
    one();
    std::vector<double> t;
    for (int i = 0; i < trials; i++)
    {
      const auto a = std::chrono::steady_clock::now();
      one();
      t.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
  }

  // Coordinate search per context, each timed alone on its own graph (the time it holds the GPU's
  // resources): delay rank, then the fused-move form (one launch per move / per delay window; real
  // orbitals, one walker per crowd), then the flush form (matrix-core pass / three LDS passes;
  // real orbitals, n <= 512).  Finally concurrent vs chained branches for the whole batch.  The
  // chains advance during tuning (equilibration moves of a valid Markov chain).
  void autotune(double tau, int trials, int* choice, double* best_ms)
  {
    trials = std::max(1, trials);
    for (size_t i = 0; i < ctxs.size(); i++)
    {
      qmi_ctx* c = ctxs[i];
      c->require_walkers();
      c->end_dmc();
      if (c->learned)
        throw QmiError(QMI_ERR_UNSUPPORTED, "batched sweeps run the drift-diffusion proposal");
      if (!c->rng_seeded)
        c->seed_rng(1u + uint32_t(i));
      c->substep_buffers(false);
    }
    auto apply = [&](qmi_ctx* c, int k, int mv, int fl) {
      c->set_delay_rank(k);
      if (c->move_one != mv || c->flush_variant != fl)
      {
        c->move_one      = mv;
        c->flush_variant = fl;
        c->invalidate_graph();
      }
    };
    for (size_t i = 0; i < ctxs.size(); i++)
    {
      qmi_ctx* c     = ctxs[i];
      const int nmax = std::max(c->sys.nup, c->sys.ndn);
      int bk = c->sys.k, bm = c->move_one, bf = c->flush_variant;
      double best = std::numeric_limits<double>::infinity();
      auto trial  = [&](int k, int mv, int fl) {
        apply(c, k, mv, fl);
        const double t = time_ctx(c, tau, trials);
        if (t < best)
        {
          best = t;
          bk   = k;
          bm   = mv;
          bf   = fl;
        }
      };
      if (c->sys.has_det)
      {
        for (int k : {1, 8, 16, 32, 64})
          if (k == 1 || k <= nmax)
            trial(k, bm, bf);
        if (!c->cplx && c->wpc == 1 && bk > 1)
          trial(bk, 1 - bm, bf);
        if (!c->cplx && bk > 1 && nmax <= 512)
          trial(bk, bm, 1 - bf);
      }
      else
        trial(bk, bm, bf);
      apply(c, bk, bm, bf);
      if (choice)
      {
        choice[3 * i + 0] = bk;
        choice[3 * i + 1] = 1 - bm; // 0: one launch per move, 1: one launch per delay window
        choice[3 * i + 2] = bf;
      }
      if (best_ms)
        best_ms[i] = best;
    }
    double tb[2];
    for (int cc = 0; cc < 2; cc++)
    {
      concurrent = cc == 1;
      invalidate();
      tb[cc] = time_substep(tau, trials);
    }
    concurrent = tb[1] <= tb[0];
    invalidate();
    if (best_ms)
    {
      best_ms[ctxs.size()]     = tb[0];
      best_ms[ctxs.size() + 1] = tb[1];
    }
  }
};

//--------------------------------------------------------------------------------------------------
namespace
{
template<typename F>
int guarded(qmi_ctx* c, F&& f)
{
  if (!c)
    return QMI_ERR_ARG;
  try
  {
    HIPCHECK(hipSetDevice(c->device));
    f();
    c->err.clear();
    return QMI_OK;
  }
  catch (const QmiError& e)
  {
    c->err = e.what();
    return e.code;
  }
  catch (const std::exception& e)
  {
    c->err = e.what();
    return QMI_ERR_DEVICE;
  }
}
} // namespace

extern "C"
{
  int qmi_ctx_create(qmi_ctx** out, int device, int precision)
  {
    if (!out)
      return QMI_ERR_ARG;
    *out = nullptr;
    if (precision < QMI_PREC_FP64 || precision > QMI_PREC_TABLE_BF16)
      return QMI_ERR_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
      return QMI_ERR_DEVICE;
    auto* c      = new qmi_ctx;
    c->device    = device;
    c->precision = precision;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess)
    {
      delete c;
      return QMI_ERR_DEVICE;
    }
    if (const char* mk = std::getenv("QMI_MOVE_KERNEL"))
      c->move_one = std::string(mk) != "segment";
    if (const char* lf = std::getenv("QMI_LEARN_FUSED"))
      c->fused_learn = std::string(lf) != "0";
    if (const char* np = std::getenv("QMI_PARTITIONS"))
      c->npart = std::max(1, std::min(8, std::atoi(np)));
    *out = c;
    return QMI_OK;
  }

  int qmi_ctx_destroy(qmi_ctx* c)
  {
    if (!c)
      return QMI_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
    c->detach_dtables();
    delete c;
    return QMI_OK;
  }

  const char* qmi_last_error(const qmi_ctx* c) { return c ? c->err.c_str() : "null context"; }

  int qmi_system_create(qmi_ctx* c, const qmi_system_desc* d)
  {
    return guarded(c, [&] { c->system_create(d); });
  }

  int qmi_walkers_load(qmi_ctx* c, int nw, const double* R, int wpc)
  {
    return guarded(c, [&] { c->walkers_load(nw, R, wpc); });
  }

  int qmi_walkers_get(qmi_ctx* c, double* Rout)
  {
    return guarded(c, [&] {
      c->require_walkers();
      const int N = c->sys.N, Np = c->sys.Npad;
      std::vector<double> t((size_t)c->nw * 3 * Np);
      HIPCHECK(hipMemcpyAsync(t.data(), c->R.p, t.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      for (int w = 0; w < c->nw; w++)
        for (int i = 0; i < N; i++)
          for (int q = 0; q < 3; q++)
            Rout[((size_t)w * N + i) * 3 + q] = t[((size_t)w * 3 + q) * Np + i];
    });
  }

  int qmi_mw_evaluate_log(qmi_ctx* c, double* logpsi, double* G, double* L)
  {
    return guarded(c, [&] {
      c->require_walkers();
      {
        qmi_ctx::Scoped t(c, "jastrow_recompute");
        HIPCHECK(launch_jastrow_recompute(c->sys, c->st, c->nw, c->stream));
      }
      c->dcount[0] = c->dcount[1] = 0;
      c->recompute_dets();
      c->eval_gl(logpsi, G, L, nullptr);
    });
  }

  int qmi_mw_eval_grad(qmi_ctx* c, int iat, double* grads)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (iat < 0 || iat >= c->sys.N || !grads)
        throw QmiError(QMI_ERR_ARG, "bad electron index");
      MoveArgs a = c->move_args(iat);
      a.out      = c->out.p;
      {
        qmi_ctx::Scoped t(c, "api_eval_grad");
        HIPCHECK(c->move_kernel(MK_API_GRAD, a));
      }
      HIPCHECK(hipMemcpyAsync(grads, c->out.p, (size_t)c->nw * 3 * c->vw() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
      c->sync();
      if (c->sys.k > 1)
        c->invrow_id = iat;
      for (int i = 0; i < 3 * c->nw * c->vw(); i++)
        if (!std::isfinite(grads[i]))
          throw QmiError(QMI_ERR_NUMERIC, "NaN gradient in mw_evalGrad");
    });
  }

  int qmi_mw_make_move(qmi_ctx* c, int iat, const double* displ)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (iat < 0 || iat >= c->sys.N || !displ)
        throw QmiError(QMI_ERR_ARG, "bad electron index");
      HIPCHECK(hipMemcpyAsync(c->hostin.p, displ, (size_t)c->nw * 3 * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
      HIPCHECK(launch_make_move(c->sys, c->st, iat, c->nw, c->hostin.p, c->stream));
    });
  }

  int qmi_mw_ratio_grad(qmi_ctx* c, int iat, double* ratios, double* grads)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (iat < 0 || iat >= c->sys.N)
        throw QmiError(QMI_ERR_ARG, "bad electron index");
      MoveArgs a = c->move_args(iat);
      a.out      = c->out.p;
      a.prepare  = (c->sys.k == 1 || c->invrow_id != iat) ? 1 : 0;
      {
        qmi_ctx::Scoped t(c, "api_ratio_grad");
        HIPCHECK(c->move_kernel(MK_API_RATIO, a));
      }
      const int vw = c->vw();
      std::vector<double> h((size_t)c->nw * 4 * vw);
      HIPCHECK(hipMemcpyAsync(h.data(), c->out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      if (c->sys.k > 1)
        c->invrow_id = iat;
      for (int w = 0; w < c->nw; w++)
      {
        for (int e = 0; e < vw; e++)
        {
          if (ratios)
            ratios[w * vw + e] = h[w * vw + e];
          if (grads)
            for (int q = 0; q < 3; q++)
              grads[(w * 3 + q) * vw + e] = h[(c->nw + w * 3 + q) * vw + e];
          if (std::isnan(h[w * vw + e]))
            throw QmiError(QMI_ERR_NUMERIC, "NaN ratio in mw_calcRatioGrad");
        }
      }
    });
  }

  int qmi_mw_component_ratio_grad(qmi_ctx* c, int iat, double* ratios, double* grad_new)
  {
    if (c && !grad_new)
      return guarded(c, [] { throw QmiError(QMI_ERR_ARG, "grad_new must not be null"); });
    std::vector<double> g(c ? (size_t)c->nw * 3 * c->vw() : 0);
    const int rc = qmi_mw_ratio_grad(c, iat, ratios, g.data());
    if (rc == QMI_OK)
      for (size_t i = 0; i < g.size(); i++)
        grad_new[i] += g[i];
    return rc;
  }

  int qmi_mw_accept_reject(qmi_ctx* c, int iat, const int* accepted, int safe_to_delay)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (iat < 0 || iat >= c->sys.N || !accepted)
        throw QmiError(QMI_ERR_ARG, "bad electron index");
      HIPCHECK(hipMemcpyAsync(c->acc.p, accepted, (size_t)c->nw * sizeof(int), hipMemcpyHostToDevice, c->stream));
      MoveArgs a  = c->move_args(iat);
      a.host_acc = c->acc.p;
      {
        qmi_ctx::Scoped t(c, "api_accept");
        HIPCHECK(c->move_kernel(MK_API_ACCEPT, a));
      }
      c->invrow_id = -1;
      const int g  = c->group(iat);
      if (c->sys.k > 1)
      {
        c->dcount[g]++;
        if (c->dcount[g] == c->sys.k || !safe_to_delay)
          c->flush(g);
      }
    });
  }

  int qmi_mw_complete_updates(qmi_ctx* c)
  {
    return guarded(c, [&] {
      c->require_walkers();
      c->flush(0);
      c->flush(1);
    });
  }

  int qmi_mw_evaluate_gl(qmi_ctx* c, int fromscratch, double* logpsi, double* G, double* L, double* ekin)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (fromscratch)
        c->recompute_dets();
      c->eval_gl(logpsi, G, L, ekin);
    });
  }

  int qmi_mw_spo_vgl(qmi_ctx* c, const double* pos, int n, double* outv)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (!c->sys.has_det || n <= 0 || n > c->sys.npad / c->vw())
        throw QmiError(QMI_ERR_ARG, "no orbital table or bad orbital count");
      DevBuf<double> dpos, dout;
      dpos.alloc((size_t)c->nw * 3);
      dout.alloc((size_t)c->nw * 5 * n * c->vw());
      HIPCHECK(hipMemcpyAsync(dpos.p, pos, dpos.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      {
        qmi_ctx::Scoped t(c, "api_spo_vgl");
        HIPCHECK(c->cplx ? launch_api_spo_vgl_c(c->sys, c->coefs.p, c->kpts.p, c->nw, n, dpos.p, dout.p, c->stream)
                         : launch_api_spo_vgl(c->sys, c->coefs.p, c->nw, n, dpos.p, dout.p, c->stream));
      }
      HIPCHECK(hipMemcpyAsync(outv, dout.p, dout.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_mw_spo_vgl_det_ratio_grads(qmi_ctx* c, const double* pos, int n, const double* inv_rows, double* phi_vgl,
                                     double* ratios, double* grads)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (!c->sys.has_det || n <= 0 || n > c->sys.npad / c->vw() || !pos || !inv_rows || !ratios || !grads)
        throw QmiError(QMI_ERR_ARG, "no orbital table, bad orbital count or null array");
      const int vw = c->vw();
      DevBuf<double> dpos, dinv, dphi, drg;
      dpos.alloc((size_t)c->nw * 3);
      dinv.alloc((size_t)c->nw * n * vw);
      if (phi_vgl)
        dphi.alloc((size_t)5 * c->nw * n * vw);
      drg.alloc((size_t)c->nw * 4 * vw);
      HIPCHECK(hipMemcpyAsync(dpos.p, pos, dpos.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(hipMemcpyAsync(dinv.p, inv_rows, dinv.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      {
        qmi_ctx::Scoped t(c, "api_vgl_ratio_grads");
        HIPCHECK(c->cplx ? launch_api_vgl_ratio_grads_c(c->sys, c->coefs.p, c->kpts.p, c->nw, n, dpos.p, dinv.p,
                                                        dphi.p, drg.p, c->stream)
                         : launch_api_vgl_ratio_grads(c->sys, c->coefs.p, c->nw, n, dpos.p, dinv.p, dphi.p, drg.p,
                                                      c->stream));
      }
      std::vector<double> h(drg.n);
      HIPCHECK(hipMemcpyAsync(h.data(), drg.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      if (phi_vgl)
        HIPCHECK(hipMemcpyAsync(phi_vgl, dphi.p, dphi.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      for (int w = 0; w < c->nw; w++)
        for (int e = 0; e < vw; e++)
        {
          ratios[w * vw + e] = h[(w * 4) * vw + e];
          for (int q = 0; q < 3; q++)
            grads[(w * 3 + q) * vw + e] = h[(w * 4 + 1 + q) * vw + e];
        }
    });
  }

  int qmi_mw_det_ratios(qmi_ctx* c, int nvp, const double* pos, int n, const double* inv_row, double* ratios)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (!c->sys.has_det || n <= 0 || n > c->sys.npad / c->vw() || nvp <= 0)
        throw QmiError(QMI_ERR_ARG, "no orbital table or bad sizes");
      DevBuf<double> dpos, dinv, dout;
      dpos.alloc((size_t)c->nw * nvp * 3);
      dinv.alloc((size_t)c->nw * n * c->vw());
      dout.alloc((size_t)c->nw * nvp * c->vw());
      HIPCHECK(hipMemcpyAsync(dpos.p, pos, dpos.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(hipMemcpyAsync(dinv.p, inv_row, dinv.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(c->cplx ? launch_api_det_ratios_c(c->sys, c->coefs.p, c->kpts.p, c->nw, nvp, n, dpos.p, dinv.p, dout.p,
                                                 c->stream)
                       : launch_api_det_ratios(c->sys, c->coefs.p, c->nw, nvp, n, dpos.p, dinv.p, dout.p, c->stream));
      HIPCHECK(hipMemcpyAsync(ratios, dout.p, dout.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_mw_invert_transpose(qmi_ctx* c, int complex_values, int n, int nmat, const double* a, double* inv_a,
                              double* log_values)
  {
    return guarded(c, [&] {
      if (n <= 0 || n > MAX_N || nmat <= 0 || !a || !inv_a)
        throw QmiError(QMI_ERR_ARG, "matrix size must be in [1, 512], at least one matrix, non-null arrays");
      const int vw = complex_values ? 2 : 1;
      const size_t per = (size_t)n * n * vw;
      // the factorisation works on A(i, c) = a[i][c] stored column-major
      std::vector<double> t(per * nmat);
      for (int m = 0; m < nmat; m++)
        for (int i = 0; i < n; i++)
          for (int j = 0; j < n; j++)
            for (int q = 0; q < vw; q++)
              t[m * per + ((size_t)j * n + i) * vw + q] = a[m * per + ((size_t)i * n + j) * vw + q];
      DevBuf<double> da, dw, dout, dlog;
      DevBuf<int> dpiv, dperm, dinfo;
      da.alloc(t.size());
      dw.alloc(t.size());
      dout.alloc(t.size());
      dlog.alloc((size_t)nmat * 2);
      dpiv.alloc((size_t)nmat * n);
      dperm.alloc((size_t)nmat * n);
      dinfo.alloc(nmat);
      HIPCHECK(hipMemcpyAsync(da.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(launch_lu_factor(complex_values != 0, n, n, nmat, da.p, dpiv.p, dperm.p, dinfo.p, c->stream));
      HIPCHECK(launch_lu_logdet(complex_values != 0, n, n, nmat, da.p, dpiv.p, dlog.p, 2, c->stream));
      HIPCHECK(launch_lu_inverse(complex_values ? LU_OUT_C128 : LU_OUT_F64, n, n, nmat, da.p, dperm.p, dw.p, dout.p,
                                 c->stream));
      std::vector<int> info(nmat);
      HIPCHECK(hipMemcpyAsync(info.data(), dinfo.p, nmat * sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHECK(hipMemcpyAsync(inv_a, dout.p, dout.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      if (log_values)
        HIPCHECK(hipMemcpyAsync(log_values, dlog.p, dlog.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      for (int m = 0; m < nmat; m++)
        if (info[m] != 0)
          throw QmiError(QMI_ERR_NUMERIC, "singular matrix " + std::to_string(m) + " (zero pivot in column " +
                                              std::to_string(info[m] - 1) + ")");
    });
  }

  // The same factorisation, inverse and log det as qmi_mw_invert_transpose (identical kernels and
  // arithmetic), on device arrays, with every intermediate in the library's static device scratch:
  // no allocation, no synchronisation.
  int qmi_mw_invert_transpose_noalloc(qmi_ctx* c, int complex_values, int n, int nmat, const double* a,
                                      double* inv_a, double* log_values, int* info)
  {
    return guarded(c, [&] {
      if (n <= 0 || n > MAX_N || nmat <= 0 || !a || !inv_a || !info)
        throw QmiError(QMI_ERR_ARG, "matrix size must be in [1, 512], at least one matrix, non-null arrays");
      const int vw      = complex_values ? 2 : 1;
      const size_t per  = (size_t)n * n * vw;                       // doubles per matrix
      const size_t need = 2 * per * nmat * sizeof(double) + 2 * (size_t)n * nmat * sizeof(int);
      if (need > (size_t)QMI_NOALLOC_SCRATCH_BYTES)
        throw QmiError(QMI_ERR_ARG, "the call needs " + std::to_string(need) + " bytes of scratch, more than " +
                                        std::to_string((size_t)QMI_NOALLOC_SCRATCH_BYTES) + " (fewer matrices per call)");
      void* base = nullptr;
      HIPCHECK(lu_noalloc_scratch(c->device, &base));
      double* lu   = static_cast<double*>(base);        // [nmat][n][n] factors, column-major
      double* work = lu + per * nmat;                    // [nmat][n][n] solve work array
      int* piv     = reinterpret_cast<int*>(work + per * nmat);
      int* perm    = piv + (size_t)n * nmat;
      HIPCHECK(launch_lu_transpose_in(n, nmat, vw, a, lu, c->stream));
      HIPCHECK(launch_lu_factor(complex_values != 0, n, n, nmat, lu, piv, perm, info, c->stream));
      if (log_values)
        HIPCHECK(launch_lu_logdet(complex_values != 0, n, n, nmat, lu, piv, log_values, 2, c->stream));
      HIPCHECK(launch_lu_inverse(complex_values ? LU_OUT_C128 : LU_OUT_F64, n, n, nmat, lu, perm, work, inv_a,
                                 c->stream));
    });
  }

  int qmi_mw_get_inverse(qmi_ctx* c, int g, double* ainv)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (!c->sys.has_det || g < 0 || g > 1 || !ainv)
        throw QmiError(QMI_ERR_ARG, "no determinants or bad spin group");
      if (c->dcount[g] != 0)
        throw QmiError(QMI_ERR_ARG, "delayed updates pending: call qmi_mw_complete_updates first");
      const int n = c->sys.n[g], ld = c->sys.ld;
      const size_t es = c->det_esize();
      std::vector<unsigned char> h((size_t)c->nw * n * ld * es);
      HIPCHECK(hipMemcpyAsync(h.data(), c->Ainv[g].p, h.size(), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      const int vw = c->vw();
      for (int w = 0; w < c->nw; w++)
        for (int i = 0; i < n; i++)
          for (int j = 0; j < n * vw; j++)
          {
            const size_t src = ((size_t)w * n + i) * ld * vw + j;
            double v;
            if (es == 4)
              v = reinterpret_cast<const float*>(h.data())[src];
            else
              v = reinterpret_cast<const double*>(h.data())[src];
            ainv[((size_t)w * n + i) * n * vw + j] = v;
          }
    });
  }

  int qmi_mw_evaluate_ratios(qmi_ctx* c, int iat, int nvp, const double* pos, double* ratios)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (iat < 0 || iat >= c->sys.N || nvp <= 0 || !pos || !ratios)
        throw QmiError(QMI_ERR_ARG, "bad electron index or virtual positions");
      if (c->dcount[c->group(iat)] != 0)
        throw QmiError(QMI_ERR_ARG, "delayed updates pending: call qmi_mw_complete_updates first");
      DevBuf<double> dpos, dout;
      dpos.alloc((size_t)c->nw * nvp * 3);
      dout.alloc((size_t)c->nw * nvp * c->vw());
      HIPCHECK(hipMemcpyAsync(dpos.p, pos, dpos.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      {
        qmi_ctx::Scoped t(c, "api_vp_ratios");
        HIPCHECK(c->cplx ? launch_api_vp_ratios_c(c->sys, c->st, c->coefs.p, c->nw, iat, nvp, dpos.p, dout.p, c->stream)
                         : launch_api_vp_ratios(c->sys, c->st, c->coefs.p, c->nw, iat, nvp, dpos.p, dout.p, c->stream));
      }
      HIPCHECK(hipMemcpyAsync(ratios, dout.p, dout.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      for (size_t i = 0; i < dout.n; i++)
        if (std::isnan(ratios[i]))
          throw QmiError(QMI_ERR_NUMERIC, "NaN ratio in mw_evaluateRatios");
    });
  }

  int qmi_coulomb_setup(qmi_ctx* c, double ion_charge, double dim_cutoff)
  {
    return guarded(c, [&] { c->coulomb_setup(ion_charge, dim_cutoff); });
  }

  int qmi_mw_structure_factor(qmi_ctx* c, int nk, const double* kvecs, double* rhok)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (nk <= 0 || !kvecs || !rhok)
        throw QmiError(QMI_ERR_ARG, "need nk > 0 wave vectors and an output array");
      if (c->sys.N > COUL_MAX_N || c->nw > 65535)
        throw QmiError(QMI_ERR_UNSUPPORTED, "more than 1024 electrons or 65535 walkers");
      DevBuf<double> dk, dout;
      dk.alloc((size_t)nk * 3);
      dout.alloc((size_t)c->nw * 2 * nk * 2);
      HIPCHECK(hipMemcpyAsync(dk.p, kvecs, dk.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      {
        qmi_ctx::Scoped t(c, "rhok");
        HIPCHECK(launch_rhok(c->sys, c->st, c->nw, nk, dk.p, dout.p, c->stream));
      }
      HIPCHECK(hipMemcpyAsync(rhok, dout.p, dout.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_system_set_lattice(qmi_ctx* c, const double* R)
  {
    return guarded(c, [&] { c->set_lattice(R); });
  }

  int qmi_coulomb_setup_handler(qmi_ctx* c, double ion_charge, double dim_cutoff, int lr_handler)
  {
    return guarded(c, [&] { c->coulomb_setup(ion_charge, dim_cutoff, lr_handler); });
  }

  int qmi_lr_breakup(const double* L, double dim_cutoff, double* coefs, int ncoefs, double* info)
  {
    if (!L || !coefs || !info || ncoefs < 45)
      return QMI_ERR_ARG;
    try
    {
      const qmi::LrBreakup b = qmi::lr_breakup(L, dim_cutoff);
      std::copy(b.t.begin(), b.t.end(), coefs);
      info[0] = b.rc;
      info[1] = b.vlr_r0();
      info[2] = b.vsr_k0();
      info[3] = b.chi2;
      return QMI_OK;
    }
    catch (const std::exception&)
    {
      return QMI_ERR_ARG;
    }
  }

  int qmi_nlpp_setup(qmi_ctx* c, int nchannel, const int* l, int ngrid, double dr, const double* v,
                     const double* v2, double rcut, int nknot, const double* knots, const double* weights)
  {
    return guarded(c, [&] { c->nlpp_setup(nchannel, l, ngrid, dr, v, v2, rcut, nknot, knots, weights); });
  }

  int qmi_mw_evaluate_nlpp(qmi_ctx* c, const double* rot, double* v)
  {
    return guarded(c, [&] { c->nlpp_evaluate(rot, v); });
  }

  int qmi_dmc_init(qmi_ctx* c, const qmi_dmc_params* p)
  {
    return guarded(c, [&] {
      if (!p)
        throw QmiError(QMI_ERR_ARG, "null parameters");
      c->dmc_init(*p);
    });
  }

  int qmi_dmc_run(qmi_ctx* c, int steps, double* stats)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (!c->dmc.on)
        throw QmiError(QMI_ERR_ARG, "call qmi_dmc_init first");
      if (steps < 0)
        throw QmiError(QMI_ERR_ARG, "negative step count");
      for (int i = 0; i < steps; i++)
        c->dmc_step(stats ? stats + (size_t)i * QMI_DMC_STATS : nullptr);
    });
  }

  int qmi_num_walkers(qmi_ctx* c) { return c ? c->nw : 0; }

  int qmi_checkpoint_save(qmi_ctx* c, void* buf, size_t capacity, size_t* bytes)
  {
    return guarded(c, [&] {
      if (!bytes)
        throw QmiError(QMI_ERR_ARG, "null size output");
      const std::vector<char> v = c->checkpoint_save();
      *bytes                    = v.size();
      if (buf)
      {
        if (capacity < v.size())
          throw QmiError(QMI_ERR_ARG, "checkpoint needs " + std::to_string(v.size()) + " bytes, buffer holds " +
                                          std::to_string(capacity));
        std::memcpy(buf, v.data(), v.size());
      }
    });
  }

  int qmi_checkpoint_load(qmi_ctx* c, const void* buf, size_t bytes)
  {
    return guarded(c, [&] { c->checkpoint_load(buf, bytes); });
  }

  int qmi_dmc_walker_data(qmi_ctx* c, double* weights, double* energies)
  {
    return guarded(c, [&] {
      if (!c->dmc.on)
        throw QmiError(QMI_ERR_ARG, "call qmi_dmc_init first");
      for (int w = 0; w < c->nw; w++)
      {
        if (weights)
          weights[w] = c->dmc.weight[w];
        if (energies)
          energies[w] = c->dmc.eloc[w];
      }
    });
  }

  int qmi_coulomb_nk(qmi_ctx* c) { return c && c->coul_ready ? c->coul.nk : 0; }

  int qmi_mw_evaluate_potential(qmi_ctx* c, double* v)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (!c->coul_ready)
        throw QmiError(QMI_ERR_ARG, "call qmi_coulomb_setup first");
      if (!v)
        throw QmiError(QMI_ERR_ARG, "null output");
      c->coul_out.alloc((size_t)c->nw * 3);
      {
        qmi_ctx::Scoped t(c, "coulomb");
        HIPCHECK(launch_coulomb(c->sys, c->st, c->coul, c->nw, c->coul_out.p, c->stream));
      }
      HIPCHECK(hipMemcpyAsync(v, c->coul_out.p, (size_t)c->nw * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_mw_distance_rows(qmi_ctx* c, const double* pos, double* ee_r, double* ee_dr, double* ei_r, double* ei_dr)
  {
    return guarded(c, [&] {
      c->require_walkers();
      const int nw = c->nw, N = c->sys.N, ni = c->sys.nion;
      DevBuf<double> dpos, a, b, cc, dd;
      dpos.alloc((size_t)nw * 3);
      a.alloc((size_t)nw * N);
      b.alloc((size_t)nw * N * 3);
      cc.alloc((size_t)nw * std::max(ni, 1));
      dd.alloc((size_t)nw * std::max(ni, 1) * 3);
      HIPCHECK(hipMemcpyAsync(dpos.p, pos, dpos.n * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(launch_api_distances(c->sys, c->st, nw, dpos.p, a.p, b.p, cc.p, dd.p, c->stream));
      if (ee_r)
        HIPCHECK(hipMemcpyAsync(ee_r, a.p, (size_t)nw * N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      if (ee_dr)
        HIPCHECK(hipMemcpyAsync(ee_dr, b.p, (size_t)nw * N * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      if (ei_r && ni)
        HIPCHECK(hipMemcpyAsync(ei_r, cc.p, (size_t)nw * ni * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      if (ei_dr && ni)
        HIPCHECK(hipMemcpyAsync(ei_dr, dd.p, (size_t)nw * ni * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_dt_create(qmi_ctx* c, int kind, qmi_dtable** out)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (!out || (kind != DT_AA && kind != DT_AB))
        throw QmiError(QMI_ERR_ARG, "kind must be 0 (electron pairs) or 1 (electron-ion), out non-null");
      if (kind == DT_AB && c->sys.nion <= 0)
        throw QmiError(QMI_ERR_ARG, "no ions for an electron-ion table");
      if (c->skew_cell)
        throw QmiError(QMI_ERR_UNSUPPORTED, "distance tables need an orthorhombic cell");
      auto t    = std::make_unique<qmi_dtable>();
      t->c      = c;
      t->device = c->device;
      t->kind   = kind;
      t->nw   = c->nw;
      t->N    = c->sys.N;
      t->ns   = kind == DT_AA ? c->sys.N : c->sys.nion;
      const size_t nw = t->nw, N = t->N, ns = t->ns;
      t->r.alloc(nw * N * ns);
      t->dr.alloc(nw * N * 3 * ns);
      t->temp_r.alloc(nw * ns);
      t->temp_dr.alloc(nw * 3 * ns);
      t->old_r.alloc(kind == DT_AA ? nw * ns : 1);
      t->old_dr.alloc(kind == DT_AA ? nw * 3 * ns : 1);
      t->rnew.alloc(nw * 3);
      t->acc.alloc(nw);
      HIPCHECK(hipMemsetAsync(t->r.p, 0, t->r.n * sizeof(double), c->stream));
      HIPCHECK(hipMemsetAsync(t->dr.p, 0, t->dr.n * sizeof(double), c->stream));
      c->sync();
      c->dtables.push_back(t.get());
      *out = t.release();
    });
  }

  int qmi_dt_destroy(qmi_dtable* t)
  {
    if (!t)
      return QMI_ERR_ARG;
    (void)hipSetDevice(t->device);
    if (t->c)
    {
      (void)hipStreamSynchronize(t->c->stream);
      auto& v = t->c->dtables;
      v.erase(std::remove(v.begin(), v.end(), t), v.end());
    }
    delete t; // device buffers are freed with the device current
    return QMI_OK;
  }

  int qmi_dt_mw_evaluate(qmi_dtable* t)
  {
    if (!t || !t->c)
      return QMI_ERR_ARG; // null, or its context was destroyed
    return guarded(t->c, [&] {
      t->check();
      HIPCHECK(launch_dt_evaluate(t->c->sys, t->c->st, t->arrays(), t->nw, t->c->stream));
      t->c->sync();
    });
  }

  int qmi_dt_mw_move(qmi_dtable* t, int iat, const double* rnew, int prepare_old)
  {
    if (!t || !t->c)
      return QMI_ERR_ARG; // null, or its context was destroyed
    return guarded(t->c, [&] {
      t->check();
      if (iat < 0 || iat >= t->N || !rnew)
        throw QmiError(QMI_ERR_ARG, "bad particle index or null positions");
      HIPCHECK(hipMemcpyAsync(t->rnew.p, rnew, t->rnew.n * sizeof(double), hipMemcpyHostToDevice, t->c->stream));
      HIPCHECK(launch_dt_move(t->c->sys, t->c->st, t->arrays(), t->nw, iat, t->rnew.p, prepare_old ? 1 : 0,
                              t->c->stream));
      t->c->sync();
    });
  }

  int qmi_dt_mw_update(qmi_dtable* t, int iat, const int* accepted, int partial)
  {
    if (!t || !t->c)
      return QMI_ERR_ARG; // null, or its context was destroyed
    return guarded(t->c, [&] {
      t->check();
      if (iat < 0 || iat >= t->N || !accepted)
        throw QmiError(QMI_ERR_ARG, "bad particle index or null mask");
      HIPCHECK(hipMemcpyAsync(t->acc.p, accepted, t->acc.n * sizeof(int), hipMemcpyHostToDevice, t->c->stream));
      HIPCHECK(launch_dt_update(t->c->sys, t->arrays(), t->nw, iat, t->acc.p, partial ? 1 : 0, t->c->stream));
      t->c->sync();
    });
  }

  int qmi_dt_get(qmi_dtable* t, double* r, double* dr, double* temp_r, double* temp_dr, double* old_r, double* old_dr)
  {
    if (!t || !t->c)
      return QMI_ERR_ARG; // null, or its context was destroyed
    return guarded(t->c, [&] {
      t->check();
      auto get = [&](double* dst, const DevBuf<double>& src) {
        if (dst)
          HIPCHECK(hipMemcpyAsync(dst, src.p, src.n * sizeof(double), hipMemcpyDeviceToHost, t->c->stream));
      };
      get(r, t->r);
      get(dr, t->dr);
      get(temp_r, t->temp_r);
      get(temp_dr, t->temp_dr);
      if (t->kind == DT_AA)
      {
        get(old_r, t->old_r);
        get(old_dr, t->old_dr);
      }
      t->c->sync();
    });
  }

  int qmi_vmc_run(qmi_ctx* c, int steps, int sub_steps, double tau, int recompute_period, uint32_t seed,
                  double* energies, double* trace)
  {
    return guarded(c, [&] { c->vmc_run(steps, sub_steps, tau, recompute_period, seed, energies, trace); });
  }

  int qmi_vmc_run_ex(qmi_ctx* c, int steps, int sub_steps, double tau, int recompute_period, uint32_t seed,
                     double* energies, double* trace, double* ltrace)
  {
    return guarded(c, [&] { c->vmc_run(steps, sub_steps, tau, recompute_period, seed, energies, trace, ltrace); });
  }

  int qmi_mlp_params(qmi_ctx* c, int set, float* params)
  {
    return guarded(c, [&] {
      if (!params)
        throw QmiError(QMI_ERR_ARG, "null parameter array");
      if (set)
        c->mlp_set(params);
      else
      {
        if (!c->mlp_ready)
          c->mlp_init(12345u);
        HIPCHECK(hipMemcpyAsync(params, c->mlp_master.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost,
                                c->stream));
        c->sync();
      }
    });
  }

  int qmi_mlp_forward(qmi_ctx* c, int B, const float* X, float* out)
  {
    return guarded(c, [&] {
      if (B <= 0 || !X || !out)
        throw QmiError(QMI_ERR_ARG, "bad batch");
      if (!c->mlp_ready)
        c->mlp_init(12345u);
      DevBuf<float> dX, dO;
      dX.alloc((size_t)B * MLP_IN);
      dO.alloc((size_t)B * MLP_OUT);
      HIPCHECK(hipMemcpyAsync(dX.p, X, dX.n * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(launch_mlp_forward(c->mlp(), B, dX.p, dO.p, nullptr, B, 0, c->stream));
      HIPCHECK(hipMemcpyAsync(out, dO.p, dO.n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_mlp_gradient(qmi_ctx* c, int B, const float* X, const float* dout, float* grad)
  {
    return guarded(c, [&] {
      if (B <= 0 || !X || !dout || !grad)
        throw QmiError(QMI_ERR_ARG, "bad batch");
      if (!c->mlp_ready)
        c->mlp_init(12345u);
      DevBuf<float> dX, dO, dD, z, za, zb, gr;
      dX.alloc((size_t)B * MLP_IN);
      dO.alloc((size_t)B * MLP_OUT);
      dD.alloc((size_t)B * MLP_OUT);
      z.alloc((size_t)3 * B * MLP_HID);
      za.alloc((size_t)B * MLP_HID);
      zb.alloc((size_t)B * MLP_HID);
      gr.alloc(MLP_NPARAM);
      HIPCHECK(hipMemcpyAsync(dX.p, X, dX.n * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(hipMemcpyAsync(dD.p, dout, dD.n * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHECK(launch_mlp_forward(c->mlp(), B, dX.p, dO.p, z.p, B, 0, c->stream));
      HIPCHECK(launch_mlp_backward(c->mlp(), B, dX.p, z.p, dD.p, gr.p, za.p, zb.p, c->stream));
      HIPCHECK(hipMemcpyAsync(grad, gr.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_proposal_config(qmi_ctx* c, int learned, int train, double lr, uint32_t init_seed)
  {
    return guarded(c, [&] {
      if (init_seed)
        c->mlp_init(init_seed);
      c->learned = learned != 0;
      c->train   = c->learned && train != 0;
      c->lr      = (float)lr;
    });
  }

  int qmi_proposal_stats(qmi_ctx* c, double* out4)
  {
    return guarded(c, [&] {
      c->require_walkers();
      std::vector<double> h((size_t)c->nw * 2);
      HIPCHECK(hipMemcpyAsync(h.data(), c->lfsum.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      double f = 0, a = 0;
      for (int w = 0; w < c->nw; w++)
      {
        f += h[w * 2];
        a += h[w * 2 + 1];
      }
      const double moves = c->learn_moves * c->nw;
      out4[0] = moves > 0 ? f / moves : 0.0;
      out4[1] = moves > 0 ? a / moves : 0.0;
      out4[2] = moves;
      out4[3] = c->adam_t;
      HIPCHECK(hipMemsetAsync(c->lfsum.p, 0, c->lfsum.n * sizeof(double), c->stream));
      c->sync();
      c->learn_moves = 0;
    });
  }

  int qmi_mw_proposal_features(qmi_ctx* c, int iat, const double* pos, const double* drift, double tau, float* out)
  {
    return guarded(c, [&] {
      c->require_walkers();
      if (iat < 0 || iat >= c->sys.N || c->sys.nion > 1280)
        throw QmiError(QMI_ERR_ARG, "bad electron index or too many ions");
      const int nw = c->nw;
      std::vector<double> h((size_t)nw * LS_SIZE, 0.0);
      for (int w = 0; w < nw; w++)
        for (int q = 0; q < 3; q++)
        {
          h[(size_t)w * LS_SIZE + LS_R + q] = pos[w * 3 + q];
          h[(size_t)w * LS_SIZE + LS_V + q] = drift[w * 3 + q];
        }
      c->learn_alloc(2 * nw);
      HIPCHECK(hipMemcpyAsync(c->lstate.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
      MoveArgs a{};
      a.iat0       = iat;
      a.nw         = nw;
      a.tau        = tau;
      LearnArgs la = c->learn_args(false);
      HIPCHECK(launch_learn(LK_FEATURES, c->sys, c->st, c->coefs.p, a, la, c->stream));
      HIPCHECK(hipMemcpyAsync(out, c->lX.p, (size_t)nw * MLP_IN * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      c->sync();
    });
  }

  int qmi_vmc_stats(qmi_ctx* c, double* stats2)
  {
    return guarded(c, [&] {
      c->require_walkers();
      std::vector<double> h((size_t)c->nw * 2);
      HIPCHECK(hipMemcpyAsync(h.data(), c->stats.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      stats2[0] = stats2[1] = 0;
      for (int w = 0; w < c->nw; w++)
      {
        stats2[0] += h[w * 2];
        stats2[1] += h[w * 2 + 1];
      }
    });
  }

  int qmi_timing_enable(qmi_ctx* c, int on)
  {
    return guarded(c, [&] {
      c->sync();
      c->timing.on = on != 0;
    });
  }

  int qmi_timing_reset(qmi_ctx* c)
  {
    return guarded(c, [&] {
      c->sync();
      std::fill(c->timing.ms.begin(), c->timing.ms.end(), 0.0);
      std::fill(c->timing.calls.begin(), c->timing.calls.end(), 0.0);
    });
  }

  int qmi_timing_get(qmi_ctx* c, char* buf, int buflen, double* ms, double* calls, int maxk, int* nk)
  {
    return guarded(c, [&] {
      c->sync();
      std::string names;
      const int n = std::min<int>(maxk, (int)c->timing.names.size());
      for (int i = 0; i < n; i++)
      {
        names += c->timing.names[i];
        names += '\n';
        ms[i]    = c->timing.ms[i];
        calls[i] = c->timing.calls[i];
      }
      if (buf && buflen > 0)
      {
        std::strncpy(buf, names.c_str(), buflen - 1);
        buf[buflen - 1] = 0;
      }
      *nk = n;
    });
  }

  // ---- tuning knobs, graph replay, batches of contexts
  int qmi_tune_set(qmi_ctx* c, int delay_rank, int move_kernel, int flush_variant)
  {
    return guarded(c, [&] {
      if (move_kernel < 0 || move_kernel > 1 || flush_variant < 0 || flush_variant > 1)
        throw QmiError(QMI_ERR_ARG, "move_kernel and flush_variant are 0 or 1");
      if (delay_rank > 0)
        c->set_delay_rank(delay_rank);
      if (c->move_one != 1 - move_kernel || c->flush_variant != flush_variant)
      {
        c->move_one      = 1 - move_kernel;
        c->flush_variant = flush_variant;
        c->invalidate_graph();
      }
    });
  }

  int qmi_tune_partitions(qmi_ctx* c, int partitions, int* current)
  {
    return guarded(c, [&] {
      if (partitions > 8)
        throw QmiError(QMI_ERR_ARG, "at most 8 walker partitions");
      if (partitions > 0 && partitions != c->npart)
      {
        c->npart = partitions;
        c->invalidate_graph();
      }
      if (current)
        *current = c->npart;
    });
  }

  int qmi_tune_get(qmi_ctx* c, int* out3)
  {
    return guarded(c, [&] {
      if (!out3)
        throw QmiError(QMI_ERR_ARG, "null output");
      out3[0] = c->sys.k;
      out3[1] = 1 - c->move_one;
      out3[2] = c->flush_variant;
    });
  }

  int qmi_graph_enable(qmi_ctx* c, int on)
  {
    return guarded(c, [&] {
      c->graph_on = on != 0;
      if (!c->graph_on)
        c->invalidate_graph();
    });
  }

  int qmi_batch_create(qmi_batch** out, qmi_ctx* const* ctxs, int nctx)
  {
    if (!out || !ctxs || nctx <= 0)
      return QMI_ERR_ARG;
    *out = nullptr;
    for (int i = 0; i < nctx; i++)
    {
      if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device)
        return QMI_ERR_ARG;
      for (int j = 0; j < i; j++)
        if (ctxs[j] == ctxs[i])
          return QMI_ERR_ARG;
    }
    auto* b   = new qmi_batch;
    b->ctxs   = std::vector<qmi_ctx*>(ctxs, ctxs + nctx);
    b->device = ctxs[0]->device;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreate(&b->origin) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess)
    {
      delete b;
      return QMI_ERR_DEVICE;
    }
    *out = b;
    return QMI_OK;
  }

  int qmi_batch_destroy(qmi_batch* b)
  {
    if (!b)
      return QMI_ERR_ARG;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->origin);
    delete b;
    return QMI_OK;
  }

  const char* qmi_batch_last_error(const qmi_batch* b) { return b ? b->err.c_str() : "null batch"; }

  static int batch_guarded(qmi_batch* b, const std::function<void()>& f)
  {
    if (!b)
      return QMI_ERR_ARG;
    try
    {
      HIPCHECK(hipSetDevice(b->device));
      f();
      b->err.clear();
      return QMI_OK;
    }
    catch (const QmiError& e)
    {
      b->err = e.what();
      return e.code;
    }
    catch (const std::exception& e)
    {
      b->err = e.what();
      return QMI_ERR_DEVICE;
    }
  }

  int qmi_batch_vmc_run(qmi_batch* b, int steps, int sub_steps, double tau, int recompute_period,
                        const uint32_t* seeds, double* const* energies)
  {
    return batch_guarded(b, [&] { b->run(steps, sub_steps, tau, recompute_period, seeds, energies); });
  }

  int qmi_batch_autotune(qmi_batch* b, double tau, int trials, int* choice, double* ms)
  {
    return batch_guarded(b, [&] {
      if (!(tau > 0))
        throw QmiError(QMI_ERR_ARG, "invalid tau");
      b->autotune(tau, trials, choice, ms);
    });
  }

  int qmi_batch_set_concurrent(qmi_batch* b, int concurrent)
  {
    return batch_guarded(b, [&] {
      b->concurrent = concurrent != 0;
      b->invalidate();
    });
  }

  int qmi_batch_get_concurrent(qmi_batch* b, int* concurrent)
  {
    return batch_guarded(b, [&] {
      if (!concurrent)
        throw QmiError(QMI_ERR_ARG, "null output");
      *concurrent = b->concurrent ? 1 : 0;
    });
  }
}
