// Krylov solvers and preconditioners on the device.
//
// The system matrix is BSR with 2x2 blocks (one block row per triangle,
// unknowns 2*ti+uv).  BiCGStab runs entirely on the device: its scalars live
// in device memory, every kernel checks a device "done" flag and returns
// early once the solve has converged, and the host only polls the flag every
// `check_every` iterations.
//
// Preconditioners
//   PC_NONE / PC_JACOBI / PC_BLOCK_JACOBI : fused into the BiCGStab vector
//       updates (block-local, no extra pass over memory)
//   multilevel aggregation (Amg) : fixed hierarchy per mesh built on the host
//       (greedy aggregation of the triangle adjacency, repeated on the
//       aggregate graph); per Picard iteration only the Galerkin values are
//       recomputed on the device, as atomic-free gather sums of fine blocks
//       into coarse blocks.  One V-cycle: damped block-Jacobi pre/post
//       smoothing, piecewise-constant prolongation, exact solve on the
//       coarsest level (dense inverse formed in LDS once per Picard
//       iteration).  It is a fixed linear operator, so plain BiCGStab applies.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

#include "kernels.h"

namespace ufe {

template <typename T> struct VecT;
template <> struct VecT<double> { using V2 = double2; using V4 = double4; };
template <> struct VecT<float> { using V2 = float2; using V4 = float4; };

template <typename T>
struct BsrT {
  int nb = 0, nnzb = 0;
  const int *ptr = nullptr, *col = nullptr;
  const typename VecT<T>::V4 *val = nullptr;
  const typename VecT<T>::V4 *dinv = nullptr;  // 2x2 inverse of each diagonal block
};

// Device buffer with RAII (hipMalloc/hipFree).
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  explicit DBuf(size_t b) { alloc(b); }
  void alloc(size_t b);
  void free();
  ~DBuf() { free(); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  template <typename U> U *as() const { return static_cast<U *>(p); }
};

// Aggregation hierarchy (pattern only; values live in AmgValues<T>).
struct AmgLevelPattern {
  int nb = 0, nnzb = 0;             // this level's matrix (level 0 = the system)
  // levels >= 1: BSR pattern, row of each entry, and the Galerkin gather
  // (entry c sums the previous level's entries glist[gptr[c] .. gptr[c+1]))
  DBuf ptr, col, crow, gptr, glist;
  // levels < last: aggregate of each block row and members of each aggregate
  DBuf agg, mem_ptr, mem_ind;
};

struct AmgHierarchy {
  std::vector<std::unique_ptr<AmgLevelPattern>> levels;  // levels[0] has nb/nnzb of the system
  int n_levels() const { return (int)levels.size(); }
};

// Build the hierarchy on the host from the block pattern of the system and
// the triangle adjacency (TriC).  coarse_target: stop when a level has at
// most this many block rows.  passes: greedy aggregation passes per level
// (each pass aggregates the previous pass's aggregates).
std::unique_ptr<AmgHierarchy> build_amg_hierarchy(int nT, const int *TriC, const int *ptr0, const int *col0,
                                                  int coarse_target, int max_levels, int passes = 1);

template <typename T>
struct AmgValues {
  std::vector<std::unique_ptr<DBuf>> val, dinv;  // levels >= 1 (level 0 values are the system's)
  std::vector<std::unique_ptr<DBuf>> r, z, zt;   // per-level vectors (levels >= 1 for r, all for z/zt)
};

struct SolveStats {
  int its = 0;
  int status = 0;       // 1 converged, 2 breakdown, 3 maxit
  double rel = 0.0;     // ||r|| / ||r0|| at exit (recurrence)
  double r0 = 0.0;
};

class Krylov {
 public:
  Krylov(int nb, hipStream_t s);
  ~Krylov();

  // Preconditioner choice for subsequent solves.
  void set_precond(int pc) { pc_ = pc; }
  void set_diag(const int *diag) { A_diag_ = diag; }
  int precond() const { return pc_; }
  void attach_amg(const AmgHierarchy *h) { amg_ = h; }
  // FP32 copy of the finest-level matrix for the FP64 V-cycle (null: use the FP64 values)
  void set_amg_fine32(const float4 *v) { fine32_ = v; }
  // Recompute Galerkin values on all levels from the level-0 matrix.
  void amg_setup64(const BsrT<double> &A0);
  void amg_setup32(const BsrT<float> &A0);
  void set_amg_options(double omega, int coarse_sweeps, double coarse_scale) {
    amg_omega_ = omega; amg_coarse_sweeps_ = coarse_sweeps; amg_coarse_scale_ = coarse_scale;
  }
  // The finest level of the V-cycle has identity diagonal blocks (the
  // block-Jacobi-scaled system): its pre-smoothing is folded into the
  // level's passes.
  void set_unit_diag0(bool on) { unit_diag0_ = on; }
  // Run the coarse levels of the V-cycle with at most max_rows block rows in
  // one workgroup (0: one launch per pass on every level).
  void set_amg_tail(int max_rows) { amg_tail_rows_ = max_rows; }
  // l1 block-Jacobi smoothing on every level (true) or plain block Jacobi;
  // the finest level's pre-smoothing is then not folded into its passes
  void set_amg_l1(bool on) { amg_l1_ = on; }
  // A preconditioner of the caller's, z = M^-1 r on the FP64 vectors, used by
  // gmres64 in place of the built-in ones while set (empty function: off).
  void set_custom_pc(std::function<void(const double *r, double *z)> f) { custom_pc_ = std::move(f); }
  // Launch BiCGStab iteration chunks as replays of captured HIP graphs
  // (same kernels in the same order: identical results).
  void set_graphs(bool on) { use_graphs_ = on; }

  // BiCGStab: x in (initial guess unless zero_guess) / out.  Stops when
  // ||r|| <= max(rtol*||r0||, abstol).
  SolveStats bicgstab64(const BsrT<double> &A, const double *b, double *x, double rtol, double abstol, int maxit,
                        bool zero_guess, int check_every);
  SolveStats bicgstab32(const BsrT<float> &A, const float *b, float *x, double rtol, double abstol, int maxit,
                        bool zero_guess, int check_every);

  // GMRES(restart), right-preconditioned with the current preconditioner:
  // x in (initial guess unless zero_guess) / out.  Stops when ||b - A x||
  // (the Arnoldi estimate inside a cycle, the true residual at each restart)
  // <= max(rtol*||r0||, abstol); iterations count Arnoldi steps.  restart <= 64.
  SolveStats gmres64(const BsrT<double> &A, const double *b, double *x, double rtol, double abstol, int maxit,
                     bool zero_guess, int restart, int check_every);
  SolveStats gmres32(const BsrT<float> &A, const float *b, float *x, double rtol, double abstol, int maxit,
                     bool zero_guess, int restart, int check_every);

  // FP64 residual r = b - A x with ||r||^2 and ||b||^2 (blocking read).
  // vmax > 0: rows whose block of x has magnitude >= vmax are left out of both
  // sums and counted in *n_clamped.
  // Norms are those of the system given (for the left-scaled system
  // D^-1 A, D^-1 b: the block-Jacobi-scaled residual, in velocity units).
  // With dblk and write_unscaled, r receives the unscaled residual D r_s.
  void residual64(const BsrT<double> &A, const double *x, const double *b, double *r, double *rr, double *bb,
                  double vmax = 0.0, int *n_clamped = nullptr, const double4 *dblk = nullptr,
                  int write_unscaled = 0);

  // Average duration (ms, HIP events on this solver's stream) of `reps`
  // back-to-back launches of the FP64 block product y = A x, and the
  // algorithmic HBM bytes of one launch (matrix blocks and column indices,
  // row pointers, x and y once each).
  double time_spmv64(const BsrT<double> &A, const double *x, double *y, int reps, double *bytes);

  // Live timing inside bicgstab64 (HIP events around each launch; counters
  // reset by set_spmv_timing(true)): kernel 0 = the FP64 products with a dot
  // epilogue, 1 = the V-cycle's finest-level prolongation + smoothing.
  void set_spmv_timing(bool on);
  void spmv_timing(int kernel, long long *launches, double *total_ms, double *bytes) const {
    *launches = tn_[kernel]; *total_ms = tms_[kernel]; *bytes = tbytes_[kernel];
  }

  hipStream_t stream() const { return s_; }
  RedSlot slot(int i) const;

  // scratch vectors usable by callers between solves
  double *tmp64(int i);
  float *tmp32(int i);

 private:
  template <typename T> SolveStats bicgstab(const BsrT<T> &A, const T *b, T *x, double rtol, double abstol,
                                            int maxit, bool zero_guess, int check_every);
  template <typename T> std::vector<uintptr_t> graph_key(const BsrT<T> &A, const T *b, const T *x, int nbx);
  template <typename T> hipGraphExec_t iteration_graph(int n, const std::function<void(int)> &launch_iter,
                                                       const BsrT<T> &A, const T *b, T *x, int nbx);
  template <typename T> SolveStats gmres(const BsrT<T> &A, const T *b, T *x, double rtol, double abstol, int maxit,
                                         bool zero_guess, int restart, int check_every);
  template <typename T> void precond_apply(const BsrT<T> &A, const T *r, T *z, hipEvent_t ev0 = nullptr,
                                           hipEvent_t ev1 = nullptr);
  template <typename T> void amg_setup(const BsrT<T> &A0);
  template <typename T> AmgValues<T> &amg_vals();

  int nb_;
  hipStream_t s_;
  std::function<void(const double *, double *)> custom_pc_;
  int pc_ = UFE_PC_BLOCK_JACOBI;
  const AmgHierarchy *amg_ = nullptr;
  const int *A_diag_ = nullptr;
  const float4 *fine32_ = nullptr;
  int amg_tail_rows_ = 300;
  bool amg_l1_ = false;
  bool unit_diag0_ = false;
  double amg_omega_ = 0.85;
  int amg_coarse_sweeps_ = 20;
  double amg_coarse_scale_ = 1.5;
  AmgValues<double> amg64_;
  AmgValues<float> amg32_;
  std::vector<BsrT<double>> lev64_;
  std::vector<BsrT<float>> lev32_;
  DBuf vec64_[10], vec32_[10];
  DBuf coarse_inv_;      // dense inverse of the coarsest level (column-major)
  DBuf scal_;            // KrylovScalars
  DBuf red_part_, red_res_;
  KrylovScalars *h_scal_ = nullptr;  // pinned host mirror
  bool time_spmv_ = false;
  std::vector<hipEvent_t> tev_;
  long long tn_[2] = {0, 0};
  double tms_[2] = {0.0, 0.0}, tbytes_[2] = {0.0, 0.0};
  int verbose_ = 0;
  // iterations of the previous solve per precision ([0] FP64, [1] FP32): the
  // first host poll of the next solve is placed there
  int last_its_[2] = {0, 0};
  int last_gmres_its_[2] = {0, 0};
  // captured iteration graphs of 1, 2, 4 and 8 BiCGStab iterations ([0] FP64, [1] FP32)
  struct GraphCache {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
    void clear();
  };
  GraphCache graphs_[2];
  bool use_graphs_ = false;
  // GMRES: basis (restart + 1 vectors), preconditioned vector, combination,
  // device scalars (Hessenberg matrix, rotations) and reduction partials
  DBuf gm_V_, gm_z_, gm_t_, gm_sc_, gm_part_;
};

}  // namespace ufe
