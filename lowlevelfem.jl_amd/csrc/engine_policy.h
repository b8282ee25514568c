// Decisions of the reliable-update engine at an FP64 update, as plain host
// functions (no HIP): which solution to keep, when a system stops, and when
// an AMG system's rescaled preconditioned residual is still usable.
//
// Reference: solveDisplacement(K, f; support, iterative=true, reltol,
// maxiter) (src/linear.jl:3711-3743) stops CG on ||r|| <= reltol ||r0||; the
// engine decides that on the true FP64 residual b - K x at each update.
#pragma once

#include <cmath>
#include <cstdint>

namespace llfem {

// Status of a system (SolveSys::status).
enum : int { kRunning = 0, kConverged = 1, kMaxIter = -1, kNonFinite = -2 };

// Bookkeeping of one system across updates.
struct UpdateBook {
  int bad = 0;              // consecutive updates whose FP32 correction was not finite
  int64_t rollbacks = 0;    // updates that kept the previous FP64 solution
  int64_t updates = 0;      // FP64 updates done (the initial residual excluded)
};

// An update folds the FP32 correction into a trial FP64 solution and forms
// its true residual norm^2, r2_trial.  A non-finite r2_trial means the
// correction overflowed: the previous solution is kept (a rollback) and the
// FP32 recursion restarts from it.  Returns true if the trial is adopted.
inline bool adopt_trial(double r2_trial, UpdateBook* bk) {
  if (std::isfinite(r2_trial)) {
    bk->bad = 0;
    return true;
  }
  ++bk->bad;
  ++bk->rollbacks;
  return false;
}

// What happens after an update with true residual norm rn (of the kept
// solution): stop as converged, stop at maxiter, stop after three
// non-finite corrections in a row, or keep iterating.
inline int update_status(double rn, double bnorm, double reltol, int64_t iters, int64_t maxiter,
                         const UpdateBook& bk) {
  if (bk.bad >= 3 || !std::isfinite(rn)) return kNonFinite;
  if (rn <= reltol * bnorm || bnorm == 0.0) return kConverged;
  if (iters >= maxiter || bk.updates > 100000) return kMaxIter;
  return kRunning;
}

// AMG systems rescale the previous z = M r with the new residual instead of
// running a V-cycle.  That z is consistent only while r.z > 0 and the true
// residual (scaled to unit norm) is close to the rescaled recursive one:
// relative drift^2 below kDrift^2.  Otherwise z is recomputed by a V-cycle.
inline bool amg_keep_rescaled(double rz, double drift2, double kDrift) {
  return rz > 0.0 && drift2 <= kDrift * kDrift;
}

}  // namespace llfem
