// replay_schedule.hpp -- host-side bookkeeping of the batched Lanczos solve.
//
// lsqed_batch_solve runs the batch as replays of a captured graph, each
// replay advancing every active model by a fixed number of Lanczos steps.  The
// host decides, model by model, which replays run it:
//   done      converged (or out of steps): inactive from the next enqueued
//             replay on;
//   paused    predicted to converge within the replay already in flight (the
//             ratio of its last two residuals): inactive from the next enqueued
//             replay; if the replay in flight does not confirm convergence, it
//             runs again from the replay after.  A paused model is untouched
//             by every kernel, so pausing changes none of its steps.
// One replay is kept in flight beyond the one being processed, so that the
// host's work on a replay's results overlaps the next replay.  Header-only
// and free of HIP so that the rules are tested on the CPU
// (tests/cpp/test_replay_schedule.cpp).
#pragma once

#include <vector>

namespace lsqed {

struct ReplaySchedule {
  int np = 0;
  std::vector<char> done, paused, cur;      // per model
  std::vector<double> prev_res;             // residual after the previous processed replay
  std::vector<std::vector<char>> act_at;    // activity of every enqueued replay
  int done_r = 0;                           // replays processed
  int ndone = 0;                            // models done
  bool dirty = false;                       // cur changed since the last enqueued replay

  explicit ReplaySchedule(int n) : np(n), done(n, 0), paused(n, 0), cur(n, 1), prev_res(n, 0.0) {}

  bool any_active() const {
    for (char c : cur)
      if (c) return true;
    return false;
  }
  // Enqueue while this holds: one replay beyond the one to process next.  A
  // model resumed after a failed prediction may find none in flight, so the
  // caller loops until there is one (at most two replays are then in flight).
  bool want_replay() const { return (int)act_at.size() <= done_r + 1 && any_active(); }
  // The caller enqueued a replay running the models now marked active.
  void enqueued() {
    act_at.push_back(cur);
    dirty = false;
  }
  bool finished() const { return ndone >= np; }
  // No replay left to process while models remain: the rules above exclude
  // it; the caller reports an error rather than waiting.
  bool stalled() const { return (int)act_at.size() <= done_r; }
  // Was model k run by the replay being processed?
  bool ran(int k) const { return !done[k] && act_at[done_r][k]; }
  // Result of model k from the replay being processed: `stop` when it
  // converged or has no steps left, else its residual `res` against `tol`.
  void result(int k, bool stop, double res, double tol) {
    const bool next_runs_k = done_r + 1 < (int)act_at.size() && act_at[done_r + 1][k];
    if (stop) {
      done[k] = 1;
      ++ndone;
      if (cur[k]) dirty = true;
      cur[k] = 0;
      paused[k] = 0;
    } else if (paused[k]) {
      // the prediction failed: resume
      paused[k] = 0;
      cur[k] = 1;
      dirty = true;
    } else if (prev_res[k] > 0.0 && res < prev_res[k] && res * (res / prev_res[k]) < 0.5 * tol && next_runs_k) {
      // paused only when the replay after this one runs the model, which then
      // settles the prediction
      paused[k] = 1;
      cur[k] = 0;
      dirty = true;
    }
    prev_res[k] = res;
  }
  void processed() { ++done_r; }
};

}  // namespace lsqed
