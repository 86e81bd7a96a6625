// The host-side replay bookkeeping of the batched Lanczos solve
// (lsquantumed_amd/csrc/replay_schedule.hpp), driven on the CPU the way
// lsqed_batch_solve drives it, with scripted residual sequences in place of
// the GPU's Lanczos histories: every model finishes, no replay is waited for
// that was never enqueued, at most two replays are in flight, and a model
// runs only in replays enqueued while it was active.
#include <random>

#include "check.hpp"
#include "replay_schedule.hpp"

namespace {

struct Outcome {
  bool stalled = false;
  int replays = 0;
  int max_in_flight = 0;
  std::vector<int> steps;  // replays that ran each model
};

// res[k][i]: residual of model k after the i-th replay that runs it (past the
// end of the list it converges); tol = 1 (converged when below).
Outcome drive(const std::vector<std::vector<double>>& res, int max_steps = 1000) {
  const int np = (int)res.size();
  lsqed::ReplaySchedule s(np);
  Outcome o;
  o.steps.assign(np, 0);
  s.enqueued();
  while (!s.finished()) {
    while (s.want_replay()) s.enqueued();
    o.max_in_flight = std::max(o.max_in_flight, (int)s.act_at.size() - s.done_r);
    if (s.stalled()) {
      o.stalled = true;
      break;
    }
    for (int k = 0; k < np; ++k) {
      if (!s.ran(k)) continue;
      const int i = o.steps[k]++;
      const double r = i < (int)res[k].size() ? res[k][i] : 0.5;
      s.result(k, r < 1.0 || o.steps[k] >= max_steps, r, 1.0);
    }
    s.processed();
  }
  o.replays = (int)s.act_at.size();
  return o;
}

}  // namespace

// Two models: the second converges early; the first is paused, resumed
// (prediction failed), and later paused again.  The resume happens while no
// replay beyond the one being processed is in flight; with one replay
// enqueued per processed replay the pipeline stayed one short from there on,
// and the second pause waited for a replay that was never enqueued.
TEST(ReplaySchedule, PauseAfterFailedPrediction) {
  const Outcome o = drive({{5, 5, 5, 1.2, 3, 8, 1.2, 8}, {1.5, 0.9}});
  EXPECT_TRUE(!o.stalled);
  EXPECT_LE(o.max_in_flight, 2);
  EXPECT_EQ(o.steps[1], 2);
  EXPECT_GE(o.steps[0], 9);
}

// A correct prediction: the paused model converges in the replay already in
// flight and no further replay runs it.
TEST(ReplaySchedule, PauseConfirmed) {
  const Outcome o = drive({{5, 1.2, 0.5}, {9, 9, 9, 9, 9, 0.5}});
  EXPECT_TRUE(!o.stalled);
  EXPECT_EQ(o.steps[0], 3);
  EXPECT_EQ(o.steps[1], 6);
}

// Randomised sequences over 1-6 models: never stalled, at most two replays in
// flight, every model run until it converges, each model's runs matching the
// replays that held it active.
TEST(ReplaySchedule, RandomSequences) {
  std::mt19937 g(7);
  const double vals[] = {0.5, 0.9, 1.1, 1.2, 1.5, 2.0, 3.0, 5.0, 8.0};
  int stalls = 0, over = 0, short_runs = 0;
  for (int t = 0; t < 20000; ++t) {
    const int np = 1 + (int)(g() % 6);
    std::vector<std::vector<double>> res(np);
    for (auto& r : res) {
      const int n = (int)(g() % 14);
      for (int i = 0; i < n; ++i) r.push_back(vals[g() % 9]);
      r.push_back(0.5);  // converges by the end of its list
    }
    const Outcome o = drive(res);
    stalls += o.stalled;
    over += o.max_in_flight > 2;
    for (int k = 0; k < np; ++k) {
      // model k converges at the first entry below 1 of its list
      int want = 0;
      while (res[k][want] >= 1.0) ++want;
      short_runs += o.steps[k] != want + 1;
    }
  }
  EXPECT_EQ(stalls, 0);
  EXPECT_EQ(over, 0);
  EXPECT_EQ(short_runs, 0);
}

int main(int argc, char** argv) { return check::run(argc, argv); }
