/* droptol-tuning.h -- golden-section search for the approximate-inverse drop tolerance,
 * driven by measured solve times (fastainv/droptol-tuning.h:4-103 of the reference), same
 * interface: probe getDroptol(), report noteTime(t), read getBestDroptol().  The search runs
 * on log2(droptol) in [lo, hi]. */
#ifndef ISPM_DROPTOL_TUNER_H
#define ISPM_DROPTOL_TUNER_H

#include <cmath>

class DroptolTuner
{
  enum { PROBE_FIRST, PROBE_LEFT, PROBE_RIGHT } phase;
  static constexpr double kRatio = 0.3819660112501051518;   /* (3 - sqrt 5) / 2 */

  static double left_point(double lo, double hi) { return lo + (hi - lo) * kRatio; }
  static double mirror(double lo, double hi, double p) { return lo + (hi - p); }

  double lo_, hi_, m1_, m2_, best_;
  double t1_ = 0, t2_ = 0, tbest_ = 0;
  int iters_;

public:
  DroptolTuner(double lo = -10, double hi = -3)
    : phase(PROBE_FIRST), lo_(lo), hi_(hi), m1_(left_point(lo, hi)), m2_(mirror(lo, hi, m1_)),
      best_(m1_), iters_(0)
  {}

  double getDroptol() const { return std::exp2(phase == PROBE_RIGHT ? m2_ : m1_); }
  double getBestDroptol() const { return std::exp2(best_); }

  void noteTime(double time)
  {
    if (phase == PROBE_FIRST) {
      tbest_ = time; best_ = m1_; t1_ = time;
      phase = PROBE_RIGHT;
      return;
    }
    if (phase == PROBE_RIGHT) {
      if (tbest_ > time) { tbest_ = time; best_ = m2_; }
      t2_ = time;
    } else {
      if (tbest_ > time) { tbest_ = time; best_ = m1_; }
      t1_ = time;
    }
    if (t1_ < t2_) {           /* minimum left of m2: drop (m2, hi] */
      hi_ = m2_; m2_ = m1_; m1_ = mirror(lo_, hi_, m2_);
      t2_ = t1_; phase = PROBE_LEFT;
    } else {                   /* drop [lo, m1) */
      lo_ = m1_; m1_ = m2_; m2_ = mirror(lo_, hi_, m1_);
      t1_ = t2_; phase = PROBE_RIGHT;
    }
    iters_++;
  }

  /* search again in a bracket of +-1 around the current left probe */
  void restart()
  {
    phase = PROBE_FIRST;
    lo_ = m1_ - 1; hi_ = m1_ + 1;
    m1_ = left_point(lo_, hi_);
    m2_ = mirror(lo_, hi_, m1_);
    iters_ = 0;
  }
  void resetPhase() { phase = PROBE_FIRST; }
  int iterations() const { return iters_; }
};

#endif
