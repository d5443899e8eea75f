/* timer.h -- event timer on the library stream (util/cuda/timer.h:6-52 of the reference):
 * start(), stop() (waits for the stop event), elapsed_seconds(). */
#ifndef ISPM_TIMER_H
#define ISPM_TIMER_H

#include "check.h"

class ispm_event_scoped
{
  void *ev_ = nullptr;
public:
  ispm_event_scoped() { ISPM_CHECK(ispm_event_create(&ev_)); }
  ~ispm_event_scoped() { ispm_event_destroy(ev_); }
  ispm_event_scoped(const ispm_event_scoped &) = delete;
  ispm_event_scoped &operator=(const ispm_event_scoped &) = delete;
  void record() { ISPM_CHECK(ispm_event_record(ev_)); }
  void await() { ISPM_CHECK(ispm_event_synchronize(ev_)); }
  void *operator()() { return ev_; }
};

struct ispm_timer
{
  ispm_event_scoped evt_start, evt_stop;
  void start() { evt_start.record(); }
  void stop()
  {
    evt_stop.record();
    evt_stop.await();
  }
  double elapsed_seconds()
  {
    float ms = 0;
    ISPM_CHECK(ispm_event_elapsed_ms(evt_start(), evt_stop(), &ms));
    return 1e-3 * ms;
  }
};

/* the reference's names (util/cuda/timer.h:6,33); events always record on the library
 * stream */
typedef ispm_event_scoped cudaEventScoped;
typedef ispm_timer cudatimer;

#endif
