/* ref_droptol_harness.cpp -- exposes the reference's own DroptolTuner
 * (fastainv/droptol-tuning.h, compiled from the reference tree where it lies) through a
 * C interface, so tests can check the oracle's restatement against the real thing.
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/libref_droptol.so
 * when the reference tree is present; the header is included, never copied. */
#include "fastainv/droptol-tuning.h"

extern "C" {
void *ref_dtt_new(double lo, double hi) { return new DroptolTuner(lo, hi); }
double ref_dtt_get(void *h) { return ((DroptolTuner *)h)->getDroptol(); }
double ref_dtt_best(void *h) { return ((DroptolTuner *)h)->getBestDroptol(); }
void ref_dtt_note(void *h, double t) { ((DroptolTuner *)h)->noteTime(t); }
void ref_dtt_restart(void *h) { ((DroptolTuner *)h)->restart(); }
void ref_dtt_reset_phase(void *h) { ((DroptolTuner *)h)->resetPhase(); }
int ref_dtt_iters(void *h) { return ((DroptolTuner *)h)->iterations(); }
void ref_dtt_free(void *h) { delete (DroptolTuner *)h; }
}
