/* diag_pattern.h -- the diagonal pattern that most slices of a plan share (host and device
 * code; no HIP dependencies).
 *
 * In a diagonal slice (fixed height S) plane d reads x[pos_d - S + r] for its rows r < S,
 * pos_d being the plane's stored diagonal position.  For a stencil matrix the positions
 * relative to the slice's first row are the same in almost every slice, and they come in
 * runs of values less than S apart, whose windows overlap (one per z-plane of a 27-point
 * stencil when a grid line is shorter than S, else one per (y, z) line).  The plan records
 * that shared list once, with one LDS window per run -- x[first row + rrel[r] + i] for
 * i < rlen[r], staged at rbase[r] -- and marks the slices whose positions equal it, so the
 * SpMV reads those slices' x from LDS without loading any position. */
#ifndef ISPM_DIAG_PATTERN_H
#define ISPM_DIAG_PATTERN_H

namespace ispm {

constexpr int kPatPlanes = 32;

struct diag_pattern {
  unsigned n;                    /* diagonal planes of the shared list */
  unsigned nrun;                 /* runs of diagonals less than S apart */
  unsigned poff[kPatPlanes];     /* LDS index of plane d's x entry for row 0 */
  int rrel[kPatPlanes];          /* run r's window starts at x[first row + rrel[r]] */
  unsigned rbase[kPatPlanes];    /* LDS index of run r's window */
  unsigned rlen[kPatPlanes];     /* window length (S + span of the run) */
  unsigned lds_len;              /* sum of the window lengths */
  unsigned diag_plane1;          /* 1 + the plane of the main diagonal (offset 0), 0: none */
  /* Set by the SpMV launcher when it restates the windows as their aligned covers (the
   * enclosing runs of whole 16-byte vectors of x, back to back in LDS): the covers span
   * x[first row + cover_lo, first row + cover_hi); read as one list of vectors, LDS vector v
   * (at index v * VX) belongs to the last run r with wvbase[r] <= v and holds
   * x[first row + wdelta[r] + v * VX]; vec_runs is nonzero when the list fits the kernel's
   * registers (wtotal vectors, at most a fixed number per thread). */
  long long cover_lo, cover_hi;
  unsigned vec_runs;
  unsigned wvbase[kPatPlanes];
  int wdelta[kPatPlanes];
  unsigned wtotal;
};

}  // namespace ispm

#endif
