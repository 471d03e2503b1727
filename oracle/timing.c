/* Timing helper for the CPU baseline -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Times or_align itself in a C loop, so the figure holds no Python call
 * overhead: the median of `repeats` alignments of one pair. */
#define _POSIX_C_SOURCE 199309L
#include <stdlib.h>
#include <time.h>

#include "oracle.h"

static int cmp_double(const void *x, const void *y) {
  const double a = *(const double *)x, b = *(const double *)y;
  return a < b ? -1 : (a > b ? 1 : 0);
}

int or_time_align(const or_params *p, const int8_t *smat, int32_t K, const uint8_t *a, int32_t m,
                  const uint8_t *b, int32_t n, int32_t repeats, char *cigar, int64_t cap, double *median_s) {
  if (repeats < 1 || !median_s) return -1;
  double *t = (double *)malloc(sizeof(double) * (size_t)repeats);
  if (!t) return -5;
  or_result res;
  for (int32_t r = 0; r < repeats; ++r) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    const int rc = or_align(p, smat, K, a, m, b, n, &res, cigar, cap);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (rc) {
      free(t);
      return rc;
    }
    t[r] = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  }
  qsort(t, (size_t)repeats, sizeof(double), cmp_double);
  *median_s = t[repeats / 2];
  free(t);
  return 0;
}
