/*
 * reference_score.c -- the binding a maintainer adds to the reference's own
 * C code to run its scalar Smith-Waterman seam on the GPU engine: it defines
 * the reference's score() (smith-waterman.h:1) on top of mnhn_score()
 * (include/mnhn_sw.h).  Link it in place of smith_waterman.o, with
 * -lmnhn_sw.
 */
#include <stddef.h>

#include "mnhn_sw.h"

unsigned int score(char *A, char *B, size_t len_a, size_t len_b, int *work) {
  return mnhn_score(A, B, len_a, len_b, work);
}
