/* ref_xerbla.c -- recording XERBLA linked into oracle/_ref (test only).
 *
 * The reference XERBLA prints and STOPs (SRC/xerbla.f:86-88); the
 * reference's own test suite replaces it at link time with one that records
 * the routine name and parameter number (TESTING/LIN/xerbla.f:74-118).  This
 * is that role in C: the oracle's error-exit tests read the last call back.
 */
#include <stddef.h>
#include <string.h>

static char g_name[16];
static int g_info;
static int g_calls;

void xerbla_(const char *srname, const int *info, size_t len) {
  size_t n = len < sizeof(g_name) - 1 ? len : sizeof(g_name) - 1;
  memcpy(g_name, srname, n);
  g_name[n] = 0;
  while (n > 0 && g_name[n - 1] == ' ') g_name[--n] = 0;
  g_info = *info;
  g_calls++;
}

int ref_xerbla_last(char *name_out, int cap) {
  if (cap > 0) {
    strncpy(name_out, g_name, (size_t)cap - 1);
    name_out[cap - 1] = 0;
  }
  return g_info;
}

int ref_xerbla_calls(void) { return g_calls; }

void ref_xerbla_reset(void) {
  g_name[0] = 0;
  g_info = 0;
  g_calls = 0;
}
