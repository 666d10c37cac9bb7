/* The C oracle under AddressSanitizer / UBSan: Glynn port and double-double
 * sums, gradient and column minors, Gray states, tensor permanent and the
 * matrix-polynomial permanent on small random inputs.  Exits 0 when every
 * call returned 0 (or -1 for the planted sum mismatch). */
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

static double rnd(unsigned *s) {
  *s = *s * 1103515245u + 12345u;
  return ((*s >> 8) & 0xffff) / 32768.0 - 1.0;
}

int main(void) {
  unsigned s = 7;
  int bad = 0;
  double a[2 * 36], out[2 * 6], g[2 * 36];
  for (int i = 0; i < 72; ++i) a[i] = rnd(&s);
  int rows[6] = {1, 2, 0, 1, 1, 1}, cols[6] = {2, 1, 1, 1, 0, 1};
  bad |= pqo_perm_port_c128(a, 6, 6, rows, cols, 3, out) != 0;
  bad |= pqo_perm_dd_c128(a, 6, 6, rows, cols, 2, out) != 0;
  bad |= pqo_grad_dd_c128(a, 6, 6, rows, cols, 2, g) != 0;
  int lr[6] = {1, 1, 0, 1, 1, 1}, lc[6] = {2, 1, 1, 1, 0, 1};
  bad |= pqo_laplace_dd_c128(a, 6, 6, lr, lc, 2, out) != 0;
  int mism[6] = {1, 1, 1, 1, 1, 0};
  bad |= pqo_perm_dd_c128(a, 6, 6, mism, cols, 1, out) != -1;
  int lim[3] = {2, 3, 2}, codes[3 * 12];
  bad |= pqo_gray_states(lim, 3, 0, 12, codes) != 0;
  double asum;
  bad |= pqo_tensor_perm_dd(a, 5, 0.4, 2, out, &asum) != 0;
  double bm[2 * 2 * 16];
  for (int i = 0; i < 64; ++i) bm[i] = rnd(&s);
  int32_t ex[6] = {1, 0, 2, 1, 0, 0};
  double pout[6], pabs[3];
  bad |= pqo_perm_poly_dd(a, bm, 4, 2, ex, 3, pout, pabs) != 0;
  printf("oracle: %s\n", bad ? "failures" : "ok");
  return bad;
}
