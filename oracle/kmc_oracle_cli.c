/*
 * kmc_oracle_cli.c -- run one start.dat deck through the CPU oracle, the way
 * mulskips.e runs it (main.f), writing I????????.xyz frames to an output
 * directory.  TEST INFRASTRUCTURE ONLY.
 *
 * usage: kmc_oracle_cli START_DAT LX LY LZ OUTDIR [NUMBERSLIST_TXT [LOG_BIN]]
 * Prints one JSON line with the final clock, iteration, counts and timing.
 */
#include "kmc_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <time.h>

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s START_DAT LX LY LZ OUTDIR [NUMBERSLIST_TXT [LOG_BIN]]\n", argv[0]);
    return 64;
  }
  orc_sim *s = orc_create(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
  if (!s) { fprintf(stderr, "allocation failed\n"); return 1; }
  if (argc >= 7) {
    FILE *f = fopen(argv[6], "r");
    if (!f) { fprintf(stderr, "cannot open %s\n", argv[6]); return 1; }
    double v[4096];
    int n = 0;
    while (n < 4096 && fscanf(f, "%lf", &v[n]) == 1) n++;
    fclose(f);
    orc_set_numberslist(s, v, n);
  }
  double t0 = now();
  if (orc_load_deck(s, argv[1])) { fprintf(stderr, "load: %s\n", orc_last_error(s)); return 2; }
  double t1 = now();
  orc_event *log = NULL;
  int64_t cap = 0, nlog = 0;
  if (argc >= 8) { cap = 1 << 24; log = (orc_event *)malloc(sizeof(orc_event) * (size_t)cap); }
  int stop = 0;
  int rc = orc_main(s, argv[5], log, cap, &nlog, &stop);
  double t2 = now();
  if (rc) { fprintf(stderr, "run: %s\n", orc_last_error(s)); return 3; }
  if (log) {
    FILE *f = fopen(argv[7], "wb");
    fwrite(log, sizeof(orc_event), (size_t)nlog, f);
    fclose(f);
  }
  int32_t a, b, c;
  orc_counts(s, &a, &b, &c);
  printf("{\"time\": %.17g, \"iter\": %lld, \"n_ad\": %d, \"n_at\": %d, \"n_vd\": %d, \"stop\": %d, "
         "\"init_s\": %.3f, \"loop_s\": %.3f}\n",
         orc_time(s), (long long)orc_iter(s), a, b, c, stop, t1 - t0, t2 - t1);
  orc_destroy(s);
  return 0;
}
