/* Harness around the reference CPUonly decoder (test infrastructure).
 * Compiled together with /root/reference/CPUonly/src/*.c into
 * oracle/_ref/libcpuonly.so; calls its openJPG(path) as
 * CPUonly/src/main.c:71-73 does and copies the decoded pixels out. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <format.h> /* CPUonly/include/format.h */

JPG *openJPG(const char *filename); /* CPUonly/src/main.c:10 */

int ref_cpuonly_decode(const char *path, unsigned char *out, long cap, int *w, int *h, int *nc) {
  JPG *jpg = openJPG(path);
  if (!jpg) return -1;
  *w = jpg->width; *h = jpg->height; *nc = jpg->num_channels;
  long n = (long)jpg->width * jpg->height * jpg->num_channels;
  if (n > cap) { delJPG(jpg); return -2; }
  memcpy(out, jpg->num_channels == 1 ? jpg->channels[0].pixels : jpg->pixels, (size_t)n);
  delJPG(jpg);
  return 0;
}

/* Decode each path `reps` times; per-decode wall times (seconds, monotonic
 * clock) go to times[i*reps + r].  Returns the number of failed decodes. */
int ref_cpuonly_time(const char **paths, int n, int reps, double *times) {
  int fails = 0;
  for (int i = 0; i < n; i++)
    for (int r = 0; r < reps; r++) {
      struct timespec a, b;
      clock_gettime(CLOCK_MONOTONIC, &a);
      JPG *jpg = openJPG(paths[i]);
      clock_gettime(CLOCK_MONOTONIC, &b);
      times[i * reps + r] = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
      if (!jpg) fails++;
      else delJPG(jpg);
    }
  return fails;
}
