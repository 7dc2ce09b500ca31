/*
 * gviz_graph_loader.c -- .edges files into a gvizGraph or straight into CSR.
 *
 * The file is read once into memory and split into lines in place; each
 * line is parsed with the reference's rules (src/utils/graphLoader.c:15-39,
 * 60-108): leading blanks skipped, empty / '\r' / '%' lines ignored, two
 * base-10 integers required (strtoll, so a sign and leading white space are
 * accepted and anything after the second number is ignored), the first data
 * line dropped when skip_header is set (it must still hold two integers),
 * negative ids rejected, at least one edge required. Ids are compacted over
 * the dense range [min, max] in order of first appearance, u before v
 * (:116-129, 790-825), and edges are applied in file order.
 */
#include "gviz/gvizGraphLoader.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void gvizEdgesFileOptionsInit(gvizEdgesFileOptions *opts) {
  opts->directed = 0;
  opts->skip_header = 0;
}

/* 1: edge parsed, 0: line to skip, -1: malformed. */
static int parse_line(const char *p, long long *u, long long *v) {
  while (*p == ' ' || *p == '\t')
    p++;
  if (*p == '\0' || *p == '\r' || *p == '%')
    return 0;
  char *end = NULL;
  const long long a = strtoll(p, &end, 10);
  if (end == p)
    return -1;
  p = end;
  while (*p == ' ' || *p == '\t')
    p++;
  const long long b = strtoll(p, &end, 10);
  if (end == p)
    return -1;
  *u = a;
  *v = b;
  return 1;
}

static char *read_all(const char *path, size_t *len) {
  FILE *f = fopen(path, "rb");
  if (!f)
    return NULL;
  size_t cap = 1 << 16, n = 0;
  char *buf = malloc(cap + 1);
  while (buf) {
    if (n == cap) {
      char *nb = realloc(buf, 2 * cap + 1);
      if (!nb) {
        free(buf);
        buf = NULL;
        break;
      }
      buf = nb;
      cap *= 2;
    }
    const size_t got = fread(buf + n, 1, cap - n, f);
    n += got;
    if (got == 0)
      break;
  }
  const int err = ferror(f);
  fclose(f);
  if (!buf || err) {
    free(buf);
    return NULL;
  }
  buf[n] = '\0';
  *len = n;
  return buf;
}

/* The file's edges over dense ids: count pairs in (from, to) order, and the
 * vertex count. 0 / -1. */
static int load_pairs(const char *path, const gvizEdgesFileOptions *opts, int64_t **pairs,
                      int64_t *count, int64_t *nv) {
  size_t len = 0;
  char *buf = read_all(path, &len);
  if (!buf)
    return -1;
  size_t cap = 1024, m = 0;
  long long *raw = malloc(sizeof(long long) * 2 * cap);
  long long lo = 0, hi = 0;
  int skip = opts->skip_header, rc = 0;
  char *line = buf;
  while (raw && line <= buf + len) {
    char *nl = memchr(line, '\n', (size_t)(buf + len - line));
    if (nl)
      *nl = '\0';
    long long u = 0, v = 0;
    const int r = parse_line(line, &u, &v);
    if (r < 0) {
      rc = -1;
      break;
    }
    if (r > 0) {
      if (skip) {
        skip = 0;
      } else if (u < 0 || v < 0) {
        rc = -1;
        break;
      } else {
        if (m == cap) {
          long long *nr = realloc(raw, sizeof(long long) * 4 * cap);
          if (!nr) {
            rc = -1;
            break;
          }
          raw = nr;
          cap *= 2;
        }
        if (m == 0)
          lo = hi = u;
        lo = u < lo ? u : lo;
        hi = u > hi ? u : hi;
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
        raw[2 * m] = u;
        raw[2 * m + 1] = v;
        m++;
      }
    }
    if (!nl)
      break;
    line = nl + 1;
  }
  free(buf);
  if (!raw || rc != 0 || m == 0) {
    free(raw);
    return -1;
  }
  /* dense map over [lo, hi] as the reference allocates it */
  const unsigned long long span = (unsigned long long)(hi - lo) + 1ull;
  if (span > (unsigned long long)(SIZE_MAX / sizeof(int64_t))) {
    free(raw);
    return -1;
  }
  int64_t *map = malloc(sizeof(int64_t) * (size_t)span);
  int64_t *out = malloc(sizeof(int64_t) * 2 * m);
  if (!map || !out) {
    free(map);
    free(out);
    free(raw);
    return -1;
  }
  memset(map, 0xff, sizeof(int64_t) * (size_t)span);
  int64_t next = 0;
  for (size_t e = 0; e < 2 * m; e++) {
    int64_t *slot = &map[raw[e] - lo];
    if (*slot < 0)
      *slot = next++;
    out[e] = *slot;
  }
  free(map);
  free(raw);
  *pairs = out;
  *count = (int64_t)m;
  *nv = next;
  return 0;
}

int gvizGraphLoadFromEdgesFile(const char *path, const gvizEdgesFileOptions *opts,
                               gvizGraph *out) {
  if (!path || !opts || !out)
    return -1;
  memset(out, 0, sizeof(*out));
  int64_t *pairs = NULL, m = 0, n = 0;
  if (load_pairs(path, opts, &pairs, &m, &n))
    return -1;
  int rc = gvizGraphInit(out, opts->directed ? 1 : 0);
  for (int64_t i = 0; rc == 0 && i < n; i++)
    rc = gvizGraphAddVertex(out, NULL, NULL, NULL);
  for (int64_t e = 0; rc == 0 && e < m; e++)
    rc = gvizGraphAddEdge(out, (size_t)pairs[2 * e], (size_t)pairs[2 * e + 1]);
  free(pairs);
  if (rc != 0) {
    gvizGraphRelease(out);
    memset(out, 0, sizeof(*out));
    return -1;
  }
  return 0;
}

int gvizGraphLoadEdgesFileCSR(const char *path, const gvizEdgesFileOptions *opts,
                              int64_t *nOut, int64_t **offsetsOut, int64_t **adjOut) {
  if (!path || !opts || !nOut || !offsetsOut || !adjOut)
    return -1;
  int64_t *pairs = NULL, m = 0, n = 0;
  if (load_pairs(path, opts, &pairs, &m, &n))
    return -1;
  const int both = !opts->directed;
  int64_t *off = calloc((size_t)n + 1, sizeof(int64_t));
  int64_t *adj = malloc(sizeof(int64_t) * (size_t)(both ? 2 * m : m));
  if (!off || !adj) {
    free(off);
    free(adj);
    free(pairs);
    return -1;
  }
  for (int64_t e = 0; e < m; e++) {
    off[pairs[2 * e] + 1]++;
    if (both)
      off[pairs[2 * e + 1] + 1]++;
  }
  for (int64_t i = 0; i < n; i++)
    off[i + 1] += off[i];
  /* rows fill in edge order: the insertion order of gvizGraphAddEdge */
  int64_t *cur = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  if (!cur) {
    free(off);
    free(adj);
    free(pairs);
    return -1;
  }
  memcpy(cur, off, sizeof(int64_t) * (size_t)n);
  for (int64_t e = 0; e < m; e++) {
    const int64_t a = pairs[2 * e], b = pairs[2 * e + 1];
    adj[cur[a]++] = b;
    if (both)
      adj[cur[b]++] = a;
  }
  free(cur);
  free(pairs);
  *nOut = n;
  *offsetsOut = off;
  *adjOut = adj;
  return 0;
}
