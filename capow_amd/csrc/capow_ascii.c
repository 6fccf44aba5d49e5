/* capow_ascii.c -- parallel reader of the periodic-box ASCII catalogue.
 *
 * Replaces the parse loop of read_periodic_serial_ascii (reference
 * src/read_data.c:32-88) with the same semantics, split over threads:
 *   - the file is read as fgets with the reference's 2000-character buffer
 *     reads it: each physical line in pieces of at most 1999 characters (the
 *     last piece keeps the newline), and every piece is one "line";
 *   - a piece starting with '#' is skipped;
 *   - a piece must hold six numbers, converted by strtod as sscanf("%lf")
 *     does (extra columns ignored), or the read fails with that piece;
 *   - v_los is column 4, 5 or 6 for LOS = 1, 2, otherwise (read_data.c:65-66);
 *   - with LOS = 0 a particle outside [MIN, MAX] on any axis is fatal.
 * The file is mapped, cut into one range per thread at line starts, and
 * every thread parses its range into its own rows; errors are reported for
 * the first failing line in file order, so the result does not depend on
 * the thread count.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "capow.h"

#define LINE_BUF 2000   /* the reference's fgets buffer (src/read_data.c:35) */

typedef struct {
  const char *base;
  size_t beg, end;          /* byte range of whole physical lines */
  int32_t los, fp64;
  int32_t raw6;             /* keep the six numbers as double (no LOS, no box check) */
  double lo[3], hi[3];
  unsigned char *rows;      /* (x, y, z, v_los) per accepted line, or the six numbers */
  size_t n, cap;
  int err;                  /* CAPOW_OK or the first error in this range */
  size_t err_off, err_len;  /* the failing piece */
  double err_xyz[3];
} chunk_t;

/* Six numbers as sscanf(buf, "%lf %lf %lf %lf %lf %lf") reads them. */
static int parse6(const char *s, double *v) {
  char *end;
  for (int q = 0; q < 6; q++) {
    v[q] = strtod(s, &end);
    if (end == s) return q;
    s = end;
  }
  return 6;
}

static int push_row(chunk_t *c, const double *v) {
  const size_t esz = c->fp64 ? 8 : 4;
  const size_t rb = c->raw6 ? 6 * sizeof(double) : 4 * esz;
  if (c->n == c->cap) {
    const size_t cap = c->cap ? 2 * c->cap : 4096;
    unsigned char *r = (unsigned char *)realloc(c->rows, cap * rb);
    if (!r) return CAPOW_ERR_NOMEM;
    c->rows = r;
    c->cap = cap;
  }
  if (c->raw6) {
    memcpy(c->rows + c->n * rb, v, rb);
    c->n++;
    return CAPOW_OK;
  }
  const double tvr = (c->los == 1) ? v[3] : (c->los == 2 ? v[4] : v[5]);
  if (c->fp64) {
    double *r = (double *)c->rows + 4 * c->n;
    r[0] = v[0]; r[1] = v[1]; r[2] = v[2]; r[3] = tvr;
  } else {
    float *r = (float *)c->rows + 4 * c->n;
    r[0] = (float)v[0]; r[1] = (float)v[1]; r[2] = (float)v[2]; r[3] = (float)tvr;
  }
  c->n++;
  return CAPOW_OK;
}

static void *parse_range(void *arg) {
  chunk_t *c = (chunk_t *)arg;
  char buf[LINE_BUF];
  size_t pos = c->beg;
  while (pos < c->end) {
    /* one fgets piece: up to LINE_BUF - 1 characters, through the newline */
    size_t len = 0;
    while (len < LINE_BUF - 1 && pos + len < c->end) {
      const char ch = c->base[pos + len];
      len++;
      if (ch == '\n') break;
    }
    const size_t off = pos;
    pos += len;
    if (c->base[off] == '#') continue;
    memcpy(buf, c->base + off, len);
    buf[len] = 0;
    double v[6];
    if (parse6(buf, v) != 6) {
      c->err = CAPOW_ERR_PARSE;
      c->err_off = off;
      c->err_len = len;
      return NULL;
    }
    if (!c->raw6 && c->los == 0 && ((v[0] < c->lo[0]) || (v[0] > c->hi[0]) || (v[1] < c->lo[1]) ||
                        (v[1] > c->hi[1]) || (v[2] < c->lo[2]) || (v[2] > c->hi[2]))) {
      c->err = CAPOW_ERR_OUT_OF_BOX;
      c->err_off = off;
      c->err_len = len;
      c->err_xyz[0] = v[0];
      c->err_xyz[1] = v[1];
      c->err_xyz[2] = v[2];
      return NULL;
    }
    if (push_row(c, v) != CAPOW_OK) {
      c->err = CAPOW_ERR_NOMEM;
      c->err_off = off;
      c->err_len = 0;
      return NULL;
    }
  }
  return NULL;
}

/* Threads for a file of `size` bytes: nthreads, or (<= 0) OMP_NUM_THREADS
 * if set, else the CPUs this process may run on; at most 16, and at least
 * 1 MB per thread. */
static int pick_threads(int32_t nthreads, size_t size) {
  int nt = nthreads;
  if (nt <= 0) {
    const char *e = getenv("OMP_NUM_THREADS");
    if (e && atoi(e) > 0) {
      nt = atoi(e);
    } else {
      cpu_set_t set;
      CPU_ZERO(&set);
      nt = sched_getaffinity(0, sizeof set, &set) == 0 ? CPU_COUNT(&set) : 1;
    }
    if (nt > 16) nt = 16;
    if (nt < 1) nt = 1;
  }
  const size_t max_nt = size / (1u << 20) + 1;
  if ((size_t)nt > max_nt) nt = (int)max_nt;
  return nt;
}

/* The parse of the text catalogue: (x, y, z, v_los) rows in the given
 * precision with the LOS = 0 box check, or (raw6) the six numbers of every
 * line as double. */
static int parse_text(const char *path, int32_t los, const double lo[3], const double hi[3],
                      int32_t fp64, int32_t raw6, int32_t nthreads, void **rows_out,
                      int64_t *n_out, char *bad_line, int32_t bad_len, double bad_xyz[3]) {
  if (!path || !lo || !hi || !rows_out || !n_out) return CAPOW_ERR_ARG;
  *rows_out = NULL;
  *n_out = 0;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return CAPOW_ERR_IO;
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return CAPOW_ERR_IO;
  }
  const size_t size = (size_t)sb.st_size;
  const size_t rb = raw6 ? 6 * sizeof(double) : 4 * (size_t)(fp64 ? 8 : 4);   /* row bytes */
  if (size == 0) {
    close(fd);
    *rows_out = malloc(rb);
    return *rows_out ? CAPOW_OK : CAPOW_ERR_NOMEM;
  }
  const char *base = (const char *)mmap(NULL, size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (base == MAP_FAILED) return CAPOW_ERR_IO;
  madvise((void *)base, size, MADV_SEQUENTIAL);

  int nt = pick_threads(nthreads, size);
  /* at least 1 MB per thread */
  chunk_t *ch = (chunk_t *)calloc((size_t)nt, sizeof(chunk_t));
  pthread_t *th = (pthread_t *)calloc((size_t)nt, sizeof(pthread_t));
  if (!ch || !th) {
    free(ch);
    free(th);
    munmap((void *)base, size);
    return CAPOW_ERR_NOMEM;
  }
  /* ranges start at the beginning of a physical line */
  size_t prev = 0;
  for (int t = 0; t < nt; t++) {
    size_t b = prev;
    size_t e = (t == nt - 1) ? size : (size_t)((unsigned __int128)size * (t + 1) / nt);
    if (e < b) e = b;
    while (e < size && e > 0 && base[e - 1] != '\n') e++;
    ch[t].base = base;
    ch[t].beg = b;
    ch[t].end = e;
    ch[t].los = los;
    ch[t].fp64 = fp64;
    ch[t].raw6 = raw6;
    for (int a = 0; a < 3; a++) {
      ch[t].lo[a] = lo[a];
      ch[t].hi[a] = hi[a];
    }
    prev = e;
  }
  int started = 0;
  for (int t = 1; t < nt; t++) {
    if (pthread_create(&th[t], NULL, parse_range, &ch[t]) != 0) break;
    started = t;
  }
  parse_range(&ch[0]);
  /* ranges whose thread could not start are parsed here */
  for (int t = started + 1; t < nt; t++) parse_range(&ch[t]);
  for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);

  int rc = CAPOW_OK;
  size_t total = 0;
  for (int t = 0; t < nt; t++) {
    if (ch[t].err != CAPOW_OK) {
      /* ranges are in file order: the first failing range holds the first
         failing line */
      rc = ch[t].err;
      if (bad_line && bad_len > 0) {
        size_t l = ch[t].err_len < (size_t)(bad_len - 1) ? ch[t].err_len : (size_t)(bad_len - 1);
        memcpy(bad_line, base + ch[t].err_off, l);
        bad_line[l] = 0;
      }
      if (bad_xyz)
        for (int a = 0; a < 3; a++) bad_xyz[a] = ch[t].err_xyz[a];
      break;
    }
    total += ch[t].n;
  }
  unsigned char *out = NULL;
  if (rc == CAPOW_OK) {
    out = (unsigned char *)malloc((total ? total : 1) * rb);
    if (!out) {
      rc = CAPOW_ERR_NOMEM;
    } else {
      size_t o = 0;
      for (int t = 0; t < nt; t++) {
        if (ch[t].n) memcpy(out + o * rb, ch[t].rows, ch[t].n * rb);
        o += ch[t].n;
      }
    }
  }
  for (int t = 0; t < nt; t++) free(ch[t].rows);
  free(ch);
  free(th);
  munmap((void *)base, size);
  if (rc == CAPOW_OK) {
    *rows_out = out;
    *n_out = (int64_t)total;
  }
  return rc;
}

int capow_read_ascii(const char *path, int32_t los, const double lo[3], const double hi[3],
                     int32_t fp64, int32_t nthreads, void **rows_out, int64_t *n_out,
                     char *bad_line, int32_t bad_len, double bad_xyz[3]) {
  return parse_text(path, los, lo, hi, fp64, 0, nthreads, rows_out, n_out, bad_line, bad_len,
                    bad_xyz);
}

void capow_free_rows(void *rows) { free(rows); }

/* ---- binary sidecar of the periodic catalogue ------------------------------
 * A 64-byte header, then the six numbers of every catalogue line (x y z vx
 * vy vz) as little-endian doubles in file order -- what the text parse
 * yields before the LOS column is picked and the box checked, so reading
 * the sidecar gives the text reader's rows bit for bit, for any LOS, box and
 * precision.  The header records the text file's size and modification time
 * at the parse; a reader given the text path refuses a sidecar whose text
 * has changed since. */
typedef struct {
  char magic[8];            /* "CAPOWCAT" */
  uint32_t version;         /* 1 */
  uint32_t ncols;           /* 6 */
  uint64_t nrows;
  uint64_t text_size;
  int64_t text_mtime_sec, text_mtime_nsec;
  uint64_t reserved[2];
} sidecar_hdr;
_Static_assert(sizeof(sidecar_hdr) == 64, "sidecar header");

static const char kSidecarMagic[8] = {'C', 'A', 'P', 'O', 'W', 'C', 'A', 'T'};

int capow_write_sidecar(const char *text_path, const char *bin_path, int32_t nthreads,
                        char *bad_line, int32_t bad_len) {
  if (!text_path || !bin_path) return CAPOW_ERR_ARG;
  struct stat sb;
  if (stat(text_path, &sb) != 0) return CAPOW_ERR_IO;
  const double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  void *rows = NULL;
  int64_t n = 0;
  int rc = parse_text(text_path, 0, lo, hi, 1, 1, nthreads, &rows, &n, bad_line, bad_len, NULL);
  if (rc != CAPOW_OK) return rc;
  sidecar_hdr h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, kSidecarMagic, 8);
  h.version = 1;
  h.ncols = 6;
  h.nrows = (uint64_t)n;
  h.text_size = (uint64_t)sb.st_size;
  h.text_mtime_sec = (int64_t)sb.st_mtim.tv_sec;
  h.text_mtime_nsec = (int64_t)sb.st_mtim.tv_nsec;
  /* written beside the target and renamed over it: a reader never sees a
     partial file */
  const size_t pl = strlen(bin_path);
  char *tmp = (char *)malloc(pl + 8);
  if (!tmp) {
    free(rows);
    return CAPOW_ERR_NOMEM;
  }
  memcpy(tmp, bin_path, pl);
  memcpy(tmp + pl, ".part", 6);
  FILE *f = fopen(tmp, "wb");
  rc = f ? CAPOW_OK : CAPOW_ERR_IO;
  if (f) {
    const size_t bytes = (size_t)n * 6 * sizeof(double);
    if (fwrite(&h, sizeof h, 1, f) != 1 || (bytes && fwrite(rows, 1, bytes, f) != bytes))
      rc = CAPOW_ERR_IO;
    if (fclose(f) != 0) rc = CAPOW_ERR_IO;
    if (rc == CAPOW_OK && rename(tmp, bin_path) != 0) rc = CAPOW_ERR_IO;
    if (rc != CAPOW_OK) unlink(tmp);
  }
  free(tmp);
  free(rows);
  return rc;
}

typedef struct {
  const double *src;        /* six numbers per row */
  size_t beg, end;
  int32_t los, fp64;
  double lo[3], hi[3];
  unsigned char *out;       /* (x, y, z, v_los) rows */
  int64_t bad;              /* first row outside the box (LOS = 0), or -1 */
} conv_t;

static void *convert_range(void *arg) {
  conv_t *c = (conv_t *)arg;
  c->bad = -1;
  for (size_t i = c->beg; i < c->end; i++) {
    const double *v = c->src + 6 * i;
    if (c->los == 0 && ((v[0] < c->lo[0]) || (v[0] > c->hi[0]) || (v[1] < c->lo[1]) ||
                        (v[1] > c->hi[1]) || (v[2] < c->lo[2]) || (v[2] > c->hi[2]))) {
      c->bad = (int64_t)i;
      return NULL;
    }
    const double tvr = (c->los == 1) ? v[3] : (c->los == 2 ? v[4] : v[5]);
    if (c->fp64) {
      double *r = (double *)c->out + 4 * i;
      r[0] = v[0]; r[1] = v[1]; r[2] = v[2]; r[3] = tvr;
    } else {
      float *r = (float *)c->out + 4 * i;
      r[0] = (float)v[0]; r[1] = (float)v[1]; r[2] = (float)v[2]; r[3] = (float)tvr;
    }
  }
  return NULL;
}

int capow_read_sidecar(const char *bin_path, const char *text_path, int32_t los,
                       const double lo[3], const double hi[3], int32_t fp64, int32_t nthreads,
                       void **rows_out, int64_t *n_out, double bad_xyz[3]) {
  if (!bin_path || !lo || !hi || !rows_out || !n_out) return CAPOW_ERR_ARG;
  *rows_out = NULL;
  *n_out = 0;
  const int fd = open(bin_path, O_RDONLY);
  if (fd < 0) return CAPOW_ERR_IO;
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return CAPOW_ERR_IO;
  }
  if ((size_t)sb.st_size < sizeof(sidecar_hdr)) {
    close(fd);
    return CAPOW_ERR_PARSE;
  }
  const size_t size = (size_t)sb.st_size;
  const unsigned char *base = (const unsigned char *)mmap(NULL, size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (base == MAP_FAILED) return CAPOW_ERR_IO;
  madvise((void *)base, size, MADV_SEQUENTIAL);
  sidecar_hdr h;
  memcpy(&h, base, sizeof h);
  int rc = CAPOW_OK;
  if (memcmp(h.magic, kSidecarMagic, 8) != 0 || h.version != 1 || h.ncols != 6 ||
      h.nrows > (size - sizeof h) / (6 * sizeof(double)) ||
      size != sizeof h + h.nrows * 6 * sizeof(double))
    rc = CAPOW_ERR_PARSE;
  if (rc == CAPOW_OK && text_path) {
    struct stat tb;
    if (stat(text_path, &tb) != 0) rc = CAPOW_ERR_IO;
    else if ((uint64_t)tb.st_size != h.text_size || (int64_t)tb.st_mtim.tv_sec != h.text_mtime_sec ||
             (int64_t)tb.st_mtim.tv_nsec != h.text_mtime_nsec)
      rc = CAPOW_ERR_STALE;
  }
  if (rc != CAPOW_OK) {
    munmap((void *)base, size);
    return rc;
  }
  const size_t n = (size_t)h.nrows;
  const size_t rb = 4 * (size_t)(fp64 ? 8 : 4);
  unsigned char *out = (unsigned char *)malloc((n ? n : 1) * rb);
  if (!out) {
    munmap((void *)base, size);
    return CAPOW_ERR_NOMEM;
  }
  const double *src = (const double *)(base + sizeof h);
  int nt = pick_threads(nthreads, n * 6 * sizeof(double));
  conv_t *cv = (conv_t *)calloc((size_t)nt, sizeof(conv_t));
  pthread_t *th = (pthread_t *)calloc((size_t)nt, sizeof(pthread_t));
  if (!cv || !th) {
    free(cv);
    free(th);
    free(out);
    munmap((void *)base, size);
    return CAPOW_ERR_NOMEM;
  }
  for (int t = 0; t < nt; t++) {
    cv[t].src = src;
    cv[t].beg = (size_t)((unsigned __int128)n * t / nt);
    cv[t].end = (size_t)((unsigned __int128)n * (t + 1) / nt);
    cv[t].los = los;
    cv[t].fp64 = fp64;
    for (int a = 0; a < 3; a++) {
      cv[t].lo[a] = lo[a];
      cv[t].hi[a] = hi[a];
    }
    cv[t].out = out;
  }
  int started = 0;
  for (int t = 1; t < nt; t++) {
    if (pthread_create(&th[t], NULL, convert_range, &cv[t]) != 0) break;
    started = t;
  }
  convert_range(&cv[0]);
  for (int t = started + 1; t < nt; t++) convert_range(&cv[t]);
  for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);
  /* ranges are in file order: the first range with a row outside the box
     holds the first such row */
  for (int t = 0; t < nt; t++)
    if (cv[t].bad >= 0) {
      rc = CAPOW_ERR_OUT_OF_BOX;
      if (bad_xyz)
        for (int a = 0; a < 3; a++) bad_xyz[a] = src[6 * cv[t].bad + a];
      break;
    }
  free(cv);
  free(th);
  munmap((void *)base, size);
  if (rc != CAPOW_OK) {
    free(out);
    return rc;
  }
  *rows_out = out;
  *n_out = (int64_t)n;
  return CAPOW_OK;
}
