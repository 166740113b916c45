/*
 * Records of Fortran's D23.15 edit descriptor, formatted and read without
 * the Fortran runtime's per-item I/O: the host's out_*.dat writer
 * (postproceso_dat, after upstream's mod_postproceso.f90 postproceso_dat
 * with its "(3D23.15)" and "(4D23.15)" records) hands whole arrays of records
 * to d23_records / d23_append, and its reader of the mesh and restart files
 * (lectura_mallas_escalonadas, after upstream's mod_malla.f90:137-162) takes
 * them from d23_read (reading: below).
 *
 * A value is written as flang writes it: 23 characters, right-justified;
 * a minus sign when the sign bit is set (-0.0 included); "0." and the 15
 * significant digits of |x| correctly rounded to nearest (ties to even);
 * the exponent of the 0.ddd form as "D+dd" / "D-dd", or "+ddd" / "-ddd"
 * when it has three digits; zero as 0.000000000000000D+00.
 *
 * The 15 digits come from y = |x| 10^(14-E), E the decimal exponent, with
 * 10^(14-E) as a double-double (hi + lo, built once in binary128) and the
 * product a * hi split exactly (Dekker), so y is known to about 1e-16 at
 * y < 1e15.  When y lies within 1e-6 of halfway between two integers, and
 * for |x| outside [1e-280, 1e290] (where a split could overflow), the C
 * library's exact conversion (snprintf "%.14e") decides instead.
 */
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define KMAX 300 /* |14 - E| for 1e-280 <= |x| <= 1e290, with room for the search; 2^27 10^KMAX
                       * must not overflow */

/* 10^k = hi + lo; hh + hl = hi, each half with at most 26 significant bits */
typedef struct {
  double hi, lo, hh, hl;
} Pow10;

static Pow10 p10[2 * KMAX + 1];
static pthread_once_t p10_once = PTHREAD_ONCE_INIT;

static void split(double x, double *h, double *l) {
  const double c = 134217729.0 * x; /* 2^27 + 1 */
  *h = c - (c - x);
  *l = x - *h;
}

static void init_p10(void) {
  __float128 up = 1, q;
  for (int k = 0; k <= KMAX; ++k) {
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      if (k == 0 && sgn < 0) continue;
      q = sgn > 0 ? up : 1 / up;
      Pow10 *p = &p10[KMAX + sgn * k];
      p->hi = (double)q;
      p->lo = (double)(q - (__float128)p->hi);
      split(p->hi, &p->hh, &p->hl);
    }
    up = up * 10;
  }
}

/* d.dddddddddddddde+XX from the exact conversion: n and E */
static void digits_exact(double a, int64_t *n_out, int *e_out) {
  char buf[40];
  snprintf(buf, sizeof buf, "%.14e", a);
  int64_t n = 0;
  const char *c = buf;
  for (; *c && *c != 'e'; ++c)
    if (*c >= '0' && *c <= '9') n = n * 10 + (*c - '0');
  *n_out = n;
  *e_out = (int)strtol(c + 1, NULL, 10);
}

/* 15 significant digits of a > 0 (finite) as n in [1e14, 1e15) and the
 * exponent q of the 0.ddd form: a ~ 0.n x 10^q */
static void digits15(double a, int64_t *n_out, int *q_out) {
  int64_t n = 0;
  int E = 0;
  int settled = 0;
  if (a >= 1e-280 && a <= 1e290) {
    uint64_t bits;
    memcpy(&bits, &a, sizeof bits);
    const int e2 = (int)((bits >> 52) & 0x7ff) - 1023; /* 2^e2 <= a < 2^(e2+1) */
    E = (e2 * 78913) >> 18; /* floor(e2 log10 2), within one of floor(log10 a) */
    double ah, al;
    split(a, &ah, &al);
    for (int t = 0; t < 4 && !settled; ++t) {
      const Pow10 *p = &p10[14 - E + KMAX];
      const double yh = a * p->hi;
      const double yl = (((ah * p->hh - yh) + ah * p->hl) + al * p->hh) + al * p->hl + a * p->lo;
      if (yh < 1.0e14) {
        --E;
      } else if (yh >= 1.0e15) {
        ++E;
      } else {
        settled = 1;
        n = (int64_t)yh;
        const double f = (yh - (double)n) + yl; /* yh - n exact: yh has a unit of 2^-6 or more */
        if (fabs(f - 0.5) < 1.0e-6) settled = 0; /* too close to call */
        else if (f > 0.5) ++n;
        break;
      }
    }
  }
  if (!settled) digits_exact(a, &n, &E);
  if (n == 1000000000000000LL) {
    n = 100000000000000LL;
    ++E;
  }
  *n_out = n;
  *q_out = E + 1;
}

static const char kTwo[] =
    "00010203040506070809101112131415161718192021222324252627282930313233343536373839"
    "40414243444546474849505152535455565758596061626364656667686970717273747576777879"
    "8081828384858687888990919293949596979899";

/* one value, 23 characters at out; 0 on success, 1 for a non-finite value.
 * Both exponent forms leave the same layout: two blanks or a blank and the
 * sign, "0.", 15 digits, then "D+dd" or "+ddd". */
static int d23_15(double x, char *out) {
  if (!isfinite(x)) return 1;
  const double a = fabs(x);
  int64_t n = 0;
  int q = 0;
  if (a != 0.0) digits15(a, &n, &q);
  out[0] = ' ';
  out[1] = signbit(x) ? '-' : ' ';
  out[2] = '0';
  out[3] = '.';
  /* the 15 digits (out[4..18]) as 7 + 8, two at a time */
  uint32_t hi = (uint32_t)(n / 100000000), lo = (uint32_t)(n % 100000000);
  for (int d = 18; d >= 12; d -= 2) {
    const uint32_t two = lo % 100;
    lo /= 100;
    out[d] = kTwo[2 * two + 1];
    out[d - 1] = kTwo[2 * two];
  }
  for (int d = 10; d >= 5; d -= 2) {
    const uint32_t two = hi % 100;
    hi /= 100;
    out[d] = kTwo[2 * two + 1];
    out[d - 1] = kTwo[2 * two];
  }
  out[4] = (char)('0' + hi);
  const int aq = q < 0 ? -q : q;
  const char sign = q < 0 ? '-' : '+';
  if (aq <= 99) {
    out[19] = 'D';
    out[20] = sign;
  } else {
    out[19] = sign;
    out[20] = (char)('0' + aq / 100);
  }
  out[21] = (char)('0' + (aq / 10) % 10);
  out[22] = (char)('0' + aq % 10);
  return 0;
}

typedef struct {
  const double *vals;
  int64_t r0, r1;
  int32_t ncol;
  char *out;
  int status;
} Chunk;

static void *format_chunk(void *arg) {
  Chunk *c = (Chunk *)arg;
  const int64_t width = 23 * (int64_t)c->ncol + 1;
  c->status = 0;
  for (int64_t r = c->r0; r < c->r1 && !c->status; ++r) {
    char *line = c->out + r * width;
    for (int32_t k = 0; k < c->ncol; ++k)
      if (d23_15(c->vals[r * c->ncol + k], line + 23 * k)) c->status = 1;
    line[23 * c->ncol] = '\n';
  }
  return NULL;
}

#define MAX_THREADS 8
#define MIN_RECORDS 16384 /* per thread */

/* nrec records of ncol values each (vals: record after record), each record
 * ncol fields of 23 characters and a newline, into out (nrec * (23 ncol + 1)
 * characters), on up to 8 threads (each a contiguous range of records).
 * Returns 0, or 1 if a value is not finite (out is then incomplete and the
 * caller writes the records itself). */
int d23_records(const double *vals, int64_t nrec, int32_t ncol, char *out) {
  pthread_once(&p10_once, init_p10);
  const long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
  int64_t nt = nrec / MIN_RECORDS;
  if (nt > MAX_THREADS) nt = MAX_THREADS;
  if (nt > ncpu) nt = ncpu;
  if (nt < 1) nt = 1;
  Chunk ch[MAX_THREADS];
  pthread_t th[MAX_THREADS];
  int started[MAX_THREADS] = {0};
  for (int64_t t = 0; t < nt; ++t) {
    ch[t] = (Chunk){vals, nrec * t / nt, nrec * (t + 1) / nt, ncol, out, 0};
    if (t > 0) started[t] = pthread_create(&th[t], NULL, format_chunk, &ch[t]) == 0;
  }
  format_chunk(&ch[0]);
  int status = ch[0].status;
  for (int64_t t = 1; t < nt; ++t) {
    if (started[t])
      pthread_join(th[t], NULL);
    else
      format_chunk(&ch[t]); /* no thread: this one does the chunk */
    status |= ch[t].status;
  }
  return status;
}

/* the same records appended to the file named by path (pathlen characters,
 * not terminated, as Fortran passes them).  Returns 0; 1 if a value is not
 * finite (nothing is appended); 2 if the file cannot be written. */
int d23_append(const char *path, int32_t pathlen, const double *vals, int64_t nrec, int32_t ncol) {
  const size_t bytes = (size_t)nrec * (size_t)(23 * ncol + 1);
  char *buf = (char *)malloc(bytes + 1);
  char *name = (char *)malloc((size_t)pathlen + 1);
  if (!buf || !name) {
    free(buf);
    free(name);
    return 2;
  }
  memcpy(name, path, (size_t)pathlen);
  name[pathlen] = 0;
  int status = d23_records(vals, nrec, ncol, buf);
  if (status == 0) {
    FILE *f = fopen(name, "ab");
    if (!f || fwrite(buf, 1, bytes, f) != bytes) status = 2;
    if (f && fclose(f) != 0) status = 2;
  }
  free(buf);
  free(name);
  return status;
}

/* ------------------------------------------------------------------ reading
 * The host's mesh and restart files (upstream's mod_malla.f90:137-162 reads
 * them with "(3D23.15)" / "(4D23.15)" after a list-directed header line)
 * parsed without the Fortran runtime's per-item input.  A field is read as
 * Fortran's D23.15 edit descriptor reads it when it holds an optional sign,
 * digits with a decimal point, and an optional exponent (D, d, E or e and
 * an optional sign, or a sign alone), with blanks only around it; the value
 * is the correctly rounded one (strtod).  Anything else -- a blank field,
 * blanks inside, no decimal point (where Fortran would scale by 10^-15), a
 * short line, too few lines, a value out of range -- makes the call return
 * 1, and the caller reads the file itself. */

static char *read_file(const char *path, int32_t pathlen, size_t *len) {
  char name[4096];
  if (pathlen < 0 || (size_t)pathlen >= sizeof name) return NULL;
  memcpy(name, path, (size_t)pathlen);
  name[pathlen] = 0;
  FILE *f = fopen(name, "rb");
  if (!f) return NULL;
  char *buf = NULL;
  size_t cap = 0, n = 0;
  for (;;) {
    if (n == cap) {
      cap = cap ? 2 * cap : (size_t)1 << 20;
      char *nb = (char *)realloc(buf, cap + 1);
      if (!nb) {
        free(buf);
        fclose(f);
        return NULL;
      }
      buf = nb;
    }
    const size_t got = fread(buf + n, 1, cap - n, f);
    n += got;
    if (got == 0) break;
  }
  fclose(f);
  buf[n] = 0;
  *len = n;
  return buf;
}

/* one 23-character field; 0 on success */
static int parse_d23(const char *f, double *out) {
  int i = 0, end = 23;
  while (i < end && f[i] == ' ') ++i;
  while (end > i && f[end - 1] == ' ') --end;
  if (i == end) return 1;
  char b[32];
  int k = 0, dot = 0, digit = 0, expo = 0;
  for (; i < end; ++i) {
    const char c = f[i];
    if (c >= '0' && c <= '9') {
      digit = 1;
      b[k++] = c;
    } else if (c == '.') {
      if (dot || expo) return 1;
      dot = 1;
      b[k++] = c;
    } else if (c == 'D' || c == 'd' || c == 'E' || c == 'e') {
      if (expo || !digit) return 1;
      expo = 1;
      b[k++] = 'e';
    } else if (c == '+' || c == '-') {
      if (k == 0 || b[k - 1] == 'e') {
        b[k++] = c;
      } else if (!expo && digit) {
        expo = 1;  /* an exponent given by its sign alone */
        b[k++] = 'e';
        b[k++] = c;
      } else {
        return 1;
      }
    } else {
      return 1;
    }
    if (k > 28) return 1;
  }
  if (!dot || !digit) return 1;
  b[k] = 0;
  char *e;
  const double v = strtod(b, &e);
  if (*e || !isfinite(v)) return 1;
  *out = v;
  return 0;
}

typedef struct {
  const char *const *line;
  int64_t r0, r1;
  int32_t ncol;
  double *vals;
  int status;
} ParseChunk;

static void *parse_chunk(void *arg) {
  ParseChunk *c = (ParseChunk *)arg;
  c->status = 0;
  for (int64_t r = c->r0; r < c->r1 && !c->status; ++r)
    for (int32_t k = 0; k < c->ncol && !c->status; ++k)
      c->status = parse_d23(c->line[r] + 23 * k, &c->vals[r * c->ncol + k]);
  return NULL;
}

/* the nrec records of ncol values after the header line of the file named
 * by path (pathlen characters, not terminated) into vals, record after
 * record, parsed on up to 8 threads.  Returns 0, or 1 if the file is not
 * of the form described above (vals is then incomplete). */
int d23_read(const char *path, int32_t pathlen, int64_t nrec, int32_t ncol, double *vals) {
  size_t len = 0;
  char *buf = read_file(path, pathlen, &len);
  if (!buf) return 1;
  const char **line = (const char **)malloc((size_t)(nrec > 0 ? nrec : 1) * sizeof *line);
  int status = line ? 0 : 1;
  char *p = buf, *endb = buf + len;
  char *nl = status ? NULL : memchr(p, '\n', (size_t)(endb - p)); /* end of the header */
  if (!nl) status = 1;
  p = nl ? nl + 1 : endb;
  for (int64_t r = 0; r < nrec && !status; ++r) {
    char *q = memchr(p, '\n', (size_t)(endb - p));
    char *e = q ? q : endb;
    if (e - p < 23 * (ptrdiff_t)ncol) {
      status = 1;
      break;
    }
    line[r] = p;
    p = q ? q + 1 : endb;
  }
  if (!status) {
    const long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    int64_t nt = nrec / MIN_RECORDS;
    if (nt > MAX_THREADS) nt = MAX_THREADS;
    if (nt > ncpu) nt = ncpu;
    if (nt < 1) nt = 1;
    ParseChunk ch[MAX_THREADS];
    pthread_t th[MAX_THREADS];
    int started[MAX_THREADS] = {0};
    for (int64_t t = 0; t < nt; ++t) {
      ch[t] = (ParseChunk){line, nrec * t / nt, nrec * (t + 1) / nt, ncol, vals, 0};
      if (t > 0) started[t] = pthread_create(&th[t], NULL, parse_chunk, &ch[t]) == 0;
    }
    parse_chunk(&ch[0]);
    status = ch[0].status;
    for (int64_t t = 1; t < nt; ++t) {
      if (started[t])
        pthread_join(th[t], NULL);
      else
        parse_chunk(&ch[t]);
      status |= ch[t].status;
    }
  }
  free(line);
  free(buf);
  return status;
}

/* the lines after the first of the file named by path that hold a
 * character other than a blank, into *n (the record count of a mesh file);
 * 0 on success, 1 if the file cannot be read */
int d23_count_lines(const char *path, int32_t pathlen, int64_t *n) {
  size_t len = 0;
  char *buf = read_file(path, pathlen, &len);
  if (!buf) return 1;
  int64_t count = 0;
  const char *p = memchr(buf, '\n', len);
  p = p ? p + 1 : buf + len;
  const char *endb = buf + len;
  while (p < endb) {
    const char *q = memchr(p, '\n', (size_t)(endb - p));
    const char *e = q ? q : endb;
    for (const char *c = p; c < e; ++c)
      if (*c != ' ') {
        ++count;
        break;
      }
    p = q ? q + 1 : endb;
  }
  free(buf);
  *n = count;
  return 0;
}
