/*
 * io_internal.h -- shared machinery of the SHAPE input readers (par, mod, obs).
 *
 * Files are read whole into memory and scanned by a cursor; the token grammar
 * is the reference's (src/util/gettstr.c, getint.c, getdouble.c,
 * src/shape/readparam.c): whitespace-separated words, comments enclosed in
 * curly braces, a word ending at whitespace or at the '{' of a comment.
 *
 * Errors never exit: a reader records a message and unwinds with longjmp to
 * its entry point, which frees everything the call allocated and returns a
 * negative code.  Every allocation of a successful call is recorded under the
 * caller's root struct (par_t, mod_t or dat_t), so shape_io_free_*() can
 * release it without the caller knowing the layout.
 */
#ifndef SHAPE_IO_INTERNAL_H
#define SHAPE_IO_INTERNAL_H

#include <setjmp.h>
#include <stddef.h>

#include "shape_abi.h"
#include "shape_io.h"

#define IO_MAXLEN 255

/* ---- allocation records ------------------------------------------------ */
typedef struct io_arena {
  void **p;
  size_t n, cap;
} io_arena;

/* ---- one read call ----------------------------------------------------- */
typedef struct io_ctx {
  jmp_buf jb;
  io_arena arena;
  char err[512];
  int code;          /* SHAPE_IO_E* of the failure */
  char *bufs[16];    /* open files' contents and other temporaries of this call */
} io_ctx;

/* A text or binary file held in memory, scanned by a cursor. */
typedef struct io_file {
  char *buf;
  size_t len, pos;
  char name[IO_MAXLEN + 1];
} io_file;

void io_fail(io_ctx *cx, const char *fmt, ...) __attribute__((noreturn, format(printf, 2, 3)));
/* io_fail with an error code other than SHAPE_IO_EPARSE */
void io_fail_code(io_ctx *cx, int code, const char *fmt, ...) __attribute__((noreturn, format(printf, 3, 4)));
/* after a failure: free the call's allocations and open files, publish the
   message for shape_io_error(); returns the error code */
int io_unwind(io_ctx *cx);

void *io_calloc(io_ctx *cx, size_t n, size_t sz);
double *io_vector(io_ctx *cx, int nl, int nh);
int *io_ivector(io_ctx *cx, int nl, int nh);
double **io_matrix(io_ctx *cx, int nrl, int nrh, int ncl, int nch);
int **io_imatrix(io_ctx *cx, int nrl, int nrh, int ncl, int nch);

/* the call's allocations move to / are freed with a root struct */
void io_arena_commit(io_ctx *cx, const void *root);
void io_arena_discard(io_ctx *cx);

/* ---- files and tokens -------------------------------------------------- */
/* Loads a whole file (a ".gz" name is decompressed through gzip -dc). */
void io_open(io_ctx *cx, io_file *f, const char *path);
void io_close(io_ctx *cx, io_file *f);
/* scratch memory of one call, released by io_temp_free or by the unwind */
void *io_temp_alloc(io_ctx *cx, size_t bytes);
void io_temp_free(io_ctx *cx, void *p);

/* next word (gettstr); returns 0 and an empty word at end of file */
int io_word(io_file *f, char *w, size_t cap);
int io_int(io_ctx *cx, io_file *f);
double io_double(io_ctx *cx, io_file *f);
/* readparam: state c / f / = and a value; 1 if 'f', 0 if 'c' or '=', -1 at end of file */
int io_param(io_ctx *cx, io_file *f, struct param_t *p);
/* calendar date yyyy mm dd hh mm ss as six integers -> Julian date */
double io_caldate(io_ctx *cx, io_file *f);
/* the rest of the current line with {comments} on it removed */
void io_restofline(io_file *f, char *line, size_t cap);
/* number of words left in the file (countdata), and whether none are left */
int io_count_words(io_file *f);
int io_no_more_words(io_file *f);

int io_str_to_int(io_ctx *cx, const char *s);
double io_str_to_double(io_ctx *cx, const char *s);

/* ---- calendar and ephemerides ------------------------------------------ */
double io_cal2jd(int yy, int mo, int dd, int hh, int mi, int ss);
/* Interpolated observing geometry at epoch t (ephem2mat.c:27-219): oe, se,
   the plane-of-sky contribution to the spin s, solar phase and azimuth;
   returns the target distance in AU.  bistatic = 0 skips the Sun. */
double io_ephem2mat(io_ctx *cx, const struct ephem_t *ast, const struct ephem_t *sun, double t,
                    double oe[3][3], double se[3][3], double s[3], double *solar_phase,
                    double *solar_azimuth, int bistatic);

/* ---- model set-up shared with the dataset reader ----------------------- */
void io_setup_sides(io_ctx *cx, struct vertices_t *vt);
void io_setup_vertices(io_ctx *cx, struct vertices_t *vt);

#endif
