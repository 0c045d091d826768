/*
 * shape_io.h -- readers of SHAPE's input files (parameter, model and obs
 * files) into the structs of shape_abi.h.  Host-only C, exported from
 * libshape_io.so.
 *
 * Two forms of each reader:
 *   * shape_read_par / shape_read_mod / shape_read_dat return an error code
 *     (negative; the message from shape_io_error()) and never exit.  What a
 *     successful call allocated is released by shape_io_free_*.
 *   * read_par / read_mod / read_dat keep the reference's names, signatures
 *     and error behaviour (src/shape/shape2.h:1513,1512,1510): a malformed
 *     file prints a message and ends the process with exit status 2, as the
 *     reference's bailout() does, so they can replace read_par.o, read_mod.o
 *     and read_dat.o in a SHAPE host build.
 *
 * Supported: every par keyword and penalty of read_par.c; model components
 * of the four types with every radar and optical law read_mod.c parses, the
 * spin state with YORP, libration and impulses; delay-Doppler, Doppler,
 * lightcurve and heliocurve datasets with ASCII, raw binary or RDF data
 * files (optionally gzip-compressed) and pixel-weighting masks.  Refused
 * with an error: plane-of-sky datasets and FITS data files.
 */
#ifndef SHAPE_IO_H
#define SHAPE_IO_H

#include "shape_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes of the shape_read_* entry points. */
#define SHAPE_IO_EFILE   -1 /* a file could not be opened or read */
#define SHAPE_IO_EPARSE  -2 /* malformed contents */
#define SHAPE_IO_EUNSUP  -3 /* a valid input this library does not read */

/* read_par.c:372  read_par(name, par): the parameter file, penalties included;
   returns 0 */
int shape_read_par(const char *name, struct par_t *par);
/* read_mod.c:204  read_mod(par, mod): reads mod->name; realizes the component
   tessellations, sides and attached-facet lists; returns the number of free
   parameters (>= 0) */
int shape_read_mod(struct par_t *par, struct mod_t *mod);
/* read_dat.c:409  read_dat(par, mod, dat): reads dat->name and every data file
   it names; returns the number of free parameters (>= 0) */
int shape_read_dat(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* message of the calling thread's last failed shape_read_* call */
const char *shape_io_error(void);

/* release what a successful shape_read_* call allocated under the struct */
void shape_io_free_par(struct par_t *par);
void shape_io_free_mod(struct mod_t *mod);
void shape_io_free_dat(struct dat_t *dat);

/* The reference's entry points (exit status 2 on malformed input). */
void read_par(char *name, struct par_t *par);
int read_mod(struct par_t *par, struct mod_t *mod);
int read_dat(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* Set-up routines of read_mod.c that other reference files call
   (shape2.h:1528-1529; convex_hull.c, merge_comps.c, split_mod.c):
   setupsides (read_mod.c:2386) numbers a component's sides and each facet's
   sides; setupvertices (read_mod.c:2501) lists each vertex's attached sides
   and facets.  Their arrays are plain calloc blocks owned by the caller. */
void setupsides(struct vertices_t *vt);
void setupvertices(struct vertices_t *vt);

/* The readers' geometry, exported for checking (the reference's own
   cal2jd and ephem2mat live in other objects, which stay as they are):
   astro/jdcal.c:10  calendar date (UTC) -> Julian date;
   shape/ephem2mat.c:27  ecliptic->observer and ecliptic->sun matrices, the
   plane-of-sky spin contribution, solar phase and azimuth at epoch t from
   linearly interpolated ephemerides; returns the distance (AU), or a negative
   value (with shape_io_error set) when t lies outside an ephemeris. */
double shape_io_cal2jd(int yy, int mo, int dd, int hh, int mi, int ss);
double shape_io_ephem2mat(const struct ephem_t *ast, const struct ephem_t *sun, double t,
                          double oe[3][3], double se[3][3], double s[3], double *solar_phase,
                          double *solar_azimuth, int bistatic);

#ifdef __cplusplus
}
#endif
#endif
