/*
 * shape_hip.h -- C ABI of libshape_hip.so, the MI355X (gfx950) implementation
 * of SHAPE's model-evaluation path.
 *
 * Part 1 keeps the reference's entry points, names and signatures, so the
 * library links in place of the reference's routines in existing SHAPE host
 * code (prototypes in the reference's src/shape/shape2.h:1443-1599).  Every
 * computation runs in HIP kernels; the caller's structs are read and written
 * exactly where the reference reads and writes them.  The reference reports
 * nothing for GPU failures, so these entries record the failure in
 * shape_hip_last_error() and leave their outputs untouched.
 *
 * Part 2 is the batched evaluation interface (no counterpart in the
 * reference): many trial parameter values, datasets and models evaluated in
 * one device-resident launch sequence replayed as a hipGraph.
 */
#ifndef SHAPE_HIP_H
#define SHAPE_HIP_H

#include "shape_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Part 1: reference entry points ---------------- */

/* realize_mod.c:153  vertex coordinates, facet normals/areas/centroids and
   vertex normals of the model's (single, VERTEX-type) component; sets
   par->baddiam and par->baddiam_logfactor. */
void realize_mod(struct par_t *par, struct mod_t *mod);

/* realize_spin.c:41  Euler matrices ae, intrinsic and total spin vectors of
   every frame view and lightcurve point (principal-axis rotators). */
void realize_spin(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* realize_delcor.c:50  delay / Doppler COM offsets of every view from the
   delay-correction polynomials (with the delcor0 modes and '=' states). */
void realize_delcor(struct dat_t *dat, double delta_delcor0, int delcor0_mode);

/* calc_fits.c:340  model fits of all delay-Doppler, Doppler and lightcurve
   datasets; writes fit images, per-frame limits, overflow terms, facet
   "seen" flags and the posbnd / badradar flags and log factors in par.
   POS arrays are written back only after shape_hip_set_strict(1). */
void calc_fits(struct par_t *par, struct mod_t *mod, struct dat_t *dat);

/* chi2.c:240  chi-square of the current fits; floats calibration factors,
   writes frame->chi2, cal.val, set chi2 and dat->chi2, returns dat->chi2. */
double chi2(struct par_t *par, struct dat_t *dat, int list_breakdown);

/* ---------------- library control ---------------- */
int shape_hip_last_error(void);       /* 0, or the code of the last failure */
void shape_hip_clear_error(void);
void shape_hip_set_strict(int write_pos); /* 1: calc_fits also writes POS arrays back */
int shape_hip_device_count(void);
void shape_hip_realize_delcor(struct par_t *par, struct mod_t *mod, struct dat_t *dat,
                              double delta_delcor0, int delcor0_mode);

#ifdef __cplusplus
}
#endif
#endif
