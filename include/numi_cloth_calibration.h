#ifndef NUMI_CLOTH_CALIBRATION_H
#define NUMI_CLOTH_CALIBRATION_H

/* Sensitivity columns for the calibration fitter.
 *
 * The reference's fitter (tools/cloth_calibration.cpp) reads one observation
 * CSV row per (trial, observable): measured value, sigma, the solver's
 * baseline prediction y0 and one column dlog_<parameter> = dy/dx_j, the
 * derivative of the prediction when parameter j is scaled by e^{x_j}
 * (loadObservations :214-300; linearised solve :568-650). Its protocol
 * takes every column as a central difference of two whole replays,
 * (y(theta e^h) - y(theta e^-h)) / 2h with h = 0.05
 * (docs/CLOTH_SPECIMEN_CALIBRATION.md:60-70), two runs per parameter and
 * observable. Here one adjoint run of a specimen batch
 * (numi_cloth_hip_batch_adjoint with NUMI_CLOTH_ADJOINT_LINEAR, the
 * observable's weights as targets) gives every column of every specimen's
 * observable at once. */

#include <stdint.h>

#include "numi_cloth_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NUMI_CLOTH_CALIBRATION_NO_COLUMN 0xffffffffu

/* Which fitter column each parameter of a specimen belongs to, in the
 * specimen's own record order (NULL: none of that family; an entry of
 * NUMI_CLOTH_CALIBRATION_NO_COLUMN: not fitted). */
typedef struct NumiClothCalibrationColumns {
    uint32_t columnCount;
    uint32_t reserved;
    const uint32_t* distanceColumn; /* per distance record: its compliance */
    const uint32_t* bendColumn;     /* per bend record: its compliance */
    const uint32_t* massColumn;     /* per knot: its mass (knots with inverse mass > 0 only) */
    const uint32_t* dampingColumn;  /* per knot: its damping coefficient */
} NumiClothCalibrationColumns;

/* dlog_j for specimen `specimen` of the batch's last adjoint run, whose own
 * scene (as passed to numi_cloth_scene_merge) is `scene`:
 *   compliance c (distance, bend):  sum of c * dy/dc
 *   mass m = 1 / w:                 sum of m * dy/dm = -sum of w * dy/dw
 *   damping coefficient c:          sum of c * dy/dc
 * over the parameters assigned to column j, with the specimen's parameter
 * values from `scene` (damping: extension.particleMotion[k].w, else
 * extension.damping). *baseline receives y (the specimen's loss). */
int numi_cloth_calibration_dlog(NumiClothHipBatch* batch, uint32_t specimen, const NumiClothHipSceneDesc* scene,
                                const NumiClothCalibrationColumns* columns, double* dlog, double* baseline);

typedef struct NumiClothCalibrationRow {
    const char* trialId;
    uint32_t heldout;      /* 0: split "calibration", 1: "heldout" */
    uint32_t reserved;
    const char* observable;
    double measured;
    double sigma;
    double baseline;
    const double* dlog;    /* columnCount values, in parameterNames order */
} NumiClothCalibrationRow;

/* Writes the fitter's observation CSV: the header
 * trial_id,split,observable,measured,sigma,baseline,dlog_<name>... and one
 * line per row, numbers with 17 significant digits (they read back as the
 * same doubles). Rejects (NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, nothing
 * written) what the fitter would reject: empty or duplicate names, empty
 * trial ids or observables, a trial in both splits, a repeated
 * (trial, observable), sigma <= 0 or a non-finite number; and names or ids
 * holding ',', '"', '#' or a line break, which its CSV reader would split
 * or drop. */
int numi_cloth_calibration_write(const char* path, uint32_t columnCount, const char* const* parameterNames,
                                 const NumiClothCalibrationRow* rows, uint32_t rowCount);

#ifdef __cplusplus
}
#endif

#endif
