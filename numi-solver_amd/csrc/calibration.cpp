// Sensitivity columns for the calibration fitter and its observation CSV
// (declared in include/numi_cloth_calibration.h). Host code: the gradients
// come from the batch adjoint through the C ABI.

#include "numi_cloth_calibration.h"

#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "common.hpp"

namespace numi_hip {
namespace {

void addColumn(const uint32_t* column, uint32_t k, uint32_t columnCount, double value, double* dlog) {
    if (column == nullptr) return;
    const uint32_t j = column[k];
    if (j == NUMI_CLOTH_CALIBRATION_NO_COLUMN) return;
    if (j >= columnCount) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "calibration column out of range");
    dlog[j] += value;
}

void dlogColumns(NumiClothHipBatch* batch, uint32_t specimen, const NumiClothHipSceneDesc& scene,
                 const NumiClothCalibrationColumns& columns, double* dlog, double* baseline) {
    uint32_t counts[3];
    batchSpecimenCounts(*batch, specimen, counts);
    if (counts[0] != scene.particleCount || counts[1] != scene.distanceCount || counts[2] != scene.bendCount)
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "scene does not match the batch specimen's record counts");
    if (scene.particleCount != 0 && scene.particles == nullptr)
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "scene has no particles");
    std::vector<float> gd(scene.distanceCount), gb(scene.bendCount), gw(scene.particleCount),
        gc(scene.particleCount);
    double loss = 0.0;
    NumiClothHipSpecimenGradients g{&loss, gd.data(), gb.data(), gw.data(), gc.data()};
    const int status = numi_cloth_hip_batch_gradients(batch, specimen, &g);
    if (status != NUMI_CLOTH_HIP_OK) fail(status, numi_cloth_hip_last_error());
    for (uint32_t j = 0; j < columns.columnCount; ++j) dlog[j] = 0.0;
    const uint32_t n = columns.columnCount;
    for (uint32_t k = 0; k < scene.distanceCount; ++k)
        addColumn(columns.distanceColumn, k, n, double(scene.distances[k].material.y) * gd[k], dlog);
    for (uint32_t k = 0; k < scene.bendCount; ++k)
        addColumn(columns.bendColumn, k, n, double(scene.bends[k].material.z) * gb[k], dlog);
    const NumiClothHipSceneExtension& ext = scene.extension;
    for (uint32_t k = 0; k < scene.particleCount; ++k) {
        const double w = scene.particles[k].positionAndInverseMass.w;
        if (w > 0.0) addColumn(columns.massColumn, k, n, -w * gw[k], dlog);  // m dy/dm = -w dy/dw
        const double c = ext.particleMotion != nullptr ? ext.particleMotion[k].w : ext.damping;
        addColumn(columns.dampingColumn, k, n, c * gc[k], dlog);
    }
    if (baseline != nullptr) *baseline = loss;
}

// What the fitter's CSV reader (reference tools/cloth_calibration.cpp:64-150)
// would split, unquote, trim or drop as a comment.
bool plainField(const char* text) {
    if (text == nullptr || *text == '\0') return false;
    const std::string s(text);
    if (s.find_first_of(",\"#\r\n") != std::string::npos) return false;
    return s.front() != ' ' && s.front() != '\t' && s.back() != ' ' && s.back() != '\t';
}

void writeRows(const char* path, uint32_t columnCount, const char* const* names, const NumiClothCalibrationRow* rows,
               uint32_t rowCount) {
    if (path == nullptr || (columnCount != 0 && names == nullptr) || (rowCount != 0 && rows == nullptr))
        fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
    std::set<std::string> unique;
    for (uint32_t j = 0; j < columnCount; ++j) {
        if (!plainField(names[j])) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "parameter names must be plain CSV fields");
        if (!unique.insert(names[j]).second) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "duplicate parameter name");
    }
    std::map<std::string, uint32_t> splits;
    std::set<std::pair<std::string, std::string>> pairs;
    for (uint32_t r = 0; r < rowCount; ++r) {
        const NumiClothCalibrationRow& row = rows[r];
        if (!plainField(row.trialId) || !plainField(row.observable))
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "trial ids and observables must be plain CSV fields");
        if (row.heldout > 1u) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "heldout must be 0 or 1");
        const auto [split, inserted] = splits.emplace(row.trialId, row.heldout);
        if (!inserted && split->second != row.heldout)
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "a trial may not cross the calibration and heldout splits");
        if (!pairs.emplace(row.trialId, row.observable).second)
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "duplicate trial and observable pair");
        if (!(row.sigma > 0.0) || !std::isfinite(row.sigma) || !std::isfinite(row.measured) ||
            !std::isfinite(row.baseline))
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "measured, sigma and baseline must be finite, sigma > 0");
        if (columnCount != 0 && row.dlog == nullptr) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "row without dlog");
        for (uint32_t j = 0; j < columnCount; ++j)
            if (!std::isfinite(row.dlog[j])) fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "non-finite sensitivity");
    }
    std::string text = "trial_id,split,observable,measured,sigma,baseline";
    for (uint32_t j = 0; j < columnCount; ++j) text += std::string(",dlog_") + names[j];
    text += '\n';
    char number[40];
    auto append = [&](double v) {
        std::snprintf(number, sizeof(number), ",%.17g", v);
        text += number;
    };
    for (uint32_t r = 0; r < rowCount; ++r) {
        const NumiClothCalibrationRow& row = rows[r];
        text += row.trialId;
        text += row.heldout ? ",heldout," : ",calibration,";
        text += row.observable;
        append(row.measured);
        append(row.sigma);
        append(row.baseline);
        for (uint32_t j = 0; j < columnCount; ++j) append(row.dlog[j]);
        text += '\n';
    }
    std::FILE* file = std::fopen(path, "wb");
    if (file == nullptr) fail(NUMI_CLOTH_HIP_ERROR_IO, std::string("cannot create ") + path);
    const bool written = std::fwrite(text.data(), 1, text.size(), file) == text.size();
    if (std::fclose(file) != 0 || !written) fail(NUMI_CLOTH_HIP_ERROR_IO, std::string("cannot write ") + path);
}

}  // namespace
}  // namespace numi_hip

using namespace numi_hip;

extern "C" {

int numi_cloth_calibration_dlog(NumiClothHipBatch* batch, uint32_t specimen, const NumiClothHipSceneDesc* scene,
                                const NumiClothCalibrationColumns* columns, double* dlog, double* baseline) {
    return guarded([&] {
        if (batch == nullptr || scene == nullptr || columns == nullptr || (columns->columnCount != 0 && dlog == nullptr))
            fail(NUMI_CLOTH_HIP_ERROR_INVALID_ARGUMENT, "null argument");
        dlogColumns(batch, specimen, *scene, *columns, dlog, baseline);
    });
}

int numi_cloth_calibration_write(const char* path, uint32_t columnCount, const char* const* parameterNames,
                                 const NumiClothCalibrationRow* rows, uint32_t rowCount) {
    return guarded([&] { writeRows(path, columnCount, parameterNames, rows, rowCount); });
}

}  // extern "C"
