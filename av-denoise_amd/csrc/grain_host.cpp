// Host side of the grain table (nl4d/grain/fit.rs): the std bucket edges, the histogram median,
// AV1 scaling points and the AR weights' quantisation.  Plain f64 arithmetic in the reference's
// order; these run once per grain segment on a few hundred numbers, so they stay on the host as
// they do in the reference.

#include <cmath>
#include <cstdint>

#include "avd/avd.h"

namespace {

constexpr int kBuckets = 64;
constexpr float kStdMin = 0.05f / 255.0f, kStdMax = 32.0f / 255.0f;

// raw_scaling (fit.rs:323-326): the scaling value one sigma needs before the scaling shift.
double raw_scaling(double sigma, double sigma_template) { return sigma * 255.0 / sigma_template; }

}  // namespace

extern "C" {

void avd_grain_bucket_edges(float *edges)
{
    if (!edges) return;
    const double ratio = (double)kStdMax / (double)kStdMin;
    for (int i = 0; i <= kBuckets; ++i) {
        const double fraction = (double)i / (double)kBuckets;
        edges[i] = (float)((double)kStdMin * std::pow(ratio, fraction));
    }
}

double avd_grain_hist_median(const uint32_t *counts, const float *edges, int32_t *found)
{
    if (found) *found = 0;
    if (!counts || !edges) return 0.0;
    uint64_t total = 0;
    for (int b = 0; b < kBuckets; ++b) total += counts[b];
    if (total == 0) return 0.0;
    const double target = 0.5 * (double)total;
    double below = 0.0;
    for (int b = 0; b < kBuckets; ++b) {
        const double count = (double)counts[b];
        if (count > 0.0 && below + count >= target) {
            const double fraction = (target - below) / count;
            const double low = (double)edges[b], high = (double)edges[b + 1];
            if (found) *found = 1;
            return low * std::pow(high / low, fraction);
        }
        below += count;
    }
    return 0.0;
}

uint32_t avd_grain_finest_scaling_shift(double peak)
{
    for (uint32_t shift : {11u, 10u, 9u})
        if (peak * (double)(1u << shift) <= 255.0) return shift;
    return 8;
}

avd_status avd_grain_scaling_points(const uint8_t *luma, const double *sigma, uint32_t count, double sigma_template,
                                    uint8_t *scaled, uint32_t *shift)
{
    if ((count && (!luma || !sigma || !scaled)) || !shift) return AVD_ERR_NULL_ARGUMENT;
    (void)luma;  // the points keep their luma; only the scaling values are computed
    double peak = 0.0;
    for (uint32_t i = 0; i < count; ++i) peak = std::fmax(peak, raw_scaling(sigma[i], sigma_template));
    const uint32_t s = avd_grain_finest_scaling_shift(peak);
    const double factor = (double)(1u << s);
    for (uint32_t i = 0; i < count; ++i) {
        const double v = std::fmin(std::round(raw_scaling(sigma[i], sigma_template) * factor), 255.0);
        scaled[i] = v > 0.0 ? (uint8_t)v : 0;  // Rust's saturating cast: negative and NaN give 0
    }
    *shift = s;
    return AVD_OK;
}

int32_t avd_grain_pooled_record(const uint32_t *source_hist, const double *group_records, const float *edges,
                                double *record)
{
    if (!source_hist || !group_records || !edges || !record) return 0;
    // overall_median (segment.rs:170-180): the 16 brightness rows merged, then their median.
    uint32_t merged[kBuckets] = {};
    for (int bin = 0; bin < 16; ++bin)
        for (int b = 0; b < kBuckets; ++b) merged[b] += source_hist[bin * kBuckets + b];
    int32_t found = 0;
    const double median = avd_grain_hist_median(merged, edges, &found);
    if (!found) return 0;
    const double band_low = 0.5 * median, band_high = 2.0 * median;
    for (int l = 0; l < 47; ++l) record[l] = 0.0;
    for (int g = 0; g < 16; ++g) {
        const double group_low = (double)edges[g * 4], group_high = (double)edges[(g + 1) * 4];
        if (group_high < band_low || group_low > band_high) continue;
        for (int l = 0; l < 47; ++l) record[l] += group_records[g * 47 + l];
    }
    return record[46] < 50000.0 ? 0 : 1;
}

void avd_grain_take_out_fine_noise(double *corrected, double floor, double fine_noise)
{
    if (!corrected) return;
    const double share = std::fmin(fine_noise, 0.95) * floor;
    corrected[0] -= share * corrected[0];
}

double avd_grain_floor_fraction(const double *weights)
{
    if (!weights) return 0.0;
    // ar_spectrum / ar_power (fit.rs:236-270): unit noise through the AR filter on a 64 x 64 grid
    // of frequencies, row by row; AR_OFFSETS order dy -3..0, dx -3..3, the causal half.
    int oy[24], ox[24], k = 0;
    for (int dy = -3; dy <= 0; ++dy)
        for (int dx = -3; dx <= 3; ++dx)
            if (dy < 0 || dx < 0) {
                oy[k] = dy;
                ox[k] = dx;
                ++k;
            }
    const double step = 6.283185307179586 / 64.0;
    double lowest = INFINITY, total = 0.0;
    for (int row = 0; row < 64; ++row)
        for (int col = 0; col < 64; ++col) {
            const double fy = (double)row * step, fx = (double)col * step;
            double re = 1.0, im = 0.0;
            for (int i = 0; i < 24; ++i) {
                const double phase = fy * (double)oy[i] + fx * (double)ox[i];
                re -= weights[i] * std::cos(phase);
                im -= weights[i] * std::sin(phase);
            }
            const double power = 1.0 / (re * re + im * im);
            lowest = std::fmin(lowest, power);
            total += power;
        }
    const double fraction = lowest / (total / 4096.0);
    if (!std::isfinite(fraction)) return 0.0;
    return std::fmin(std::fmax(fraction, 0.0), 1.0);
}

void avd_grain_quantise_ar(const double *weights, int32_t *quantised, uint32_t *shift)
{
    if (!weights || !quantised || !shift) return;
    uint32_t s = 6;
    for (uint32_t cand : {9u, 8u, 7u}) {
        bool fits = true;
        for (int i = 0; i < 24; ++i) {
            const double v = std::round(weights[i] * (double)(1u << cand));
            fits = fits && v >= -128.0 && v <= 127.0;
        }
        if (fits) {
            s = cand;
            break;
        }
    }
    for (int i = 0; i < 24; ++i)
        quantised[i] = (int32_t)std::fmin(std::fmax(std::round(weights[i] * (double)(1u << s)), -128.0), 127.0);
    *shift = s;
}

}  // extern "C"
