// pk_env.cpp — environment lookup and importance-sampling tables (host side).
//
// Bins sampler: render/env_sampler.js:31-157 (EnvironmentGenerator).
//   radiance r = 2^(a-128) * luma(rgb) / 255 per RGBE8 texel (:43-48), or
//   luma(rgb) for float maps; stable sort by r descending (:59-72);
//   numBins = min(256, max(32, 2*ceil(total/brightest))) (:77); greedy bins
//   closing when the running sum reaches total/numBins, tested before adding
//   the texel (:79-92); texel pdf = (1/numBins) * (T/binCount) / (2 pi^2)
//   (:124-157).  Decoded lookup radiance follows shade_miss.wgsl:41-42.
// CDF sampler (new, not in the reference): texel weights w = luma * sin(theta)
//   of the texel centre; a row marginal CDF and per-row conditional CDFs,
//   searched on the device by binary search; the texel pdf table then holds
//   p_texel * T / (2 pi^2) so the shared lookup pdf(dir) = pdf[texel]/sin(theta)
//   applies to both samplers.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "pk_internal.h"
#include "pk_math.h"

int pk_build_env_tables(const pk_scene_desc* d, EnvTables* e, std::string* err) {
  const uint32_t W = d->env_width, H = d->env_height;
  if (d->env_format == PK_ENV_NONE) {
    *e = EnvTables();
    return PK_OK;
  }
  if (W == 0 || H == 0 || !d->env_data) {
    *err = "environment map has zero size or no data";
    return PK_E_INVALID;
  }
  if (d->env_format != PK_ENV_RGBE8 && d->env_format != PK_ENV_RGB32F) {
    *err = "unknown environment format";
    return PK_E_INVALID;
  }
  const size_t T = (size_t)W * H;
  if (T >= (1u << 31)) {
    *err = "environment map larger than 2^31 texels";
    return PK_E_LIMIT;
  }
  e->width = W;
  e->height = H;
  e->radiance.resize(T * 3);
  std::vector<double> rad(T);
  for (size_t i = 0; i < T; ++i) {
    if (d->env_format == PK_ENV_RGBE8) {
      const uint8_t* c = (const uint8_t*)d->env_data + i * 4;
      double power = std::pow(2.0, (double)c[3] - 128);
      double luma = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];
      rad[i] = power * luma / 255.0;
      float a = (float)c[3] / 255.0f;
      float p2 = pkm::exp2f_(a * 255.0f - 128.0f);
      for (int k = 0; k < 3; ++k) e->radiance[i * 3 + k] = ((float)c[k] / 255.0f) * p2;
    } else {
      const float* c = (const float*)d->env_data + i * 3;
      rad[i] = 0.2126 * (double)c[0] + 0.7152 * (double)c[1] + 0.0722 * (double)c[2];
      for (int k = 0; k < 3; ++k) e->radiance[i * 3 + k] = c[k];
    }
  }
  // ---- equal-radiance bins -------------------------------------------------
  double total = 0, brightest = 0;
  for (size_t i = 0; i < T; ++i) {
    brightest = std::max(rad[i], brightest);
    total += rad[i];
  }
  std::vector<uint32_t> order(T);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return rad[b] < rad[a]; });
  double ratio = total / brightest;
  // Math.min/Math.max propagate NaN (all-black map): no bin closes early.
  double bin_size = std::isnan(ratio) ? ratio : total / std::min(256.0, std::max(32.0, 2 * std::ceil(ratio)));
  e->bins.clear();
  double current = 0;
  for (size_t i = 0; i < T; ++i) {
    if (current >= bin_size) {
      int32_t prev = e->bins.empty() ? 0 : e->bins[e->bins.size() - 1];
      e->bins.push_back(prev);
      e->bins.push_back((int32_t)i);
      current = 0;
    }
    current += rad[order[i]];
  }
  int32_t last = e->bins.empty() ? 0 : e->bins[e->bins.size() - 1];
  e->bins.push_back(last);
  e->bins.push_back((int32_t)T);
  e->coords.resize(T * 2);
  for (size_t i = 0; i < T; ++i) {
    e->coords[i * 2] = (int32_t)(order[i] % W);
    e->coords[i * 2 + 1] = (int32_t)(order[i] / W);
  }
  e->pdf.assign(T, 0.0f);
  const size_t nb = e->bins.size() / 2;
  if (d->env_sampler == PK_ENV_SAMPLER_BINS) {
    for (size_t b = 0; b < nb; ++b) {
      int32_t h0 = e->bins[b * 2], h1 = e->bins[b * 2 + 1];
      for (int32_t i = h0; i < h1; ++i) {
        size_t idx = (size_t)e->coords[(size_t)i * 2] + (size_t)W * e->coords[(size_t)i * 2 + 1];
        double bin_pdf = 1 / (double)nb;
        double coord_pdf = (double)T / (double)(h1 - h0);
        double sphere_pdf = 1 / (3.141592653589793 * 3.141592653589793 * 2);
        e->pdf[idx] = (float)(bin_pdf * coord_pdf * sphere_pdf);
      }
    }
    e->row_cdf.clear();
    e->col_cdf.clear();
    e->row_guide.clear();
    e->col_guide.clear();
    return PK_OK;
  }
  // ---- marginal / conditional CDF -----------------------------------------
  std::vector<double> w(T);
  std::vector<double> row(H, 0.0);
  for (uint32_t y = 0; y < H; ++y) {
    double s = std::sin(3.141592653589793 * ((double)y + 0.5) / (double)H);
    for (uint32_t x = 0; x < W; ++x) {
      double v = rad[(size_t)y * W + x] * s;
      w[(size_t)y * W + x] = v;
      row[y] += v;
    }
  }
  double sum = std::accumulate(row.begin(), row.end(), 0.0);
  if (!(sum > 0)) {
    *err = "environment map has no positive radiance for the CDF sampler";
    return PK_E_INVALID;
  }
  e->row_cdf.assign(H + 1, 0.0f);
  e->col_cdf.assign((size_t)H * (W + 1), 0.0f);
  double acc = 0;
  for (uint32_t y = 0; y < H; ++y) {
    e->row_cdf[y] = (float)(acc / sum);
    acc += row[y];
    double ra = 0;
    for (uint32_t x = 0; x < W; ++x) {
      e->col_cdf[(size_t)y * (W + 1) + x] = row[y] > 0 ? (float)(ra / row[y]) : (float)x / (float)W;
      ra += w[(size_t)y * W + x];
    }
    e->col_cdf[(size_t)y * (W + 1) + W] = 1.0f;
  }
  e->row_cdf[H] = 1.0f;
  // Cutpoint guides (Chen & Asau 1974): for cell k of [0, 1) the answer of
  // "largest index with cdf <= u" at u = k/n.  A lookup computes k = (int)(u*n)
  // and searches only between the guides of cells k-1 and k+2, which bracket
  // the answer for every u that can round into cell k; the result equals the
  // full binary search's.
  auto guide = [](const float* cdf, uint32_t n, uint32_t* out) {
    uint32_t y = 0;
    for (uint32_t k = 0; k <= n; ++k) {
      const float t = (float)k / (float)n;
      while (y + 1 < n && cdf[y + 1] <= t) ++y;
      out[k] = y;
    }
  };
  e->row_guide.assign(H + 1, 0u);
  guide(e->row_cdf.data(), H, e->row_guide.data());
  e->col_guide.assign((size_t)H * (W + 1), 0u);
  for (uint32_t y = 0; y < H; ++y)
    guide(&e->col_cdf[(size_t)y * (W + 1)], W, &e->col_guide[(size_t)y * (W + 1)]);
  // The sampled texel's probability is the product of the two CDF steps the
  // device search will see (f32 table values), so pdf and sampling agree.
  for (uint32_t y = 0; y < H; ++y) {
    double py = (double)e->row_cdf[y + 1] - (double)e->row_cdf[y];
    for (uint32_t x = 0; x < W; ++x) {
      const float* cc = &e->col_cdf[(size_t)y * (W + 1)];
      double px = (double)cc[x + 1] - (double)cc[x];
      e->pdf[(size_t)y * W + x] = (float)(py * px * (double)T / (2 * 3.141592653589793 * 3.141592653589793));
    }
  }
  return PK_OK;
}

extern "C" int pk_env_tables(const pk_scene_desc* d, int64_t* info4, int32_t* bins2, int32_t* coords2, float* pdf) {
  return pk_guard("pk_env_tables", [&]() -> int {
    if (!d || !info4) {
      pk_set_error("pk_env_tables: null argument");
      return PK_E_INVALID;
    }
    EnvTables e;
    std::string err;
    int rc = pk_build_env_tables(d, &e, &err);
    if (rc) {
      pk_set_error("pk_env_tables: " + err);
      return rc;
    }
    info4[0] = (int64_t)(e.bins.size() / 2);
    info4[1] = (int64_t)(e.coords.size() / 2);
    info4[2] = e.width;
    info4[3] = e.height;
    if (bins2) std::copy(e.bins.begin(), e.bins.end(), bins2);
    if (coords2) std::copy(e.coords.begin(), e.coords.end(), coords2);
    if (pdf) std::copy(e.pdf.begin(), e.pdf.end(), pdf);
    return PK_OK;
  });
}
