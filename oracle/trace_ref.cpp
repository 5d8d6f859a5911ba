// ORACLE — test infrastructure only. Restatement of the reference's sphere
// tracer (src/prtcl/util/sphere_tracer.hpp:46-268) and pinhole camera
// (src/prtcl/geometry/pinhole_camera.hpp:9-66) in plain C++, built into
// oracle/_build/libprtcl_oracle.so and used by tests/ as the checker of the
// device tracer (prtcl_hip_scene_trace).
//
// Real is double (PinholeCamera::Real). The occupied-cell grid of ungrd
// (external/ungrd, an empty submodule here) is restated as a map from the
// integer cell floor(x / grid_diameter) to the particles in it, in input
// order; the reference only iterates its occupied cells
// (foreach_position) and the entries of one cell (foreach_entry_at_position),
// so the order of entries inside a cell is the one choice made here: input
// order, which the device path uses too. Vector sums of three terms follow
// Eigen's fixed-size reduction, x0 + (x1 + x2).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <vector>

#include "cpu_ref.h"

namespace {

using V3 = double[3];

double dot3(const double *a, const double *b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }
double norm3(const double *a) { return std::sqrt(dot3(a, a)); }

// Eigen normalized(): v / sqrt(squaredNorm) when the norm is positive
void normalized(const double *a, double *out) {
  const double z = dot3(a, a);
  for (int k = 0; k < 3; ++k)
    out[k] = z > 0 ? a[k] / std::sqrt(z) : a[k];
}

double cpow(double x, int n) {  // util/constpow.hpp:9-18
  double r = 1;
  for (; n > 0; --n)
    r *= x;
  return r;
}

// cubic_spline_kernel.hpp:43-78 with kernel_facade.hpp:95-124, N = 3
struct Spline {
  double alpha;
  double evalq(double q) const {
    double r = 0;
    if (q < 1)
      r -= 4 * cpow(1 - q, 3);
    if (q < 2)
      r += cpow(2 - q, 3);
    return alpha * r;
  }
  double evaldq(double q) const {
    double r = 0;
    if (q < 1)
      r += 12 * cpow(1 - q, 2);
    if (q < 2)
      r -= 3 * cpow(2 - q, 2);
    return alpha * r;
  }
  double evalr(double r, double h) const { return evalq(std::abs(r) / h) / cpow(h, 3); }
  double evaldr(double r, double h) const {
    return -std::copysign(evaldq(std::abs(r) / h) / cpow(h, 4), r);
  }
  void evalgrad(const double *v, double h, double *out) const {
    const double r = norm3(v);
    if (r < std::sqrt(std::numeric_limits<double>::epsilon())) {
      out[0] = out[1] = out[2] = 0;
      return;
    }
    const double s = evaldr(r, h) / r;
    for (int k = 0; k < 3; ++k)
      out[k] = s * v[k];
  }
  double lipschitz(double h) const { return -evaldq(2.0 / 3.0) / cpow(h, 4); }
};

}  // namespace

extern "C" {

double oracle_spline_alpha3() {
  // cubic_spline_kernel.hpp:98: 1/4 * (1/pi) in long double, then real
  return static_cast<double>(1.L / 4.L * (1.L / 3.141592653589793238462643383279502884L));
}

int oracle_sphere_trace(const double *x, int64_t n, double smoothing_scale,
                        const oracle_camera *cam, double threshold, int64_t max_steps,
                        double *image, double *parameter) {
  if (!cam || !image || cam->width <= 0 || cam->height <= 0 || n < 0)
    return 1;
  const double gd = 4 * smoothing_scale;  // sphere_tracer.hpp:74-75
  const double half = gd / 2;
  Spline W{oracle_spline_alpha3()};

  // occupied cells and their entries in input order (sphere_tracer.hpp:87-100)
  std::map<std::vector<int32_t>, std::vector<int64_t>> cells;
  for (int64_t i = 0; i < n; ++i) {
    std::vector<int32_t> c(3);
    for (int k = 0; k < 3; ++k)
      c[k] = static_cast<int32_t>(std::floor(x[3 * i + k] / gd));
    cells[c].push_back(i);
  }
  std::vector<double> centers;
  for (const auto &kv : cells)
    for (int k = 0; k < 3; ++k)
      centers.push_back(gd * kv.first[k] + half);
  const size_t nc = cells.size();

  // pinhole_camera.hpp:33-65
  double v[3], p[3], hh[3], vp[3];
  normalized(cam->up, v);
  normalized(cam->principal, p);
  vp[0] = v[1] * p[2] - v[2] * p[1];
  vp[1] = v[2] * p[0] - v[0] * p[2];
  vp[2] = v[0] * p[1] - v[1] * p[0];
  normalized(vp, hh);
  const double pixel_size = 1.0 / static_cast<double>(cam->width);
  const double sox = -static_cast<double>(cam->width - 1) / 2;
  const double soy = -static_cast<double>(cam->height - 1) / 2;
  const double *O = cam->origin;
  const double f = cam->focal_length;

  const double max_parameter = gd * 10000;
  const double phi0 = threshold * W.evalr(0, half);
  const double L = W.lipschitz(half);

  for (int64_t iy = 0; iy < cam->height; ++iy)
    for (int64_t ix = 0; ix < cam->width; ++ix) {
      const double px = (sox + static_cast<double>(ix)) * pixel_size;
      const double py = (soy + static_cast<double>(iy)) * pixel_size;
      double pp[3], dir0[3], d[3];
      for (int k = 0; k < 3; ++k) {
        pp[k] = ((O[k] + px * hh[k]) + py * v[k]) - f * p[k];
        dir0[k] = O[k] - pp[k];
      }
      normalized(dir0, d);

      double t = 0, normal[3] = {0, 0, 0}, sdf = 0;
      int64_t steps = 0;
      for (; steps < max_steps; ++steps) {
        double rx[3];
        for (int k = 0; k < 3; ++k)
          rx[k] = O[k] + t * d[k];
        sdf = std::numeric_limits<double>::max();
        for (size_t c = 0; c < nc; ++c) {
          double q[3], qm[3];
          for (int k = 0; k < 3; ++k) {
            q[k] = std::abs(rx[k] - centers[3 * c + k]) - gd;
            qm[k] = std::max(q[k], 0.0);
          }
          const double mc = std::max(q[0], std::max(q[1], q[2]));
          const double l = norm3(qm) + std::min(mc, 0.0) - gd;
          sdf = std::min(sdf, l);
        }
        if (sdf < gd) {
          int32_t rc[3];
          for (int k = 0; k < 3; ++k)
            rc[k] = static_cast<int32_t>(std::floor(rx[k] / gd));
          double phi = phi0, grad[3] = {0, 0, 0};
          int64_t count = 0;
          for (int cx = -1; cx <= 1; ++cx)
            for (int cy = -1; cy <= 1; ++cy)
              for (int cz = -1; cz <= 1; ++cz) {
                auto it = cells.find({rc[0] + cx, rc[1] + cy, rc[2] + cz});
                if (it == cells.end())
                  continue;
                for (int64_t e : it->second) {
                  double dv[3];
                  for (int k = 0; k < 3; ++k)
                    dv[k] = rx[k] - x[3 * e + k];
                  const double w = W.evalr(norm3(dv), half);
                  if (w > 0) {
                    double g[3];
                    W.evalgrad(dv, half, g);
                    phi -= w;
                    for (int k = 0; k < 3; ++k)
                      grad[k] -= g[k];
                    ++count;
                  }
                }
              }
          if (count > 0) {
            const double s = L * static_cast<double>(count);
            phi /= s;
            for (int k = 0; k < 3; ++k)
              normal[k] = grad[k] / s;
            sdf = std::max(sdf, phi);
          } else {
            sdf = std::max(sdf, gd / 4);
          }
        }
        if (sdf < 1e-6)
          break;
        t += sdf;
        if (t >= max_parameter)
          break;
      }
      if (steps >= max_steps)
        t = max_parameter;
      double pix;
      if (t == max_parameter) {
        pix = 0;
      } else if (t > max_parameter) {
        pix = 0.3;
      } else {
        double dn[3], nn[3];
        normalized(d, dn);
        normalized(normal, nn);
        pix = -dot3(dn, nn);
      }
      image[iy * cam->width + ix] = pix;
      if (parameter)
        parameter[iy * cam->width + ix] = t;
    }
  return 0;
}

}  // extern "C"
