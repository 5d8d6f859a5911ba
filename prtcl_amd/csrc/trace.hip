// Sphere tracer of a scene group on the device, the algorithm of the
// reference's SphereTracer::Trace (util/sphere_tracer.hpp:46-268):
//
// * cells of edge grid_diameter = 4 h (floor(x / grid_diameter)), in a dense
//   box around the group; a counting sort puts every particle in its cell,
//   ordered by host index inside a cell, and the occupied cells are
//   compacted into a list of cell centres;
// * one thread per pixel marches its ray (pinhole camera,
//   geometry/pinhole_camera.hpp:33-65). At every step the distance bound is
//   the minimum over all occupied cells of the rounded-box distance
//   |max(q, 0)| + min(max_k q_k, 0) - grid_diameter, q = |x - centre| -
//   grid_diameter. The reference evaluates every cell; here cells are
//   grouped by 4^3, a group's box bounds the distances of its cells from
//   below, and groups that cannot lower the minimum are skipped, so the
//   minimum is the same value at a fraction of the work. Below
//   grid_diameter the implicit function of the particles
//   in the 27 cells around the ray point (cubic spline with h =
//   grid_diameter / 2, divided by its Lipschitz constant and the particle
//   count) tightens the bound and gives the normal;
// * the pixel is 0 when the march ran out of steps, 0.3 when the ray left
//   (parameter beyond 10000 grid diameters), else -dot(ray, normal).
//
// Double precision throughout (PinholeCamera::Real), no contraction of
// multiply-adds, and every expression in the reference's order, so the image
// is that of the same arithmetic on the CPU.
#include "trace.hpp"

#include <cfloat>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

namespace prtcl_amd {
namespace {

constexpr int kTB = 256;

struct CellBox {
  int32_t ox, oy, oz, nx, ny, nz;
  double gd;
};

struct TraceParams {
  double O[3], v[3], p[3], hh[3];
  double f, pixel_size, sox, soy;
  double gd, half, phi0, L, alpha, h3, h4, grad_eps, max_parameter;
  int64_t width, height, max_steps;
};

__host__ __device__ inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

__host__ __device__ inline void normalized3(const double *a, double *out) {
  const double z = dot3(a, a);
  for (int k = 0; k < 3; ++k)
    out[k] = z > 0 ? a[k] / sqrt(z) : a[k];
}

__host__ __device__ inline double cpow(double x, int n) {
  double r = 1;
  for (; n > 0; --n)
    r *= x;
  return r;
}

// cubic spline, N = 3 (cubic_spline_kernel.hpp:43-78)
__host__ __device__ inline double evalq(double alpha, double q) {
  double r = 0;
  if (q < 1)
    r -= 4 * cpow(1 - q, 3);
  if (q < 2)
    r += cpow(2 - q, 3);
  return alpha * r;
}

__host__ __device__ inline double evaldq(double alpha, double q) {
  double r = 0;
  if (q < 1)
    r += 12 * cpow(1 - q, 2);
  if (q < 2)
    r -= 3 * cpow(2 - q, 2);
  return alpha * r;
}

// std::min / std::max of the reference, NaN behaviour included
__device__ inline double std_min(double a, double b) { return (b < a) ? b : a; }
__device__ inline double std_max(double a, double b) { return (a < b) ? b : a; }

__device__ inline bool cell_index(const CellBox &B, int32_t cx, int32_t cy, int32_t cz,
                                  int64_t *lin) {
  const int64_t x = static_cast<int64_t>(cx) - B.ox, y = static_cast<int64_t>(cy) - B.oy,
                z = static_cast<int64_t>(cz) - B.oz;
  if (x < 0 || y < 0 || z < 0 || x >= B.nx || y >= B.ny || z >= B.nz)
    return false;
  *lin = (z * B.ny + y) * B.nx + x;
  return true;
}

template <class T>
__global__ void k_tr_to_double(const T *src, double *dst, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (i < n)
    dst[i] = static_cast<double>(src[i]);
}

// per block: min x, y, z and max x, y, z of its particles
__global__ __launch_bounds__(kTB) void k_tr_bbox(const double *x, int64_t n, double *part) {
  __shared__ double s[6][kTB];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  for (int k = 0; k < 3; ++k) {
    s[k][threadIdx.x] = i < n ? x[3 * i + k] : DBL_MAX;
    s[3 + k][threadIdx.x] = i < n ? x[3 * i + k] : -DBL_MAX;
  }
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        s[3 + k][threadIdx.x] = fmax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6)
    part[6 * static_cast<int64_t>(blockIdx.x) + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_tr_count(const double *x, int64_t n, CellBox B, uint32_t *count,
                           int64_t *lin_of) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (i >= n)
    return;
  int64_t lin = -1;
  const double *p = x + 3 * i;
  if (isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2])) {
    // inside the box, which was planned from the finite positions
    const int32_t cx = static_cast<int32_t>(floor(p[0] / B.gd));
    const int32_t cy = static_cast<int32_t>(floor(p[1] / B.gd));
    const int32_t cz = static_cast<int32_t>(floor(p[2] / B.gd));
    if (!cell_index(B, cx, cy, cz, &lin))
      lin = -1;
  }
  if (lin >= 0)
    atomicAdd(&count[lin], 1u);
  lin_of[i] = lin;  // -1: not placed (a non-finite position)
}

// exclusive scan of uint32 counts into uint32 starts, in tiles of kTB
// values per block (block sums scanned by one block, then added back)
__device__ uint32_t tr_block_scan(uint32_t v, uint32_t *total) {
  __shared__ uint32_t wsum[kTB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = v;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d)
      x += y;
  }
  if (lane == 63)
    wsum[wv] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int k = 0; k < kTB / 64; ++k) {
    if (k < wv)
      before += wsum[k];
    all += wsum[k];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

__global__ __launch_bounds__(kTB) void k_tr_scan_tiles(const uint32_t *in, int64_t n,
                                                       uint32_t *out, uint32_t *bsum) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  uint32_t tot;
  const uint32_t ex = tr_block_scan(i < n ? in[i] : 0u, &tot);
  if (i < n)
    out[i] = ex;
  if (threadIdx.x == 0)
    bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kTB) void k_tr_scan_top(uint32_t *bsum, int64_t nb) {
  uint32_t carry = 0;
  for (int64_t base = 0; base < nb; base += kTB) {
    const int64_t i = base + threadIdx.x;
    uint32_t tot;
    const uint32_t ex = tr_block_scan(i < nb ? bsum[i] : 0u, &tot);
    if (i < nb)
      bsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0)
    bsum[nb] = carry;
}

__global__ void k_tr_scan_add(uint32_t *out, int64_t n, const uint32_t *bsum) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (i < n)
    out[i] += bsum[blockIdx.x];
  if (i == n)
    out[n] = bsum[gridDim.x];  // the total after the last value
}

__global__ void k_tr_place(const int64_t *lin_of, int64_t n, const uint32_t *start,
                           uint32_t *cursor, int64_t *ent) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (i >= n)
    return;
  const int64_t lin = lin_of[i];
  if (lin >= 0)
    ent[start[lin] + atomicAdd(&cursor[lin], 1u)] = i;
}

// host order inside every cell (insertion sort of the cell's entries)
__global__ void k_tr_sort_cells(const uint32_t *start, int64_t ncells, int64_t *ent) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (c >= ncells)
    return;
  const uint32_t b = start[c], e = start[c + 1];
  for (uint32_t k = b + 1; k < e; ++k) {
    const int64_t v = ent[k];
    uint32_t j = k;
    while (j > b && ent[j - 1] > v) {
      ent[j] = ent[j - 1];
      --j;
    }
    ent[j] = v;
  }
}

__global__ void k_tr_pack(const double *x, const int64_t *ent, int64_t n, double *ex) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (k >= n)
    return;
  const int64_t i = ent[k];
  for (int d = 0; d < 3; ++d)
    ex[3 * k + d] = x[3 * i + d];
}

__global__ void k_tr_flags(const uint32_t *count, int64_t ncells, uint32_t *flag) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (c < ncells)
    flag[c] = count[c] > 0 ? 1u : 0u;
}

// centre of every occupied cell, in cell order (sphere_tracer.hpp:125-132)
// cells are grouped by kGroup^3 for the distance bound of the march
constexpr int kGroup = 4;

__global__ void k_tr_centres(const uint32_t *count, const uint32_t *occ, int64_t ncells,
                             CellBox B, double half, double *centres, uint32_t *gidx) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (c >= ncells || count[c] == 0)
    return;
  const int64_t x = c % B.nx, y = (c / B.nx) % B.ny, z = c / (static_cast<int64_t>(B.nx) * B.ny);
  const int64_t cp[3] = {x + B.ox, y + B.oy, z + B.oz};
  const int64_t k0 = occ[c];
  for (int k = 0; k < 3; ++k)
    centres[3 * k0 + k] = B.gd * static_cast<double>(cp[k]) + half;
  const int64_t gx = (B.nx + kGroup - 1) / kGroup, gy = (B.ny + kGroup - 1) / kGroup;
  gidx[k0] = static_cast<uint32_t>(((z / kGroup) * gy + y / kGroup) * gx + x / kGroup);
}

__global__ void k_tr_gcount(const uint32_t *gidx, int64_t nocc, uint32_t *gcount) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (k < nocc)
    atomicAdd(&gcount[gidx[k]], 1u);
}

// cell centres grouped (any order inside a group: only their minimum is used)
__global__ void k_tr_gplace(const uint32_t *gidx, const double *centres, int64_t nocc,
                            const uint32_t *gstart, uint32_t *gcursor, double *gcent) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (k >= nocc)
    return;
  const uint32_t g = gidx[k];
  const int64_t pos = gstart[g] + atomicAdd(&gcursor[g], 1u);
  for (int d = 0; d < 3; ++d)
    gcent[3 * pos + d] = centres[3 * k + d];
}

// the non-empty groups: box (centre, half extents) containing the boxes of
// half extent gd around the centres of all its cells, and the cell range
__global__ void k_tr_glist(const uint32_t *gcount, const uint32_t *gocc, const uint32_t *gstart,
                           int64_t ng, CellBox B, double half, double *gbox, uint32_t *grange) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (g >= ng || gcount[g] == 0)
    return;
  const int64_t gx = (B.nx + kGroup - 1) / kGroup, gy = (B.ny + kGroup - 1) / kGroup;
  const int64_t gc[3] = {g % gx, (g / gx) % gy, g / (gx * gy)};
  const int64_t o[3] = {B.ox, B.oy, B.oz};
  const int64_t j = gocc[g];
  for (int k = 0; k < 3; ++k) {
    const double c0 = static_cast<double>(o[k] + kGroup * gc[k]);
    const double lo = B.gd * c0 + half - B.gd;
    const double hi = B.gd * (c0 + (kGroup - 1)) + half + B.gd;
    gbox[6 * j + k] = 0.5 * (lo + hi);
    gbox[6 * j + 3 + k] = 0.5 * (hi - lo);
  }
  grange[2 * j] = gstart[g];
  grange[2 * j + 1] = gstart[g + 1];
}

// the rounded-box distance of the reference (sphere_tracer.hpp:134-139) to
// the box of centre c and half extents e, minus gd
__device__ inline double box_sdf(const double *x, const double *c, const double *e, double gd) {
  double q[3], qm[3];
  for (int k = 0; k < 3; ++k) {
    q[k] = fabs(x[k] - c[k]) - e[k];
    qm[k] = std_max(q[k], 0.0);
  }
  const double mc = std_max(q[0], std_max(q[1], q[2]));
  return (sqrt(dot3(qm, qm)) + std_min(mc, 0.0)) - gd;
}

// minimum of the cell distances of group g into sdf (cells of half extent gd)
__device__ inline void group_min(const double *x, const uint32_t *grange, const double *gcent,
                                 int64_t g, double gd, double *sdf) {
  const double e[3] = {gd, gd, gd};
  for (uint32_t c = grange[2 * g]; c < grange[2 * g + 1]; ++c)
    *sdf = std_min(*sdf, box_sdf(x, gcent + 3 * static_cast<int64_t>(c), e, gd));
}

__global__ __launch_bounds__(kTB) void k_tr_march(TraceParams P, const double *gbox,
                                                  const uint32_t *grange, const double *gcent,
                                                  int64_t ngroups, const uint32_t *start,
                                                  const double *ex, CellBox B, double *image,
                                                  double *par) {
  const int64_t pid = static_cast<int64_t>(blockIdx.x) * kTB + threadIdx.x;
  if (pid >= P.width * P.height)
    return;
  const int64_t ix = pid % P.width, iy = pid / P.width;

  // pinhole_camera.hpp:53-62
  double d[3];
  {
    const double px = (P.sox + static_cast<double>(ix)) * P.pixel_size;
    const double py = (P.soy + static_cast<double>(iy)) * P.pixel_size;
    double dir0[3];
    for (int k = 0; k < 3; ++k) {
      const double pp = ((P.O[k] + px * P.hh[k]) + py * P.v[k]) - P.f * P.p[k];
      dir0[k] = P.O[k] - pp;
    }
    normalized3(dir0, d);
  }

  double t = 0, normal[3] = {0, 0, 0};
  bool done = false;  // broke out of the march
  for (int64_t step = 0; step < P.max_steps && !done; ++step) {
    double rx[3];
    for (int k = 0; k < 3; ++k)
      rx[k] = P.O[k] + t * d[k];
    // the minimum over every occupied cell of the reference, exactly: a
    // group's box contains its cells' boxes, so its distance bounds theirs
    // from below, and a group whose bound exceeds the minimum found so far
    // (beyond a margin far above the rounding of either) cannot lower it.
    // The group with the smallest bound is evaluated first.
    double sdf = DBL_MAX;
    int64_t first = -1;
    double best_lb = DBL_MAX;
    for (int64_t g = 0; g < ngroups; ++g) {
      const double lb = box_sdf(rx, gbox + 6 * g, gbox + 6 * g + 3, P.gd);
      if (lb < best_lb) {
        best_lb = lb;
        first = g;
      }
    }
    if (first >= 0)
      group_min(rx, grange, gcent, first, P.gd, &sdf);
    for (int64_t g = 0; g < ngroups; ++g) {
      if (g == first)
        continue;
      const double lb = box_sdf(rx, gbox + 6 * g, gbox + 6 * g + 3, P.gd);
      if (lb > sdf + 1e-9 * (fabs(sdf) + P.gd))
        continue;
      group_min(rx, grange, gcent, g, P.gd, &sdf);
    }
    if (sdf < P.gd) {
      int32_t rc[3];
      for (int k = 0; k < 3; ++k)
        rc[k] = static_cast<int32_t>(floor(rx[k] / P.gd));
      double phi = P.phi0, grad[3] = {0, 0, 0};
      int64_t count = 0;
      for (int cx = -1; cx <= 1; ++cx)
        for (int cy = -1; cy <= 1; ++cy)
          for (int cz = -1; cz <= 1; ++cz) {
            int64_t lin;
            if (!cell_index(B, rc[0] + cx, rc[1] + cy, rc[2] + cz, &lin))
              continue;
            for (uint32_t e = start[lin]; e < start[lin + 1]; ++e) {
              double dv[3];
              for (int k = 0; k < 3; ++k)
                dv[k] = rx[k] - ex[3 * static_cast<int64_t>(e) + k];
              const double r = sqrt(dot3(dv, dv));
              const double w = evalq(P.alpha, fabs(r) / P.half) / P.h3;
              if (w > 0) {
                double g[3] = {0, 0, 0};
                if (!(r < P.grad_eps)) {
                  const double dr =
                      -copysign(evaldq(P.alpha, fabs(r) / P.half) / P.h4, r);
                  const double s = dr / r;
                  for (int k = 0; k < 3; ++k)
                    g[k] = s * dv[k];
                }
                phi -= w;
                for (int k = 0; k < 3; ++k)
                  grad[k] -= g[k];
                ++count;
              }
            }
          }
      if (count > 0) {
        const double s = P.L * static_cast<double>(count);
        phi /= s;
        for (int k = 0; k < 3; ++k)
          normal[k] = grad[k] / s;
        sdf = std_max(sdf, phi);
      } else {
        sdf = std_max(sdf, P.gd / 4);
      }
    }
    if (sdf < 1e-6) {
      done = true;
      break;
    }
    t += sdf;
    if (t >= P.max_parameter)
      done = true;
  }
  if (!done)
    t = P.max_parameter;  // ran out of steps
  double pix;
  if (t == P.max_parameter) {
    pix = 0;
  } else if (t > P.max_parameter) {
    pix = 0.3;
  } else {
    double dn[3], nn[3];
    normalized3(d, dn);
    normalized3(normal, nn);
    pix = -dot3(dn, nn);
  }
  image[pid] = pix;
  if (par)
    par[pid] = t;
}

struct Dev {
  void *p = nullptr;
  explicit Dev(size_t bytes) { PRTCL_HIP_CHECK(hipMalloc(&p, bytes < 64 ? 64 : bytes)); }
  ~Dev() { (void)hipFree(p); }
  template <class U>
  U *as() const {
    return static_cast<U *>(p);
  }
};

unsigned blocks(int64_t n) { return static_cast<unsigned>((n + kTB - 1) / kTB); }

void scan_u32(hipStream_t st, const uint32_t *in, int64_t n, uint32_t *out) {
  // out has n + 1 entries; out[n] = total
  const int64_t nb = blocks(n + 1);
  Dev bsum(sizeof(uint32_t) * static_cast<size_t>(nb + 1));
  hipLaunchKernelGGL(k_tr_scan_tiles, dim3(static_cast<unsigned>(nb)), dim3(kTB), 0, st, in, n,
                     out, bsum.as<uint32_t>());
  hipLaunchKernelGGL(k_tr_scan_top, dim3(1), dim3(kTB), 0, st, bsum.as<uint32_t>(), nb);
  hipLaunchKernelGGL(k_tr_scan_add, dim3(static_cast<unsigned>(nb)), dim3(kTB), 0, st, out, n,
                     bsum.as<uint32_t>());
  PRTCL_HIP_CHECK(hipGetLastError());
  PRTCL_HIP_CHECK(hipStreamSynchronize(st));  // bsum is freed on return
}

}  // namespace

void trace_scene(SceneBase &sc, CtxResources &res, int group, const prtcl_hip_camera &cam,
                 double threshold, int32_t max_steps, double *image, double *parameter) {
  if (group != PRTCL_HIP_FLUID && group != PRTCL_HIP_BOUNDARY)
    throw std::invalid_argument("group must be PRTCL_HIP_FLUID or PRTCL_HIP_BOUNDARY");
  if (cam.width <= 0 || cam.height <= 0 || cam.width > (1 << 16) || cam.height > (1 << 16))
    throw std::invalid_argument("sensor width and height must be in 1..65536");
  if (max_steps < 0)
    throw std::invalid_argument("max_steps must not be negative");
  sc.ensure_resident();
  const prtcl_hip_scene_desc &d = sc.desc;
  hipStream_t st = res.stream;
  const int64_t n = group == PRTCL_HIP_FLUID ? d.n_fluid : d.n_boundary;

  TraceParams P{};
  // the model's smoothing_scale is a global of the scene's real type
  const double h = d.real_type == 1 ? d.smoothing_scale
                                    : static_cast<double>(static_cast<float>(d.smoothing_scale));
  P.gd = 4 * h;
  P.half = P.gd / 2;
  P.alpha = static_cast<double>(1.L / 4.L * (1.L / 3.141592653589793238462643383279502884L));
  P.h3 = cpow(P.half, 3);
  P.h4 = cpow(P.half, 4);
  P.phi0 = threshold * (evalq(P.alpha, std::abs(0.0) / P.half) / P.h3);
  P.L = -evaldq(P.alpha, 2.0 / 3.0) / P.h4;
  P.grad_eps = std::sqrt(std::numeric_limits<double>::epsilon());
  P.max_parameter = P.gd * 10000;
  P.width = cam.width;
  P.height = cam.height;
  P.max_steps = max_steps;
  // pinhole_camera.hpp:35-47
  normalized3(cam.up, P.v);
  normalized3(cam.principal, P.p);
  const double vp[3] = {P.v[1] * P.p[2] - P.v[2] * P.p[1], P.v[2] * P.p[0] - P.v[0] * P.p[2],
                        P.v[0] * P.p[1] - P.v[1] * P.p[0]};
  normalized3(vp, P.hh);
  P.pixel_size = 1.0 / static_cast<double>(cam.width);
  P.sox = -static_cast<double>(cam.width - 1) / 2;
  P.soy = -static_cast<double>(cam.height - 1) / 2;
  for (int k = 0; k < 3; ++k)
    P.O[k] = cam.origin[k];
  P.f = cam.focal_length;

  // positions in host order, as double
  const int64_t n3 = 3 * n;
  Dev x(sizeof(double) * static_cast<size_t>(n3));
  PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  if (n > 0) {
    Dev raw(sizeof(double) * static_cast<size_t>(n3));
    sc.gather_device(group, PRTCL_HIP_F_POSITION, raw.p, static_cast<size_t>(n3));
    if (d.real_type == 1)
      hipLaunchKernelGGL(k_tr_to_double<double>, dim3(blocks(n3)), dim3(kTB), 0, st,
                         raw.as<const double>(), x.as<double>(), n3);
    else
      hipLaunchKernelGGL(k_tr_to_double<float>, dim3(blocks(n3)), dim3(kTB), 0, st,
                         raw.as<const float>(), x.as<double>(), n3);
    PRTCL_HIP_CHECK(hipGetLastError());
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  }

  // the dense box of cells around the particles (one empty cell of margin)
  CellBox B{0, 0, 0, 1, 1, 1, P.gd};
  if (n > 0) {
    const int64_t nb = blocks(n);
    Dev part(sizeof(double) * 6 * static_cast<size_t>(nb));
    hipLaunchKernelGGL(k_tr_bbox, dim3(static_cast<unsigned>(nb)), dim3(kTB), 0, st,
                       x.as<const double>(), n, part.as<double>());
    std::vector<double> hp(6 * static_cast<size_t>(nb));
    PRTCL_HIP_CHECK(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * hp.size(),
                                   hipMemcpyDeviceToHost, st));
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t b = 0; b < nb; ++b)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], hp[6 * b + k]);
        hi[k] = std::max(hi[k], hp[6 * b + 3 + k]);
      }
    int32_t o[3], e[3];
    for (int k = 0; k < 3; ++k) {
      if (!(lo[k] <= hi[k]) || !std::isfinite(lo[k]) || !std::isfinite(hi[k])) {
        lo[k] = hi[k] = 0;  // no finite position: an empty box
      }
      const double a = std::floor(lo[k] / P.gd) - 1, b = std::floor(hi[k] / P.gd) + 1;
      if (a < -1e9 || b > 1e9 || b - a + 1 > 4096)
        throw std::invalid_argument("trace: the particles span more than 4096 cells per axis");
      o[k] = static_cast<int32_t>(a);
      e[k] = static_cast<int32_t>(b - a + 1);
    }
    if (static_cast<int64_t>(e[0]) * e[1] * e[2] > (int64_t(1) << 28))
      throw std::invalid_argument("trace: more than 2^28 cells around the particles");
    B = CellBox{o[0], o[1], o[2], e[0], e[1], e[2], P.gd};
  }
  const int64_t ncells = static_cast<int64_t>(B.nx) * B.ny * B.nz;

  Dev count(sizeof(uint32_t) * static_cast<size_t>(ncells + 1));
  Dev cursor(sizeof(uint32_t) * static_cast<size_t>(ncells + 1));
  Dev startb(sizeof(uint32_t) * static_cast<size_t>(ncells + 1));
  Dev occ(sizeof(uint32_t) * static_cast<size_t>(ncells + 1));
  Dev flag(sizeof(uint32_t) * static_cast<size_t>(ncells + 1));
  Dev lin_of(sizeof(int64_t) * static_cast<size_t>(n > 0 ? n : 1));
  Dev ent(sizeof(int64_t) * static_cast<size_t>(n > 0 ? n : 1));
  Dev ex(sizeof(double) * static_cast<size_t>(n3 > 0 ? n3 : 1));
  PRTCL_HIP_CHECK(hipMemsetAsync(count.p, 0, sizeof(uint32_t) * (ncells + 1), st));
  PRTCL_HIP_CHECK(hipMemsetAsync(cursor.p, 0, sizeof(uint32_t) * (ncells + 1), st));
  if (n > 0)
    hipLaunchKernelGGL(k_tr_count, dim3(blocks(n)), dim3(kTB), 0, st, x.as<const double>(), n,
                       B, count.as<uint32_t>(), lin_of.as<int64_t>());
  PRTCL_HIP_CHECK(hipGetLastError());
  scan_u32(st, count.as<const uint32_t>(), ncells, startb.as<uint32_t>());
  uint32_t placed = 0;
  PRTCL_HIP_CHECK(hipMemcpyAsync(&placed, startb.as<uint32_t>() + ncells, sizeof(placed),
                                 hipMemcpyDeviceToHost, st));
  PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  if (placed > 0) {
    hipLaunchKernelGGL(k_tr_place, dim3(blocks(n)), dim3(kTB), 0, st, lin_of.as<const int64_t>(),
                       n, startb.as<const uint32_t>(), cursor.as<uint32_t>(), ent.as<int64_t>());
    hipLaunchKernelGGL(k_tr_sort_cells, dim3(blocks(ncells)), dim3(kTB), 0, st,
                       startb.as<const uint32_t>(), ncells, ent.as<int64_t>());
    hipLaunchKernelGGL(k_tr_pack, dim3(blocks(placed)), dim3(kTB), 0, st, x.as<const double>(),
                       ent.as<const int64_t>(), static_cast<int64_t>(placed), ex.as<double>());
  }
  hipLaunchKernelGGL(k_tr_flags, dim3(blocks(ncells)), dim3(kTB), 0, st,
                     count.as<const uint32_t>(), ncells, flag.as<uint32_t>());
  PRTCL_HIP_CHECK(hipGetLastError());
  scan_u32(st, flag.as<const uint32_t>(), ncells, occ.as<uint32_t>());
  uint32_t nocc = 0;
  PRTCL_HIP_CHECK(hipMemcpyAsync(&nocc, occ.as<uint32_t>() + ncells, sizeof(nocc),
                                 hipMemcpyDeviceToHost, st));
  PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  Dev centres(sizeof(double) * 3 * static_cast<size_t>(nocc > 0 ? nocc : 1));
  Dev gidx(sizeof(uint32_t) * static_cast<size_t>(nocc > 0 ? nocc : 1));
  hipLaunchKernelGGL(k_tr_centres, dim3(blocks(ncells)), dim3(kTB), 0, st,
                     count.as<const uint32_t>(), occ.as<const uint32_t>(), ncells, B, P.half,
                     centres.as<double>(), gidx.as<uint32_t>());
  // groups of kGroup^3 cells
  const int64_t ng = ((B.nx + kGroup - 1) / kGroup) * ((B.ny + kGroup - 1) / kGroup) *
                     ((B.nz + kGroup - 1) / kGroup);
  Dev gcount(sizeof(uint32_t) * static_cast<size_t>(ng + 1));
  Dev gcursor(sizeof(uint32_t) * static_cast<size_t>(ng + 1));
  Dev gstart(sizeof(uint32_t) * static_cast<size_t>(ng + 1));
  Dev gflag(sizeof(uint32_t) * static_cast<size_t>(ng + 1));
  Dev gocc(sizeof(uint32_t) * static_cast<size_t>(ng + 1));
  PRTCL_HIP_CHECK(hipMemsetAsync(gcount.p, 0, sizeof(uint32_t) * (ng + 1), st));
  PRTCL_HIP_CHECK(hipMemsetAsync(gcursor.p, 0, sizeof(uint32_t) * (ng + 1), st));
  if (nocc > 0)
    hipLaunchKernelGGL(k_tr_gcount, dim3(blocks(nocc)), dim3(kTB), 0, st,
                       gidx.as<const uint32_t>(), static_cast<int64_t>(nocc),
                       gcount.as<uint32_t>());
  PRTCL_HIP_CHECK(hipGetLastError());
  scan_u32(st, gcount.as<const uint32_t>(), ng, gstart.as<uint32_t>());
  Dev gcent(sizeof(double) * 3 * static_cast<size_t>(nocc > 0 ? nocc : 1));
  if (nocc > 0)
    hipLaunchKernelGGL(k_tr_gplace, dim3(blocks(nocc)), dim3(kTB), 0, st,
                       gidx.as<const uint32_t>(), centres.as<const double>(),
                       static_cast<int64_t>(nocc), gstart.as<const uint32_t>(),
                       gcursor.as<uint32_t>(), gcent.as<double>());
  hipLaunchKernelGGL(k_tr_flags, dim3(blocks(ng)), dim3(kTB), 0, st, gcount.as<const uint32_t>(),
                     ng, gflag.as<uint32_t>());
  PRTCL_HIP_CHECK(hipGetLastError());
  scan_u32(st, gflag.as<const uint32_t>(), ng, gocc.as<uint32_t>());
  uint32_t ngroups = 0;
  PRTCL_HIP_CHECK(hipMemcpyAsync(&ngroups, gocc.as<uint32_t>() + ng, sizeof(ngroups),
                                 hipMemcpyDeviceToHost, st));
  PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  Dev gbox(sizeof(double) * 6 * static_cast<size_t>(ngroups > 0 ? ngroups : 1));
  Dev grange(sizeof(uint32_t) * 2 * static_cast<size_t>(ngroups > 0 ? ngroups : 1));
  hipLaunchKernelGGL(k_tr_glist, dim3(blocks(ng)), dim3(kTB), 0, st, gcount.as<const uint32_t>(),
                     gocc.as<const uint32_t>(), gstart.as<const uint32_t>(), ng, B, P.half,
                     gbox.as<double>(), grange.as<uint32_t>());

  const int64_t npx = cam.width * cam.height;
  Dev img(sizeof(double) * static_cast<size_t>(npx));
  Dev par(sizeof(double) * static_cast<size_t>(npx));
  hipLaunchKernelGGL(k_tr_march, dim3(blocks(npx)), dim3(kTB), 0, st, P,
                     gbox.as<const double>(), grange.as<const uint32_t>(),
                     gcent.as<const double>(), static_cast<int64_t>(ngroups),
                     startb.as<const uint32_t>(), ex.as<const double>(), B, img.as<double>(),
                     par.as<double>());
  PRTCL_HIP_CHECK(hipGetLastError());
  PRTCL_HIP_CHECK(hipMemcpyAsync(image, img.p, sizeof(double) * npx, hipMemcpyDeviceToHost, st));
  if (parameter)
    PRTCL_HIP_CHECK(hipMemcpyAsync(parameter, par.p, sizeof(double) * npx,
                                   hipMemcpyDeviceToHost, st));
  PRTCL_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace prtcl_amd
