// Legacy-VTK output of a scene group with the numbers formatted on the
// device. Per section (POINTS, then each POINT_DATA array) three kernels run:
// the length of every item's text line and per-block sums, one block scanning
// the block sums into offsets, and the write of every line at its offset
// (lines are in host order, so a wavefront writes one contiguous run of
// text). The text then goes to one of two page-locked buffers, and the file
// write of a section overlaps the formatting of the next.
//
// Every real is printed as the C library's "%.6f" (what std::fixed with the
// default precision 6 prints, util/save_vtk.cpp:14,27): the exact binary
// value rounded half to even at the sixth decimal, a '-' whenever the sign
// bit is set (also for -0 and for values that round to zero), "nan"/"-nan"
// and "inf"/"-inf" for the non-finite values.
#include "vtk.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace prtcl_amd {
namespace {

constexpr int kVtkBlock = 256;

// a line sink: counts characters, and writes them when kWrite
template <bool kWrite>
struct Sink {
  char *p;
  uint32_t n;
  __device__ void put(char c) {
    if (kWrite)
      p[n] = c;
    ++n;
  }
};

__device__ __forceinline__ int digits_u64(uint64_t v) {
  int d = 1;
  while (v >= 10) {
    v /= 10;
    ++d;
  }
  return d;
}

// v in decimal, at least `width` digits (zero-padded)
template <bool kW>
__device__ void put_u64(Sink<kW> &s, uint64_t v, int width = 1) {
  const int d = max(digits_u64(v), width);
  if (kW) {
    uint64_t x = v;
    for (int i = d - 1; i >= 0; --i) {
      s.p[s.n + i] = static_cast<char>('0' + x % 10);
      x /= 10;
    }
  }
  s.n += d;
}

// m * 2^e (m < 2^53, e >= 0, at least 2^64) in decimal: the integer as
// 32-bit limbs, cut into base-10^9 chunks by long division
template <bool kW>
__device__ void put_big(Sink<kW> &s, uint64_t m, int e) {
  uint32_t w[34];
  for (int i = 0; i < 34; ++i)
    w[i] = 0;
  const int q = e >> 5, r = e & 31;
  const uint64_t lo = m << r;
  const uint64_t hi = r ? (m >> (64 - r)) : 0;
  w[q] = static_cast<uint32_t>(lo);
  w[q + 1] = static_cast<uint32_t>(lo >> 32);
  w[q + 2] = static_cast<uint32_t>(hi);
  int nw = q + 3;
  while (nw > 0 && w[nw - 1] == 0)
    --nw;
  uint32_t c[36];
  int nc = 0;
  while (nw > 0) {
    uint64_t rem = 0;
    for (int i = nw - 1; i >= 0; --i) {
      const uint64_t cur = (rem << 32) | w[i];
      w[i] = static_cast<uint32_t>(cur / 1000000000u);
      rem = cur % 1000000000u;
    }
    c[nc++] = static_cast<uint32_t>(rem);
    while (nw > 0 && w[nw - 1] == 0)
      --nw;
  }
  put_u64(s, c[nc - 1]);
  for (int i = nc - 2; i >= 0; --i)
    put_u64(s, c[i], 9);
}

template <class T>
struct Bits;
template <>
struct Bits<float> {
  using U = uint32_t;
  static constexpr int kMant = 23, kBias = 127, kExpMax = 255;
};
template <>
struct Bits<double> {
  using U = uint64_t;
  static constexpr int kMant = 52, kBias = 1023, kExpMax = 2047;
};

// "%.6f" of v
template <class T, bool kW>
__device__ void put_fixed6(Sink<kW> &s, T v) {
  using B = Bits<T>;
  typename B::U u;
  static_assert(sizeof(u) == sizeof(v), "");
  __builtin_memcpy(&u, &v, sizeof(v));
  const bool neg = (u >> (8 * sizeof(u) - 1)) != 0;
  const int E = static_cast<int>((u >> B::kMant) & B::kExpMax);
  const uint64_t F = static_cast<uint64_t>(u) & ((uint64_t(1) << B::kMant) - 1);
  if (E == B::kExpMax) {
    if (neg)
      s.put('-');
    if (F) {
      s.put('n');
      s.put('a');
      s.put('n');
    } else {
      s.put('i');
      s.put('n');
      s.put('f');
    }
    return;
  }
  if (neg)
    s.put('-');
  // v = m 2^e
  const uint64_t m = E ? (F | (uint64_t(1) << B::kMant)) : F;
  const int e = (E ? E : 1) - B::kBias - B::kMant;
  uint64_t ip = 0;
  uint32_t fr = 0;
  if (m != 0 && e >= 0) {
    if (64 - __clzll(m) + e > 64) {
      put_big(s, m, e);
      s.put('.');
      for (int i = 0; i < 6; ++i)
        s.put('0');
      return;
    }
    ip = m << e;
  } else if (m != 0) {
    const int k = -e;  // v = m / 2^k
    // integer part, and the fraction's bits f (f < 2^k)
    uint64_t f;
    if (k < 64) {
      ip = m >> k;
      f = m & ((uint64_t(1) << k) - 1);
    } else {
      ip = 0;
      f = m;
    }
    // q = round-half-even(f 10^6 / 2^k), with f 10^6 as (hi, lo) 64-bit words
    const uint64_t plo = f * 1000000u;
    const uint64_t phi = __umul64hi(f, 1000000u);
    uint64_t q = 0;
    bool up = false;
    if (k < 64) {
      q = (plo >> k) | (phi << (64 - k));
      const uint64_t rem = plo & ((uint64_t(1) << k) - 1);
      const uint64_t half = uint64_t(1) << (k - 1);
      up = rem > half || (rem == half && (q & 1));
    } else if (k < 128) {
      const int k2 = k - 64;
      q = k2 ? (phi >> k2) : phi;
      const uint64_t rhi = k2 ? (phi & ((uint64_t(1) << k2) - 1)) : 0;
      // half = 2^(k-1) as (hi, lo)
      const uint64_t hhi = k2 ? (uint64_t(1) << (k2 - 1)) : 0;
      const uint64_t hlo = k2 ? 0 : (uint64_t(1) << 63);
      const bool gt = rhi > hhi || (rhi == hhi && plo > hlo);
      const bool eq = rhi == hhi && plo == hlo;
      up = gt || (eq && (q & 1));
    }  // k >= 128: f 10^6 < 2^73 < 2^(k-1), q = 0, no rounding up
    if (up)
      ++q;
    if (q == 1000000u) {
      q = 0;
      ++ip;
    }
    fr = static_cast<uint32_t>(q);
  }
  put_u64(s, ip);
  s.put('.');
  put_u64(s, fr, 6);
}

// one text line: the host index (index_mode) or ncomp space-separated reals
template <class T, bool kW>
__device__ void put_line(Sink<kW> &s, const T *vals, int ncomp, int64_t h, int index_mode) {
  if (index_mode) {
    put_u64(s, static_cast<uint64_t>(h));
  } else {
    for (int c = 0; c < ncomp; ++c) {
      if (c)
        s.put(' ');
      put_fixed6(s, vals[h * ncomp + c]);
    }
  }
  s.put('\n');
}

// exclusive scan of one value per thread over the block; *total gets the sum
__device__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t *total) {
  __shared__ uint64_t wsum[kVtkBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t x = v;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, 64);
    if (lane >= d)
      x += y;
  }
  if (lane == 63)
    wsum[wv] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int k = 0; k < kVtkBlock / 64; ++k) {
    if (k < wv)
      before += wsum[k];
    all += wsum[k];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

template <class T>
__global__ __launch_bounds__(kVtkBlock) void k_vtk_len(const T *vals, int ncomp, int64_t n,
                                                       int index_mode, uint64_t *bsum) {
  const int64_t h = static_cast<int64_t>(blockIdx.x) * kVtkBlock + threadIdx.x;
  uint32_t len = 0;
  if (h < n) {
    Sink<false> s{nullptr, 0};
    put_line(s, vals, ncomp, h, index_mode);
    len = s.n;
  }
  uint64_t total;
  block_excl_scan_u64(len, &total);
  if (threadIdx.x == 0)
    bsum[blockIdx.x] = total;
}

// bsum[0..nb) -> exclusive prefix sums, bsum[nb] = total (one block)
__global__ __launch_bounds__(kVtkBlock) void k_vtk_scan(uint64_t *bsum, int64_t nb) {
  uint64_t carry = 0;
  for (int64_t base = 0; base < nb; base += kVtkBlock) {
    const int64_t i = base + threadIdx.x;
    const uint64_t v = i < nb ? bsum[i] : 0;
    uint64_t total;
    const uint64_t ex = block_excl_scan_u64(v, &total);
    if (i < nb)
      bsum[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0)
    bsum[nb] = carry;
}

template <class T>
__global__ __launch_bounds__(kVtkBlock) void k_vtk_write(const T *vals, int ncomp, int64_t n,
                                                         int index_mode, const uint64_t *bsum,
                                                         char *out) {
  const int64_t h = static_cast<int64_t>(blockIdx.x) * kVtkBlock + threadIdx.x;
  uint32_t len = 0;
  if (h < n) {
    Sink<false> s{nullptr, 0};
    put_line(s, vals, ncomp, h, index_mode);
    len = s.n;
  }
  uint64_t total;
  const uint64_t off = bsum[blockIdx.x] + block_excl_scan_u64(len, &total);
  if (h < n) {
    Sink<true> s{out + off, 0};
    put_line(s, vals, ncomp, h, index_mode);
  }
}

struct Section {
  std::string header;  // text before the lines
  int field;           // -1: the host index
  int ncomp;
};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool pinned = false;
  void ensure(size_t b) {
    if (b <= bytes)
      return;
    release();
    if (pinned)
      PRTCL_HIP_CHECK(hipHostMalloc(&p, b, hipHostMallocDefault));
    else
      PRTCL_HIP_CHECK(hipMalloc(&p, b));
    bytes = b;
  }
  void release() {
    if (p)
      (void)(pinned ? hipHostFree(p) : hipFree(p));
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
};

template <class T>
uint64_t save_vtk_impl(SceneBase &sc, CtxResources &res, int group, const char *path) {
  const prtcl_hip_scene_desc &d = sc.desc;
  const int64_t n = group == PRTCL_HIP_FLUID ? d.n_fluid : d.n_boundary;
  const char *gname = group == PRTCL_HIP_FLUID ? "f" : "b";
  const char *gtype = group == PRTCL_HIP_FLUID ? "fluid" : "boundary";
  const std::string ns = std::to_string(n);

  // the sections of util/save_vtk.cpp:57-129, for the fields the group holds
  std::vector<Section> secs;
  secs.push_back({std::string("# vtk DataFile Version 2.0\n") + gname + " " + gtype +
                      "\nASCII\nDATASET POLYDATA\nPOINTS " + ns + " float\n",
                  PRTCL_HIP_F_POSITION, 3});
  std::string pending = "POINT_DATA " + ns + "\n";
  auto add = [&](const char *name, int field, bool vector) {
    if (field >= 0 && sc.field_components(group, field) == 0)
      return;
    std::string h = pending + (vector ? std::string("VECTORS ") + name + " float\n"
                                      : std::string("SCALARS ") + name +
                                            " float 1\nLOOKUP_TABLE default\n");
    pending.clear();
    secs.push_back({h, field, vector ? 3 : 1});
  };
  add("velocity", PRTCL_HIP_F_VELOCITY, true);
  add("acceleration", PRTCL_HIP_F_ACCELERATION, true);
  add("_index", -1, false);
  add("mass", PRTCL_HIP_F_MASS, false);
  add("volume", PRTCL_HIP_F_VOLUME, false);
  add("density", PRTCL_HIP_F_DENSITY, false);
  add("pressure", PRTCL_HIP_F_PRESSURE, false);
  add("time_of_birth", PRTCL_HIP_F_TIME_OF_BIRTH, false);

  FILE *f = std::fopen(path, "wb");
  if (!f)
    throw std::invalid_argument(std::string("vtk: cannot open for writing: ") + path);
  struct Close {
    FILE *&f;
    ~Close() {
      if (f)
        std::fclose(f);
    }
  } close_guard{f};
  uint64_t written = 0;
  auto out = [&](const void *p, size_t b) {
    if (b && std::fwrite(p, 1, b, f) != b)
      throw std::runtime_error("vtk: write failed");
    written += b;
  };

  hipStream_t st = res.stream;
  PRTCL_HIP_CHECK(hipStreamSynchronize(st));
  const int64_t nb = (n + kVtkBlock - 1) / kVtkBlock;
  DevBuf vals, bsum, dev[2], host[2];
  host[0].pinned = host[1].pinned = true;
  vals.ensure(std::max<size_t>(static_cast<size_t>(n) * 3 * sizeof(T), 64));
  bsum.ensure(sizeof(uint64_t) * static_cast<size_t>(nb + 1));
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (hipEvent_t &e : ev)
    PRTCL_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct Events {
    hipEvent_t *ev;
    ~Events() {
      for (int k = 0; k < 2; ++k)
        (void)hipEventDestroy(ev[k]);
    }
  } ev_guard{ev};
  std::vector<uint64_t> bytes(secs.size(), 0);

  // lengths, offsets and text of section k into dev[k&1], copied to host[k&1]
  auto prepare = [&](size_t k) {
    if (n == 0)
      return;
    const Section &S = secs[k];
    const T *v = static_cast<const T *>(vals.p);
    if (S.field >= 0)
      sc.gather_device(group, S.field, vals.p, static_cast<size_t>(n) * S.ncomp);
    const int idx = S.field < 0 ? 1 : 0;
    auto *bs = static_cast<uint64_t *>(bsum.p);
    hipLaunchKernelGGL(k_vtk_len<T>, dim3(static_cast<unsigned>(nb)), dim3(kVtkBlock), 0, st, v,
                       S.ncomp, n, idx, bs);
    hipLaunchKernelGGL(k_vtk_scan, dim3(1), dim3(kVtkBlock), 0, st, bs, nb);
    PRTCL_HIP_CHECK(hipGetLastError());
    uint64_t total = 0;
    PRTCL_HIP_CHECK(hipMemcpyAsync(&total, bs + nb, sizeof(total), hipMemcpyDeviceToHost, st));
    PRTCL_HIP_CHECK(hipStreamSynchronize(st));
    const int b = static_cast<int>(k & 1);
    dev[b].ensure(std::max<uint64_t>(total, 64));
    host[b].ensure(std::max<uint64_t>(total, 64));
    hipLaunchKernelGGL(k_vtk_write<T>, dim3(static_cast<unsigned>(nb)), dim3(kVtkBlock), 0, st,
                       v, S.ncomp, n, idx, bs, static_cast<char *>(dev[b].p));
    PRTCL_HIP_CHECK(hipGetLastError());
    PRTCL_HIP_CHECK(hipMemcpyAsync(host[b].p, dev[b].p, total, hipMemcpyDeviceToHost, st));
    PRTCL_HIP_CHECK(hipEventRecord(ev[b], st));
    bytes[k] = total;
  };

  prepare(0);
  for (size_t k = 0; k < secs.size(); ++k) {
    if (k + 1 < secs.size())
      prepare(k + 1);  // the device formats the next section meanwhile
    out(secs[k].header.data(), secs[k].header.size());
    if (n > 0) {
      const int b = static_cast<int>(k & 1);
      PRTCL_HIP_CHECK(hipEventSynchronize(ev[b]));
      out(host[b].p, bytes[k]);
    }
  }
  if (!pending.empty())
    out(pending.data(), pending.size());
  FILE *ff = f;
  f = nullptr;
  if (std::fclose(ff) != 0)
    throw std::runtime_error("vtk: close failed");
  return written;
}

}  // namespace

uint64_t save_scene_vtk(SceneBase &sc, CtxResources &res, int group, const char *path) {
  if (group != PRTCL_HIP_FLUID && group != PRTCL_HIP_BOUNDARY)
    throw std::invalid_argument("group must be PRTCL_HIP_FLUID or PRTCL_HIP_BOUNDARY");
  sc.ensure_resident();
  return sc.desc.real_type == 1 ? save_vtk_impl<double>(sc, res, group, path)
                                : save_vtk_impl<float>(sc, res, group, path);
}

}  // namespace prtcl_amd
