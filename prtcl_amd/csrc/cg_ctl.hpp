// FP64 control of the preconditioned CG of cg_openmp.hpp:32-215 on the
// device, shared by every solver that uses it (WKBB18's 3-vector system, the
// scalar systems of pt16): per-block partial sums of the dot products, the
// set-up scalars, alpha with the |p.q| < tol break, beta with the
// |r.y|_prev < tol break and ok(k, |r|^2), all in one block reading the
// partials in block order (deterministic), and the progress word the host
// polls.
#pragma once

#include "device_common.hpp"
#include "scene.hpp"

namespace prtcl_amd {
namespace {

// per-block partial sums of up to three FP64 dot products, stored at the
// block's logical index b (the partials are summed in logical block order)
template <int K>
__device__ __forceinline__ void store_partials(double *partials, unsigned nblocks,
                                               const double (&v)[K], unsigned b) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double s = block_sum<kBlock>(v[k]);
    if (threadIdx.x == 0)
      partials[k * nblocks + b] = s;
  }
}

template <int K>
__device__ __forceinline__ void store_partials(double *partials, unsigned nblocks,
                                               const double (&v)[K]) {
  store_partials<K>(partials, nblocks, v, blockIdx.x);
}

// the sum of n partials in block order (block_sum: valid in thread 0's
// wavefront only)
__device__ __forceinline__ double sum_partials(const double *p, unsigned n) {
  return block_sum<kBlock>(strided_sum<kBlock>(p, n));
}

// the same sum (thread 0's value) in every thread of the block
__device__ __forceinline__ double sum_partials_all(const double *p, unsigned n) {
  __shared__ double total;
  const double s = sum_partials(p, n);
  if (threadIdx.x == 0)
    total = s;
  __syncthreads();
  const double r = total;
  __syncthreads();  // a later call writes total again
  return r;
}

__device__ __forceinline__ void publish_cg(Ctl *c) {
  if (c->host_mirror) {
    const unsigned long long w = (static_cast<unsigned long long>(c->seq) << 32) |
                                 (static_cast<unsigned long long>(c->cg_done ? 1u : 0u) << 31) |
                                 static_cast<unsigned long long>(static_cast<unsigned>(c->cg_iter));
    __hip_atomic_store(c->host_mirror + 2, w, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__device__ __forceinline__ void cg_setup_fin_body(Ctl *c, const double *part, unsigned nb,
                                                  double tol, int max_k, unsigned seq) {
  const double bb = sum_partials(part, nb);
  const double rr = sum_partials(part + nb, nb);
  const double ry = sum_partials(part + 2 * nb, nb);
  if (threadIdx.x == 0) {
    c->cg_tol = tol;
    c->cg_max_k = max_k;
    c->cg_b_nsq = bb;
    // thr = max(tol^2 |b|^2, FLT_MIN)   (cg_openmp.hpp:74)
    const double thr = tol * tol * bb;
    c->cg_thr = thr > 1.1754943508222875e-38 ? thr : 1.1754943508222875e-38;
    c->cg_r_nsq = rr;
    c->cg_ry = ry;
    c->cg_iter = 0;
    c->cg_done = 0;  // ok(0) holds: k < min_k = 1
    c->seq = seq;
    publish_cg(c);
  }
}

__global__ __launch_bounds__(kBlock) void k_cg_setup_fin(Ctl *c, const double *part,
                                                         unsigned nb, double tol, int max_k,
                                                         unsigned seq) {
  cg_setup_fin_body(c, part, nb, tol, max_k, seq);
}

// |p.q| < tol breaks (cg_openmp.hpp:147); alpha = (r.y) / (p.q)
__device__ __forceinline__ void cg_alpha_body(Ctl *c, const double *part, unsigned nb) {
  if (c->cg_done)
    return;
  const double pq = sum_partials(part, nb);
  if (threadIdx.x == 0) {
    c->cg_ry_prev = c->cg_ry;
    c->cg_pq = pq;
    if (fabs(pq) < c->cg_tol) {
      c->cg_done = 1;
      publish_cg(c);
    } else {
      c->cg_alpha = c->cg_ry_prev / pq;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_cg_alpha(Ctl *c, const double *part, unsigned nb) {
  cg_alpha_body(c, part, nb);
}

// |r.y|_prev < tol breaks (:187); beta; ++k; ok(k, |r|^2) (:100-111)
__device__ __forceinline__ void cg_beta_body(Ctl *c, const double *part, unsigned nb) {
  if (c->cg_done)
    return;
  const double rr = sum_partials(part, nb);
  const double ry = sum_partials(part + nb, nb);
  if (threadIdx.x == 0) {
    c->cg_r_nsq = rr;
    c->cg_ry = ry;
    if (fabs(c->cg_ry_prev) < c->cg_tol) {
      c->cg_done = 1;
    } else {
      c->cg_beta = ry / c->cg_ry_prev;
      const int k = c->cg_iter + 1;
      c->cg_iter = k;
      if (k > c->cg_max_k || rr < c->cg_thr)
        c->cg_done = 1;
    }
    publish_cg(c);
  }
}

__global__ __launch_bounds__(kBlock) void k_cg_beta(Ctl *c, const double *part, unsigned nb) {
  cg_beta_body(c, part, nb);
}

// The same decisions inside the vector kernels of an iteration, for
// launches of at most kFusedCgBlocks blocks: every block sums all partials of
// the kernel before it (sum_partials_all: the order of the one-block
// kernels, so every block derives the same scalars and decisions), and block
// 0 writes them to the control block. The iteration count the stop test
// needs is staged in cg_iter_next beside alpha, so no value a block reads is
// written in the same kernel except cg_done, which a block reads once
// (cg_done_for_block): a block that sees it set would have decided the same.
// (kFusedCgBlocks: scene.hpp)

// cg_done as one value for the whole block. Block 0 of a folded kernel may
// set it while the wavefronts of another block start, so each wavefront
// reading it itself could see different values; a block that split would
// reduce with missing wavefronts. Thread 0 reads it for the block.
__device__ __forceinline__ bool cg_done_for_block(const Ctl *c) {
  __shared__ int flag;
  if (threadIdx.x == 0)
    flag = c->cg_done;
  __syncthreads();
  const bool done = flag != 0;
  __syncthreads();  // a later call writes flag again
  return done;
}

// k_cg_alpha in every block of the update kernel: false when the solve
// breaks (|p.q| < tol), else alpha
// (lb: the block's logical index; block 0 writes the control block)
__device__ __forceinline__ bool cg_alpha_in_block(Ctl *c, const double *pq_part, unsigned nb,
                                                  double *alpha, unsigned lb = blockIdx.x) {
  const double pq = sum_partials_all(pq_part, nb);
  const double ry_prev = c->cg_ry;
  const bool brk = fabs(pq) < c->cg_tol;
  if (lb == 0 && threadIdx.x == 0) {
    c->cg_ry_prev = ry_prev;
    c->cg_pq = pq;
    c->cg_iter_next = c->cg_iter + 1;
    if (brk) {
      c->cg_done = 1;
      publish_cg(c);
    } else {
      c->cg_alpha = ry_prev / pq;
    }
  }
  *alpha = brk ? 0.0 : ry_prev / pq;
  return !brk;
}

// k_cg_beta in every block of the direction kernel (r.r partials at part,
// r.y at part + nb): false when the solve is done, else beta
__device__ __forceinline__ bool cg_beta_in_block(Ctl *c, const double *part, unsigned nb,
                                                 double *beta, unsigned lb = blockIdx.x) {
  const double rr = sum_partials_all(part, nb);
  const double ry = sum_partials_all(part + nb, nb);
  const double ry_prev = c->cg_ry_prev;
  const bool brk = fabs(ry_prev) < c->cg_tol;
  const double b = brk ? 0.0 : ry / ry_prev;
  const int k = c->cg_iter_next;
  const bool done = brk || k > c->cg_max_k || rr < c->cg_thr;
  if (lb == 0 && threadIdx.x == 0) {
    c->cg_r_nsq = rr;
    c->cg_ry = ry;
    if (!brk) {
      c->cg_beta = b;
      c->cg_iter = k;
    }
    if (done)
      c->cg_done = 1;
    publish_cg(c);
  }
  *beta = b;
  return !done;
}

}  // namespace
}  // namespace prtcl_amd
