// NumPy-order float64 reductions over a PCG partition (see PcgArgs in
// hoc_kernels.h): per-part pairwise sums and the in-order total.
#pragma once

#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 8); }

struct Two {
  double a, b;
};

// The elements of one part: [base, base + nel), nleaves leaves of leaf_len.
// Parts: full 8192-element chunks are cut into `sub` aligned subtrees of
// 8192/sub elements (each a perfect tree of 128-element leaves); the
// trailing partial chunk contributes one part per leaf of its plan.
struct PartRange {
  long long base;
  int nel, nleaves, leaf_len;
};

__device__ __forceinline__ PartRange part_range(const PcgArgs& a, int part) {
  const long long nfp = a.nfull * a.sub;
  PartRange r;
  if (part < nfp) {
    const long long c = part / a.sub;
    const int s = part % a.sub;
    r.nel = (int)(PW_CHUNK / a.sub);
    r.base = c * PW_CHUNK + (long long)s * r.nel;
    r.nleaves = r.nel / 128;
    r.leaf_len = 128;
  } else {
    const PwPlan* tp = &a.plans[1];
    const int l = (int)(part - nfp);
    r.base = a.nfull * PW_CHUNK + tp->leaf_off[l];
    r.nel = tp->leaf_len[l];
    r.nleaves = 1;
    r.leaf_len = r.nel;
  }
  return r;
}

// Pairwise sum(s) in NumPy's order over values staged in lds[0..nel)
// (and lds[PCG_STAGE..] for the second sum): eight lanes per leaf
// accumulate with NumPy's eight interleaved accumulators, then the leaves
// are combined pairwise. Starts with a barrier (the staging must be done).
template <int NS>
__device__ Two part_reduce(const PartRange& pr, double* lds, double* leafs /* [2][64] */) {
  const int tid = threadIdx.x;
  const int nleaves = pr.nleaves, leaf_len = pr.leaf_len;
  __syncthreads();
  const int g = tid >> 3, lane = tid & 7;
  if (g < nleaves) {
    const int lo = g * leaf_len;
    const int len = leaf_len;
    Two res{0.0, 0.0};
    if (len >= 8) {
      const int fl = len - (len % 8);
      Two acc{lds[lo + lane], NS > 1 ? lds[PCG_STAGE + lo + lane] : 0.0};
      for (int t = 8; t < fl; t += 8) {
        acc.a = acc.a + lds[lo + t + lane];
        if (NS > 1) acc.b = acc.b + lds[PCG_STAGE + lo + t + lane];
      }
      for (int m = 1; m < 8; m <<= 1) {
        acc.a = acc.a + shfl_xor_d(acc.a, m);
        if (NS > 1) acc.b = acc.b + shfl_xor_d(acc.b, m);
      }
      res = acc;
      if (lane == 0) {
        for (int t = fl; t < len; ++t) {
          res.a = res.a + lds[lo + t];
          if (NS > 1) res.b = res.b + lds[PCG_STAGE + lo + t];
        }
      }
    } else if (lane == 0) {
      for (int t = 0; t < len; ++t) {
        res.a = res.a + lds[lo + t];
        if (NS > 1) res.b = res.b + lds[PCG_STAGE + lo + t];
      }
    }
    if (lane == 0) {
      leafs[g] = res.a;
      if (NS > 1) leafs[64 + g] = res.b;
    }
  }
  __syncthreads();
  // perfect pairwise tree over the part's leaves (1, 2, 4, ... leaves)
  for (int w = 1; w < nleaves; w <<= 1) {
    if (tid < nleaves / (2 * w)) {
      leafs[2 * w * tid] = leafs[2 * w * tid] + leafs[2 * w * tid + w];
      if (NS > 1) leafs[64 + 2 * w * tid] = leafs[64 + 2 * w * tid] + leafs[64 + 2 * w * tid + w];
    }
    __syncthreads();
  }
  Two out{leafs[0], NS > 1 ? leafs[64] : 0.0};
  __syncthreads();
  return out;
}

// This workgroup's part and the pairwise sum(s) over it in NumPy's order.
// Each thread evaluates elements independently (`elem(e)` may store
// per-element results; it runs exactly once per index) and the values are
// staged in LDS for part_reduce.
template <int NS, class F>
__device__ Two part_sum(const PcgArgs& a, int part, F elem, double* lds /* [2][PCG_STAGE] */,
                        double* leafs /* [2][64] */) {
  const PartRange pr = part_range(a, part);
  for (int t = threadIdx.x; t < pr.nel; t += PCG_WG) {
    const Two v = elem(pr.base + t);
    lds[t] = v.a;
    if (NS > 1) lds[PCG_STAGE + t] = v.b;
  }
  return part_reduce<NS>(pr, lds, leafs);
}

// In-order float64 sums of whole vectors from their per-part partials: each
// full chunk's top tree levels over its `sub` subtree sums, the trailing
// chunk's tree from its plan over its leaf sums, then the chunk sums added
// sequentially. Every workgroup computes them identically; NV vectors
// (1 or 2) are summed in one pass.
template <int NV>
__device__ void totals_from_parts(const PcgArgs& a, const double* const* part, double* cs,
                                  double* out) {
  const int tid = threadIdx.x;
  __shared__ double s_tot[2];
  __shared__ double tail_v[2][256];
  const long long nfp = a.nfull * a.sub;
  for (int c = tid; c < a.nfull; c += PCG_WG) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double t[PCG_MAX_SUB];
#pragma unroll
      for (int s = 0; s < PCG_MAX_SUB; ++s)
        if (s < a.sub) t[s] = part[q][(long long)c * a.sub + s];
#pragma unroll
      for (int w = 1; w < PCG_MAX_SUB; w <<= 1)
#pragma unroll
        for (int s = 0; s + w < PCG_MAX_SUB; s += 2 * w)
          if (s + w < a.sub) t[s] = t[s] + t[s + w];
      cs[q * PCG_MAX_CHUNKS + c] = t[0];
    }
  }
  if (a.tail) {
    const PwPlan* tp = &a.plans[1];
    const int L = tp->nleaves;
    for (int l = tid; l < L; l += PCG_WG)
#pragma unroll
      for (int q = 0; q < NV; ++q) tail_v[q][l] = part[q][nfp + l];
    __syncthreads();
    int start = 0;
    for (int lv = 0; lv < tp->nlevels; ++lv) {
      const int end = tp->level_end[lv];
      for (int n = start + tid; n < end; n += PCG_WG)
#pragma unroll
        for (int q = 0; q < NV; ++q)
          tail_v[q][L + n] = tail_v[q][tp->node_l[n]] + tail_v[q][tp->node_r[n]];
      start = end;
      __syncthreads();
    }
    if (tid < NV)
      cs[tid * PCG_MAX_CHUNKS + a.nfull] =
          tp->nnodes > 0 ? tail_v[tid][L + tp->nnodes - 1] : tail_v[tid][0];
  }
  __syncthreads();
  if (tid < NV) {
    double s = 0.0;
    for (int c = 0; c < a.nchunks; ++c) s = s + cs[tid * PCG_MAX_CHUNKS + c];
    s_tot[tid] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q] = s_tot[q];
  __syncthreads();
}

__device__ __forceinline__ double total_from_parts(const PcgArgs& a, const double* part,
                                                   double* cs) {
  const double* p[1] = {part};
  double out[1];
  totals_from_parts<1>(a, p, cs, out);
  return out[0];
}

}  // namespace hoc
