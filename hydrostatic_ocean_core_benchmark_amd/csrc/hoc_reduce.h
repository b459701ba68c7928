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

// This workgroup's part and the pairwise sum(s) over it in NumPy's order.
// Parts: full 8192-element chunks are cut into `sub` aligned subtrees of
// 8192/sub elements (each a perfect tree of 128-element leaves); the
// trailing partial chunk contributes one part per leaf of its plan.
// Each thread evaluates elements independently (`elem(e)` may store
// per-element results; it runs exactly once per index), the products are
// staged in LDS, then eight lanes per leaf accumulate them with NumPy's
// eight interleaved accumulators and the leaves are combined pairwise.
template <int NS, class F>
__device__ Two part_sum(const PcgArgs& a, int part, F elem, double* lds /* [2][PCG_STAGE] */,
                        double* leafs /* [2][64] */) {
  const int tid = threadIdx.x;
  const long long nfp = a.nfull * a.sub;
  long long base;
  int nel, nleaves, leaf_len;
  if (part < nfp) {
    const long long c = part / a.sub;
    const int s = part % a.sub;
    nel = (int)(PW_CHUNK / a.sub);
    base = c * PW_CHUNK + (long long)s * nel;
    nleaves = nel / 128;
    leaf_len = 128;
  } else {
    const PwPlan* tp = &a.plans[1];
    const int l = (int)(part - nfp);
    base = a.nfull * PW_CHUNK + tp->leaf_off[l];
    nel = tp->leaf_len[l];
    nleaves = 1;
    leaf_len = nel;
  }
  for (int t = tid; t < nel; t += PCG_WG) {
    const Two v = elem(base + t);
    lds[t] = v.a;
    if (NS > 1) lds[PCG_STAGE + t] = v.b;
  }
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

// In-order float64 sum of a whole vector from the per-part partials: each
// full chunk's top tree levels over its `sub` subtree sums, the trailing
// chunk's tree from its plan over its leaf sums, then the chunk sums added
// sequentially. Every workgroup computes it identically.
__device__ double total_from_parts(const PcgArgs& a, const double* part, double* cs) {
  const int tid = threadIdx.x;
  __shared__ double s_tot;
  __shared__ double tail_v[256];
  const long long nfp = a.nfull * a.sub;
  for (int c = tid; c < a.nfull; c += PCG_WG) {
    double t[PCG_MAX_SUB];
#pragma unroll
    for (int s = 0; s < PCG_MAX_SUB; ++s)
      if (s < a.sub) t[s] = part[(long long)c * a.sub + s];
#pragma unroll
    for (int w = 1; w < PCG_MAX_SUB; w <<= 1)
#pragma unroll
      for (int s = 0; s + w < PCG_MAX_SUB; s += 2 * w)
        if (s + w < a.sub) t[s] = t[s] + t[s + w];
    cs[c] = t[0];
  }
  if (a.tail) {
    const PwPlan* tp = &a.plans[1];
    const int L = tp->nleaves;
    for (int l = tid; l < L; l += PCG_WG) tail_v[l] = part[nfp + l];
    __syncthreads();
    int start = 0;
    for (int lv = 0; lv < tp->nlevels; ++lv) {
      const int end = tp->level_end[lv];
      for (int n = start + tid; n < end; n += PCG_WG)
        tail_v[L + n] = tail_v[tp->node_l[n]] + tail_v[tp->node_r[n]];
      start = end;
      __syncthreads();
    }
    if (tid == 0) cs[a.nfull] = tp->nnodes > 0 ? tail_v[L + tp->nnodes - 1] : tail_v[0];
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int c = 0; c < a.nchunks; ++c) s = s + cs[c];
    s_tot = s;
  }
  __syncthreads();
  return s_tot;
}

}  // namespace hoc
