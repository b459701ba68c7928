// Barotropic Helmholtz solve  eta - coef D[K G eta] = rhs  on the masked
// C-grid (libs/core/solvers_var.py:22-197), PCG with device-resident
// scalars: no host synchronisation inside an iteration; the host only polls
// a `done` flag between batches of iterations, and every kernel of an
// iteration past convergence returns immediately, so the iterate and the
// iteration count are exactly those of the sequential algorithm.
//
// Dot products reproduce NumPy's float64 `sum` bit for bit: the vector is cut
// into 8192-element chunks; each chunk is summed pairwise (recursive halving
// to blocks of <= 128 elements, each block with eight interleaved
// accumulators combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then its
// remainder); the chunk sums are then added in order. One workgroup owns one
// chunk: eight lanes per block accumulate, a 3-step xor-shuffle combines the
// eight accumulators in exactly that order, and the blocks' tree is walked
// level by level in LDS from a host-built plan.
#include "hoc_device.h"
#include "hoc_kernels.h"
#include "hoc_reduce.h"

namespace hoc {

constexpr int PW_THREADS = 512;  // 64 groups of 8 lanes


// Pairwise sum(s) of one chunk. `elem(e)` returns the product(s) summed at
// flat index e and may also store per-element results; every index of the
// chunk is visited exactly once. Returns the chunk sums in lane 0 of the
// block (valid for threadIdx.x == 0).
template <int NS, class F>
__device__ Two chunk_sum(const PwPlan* __restrict__ plan, long long base, F elem,
                         double* lds /* [2][256] */) {
  const int tid = threadIdx.x;
  const int g = tid >> 3, lane = tid & 7;
  const int L = plan->nleaves;
  for (int leaf = g; leaf < L; leaf += PW_THREADS / 8) {
    const int off = plan->leaf_off[leaf], len = plan->leaf_len[leaf];
    Two res{0.0, 0.0};
    if (len == 128) {
      // 16 elements per lane: evaluate them all first so their loads are in
      // flight together, then accumulate in NumPy's order
      Two v[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) v[t] = elem(base + off + 8 * t + lane);
      Two acc = v[0];
#pragma unroll
      for (int t = 1; t < 16; ++t) {
        acc.a = acc.a + v[t].a;
        if (NS > 1) acc.b = acc.b + v[t].b;
      }
      for (int m = 1; m < 8; m <<= 1) {
        acc.a = acc.a + shfl_xor_d(acc.a, m);
        if (NS > 1) acc.b = acc.b + shfl_xor_d(acc.b, m);
      }
      res = acc;
    } else if (len >= 8) {
      const int full = len - (len % 8);
      Two acc = elem(base + off + lane);
      for (int t = 8; t < full; t += 8) {
        const Two v = elem(base + off + t + lane);
        acc.a = acc.a + v.a;
        if (NS > 1) acc.b = acc.b + v.b;
      }
      for (int m = 1; m < 8; m <<= 1) {
        acc.a = acc.a + shfl_xor_d(acc.a, m);
        if (NS > 1) acc.b = acc.b + shfl_xor_d(acc.b, m);
      }
      res = acc;
      if (lane == 0) {
        for (int t = full; t < len; ++t) {
          const Two v = elem(base + off + t);
          res.a = res.a + v.a;
          if (NS > 1) res.b = res.b + v.b;
        }
      }
    } else if (lane == 0) {
      for (int t = 0; t < len; ++t) {
        const Two v = elem(base + off + t);
        res.a = res.a + v.a;
        if (NS > 1) res.b = res.b + v.b;
      }
    }
    if (lane == 0) {
      lds[leaf] = res.a;
      if (NS > 1) lds[256 + leaf] = res.b;
    }
  }
  __syncthreads();
  int start = 0;
  for (int lv = 0; lv < plan->nlevels; ++lv) {
    const int end = plan->level_end[lv];
    for (int n = start + tid; n < end; n += PW_THREADS) {
      const int l = plan->node_l[n], r = plan->node_r[n];
      lds[L + n] = lds[l] + lds[r];
      if (NS > 1) lds[256 + L + n] = lds[256 + l] + lds[256 + r];
    }
    start = end;
    __syncthreads();
  }
  const int root = plan->nnodes > 0 ? L + plan->nnodes - 1 : 0;
  Two out{lds[root], NS > 1 ? lds[256 + root] : 0.0};
  return out;
}

__device__ __forceinline__ const PwPlan* plan_of(const PwPlan* plans, long long chunk,
                                                 long long nfull) {
  return chunk < nfull ? &plans[0] : &plans[1];
}
__device__ __forceinline__ long long chunk_len(long long chunk, long long n) {
  const long long b = chunk * PW_CHUNK;
  return (n - b) < PW_CHUNK ? (n - b) : PW_CHUNK;
}

// D[K G x] at flat index c over the masked field: the five-point
// variable-coefficient Laplacian with face coefficients Ku (east) and Kv
// (north). Returns x[c] * mask[c] through xc_out.
__device__ __forceinline__ double divgrad_at(const Dims& d, const HelmOp& op, const double* x,
                                             long long c, double* xc_out) {
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i), cs = at2(d, wrap(j, -1, d.ny), i);
  const double mc = mk(op.mask[c]);
  const double xc = x[c] * mc;
  const double xe = x[ce] * mk(op.mask[ce]);
  const double xw = x[cw] * mk(op.mask[cw]);
  const double xn = x[cn] * mk(op.mask[cn]);
  const double xs = x[cs] * mk(op.mask[cs]);
  const double fx = op.Ku[c] * (xe - xc);
  const double fxw = op.Ku[cw] * (xc - xw);
  const double fy = op.Kv[c] * (xn - xc);
  const double fys = op.Kv[cs] * (xc - xs);
  *xc_out = xc;
  return (fx - fxw) / op.dx2 + (fy - fys) / op.dy2;
}

// Masked Helmholtz operator A x at flat index c (libs/core/solvers_var.py:47-53).
__device__ __forceinline__ double apply_A(const Dims& d, const HelmOp& op, const double* x,
                                          long long c) {
  double xc;
  const double lap = divgrad_at(d, op, x, c, &xc);
  const double mc = mk(op.mask[c]);
  return (xc - op.coef * lap) * mc + xc * (1.0 - mc);
}

// out = mask * D[K G x] (the operator whose truncation order the reference
// measures, tools/verify_v05.py:113-176).
__global__ void k_helm_divgrad(Dims d, HelmOp op, const double* __restrict__ x,
                               double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  double xc;
  const double lap = divgrad_at(d, op, x, c, &xc);
  out[c] = lap * mk(op.mask[c]);
}

// diag = 1 + cx (Ku + Ku_w) + cy (Kv + Kv_s), 1 on dry cells; inv = 1/diag.
__global__ void k_helm_diag(Dims d, HelmOp op, double* __restrict__ inv_diag) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long cw = at2(d, j, wrap(i, -1, d.nx)), cs = at2(d, wrap(j, -1, d.ny), i);
  double dg = 1.0 + op.cx * (op.Ku[c] + op.Ku[cw]) + op.cy * (op.Kv[c] + op.Kv[cs]);
  dg = op.mask[c] ? dg : 1.0;
  inv_diag[c] = 1.0 / dg;
}

// Generic pairwise dot (used by tests of the reduction): part[chunk] = sum(a*b)
__global__ __launch_bounds__(PW_THREADS) void k_pw_dot(const double* __restrict__ a,
                                                       const double* __restrict__ b,
                                                       const PwPlan* __restrict__ plans,
                                                       long long nfull, double* __restrict__ part) {
  __shared__ double lds[512];
  const long long chunk = blockIdx.x;
  const long long base = chunk * PW_CHUNK;
  const Two s = chunk_sum<1>(plan_of(plans, chunk, nfull), base, [&](long long e) -> Two {
    return Two{a[e] * b[e], 0.0};
  }, lds);
  if (threadIdx.x == 0) part[chunk] = s.a;
}

__global__ void k_seq_sum(const double* __restrict__ part, long long nchunks, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (long long c = 0; c < nchunks; ++c) s = s + part[c];
  out[0] = s;
}

}  // namespace hoc

// ====================================================== persistent PCG
namespace hoc {

constexpr long long SPIN_LIMIT = 1ll << 22;

// Grid barrier over all workgroups of the cooperative launch. Bounded: a
// wait longer than SPIN_LIMIT polls raises the error flag and every
// workgroup leaves, so the kernel always finishes. Returns false on error.
//
// bar holds GRID_BAR_WORDS words, each counter on a 256-byte line of its
// own: the arrival counters of the 8 workgroup groups (blockIdx.x mod 8,
// the XCD a workgroup is dispatched to), the counter of the groups' last
// arrivals, and the generation word n (barriers completed). The counters
// only ever grow: in barrier n the last arrival of a group of g workgroups
// is the one that takes its counter to (n + 1) g, so nothing is reset and
// every launch on one bar block must use the same grid size. Each wave
// waits for its own memory operations before the workgroup barrier;
// thread 0 then makes them visible with one agent-scope release fence,
// arrives with a relaxed add (the groups' last arrivals continue with
// acquire-release adds, which chain every arrival to the generation
// increment), polls n with relaxed loads and takes one acquire fence after
// the wait. Measured at 256 / 512 / 1024 workgroups: 6.6 / 9.9 / 17.5 us,
// against 22 / 40 / 82 us for one counter polled with acquire loads (each
// of which invalidates the XCD's L2) and 2.4 us for a kernel boundary.
__device__ bool grid_barrier(unsigned* bar, int* err, unsigned nblocks) {
  __shared__ int s_ok;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned* gen = bar + 9 * GRID_BAR_LINE;
    const unsigned n = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned grp = blockIdx.x & 7u;
    const unsigned gsize = (nblocks - grp + 7u) / 8u;
    const unsigned ngroups = nblocks < 8u ? nblocks : 8u;
    bool last = false;
    const unsigned prev = __hip_atomic_fetch_add(bar + grp * GRID_BAR_LINE, 1u,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev + 1u == (n + 1u) * gsize) {
      const unsigned p2 = __hip_atomic_fetch_add(bar + 8 * GRID_BAR_LINE, 1u, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_AGENT);
      if (p2 + 1u == (n + 1u) * ngroups) {
        __hip_atomic_fetch_add(gen, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = true;
      }
    }
    int ok = 1;
    if (!last) {
      long long spins = 0;
      while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == n) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > SPIN_LIMIT) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        if ((spins & 255) == 0 &&
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          ok = 0;
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Per-element work of the A phase: p_new at the five stencil points (z +
// beta p_old; p itself on the first iteration), A p_new at e, store p_new
// and Ap; returns p_new * Ap for the p.Ap sum.
__device__ __forceinline__ Two pcg_apply_elem(const PcgArgs& a, long long e, bool first,
                                              double bt, double* pnew, const double* pold) {
  const Dims& d = a.d;
  const HelmOp& op = a.op;
  const int i = (int)(e % d.nx), j = (int)(e / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i), cs = at2(d, wrap(j, -1, d.ny), i);
  auto pv = [&](long long q) -> double { return first ? pnew[q] : a.z[q] + bt * pold[q]; };
  const double pc = pv(e);
  const double mc = mk(op.mask[e]);
  const double xc = pc * mc;
  const double xe = pv(ce) * mk(op.mask[ce]);
  const double xw = pv(cw) * mk(op.mask[cw]);
  const double xn = pv(cn) * mk(op.mask[cn]);
  const double xs = pv(cs) * mk(op.mask[cs]);
  const double fx = op.Ku[e] * (xe - xc);
  const double fxw = op.Ku[cw] * (xc - xw);
  const double fy = op.Kv[e] * (xn - xc);
  const double fys = op.Kv[cs] * (xc - xs);
  const double lap = (fx - fxw) / op.dx2 + (fy - fys) / op.dy2;
  const double ap = (xc - op.coef * lap) * mc + xc * (1.0 - mc);
  if (!first) pnew[e] = pc;
  a.Ap[e] = ap;
  return Two{pc * ap, 0.0};
}

// Per-element work of the update phase: x += alpha p, r -= alpha Ap,
// z = M^-1 r; returns (r*r, r*z).
__device__ __forceinline__ Two pcg_update_elem(const PcgArgs& a, long long e, double al,
                                               const double* pnew) {
  a.x[e] = a.x[e] + al * pnew[e];
  const double r = a.r[e] - al * a.Ap[e];
  a.r[e] = r;
  const double z = a.inv_diag[e] * r;
  a.z[e] = z;
  return Two{r * r, r * z};
}

// Loads of one element of the A phase that do not depend on beta, issued
// before the scalar reductions so their latency overlaps them.
struct ApLoads {
  double z0, z1, z2, z3, z4, p0, p1, p2, p3, p4;  // centre, e, w, n, s
  double mc, me, mw, mn, ms, ku, kuw, kv, kvs;
};

__device__ __forceinline__ ApLoads ap_load(const PcgArgs& a, long long e, bool first,
                                           const double* pnew, const double* pold) {
  const Dims& d = a.d;
  const HelmOp& op = a.op;
  const int i = (int)(e % d.nx), j = (int)(e / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i), cs = at2(d, wrap(j, -1, d.ny), i);
  ApLoads L;
  const double* pp = first ? pnew : pold;
  L.p0 = pp[e];
  L.p1 = pp[ce];
  L.p2 = pp[cw];
  L.p3 = pp[cn];
  L.p4 = pp[cs];
  if (first) {
    L.z0 = L.z1 = L.z2 = L.z3 = L.z4 = 0.0;
  } else {
    L.z0 = a.z[e];
    L.z1 = a.z[ce];
    L.z2 = a.z[cw];
    L.z3 = a.z[cn];
    L.z4 = a.z[cs];
  }
  L.mc = mk(op.mask[e]);
  L.me = mk(op.mask[ce]);
  L.mw = mk(op.mask[cw]);
  L.mn = mk(op.mask[cn]);
  L.ms = mk(op.mask[cs]);
  L.ku = op.Ku[e];
  L.kuw = op.Ku[cw];
  L.kv = op.Kv[e];
  L.kvs = op.Kv[cs];
  return L;
}

// pcg_apply_elem from preloaded operands (same arithmetic).
__device__ __forceinline__ double ap_finish(const PcgArgs& a, long long e, bool first, double bt,
                                            double* pnew, const ApLoads& L) {
  const HelmOp& op = a.op;
  const double pc = first ? L.p0 : L.z0 + bt * L.p0;
  const double pe = first ? L.p1 : L.z1 + bt * L.p1;
  const double pw = first ? L.p2 : L.z2 + bt * L.p2;
  const double pn = first ? L.p3 : L.z3 + bt * L.p3;
  const double ps = first ? L.p4 : L.z4 + bt * L.p4;
  const double xc = pc * L.mc;
  const double xe = pe * L.me;
  const double xw = pw * L.mw;
  const double xn = pn * L.mn;
  const double xs = ps * L.ms;
  const double fx = L.ku * (xe - xc);
  const double fxw = L.kuw * (xc - xw);
  const double fy = L.kv * (xn - xc);
  const double fys = L.kvs * (xc - xs);
  const double lap = (fx - fxw) / op.dx2 + (fy - fys) / op.dy2;
  const double ap = (xc - op.coef * lap) * L.mc + xc * (1.0 - L.mc);
  if (!first) pnew[e] = pc;
  a.Ap[e] = ap;
  return pc * ap;
}

__device__ __forceinline__ int logical_block(int nb) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous range of parts so its L2 holds one slab of the vectors.
  const int w = blockIdx.x;
  if (nb % 8 != 0) return w;
  return (w % 8) * (nb / 8) + (w / 8);
}

__global__ __launch_bounds__(PCG_WG) void k_pcg_persistent(PcgArgs a) {
  __shared__ double lds[2 * PCG_STAGE];
  __shared__ double leafs[128];
  __shared__ double cs[PCG_MAX_CHUNKS];
  const unsigned nb = gridDim.x;
  const int part = logical_block((int)nb);
  const HelmOp& op = a.op;
  const Dims& d = a.d;
  bool ok = true;
  int it = 0, done = 0, conv = 1;
  double rz = 0.0, alpha = 0.0, beta = 0.0, tol = 0.0, res = 0.0;

  // r = rhs * mask, x = 0; ||rhs||
  {
    const Two s = part_sum<1>(a, part, [&](long long e) -> Two {
      const double r = a.rhs[e] * mk(op.mask[e]);
      a.r[e] = r;
      a.x[e] = 0.0;
      return Two{r * r, 0.0};
    }, lds, leafs);
    if (threadIdx.x == 0) a.part_a[part] = s.a;
  }
  ok = grid_barrier(a.bar, a.err, nb);
  if (ok) {
    const double nb2 = total_from_parts(a, a.part_a, cs);
    const double norm_b = sqrt(nb2);
    tol = a.rtol * norm_b;
    if (norm_b == 0.0) done = 1;
  }
  if (ok && !done) {
    // z = M^-1 r, p = z (into pb1, the buffer iteration 1 uses); r.z
    const Two s = part_sum<1>(a, part, [&](long long e) -> Two {
      const double z = a.inv_diag[e] * a.r[e];
      a.z[e] = z;
      a.pb1[e] = z;
      return Two{a.r[e] * z, 0.0};
    }, lds, leafs);
    if (threadIdx.x == 0) a.part_b[part] = s.a;
    ok = grid_barrier(a.bar, a.err, nb);
    if (ok) rz = total_from_parts(a, a.part_b, cs);
  }
  while (ok && !done) {
    it += 1;
    double* pnew = (it & 1) ? a.pb1 : a.pb0;
    const double* pold = (it & 1) ? a.pb0 : a.pb1;
    const bool first = it == 1;
    const double bt = beta;
    {
      const Two s = part_sum<1>(a, part, [&](long long e) -> Two {
        return pcg_apply_elem(a, e, first, bt, pnew, pold);
      }, lds, leafs);
      if (threadIdx.x == 0) a.part_a[part] = s.a;
    }
    ok = grid_barrier(a.bar, a.err, nb);
    if (!ok) break;
    alpha = rz / total_from_parts(a, a.part_a, cs);
    {
      const double al = alpha;
      const Two s = part_sum<2>(a, part, [&](long long e) -> Two {
        return pcg_update_elem(a, e, al, pnew);
      }, lds, leafs);
      if (threadIdx.x == 0) {
        a.part_b[part] = s.a;
        a.part_c[part] = s.b;
      }
    }
    ok = grid_barrier(a.bar, a.err, nb);
    if (!ok) break;
    res = sqrt(total_from_parts(a, a.part_b, cs));
    if (res <= tol) {
      done = 1;
      conv = 1;
      break;
    }
    const double rz_new = total_from_parts(a, a.part_c, cs);
    beta = rz_new / rz;
    rz = rz_new;
    if (it >= a.max_iter) {
      done = 1;
      conv = 0;
      break;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc->it = it;
    a.sc->iters = ok ? it : 0;
    a.sc->done = 1;
    a.sc->converged = ok ? conv : 0;
    a.sc->res = res;
    if (ok) {
      a.counters[0] += it;
      if (!conv) a.counters[1] += 1;
    }
  }
}


// ------------------------------------------- per-phase launches (path 1)
// The same algorithm and reductions as k_pcg_persistent in two launches per
// iteration. Every workgroup of a launch recomputes the scalar the previous
// phase's partial sums define (identically, so all agree); the host passes
// the iteration number, and rz is double-buffered by iteration parity, so
// no workgroup reads a scalar that another workgroup of the same launch
// writes.

// r = rhs * mask, x = 0; partial sums of r*r
__global__ __launch_bounds__(PCG_WG) void k_pcg_init(PcgArgs a) {
  __shared__ double lds[2 * PCG_STAGE];
  __shared__ double leafs[128];
  const int part = blockIdx.x;
  const Two s = part_sum<1>(a, part, [&](long long e) -> Two {
    const double r = a.rhs[e] * mk(a.op.mask[e]);
    a.r[e] = r;
    a.x[e] = 0.0;
    return Two{r * r, 0.0};
  }, lds, leafs);
  if (threadIdx.x == 0) a.part_a[part] = s.a;
}

// ||rhs|| (zero: done in 0 iterations); z = M^-1 r, p = z; partials of r*z
__global__ __launch_bounds__(PCG_WG) void k_pcg_init_z(PcgArgs a) {
  __shared__ double lds[2 * PCG_STAGE];
  __shared__ double leafs[128];
  __shared__ double cs[PCG_MAX_CHUNKS];
  const double nb = sqrt(total_from_parts(a, a.part_a, cs));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc->norm_b = nb;
    a.sc->tol = a.rtol * nb;
    a.sc->iters = 0;
    a.sc->converged = 1;
    a.sc->done = (nb == 0.0) ? 1 : 0;
    a.sc->res = 0.0;
  }
  if (nb == 0.0) return;
  const int part = blockIdx.x;
  const Two s = part_sum<1>(a, part, [&](long long e) -> Two {
    const double z = a.inv_diag[e] * a.r[e];
    a.z[e] = z;
    a.pb1[e] = z;
    return Two{a.r[e] * z, 0.0};
  }, lds, leafs);
  if (threadIdx.x == 0) a.part_b[part] = s.a;
}

// The scalars the next phase needs from the last phase's partial sums,
// formed once for the grid by one workgroup (grids with many chunks, where
// every workgroup forming them itself would repeat the in-order chunk sum
// thousands of times); the same reductions as in the phases, so the
// iterates do not change. which 0: tot[0] = sum part_a (p.Ap); 1: tot[1] =
// sum part_b (the first rz); 2: tot[1], tot[2] = sums of part_b (r.r) and
// part_c (r.z).
__global__ __launch_bounds__(PCG_WG) void k_pcg_totals(PcgArgs a, int which) {
  if (a.sc->done) return;
  __shared__ double cs[2 * PCG_MAX_CHUNKS];
  if (which == 0) {
    const double t = total_from_parts(a, a.part_a, cs);
    if (threadIdx.x == 0) a.sc->tot[0] = t;
  } else if (which == 1) {
    const double t = total_from_parts(a, a.part_b, cs);
    if (threadIdx.x == 0) a.sc->tot[1] = t;
  } else {
    const double* parts[2] = {a.part_b, a.part_c};
    double tot[2];
    totals_from_parts<2>(a, parts, cs, tot);
    if (threadIdx.x == 0) {
      a.sc->tot[1] = tot[0];
      a.sc->tot[2] = tot[1];
    }
  }
}

// Iteration `it` (1-based), phase A. For it > 1 it first finishes
// iteration it-1: the convergence test on ||r||, then beta and the new rz.
__global__ __launch_bounds__(PCG_WG) void k_pcg_ap(PcgArgs a, int it) {
  if (a.sc->done) return;
  __shared__ double lds[2 * PCG_STAGE];
  __shared__ double leafs[128];
  __shared__ double cs[2 * PCG_MAX_CHUNKS];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const bool first = it == 1;
  double* pnew = (it & 1) ? a.pb1 : a.pb0;
  const double* pold = (it & 1) ? a.pb0 : a.pb1;
  const int part = blockIdx.x;
  const PartRange pr = part_range(a, part);
  // operands of this part's (at most two per thread) elements first
  const int t0 = threadIdx.x, t1 = threadIdx.x + PCG_WG;
  const bool h0 = t0 < pr.nel, h1 = t1 < pr.nel;
  const ApLoads L0 = ap_load(a, pr.base + (h0 ? t0 : 0), first, pnew, pold);
  const ApLoads L1 = ap_load(a, pr.base + (h1 ? t1 : 0), first, pnew, pold);
  double beta = 0.0;
  if (first) {
    const double rz = a.central ? a.sc->tot[1] : total_from_parts(a, a.part_b, cs);
    if (lead) a.sc->rz2[1] = rz;
  } else {
    const double* parts[2] = {a.part_b, a.part_c};
    double tot[2];
    if (a.central) {
      tot[0] = a.sc->tot[1];
      tot[1] = a.sc->tot[2];
    } else {
      totals_from_parts<2>(a, parts, cs, tot);
    }
    const double res = sqrt(tot[0]);
    if (res <= a.sc->tol) {
      if (lead) {
        a.sc->res = res;
        a.sc->done = 1;
        a.sc->converged = 1;
        a.sc->iters = it - 1;
        a.counters[0] += it - 1;
      }
      return;
    }
    if (it - 1 >= a.max_iter) {
      if (lead) {
        a.sc->res = res;
        a.sc->done = 1;
        a.sc->converged = 0;
        a.sc->iters = a.max_iter;
        a.counters[0] += a.max_iter;
        a.counters[1] += 1;
      }
      return;
    }
    const double rz_new = tot[1];
    beta = rz_new / a.sc->rz2[(it - 1) & 1];
    if (lead) a.sc->rz2[it & 1] = rz_new;
  }
  if (h0) lds[t0] = ap_finish(a, pr.base + t0, first, beta, pnew, L0);
  if (h1) lds[t1] = ap_finish(a, pr.base + t1, first, beta, pnew, L1);
  const Two s = part_reduce<1>(pr, lds, leafs);
  if (threadIdx.x == 0) a.part_a[part] = s.a;
}

// Iteration `it`, phase B: alpha = rz / (p.Ap); x += alpha p; r -= alpha Ap;
// with the Jacobi preconditioner also z = inv_diag r and the partial sums
// of r*r and r*z, otherwise only r*r (the preconditioner follows in its own
// launches and then k_pcg_rz).
__global__ __launch_bounds__(PCG_WG) void k_pcg_update(PcgArgs a, int it, int jacobi) {
  if (a.sc->done) return;
  __shared__ double lds[2 * PCG_STAGE];
  __shared__ double leafs[128];
  __shared__ double cs[PCG_MAX_CHUNKS];
  const double* pnew = (it & 1) ? a.pb1 : a.pb0;
  const int part = blockIdx.x;
  const PartRange pr = part_range(a, part);
  // element operands first (independent of alpha); element base + t for
  // t = tid and tid + PCG_WG (a part has at most 2 * PCG_WG elements)
  const int t0 = threadIdx.x, t1 = threadIdx.x + PCG_WG;
  const bool h0 = t0 < pr.nel, h1 = t1 < pr.nel;
  const long long e0 = pr.base + (h0 ? t0 : 0), e1 = pr.base + (h1 ? t1 : 0);
  const double x0 = a.x[e0], p0 = pnew[e0], r0 = a.r[e0], ap0 = a.Ap[e0];
  const double x1 = a.x[e1], p1 = pnew[e1], r1 = a.r[e1], ap1 = a.Ap[e1];
  const double d0 = jacobi ? a.inv_diag[e0] : 0.0, d1 = jacobi ? a.inv_diag[e1] : 0.0;
  const double al =
      a.sc->rz2[it & 1] / (a.central ? a.sc->tot[0] : total_from_parts(a, a.part_a, cs));
  auto fin = [&](bool h, int t, long long e, double xv, double pv, double rv, double av,
                 double dv) {
    if (!h) return;
    a.x[e] = xv + al * pv;
    const double r = rv - al * av;
    a.r[e] = r;
    lds[t] = r * r;
    if (jacobi) {
      const double z = dv * r;
      a.z[e] = z;
      lds[PCG_STAGE + t] = r * z;
    }
  };
  fin(h0, t0, e0, x0, p0, r0, ap0, d0);
  fin(h1, t1, e1, x1, p1, r1, ap1, d1);
  if (jacobi) {
    const Two s = part_reduce<2>(pr, lds, leafs);
    if (threadIdx.x == 0) {
      a.part_b[part] = s.a;
      a.part_c[part] = s.b;
    }
  } else {
    const Two s = part_reduce<1>(pr, lds, leafs);
    if (threadIdx.x == 0) a.part_b[part] = s.a;
  }
}

// Partial sums of r*z after a preconditioner application; into_b = 1 on the
// first application (also p = z into pb1), 0 in the iterations (part_c).
__global__ __launch_bounds__(PCG_WG) void k_pcg_rz(PcgArgs a, int into_b) {
  if (a.sc->done) return;
  __shared__ double lds[2 * PCG_STAGE];
  __shared__ double leafs[128];
  const int part = blockIdx.x;
  const Two s = part_sum<1>(a, part, [&](long long e) -> Two {
    const double z = a.z[e];
    if (into_b) a.pb1[e] = z;
    return Two{a.r[e] * z, 0.0};
  }, lds, leafs);
  if (threadIdx.x == 0) (into_b ? a.part_b : a.part_c)[part] = s.a;
}

// After a sync-free solve of n_it launched iterations: a solve whose done
// flag is still clear did not converge within the budget; it is counted as
// a failure with n_it iterations.
__global__ void k_pcg_budget_check(PcgArgs a, int n_it) {
  if (threadIdx.x != 0 || a.sc->done) return;
  a.counters[0] += n_it;
  a.counters[1] += 1;
}

}  // namespace hoc
