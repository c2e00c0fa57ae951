// ndg_kernels.hip -- CDNA4 (gfx950) kernels of the MD step.
//
// Every per-atom kernel walks 128-atom tiles of one or many replicas with a persistent
// grid; workgroup b runs on XCD (b % 8) and takes tiles from that XCD's contiguous
// share, so neighbouring tiles (spatially neighbouring atoms after the cell sort) share
// an L2.  All reductions use a fixed tree, so results do not depend on how replicas are
// batched.  Built with -ffp-contract=off: FP64 expressions round in written order (the
// reference's order); FP32 hot loops use explicit fmaf.
#include "ndg_internal.h"

#include <math.h>

namespace ndg {

// List entries the force kernel reads per lane and iteration.  The list builder pads every
// atom's list with absent entries up to a whole number of such groups per lane, so the
// kernel reads them without a bounds select.
constexpr int kUnroll = 4;

// ------------------------------------------------------------------ helpers

__device__ inline void tile_range(int t0, int nt, int &begin, int &end, int &step) {
  const int nb = gridDim.x, b = blockIdx.x;
  if (nb < 8) {
    begin = t0 + b;
    end = t0 + nt;
    step = nb;
    return;
  }
  const int x = b & 7;
  const int nwx = (nb - x + 7) >> 3;
  const long long lo = (long long)nt * x / 8, hi = (long long)nt * (x + 1) / 8;
  begin = t0 + (int)lo + (b >> 3);
  end = t0 + (int)hi;
  step = nwx;
}

// Global-memory view of a pointer read from the parameter block (global_* instead of
// flat_* instructions).
template <class T>
__device__ inline __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// A pointer that is the same in every lane, moved to scalar registers.
template <class T>
__device__ inline T *uniform_ptr(T *p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (T *)(((unsigned long long)hi << 32) | lo);
}

__device__ inline double shfl_xor_d(double v, int m) {
  int2 a = *reinterpret_cast<int2 *>(&v);
  a.x = __shfl_xor(a.x, m);
  a.y = __shfl_xor(a.y, m);
  return *reinterpret_cast<double *>(&a);
}

// Sum over the 64 lanes of a wave (xor tree); the result is in every lane.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_d(v, m);
  return v;
}

// Deterministic block sum of NV doubles per thread; result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += shfl_xor_d(v[q], m);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[wid * NV + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w)
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] += sh[w * NV + q];
  }
  __syncthreads();
}

__device__ inline void top2_merge(double &a1, double &a2, double b1, double b2) {
  double hi = fmax(a1, b1), lo = fmin(a1, b1);
  a2 = fmax(lo, fmax(a2, b2));
  a1 = hi;
}

// Block sums of NV doubles per thread and a block top-2 in one LDS exchange; valid in
// thread 0 (the same order of additions as block_sum).  The caller synchronises before
// sh is written again.
template <int NV>
__device__ inline void block_sum_top2(double (&v)[NV], double &t1, double &t2, double *sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += shfl_xor_d(v[q], m);
    const double b1 = shfl_xor_d(t1, m), b2 = shfl_xor_d(t2, m);
    top2_merge(t1, t2, b1, b2);
  }
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[wid * (NV + 2) + q] = v[q];
    sh[wid * (NV + 2) + NV] = t1;
    sh[wid * (NV + 2) + NV + 1] = t2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] += sh[w * (NV + 2) + q];
      top2_merge(t1, t2, sh[w * (NV + 2) + NV], sh[w * (NV + 2) + NV + 1]);
    }
  }
}

// Block top-2 (largest, second largest); valid in thread 0.
__device__ inline void block_top2(double &t1, double &t2, double *sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    double b1 = shfl_xor_d(t1, m), b2 = shfl_xor_d(t2, m);
    top2_merge(t1, t2, b1, b2);
  }
  __syncthreads();
  if (lane == 0) {
    sh[2 * wid] = t1;
    sh[2 * wid + 1] = t2;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < nw; ++w) top2_merge(t1, t2, sh[2 * w], sh[2 * w + 1]);
  __syncthreads();
}

// Orders one wave's LDS writes before its later LDS reads (no block barrier).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int active_flags(const EngineDev &E, int rep, int xflags) {
  if (xflags) return xflags;
  return step_flags(E.reps[rep], E.st[rep]);
}

// Cell-anchored coordinate (A) of a scaled position: ((p+1/2) M - c - 1/2) w, |x| <~ w/2.
// Pair vectors are x_i - x_j + o w with a per-pair integer cell/image offset o.
__device__ inline double anchored(double p, int c, int m, double w) {
  return ((p + 0.5) * (double)m - (double)c - 0.5) * w;
}

__device__ inline double type_bits(int t) { return __longlong_as_double((long long)t); }
__device__ inline int bits_type(double w) { return (int)__double_as_longlong(w); }

// 16-byte form of a cell-anchored position, read by the mixed force kernel (one gather
// per neighbour instead of two).  Each coordinate x is the 32-bit integer n = rint(x / q)
// with q = 2^(e-30), where 2^e >= w/2 + skin + 1 A bounds |x| between rebuilds, so
// |n| < 2^30 and a difference of two never overflows.  Word 3 holds type - 1.  The pair
// displacement is then (n_i - n_j) q + offset, exact but for the rounding of x to q/2
// (~4e-9 A for Cu cells).
__device__ inline double xq_unit(const RepState &S, double skin) {
  double w = 0.0;
  for (int a = 0; a < 3; ++a) w = fmax(w, S.box[a] / S.mcell[a]);
  int e;
  frexp(0.5 * w + skin + 1.0, &e);   // 2^(e-1) <= bound < 2^e
  return ldexp(1.0, e - 30);
}

constexpr double kXqLimit = 1073741823.0;   // 2^30 - 1

// An anchored coordinate the packed record cannot hold (|x| / q beyond 2^30 - 1): only an
// atom that moved more than skin + 1 A from its list-build position without a rebuild.
__device__ inline bool xq_out_of_range(double x, double y, double z, double q) {
  return fmax(fabs(x), fmax(fabs(y), fabs(z))) > kXqLimit * q;
}

// Values beyond the range are clamped (so the integer differences never overflow); the
// drift kernel flags NDG_ERR_RANGE for them.
__device__ inline uint4 xq_pack(double x, double y, double z, int ty, double q) {
  const double lim = kXqLimit;
  const double iq = 1.0 / q;         // a power of two: exact
  const int nx = (int)fmin(lim, fmax(-lim, rint(x * iq)));
  const int ny = (int)fmin(lim, fmax(-lim, rint(y * iq)));
  const int nz = (int)fmin(lim, fmax(-lim, rint(z * iq)));
  return make_uint4((unsigned)nx, (unsigned)ny, (unsigned)nz, (unsigned)(ty - 1));
}

// Reference binning (src/nd.f90:1240-1246): sr = p - anint(p); I = int((sr+0.4999) M).
__device__ inline void bin_axis(double p, int m, int brute, int &k, int &c) {
  double n = round(p);
  if (brute) {
    k = 0;
    c = (int)n;
    return;
  }
  double sr = p - n;
  int kk = (int)((sr + 0.4999) * (double)m);
  kk = kk < 0 ? 0 : (kk >= m ? m - 1 : kk);
  k = kk;
  c = (int)n * m + kk;
}

// ------------------------------------------------------------------ integrator (mixed)

// integrate1 (src/nd.f90:1029-1046) with the pending thermostat scale of the previous
// step folded in; writes the packed cell-anchored copy (xq) and the displacement top-2.
// integrate1's displacement conditions (src/nd.f90:1055-1064): a held atom whose hold
// group has Disp moves by the prescribed scaled displacement along the constrained axes
__device__ inline void hold_displacement(const RepParams &R, int i, double &du0, double &du1,
                                         double &du2) {
  const int b = R.hbc[i];
  if (b > 0) {
    const int d = R.bc_dir[b - 1];
    if (d & 1) du0 = R.bc_du[b - 1][0];
    if (d & 2) du1 = R.bc_du[b - 1][1];
    if (d & 4) du2 = R.bc_du[b - 1][2];
  }
}

// A value read or written once per step; NT: past the last-level cache.
template <bool NT, class T>
__device__ inline T stream_load(const T *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT, class T>
__device__ inline void stream_store(T v, T *p) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
typedef int int4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

// NT: the state is larger than the last-level cache; the positions, anchors and build-time
// copies, which this kernel alone touches each step, are then streamed past it (0.5-1.5 %
// of a step on the 2M-atom workloads), and the velocities and forces that the force
// kernel's kick reads next are left to it.
template <bool NT>
__global__ __launch_bounds__(kTile) void k_drift_mixed(EngineDev E, int t0, int nt, int xflags) {
  __shared__ double sh[8];
  int tb, te, ts;
  tile_range(t0, nt, tb, te, ts);
  for (int t = tb; t < te; t += ts) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    const int f = active_flags(E, T.rep, xflags);
    if (!(f & F_ACTIVE)) continue;
    const int i = T.atom0 + threadIdx.x;
    double t1 = -1.0, t2 = -1.0;
    if (i < R.np) {
      const double dt = R.dt;
      const double l1 = S.lam1, l2 = S.lam2;
      double vx = R.vx[i], vy = R.vy[i], vz = R.vz[i];
      if (l1 != 1.0 || l2 != 1.0) {
        vx = vx * l1; vy = vy * l1; vz = vz * l1;
        vx = vx * l2; vy = vy * l2; vz = vz * l2;
        R.vx[i] = vx; R.vy[i] = vy; R.vz[i] = vz;
      }
      const float4 Fo = (S.parity ? R.f32b : R.f32a)[i];
      const int ty = R.typ[i];
      const double rm = 1.0 / R.mass[ty - 1];
      const double ax = (double)Fo.x * rm, ay = (double)Fo.y * rm, az = (double)Fo.z * rm;
      const double li0 = 1.0 / S.box[0], li1 = 1.0 / S.box[1], li2 = 1.0 / S.box[2];
      double du0 = li0 * (vx * dt + 0.5 * ax * dt * dt);
      double du1 = li1 * (vy * dt + 0.5 * ay * dt * dt);
      double du2 = li2 * (vz * dt + 0.5 * az * dt * dt);
      if (R.nbc) hold_displacement(R, i, du0, du1, du2);
      const double px = stream_load<NT>(&R.px[i]) + du0, py = stream_load<NT>(&R.py[i]) + du1,
                   pz = stream_load<NT>(&R.pz[i]) + du2;
      stream_store<NT>(px, &R.px[i]);
      stream_store<NT>(py, &R.py[i]);
      stream_store<NT>(pz, &R.pz[i]);
      const int4v Av = stream_load<NT>((const int4v *)&R.anchor[i]);
      const int4 A = make_int4(Av.x, Av.y, Av.z, Av.w);
      const int m0 = S.mcell[0], m1 = S.mcell[1], m2 = S.mcell[2];
      const double xl = anchored(px, A.x, m0, S.box[0] / m0);
      const double yl = anchored(py, A.y, m1, S.box[1] / m1);
      const double zl = anchored(pz, A.z, m2, S.box[2] / m2);
      // x64a (FP64 anchored) is written only at a rebuild, for the list builder; the
      // force kernel reads the packed copy
      const double qu = xq_unit(S, R.skin);
      R.xq[i] = xq_pack(xl, yl, zl, ty, qu);
      // the run stops with NDG_ERR_RANGE rather than compute pairs from clamped records
      if (xq_out_of_range(xl, yl, zl, qu)) E.st[T.rep].err = NDG_ERR_RANGE;
      const float4v Bv = stream_load<NT>((const float4v *)&R.x32b[i]);
      const float4 B = make_float4(Bv.x, Bv.y, Bv.z, Bv.w);
      const double dx = (double)((float)xl - B.x), dy = (double)((float)yl - B.y),
                   dz = (double)((float)zl - B.z);
      if (R.trigger == NDG_TRIGGER_COMPAT) {
        const double sx = dx * li0, sy = dy * li1, sz = dz * li2;
        t1 = sx * sx + sy * sy + sz * sz;
      } else {
        t1 = dx * dx + dy * dy + dz * dz;
      }
    }
    block_top2(t1, t2, sh);
    if (threadIdx.x == 0) {
      R.partials[kPartD1 * R.ntiles + T.ltile] = t1;
      R.partials[kPartD2 * R.ntiles + T.ltile] = t2;
    }
  }
}

// integrate2 (src/nd.f90:1091-1117): kick with the new force, kinetic energy partials (one
// per wave of the tile; the slots of a TPA > 1 layout beyond the first two are cleared).
__global__ __launch_bounds__(kTile) void k_kick_mixed(EngineDev E, int t0, int nt, int xflags) {
  int tb, te, ts;
  tile_range(t0, nt, tb, te, ts);
  for (int t = tb; t < te; t += ts) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    const int f = active_flags(E, T.rep, xflags);
    if (!(f & F_ACTIVE)) continue;
    const int i = T.atom0 + threadIdx.x;
    double kt[1] = {0.0};
    if (i < R.np) {
      const float4 Fo = (S.parity ? R.f32b : R.f32a)[i];
      const float4 Fn = (S.parity ? R.f32a : R.f32b)[i];
      const int ty = R.typ[i];
      const double m = R.mass[ty - 1];
      const double rm = 1.0 / m;
      const double hdt = 0.5 * R.dt;
      double vx = R.vx[i] + hdt * ((double)Fn.x * rm + (double)Fo.x * rm);
      double vy = R.vy[i] + hdt * ((double)Fn.y * rm + (double)Fo.y * rm);
      double vz = R.vz[i] + hdt * ((double)Fn.z * rm + (double)Fo.z * rm);
      R.vx[i] = vx; R.vy[i] = vy; R.vz[i] = vz;
      if (f & F_CK) {
        double k = vx * vx;
        k = k + vy * vy;
        k = k + vz * vz;
        kt[0] = k * m;
      }
    }
    if (f & F_CK) {
      const double w = wave_sum(kt[0]);
      const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
      if (lane == 0) R.partials[(kPartKeW + wid) * R.ntiles + T.ltile] = w;
      if (lane >= 2 && lane < 2 * R.tpa && wid == 0)
        R.partials[(kPartKeW + lane) * R.ntiles + T.ltile] = 0.0;
    }
  }
}

// ------------------------------------------------------------------ force (mixed)

// compute0 (src/nd.f90:262-333) on a full list, mixed precision.  Pair geometry in FP64
// on cell-anchored coordinates (no rounding of minimum images), so the cutoff test and
// the single-precision table index rk are the reference's own; table values, the linear
// interpolation and 1/d2 in FP32; per-atom force, energy and virial sums in FP64 (full
// rate on CDNA).  The distinct FP32 force tables of all replicas are staged in LDS once
// per persistent workgroup (two-dword LDS read per pair); otherwise {T_k, T_k+1} pairs
// are read from global memory (same arithmetic, same result).  Each lane walks U list entries per iteration with
// every load issued first and the cutoff as a select.  A workgroup of 512 (1024 for
// TPA 8) threads runs 512/(128 TPA) tiles at a time; TPA lanes share one atom and split
// its list; lane sums combine in a fixed order.
template <int TPA>
struct ForceShape {
  static constexpr int kSub = kTile * TPA;                 // threads per tile
  static constexpr int kThreads = TPA >= 8 ? 1024 : 512;   // threads per workgroup
  static constexpr int kNSub = kThreads / kSub;            // tiles per pass
};

// A list entry; NT: past the last-level cache (the lists are streamed once per step and
// would otherwise evict the neighbour records that the gathers reuse).
template <bool NT>
__device__ inline unsigned list_load(const __attribute__((address_space(1))) unsigned *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

template <int TPA, bool FULL, bool NT, bool LDS_TAB, bool ONE, bool ASYM = false,
          bool SPRING = false>
__device__ inline void force_atom(const RepParams &R, int i, int sub, const double4 *shift,
                                  const double *rc2, const float *vdr2, const int *pm,
                                  const int *lt, const float *stab, int ptoff, const unsigned *L,
                                  double qunit, double &fx, double &fy, double &fz,
                                  double (&acc)[7], const double *ibox = nullptr) {
  // the energy/virial variant holds seven more sums: an unroll of 2 keeps it at four
  // waves per SIMD (1.37 vs 1.52 ms per launch on 2M atoms)
  constexpr int U = FULL ? 2 : kUnroll;
  static_assert(U <= kListPad, "the list holds kListPad spare chunks for the unrolled reads");
  // the replica's arrays are the same for the whole wave (a wave never spans two tiles)
  const auto XQ = gmem((const unsigned *)uniform_ptr(R.xq));
  const auto Lg = gmem(L);
  const auto TF = gmem((const float *)R.tabF2);
  const auto TP = gmem((const float *)R.tabP2);
  const int nix = (int)XQ[4 * i], niy = (int)XQ[4 * i + 1], niz = (int)XQ[4 * i + 2];
  const int ti = (int)XQ[4 * i + 3] + 1;
  const int n = gmem(R.nnb)[i];
  const int nmine = n > sub ? (n - sub + TPA - 1) / TPA : 0;  // entries of this lane
  const int tab = R.ptsize + 1;
  const float kmaxf = (float)R.ptsize;
  const double ptrmin = R.ptrmin;
  // single pair potential: its parameters stay in registers
  const double rc2_0 = rc2[0];
  const float vdr2_0 = vdr2[0];
  const int lt_0 = LDS_TAB ? lt[0] : 0;
  // absent entries read the atom itself with the far offset of code 63, so the cutoff
  // select drops them
  const unsigned pad = (unsigned)i | (63u << kCodeShift);
  // the list entries of the next iteration are loaded one iteration ahead, so only the
  // position gathers stand between an iteration's loads and its arithmetic; every list is
  // padded with absent entries to a whole number of kUnroll chunks per lane (U divides
  // kUnroll), so the entries of an iteration need no bounds select
  static_assert(kUnroll % U == 0, "lists are padded to whole groups of kUnroll chunks");
  unsigned nxt[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
    nxt[u] = nmine > 0 ? list_load<NT>(Lg + (size_t)u * (kTile * TPA)) : pad;
  for (int c = 0; c < nmine; c += U) {
    unsigned e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) e[u] = nxt[u];
    uint4 pj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const unsigned j = e[u] & kIndexMask;
      pj[u] = make_uint4(XQ[4 * j], XQ[4 * j + 1], XQ[4 * j + 2], XQ[4 * j + 3]);
    }
    if (c + U < nmine) {
#pragma unroll
      for (int u = 0; u < U; ++u) nxt[u] = list_load<NT>(Lg + (size_t)(c + U + u) * (kTile * TPA));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double4 sft = shift[e[u] >> kCodeShift];
      const double dx = fma((double)(nix - (int)pj[u].x), qunit, sft.x);
      const double dy = fma((double)(niy - (int)pj[u].y), qunit, sft.y);
      const double dz = fma((double)(niz - (int)pj[u].z), qunit, sft.z);
      const double d2 = fma(dx, dx, fma(dy, dy, dz * dz));
      int pid = 0;
      double rc2p = rc2_0;
      float vd = vdr2_0;
      int ltp = lt_0;
      bool feel = true, bal = true;   // ASYM: the neighbour acts on this atom; balanced pair
      if (!ONE) {
        const int pij = pm[(ti - 1) + (int)pj[u].w * 8];
        pid = pij - 1;
        if (ASYM) {   // compute1 (src/nd.f90:445-498): the larger id of the two directions
          const int pji = pm[(int)pj[u].w + (ti - 1) * 8];
          pid = (pij > pji ? pij : pji) - 1;
          feel = pji > 0;
          bal = pij == pji;
        }
        // padding entries (code 63) read the atom itself, whose type may have no pair with
        // itself (id 0): keep the parameter and table reads in bounds (d2 masks the entry)
        pid = pid < 0 ? 0 : pid;
        rc2p = rc2[pid];
        vd = vdr2[pid];
        if (LDS_TAB) ltp = lt[pid];
      }
      bool ok = !(d2 > rc2p) && feel;
      double dsp = 0.0, psp = 0.0;   // SPRING: force over r and pair energy of a spring pair
      const bool spr = SPRING && ((R.spring_mask >> pid) & 1);
      if (spr) {
        // compute2/3 spring (src/nd.f90:622-660): strain against the initial positions in
        // FP64, no distance cutoff, nothing past the strain limit PairParams(3)
        const int j = (int)(e[u] & kIndexMask);
        double u0 = R.ipx[i] - R.ipx[j], u1 = R.ipy[i] - R.ipy[j], u2 = R.ipz[i] - R.ipz[j];
        u0 = u0 - round(u0);
        u1 = u1 - round(u1);
        u2 = u2 - round(u2);
        const double q0 = ibox[0] * u0, q1 = ibox[1] * u1, q2 = ibox[2] * u2;
        const double rr = sqrt(d2), rr0 = sqrt(fma(q0, q0, fma(q1, q1, q2 * q2)));
        const double sij = rr / rr0 - 1.0;
        ok = feel && (e[u] >> kCodeShift) != 63u && !(sij > R.spr_s0[pid]);
        const double Vi = R.spr_vi[pid];
        const double fs = R.spr_c[pid] * sij * Vi * Vi;
        dsp = ok ? -fs / rr : 0.0;
        psp = 0.5 * fs * sij * rr0;
      }
      // the reference's table position, in its single-precision operations
      // (src/nd.f90:299): rk = real(d2 - PTrMin) * vPTdr2 + 1
      const float rk = (float)(d2 - ptrmin) * vd + 1.0f;
      // k = clamp(int(rk), 1, kmax) by a median in FP32, then the fraction
      const float kf = floorf(__builtin_amdgcn_fmed3f(rk, 1.0f, kmaxf));
      const int k = (int)kf;
      const float w = rk - kf;
      float fv;
      if (LDS_TAB) {
        const float *Tm = stab + (ltp - 1);   // Tm[k] = T_{k-1}: address k * 4 + base
        const float F0 = Tm[k], F1 = Tm[k + 1];
        fv = fmaf(w, F1 - F0, F0);
      } else {
        const int q = 2 * (pid * tab + k - 1);
        const float F0 = TF[q], F1 = TF[q + 1];
        fv = fmaf(w, F1 - F0, F0);
      }
      const double df = spr ? dsp : (double)(ok ? fv * __builtin_amdgcn_rcpf((float)d2) : 0.0f);
      fx = fma(dx, df, fx);
      fy = fma(dy, df, fy);
      fz = fma(dz, df, fz);
      if (FULL) {
        float P0, P1;
        if (LDS_TAB && ptoff >= 0) {   // P half of the staged table set (uniform choice)
          const float *Pm = stab + (ptoff + ltp - 1);
          P0 = Pm[k];
          P1 = Pm[k + 1];
        } else {
          const int q = 2 * (pid * tab + k - 1);
          P0 = TP[q];
          P1 = TP[q + 1];
        }
        const double pv = ok && bal ? (spr ? psp : (double)fmaf(w, P1 - P0, P0)) : 0.0;
        // energy and virial sums without their factor 1/2, applied once per atom (an
        // exact scaling); an unbalanced pair of an asymmetric matrix adds to neither
        const double dv = !ASYM || bal ? df : 0.0;
        const double gx = dx * dv, gy = dy * dv, gz = dz * dv;
        acc[0] += pv;
        acc[1] = fma(dx, gx, acc[1]);
        acc[2] = fma(dy, gy, acc[2]);
        acc[3] = fma(dz, gz, acc[3]);
        acc[4] = fma(dx, gy, acc[4]);
        acc[5] = fma(dx, gz, acc[5]);
        acc[6] = fma(dy, gz, acc[6]);
      }
    }
  }
}

// kick != 0: integrate2 follows in the same pass (v += dt/2 (F_new + F_old)/m, kinetic
// energy partials), with the arithmetic and the summation order of k_kick_mixed.
// The force-only instances fit five waves per SIMD (at most 96 VGPRs, no spills) and run
// ~2 % faster than at four; the other instances keep the registers they need.
template <int TPA, bool FULL, bool SPR, bool NT>
__global__ __launch_bounds__(ForceShape<TPA>::kThreads)
__attribute__((amdgpu_waves_per_eu(FULL || SPR || TPA > 1 ? 1 : 5))) void k_force_mixed(EngineDev E, int t0,
                                                                           int nt, int xflags,
                                                                           int kick) {
  using SH = ForceShape<TPA>;
  constexpr int NSUB = SH::kNSub, SUB = SH::kSub, NW = SH::kThreads / 64;
  extern __shared__ float sh_tab[];
  __shared__ double4 sh_shift[NSUB][64];
  __shared__ double sh_rc2[NSUB][16];
  __shared__ float sh_vdr2[NSUB][16];
  __shared__ int sh_pm[NSUB][64];
  __shared__ int sh_lt[NSUB][16];
  __shared__ int sh_uselds[NSUB];
  __shared__ double sh_red[NW][8];
  __shared__ int sh_flags[NSUB], sh_act[NSUB];
  __shared__ double sh_qu[NSUB];
  // the F half of the table set; the energy/virial variant also the P half when it fits
  const int ptoff = FULL && E.tabset_p ? E.tabset_n : -1;
  const int nstage = ptoff >= 0 ? 2 * E.tabset_n : E.tabset_n;
  for (int q = threadIdx.x; q < nstage; q += blockDim.x) sh_tab[q] = E.tabset[q];
  // CACHE (TPA >= 2): the sub-group index is made wave-uniform (a wave never spans two
  // sub-groups), so tile and replica reads use scalar addresses, and a replica's step
  // flags and record unit are evaluated once, when its parameters are staged, instead of
  // by a chain of dependent reads in every pass (c2 force launch 0.752 -> 0.717 ms, c1
  // 0.0505 -> 0.0480 ms).  The TPA 1 instance keeps the per-pass form: made uniform it
  // ran 1 % slower on the replica mix, and the LDS copies spill at its 96-VGPR budget.
  constexpr bool CACHE = TPA >= 2;
  const int sg = CACHE ? __builtin_amdgcn_readfirstlane(threadIdx.x / SUB) : threadIdx.x / SUB;
  const int lt = threadIdx.x % SUB;
  const int a = lt / TPA, sub = lt % TPA;
  const int nblk = (nt + NSUB - 1) / NSUB;
  int bb, be, bs;
  tile_range(0, nblk, bb, be, bs);
  // replica whose parameters each sub-group holds in LDS (box, grid and step counter do
  // not change during a launch, so a sub-group keeps them while its tiles stay in one
  // replica)
  int held[NSUB];
#pragma unroll
  for (int q = 0; q < NSUB; ++q) held[q] = -1;
  for (int blk = bb; blk < be; blk += bs) {
    const int t = t0 + blk * NSUB + sg;
    bool have = t < t0 + nt;
    Tile T = have ? E.tiles[t] : Tile{0, 0, 0, 0};
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    int f = 0;
    if (!CACHE && have) {
      f = active_flags(E, T.rep, xflags);
      const bool full = (f & (F_CU | F_CV)) != 0;
      have = (f & F_ACTIVE) && R.tpa == TPA && full == FULL;
    }
    // the same decision in every thread: the tiles of this pass, read by all
    bool renew = false;
#pragma unroll
    for (int q = 0; q < NSUB; ++q) {
      const int tq = t0 + blk * NSUB + q;
      const int rq = tq < t0 + nt ? E.tiles[tq].rep : -1;
      renew = renew || (rq != held[q]);
      held[q] = rq;
    }
    if (renew) __syncthreads();
    // written for every existing tile, active or not, so that `held` stays true
    if (renew && t < t0 + nt && lt < 64) {
      const int c = lt;
      // codes with a field of 3 never occur in a list; 63 marks an absent entry and
      // places it far beyond every cutoff
      sh_shift[sg][c] = c == 63 ? make_double4(1e30, 0.0, 0.0, 0.0)
                                : make_double4((double)((c & 3) - 1) * (S.box[0] / S.mcell[0]),
                                               (double)(((c >> 2) & 3) - 1) * (S.box[1] / S.mcell[1]),
                                               (double)(((c >> 4) & 3) - 1) * (S.box[2] / S.mcell[2]), 0.0);
      if (c < R.npairs) {
        sh_rc2[sg][c] = R.rc2d[c];
        sh_vdr2[sg][c] = R.vdr2[c];
        sh_lt[sg][c] = R.ltab[c];
      }
      // PairMat with a row stride of 8 (types are at most 8): indexed by shifts
      {
        const int ta = c & 7, tb = c >> 3;
        sh_pm[sg][c] = ta < R.ntypes && tb < R.ntypes ? R.pairmat[ta + tb * R.ntypes] : 0;
      }
      if (CACHE && c == 0) {
        const int fl = active_flags(E, T.rep, xflags);
        const bool full = (fl & (F_CU | F_CV)) != 0;
        sh_flags[sg] = fl;
        sh_act[sg] = (fl & F_ACTIVE) && R.tpa == TPA && full == FULL;
        sh_qu[sg] = xq_unit(S, R.skin);
      }
      if (c == 0) {
        int use = E.tabset_n > 0;
        for (int p = 0; p < R.npairs; ++p) use = use && R.ltab[p] >= 0;
        sh_uselds[sg] = use;
      }
    }
    if (renew) __syncthreads();
    if (CACHE && have) {
      f = sh_flags[sg];
      have = sh_act[sg] != 0;
    }
    const int i = T.atom0 + a;
    const double qunit = have ? (CACHE ? sh_qu[sg] : xq_unit(S, R.skin)) : 1.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // e_i, vir xx yy zz xy xz yz (FULL)
    if (have && i < R.np) {
      const unsigned *L = R.list + ((size_t)T.ltile * R.list_chunks) * (kTile * TPA) + lt;
      // both choices are the same for every wave of the tile
      const bool asym = R.asym != 0;
      // one pair potential (or several identical ones covering every type pair): its
      // parameters stay in registers
      const bool lds = sh_uselds[sg] != 0,
                 one = (R.npairs == 1 || R.uniform) && !asym && !R.spring_mask;
      // SPR: the instance launched when some replica has spring pairs (the spring
      // variant's registers stay out of the common instance)
      if (SPR && R.spring_mask)
        force_atom<TPA, FULL, NT, false, false, true, true>(R, i, sub, sh_shift[sg], sh_rc2[sg],
                                                        sh_vdr2[sg], sh_pm[sg], sh_lt[sg], sh_tab,
                                                        ptoff, L, qunit, fx, fy, fz, acc,
                                                        E.st[T.rep].ibox);
      else if (asym && lds)
        force_atom<TPA, FULL, NT, true, false, true>(R, i, sub, sh_shift[sg], sh_rc2[sg], sh_vdr2[sg],
                                                 sh_pm[sg], sh_lt[sg], sh_tab, ptoff, L, qunit, fx, fy, fz, acc);
      else if (asym)
        force_atom<TPA, FULL, NT, false, false, true>(R, i, sub, sh_shift[sg], sh_rc2[sg], sh_vdr2[sg],
                                                  sh_pm[sg], sh_lt[sg], sh_tab, ptoff, L, qunit, fx, fy, fz, acc);
      else if (lds && one)
        force_atom<TPA, FULL, NT, true, true>(R, i, sub, sh_shift[sg], sh_rc2[sg], sh_vdr2[sg],
                                          sh_pm[sg], sh_lt[sg], sh_tab, ptoff, L, qunit, fx, fy, fz, acc);
      else if (lds)
        force_atom<TPA, FULL, NT, true, false>(R, i, sub, sh_shift[sg], sh_rc2[sg], sh_vdr2[sg],
                                           sh_pm[sg], sh_lt[sg], sh_tab, ptoff, L, qunit, fx, fy, fz, acc);
      else if (one)
        force_atom<TPA, FULL, NT, false, true>(R, i, sub, sh_shift[sg], sh_rc2[sg], sh_vdr2[sg],
                                           sh_pm[sg], sh_lt[sg], sh_tab, ptoff, L, qunit, fx, fy, fz, acc);
      else
        force_atom<TPA, FULL, NT, false, false>(R, i, sub, sh_shift[sg], sh_rc2[sg], sh_vdr2[sg],
                                            sh_pm[sg], sh_lt[sg], sh_tab, ptoff, L, qunit, fx, fy, fz, acc);
      if (FULL)
#pragma unroll
        for (int q = 0; q < 7; ++q) acc[q] *= 0.5;
    }
    if (TPA > 1) {
#pragma unroll
      for (int m = 1; m < TPA; m <<= 1) {
        fx += shfl_xor_d(fx, m);
        fy += shfl_xor_d(fy, m);
        fz += shfl_xor_d(fz, m);
        if (FULL)
#pragma unroll
          for (int q = 0; q < 7; ++q) acc[q] += shfl_xor_d(acc[q], m);
      }
    }
    const float4 Fn = make_float4((float)fx, (float)fy, (float)fz, FULL ? (float)acc[0] : 0.f);
    if (have && i < R.np && sub == 0) {
      float4 *Fout = S.parity ? R.f32a : R.f32b;
      Fout[i] = Fn;
    }
    if (kick) {
      double ke = 0.0;
      if (have && i < R.np && sub == 0) {
        const float4 Fo = (S.parity ? R.f32b : R.f32a)[i];
        const int ty = R.typ[i];
        const double m = R.mass[ty - 1];
        const double rm = 1.0 / m;
        const double hdt = 0.5 * R.dt;
        double vx = R.vx[i] + hdt * ((double)Fn.x * rm + (double)Fo.x * rm);
        double vy = R.vy[i] + hdt * ((double)Fn.y * rm + (double)Fo.y * rm);
        double vz = R.vz[i] + hdt * ((double)Fn.z * rm + (double)Fo.z * rm);
        R.vx[i] = vx; R.vy[i] = vy; R.vz[i] = vz;
        double k = vx * vx;
        k = k + vy * vy;
        k = k + vz * vz;
        ke = k * m;
      }
      // a wave lies inside one tile, so the condition is the same in all its lanes
      if (have && (f & F_CK)) {
        const double w = wave_sum(ke);
        if ((lt & 63) == 0) R.partials[(kPartKeW + (lt >> 6)) * R.ntiles + T.ltile] = w;
      }
    }
    if (FULL) {
      // fixed-order sum over the tile's waves
      if (!have || sub != 0 || i >= R.np)
#pragma unroll
        for (int q = 0; q < 7; ++q) acc[q] = 0.0;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int q = 0; q < 7; ++q) acc[q] += shfl_xor_d(acc[q], m);
      __syncthreads();   // the previous pass has read sh_red
      if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 7; ++q) sh_red[threadIdx.x >> 6][q] = acc[q];
      __syncthreads();
      if (have && lt == 0) {
        const int w0 = threadIdx.x >> 6;
        for (int q = 0; q < 7; ++q) {
          double v = 0.0;
          for (int w = 0; w < SUB / 64; ++w) v += sh_red[w0 + w][q];
          R.partials[(kPartPe + q) * R.ntiles + T.ltile] = v;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ FP64 plain path

__global__ __launch_bounds__(kTile) void k_drift_fp64(EngineDev E, int t0, int nt, int xflags) {
  __shared__ double sh[8];
  int tb, te, ts;
  tile_range(t0, nt, tb, te, ts);
  for (int t = tb; t < te; t += ts) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    const int f = active_flags(E, T.rep, xflags);
    if (!(f & F_ACTIVE)) continue;
    const int i = T.atom0 + threadIdx.x;
    double t1 = -1.0, t2 = -1.0;
    if (i < R.np) {
      const double dt = R.dt;
      const double l1 = S.lam1, l2 = S.lam2;
      double vx = R.vx[i], vy = R.vy[i], vz = R.vz[i];
      if (l1 != 1.0 || l2 != 1.0) {
        vx = vx * l1; vy = vy * l1; vz = vz * l1;
        vx = vx * l2; vy = vy * l2; vz = vz * l2;
        R.vx[i] = vx; R.vy[i] = vy; R.vz[i] = vz;
      }
      const double li0 = 1.0 / S.box[0], li1 = 1.0 / S.box[1], li2 = 1.0 / S.box[2];
      double du0 = li0 * (vx * dt + 0.5 * R.ax[i] * dt * dt);
      double du1 = li1 * (vy * dt + 0.5 * R.ay[i] * dt * dt);
      double du2 = li2 * (vz * dt + 0.5 * R.az[i] * dt * dt);
      if (R.nbc) hold_displacement(R, i, du0, du1, du2);
      const double px = R.px[i] + du0, py = R.py[i] + du1, pz = R.pz[i] + du2;
      R.px[i] = px; R.py[i] = py; R.pz[i] = pz;
      const double dx = px - R.pbx[i], dy = py - R.pby[i], dz = pz - R.pbz[i];
      if (R.trigger == NDG_TRIGGER_COMPAT) {
        t1 = dx * dx;
        t1 = t1 + dy * dy;
        t1 = t1 + dz * dz;
      } else {
        const double ex = dx * S.box[0], ey = dy * S.box[1], ez = dz * S.box[2];
        t1 = ex * ex + ey * ey + ez * ez;
      }
    }
    block_top2(t1, t2, sh);
    if (threadIdx.x == 0) {
      R.partials[kPartD1 * R.ntiles + T.ltile] = t1;
      R.partials[kPartD2 * R.ntiles + T.ltile] = t2;
    }
  }
}

// compute0 in FP64 on the full list: the reference's minimum image, d2 order, single-
// precision table index and FP64 interpolation (src/nd.f90:283-330).  Same walk as the
// mixed kernel (U entries in flight per lane, cutoff as a select), one atom per lane.
template <bool FULL, bool SPR>
__global__ __launch_bounds__(kTile) void k_force_fp64(EngineDev E, int t0, int nt, int xflags) {
  // SPR: the instance launched when some replica has spring pairs (its registers stay out
  // of the common instance)
  __shared__ double sh_red[8 * 2];
  constexpr int U = kUnroll;
  int tb, te, ts;
  tile_range(t0, nt, tb, te, ts);
  for (int t = tb; t < te; t += ts) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    const int f = active_flags(E, T.rep, xflags);
    if (!(f & F_ACTIVE)) continue;
    const bool full = (f & (F_CU | F_CV)) != 0;
    if (full != FULL) continue;
    const int i = T.atom0 + threadIdx.x;
    const int tab = R.ptsize + 1;
    const int kmax = R.ptsize;
    const double b0 = S.box[0], b1 = S.box[1], b2 = S.box[2];
    double fx = 0, fy = 0, fz = 0;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < R.np) {
      const double xi = R.px[i], yi = R.py[i], zi = R.pz[i];
      const int ti = R.typ[i];
      const int n = R.nnb[i];
      const unsigned *L = R.list + ((size_t)T.ltile * R.list_chunks) * kTile + threadIdx.x;
      for (int c = 0; c < n; c += U) {
        int j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = (c + u < n) ? (int)(L[(size_t)(c + u) * kTile] & kIndexMask) : i;
        double qx[U], qy[U], qz[U];
        int tj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          qx[u] = R.px[j[u]];
          qy[u] = R.py[j[u]];
          qz[u] = R.pz[j[u]];
          tj[u] = R.typ[j[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          // compute0: the pair's own id; compute1 (asymmetric matrix, src/nd.f90:445-498):
          // the larger id of the two directions, the force only if the neighbour's type
          // acts on this atom's type, energy and virial only for balanced pairs
          const int pij = R.pairmat[(ti - 1) + (tj[u] - 1) * R.ntypes];
          int pid = pij - 1;
          bool feel = true, bal = true;
          if (R.asym) {
            const int pji = R.pairmat[(tj[u] - 1) + (ti - 1) * R.ntypes];
            pid = (pij > pji ? pij : pji) - 1;
            feel = pji > 0;
            bal = pij == pji;
          }
          // list entries past the end read the atom itself, whose type may have no pair
          // with itself (id 0): keep the table reads in bounds (the entry is masked anyway)
          pid = pid < 0 ? 0 : pid;
          double s0 = xi - qx[u], s1 = yi - qy[u], s2 = zi - qz[u];
          s0 = s0 - round(s0);
          s1 = s1 - round(s1);
          s2 = s2 - round(s2);
          const double r0 = b0 * s0, r1 = b1 * s1, r2 = b2 * s2;
          double d2 = r0 * r0;
          d2 = d2 + r1 * r1;
          d2 = d2 + r2 * r2;
          double df, dp = 0.0;
          if (SPR && ((R.spring_mask >> pid) & 1)) {
            // compute2/3 spring (src/nd.f90:622-660): strain against the initial positions,
            // no distance cutoff, nothing past the strain limit PairParams(3)
            const int jj = j[u];
            double u0 = R.ipx[i] - R.ipx[jj], u1 = R.ipy[i] - R.ipy[jj], u2 = R.ipz[i] - R.ipz[jj];
            u0 = u0 - round(u0);
            u1 = u1 - round(u1);
            u2 = u2 - round(u2);
            const double q0 = S.ibox[0] * u0, q1 = S.ibox[1] * u1, q2 = S.ibox[2] * u2;
            double d20 = q0 * q0;
            d20 = d20 + q1 * q1;
            d20 = d20 + q2 * q2;
            const double rr = sqrt(d2), rr0 = sqrt(d20);
            const double sij = rr / rr0 - 1.0;
            const bool on = (c + u < n) && feel && !(sij > R.spr_s0[pid]);
            const double Vi = R.spr_vi[pid];
            const double fs = R.spr_c[pid] * sij * Vi * Vi;
            if (FULL) dp = on && bal ? 0.5 * fs * sij * rr0 : 0.0;
            df = on ? -fs / rr : 0.0;
          } else {
            const bool ok = (c + u < n) && !(d2 > R.rc2d[pid]) && feel;
            const float rk = (float)(d2 - R.ptrmin) * R.vdr2[pid] + 1.0f;
            int k = (int)rk;
            k = k < 1 ? 1 : (k > kmax ? kmax : k);
            const double w = (double)rk - (double)k;
            const double *Ft = R.tabF64 + (size_t)pid * tab;
            df = w * Ft[k] + (1.0 - w) * Ft[k - 1];
            df = ok ? df / d2 : 0.0;
            if (FULL) {
              const double *Pt = R.tabP64 + (size_t)pid * tab;
              dp = ok && bal ? w * Pt[k] + (1.0 - w) * Pt[k - 1] : 0.0;
            }
          }
          fx = fx + r0 * df;
          fy = fy + r1 * df;
          fz = fz + r2 * df;
          if (FULL) {
            const double dv = bal ? df : 0.0;
            acc[0] += 0.5 * dp;
            acc[1] += 0.5 * (r0 * r0 * dv);
            acc[2] += 0.5 * (r1 * r1 * dv);
            acc[3] += 0.5 * (r2 * r2 * dv);
            acc[4] += 0.5 * (r0 * r1 * dv);
            acc[5] += 0.5 * (r0 * r2 * dv);
            acc[6] += 0.5 * (r1 * r2 * dv);
          }
        }
      }
      R.fx[i] = fx;
      R.fy[i] = fy;
      R.fz[i] = fz;
      if (FULL) R.pene[i] = acc[0];
    }
    if (FULL) {
      for (int q = 0; q < 7; q += 2) {
        double pr[2] = {acc[q], q + 1 < 7 ? acc[q + 1] : 0.0};
        block_sum<2>(pr, sh_red);
        if (threadIdx.x == 0) {
          R.partials[(kPartPe + q) * R.ntiles + T.ltile] = pr[0];
          if (q + 1 < 7) R.partials[(kPartPe + q + 1) * R.ntiles + T.ltile] = pr[1];
        }
      }
    }
  }
}

__global__ __launch_bounds__(kTile) void k_kick_fp64(EngineDev E, int t0, int nt, int xflags) {
  __shared__ double sh[2];
  int tb, te, ts;
  tile_range(t0, nt, tb, te, ts);
  for (int t = tb; t < te; t += ts) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const int f = active_flags(E, T.rep, xflags);
    if (!(f & F_ACTIVE)) continue;
    const int i = T.atom0 + threadIdx.x;
    double kt[1] = {0.0};
    if (i < R.np) {
      const int ty = R.typ[i];
      const double m = R.mass[ty - 1];
      const double rm = 1.0 / m;
      const double hdt = 0.5 * R.dt;
      const double gx = R.fx[i] * rm, gy = R.fy[i] * rm, gz = R.fz[i] * rm;
      const double vx = R.vx[i] + hdt * (gx + R.ax[i]);
      const double vy = R.vy[i] + hdt * (gy + R.ay[i]);
      const double vz = R.vz[i] + hdt * (gz + R.az[i]);
      R.vx[i] = vx; R.vy[i] = vy; R.vz[i] = vz;
      R.ax[i] = gx; R.ay[i] = gy; R.az[i] = gz;
      if (f & F_CK) {
        double k = vx * vx;
        k = k + vy * vy;
        k = k + vz * vz;
        kt[0] = k * m;
      }
    }
    if (f & F_CK) {
      block_sum<1>(kt, sh);
      if (threadIdx.x == 0) R.partials[kPartKe * R.ntiles + T.ltile] = kt[0];
    }
  }
}

// ------------------------------------------------------------------ reduction + ensembles

enum : int { RED_FORCE = 1, RED_KICK = 2, RED_FLIP = 4, RED_STEP = 8, RED_CHECK = 16 };

__device__ inline void regrid(const RepParams &R, RepState &S) {
  int m[3];
  for (int a = 0; a < 3; ++a) m[a] = (int)(S.box[a] / (R.rcut + R.skin));
  while ((long long)m[0] * m[1] * m[2] > R.ncell_cap) {
    int a = (m[0] >= m[1] && m[0] >= m[2]) ? 0 : (m[1] >= m[2] ? 1 : 2);
    m[a] -= 1;
  }
  S.brute = (m[0] < 3 || m[1] < 3 || m[2] < 3) ? 1 : 0;
  for (int a = 0; a < 3; ++a) S.mcell[a] = S.brute ? 1 : m[a];
}

// Rebuild criterion of CheckListUpdate (src/nd.f90:1488-1496) from the top-2 metric.
__device__ inline int rebuild_test(const RepParams &R, RepState &S, int trigger, double d1,
                                   double d2) {
  if (d1 < 0) return 0;
  if (d2 < 0) d2 = 0;
  if (trigger == NDG_TRIGGER_COMPAT) {
    const double dsum = d1 + d2;
    if (dsum > R.skin * R.skin) {
      if (dsum > 1.0) {
        S.err = NDG_ERR_OUTSIDE;
        return 0;
      }
      return 1;
    }
    return 0;
  }
  const double bmin = fmin(S.box[0], fmin(S.box[1], S.box[2]));
  if (sqrt(d1) > 0.5 * bmin) {
    S.err = NDG_ERR_OUTSIDE;
    return 0;
  }
  double strain = 0.0;
  for (int a = 0; a < 3; ++a) strain = fmax(strain, fabs(S.box[a] / S.box_build[a] - 1.0));
  return (sqrt(d1) + sqrt(d2) + (R.rcut + R.skin) * strain > R.skin) ? 1 : 0;
}

// End of a step for one replica (src/nd.f90:164-178): e0, thermo record, mkIsoKE /
// mkIsoP / mkIsoE, BoxVel, and the list-update decision for the next step.
__device__ void step_epilogue(const RepParams &R, RepState &S, int f, double t1, double t2) {
  const int s = S.step + 1;
  const int tt = s - R.run_step0;
  if (tt == 1) S.e0 = S.pe + S.ke;
  if (f & F_OUT) {
    ndg_thermo &rec = R.ring[S.nrec % kRing];
    rec.step = s;
    rec.pad_ = 0;
    rec.pe = S.pe;
    rec.ke = S.ke;
    rec.eerr = (S.pe + S.ke - S.e0) / S.e0;
    rec.pressure = (S.vir[0] + S.vir[1] + S.vir[2]) / (S.box[0] * S.box[1] * S.box[2] * 3.0);
    rec.volume = S.box[0] * S.box[1] * S.box[2];
    for (int q = 0; q < 6; ++q) rec.virial[q] = S.vir[q];
    for (int q = 0; q < 3; ++q) rec.box[q] = S.box[q];
    S.nrec += 1;
  }
  const double dt = R.dt;
  double lam1 = 1.0, lam2 = 1.0;
  if (R.isokinetic && s % R.ke_freq == 0) {
    lam1 = 1.0 + dt / R.ke_tau * (R.ke_target * (double)R.np / S.ke - 1.0);
    lam1 = sqrt(lam1);
  }
  if (R.isobaric && s % R.p_freq == 0) {
    const double vol = S.box[0] * S.box[1] * S.box[2];
    double pr[6], lm[6];
    for (int q = 0; q < 6; ++q) pr[q] = S.vir[q] / vol;
    for (int q = 0; q < 3; ++q) lm[q] = 1.0 - dt / R.p_tau * (R.p_target[q] - pr[q]);
    lm[3] = dt / R.p_tau * pr[3];
    lm[4] = dt / R.p_tau * pr[4];
    lm[5] = dt / R.p_tau * pr[5];
    S.box[0] = S.box[0] * lm[0] + S.box[1] * lm[5] + S.box[2] * lm[4];
    S.box[1] = S.box[0] * lm[5] + S.box[1] * lm[1] + S.box[2] * lm[3];
    S.box[2] = S.box[0] * lm[4] + S.box[1] * lm[3] + S.box[2] * lm[2];
  }
  if (R.isoenergetic && s % R.e_freq == 0)
    lam2 = 1.0 - dt / R.e_tau * (S.e0 / (S.pe + S.ke) - 1.0);
  if (R.box_load)
    for (int q = 0; q < 3; ++q) S.box[q] = S.box[q] + R.box_vel[q] * dt;
  S.lam1 = lam1;
  S.lam2 = lam2;
  S.parity ^= 1;
  S.fnew = 0;
  S.step = s;
  S.rebuild = 0;
  if (tt >= R.run_nsteps) {
    S.done = 1;
    return;
  }
  if (R.renei && (s + 1) % R.renei_freq == 0) {
    if (rebuild_test(R, S, R.trigger, t1, t2)) {
      S.nupdates += 1;
      if (R.trigger == NDG_TRIGGER_SKIN) regrid(R, S);
      S.rebuild = 1;
    }
  }
}

// Kinetic-energy partials (NKW slots from kbase) and displacement top-2 of a thread's
// tiles, in tile order; the loads of a tile are issued together, four tiles at a time, so
// the sums cost one round trip to memory per four tiles.
template <int NKW>
__device__ inline void ke_top2_sum(const __attribute__((address_space(1))) double *p, int kbase,
                                   int ntl, int t0, int nthr, double &ke, double &t1,
                                   double &t2) {
#pragma unroll 4
  for (int lt = t0; lt < ntl; lt += nthr) {
    double x[NKW];
#pragma unroll
    for (int w = 0; w < NKW; ++w) x[w] = p[(size_t)(kbase + w) * ntl + lt];
    const double d1 = p[(size_t)kPartD1 * ntl + lt], d2 = p[(size_t)kPartD2 * ntl + lt];
#pragma unroll
    for (int w = 0; w < NKW; ++w) ke += x[w];
    top2_merge(t1, t2, d1, d2);
  }
}

// One workgroup per replica: fixed-order sums of the tile partials, then (by mode) the
// force / kick results, the drop-in rebuild check, or the full step epilogue.
// A replica's partials are summed by the first T threads, T = 1024 for a replica of more
// than 1024 tiles (2M atoms: 16 tiles per thread) and 256 otherwise; T depends only on the
// replica, and the other threads of the workgroup add exact zeros, so the sums do not
// depend on the launch's block size.  The kernel is a chain of round trips to memory (a
// replica of 2000 tiles is summed by one workgroup): the kinetic-energy and displacement
// partials, which every step reads, are loaded before the step flags are known (a
// released slot has no tiles), and the block reduction of the sums and the top-2 is one
// LDS exchange.
__global__ __launch_bounds__(1024) void k_reduce(EngineDev E, int r0, int nr, int mode,
                                                 int xflags, int check_trigger) {
  __shared__ double sh[16 * 10];
  for (int r = r0 + blockIdx.x; r < r0 + nr; r += gridDim.x) {
    const RepParams &R = E.reps[r];
    RepState &S = E.st[r];
    const int ntl = R.ntiles;
    const auto p = gmem((const double *)R.partials);
    const int nthr = ntl > 1024 ? 1024 : 256;
    const int t0 = (int)threadIdx.x < nthr ? (int)threadIdx.x : ntl;   // idle threads: no tiles
    double ke = 0.0, t1 = -1.0, t2 = -1.0;
    if (R.path == NDG_PATH_MIXED) {
      // kinetic-energy slots of a mixed-path tile: one per wave of the force kernel's tile
      switch (R.tpa) {
        case 1: ke_top2_sum<2>(p, kPartKeW, ntl, t0, nthr, ke, t1, t2); break;
        case 2: ke_top2_sum<4>(p, kPartKeW, ntl, t0, nthr, ke, t1, t2); break;
        case 4: ke_top2_sum<8>(p, kPartKeW, ntl, t0, nthr, ke, t1, t2); break;
        default: ke_top2_sum<16>(p, kPartKeW, ntl, t0, nthr, ke, t1, t2); break;
      }
    } else {
      ke_top2_sum<1>(p, kPartKe, ntl, t0, nthr, ke, t1, t2);
    }
    const int f = xflags ? xflags : step_flags(R, S);
    if (!(f & F_ACTIVE)) continue;
    double v[8] = {(f & F_CK) ? ke : 0.0, 0, 0, 0, 0, 0, 0, 0};
    if (f & (F_CU | F_CV)) {
#pragma unroll 4
      for (int lt = t0; lt < ntl; lt += nthr)
#pragma unroll
        for (int q = 1; q < 8; ++q) v[q] += p[q * ntl + lt];
    }
    block_sum_top2<8>(v, t1, t2, sh);
    if (threadIdx.x == 0) {
      if (mode & RED_FORCE) {
        S.fnew = 1;
        if (f & F_CU) S.pe = v[1];
        if (f & F_CV)
          for (int q = 0; q < 6; ++q) S.vir[q] = v[2 + q];
      }
      if (mode & RED_KICK) S.ke = (f & F_CK) ? 0.5 * v[0] : 0.0;
      if (mode & RED_FLIP) {
        S.parity ^= 1;
        S.fnew = 0;
      }
      if (mode & RED_CHECK) S.rebuild = rebuild_test(R, S, check_trigger, t1, t2);
      if (mode & RED_STEP) step_epilogue(R, S, f, t1, t2);
    }
    __syncthreads();
  }
}

// Applies a pending velocity scale (end of a Run: the last step's mkIsoKE/mkIsoE).
__global__ __launch_bounds__(256) void k_apply_pending(EngineDev E, int r0, int nr) {
  for (int r = r0 + blockIdx.x; r < r0 + nr; r += gridDim.x) {
    const RepParams &R = E.reps[r];
    RepState &S = E.st[r];
    const double l1 = S.lam1, l2 = S.lam2;
    if (l1 == 1.0 && l2 == 1.0) continue;
    for (int i = threadIdx.x; i < R.np; i += blockDim.x) {
      double vx = R.vx[i], vy = R.vy[i], vz = R.vz[i];
      vx = vx * l1; vy = vy * l1; vz = vz * l1;
      vx = vx * l2; vy = vy * l2; vz = vz * l2;
      R.vx[i] = vx; R.vy[i] = vy; R.vz[i] = vz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      S.lam1 = 1.0;
      S.lam2 = 1.0;
    }
  }
}

// Caller-order copy of one replica's output state, for a single copy to the host:
// pos(3,np) scaled, vel(3,np), force(3,np), per-atom energy(np), then box(3) and the step.
// Velocities are as the last step left them (a pending thermostat scale is applied by the
// next drift, as the reference writes its frame before mkIsoKE, src/nd.f90:172-176).
__global__ __launch_bounds__(256) void k_snapshot(EngineDev E, int r, double *out) {
  const RepParams &R = E.reps[r];
  const RepState &S = E.st[r];
  const int np = R.np;
  const bool mixed = R.path == NDG_PATH_MIXED;
  // latest forces: the buffer a force evaluation wrote if no kick has consumed it yet
  const float4 *fl = (S.parity ^ S.fnew) ? R.f32b : R.f32a;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    const int o = R.oid[i];
    double *p = out + 3 * (size_t)o;
    p[0] = R.px[i];
    p[1] = R.py[i];
    p[2] = R.pz[i];
    double *v = out + 3 * ((size_t)np + o);
    v[0] = R.vx[i];
    v[1] = R.vy[i];
    v[2] = R.vz[i];
    double *f = out + 3 * (2 * (size_t)np + o);
    if (mixed) {
      const float4 q = fl[i];
      f[0] = q.x;
      f[1] = q.y;
      f[2] = q.z;
      out[9 * (size_t)np + o] = q.w;
    } else {
      f[0] = R.fx[i];
      f[1] = R.fy[i];
      f[2] = R.fz[i];
      out[9 * (size_t)np + o] = R.pene[i];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 4) {
    const int a = threadIdx.x;
    out[10 * (size_t)np + a] = a < 3 ? S.box[a] : (double)S.step;
  }
}

// ------------------------------------------------------------------ list rebuild
// Counting sort of atoms by reference cell (canonical order: cell, then caller index),
// then a full Verlet list from the 27 surrounding cells in a fixed stencil order.

__device__ inline void derive_atom(const RepParams &R, const RepState &S, int dst, double px,
                                   double py, double pz, int ty);

// Copies atom src of the state to slot dst of the tmp arrays.  DERIVED: also writes the
// arrays derived from its position at slot dst directly (nothing reads them between this
// kernel and the list build, which then places the state arrays from tmp itself);
// otherwise k_rb_place writes both.
template <bool DERIVED>
__device__ inline void permute_atom(const RepParams &R, const RepState &S, int src, int dst) {
  const double px = R.px[src], py = R.py[src], pz = R.pz[src];
  const int ty = R.typ[src];
  R.tmpd[dst] = px;
  R.tmpd[R.np + dst] = py;
  R.tmpd[2 * R.np + dst] = pz;
  R.tmpd[3 * R.np + dst] = R.vx[src];
  R.tmpd[4 * R.np + dst] = R.vy[src];
  R.tmpd[5 * R.np + dst] = R.vz[src];
  if (R.path == NDG_PATH_FP64) {
    R.tmpd[6 * R.np + dst] = R.ax[src];
    R.tmpd[7 * R.np + dst] = R.ay[src];
    R.tmpd[8 * R.np + dst] = R.az[src];
  } else {
    R.tmpf[dst] = R.f32a[src];
    R.tmpf[R.np + dst] = R.f32b[src];
  }
  if (R.spring_mask) {
    R.tmpd[9 * R.np + dst] = R.ipx[src];
    R.tmpd[10 * R.np + dst] = R.ipy[src];
    R.tmpd[11 * R.np + dst] = R.ipz[src];
  }
  if (R.nbc) R.tmpi[2 * R.np + dst] = R.hbc[src];
  R.tmpi[dst] = ty;
  R.tmpi[R.np + dst] = R.oid[src];
  if (DERIVED) derive_atom(R, S, dst, px, py, pz, ty);
}

// The arrays derived from the position of slot dst: anchor, anchored and packed records,
// build-time copies.
__device__ inline void derive_atom(const RepParams &R, const RepState &S, int dst, double px,
                                   double py, double pz, int ty) {
  if (R.path == NDG_PATH_FP64) {
    R.pbx[dst] = px; R.pby[dst] = py; R.pbz[dst] = pz;
  }
  int kx, ky, kz, cxa, cya, cza;
  bin_axis(px, S.mcell[0], S.brute, kx, cxa);
  bin_axis(py, S.mcell[1], S.brute, ky, cya);
  bin_axis(pz, S.mcell[2], S.brute, kz, cza);
  const int cell = cell_rank(kx, ky, kz, S.mcell[0], S.mcell[1], S.mcell[2], R.cb_y, R.cb_z);
  R.anchor[dst] = make_int4(cxa, cya, cza, cell);
  const double4 x = make_double4(anchored(px, cxa, S.mcell[0], S.box[0] / S.mcell[0]),
                                 anchored(py, cya, S.mcell[1], S.box[1] / S.mcell[1]),
                                 anchored(pz, cza, S.mcell[2], S.box[2] / S.mcell[2]),
                                 type_bits(ty));
  R.x64a[dst] = x;
  R.xq[dst] = xq_pack(x.x, x.y, x.z, ty, xq_unit(S, R.skin));
  R.x32b[dst] = make_float4((float)x.x, (float)x.y, (float)x.z, 0.f);
}

// Slot p of the state arrays from tmp (the permuted copy written by permute_atom).
__device__ inline void place_state(const RepParams &R, int p) {
  R.px[p] = R.tmpd[p];
  R.py[p] = R.tmpd[R.np + p];
  R.pz[p] = R.tmpd[2 * R.np + p];
  R.vx[p] = R.tmpd[3 * R.np + p];
  R.vy[p] = R.tmpd[4 * R.np + p];
  R.vz[p] = R.tmpd[5 * R.np + p];
  R.typ[p] = R.tmpi[p];
  R.oid[p] = R.tmpi[R.np + p];
  if (R.nbc) R.hbc[p] = R.tmpi[2 * R.np + p];
  if (R.spring_mask) {
    R.ipx[p] = R.tmpd[9 * R.np + p];
    R.ipy[p] = R.tmpd[10 * R.np + p];
    R.ipz[p] = R.tmpd[11 * R.np + p];
  }
  if (R.path == NDG_PATH_FP64) {
    R.ax[p] = R.tmpd[6 * R.np + p];
    R.ay[p] = R.tmpd[7 * R.np + p];
    R.az[p] = R.tmpd[8 * R.np + p];
  } else {
    R.f32a[p] = R.tmpf[p];
    R.f32b[p] = R.tmpf[R.np + p];
  }
}

__device__ inline int cell_of(const RepParams &R, const RepState &S, int i) {
  int kx, ky, kz, c;
  bin_axis(R.px[i], S.mcell[0], S.brute, kx, c);
  bin_axis(R.py[i], S.mcell[1], S.brute, ky, c);
  bin_axis(R.pz[i], S.mcell[2], S.brute, kz, c);
  return cell_rank(kx, ky, kz, S.mcell[0], S.mcell[1], S.mcell[2], R.cb_y, R.cb_z);
}

// Order of atoms inside a cell: Morton index of the atom's 4 x 4 x 4 sub-cell, then
// caller index.  Neighbouring slots are then close in space, so the lists of consecutive
// atoms overlap (gathers of a wave share cache lines) and a run of consecutive candidates
// spans a small box.  Key = sub-cell << 26 | caller index (caller indices < 2^26).
__device__ inline unsigned order_key(const RepParams &R, const RepState &S, int i) {
  const double p[3] = {R.px[i], R.py[i], R.pz[i]};
  unsigned m = 0;
  for (int a = 0; a < 3; ++a) {
    const double sr = p[a] - round(p[a]);
    const double f = (sr + 0.4999) * (double)S.mcell[a];
    const double fr = f - floor(f);
    const int q = min(3, max(0, (int)(fr * 4.0)));
    m |= (unsigned)((q & 1) | ((q & 2) << 2)) << a;   // bits a and a + 3
  }
  return (m << kCodeShift) | (unsigned)R.oid[i];
}

// Final order of one cell's slots [lo, hi): ascending order_key.  Cells of up to 64
// atoms are sorted inside one wave (bitonic network over shuffles); cells up to
// kCellLds atoms by rank counts over a per-wave LDS copy; larger cells by rank counts over
// global memory.  The atoms are then copied to their final slots in tmp.
constexpr int kCellLds = 256;

template <bool DERIVED>
__device__ inline void sort_cell_wave(const RepParams &R, const RepState &S, int lo, int hi,
                                      int *wsrc, unsigned *wkey) {
  const int lane = threadIdx.x & 63;
  const int cnt = hi - lo;
  if (cnt <= 64) {
    int u = lane < cnt ? R.slot_src[lo + lane] : -1;
    unsigned key = lane < cnt ? order_key(R, S, u) : 0xffffffffu;
    for (int k = 2; k <= 64; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        const unsigned okey = __shfl_xor(key, j);
        const int ou = __shfl_xor(u, j);
        const bool up = (lane & k) == 0, lower = (lane & j) == 0;
        const bool take = lower ? (up ? okey < key : okey > key) : (up ? okey > key : okey < key);
        if (take) {
          key = okey;
          u = ou;
        }
      }
    if (lane < cnt) permute_atom<DERIVED>(R, S, u, lo + lane);
  } else if (cnt <= kCellLds) {
    for (int v = lane; v < cnt; v += 64) {
      const int u = R.slot_src[lo + v];
      wsrc[v] = u;
      wkey[v] = order_key(R, S, u);
    }
    wave_sync();
    for (int v = lane; v < cnt; v += 64) {
      const unsigned me = wkey[v];
      int rank = 0;
      for (int w = 0; w < cnt; ++w) rank += wkey[w] < me;
      permute_atom<DERIVED>(R, S, wsrc[v], lo + rank);
    }
    wave_sync();
  } else {
    for (int v = lo + lane; v < hi; v += 64) {
      const int i = R.slot_src[v];
      const unsigned me = order_key(R, S, i);
      int rank = 0;
      for (int w = lo; w < hi; ++w) rank += order_key(R, S, R.slot_src[w]) < me;
      permute_atom<DERIVED>(R, S, i, lo + rank);
    }
  }
}

// Work list of a rebuild: the tiles of every replica in [r0, r0 + nr) that rebuilds in
// this step, in slot order (rb[0] = count, rb[1 ..] = tile indices).  The rebuild kernels
// below walk this list, so their workgroups never visit the tiles of replicas that keep
// their lists (on a batch where a few replicas rebuild per step, those would be most).
__global__ __launch_bounds__(1024) void k_rb_plan(EngineDev E, int r0, int nr) {
  __shared__ int sh_n[1024], sh_t0[1024];
  __shared__ int sh_base;
  const int tid = threadIdx.x, nb = blockDim.x;
  if (tid == 0) sh_base = 0;
  __syncthreads();
  for (int c0 = r0; c0 < r0 + nr; c0 += nb) {
    const int r = c0 + tid;
    int n = 0, tb = 0;
    if (r < r0 + nr) {
      const int t0 = E.tile0[r];
      const RepState &S = E.st[r];
      if (t0 >= 0 && S.rebuild && !S.err) {
        n = E.reps[r].ntiles;
        tb = t0;
      }
    }
    sh_n[tid] = n;
    sh_t0[tid] = tb;
    __syncthreads();
    // inclusive scan of the chunk's tile counts
    for (int off = 1; off < nb; off <<= 1) {
      const int v = tid >= off ? sh_n[tid - off] : 0;
      __syncthreads();
      sh_n[tid] += v;
      __syncthreads();
    }
    const int base = sh_base, tot = sh_n[nb - 1];
    // output position q of the chunk belongs to the first replica j with sh_n[j] > q
    for (int q = tid; q < tot; q += nb) {
      int lo = 0, hi = nb - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sh_n[mid] > q) hi = mid; else lo = mid + 1;
      }
      const int first = lo == 0 ? 0 : sh_n[lo - 1];
      E.rb[1 + base + q] = sh_t0[lo] + (q - first);
    }
    __syncthreads();
    if (tid == 0) sh_base = base + tot;
    __syncthreads();
  }
  if (tid == 0) E.rb[0] = sh_base;
}

// Tile k of the rebuild work list for this workgroup (uniform), or -1 past its end.
__device__ inline int rb_tile(const EngineDev &E, int k) {
  return k < __builtin_amdgcn_readfirstlane(E.rb[0]) ? __builtin_amdgcn_readfirstlane(E.rb[1 + k]) : -1;
}

// Cell sort of the rebuilding replicas in five grid-wide phases over their tiles:
// histogram, per-replica scan, scatter, per-cell order (one wave per cell), placement.
// (A whole replica sorted inside one workgroup, with its bookkeeping in LDS, was slower
// on the replica mix at every size: one workgroup serialises what the tiles share.)
// Runs of equal cells within a wave (atoms stay nearly sorted by cell between rebuilds):
// c = -1 marks a lane without an atom.  Returns the lane's run leader and the run length
// (valid in the leader), so that one atomic serves the run.
__device__ inline int cell_run(int c, int &len) {
  const int lane = threadIdx.x & 63;
  const int cp = __shfl_up(c, 1);
  const unsigned long long lead = __ballot(lane == 0 || c != cp);
  const unsigned long long after = lane == 63 ? 0ull : lead & (~0ull << (lane + 1));
  const int next = after ? __builtin_ctzll(after) : 64;
  len = next - lane;
  return 63 - __builtin_clzll(lead & (lane == 63 ? ~0ull : ((2ull << lane) - 1)));
}

// Histogram of one tile: the cell of each atom, one atomic per run of equal cells.
__device__ inline void bin_tile(const EngineDev &E, int t) {
  const Tile T = E.tiles[t];
  const RepParams &R = E.reps[T.rep];
  const RepState &S = E.st[T.rep];
  if (!S.rebuild || S.err) return;
  const int i = T.atom0 + threadIdx.x;
  const int c = i < R.np ? cell_of(R, S, i) : -1;
  if (c >= 0) R.key[i] = c;
  int len;
  const int ld = cell_run(c, len);
  if (c >= 0 && ld == (int)(threadIdx.x & 63)) atomicAdd(&R.cell_count[c], len);
}

__global__ __launch_bounds__(kTile) void k_rb_bin(EngineDev E) {
  for (int k = blockIdx.x, t; (t = rb_tile(E, k)) >= 0; k += gridDim.x) bin_tile(E, t);
}

// The histogram with the work list planned in the same launch: every workgroup scans the
// rebuild flags of the replicas [r0, r0 + nr) into LDS (prefix of the rebuilding
// replicas' tile counts, in slot order, as k_rb_plan) and maps its work indices to
// tiles itself; each workgroup writes the list entries it visits and workgroup 0 the
// count, for the kernels that follow.  One launch less per step than k_rb_plan + k_rb_bin
// (an empty launch costs ~3 us).  Dynamic LDS: 2 nr + 1 ints, nr <= kPlanLds.
constexpr int kPlanLds = 4096;

__global__ __launch_bounds__(kTile) void k_rb_bin_plan(EngineDev E, int r0, int nr) {
  extern __shared__ int sh_plan[];
  __shared__ int sh_scan[kTile];
  int *pre = sh_plan, *first = sh_plan + nr + 1;   // pre[j]: tiles of replicas before j
  const int tid = threadIdx.x;
  for (int j = tid; j < nr; j += kTile) {
    const int r = r0 + j;
    const int t0 = E.tile0[r];
    const RepState &S = E.st[r];
    pre[j + 1] = t0 >= 0 && S.rebuild && !S.err ? E.reps[r].ntiles : 0;
    first[j] = t0;
  }
  if (tid == 0) pre[0] = 0;
  __syncthreads();
  // inclusive prefix over chunks of consecutive replicas, one chunk per thread
  const int per = (nr + kTile - 1) / kTile;
  const int c0 = min(nr, tid * per), c1 = min(nr, c0 + per);
  int loc = 0;
  for (int j = c0; j < c1; ++j) loc += pre[j + 1];
  sh_scan[tid] = loc;
  __syncthreads();
  for (int off = 1; off < kTile; off <<= 1) {
    const int v = tid >= off ? sh_scan[tid - off] : 0;
    __syncthreads();
    sh_scan[tid] += v;
    __syncthreads();
  }
  int run = sh_scan[tid] - loc;
  for (int j = c0; j < c1; ++j) {
    run += pre[j + 1];
    pre[j + 1] = run;
  }
  __syncthreads();
  const int total = pre[nr];
  if (blockIdx.x == 0 && tid == 0) E.rb[0] = total;
  for (int k = blockIdx.x; k < total; k += gridDim.x) {
    // the last replica j with pre[j] <= k (it has tiles: pre[j + 1] > k)
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const int t = first[lo] + (k - pre[lo]);
    if (tid == 0) E.rb[1 + k] = t;
    bin_tile(E, t);
  }
}

__global__ __launch_bounds__(1024) void k_rb_scan(EngineDev E, int r0, int nr) {
  __shared__ int sh_scan[1024];
  for (int r = r0 + blockIdx.x; r < r0 + nr; r += gridDim.x) {
    const RepParams &R = E.reps[r];
    RepState &S = E.st[r];
    if (!S.rebuild || S.err) continue;
    const int nb = blockDim.x, tid = threadIdx.x;
    const int ncell = S.mcell[0] * S.mcell[1] * S.mcell[2];
    const int per = (ncell + nb - 1) / nb;
    const int c0 = tid * per, c1 = min(ncell, c0 + per);
    int loc = 0;
    for (int c = c0; c < c1; ++c) loc += R.cell_count[c];
    sh_scan[tid] = loc;
    __syncthreads();
    for (int off = 1; off < nb; off <<= 1) {
      int v = tid >= off ? sh_scan[tid - off] : 0;
      __syncthreads();
      sh_scan[tid] += v;
      __syncthreads();
    }
    int run = sh_scan[tid] - loc;
    for (int c = c0; c < c1; ++c) {
      R.cell_start[c] = run;
      run += R.cell_count[c];
      R.cell_count[c] = 0;
      R.cell_fill[c] = 0;
    }
    if (tid == nb - 1) R.cell_start[ncell] = sh_scan[nb - 1];
    if (tid == 0) {
      S.nbr_sum = 0;
      for (int a = 0; a < 3; ++a) S.box_build[a] = S.box[a];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kTile) void k_rb_scatter(EngineDev E) {
  for (int k = blockIdx.x, t; (t = rb_tile(E, k)) >= 0; k += gridDim.x) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    if (!S.rebuild || S.err) continue;
    const int i = T.atom0 + threadIdx.x;
    const int c = i < R.np ? R.key[i] : -1;
    int len;
    const int ld = cell_run(c, len);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (c >= 0 && ld == lane) base = atomicAdd(&R.cell_fill[c], len);
    base = __shfl(base, ld);
    if (c >= 0) R.slot_src[R.cell_start[c] + base + (lane - ld)] = i;
  }
}

// Each tile sorts the cells whose first slot lies in it (one wave per cell, kRankWaves
// waves per tile, so that a tile's few cells are sorted side by side).
constexpr int kRankWaves = 8;
// BIG: the derived arrays are written here (permute_atom<true>).
template <bool BIG>
__global__ __launch_bounds__(64 * kRankWaves) void k_rb_rank(EngineDev E) {
  __shared__ int sh_sort[kRankWaves][2 * kCellLds];
  for (int k = blockIdx.x, t; (t = rb_tile(E, k)) >= 0; k += gridDim.x) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    if (!S.rebuild || S.err) continue;
    if (S.brute) {
      const int i = T.atom0 + threadIdx.x;
      if (threadIdx.x < kTile && i < R.np) permute_atom<BIG>(R, S, i, R.oid[i]);
      continue;
    }
    const int ncell = S.mcell[0] * S.mcell[1] * S.mcell[2];
    int c = R.key[R.slot_src[T.atom0]];
    if (R.cell_start[c] < T.atom0) ++c;
    const int wid = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (c += wid; c < ncell && R.cell_start[c] < T.atom0 + kTile; c += nwave) {
      const int lo = R.cell_start[c], hi = R.cell_start[c + 1];
      if (hi > lo)
        sort_cell_wave<BIG>(R, S, lo, hi, sh_sort[wid], (unsigned *)(sh_sort[wid] + kCellLds));
    }
  }
}

// The state arrays from tmp and the derived arrays, for engines whose order kernel and list
// build do not write them (see list_tile).
__global__ __launch_bounds__(kTile) void k_rb_place(EngineDev E) {
  for (int k = blockIdx.x, t; (t = rb_tile(E, k)) >= 0; k += gridDim.x) {
    const Tile T = E.tiles[t];
    const RepParams &R = E.reps[T.rep];
    const RepState &S = E.st[T.rep];
    if (!S.rebuild || S.err) continue;
    const int p = T.atom0 + threadIdx.x;
    if (p >= R.np) continue;
    place_state(R, p);
    derive_atom(R, S, p, R.tmpd[p], R.tmpd[R.np + p], R.tmpd[2 * R.np + p], R.tmpi[p]);
  }
}

// Full Verlet list (both directions) of each atom: the reference's pair test
// (src/nd.f90:1358-1377,1404-1420) over the 27 surrounding cells (or all atoms in brute
// mode).  Work unit: one wave per chunk of up to kChunk home atoms of one cell, staged in
// LDS.  The wave walks the cell's candidate stream (stencil cell (dz,dy,dx)
// lexicographic, then slot) 64 candidates at a time, lane = candidate.  On the anchored
// path a batch is handled in two phases: every home atom's test leaves one ballot mask in
// LDS, then lane = home atom appends the set bits in lane order, so every list is in
// canonical order.  The exact path tests and appends one home atom at a time.
//
// Test.  The reference computes d2 from scaled positions with anint images.  When the
// grid guarantees that the stencil representation of a pair is its minimum image for
// every pair inside the cutoff ((M-2) w - 2.5e-4 box > cut on each axis; the 2.5e-4
// covers the 0.4999 of the binning), d2 is formed instead from the cell-anchored
// coordinates in FP32, candidate moved by its stencil offset into the home cell's frame:
// three subtractions and three multiply-adds per pair.  That d2 is within ~2e-6 c2 of
// the reference's; a pair whose FP32 d2 lies within test_band (1e-3 A^2 + 1e-5 c2) of the
// list cutoff (or of PTrMin for same-cell pairs) is decided by the reference's own
// formula, so the accepted set is the reference's.  Candidates farther than the cutoff
// from the chunk's bounding box are culled before the test.  Other grids
// (brute mode, M = 3 with w close to the cutoff) use the reference's formula for every
// pair.  Entries carry the cell/image offset code of the mixed force kernel (ox = -dx).
constexpr int kListWaves = 8;    // waves per workgroup (one workgroup per tile)
// Home atoms per wave task: at most kChunk; 64 when every replica of the engine has at least
// kBigTiles tiles (the k_rb_list<true> instance), 32 otherwise (16 in brute mode).  A large
// replica rebuilds many tiles at once, enough to fill the GPU with either chunk size, and a
// cell's one chunk then streams its 27 cells once instead of twice (c2 +4 %, c1 +3 %
// atom-steps/s); the replica mix, whose few rebuilding tiles per step need the waves of the
// smaller chunks to hide their latency, runs 6 % slower with 64.
constexpr int kChunk = 64;

struct alignas(16) HomeRec {
  float x, y, z;                 // cell-anchored position (FP32, for the banded test)
  int ih, th;                    // slot, type
  double px, py, pz;             // scaled position (reference test)
  int ax, ay, az, pad;           // anchor
};

// Candidates that passed the chunk's bounding-box cull, in stream order, waiting for a
// full batch of 64: a ring of 128 (at most 63 wait while a raw batch adds up to 64 more).
constexpr int kQueue = 128;

struct ListWave {
  HomeRec home[kChunk];
  int lo[27], pref[28];
  // two tested batches, appended together (lanes 0-31 the first, 32-63 the second) when a
  // chunk has at most 32 homes
  unsigned long long mask[2][kChunk];   // accepted candidates of the batch, per home atom
  unsigned ent[2][64];                  // list entry of each candidate of the batch
  float qx[kQueue], qy[kQueue], qz[kQueue];    // candidate in the frame of the home cell
  int qu[kQueue];                    // slot
  unsigned qent[kQueue];             // list entry
  int qtj[kQueue];                   // type | 256 if in the home cell
};

__device__ inline unsigned long long ballot64(bool p) { return __ballot(p); }

// A double that is the same in every lane, held in scalar registers.
__device__ inline double uniform_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline int lane_prefix(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// The reference's d2 (scaled difference, anint image, box scale, sum in order).
__device__ inline double ref_d2(double xi, double yi, double zi, double xj, double yj,
                                double zj, double b0, double b1, double b2, double &nx,
                                double &ny, double &nz) {
  const double dpx = xi - xj, dpy = yi - yj, dpz = zi - zj;
  nx = round(dpx);
  ny = round(dpy);
  nz = round(dpz);
  const double r0 = b0 * (dpx - nx), r1 = b1 * (dpy - ny), r2 = b2 * (dpz - nz);
  double d2 = r0 * r0;
  d2 = d2 + r1 * r1;
  d2 = d2 + r2 * r2;
  return d2;
}

// list cutoff^2 per type pair (8 x 8, -1: no pair) of the workgroup's replica
__shared__ double s_lcut2[64];


struct ListCtx {
  unsigned *list;
  size_t kstride, tstride;       // entry stride (kTile tpa), tile stride of the list
  int tpa, tsh, neimax;
  double ptrmin;
  const double *px, *py, *pz;
  bool uni;                      // one list cutoff for every type pair
  double c2u;
};

// Append the accepted lanes of batch mask `acc` to the list of home h (slot ih).
__device__ inline void append(const ListCtx &C, int &nv, int h, int ih, unsigned long long acc,
                              unsigned ent, bool bad_any) {
  const int lane = threadIdx.x & 63;
  int nov = __builtin_amdgcn_readlane(nv, h);
  if (nov >= 0) {
    const int nacc = __popcll(acc);
    if (bad_any || nov + nacc > C.neimax) {
      nov = -1;
    } else if (nacc) {
      if ((acc >> lane) & 1ull) {
        auto bh = gmem(C.list + (size_t)(ih >> 7) * C.tstride + (ih & (kTile - 1)) * C.tpa);
        const unsigned k = (unsigned)(nov + lane_prefix(acc));
        bh[((k >> C.tsh) << (7 + C.tsh)) | (k & (unsigned)(C.tpa - 1))] = ent;
      }
      nov += nacc;
    }
    if (lane == h) nv = nov;
  }
}

// Cutoff bands and box of the anchored test.
struct BandCtx {
  double c2u;                    // uniform (or largest) list cutoff^2
  float c2lo_u, c2hi_u;          // its band
  float pmlo, pmhi;              // PTrMin band (same-cell pairs)
  double ptrmin;
  double b0, b1, b2;             // box (reference re-test)
};

// Half-width (A^2) of the bands around a list cutoff and PTrMin inside which the FP32
// test defers to the reference's formula.  In the home cell's frame a coordinate is at
// most X ~ 2 cut + skin + 1 (cells are at most 4/3 of the cut wide on the anchored path);
// rounding it, the difference and the sums to FP32 leaves d2 within ~2e-6 cut^2 (+ a few
// 1e-5 A^2 for small cuts) of the exact value, so a pair outside the band is decided the
// same way by both.
__device__ inline double test_band(double c2) { return 1e-5 * fabs(c2) + 1e-3; }

// Test phase of one candidate batch on the anchored path: mrow[h] = the lanes
// (candidates) that enter the list of home h, for every home of the chunk.  Homes are
// taken kH at a time: their LDS records are read together, then kH independent FP64
// chains run; a candidate within the band of a cutoff is decided by the reference's
// formula (rare, after the group).  SAME: the batch holds same-cell candidates (PTrMin
// band); UNI: one list cutoff for every type pair.  Two homes per group: with the
// candidate queue's registers, four spill 8 VGPRs at the 128-register budget (0.3 %
// slower on c3, c2 and c1).
constexpr int kH = 2;

template <bool SAME, bool UNI>
__device__ inline void test_batch(ListWave &L, unsigned long long *mrow,
                                  const __attribute__((address_space(1))) double *gpx,
                                  const __attribute__((address_space(1))) double *gpy,
                                  const __attribute__((address_space(1))) double *gpz,
                                  const BandCtx &B, int lo, int cnt, int u, bool same,
                                  float ux, float uy, float uz, int tj,
                                  unsigned long long valid, unsigned long long samem) {
  const int lane = threadIdx.x & 63;
  for (int h0 = 0; h0 < cnt; h0 += kH) {
    float hx[kH], hy[kH], hz[kH];
    int hth[kH];
#pragma unroll
    for (int k = 0; k < kH; ++k) {
      const HomeRec &hr = L.home[min(h0 + k, cnt - 1)];
      hx[k] = hr.x;
      hy[k] = hr.y;
      hz[k] = hr.z;
      hth[k] = hr.th;
    }
    unsigned long long acc[kH], amb[kH];
#pragma unroll
    for (int k = 0; k < kH; ++k) {
      const int ih = lo + min(h0 + k, cnt - 1);
      const float dx = hx[k] - ux, dy = hy[k] - uy, dz = hz[k] - uz;
      const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      unsigned long long live = valid & ~ballot64(u == ih);
      float c2lo = B.c2lo_u, c2hi = B.c2hi_u;
      if (!UNI) {
        const double c2 = s_lcut2[(hth[k] - 1) * 8 + tj - 1];
        c2lo = (float)(c2 - test_band(c2));
        c2hi = (float)(c2 + test_band(c2));
        live &= ballot64(c2 >= 0.0);
      }
      const unsigned long long lo_m = ballot64(d2 < c2lo), hi_m = ballot64(d2 <= c2hi);
      unsigned long long a = live & lo_m, m = live & hi_m & ~lo_m;
      if (SAME) {
        const unsigned long long plo = ballot64(d2 < B.pmlo), phi = ballot64(d2 <= B.pmhi);
        m = (m & ~(samem & plo)) | (a & samem & phi & ~plo);
        a &= ~(samem & phi);
      }
      acc[k] = a;
      amb[k] = m;
    }
    unsigned long long any = 0;
#pragma unroll
    for (int k = 0; k < kH; ++k) any |= amb[k];
    if (any) {
      for (int k = 0; k < kH && h0 + k < cnt; ++k) {
        if (!amb[k]) continue;
        const HomeRec &hr = L.home[h0 + k];
        bool ok = false;
        if ((amb[k] >> lane) & 1ull) {
          double c2 = B.c2u;
          if (!UNI) c2 = s_lcut2[(hr.th - 1) * 8 + tj - 1];
          double nx, ny, nz;
          const double e2 = ref_d2(hr.px, hr.py, hr.pz, gpx[u], gpy[u], gpz[u], B.b0, B.b1,
                                   B.b2, nx, ny, nz);
          ok = !(same && e2 < B.ptrmin) && !(e2 > c2);
        }
        acc[k] |= ballot64(ok);
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < kH; ++k)
        if (h0 + k < cnt) mrow[h0 + k] = acc[k];
    }
  }
}

__device__ long long list_chunk(const RepParams &R, RepState &S, ListWave &L, const ListCtx &C,
                                bool anch, int lo, int cnt, int cell, bool pair) {
  const int lane = threadIdx.x & 63;
  lo = __builtin_amdgcn_readfirstlane(lo);
  cnt = __builtin_amdgcn_readfirstlane(cnt);
  const bool brute = cell < 0;
  const int mx = S.mcell[0], my = S.mcell[1], mz = S.mcell[2];
  const double b0 = S.box[0], b1 = S.box[1], b2 = S.box[2];
  const double w0 = b0 / mx, w1 = b1 / my, w2 = b2 / mz;
  // bounding box of the chunk's homes (cell-anchored frame), for the candidate cull
  double bl0 = 1e300, bl1 = 1e300, bl2 = 1e300, bh0 = -1e300, bh1 = -1e300, bh2 = -1e300;
  if (lane < cnt) {
    const int ih = lo + lane;
    const double4 xa = R.x64a[ih];
    const int4 A = R.anchor[ih];
    bl0 = bh0 = xa.x;
    bl1 = bh1 = xa.y;
    bl2 = bh2 = xa.z;
    HomeRec &hr = L.home[lane];
    hr.x = (float)xa.x;
    hr.y = (float)xa.y;
    hr.z = (float)xa.z;
    hr.ih = ih;
    hr.th = bits_type(xa.w);
    hr.px = C.px[ih];
    hr.py = C.py[ih];
    hr.pz = C.pz[ih];
    hr.ax = A.x;
    hr.ay = A.y;
    hr.az = A.z;
  }
  int total;
  if (brute) {
    total = R.np;
  } else {
    int q_cnt = 0;
    if (lane < 27) {
      int cx, cy, cz;
      cell_coords(cell, mx, my, mz, R.cb_y, R.cb_z, cx, cy, cz);
      const int dz = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dx = lane % 3 - 1;
      int ux = cx + dx, uy = cy + dy, uz = cz + dz;
      ux += ux < 0 ? mx : (ux >= mx ? -mx : 0);
      uy += uy < 0 ? my : (uy >= my ? -my : 0);
      uz += uz < 0 ? mz : (uz >= mz ? -mz : 0);
      const int cu = cell_rank(ux, uy, uz, mx, my, mz, R.cb_y, R.cb_z);
      L.lo[lane] = R.cell_start[cu];
      q_cnt = R.cell_start[cu + 1] - R.cell_start[cu];
    }
    int run = q_cnt;
    for (int off = 1; off < 32; off <<= 1) {
      const int v = __shfl_up(run, off);
      if (lane >= off) run += v;
    }
    if (lane < 27) L.pref[lane + 1] = run;
    if (lane == 0) L.pref[0] = 0;
    total = __shfl(run, 26);
  }
  total = __builtin_amdgcn_readfirstlane(total);
  wave_sync();
  const double ptrmin = C.ptrmin;
  const bool uni = C.uni;
  const double c2u = C.c2u;
  BandCtx B;
  B.c2u = c2u;
  B.c2lo_u = (float)(c2u - test_band(c2u));
  B.c2hi_u = (float)(c2u + test_band(c2u));
  // PTrMin band: the same width as the cutoff band of the largest cutoff
  B.pmlo = (float)(ptrmin - test_band(c2u));
  B.pmhi = (float)(ptrmin + test_band(c2u));
  B.ptrmin = ptrmin;
  B.b0 = b0;
  B.b1 = b1;
  B.b2 = b2;
  const auto gx64a = gmem((const double *)R.x64a);
  const auto gpx = gmem(C.px), gpy = gmem(C.py), gpz = gmem(C.pz);
  // slot u and stencil cell q of stream position g (u = lo, q = 13 past the end)
  auto locate = [&](int g, int &u, int &q) {
    u = lo;
    q = 13;
    if (g < total) {
      if (brute) {
        u = g;
      } else {
        int a_ = 0, b_ = 26;   // last stencil cell q with pref[q] <= g
        while (a_ < b_) {
          const int mid = (a_ + b_ + 1) >> 1;
          if (L.pref[mid] <= g) a_ = mid; else b_ = mid - 1;
        }
        q = a_;
        u = L.lo[q] + (g - L.pref[q]);
      }
    }
  };
  auto load_x = [&](int u) {
    return make_double4(gx64a[4 * u], gx64a[4 * u + 1], gx64a[4 * u + 2], gx64a[4 * u + 3]);
  };
  int nv = 0;   // lane h: list length of home h, -1 after an overflow
  // test of one batch of 64 candidates into mask/entry slot ts (lane = candidate: slot bu,
  // position bx, by, bz in the frame of the home cell, type btj, entry bent)
  auto test_slot = [&](int ts, int bu, bool bsame, float bx, float by, float bz, int btj,
                       unsigned bent, unsigned long long bvalid, unsigned long long bsamem) {
    L.ent[ts][lane] = bent;
    unsigned long long *mrow = L.mask[ts];
    // one ballot of the batch per home atom (no appends here)
    if (uni) {
      if (bsamem) test_batch<true, true>(L, mrow, gpx, gpy, gpz, B, lo, cnt, bu, bsame, bx, by, bz, btj, bvalid, bsamem);
      else test_batch<false, true>(L, mrow, gpx, gpy, gpz, B, lo, cnt, bu, bsame, bx, by, bz, btj, bvalid, bsamem);
    } else {
      if (bsamem) test_batch<true, false>(L, mrow, gpx, gpy, gpz, B, lo, cnt, bu, bsame, bx, by, bz, btj, bvalid, bsamem);
      else test_batch<false, false>(L, mrow, gpx, gpy, gpz, B, lo, cnt, bu, bsame, bx, by, bz, btj, bvalid, bsamem);
    }
  };
  // append of the ns (1 or 2) tested batches: lane = home atom h of batch half = lane / 32
  // (the second batch's entries follow the first's), the set bits of its mask in lane
  // (= stream) order, up to four entries per step (their LDS reads in flight together)
  auto append_slots = [&](int ns) {
    wave_sync();
    const int half = pair ? lane >> 5 : 0, h = pair ? lane & 31 : lane;
    const int nvh = __shfl(nv, h);   // the home's list length, held by lane h
    const bool act = h < cnt && half < ns;
    const unsigned long long m0 = act ? L.mask[0][h] : 0ull;
    const unsigned long long m1 = act && ns > 1 ? L.mask[1][h] : 0ull;
    const int tot = __popcll(m0) + __popcll(m1);
    const bool ovf = nvh < 0 || nvh + tot > C.neimax;
    if (act && !ovf) {
      unsigned long long m = half ? m1 : m0;
      const int ih = lo + h;
      const auto bh = gmem(C.list + (size_t)(ih >> 7) * C.tstride + (ih & (kTile - 1)) * C.tpa);
      const unsigned msk = (unsigned)(C.tpa - 1), sh = (unsigned)C.tsh;
      const unsigned *er = L.ent[half];
      unsigned k = (unsigned)(nvh + (half ? __popcll(m0) : 0));
      while (m) {
        int b[4];
        int nb = 0;
#pragma unroll
        for (int z = 0; z < 4; ++z) {
          b[z] = m ? __builtin_ctzll(m) : 0;
          nb += m ? 1 : 0;
          m &= m - 1;
        }
        unsigned ev[4];
#pragma unroll
        for (int z = 0; z < 4; ++z) ev[z] = er[b[z]];
#pragma unroll
        for (int z = 0; z < 4; ++z)
          if (z < nb) bh[(((k + z) >> sh) << (7 + sh)) | ((k + z) & msk)] = ev[z];
        k += (unsigned)nb;
      }
    }
    if (lane < cnt) nv = ovf ? -1 : nvh + tot;
    wave_sync();   // the next batches rewrite mask and ent
  };
  int ns = 0;   // tested batches waiting for their append (wave-uniform)
  auto run_batch = [&](int bu, bool bsame, float bx, float by, float bz, int btj,
                       unsigned bent, unsigned long long bvalid, unsigned long long bsamem) {
    test_slot(ns, bu, bsame, bx, by, bz, btj, bent, bvalid, bsamem);
    if (!pair) {
      append_slots(1);   // chunks of more than 32 homes: one batch per append
    } else if (ns == 1) {
      append_slots(2);
      ns = 0;
    } else {
      ns = 1;
    }
  };
  // reduce the homes' bounding box to every lane; a candidate farther than the largest list
  // cutoff from it cannot enter any of the chunk's lists (the margin covers the rounding of
  // the box distance and the reference re-test's ~1e-12 A^2), so on the anchored path only
  // the survivors, about half of the 27-cell stream, are tested, in full batches of 64
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    bl0 = fmin(bl0, shfl_xor_d(bl0, off));
    bl1 = fmin(bl1, shfl_xor_d(bl1, off));
    bl2 = fmin(bl2, shfl_xor_d(bl2, off));
    bh0 = fmax(bh0, shfl_xor_d(bh0, off));
    bh1 = fmax(bh1, shfl_xor_d(bh1, off));
    bh2 = fmax(bh2, shfl_xor_d(bh2, off));
  }
  bl0 = uniform_d(bl0);
  bl1 = uniform_d(bl1);
  bl2 = uniform_d(bl2);
  bh0 = uniform_d(bh0);
  bh1 = uniform_d(bh1);
  bh2 = uniform_d(bh2);
  // FP32 box distance of the FP32 candidate: within the same bound as the test's d2
  const float cull2 = (float)uniform_d(c2u + test_band(c2u));
  const float fl0 = (float)bl0, fl1 = (float)bl1, fl2 = (float)bl2;
  const float fh0 = (float)bh0, fh1 = (float)bh1, fh2 = (float)bh2;
  int nq = 0, qh = 0;   // queued candidates and the ring position of the first (wave-uniform)
  // the next batch's candidates are located and loaded while this batch is tested
  int u, q;
  locate(lane, u, q);
  double4 xu = load_x(u);
  for (int g0 = 0; g0 < total; g0 += 64) {
    const bool inb = g0 + lane < total;
    const bool same = !brute && q == 13;
    const unsigned long long valid = ballot64(inb);
    const unsigned long long samem = ballot64(inb && same);
    int un = lo, qn = 13;
    double4 xn = xu;
    if (g0 + 64 < total) {
      locate(g0 + 64 + lane, un, qn);
      xn = load_x(un);
    }
    const int tj = bits_type(xu.w);
    if (anch) {
      const int dxq = q % 3 - 1, dyq = (q / 3) % 3 - 1, dzq = q / 9 - 1;
      // the candidate in the frame of the home cell
      const float ux_ = (float)(xu.x + dxq * w0), uy_ = (float)(xu.y + dyq * w1),
                  uz_ = (float)(xu.z + dzq * w2);
      const unsigned ent = (unsigned)u |
          ((unsigned)((1 - dxq) | ((1 - dyq) << 2) | ((1 - dzq) << 4)) << kCodeShift);
      const float gx = fmaxf(0.f, fmaxf(fl0 - ux_, ux_ - fh0));
      const float gy = fmaxf(0.f, fmaxf(fl1 - uy_, uy_ - fh1));
      const float gz = fmaxf(0.f, fmaxf(fl2 - uz_, uz_ - fh2));
      const bool keep = inb && !(fmaf(gx, gx, fmaf(gy, gy, gz * gz)) > cull2);
      const unsigned long long km = ballot64(keep);
      if (keep) {
        const int k = (qh + nq + lane_prefix(km)) & (kQueue - 1);
        L.qx[k] = ux_;
        L.qy[k] = uy_;
        L.qz[k] = uz_;
        L.qu[k] = u;
        L.qent[k] = ent;
        L.qtj[k] = tj | (same ? 256 : 0);
      }
      nq += __popcll(km);
      // full batches of queued candidates, and after the last raw batch the rest
      // (absent lanes of a partial batch: the first home, far away)
      const bool last = g0 + 64 >= total;
      while (nq >= 64 || (last && nq > 0)) {
        wave_sync();
        const int k = (qh + lane) & (kQueue - 1);
        const bool inq = lane < nq;
        const int bt = inq ? L.qtj[k] : 0;
        const bool bs = inq && (bt & 256) != 0;
        run_batch(inq ? L.qu[k] : lo, bs, inq ? L.qx[k] : 1e30f, inq ? L.qy[k] : 0.f,
                  inq ? L.qz[k] : 0.f, inq ? (bt & 255) : 1, inq ? L.qent[k] : 0u,
                  ballot64(inq), ballot64(bs));
        const int used = nq < 64 ? nq : 64;
        qh = (qh + used) & (kQueue - 1);
        nq -= used;
      }
    } else {
      const double pxu = gpx[u], pyu = gpy[u], pzu = gpz[u];
      const auto ga = gmem((const int *)R.anchor) + 4 * u;
      const int4 Au = make_int4(ga[0], ga[1], ga[2], ga[3]);
      for (int h = 0; h < cnt; ++h) {
        const HomeRec &hr = L.home[h];
        const int ih = lo + h;
        double c2 = c2u;
        if (!uni) c2 = s_lcut2[(__builtin_amdgcn_readfirstlane(hr.th) - 1) * 8 + tj - 1];
        double nx, ny, nz;
        const double d2 = ref_d2(hr.px, hr.py, hr.pz, pxu, pyu, pzu, b0, b1, b2, nx, ny, nz);
        const bool ok = ((valid >> lane) & 1ull) && u != ih && c2 >= 0.0 &&
                        !(same && d2 < ptrmin) && !(d2 > c2);
        const int ox = (hr.ax - Au.x) - (int)nx * mx;
        const int oy = (hr.ay - Au.y) - (int)ny * my;
        const int oz = (hr.az - Au.z) - (int)nz * mz;
        // offsets beyond one cell cannot occur for mcell >= 3 or brute grids
        const bool bad = ok && (ox < -1 || ox > 1 || oy < -1 || oy > 1 || oz < -1 || oz > 1);
        const unsigned ent = (unsigned)u |
            ((unsigned)(((ox + 1) & 3) | (((oy + 1) & 3) << 2) | (((oz + 1) & 3) << 4)) << kCodeShift);
        append(C, nv, h, ih, ballot64(ok), ent, ballot64(bad) != 0);
      }
    }
    u = un;
    q = qn;
    xu = xn;
  }
  if (anch && ns > 0) append_slots(ns);
  wave_sync();   // the home records are read by every lane above
  const bool act = lane < cnt;
  if (act) R.nnb[lo + lane] = nv < 0 ? 0 : nv;
  if (act && nv >= 0) {
    // absent entries (the atom itself with the far offset code 63) up to a multiple of
    // kUnroll chunks, so that every unrolled read of the force kernel is a list entry
    const int ih = lo + lane;
    const auto bh = gmem(C.list + (size_t)(ih >> 7) * C.tstride + (ih & (kTile - 1)) * C.tpa);
    const unsigned msk = (unsigned)(C.tpa - 1), sh = (unsigned)C.tsh;
    const int chunks = (nv + C.tpa - 1) / C.tpa;
    const int end = (chunks + kUnroll - 1) / kUnroll * kUnroll * C.tpa;
    const unsigned pad = (unsigned)ih | (63u << kCodeShift);
    for (unsigned k = (unsigned)nv; k < (unsigned)end; ++k)
      bh[((k >> sh) << (7 + sh)) | (k & msk)] = pad;
  }
  if (ballot64(act && nv < 0) && lane == 0) atomicMax(&S.err, NDG_ERR_NEIMAX);
  long long mine = act && nv > 0 ? nv : 0;
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
  return mine;
}

// One tile of the list build, by the whole workgroup: the chunks of the cells whose first
// slot lies in the tile (the tile's own slots in brute mode), dealt to its waves in turn.
template <bool BIG>
__device__ inline void list_tile(const EngineDev &E, int t, ListWave &L, double &s_cmax,
                                 double &s_cmin) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const Tile T = E.tiles[t];
  const RepParams &R = E.reps[T.rep];
  RepState &S = E.st[T.rep];
  // S.rebuild does not change during this kernel (S.err can be raised meanwhile)
  if (!S.rebuild) return;
  // BIG: the tile's slots of the state arrays from tmp, here instead of in k_rb_place
  // (one launch less per step; c1 +6.4 % atom-steps/s).  Every replica on the work list
  // had no error when it was planned, and no kernel since raises one, so the order kernel
  // has permuted all of its atoms into tmp.  A mix of small replicas keeps k_rb_place: its
  // list build is bound by the latency of each workgroup's chain, which the copy would
  // lengthen (-0.5 %), and its order kernel leaves the derived arrays to k_rb_place
  // (-0.3 % when written there).  The pair tests read the positions from tmp either way.
  if (BIG) {
    const int p = T.atom0 + (int)threadIdx.x;
    if (threadIdx.x < kTile && p < R.np) place_state(R, p);
  }
  if (threadIdx.x < 64) {
    // the reference's per-pair list cutoff (src/nd.f90:1361,1375-1377), per type pair
    const int ti = lane / 8 + 1, tj = lane % 8 + 1;
    double c2 = -1.0;
    const bool in = ti <= R.ntypes && tj <= R.ntypes;
    if (in) {
      const int pij = R.pairmat[(ti - 1) + (tj - 1) * R.ntypes];
      const int pji = R.pairmat[(tj - 1) + (ti - 1) * R.ntypes];
      const int pm = pij > pji ? pij : pji;
      if (pm > 0) {
        const double rcs = R.skin + R.rc_pair[pm - 1];
        c2 = rcs * rcs;
      }
    }
    s_lcut2[lane] = c2;
    double mxv = c2, mnv = in ? c2 : 1e300;
    for (int off = 32; off > 0; off >>= 1) {
      mxv = fmax(mxv, shfl_xor_d(mxv, off));
      mnv = fmin(mnv, shfl_xor_d(mnv, off));
    }
    if (lane == 0) {
      s_cmax = mxv;
      s_cmin = mnv;
    }
  }
  __syncthreads();
  ListCtx C;
  C.list = R.list;
  C.tpa = R.tpa;
  C.tsh = C.tpa == 1 ? 0 : (C.tpa == 2 ? 1 : (C.tpa == 4 ? 2 : 3));
  C.kstride = (size_t)kTile * C.tpa;
  C.tstride = (size_t)R.list_chunks * C.kstride;
  C.neimax = R.neimax;
  C.ptrmin = R.ptrmin;
  C.px = R.tmpd;
  C.py = R.tmpd + R.np;
  C.pz = R.tmpd + 2 * R.np;
  C.uni = s_cmax == s_cmin;
  C.c2u = s_cmax;
  long long mine = 0;
  if (S.brute) {
    constexpr int kb = 16;
    for (int k = wid; k < kTile / kb; k += kListWaves) {
      const int lo = T.atom0 + kb * k;
      if (lo < R.np) mine += list_chunk(R, S, L, C, false, lo, min(kb, R.np - lo), -1, true);
    }
  } else {
    const int mx = S.mcell[0], my = S.mcell[1], mz = S.mcell[2];
    const double cut = sqrt(s_cmax) * (1.0 + 1e-6);
    const bool anch = (mx - 2) * (S.box[0] / mx) - 2.5e-4 * S.box[0] > cut &&
                      (my - 2) * (S.box[1] / my) - 2.5e-4 * S.box[1] > cut &&
                      (mz - 2) * (S.box[2] / mz) - 2.5e-4 * S.box[2] > cut;
    const int ncell = mx * my * mz;
    constexpr int chunk = BIG ? 64 : 32;
    int c = __builtin_amdgcn_readfirstlane(R.anchor[T.atom0].w);
    if (R.cell_start[c] < T.atom0) ++c;
    int k = 0;
    for (; c < ncell; ++c) {
      const int c_lo = __builtin_amdgcn_readfirstlane(R.cell_start[c]);
      if (c_lo >= T.atom0 + kTile) break;
      const int c_hi = __builtin_amdgcn_readfirstlane(R.cell_start[c + 1]);
      const int nch = (c_hi - c_lo + chunk - 1) / chunk;
      for (int j = 0; j < nch; ++j, ++k) {
        if (k % kListWaves != wid) continue;
        // near-equal chunks of the cell
        const int a0 = c_lo + (int)((long long)(c_hi - c_lo) * j / nch);
        const int a1 = c_lo + (int)((long long)(c_hi - c_lo) * (j + 1) / nch);
        mine += list_chunk(R, S, L, C, anch, a0, a1 - a0, c, !BIG);
      }
    }
  }
  if (lane == 0 && mine) atomicAdd((unsigned long long *)&S.nbr_sum, (unsigned long long)mine);
}

// At most 128 VGPRs (four waves per SIMD): two workgroups of a CU build lists at once.
// BIG: chunks of up to 64 home atoms, and the state arrays placed from tmp here (every
// replica has at least kBigTiles tiles).
template <bool BIG>
__global__ __launch_bounds__(64 * kListWaves) __attribute__((amdgpu_waves_per_eu(4)))
void k_rb_list(EngineDev E) {
  __shared__ ListWave s_wave[kListWaves];
  __shared__ double s_cmax, s_cmin;
  // one workgroup per entry of the work list (the grid covers every tile): the hardware
  // hands tiles to CUs as they free up, and a wave without a chunk ends at once
  const int t = rb_tile(E, blockIdx.x);
  if (t >= 0) list_tile<BIG>(E, t, s_wave[threadIdx.x >> 6], s_cmax, s_cmin);
}

// ------------------------------------------------------------------ bandwidth probe

// Plain streaming copy: the reference point for HBM bandwidth and for scaling the
// FETCH_SIZE / WRITE_SIZE counters.  Each lane has four 16-byte loads in flight (grid-stride
// by four), all issued before the stores; loads and stores are non-temporal (the data is
// touched once).  Measured on MI355X with 1 GiB: 4.5 TB/s with ordinary loads and 8192
// workgroups, 5.2-5.3 TB/s in this form.
typedef float copy4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy(const copy4 *__restrict__ src,
                                              copy4 *__restrict__ dst, size_t n) {
  const size_t S = (size_t)gridDim.x * blockDim.x;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * S < n; i += 4 * S) {
    copy4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = __builtin_nontemporal_load(&src[i + k * S]);
#pragma unroll
    for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(v[k], &dst[i + k * S]);
  }
  for (; i < n; i += S) dst[i] = src[i];
}

void launch_copy(const float4 *src, float4 *dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(256 * 128), dim3(256), 0, s, (const copy4 *)src, (copy4 *)dst, n);
}

// ------------------------------------------------------------------ launchers

static int grid_for(int nt, int per_cu) {
  int g = 256 * per_cu;
  if (nt < g) g = nt;
  if (g < 1) g = 1;
  return g;
}

// Per-atom streaming kernels: one workgroup per tile (a grid of every tile keeps enough
// waves in flight per CU to cover the memory latency; 0.083 vs 0.110 ms per drift over 20k
// tiles with eight workgroups per CU looping over their tiles).
void launch_drift(const EngineDev &E, int path, int t0, int nt, int xflags, hipStream_t s) {
  if (nt <= 0) return;
  if (path == NDG_PATH_FP64)
    hipLaunchKernelGGL(k_drift_fp64, dim3(nt), dim3(kTile), 0, s, E, t0, nt, xflags);
  else
    if (E.list_nt)
      hipLaunchKernelGGL(k_drift_mixed<true>, dim3(nt), dim3(kTile), 0, s, E, t0, nt, xflags);
    else
      hipLaunchKernelGGL(k_drift_mixed<false>, dim3(nt), dim3(kTile), 0, s, E, t0, nt, xflags);
}

void launch_kick(const EngineDev &E, int path, int t0, int nt, int xflags, hipStream_t s) {
  if (nt <= 0) return;
  if (path == NDG_PATH_FP64)
    hipLaunchKernelGGL(k_kick_fp64, dim3(nt), dim3(kTile), 0, s, E, t0, nt, xflags);
  else
    hipLaunchKernelGGL(k_kick_mixed, dim3(nt), dim3(kTile), 0, s, E, t0, nt, xflags);
}

// The force kernel with list loads past the last-level cache when the lists exceed it.
template <int TPA, bool FULL, bool SPR>
static void launch_force_nt(const EngineDev &E, int g, size_t lds, int t0, int nt, int xflags,
                            int kick, hipStream_t s) {
  using SH = ForceShape<TPA>;
  if (E.list_nt)
    hipLaunchKernelGGL((k_force_mixed<TPA, FULL, SPR, true>), dim3(g), dim3(SH::kThreads), lds, s, E, t0, nt, xflags, kick);
  else
    hipLaunchKernelGGL((k_force_mixed<TPA, FULL, SPR, false>), dim3(g), dim3(SH::kThreads), lds, s, E, t0, nt, xflags, kick);
}

template <int TPA>
static void launch_force_tpa(const EngineDev &E, bool full, int t0, int nt, int xflags, int kick,
                             hipStream_t s) {
  using SH = ForceShape<TPA>;
  const int nblk = (nt + SH::kNSub - 1) / SH::kNSub;
  const int g = grid_for(nblk, 2);
  const size_t lds = sizeof(float) * (size_t)E.tabset_n * (full && E.tabset_p ? 2 : 1);
  if (full && E.springs)
    launch_force_nt<TPA, true, true>(E, g, lds, t0, nt, xflags, kick, s);
  else if (full)
    launch_force_nt<TPA, true, false>(E, g, lds, t0, nt, xflags, kick, s);
  else if (E.springs)
    launch_force_nt<TPA, false, true>(E, g, lds, t0, nt, xflags, kick, s);
  else
    launch_force_nt<TPA, false, false>(E, g, lds, t0, nt, xflags, kick, s);
}

void launch_force(const EngineDev &E, int path, int tpa, bool full, int t0, int nt, int xflags,
                  hipStream_t s, int kick) {
  if (nt <= 0) return;
  if (path == NDG_PATH_FP64) {
    const dim3 g(grid_for(nt, 8));
    if (full && E.springs)
      hipLaunchKernelGGL((k_force_fp64<true, true>), g, dim3(kTile), 0, s, E, t0, nt, xflags);
    else if (full)
      hipLaunchKernelGGL((k_force_fp64<true, false>), g, dim3(kTile), 0, s, E, t0, nt, xflags);
    else if (E.springs)
      hipLaunchKernelGGL((k_force_fp64<false, true>), g, dim3(kTile), 0, s, E, t0, nt, xflags);
    else
      hipLaunchKernelGGL((k_force_fp64<false, false>), g, dim3(kTile), 0, s, E, t0, nt, xflags);
    return;
  }
  switch (tpa) {
    case 1: launch_force_tpa<1>(E, full, t0, nt, xflags, kick, s); break;
    case 2: launch_force_tpa<2>(E, full, t0, nt, xflags, kick, s); break;
    case 4: launch_force_tpa<4>(E, full, t0, nt, xflags, kick, s); break;
    default: launch_force_tpa<8>(E, full, t0, nt, xflags, kick, s); break;
  }
}

void launch_reduce(const EngineDev &E, int r0, int nr, int mode, int xflags, int trigger,
                   hipStream_t s) {
  if (nr <= 0) return;
  const int threads = E.max_rep_tiles > 1024 ? 1024 : 256;
  hipLaunchKernelGGL(k_reduce, dim3(nr < 1024 ? nr : 1024), dim3(threads), 0, s, E, r0, nr, mode,
                     xflags, trigger);
}

void launch_apply_pending(const EngineDev &E, int r0, int nr, hipStream_t s) {
  if (nr <= 0) return;
  hipLaunchKernelGGL(k_apply_pending, dim3(nr < 1024 ? nr : 1024), dim3(256), 0, s, E, r0, nr);
}

void launch_snapshot(const EngineDev &E, int r, int np, double *out, hipStream_t s) {
  int g = (np + 255) / 256;
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_snapshot, dim3(g), dim3(256), 0, s, E, r, out);
}

void launch_rebuild(const EngineDev &E, int nt, int r0, int nr, hipStream_t s) {
  // nt: tiles held by the replicas [r0, r0 + nr), a bound on the work list's length
  if (nt <= 0 || nr <= 0) return;
  const int gr = nr < 1024 ? nr : 1024;
  // the sort kernels stride over the work list: at most 2048 workgroups (eight per CU), so
  // that a step in which few tiles rebuild does not dispatch a workgroup per tile (c2, 16000
  // tiles rebuilt every ~45 steps: 1.017 -> 1.007 ms per step; c3 and c1 unchanged)
  const int g = nt < 2048 ? nt : 2048;
  if (nr <= kPlanLds) {
    hipLaunchKernelGGL(k_rb_bin_plan, dim3(g), dim3(kTile), sizeof(int) * (2 * (size_t)nr + 1), s,
                       E, r0, nr);
  } else {
    hipLaunchKernelGGL(k_rb_plan, dim3(1), dim3(1024), 0, s, E, r0, nr);
    hipLaunchKernelGGL(k_rb_bin, dim3(g), dim3(kTile), 0, s, E);
  }
  hipLaunchKernelGGL(k_rb_scan, dim3(gr), dim3(1024), 0, s, E, r0, nr);
  hipLaunchKernelGGL(k_rb_scatter, dim3(g), dim3(kTile), 0, s, E);
  if (E.list_big) {
    hipLaunchKernelGGL(k_rb_rank<true>, dim3(g), dim3(64 * kRankWaves), 0, s, E);
    hipLaunchKernelGGL(k_rb_list<true>, dim3(nt), dim3(64 * kListWaves), 0, s, E);
  } else {
    hipLaunchKernelGGL(k_rb_rank<false>, dim3(g), dim3(64 * kRankWaves), 0, s, E);
    hipLaunchKernelGGL(k_rb_place, dim3(g), dim3(kTile), 0, s, E);
    hipLaunchKernelGGL(k_rb_list<false>, dim3(nt), dim3(64 * kListWaves), 0, s, E);
  }
}

}  // namespace ndg
