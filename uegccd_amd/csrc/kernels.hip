// CDNA4 (gfx950) kernels of the CCD amplitude iteration.
//
// One residual pass (TermsWrapper, source/CCD.f90:146-181) plus the update
// (SolveCC/GetRes/CnvrgCC/CCEnergy, source/CCD.f90:186-462) is four launches:
//   k_slice         per virtual slice: the pass's amplitudes (the last update
//                   or its DIIS extrapolation, source/DIISCC.f90:52-131),
//                   linear ring / crossed-ring products, quadratic ring
//                   intermediates, mosaic partial sums
//                   (source/CCD.f90:507-585, 689-940)
//   k_contract      the particle-particle ladder, the hole-hole W rows and
//                   ladder as pair-momentum-blocked MFMA tiles
//                   (source/CCD.f90:611-681), the quadratic ring /
//                   crossed-ring products per momentum transfer, mosaic Joo
//   k_main          residual assembly, R = D*T - G, damped update, dT, energy,
//                   DIIS dot products
//   k_finalize      deterministic reductions, convergence test, DIIS solve
// Every kernel resolves its solve from a block-prefix table, so one launch
// covers any number of independent solves (twists, r_s values, bases).
//
// Interleaved layouts keep each gathered access one 16/32/64-byte load:
//   tcur, lad, histT/histE entries   double2  (abab, abba) or (S, P) per (i,j,a)
//   tk, hh                           double2  the valid elements by K block (column
//                                    ci of block K holds its pairs contiguously)
//   tv                               double2  (TV_ab + VT_ab, TV_ba + VT_ba) per (i,j,a)
//   jring / jxring                   double2  (J2, J3) per (row of P(q), l)
//   histT / histE                    valid elements only, in element order (compact
//                                    index cpos[x]), slot-major: k*nvalid + cpos[x]
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>

#include <algorithm>

#include "uegccd_internal.h"
#include "../../include/uegccd.h"

namespace uegccd {

// The descriptor as the kernels read it: its pointers are global memory, so
// loads and stores through them compile to global (not flat) instructions.
// (The host pass only checks device bodies; there the pointers stay plain.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using GlobalPtr = T __attribute__((address_space(1)))*;
#else
template <typename T>
using GlobalPtr = T*;
#endif
using SolveG = SolveDevT<GlobalPtr>;
static_assert(sizeof(SolveG) == sizeof(SolveDev), "one layout on host and device");
__device__ __forceinline__ const SolveG* as_global(const SolveDev* p) {
  return reinterpret_cast<const SolveG*>(p);
}
__device__ __forceinline__ double2 ld(GlobalPtr<const double2> p) { return *p; }
__device__ __forceinline__ double4 ld(GlobalPtr<const double4> p) { return *p; }
__device__ __forceinline__ int4 ld(GlobalPtr<const int4> p) { return *p; }
__device__ __forceinline__ uint2 ld(GlobalPtr<const uint2> p) { return *p; }
__device__ __forceinline__ void st2(GlobalPtr<double2> p, double2 v) { *p = v; }

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kStateWords = sizeof(State) / sizeof(double);
static_assert(kStateWords <= 2 * kFinThreads, "two state words per thread");

__device__ __forceinline__ int find_solve(const int* __restrict__ prefix, int n, int blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Element kernels renumber their blocks so that XCD x owns one contiguous
// range of amplitudes (a bijection on [0, nblocks)); the per-element streams
// (amplitudes, histories, slice products of the element's own slice) are then
// fetched into one L2 instead of all eight.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
  constexpr int kXcd = 8;
  const int per = nblocks / kXcd, rem = nblocks % kXcd;
  const int x = bid % kXcd, idx = bid / kXcd;
  return x * per + min(x, rem) + idx;
}

__device__ __forceinline__ int d2(int4 p, int4 q) {
  int a = p.x - q.x, b = p.y - q.y, c = p.z - q.z;
  return a * a + b * b + c * c;
}

// Lane `lane`'s x in every lane (a scalar read, no trip through the LDS
// crossbar as a shuffle takes; lane must be uniform).
__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, lane);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Block-wide reductions (wave64 shuffles, then LDS), fixed order.
__device__ __forceinline__ double wave_sum(double x) {
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
  for (int m = 32; m >= 1; m >>= 1) x = fmax(x, __shfl_xor(x, m, 64));
  return x;
}
__device__ double block_sum(double x, double* red) {
  x = wave_sum(x);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = x;
  __syncthreads();
  double r = 0.0;
  const int nw = blockDim.x >> 6;
  for (int k = 0; k < nw; ++k) r += red[k];
  return r;
}
__device__ double block_max(double x, double* red) {
  x = wave_max(x);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = x;
  __syncthreads();
  double r = 0.0;
  const int nw = blockDim.x >> 6;
  for (int k = 0; k < nw; ++k) r = fmax(r, red[k]);
  return r;
}

// Reduces NQ values per thread (NQ a power of two, 2..64) over the block by
// a reduce-scatter butterfly inside each wave: at the step with partner
// distance m a lane keeps half of the values it holds and hands the other
// half to its partner, so a wave exchanges NQ - 1 + log2(64/NQ) values instead
// of 6 NQ; then one LDS exchange combines the waves in wave order (a fixed
// order).  Quantities q < NMAX are maxima, the rest sums.  Threads q < NQ
// return quantity q of the block; the others return 0.  red must hold
// (blockDim/64)*NQ doubles.
// One halving step of the butterfly (W values held, partner distance
// 32 W / NQ), then the next; unrolled at compile time so v stays in registers.
template <int NQ, int NMAX, int W>
__device__ __forceinline__ void reduce_scatter_step(double (&v)[NQ], int lane, int& base) {
  if constexpr (W > 1) {
    constexpr int h = W / 2, m = 32 * W / NQ;
    const bool upper = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      const double send = upper ? v[i] : v[i + h];
      const double keep = upper ? v[i + h] : v[i];
      const double recv = __shfl_xor(send, m, 64);
      const int q = base + (upper ? h : 0) + i;
      v[i] = q < NMAX ? fmax(keep, recv) : keep + recv;
    }
    if (upper) base += h;
    reduce_scatter_step<NQ, NMAX, W / 2>(v, lane, base);
  }
}

template <int NQ, int NMAX>
__device__ __forceinline__ double block_reduce_scatter(double (&v)[NQ], double* red) {
  static_assert(NQ >= 2 && NQ <= 64 && (NQ & (NQ - 1)) == 0, "NQ a power of two");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int base = 0;  // first quantity of the block of values this lane holds
  reduce_scatter_step<NQ, NMAX, NQ>(v, lane, base);
  // lanes that share the quantity differ in their low bits: finish over them
  double x = v[0];
#pragma unroll
  for (int m = 32 / NQ; m >= 1; m >>= 1) {
    const double o = __shfl_xor(x, m, 64);
    x = base < NMAX ? fmax(x, o) : x + o;
  }
  __syncthreads();
  if ((lane & (64 / NQ - 1)) == 0) red[w * NQ + base] = x;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < NQ) {
    const int q = threadIdx.x;
    r = red[q];
    for (int k = 1; k < nw; ++k) r = q < NMAX ? fmax(r, red[k * NQ + q]) : r + red[k * NQ + q];
  }
  return r;
}

// The slice blocks' LDS tables: the occupied n-vectors and the Coulomb table
// v(|dn|^2) up to slice_n2(S) (zeros above it).
struct Lds {
  int4 nv[kMaxOcc];
  double vt[kMaxN2];
  double red[16];
};

// The last Coulomb-table entry a slice block stages: the table's own end, or
// kMaxN2 - 2 for a larger basis (the slice blocks read it at the differences
// of a virtual and an occupied momentum, which the host keeps below that).
__device__ __forceinline__ int slice_n2(const SolveG& S) { return min(S.n2max, kMaxN2 - 2); }

// Whether a pass has nothing to compute for a solve: it has finished, or it
// has converged in DIIS mode, where the last residual pass already holds the
// converged amplitudes' residual and the final pass only lays out the
// closing plain step (k_slice writes those amplitudes to tcur; the
// reference-mode final pass is the reference's residual check and runs).
__device__ __forceinline__ bool pass_idle(const SolveG& S, int phase) {
  return phase == PH_DONE || (phase == PH_FINAL && S.mode == UEGCCD_MODE_DIIS);
}

// Stages the occupied n-vectors and the Coulomb table of one slot.
__device__ __forceinline__ void stage(const SolveG& S, Lds& L, int slot) {
  const auto vt = S.vtab + (size_t)slot * (S.n2max + 1);
  const int n2 = slice_n2(S);
  for (int k = threadIdx.x; k < S.o; k += blockDim.x) L.nv[k] = S.nv4[k];
  for (int k = threadIdx.x; k <= n2; k += blockDim.x) L.vt[k] = vt[k];
  __syncthreads();
}

// Slot-major block partials: quantity q of element block lb.
__device__ __forceinline__ GlobalPtr<double> partial(const SolveG& S, int q, int lb) {
  return S.partials + (size_t)q * S.nblk_elem + lb;
}

// --------------------------------------------------------------------------
// Initial state: E(T0) and max|T0| (source/CCD.f90:62-64); also lays the
// starting amplitudes out by K block (tk), as k_slice does for later passes.
__global__ __launch_bounds__(kElemPerBlock) void k_init(const SolveDev* __restrict__ solves,
                                                         const int* __restrict__ prefix, int n,
                                                         int drpa) {
  __shared__ double red[16];
  const int gb = xcd_block(blockIdx.x, gridDim.x);
  const int s = find_solve(prefix, n, gb);
  const SolveG& S = as_global(solves)[s];
  const int lb = gb - prefix[s];
  const int o = S.o, o2 = o * o;
  // v(n_p - n_k) of a virtual p and an occupied k, from the energy's table
  const auto vs = S.vso + (size_t)S.rslot[RO_ENERGY] * S.v * o;
  const int c = lb * blockDim.x + threadIdx.x;  // compact index of a valid element
  double e = 0.0, tmax = 0.0, bad = 0.0;
  if (c < S.nvalid) {
    const uint2 cm = ld(S.cmap + c);
    const int x = elem_x(S.ecode, cm);
    const int i = x % o, j = (x / o) % o, av = x / o2;
    const double2 t = S.tcur[x];
    st2(S.tk + elem_pos(S.ecode, cm), t);
    const double taa = t.x + t.y;
    const double vai = vs[av * o + i];
    if (drpa == 1)  // the dRPA module's CCEnergy with CSosex 0 (source/dRPA.f90:200-208)
      e = vai * t.x * 2;
    else if (drpa == 2)  // the same with CSosex 1: Ec(SOSEX)
      e = vai * t.x * 2 - vs[(elem_partner(S.ecode, cm) - o) * o + i] * t.x;
    else
      e = 0.5 * (vai * (taa + t.x) - vs[av * o + j] * (taa + t.y));
    tmax = fmax(fabs(t.x), fabs(t.y));
    if (!isfinite(t.x) || !isfinite(t.y)) bad = 1.0;
  }
  e = block_sum(e, red);
  tmax = block_max(tmax, red);
  bad = block_sum(bad, red);
  if (threadIdx.x == 0) {
    *partial(S, 0, lb) = e;
    *partial(S, 1, lb) = tmax;
    *partial(S, 2, lb) = 0.0;
    *partial(S, 3, lb) = bad;
  }
}

// DrvMBPT amplitudes and E(2) (source/MP2.f90:56-74); writes tcur.
__global__ __launch_bounds__(kElemPerBlock) void k_mp2(const SolveDev* __restrict__ solves,
                                                        const int* __restrict__ prefix, int n,
                                                        int drpa) {
  __shared__ double red[16];
  const int gb = xcd_block(blockIdx.x, gridDim.x);
  const int s = find_solve(prefix, n, gb);
  const SolveG& S = as_global(solves)[s];
  const int lb = gb - prefix[s];
  const int o = S.o, o2 = o * o;
  // DrvMBPT's ERI calls pass no flag (source/MP2.f90:61-62): table slot 0
  const auto vs = S.vso;
  const int c = lb * blockDim.x + threadIdx.x;  // compact index of a valid element
  double e = 0.0;
  if (c < S.nvalid) {
    const uint2 cm = ld(S.cmap + c);
    const int x = elem_x(S.ecode, cm), b = elem_partner(S.ecode, cm);
    const int i = x % o, j = (x / o) % o, av = x / o2, a = o + av;
    const double vab = vs[av * o + i], vba = vs[(b - o) * o + i];
    const double den = S.eig[i] + S.eig[j] - S.eig[a] - S.eig[b];
    const double tab = vab / den, tba = -vba / den;
    const double taa = tab + tba;
    e = 0.5 * taa * (vab - vba) + 0.5 * tab * vab - 0.5 * tba * vba;
    // dRPA starts from T2aaaa = T2abab + T2abba (source/Main.f90:85,91)
    S.tcur[x] = drpa ? make_double2(taa, 0.0) : make_double2(tab, tba);
  }
  e = block_sum(e, red);
  if (threadIdx.x == 0) {
    *partial(S, 0, lb) = e;
    *partial(S, 1, lb) = 0.0;
    *partial(S, 2, lb) = 0.0;
    *partial(S, 3, lb) = 0.0;
  }
}

// FP64/FP32 16x16x4 MFMA (the slice products, ladders and contractions).
template <typename Real>
struct MfmaTraits;
// Operands of both forms: lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result rows differ (measured on gfx950): the f64 form returns
// D[(l/16) + 4r][l%16], the f32 form D[4(l/16) + r][l%16], r = 0..3.
template <>
struct MfmaTraits<double> {
  typedef v4d acc_t;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int kq, int r) { return kq + 4 * r; }
};
template <>
struct MfmaTraits<float> {
  typedef v4f acc_t;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int kq, int r) { return 4 * kq + r; }
};

// --------------------------------------------------------------------------
// Per-virtual-slice work, one workgroup per slice s (LDS-staged).
//
// With T(:,:,s) of both spins and V(k,j) = v(n_k - n_j) in LDS, and using that
// invalid amplitudes are exactly zero (so the reference's FindIndex skips
// only skip zero terms), every linear ring / crossed-ring term of
// source/CCD.f90:716-763, 845-892 is a product of a slice with V:
//   TV[s](i,j) = sum_k T(i,k,s) V(k,j)      VT[s](j,i) = sum_k V(j,k) T(k,i,s)
//   rsum[s](i) = sum_k (T^aa + T^ab)(i,k,s) csum[s](i) = sum_k (T^aa + T^ba)(k,i,s)
// The quadratic intermediates (source/CCD.f90:771-797, 899-925) share one
// Coulomb factor Y(k,l) = v(n_i + n_k - n_s - n_l) between rings (a = s) and
// crossed rings (b = s), with q = n_i - n_s:
//   J1 = v(q) rsum(i) - sum_k T^aa(i,k,s) Y    J1x = v(q) csum(i) - sum_k T^aa(k,i,s) Y
//   J2 = v(q) rsum(i) - sum_k T^ab(i,k,s) Y    J2x = v(q) csum(i) - sum_k T^ba(k,i,s) Y
//   J3 =              - sum_k T^ba(i,k,s) Y    J3x =                sum_k T^ab(k,i,s) Y
// (the momentum test on d is applied where J is contracted).  T^aa = T^ab +
// T^ba makes J1 = J2 + J3 and J1x = J2x - J3x, so only J2, J3 are stored.
// A residual element takes the sum TV + VT of each spin at (i,j) of its slice
// a and at (j,i) of its partner's slice b, so that sum is what is stored.
// The mosaic
// sums of source/CCD.f90:548-567 are per-slice reductions as well:
//   Jvv(s) = sum_k P_s(k),  Joo(k) = sum_s P_s(k)  (k_joo),
//   P_s(k) = sum_l v(n_s-n_k)(T^aa+T^ab)(k,l,s) - v(n_s-n_l)(T^aa+T^ba)(k,l,s).
// --------------------------------------------------------------------------
// Small dense solve for the DIIS coefficients (FP64, partial pivoting), by one
// wave64: lane r owns row r of the bordered system
//   [ B/scale  -1 ] [c]   [ 0]
//   [ -1^T      0 ] [l] = [-1]
// over the nh newest history slots; the rows live in registers (columns
// unrolled), pivots and row swaps go through cross-lane shuffles.
// M is the unrolled system size (at least nh + 1).
template <int M>
__device__ bool diis_coefficients(const double* bmat, int nh, int slot_cur, int ndiis, double* coef) {
  const int lane = threadIdx.x & 63;
  const int m = nh + 1;
  auto slot_of = [&](int r) { return ((slot_cur - (nh - 1 - r)) % ndiis + ndiis) % ndiis; };
  const int slr = lane < nh ? slot_of(lane) : 0;
  double scale = lane < nh ? fabs(bmat[slr * kMaxDiis + slr]) : 0.0;
  scale = wave_max(scale);
  if (!(scale > 0.0) || !isfinite(scale)) return false;
  double row[M];
#pragma unroll
  for (int c = 0; c < M; ++c) {
    double val = 0.0;
    if (lane < nh) {
      if (c < nh) val = bmat[slr * kMaxDiis + slot_of(c)] / scale;
      else if (c == nh) val = -1.0;
    } else if (lane == nh && c < nh) {
      val = -1.0;
    }
    row[c] = val;
  }
  double rhs = lane == nh ? -1.0 : 0.0;
  bool ok = true;
#pragma unroll
  for (int c = 0; c < M; ++c) {
    if (c >= m || !ok) continue;
    double best = (lane >= c && lane < m) ? fabs(row[c]) : -1.0;
    int bl = lane;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int ol = __shfl_xor(bl, off, 64);
      if (ob > best || (ob == best && ol < bl)) { best = ob; bl = ol; }
    }
    if (!(best >= 1e-14)) { ok = false; continue; }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double vc = __shfl(row[k], c, 64), vp = __shfl(row[k], bl, 64);
      if (lane == c) row[k] = vp; else if (lane == bl) row[k] = vc;
    }
    {
      const double vc = __shfl(rhs, c, 64), vp = __shfl(rhs, bl, 64);
      if (lane == c) rhs = vp; else if (lane == bl) rhs = vc;
    }
    const double piv = __shfl(row[c], c, 64);
    const double f = (lane > c && lane < m) ? row[c] / piv : 0.0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double pk = __shfl(row[k], c, 64);
      if (lane > c && lane < m) row[k] -= f * pk;
    }
    const double pr = __shfl(rhs, c, 64);
    if (lane > c && lane < m) rhs -= f * pr;
  }
  if (!ok) return false;
  double xs = 0.0;
#pragma unroll
  for (int r = M - 1; r >= 0; --r) {
    if (r >= m) continue;
    double acc = rhs;
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (k > r && k < m) acc -= row[k] * __shfl(xs, k, 64);
    if (lane == r) xs = acc / row[r];
  }
  const bool fin = isfinite(xs) || lane >= nh;
  if (!__all(fin)) return false;
  if (lane < ndiis) coef[lane] = 0.0;
  if (lane < nh) coef[slr] = xs;
  return true;
}


// The same coefficients for histories of up to 8 entries, by one wave:
// c = B^-1 1 / (1^T B^-1 1) (the solution of the bordered system) with B
// scaled by its largest diagonal and padded to 8 x 8 with the identity.
// Lane 8r + c holds B[r][c]; Gauss-Jordan elimination without pivoting (B is
// symmetric positive definite, so its pivots are those of a Cholesky
// factorisation) brings B to the identity and 1 to y = B^-1 1, one pivot per
// step, its row and column broadcast by shuffles.  A pivot below 1e-14 fails
// the extrapolation (the plain update is used).  Every lane of the wave must
// call it; the result is the same on all of them.
__device__ bool diis_gauss_jordan8(const double* bmat, int nh, int slot_cur, int ndiis,
                                   double* coef) {
  constexpr int M = 8;
  const int lane = threadIdx.x & 63, r = lane >> 3, c = lane & 7;
  // history slot of row q, oldest first
  auto slot_of = [&](int q) {
    int sv = slot_cur - (nh - 1) + q;
    if (sv < 0) sv += ndiis;
    return sv >= ndiis ? sv - ndiis : sv;
  };
  const int sr = slot_of(r), sc = slot_of(c);
  double scale = (lane < nh) ? fabs(bmat[slot_of(lane) * (kMaxDiis + 1)]) : 0.0;
  scale = wave_max(scale);
  if (!(scale > 0.0) || !isfinite(scale)) return false;
  const double inv_scale = 1.0 / scale;
  double a = (r < nh && c < nh) ? bmat[sr * kMaxDiis + sc] * inv_scale : (r == c ? 1.0 : 0.0);
  double b = 1.0;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    // (the pivot and its right-hand side are read as scalars, so that the
    // division runs while the row and column shuffles are on their way)
    const double p = readlane_f64(a, k * M + k);
    if (!(p > 1e-14)) return false;
    const double inv_p = 1.0 / p;
    const double rk = __shfl(a, k * M + c, 64);  // pivot row, this lane's column
    const double ck = __shfl(a, r * M + k, 64);  // this lane's row, pivot column
    const double bk = readlane_f64(b, k * M);
    if (r == k) {
      a *= inv_p;
      b *= inv_p;
    } else {
      const double f = ck * inv_p;
      a -= f * rk;
      b -= f * bk;
    }
  }
  // y_r on lane 8r; the sum over the filled rows in row order
  double sum = 0.0;
#pragma unroll
  for (int q = 0; q < M; ++q) {
    const double yq = readlane_f64(b, q * M);
    if (q < nh) sum += yq;
  }
  if (!(fabs(sum) > 0.0) || !isfinite(sum)) return false;
  const double cr = b * (1.0 / sum);
  if (!__all(isfinite(cr) || r >= nh)) return false;
  if (lane < ndiis) coef[lane] = 0.0;
  if (c == 0 && r < nh) coef[sr] = cr;
  return true;
}

// A value the same in every lane, as a scalar (so that branches on it are
// scalar branches).
__device__ __forceinline__ double uniform(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Per-block LDS of k_slice for the DIIS solve: the B matrix and the
// coefficients of the slab's amplitudes by history slot.
struct SliceDiis {
  double bm[kMaxDiis * kMaxDiis];
  double coef[kMaxDiis];
};

// Amplitude of one valid element for the pass, sum_k coef[k] histT[k] in
// slot order over the nonzero coefficients (the DIIS extrapolation, or one
// coefficient 1 for the plain update, which the sum reproduces exactly), from
// its preloaded history entries hm and the coefficients cf of slots below P;
// later slots are loaded here from HC (the element's entry of slot 0).
template <int P>
__device__ __forceinline__ double2 combine_history(const double (&cf)[P], const double2 (&hm)[P],
                                                   const double* coef, int ndiis,
                                                   GlobalPtr<const double2> HC, size_t nval) {
  double ab = 0.0, ba = 0.0;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const double c = cf[k];
    if (c == 0.0) continue;
    ab += c * hm[k].x;
    ba += c * hm[k].y;
  }
  for (int k = P; k < ndiis; ++k) {
    const double c = uniform(coef[k]);
    if (c == 0.0) continue;
    const double2 hk = ld(HC + (size_t)k * nval);
    ab += c * hk.x;
    ba += c * hk.y;
  }
  return make_double2(ab, ba);
}

// A valid element's amplitude to its slab entry in LDS, to tcur and to tk;
// cm = elem_map(element, position in tk, partner), xs the slab's first element.
__device__ __forceinline__ void slab_place(const SolveG& S, double* Tab, double* Tba, int xs, uint2 cm,
                                           double2 t) {
  const int x = elem_x(S.ecode, cm);
  Tab[x - xs] = t.x;
  Tba[x - xs] = t.y;
  st2(S.tcur + x, t);
  st2(S.tk + elem_pos(S.ecode, cm), t);
}

// One block of BT threads per virtual slice s (BT = slice_threads of the
// largest NOcc among the launch's solves): one slab entry and one J output per thread
// (a larger slab loops over the rest).
// SPLIT: the ring and crossed-ring roles read different Coulomb tables
// (range flags): a second V slab Vx for the crossed-ring products, v(n_s -
// n_k) of the quadratic ring and crossed-ring directs (vsr, vsx) beside the
// mosaic's (vs), and a second Y table vt2 for the crossed rings.
// fin_prev: the block first runs the reductions and state update of the
// previous pass (finalize_state, the same in every block; the block of slice
// 0 writes the new state to st2), with fvals, fred ((BT/64)*32 doubles) and
// fsst as its LDS.
template <int MAXQ>
__device__ __forceinline__ int finalize_state(const SolveG& S, int kind, int prec64, double* vals, double* red,
                                              State& sst, bool write_trace);
__device__ __forceinline__ void put_state(const State& sst, GlobalPtr<State> to);
// The loads of a state update, requested ahead of its reductions: a thread's
// state words (words t and t + kFinThreads), and its first round of block
// partials (blocks k and k + kFinThreads of thread k; the 32-wide form one
// block a round, for registers).
struct StateLoads {
  double sw0, sw1;
};
template <int NQ>
struct PartialLoads {
  double p[NQ], p2[NQ];
};
__device__ __forceinline__ void state_load(const SolveG& S, StateLoads& L);
__device__ __forceinline__ int finalize_dots(const SolveG& S, int kind);
template <int NQ>
__device__ __forceinline__ void partials_load(const SolveG& S, int nd, int k, PartialLoads<NQ>& L);
template <int NQ>
__device__ __forceinline__ void partials_finish(const SolveG& S, int nd, PartialLoads<NQ>& L, double* vals,
                                                double* red);
__device__ __forceinline__ int finalize_apply(const SolveG& S, int kind, int prec64, const double* vals,
                                              State& sst, bool write_trace, const StateLoads& SL);

template <int BT, bool SPLIT>
__device__ __forceinline__ void slice_block(const SolveDev* __restrict__ solves, const int* __restrict__ prefix,
                            int n, int bid, Lds& L, SliceDiis& D, double* dyn, double* vt2, int fin_prev, FoldArgs fa,
                            double* fvals, double* fred, State* fsst) {
  const int s = find_solve(prefix, n, bid);
  const SolveG& S = as_global(solves)[s];
  const int sv = bid - prefix[s];
  const int o = S.o, o2 = o * o, so = o + sv;
  double* Tab = dyn;
  double* Tba = dyn + o2;
  double* Vo = dyn + 2 * o2;
  double* rs = dyn + 3 * o2;
  double* cs = rs + o;
  double* jp = cs + o;
  double* vs = jp + o;  // v(n_s - n_k)
  double* Vx = SPLIT ? vs + o : Vo;
  double* vsr = SPLIT ? Vx + o2 : vs;
  double* vsx = SPLIT ? vsr + o : vs;
  const size_t vo = (size_t)S.v * o;
  const auto voo_r = S.voo + (size_t)S.rslot[RO_RING_LX] * o2;
  const auto vso_m = S.vso + S.rslot[RO_MOSAIC] * vo;
  // The pass's amplitudes on the slab (source/CCD.f90:80-99 with
  // source/DIISCC.f90:52-131): tcur as it stands after a start, else
  // sum_k c_k histT[k]: the DIIS extrapolation, whose coefficients wave 0 of
  // every block solves for (the same inputs and instructions in every block,
  // so the same coefficients) while the history is on its way, or the last
  // update histT[slot_cur] alone.  The slab's valid elements are one
  // contiguous run of the history (thread t takes the t-th); the block writes
  // them back to tcur and to the K-blocked copy tk, which the later kernels
  // of the pass read, and zeros stand at the slab's invalid entries.  The
  // history (up to kPre slots), the occupied Coulomb block V(k, j), the state
  // and its B matrix go out in the same round of loads as the tables.
  constexpr int kPre = 8;
  const int ndiis = S.ndiis;
  const size_t nval = S.nvalid;
  const int xs = o2 * sv;  // the slab's first element
  const int c0 = S.slab_coff[sv], nvs = S.slab_coff[sv + 1] - c0;
  const int t0 = threadIdx.x;
  // the previous pass's state words and block partials go out first, so that
  // its update waits on them alone, not on the history behind them
  StateLoads fsl{0.0, 0.0};
  PartialLoads<16> fpl;
  if (fin_prev) {
    // (state_load and partials_load<16> from the kernel arguments, so that
    // the loads wait on no descriptor read)
    const auto stw = (GlobalPtr<const double>)(uintptr_t)fa.st;
    const int w1 = t0 + kFinThreads;
    fsl.sw0 = t0 < kStateWords ? stw[t0] : 0.0;
    fsl.sw1 = t0 < kFinThreads && w1 < kStateWords ? stw[w1] : 0.0;
    const auto pp = (GlobalPtr<const double>)(uintptr_t)fa.partials;
    const int k2 = t0 + kFinThreads;
    const bool one = t0 < fa.nblk && t0 < kFinThreads, two = one && k2 < fa.nblk;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int src = q == 0 ? 1 : q == 1 ? 2 : q == 2 ? 0 : q;
      fpl.p[q] = src < 4 + fa.nd && one ? pp[(size_t)src * fa.nblk + t0] : 0.0;
      fpl.p2[q] = src < 4 + fa.nd && two ? pp[(size_t)src * fa.nblk + k2] : 0.0;
    }
  }
  const auto HC = S.histT + c0 + t0;
  const uint2 cm = t0 < nvs ? ld(S.cmap + c0 + t0) : make_uint2(0u, 0u);
  const double vr = t0 < o2 ? voo_r[t0] : 0.0;
  double2 h[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k)
    h[k] = (t0 < nvs && k < ndiis) ? ld(HC + (size_t)k * nval) : make_double2(0.0, 0.0);
  const double vsm = t0 < o ? vso_m[sv * o + t0] : 0.0;
  // the LDS tables' entries (the occupied n-vectors and the Y table up to
  // slice_n2, zeros above it), requested with the history too and stored
  // once the state is known
  // (two entries a thread in blocks of 512 and 768 threads, whose registers
  // have room: the whole table; 256-thread blocks load theirs after the update)
  constexpr int kVt = BT >= 512 ? 2 : 0;
  const auto vtq = S.vtab + (size_t)S.rslot[RO_RING_QX] * (S.n2max + 1);
  const int n2s = slice_n2(S);
  double vtr[kVt > 0 ? kVt : 1];
#pragma unroll
  for (int m = 0; m < kVt; ++m) vtr[m] = t0 + m * BT <= n2s ? vtq[t0 + m * BT] : 0.0;
  const int4 nvo = kVt > 0 && t0 < o ? ld(S.nv4 + t0) : make_int4(0, 0, 0, 0);
  // the state of this pass (requested after the history, which does not
  // depend on it)
  int phase, src, nh, slot_cur;
  double bw;
  if (fin_prev) {
    partials_finish<16>(S, finalize_dots(S, FIN_ITER), fpl, fvals, fred);
    finalize_apply(S, FIN_ITER, 1, fvals, *fsst, sv == 0, fsl);
    if (sv == 0) put_state(*fsst, S.st2);
    phase = fsst->phase;
    src = fsst->next_src;
    nh = fsst->nhist;
    slot_cur = fsst->slot_cur;
    bw = t0 < kMaxDiis * kMaxDiis ? fsst->bmat[t0] : 0.0;
  } else {
    const State& st = *S.st;
    phase = st.phase;
    src = st.next_src;
    nh = st.nhist;
    slot_cur = st.slot_cur;
    bw = t0 < kMaxDiis * kMaxDiis ? st.bmat[t0] : 0.0;
  }
  if (t0 < kMaxDiis * kMaxDiis) D.bm[t0] = bw;
  if (t0 < kMaxDiis) D.coef[t0] = t0 == slot_cur ? 1.0 : 0.0;
  for (int idx = t0; idx < o2; idx += BT) {
    Tab[idx] = 0.0;
    Tba[idx] = 0.0;
    Vo[idx] = idx == t0 ? vr : voo_r[idx];
  }
  if (SPLIT) {
    const auto voo_x = S.voo + (size_t)S.rslot[RO_XR_LX] * o2;
    const auto vso_r = S.vso + S.rslot[RO_RING_QD] * vo, vso_x = S.vso + S.rslot[RO_XR_QD] * vo;
    const auto vt_x = S.vtab + (size_t)S.rslot[RO_XR_QX] * (S.n2max + 1);
    for (int idx = t0; idx < o2; idx += BT) Vx[idx] = voo_x[idx];
    for (int k = t0; k < o; k += BT) {
      vsr[k] = vso_r[sv * o + k];
      vsx[k] = vso_x[sv * o + k];
    }
    for (int k = t0; k < kMaxN2; k += BT) vt2[k] = k <= slice_n2(S) ? vt_x[k] : 0.0;
  }
#pragma unroll
  for (int m = 0; m < kVt; ++m)
    if (t0 + m * BT < kMaxN2) L.vt[t0 + m * BT] = vtr[m];
  for (int k = t0 + kVt * BT; k < kMaxN2; k += BT) L.vt[k] = k <= n2s ? vtq[k] : 0.0;
  if (t0 < o) L.nv[t0] = kVt > 0 ? nvo : ld(S.nv4 + t0);
  __syncthreads();
  if (phase == PH_DONE) return;  // the whole block: the solve has finished
  if (src == SRC_DIIS) {  // a failed solve leaves the plain update
    if (t0 < 64) {
      if (ndiis <= 8)
        diis_gauss_jordan8(D.bm, nh, slot_cur, ndiis, D.coef);
      else
        diis_coefficients<kMaxDiis + 1>(D.bm, nh, slot_cur, ndiis, D.coef);
    }
    __syncthreads();
  }
  // (Y below is looked up at n_i + n_k - n_s - n_l = n_b - n_l, b the
  // partner of (i,k,s): a virtual-occupied difference, within the staged
  // table.  Without a partner the index can leave the table; those terms
  // multiply a zero amplitude, so the table reads as zero there: the last
  // entry is always past the staged ones.)
  if (src == SRC_TCUR) {
    for (int idx = t0; idx < o2; idx += BT) {
      const double2 t = ld(S.tcur + xs + idx);
      Tab[idx] = t.x;
      Tba[idx] = t.y;
    }
  } else {
    double cf[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) cf[k] = k < ndiis ? uniform(D.coef[k]) : 0.0;
    if (t0 < nvs) slab_place(S, Tab, Tba, xs, cm, combine_history<kPre>(cf, h, D.coef, ndiis, HC, nval));
    for (int c = t0 + BT; c < nvs; c += BT) {  // slabs of more than BT valid elements
      const auto HCc = S.histT + c0 + c;
      double2 hm[kPre];
#pragma unroll
      for (int k = 0; k < kPre; ++k) hm[k] = k < ndiis ? ld(HCc + (size_t)k * nval) : make_double2(0.0, 0.0);
      slab_place(S, Tab, Tba, xs, ld(S.cmap + c0 + c), combine_history<kPre>(cf, hm, D.coef, ndiis, HCc, nval));
    }
  }
  if (pass_idle(S, phase)) return;  // the closing step's amplitudes are in tcur (DIIS mode)
  if ((int)threadIdx.x < o) vs[threadIdx.x] = vsm;
  __syncthreads();
  // Row/column sums (for J) by groups of 8 lanes per row i, each lane a
  // stride-8 share of k, then a 3-step shuffle tree; stored with the mosaic
  // partials P_s(i) = sum_k v(s-i)(2T^ab + T^ba)(i,k) - v(s-k)(T^ab + 2T^ba)(i,k).
  for (int i0 = 0; i0 < o; i0 += BT / 8) {
    const int i = i0 + (int)threadIdx.x / 8, sub = threadIdx.x & 7;
    double r = 0.0, c = 0.0, p = 0.0;
    if (i < o) {
      const double vsi = vs[i];
      for (int k = sub; k < o; k += 8) {
        const double ta = Tab[i + o * k], tb = Tba[i + o * k];
        r += 2.0 * ta + tb;
        c += Tab[k + o * i] + 2.0 * Tba[k + o * i];
        p += vsi * (2.0 * ta + tb) - vs[k] * (ta + 2.0 * tb);
      }
    }
#pragma unroll
    for (int m = 4; m >= 1; m >>= 1) {
      r += __shfl_xor(r, m, 64);
      c += __shfl_xor(c, m, 64);
      p += __shfl_xor(p, m, 64);
    }
    if (i < o && sub == 0) {
      rs[i] = r;
      cs[i] = c;
      jp[i] = p;
      S.rsum[sv * o + i] = r;
      S.csum[sv * o + i] = c;
      S.joo_part[sv * o + i] = p;
    }
  }
  // The slice products TV = T V and VT = V T as 16x16 FP64 MFMA tiles over
  // (r, c), one spin per work unit (T^ab units write the first component of
  // tv, T^ba units the second), stored as the sum of the products the ring
  // flags turn on.  The units go to the block's last waves, so that they run
  // beside the row sums above (its first waves).
  {
    const double fr = S.do_ring ? 1.0 : 0.0, fx = S.do_xring ? 1.0 : 0.0;
    typedef MfmaTraits<double>::acc_t acc_t;
    const int nt = (o + 15) >> 4, nunits = 2 * nt * nt;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
    constexpr int nw = BT / 64;
    const auto tvd = reinterpret_cast<GlobalPtr<double>>(S.tv) + 2 * (size_t)o2 * sv;
    for (int u = nw - 1 - wave; u < nunits; u += nw) {
      const int t = u >> 1, half = u & 1;
      const double* TT = half ? Tba : Tab;
      const int rt = t / nt, ct = t % nt;
      const int ra = rt * 16 + r16, cb = ct * 16 + r16;
      const int rac = min(ra, o - 1), cbc = min(cb, o - 1);
      acc_t tv = {0, 0, 0, 0}, vt = {0, 0, 0, 0};
      for (int k0 = 0; k0 < o; k0 += 4) {
        // in-range LDS addresses, masked values: no branch between the loads
        const int k = k0 + kq, kc = min(k, o - 1);
        const bool ak = ra < o && k < o, bk = cb < o && k < o;
        const double t_rk = TT[rac + o * kc], v_rk = Vx[rac + o * kc];
        const double v_kc = Vo[kc + o * cbc], t_kc = TT[kc + o * cbc];
        tv = MfmaTraits<double>::mma(ak ? t_rk : 0.0, bk ? v_kc : 0.0, tv);
        vt = MfmaTraits<double>::mma(ak ? v_rk : 0.0, bk ? t_kc : 0.0, vt);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + MfmaTraits<double>::row(kq, r);
        if (row < o && cb < o) tvd[2 * (row + o * cb) + half] = fr * tv[r] + fx * vt[r];
      }
    }
  }
  __syncthreads();
  // Jvv(s) = sum_i P_s(i) by one wave (o <= 64), a fixed tree
  if (threadIdx.x < 64) {
    double jv = (int)threadIdx.x < o ? jp[threadIdx.x] : 0.0;
    jv = wave_sum(jv);
    if (threadIdx.x == 0) S.jvv[sv] = jv;
  }
  // The quadratic intermediates J(i, l), kJL values of l per thread (l = g,
  // g + G, ...): the four amplitudes of (i, k) and (k, i) are read once per k
  // for all of them.  Consecutive threads take consecutive g of one i, so
  // their stores fill runs of one J row.
  constexpr int kJL = 3;
  const int G = (o + kJL - 1) / kJL;
  const int4 ns = S.nv4[so];
  for (int idx = threadIdx.x; idx < o * G; idx += BT) {
    const int i = idx / G, g = idx - i * G;
    const int4 ni = L.nv[i];
    int bx[kJL], by[kJL], bz[kJL];
#pragma unroll
    for (int m = 0; m < kJL; ++m) {
      const int4 nl = L.nv[min(g + m * G, o - 1)];
      bx[m] = ni.x - ns.x - nl.x;
      by[m] = ni.y - ns.y - nl.y;
      bz[m] = ni.z - ns.z - nl.z;
    }
    double a2[kJL], a3[kJL], a5[kJL], a6[kJL];
#pragma unroll
    for (int m = 0; m < kJL; ++m) a2[m] = a3[m] = a5[m] = a6[m] = 0.0;
    // (unrolled so that several table lookups, each waiting on its n_k, are
    // in flight at once; |d| < 2^11, so 24-bit multiplies are exact)
#pragma unroll 2
    for (int k = 0; k < o; ++k) {
      const int4 nk = L.nv[k];
      const double ta = Tab[i + o * k], tb = Tba[i + o * k];
      const double ua = Tab[k + o * i], ub = Tba[k + o * i];
#pragma unroll
      for (int m = 0; m < kJL; ++m) {
        const int dx = bx[m] + nk.x, dy = by[m] + nk.y, dz = bz[m] + nk.z;
        const int yi = min(__mul24(dx, dx) + __mul24(dy, dy) + __mul24(dz, dz), kMaxN2 - 1);
        const double y = L.vt[yi];
        const double yx = SPLIT ? vt2[yi] : y;
        a2[m] += ta * y;
        a3[m] += tb * y;
        a5[m] += ub * yx;
        a6[m] += ua * yx;
      }
    }
    const double vq = vsr[i], vqx = vsx[i];
    const size_t jrow = (size_t)S.qrow[sv * o + i] * o;
#pragma unroll
    for (int m = 0; m < kJL; ++m) {
      const int l = g + m * G;
      if (l >= o) continue;
      st2(S.jring + jrow + l, make_double2(vq * rs[i] - a2[m], -a3[m]));
      st2(S.jxring + jrow + l, make_double2(vqx * cs[i] - a5[m], a6[m]));
    }
  }
}

// --------------------------------------------------------------------------
// Ladders as GEMMs per pair-momentum block K (source/CCD.f90:611-681).
//
// With S = T^ab + T^ba and P = T^ab - T^ba over the block's pairs p and
// virtual columns c, the particle-particle ladder is
//   lad_S[p,a] = 1/2 sum_c S[p,c] (v(c-a) - v(c-b(a)))
//   lad_P[p,a] = 1/2 sum_c P[p,c] (v(c-a) + v(c-b(a)))
// and the hole-hole W intermediates are the same product with the pair (k,l)
// in place of (a,b):
//   W_S[p,q] = v(i-k) - v(i-l) + 1/2 sum_c S[p,c] (v(c-k) - v(c-l))
//   W_P[p,q] = v(i-k) + v(i-l) + 1/2 sum_c P[p,c] (v(c-k) + v(c-l))
// so one tile routine serves both: rows = pairs of K (16 per tile), columns
// = the block's a list (particle-particle tiles) or its pair list (W).  The
// B operand is never stored: each lane synthesises its element from two
// n-vectors and the Coulomb table.  One wave64 per 16x16 tile,
// v_mfma_f64_16x16x4f64 (or the f32 form while the mixed-precision schedule
// is in its FP32 phase).  The W columns of a row tile are reduced inside one
// block, the hole-hole unit, which keeps its W rows in LDS and contracts the
// hole-hole ladder from there (source/CCD.f90:651-656):
//   hh_S[p,a] = sum_q W_S[p,q] S[q,a],  hh_P[p,a] = sum_q W_P[p,q] P[q,a].

// The part of block K's virtual list from entry base on, in LDS for the
// tile loops: a list of up to kClsChunk entries is staged once with the
// ladder tables (stage_ladder); a longer one part by part here, behind block
// barriers (every thread of the block calls this, base uniform).  Returns
// the entries of the part.
__device__ __forceinline__ int restage_cls(const SolveG& S, int K, int base, int nvv, const int4* cls) {
  if (nvv > kClsChunk) {
    const int voff = S.kb_voff[K], n = min(kClsChunk, nvv - base);
    int4* const part = const_cast<int4*>(cls);  // (the block's own LDS)
    __syncthreads();  // every wave is done with the part before
    for (int k = threadIdx.x; k < n; k += blockDim.x) part[k] = ld(S.kb_nva + voff + base + k);
    __syncthreads();
  }
  return min(nvv - base, kClsChunk);
}

// The sums of one 16x16 tile (column lane r16 synthesises its B column from
// the orbitals cx, cy and the tables vtx, vty), the reduction over the K
// block's virtual list split over the block's four waves (wave w takes steps
// w, w+4, ...) and combined in a fixed order.  cls is the block's virtual list
// (n_c, c) in LDS, so each group of steps costs one round of global loads
// (the amplitudes its list entries select).  Every thread of the block calls
// it; wave 0 returns true with the tiles' sums in accs, accp.  NC column
// tiles (columns colv[c] apart, each synthesised from its orbitals cx[c],
// cy[c]) share each k-step's amplitude loads, every tile with the same MFMA
// sequence as a tile of its own (the same bits for every NC).
template <typename Real, int G, int NC>
__device__ __forceinline__ bool ladder_sums(const SolveG& S, int K, int row0, const int (&cx)[NC],
                                            const int (&cy)[NC], const bool (&colv)[NC], const double* vtx,
                                            const double* vty, const int4* cls, double* red,
                                            typename MfmaTraits<Real>::acc_t (&accs)[NC],
                                            typename MfmaTraits<Real>::acc_t (&accp)[NC]) {
  const int poff = S.kb_poff[K], noo = S.kb_poff[K + 1] - poff;
  const int voff = S.kb_voff[K], nvv = S.kb_voff[K + 1] - voff;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  // A operand from the K-blocked amplitudes: column ci of the block holds its
  // noo pairs contiguously, so a step's 16 rows are one contiguous run
  const auto TK = S.tk + S.kb_toff[K];
  const int rowA = row0 + r16;
  int4 nx[NC], ny[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    nx[c] = S.nv4[cx[c]];
    ny[c] = S.nv4[cy[c]];
    accs[c] = {0, 0, 0, 0};
    accp[c] = {0, 0, 0, 0};
  }
  // this wave's k-steps are c0 = 4*wave + 16*m; groups of G steps issue all
  // their amplitude loads (unconditional, in-range addresses, masked values)
  // before the MFMAs, which run for every step (masked steps add zeros).
  // G comes from the tuned tile table (uegccd_amd/tuned_tiles.json).  A list
  // longer than kClsChunk is staged a part at a time (parts a multiple of
  // every group long, so each wave takes the same steps in the same order,
  // and no group straddles two parts).
  for (int base = 0; base < nvv; base += kClsChunk) {
    // (indices from base: entries left, entries of this part, their amplitudes)
    const int nleft = nvv - base, end = restage_cls(S, K, base, nvv, cls);
    const auto TKp = TK + (size_t)base * noo;
    for (int g0 = 4 * wave; g0 < end; g0 += 16 * G) {
      double2 t[G];
#pragma unroll
      for (int st = 0; st < G; ++st) {
        const int ci = g0 + 16 * st + kq;
        const bool ok = ci < nleft && rowA < noo;
        const double2 v = ld(TKp + (ok ? ci * noo + rowA : 0));
        t[st] = ok ? v : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int st = 0; st < G; ++st) {
        const int ci = g0 + 16 * st + kq;
        const int4 nc = cls[min(ci, end - 1)];
        const Real as = (Real)(t[st].x + t[st].y), ap = (Real)(t[st].x - t[st].y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          Real bs = 0, bp = 0;
          if (ci < nleft && colv[c]) {
            const Real vx = (Real)vtx[d2(nc, nx[c])];
            const Real vy = (Real)vty[d2(nc, ny[c])];
            bs = vx - vy;
            bp = vx + vy;
          }
          accs[c] = MfmaTraits<Real>::mma(as, bs, accs[c]);
          accp[c] = MfmaTraits<Real>::mma(ap, bp, accp[c]);
        }
      }
    }
  }
  // each tile's four wave sums, combined in wave order (red holds 3 waves'
  // partials of one tile at a time)
  bool lead = true;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (wave > 0) {
      double* mine = red + ((wave - 1) * 64 + lane) * 8;
      for (int r = 0; r < 4; ++r) {
        mine[r] = (double)accs[c][r];
        mine[4 + r] = (double)accp[c][r];
      }
    }
    __syncthreads();
    if (wave == 0) {
      for (int w = 1; w < 4; ++w) {
        const double* theirs = red + ((w - 1) * 64 + lane) * 8;
        for (int r = 0; r < 4; ++r) {
          accs[c][r] = accs[c][r] + (Real)theirs[r];
          accp[c][r] = accp[c][r] + (Real)theirs[4 + r];
        }
      }
    } else {
      lead = false;
    }
    if (c + 1 < NC) __syncthreads();  // red is reused by the next tile
  }
  return lead;
}

// NC 16x16 particle-particle tiles (rows row0.., columns col0 + 16c.. of
// the a list) into lad.  vt holds the solve's distinct ladder tables
// (n2max + 1 entries each); lad_tab[] picks the table of each role: these
// columns read the linear direct (v(c-a)) and exchange (v(c-b)) tables.
template <typename Real, int G, int NC>
__device__ __forceinline__ void ladder_tile(const SolveG& S, int K, int row0, int col0, const double* vt,
                            const int4* cls, double* red) {
  typedef typename MfmaTraits<Real>::acc_t acc_t;
  const int o = S.o, o2 = o * o;
  const int poff = S.kb_poff[K], noo = S.kb_poff[K + 1] - poff;
  const int voff = S.kb_voff[K], nvv = S.kb_voff[K + 1] - voff;
  const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  int cx[NC], cy[NC];
  bool colv[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int col = col0 + 16 * c + r16;
    colv[c] = col < nvv;
    cx[c] = colv[c] ? S.kb_va[voff + col] : 0;
    cy[c] = colv[c] ? S.kb_vb[voff + col] : 0;
  }
  // the output rows' pairs, requested now (used after the reduction)
  int xpo[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xpo[r] = S.kb_pairs[poff + min(row0 + MfmaTraits<Real>::row(kq, r), noo - 1)];
  const int nt = S.n2max + 1;
  acc_t accs[NC], accp[NC];
  if (!ladder_sums<Real, G, NC>(S, K, row0, cx, cy, colv, vt + nt * S.lad_tab[0], vt + nt * S.lad_tab[1], cls,
                                red, accs, accp))
    return;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (!colv[c]) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rowo = row0 + MfmaTraits<Real>::row(kq, r);
      if (rowo >= noo) continue;
      st2(S.lad + (size_t)xpo[r] + (size_t)o2 * (cx[c] - o),
          make_double2(0.5 * (double)accs[c][r], 0.5 * (double)accp[c][r]));
    }
  }
}

// The W rows row0..row0+15 of block K over NC tiles of 16 pair columns at
// once (columns cbase + 16c + r16, c < NC): one pass over the block's virtual list,
// each group of amplitude loads shared by the NC tiles, every tile with the
// same MFMA sequence as a tile of its own; then each tile's four wave sums
// are combined in a fixed order and W = v(i-k) -+ v(i-l) + 1/2 sums goes to
// Wl (16 x noo, (W_S, W_P); zeros in rows past the block).  The W columns
// read the quadratic tables v(c-k), v(c-l), and v(i-k), v(i-l) the linear
// ones.  Every thread of the block calls it.
template <typename Real, int G, int NC>
__device__ __forceinline__ void w_rows(const SolveG& S, int K, int row0, int cbase, int qend, const double* vt,
                                       const int4* cls, double* red, double2* Wl) {
  typedef typename MfmaTraits<Real>::acc_t acc_t;
  const int o = S.o;
  const int poff = S.kb_poff[K], noo = S.kb_poff[K + 1] - poff;
  const int voff = S.kb_voff[K], nvv = S.kb_voff[K + 1] - voff;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const int nt = S.n2max + 1;
  const double* vtx = vt + nt * S.lad_tab[2];
  const double* vty = vt + nt * S.lad_tab[3];
  const auto TK = S.tk + S.kb_toff[K];
  const int rowA = row0 + r16;
  int4 nx[NC], ny[NC];
  bool colv[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int q = cbase + 16 * c + r16;
    colv[c] = q < qend;
    const int xq = colv[c] ? S.kb_pairs[poff + q] : 0;
    nx[c] = S.nv4[xq % o];
    ny[c] = S.nv4[xq / o];
  }
  acc_t accs[NC], accp[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    accs[c] = {0, 0, 0, 0};
    accp[c] = {0, 0, 0, 0};
  }
  for (int base = 0; base < nvv; base += kClsChunk) {  // (list parts as in ladder_sums)
    // (indices from base: entries left, entries of this part, their amplitudes)
    const int nleft = nvv - base, end = restage_cls(S, K, base, nvv, cls);
    const auto TKp = TK + (size_t)base * noo;
    for (int g0 = 4 * wave; g0 < end; g0 += 16 * G) {
      double2 t[G];
#pragma unroll
      for (int st = 0; st < G; ++st) {
        const int ci = g0 + 16 * st + kq;
        const bool ok = ci < nleft && rowA < noo;
        const double2 v = ld(TKp + (ok ? ci * noo + rowA : 0));
        t[st] = ok ? v : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int st = 0; st < G; ++st) {
        const int ci = g0 + 16 * st + kq;
        const int4 nc = cls[min(ci, end - 1)];
        const Real as = (Real)(t[st].x + t[st].y), ap = (Real)(t[st].x - t[st].y);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          Real bs = 0, bp = 0;
          if (ci < nleft && colv[c]) {
            const Real vx = (Real)vtx[d2(nc, nx[c])];
            const Real vy = (Real)vty[d2(nc, ny[c])];
            bs = vx - vy;
            bp = vx + vy;
          }
          accs[c] = MfmaTraits<Real>::mma(as, bs, accs[c]);
          accp[c] = MfmaTraits<Real>::mma(ap, bp, accp[c]);
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (cbase + 16 * c >= qend) break;
    if (wave > 0) {
      double* mine = red + ((wave - 1) * 64 + lane) * 8;
      for (int r = 0; r < 4; ++r) {
        mine[r] = (double)accs[c][r];
        mine[4 + r] = (double)accp[c][r];
      }
    }
    __syncthreads();
    if (wave == 0) {
      acc_t ss = accs[c], pp = accp[c];
      for (int w = 1; w < 4; ++w) {
        const double* theirs = red + ((w - 1) * 64 + lane) * 8;
        for (int r = 0; r < 4; ++r) {
          ss[r] = ss[r] + (Real)theirs[r];
          pp[r] = pp[r] + (Real)theirs[4 + r];
        }
      }
      if (colv[c]) {
        const int q = cbase + 16 * c + r16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int lr = MfmaTraits<Real>::row(kq, r), rowo = row0 + lr;
          double2 wv = make_double2(0.0, 0.0);
          if (rowo < noo) {
            const int4 ni = S.nv4[S.kb_pairs[poff + rowo] % o];
            const double vik = vt[nt * S.lad_tab[0] + d2(ni, nx[c])], vil = vt[nt * S.lad_tab[1] + d2(ni, ny[c])];
            wv = make_double2(vik - vil + 0.5 * (double)ss[r], vik + vil + 0.5 * (double)pp[r]);
          }
          Wl[lr * noo + q] = wv;
        }
      }
    }
    __syncthreads();  // red is reused by the next tile; Wl is complete after the last
  }
}

// The hole-hole unit of rows row0..row0+15 of block K over part `part` of its
// W columns (0: all of them; 1, 2: the first and second half of their 16-pair
// tiles): the W rows over those pair columns into Wl (w_rows, two column
// tiles per pass over the virtual list), then their part of hh = W . (S, P)
// as 16x16 FP64 MFMA tiles over the block's a columns, dealt over the waves,
// A from Wl and B from the K-blocked amplitudes (a column's pairs
// contiguous), into hh (parts 0 and 1) or hh2 (part 2) at the amplitudes'
// K-blocked positions; k_main adds hh2 in launches of halved units.  HS:
// k-steps of 4 pairs per round of loads (from the tuned tile table).
constexpr int kWGroup = 8;
template <typename Real, int G, int HS>
__device__ __forceinline__ void hh_unit(const SolveG& S, int K, int row0, int part, const double* vt, const int4* cls,
                                        double* red, double2* Wl) {
  typedef MfmaTraits<double>::acc_t acc64_t;
  const int poff = S.kb_poff[K], noo = S.kb_poff[K + 1] - poff;
  const int nvv = S.kb_voff[K + 1] - S.kb_voff[K];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  // the unit's W column tiles [tlo, thi) and pair columns [qlo, qhi)
  const int ntq = (noo + 15) >> 4, hsplit = (ntq + 1) >> 1;
  const int tlo = part == 2 ? hsplit : 0, thi = part == 0 ? ntq : (part == 1 ? hsplit : ntq);
  const int qlo = 16 * tlo, qhi = min(noo, 16 * thi);
  // (the W pass is the unit's longest chain of loads: kWGroup steps per round)
  if (thi - tlo == 1)
    w_rows<Real, kWGroup, 1>(S, K, row0, qlo, qhi, vt, cls, red, Wl);
  else
    for (int cb = qlo; cb < qhi; cb += 32) w_rows<Real, kWGroup, 2>(S, K, row0, cb, qhi, vt, cls, red, Wl);
  const int rowA = row0 + r16;
  const bool rv = rowA < noo;
  const double2* Wrow = Wl + r16 * noo;
  const auto TKb = S.tk + S.kb_toff[K];
  const auto HH = (part == 2 ? S.hh2 : S.hh) + S.kb_toff[K];
  for (int c0 = 16 * wave; c0 < nvv; c0 += 64) {
    const int colB = c0 + r16;
    const bool cv = colB < nvv;
    const auto TK = TKb + (size_t)(cv ? colB : 0) * noo;
    acc64_t accs = {0, 0, 0, 0}, accp = {0, 0, 0, 0};
    for (int g0 = qlo; g0 < qhi; g0 += 4 * HS) {
      double ws[HS], wp[HS];
#pragma unroll
      for (int st = 0; st < HS; ++st) {
        const int q = g0 + 4 * st + kq;
        const double2 w = Wrow[min(q, qhi - 1)];
        const bool ok = q < qhi && rv;
        ws[st] = ok ? w.x : 0.0;
        wp[st] = ok ? w.y : 0.0;
      }
      double bs[HS], bp[HS];
#pragma unroll
      for (int st = 0; st < HS; ++st) {
        const int q = g0 + 4 * st + kq;
        const double2 t = ld(TK + min(q, qhi - 1));
        const bool ok = q < qhi && cv;
        bs[st] = ok ? t.x + t.y : 0.0;
        bp[st] = ok ? t.x - t.y : 0.0;
      }
      // (all steps, the masked ones add zeros: a branch here would let the
      // compiler sink each step's loads behind the previous step's MFMAs)
#pragma unroll
      for (int st = 0; st < HS; ++st) {
        accs = MfmaTraits<double>::mma(ws[st], bs[st], accs);
        accp = MfmaTraits<double>::mma(wp[st], bp[st], accp);
      }
    }
    if (!cv) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int p = row0 + MfmaTraits<double>::row(kq, r);
      if (p < noo) st2(HH + (size_t)colB * noo + p, make_double2(0.5 * accs[r], 0.5 * accp[r]));
    }
  }
}

// Dynamic LDS of a ladder-tile or hole-hole block: the solve's distinct
// ladder tables (n2max + 1 doubles each), then (16-byte aligned) a K block's
// virtual list, then the hole-hole unit's W rows.
struct LadderLds {
  double* vt;
  int4* cls;
  double2* wl;
};
__device__ __forceinline__ LadderLds ladder_lds(double* dyn, int tab_doubles, int cls_max) {
  LadderLds l;
  l.vt = dyn;
  l.cls = reinterpret_cast<int4*>(dyn + ((tab_doubles + 1) & ~1));
  l.wl = reinterpret_cast<double2*>(l.cls + cls_max);
  return l;
}

// Stages the solve's ladder tables and block K's virtual list in LDS; returns
// the solve's state fields the block needs (phase, precision).
__device__ __forceinline__ int2 stage_ladder(const SolveG& S, int K, const LadderLds& l) {
  const int phase = S.st2->phase, prec64 = S.st2->prec64;  // requested with the tables, tested after them
  const int voff = S.kb_voff[K], nvv = S.kb_voff[K + 1] - voff;
  const int nt = S.n2max + 1;
  for (int m = 0; m < S.lad_ntab; ++m) {
    const auto src = S.vtab + (size_t)S.lad_slot[m] * nt;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) l.vt[m * nt + k] = src[k];
  }
  if (nvv <= kClsChunk)  // (a longer list is staged part by part by the tile loops)
    for (int k = threadIdx.x; k < nvv; k += blockDim.x) l.cls[k] = ld(S.kb_nva + voff + k);
  __syncthreads();
  return make_int2(phase, prec64);
}

// A particle-particle block: ladder item `item`, NC column tiles (the solve's
// 16-column items for NC = 1, its 32-column items for NC = 2).
template <int G, int NC>
__device__ __forceinline__ void ladder_block(const SolveDev* __restrict__ solves, const int* __restrict__ prefix,
                             int n, int item, const LadderLds& l, double* red) {
  static_assert(NC == 1 || NC == 2, "one or two column tiles");
  const int s = find_solve(prefix, n, item);
  const SolveG& S = as_global(solves)[s];
  const int4 it = ld((NC == 2 ? S.lad_items2 : S.lad_items) + (item - prefix[s]));
  const int2 st = stage_ladder(S, it.y, l);
  if (pass_idle(S, st.x)) return;
  if (st.y)
    ladder_tile<double, G, NC>(S, it.y, it.z, it.w, l.vt, l.cls, red);
  else
    ladder_tile<float, G, NC>(S, it.y, it.z, it.w, l.vt, l.cls, red);
}

template <int G, int HS>
__device__ __forceinline__ void hh_block(const SolveDev* __restrict__ solves, const int* __restrict__ prefix,
                         int n, int blk, int split, const LadderLds& l, double* red) {
  const int s = find_solve(prefix, n, blk);
  const SolveG& S = as_global(solves)[s];
  const int4 it = ld((split ? S.hh_items2 : S.hh_items) + (blk - prefix[s]));
  const int2 st = stage_ladder(S, it.y, l);
  if (pass_idle(S, st.x)) return;
  if (st.y)
    hh_unit<double, G, HS>(S, it.y, it.z, it.w, l.vt, l.cls, red, l.wl);
  else
    hh_unit<float, G, HS>(S, it.y, it.z, it.w, l.vt, l.cls, red, l.wl);
}

template <int BT, bool SPLIT>
__global__ __launch_bounds__(BT) void k_slice(const SolveDev* __restrict__ solves,
                                              const int* __restrict__ prefix, int n, int fin_prev, FoldArgs fa) {
  __shared__ Lds L;
  __shared__ SliceDiis D;
  __shared__ double vt2[SPLIT ? kMaxN2 : 1];
  __shared__ double fvals[kPartialStride];
  __shared__ double fred[(BT / 64) * 32];
  __shared__ State fsst;
  extern __shared__ double dyn[];
  slice_block<BT, SPLIT>(solves, prefix, n, blockIdx.x, L, D, dyn, vt2, fin_prev, fa, fvals, fred, &fsst);
}

// --------------------------------------------------------------------------
// Contractions that follow the slice and ladder passes (k_contract).
//
// Quadratic rings and crossed rings per momentum transfer q.  With
// P(q) = {(i,p): n_p - n_i = q}, a valid element (i,j,a) with partner b is
//   ring:         row (i,a) of P(q),  q  = n_a - n_i, column (j,b) of P(-q)
//   crossed ring: row (i,b) of P(q'), q' = n_b - n_i, column (j,a) of P(-q')
// and every (row, column) pair of P(q) x P(-q) is a valid element.  So
//   ring(i,j,a)  = sum_l J(a,i,l) . T(l,j,d_q(l)),  n_d = n_l + q  (source/CCD.f90:801-808)
//   xring(i,j,a) = sum_l Jx(b,i,l) . T(l,j,a)                      (source/CCD.f90:927-934)
// are, per q, one product [rows x o] . [o x columns] each (A from the J rows,
// B gathered from T), each output written once (ring to quad.xy, crossed ring
// to quad.zw of its element).
//
// Rings and crossed rings are separate work items (one wave each): half the
// registers per wave, so more waves are resident.  A tile's loads go out in
// two rounds with no branch between them (masked values, clamped in-range
// addresses): the J rows with every index (the column's pair, the four output
// rows' pairs, the partner slices d_q(l)), then the T columns they select.
// QS: k-steps of 4 l values per round of loads (from the tuned tile table).
template <bool XRING, int QS>
__device__ __forceinline__ void quad_tile(const SolveG& S, int qb, int tile) {
  typedef MfmaTraits<double>::acc_t acc_t;
  const int4 info = S.qb_info[qb];  // rows of q: [r0, r0+n); columns (rows of -q): [c0, c0+m)
  const int r0 = info.x, n = info.y, c0 = info.z, m = info.w;
  const int ntc = (m + 15) >> 4;
  if (m == 0 || tile >= ((n + 15) >> 4) * ntc) return;
  const int o = S.o, o2 = o * o;
  const int row0 = (tile / ntc) * 16, col0 = (tile % ntc) * 16;
  const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  // lane supplies A[row r16][l] (J rows, contiguous per q) and B[l][column r16]
  const int rowA = row0 + r16, colB = col0 + r16;
  const bool rv = rowA < n, cv = colB < m;
  const auto J = (XRING ? S.jxring : S.jring) + (size_t)(r0 + (rv ? rowA : 0)) * o;
  const auto dq = S.qb_d + (size_t)qb * o;
  const auto T = S.tcur;
  acc_t cab = {0, 0, 0, 0}, cba = {0, 0, 0, 0};
  int prow[4], j = 0, pcv = 0;
  for (int g0 = 0; g0 < o; g0 += 4 * QS) {
    double a2[QS], a3[QS];
    int dv[QS];
#pragma unroll
    for (int st = 0; st < QS; ++st) {
      const int l = g0 + 4 * st + kq;
      const double2 a = ld(J + min(l, o - 1));
      const bool ok = l < o && rv;
      a2[st] = ok ? a.x : 0.0;
      a3[st] = ok ? a.y : 0.0;
      if (!XRING) dv[st] = dq[min(l, o - 1)];
    }
    if (g0 == 0) {
      const int pc = S.qb_rows[c0 + (cv ? colB : 0)];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        prow[r] = S.qb_rows[r0 + min(row0 + MfmaTraits<double>::row(kq, r), n - 1)];
      j = pc % o;
      pcv = pc / o;
    }
    const size_t tcol = (size_t)o * j;
    double2 b[QS];
#pragma unroll
    for (int st = 0; st < QS; ++st) {
      const int l = g0 + 4 * st + kq;
      const int slice = XRING ? pcv : max(dv[st], 0);
      const bool ok = l < o && cv && (XRING || dv[st] >= 0);
      const double2 t = ld(T + min(l, o - 1) + tcol + (size_t)o2 * slice);
      b[st] = ok ? t : make_double2(0.0, 0.0);
    }
    // (all steps, the masked ones add zeros: a branch here would let the
    // compiler sink each step's loads behind the previous step's MFMAs)
#pragma unroll
    for (int st = 0; st < QS; ++st) {
      if (!XRING) {
        // ring:  ab += J2 (T^ab + T^ba) + J1 T^ab = J2 (2 T^ab + T^ba) + J3 T^ab,
        //        ba += J3 T^ba
        cab = MfmaTraits<double>::mma(a2[st], 2.0 * b[st].x + b[st].y, cab);
        cab = MfmaTraits<double>::mma(a3[st], b[st].x, cab);
        cba = MfmaTraits<double>::mma(a3[st], b[st].y, cba);
      } else {
        // crossed ring:  ba -= J1 T^ba + J2 (T^ab + T^ba) = J2 (T^ab + 2 T^ba) - J3 T^ba,
        //                ab += J3 T^ab
        cba = MfmaTraits<double>::mma(-a2[st], b[st].x + 2.0 * b[st].y, cba);
        cba = MfmaTraits<double>::mma(a3[st], b[st].y, cba);
        cab = MfmaTraits<double>::mma(a3[st], b[st].x, cab);
      }
    }
  }
  if (!cv) return;
  const size_t tcol = (size_t)o * j;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + MfmaTraits<double>::row(kq, r);
    if (row >= n) continue;
    // ring sums go to element (i, j, a) with a the row's particle (quad.xy),
    // crossed-ring sums to (i, j, a') with a' the column's particle (quad.zw)
    const int pr = prow[r];
    const auto q = S.quad + (size_t)(pr % o) + tcol + (size_t)o2 * (XRING ? pcv : pr / o);
    st2(reinterpret_cast<GlobalPtr<double2>>(q) + (XRING ? 1 : 0), make_double2(cab[r], cba[r]));
  }
}

// Mosaic Joo(k) = sum over slices of the slice blocks' partials
// (source/CCD.f90:563-567), split over kJooParts waves: wave `part` sums the
// slices of its range in order (lane k, column k) into joo[part][k]; k_main
// adds the parts in order.
__device__ __forceinline__ void joo_wave(const SolveG& S, int part) {
  if (!S.do_mosaic) return;
  const int o = S.o, lane = threadIdx.x & 63;
  const int s0 = (S.v * part) / kJooParts, s1 = (S.v * (part + 1)) / kJooParts;
  for (int k = lane; k < o; k += 64) {
    double a = 0.0;
    for (int r = s0; r < s1; ++r) a += S.joo_part[r * o + k];
    S.joo[part * o + k] = a;
  }
}

// One ring / crossed-ring / Joo work item, by one wave.
template <int QS>
__device__ __forceinline__ void contract_item(const SolveDev* __restrict__ solves, const int* __restrict__ prefix, int n,
                              int item) {
  const int s = find_solve(prefix, n, item);
  const SolveG& S = as_global(solves)[s];
  const int4 it = ld(S.con_items + (item - prefix[s]));
  // (the phase is requested beside the item, so the test adds no round trip)
  if (pass_idle(S, S.st2->phase)) return;
  if (it.y == CK_RING)
    quad_tile<false, QS>(S, it.z, it.w);
  else if (it.y == CK_XRING)
    quad_tile<true, QS>(S, it.z, it.w);
  else
    joo_wave(S, it.z);
}

// The contraction pass, one launch of three kinds of 256-thread blocks:
//   [0, nhh_pad)             hole-hole units (the longest blocks, so they go
//                            first and, not renumbered, round-robin over the XCDs)
//   [nhh_pad, + nlad_pad)    particle-particle ladder blocks, NC column tiles
//                            each (nlad blocks over the ladder items)
//   the rest                 four ring / crossed-ring / Joo items per block
// The pads are multiples of 8, so a block's XCD is still blockIdx mod 8 inside
// each range, which the ladder tiles and items renumber XCD-contiguously.
// (held to 128 VGPRs, four waves per SIMD: its tiles wait on gathers)
template <int G, int QS, int HS, int NC>
__global__ __launch_bounds__(256, 4) void k_contract(const SolveDev* __restrict__ solves,
                                                  const int* __restrict__ hh_prefix, int nhh, int nhh_pad,
                                                  const int* __restrict__ lad_prefix, int nlad, int nlad_pad,
                                                  const int* __restrict__ con_prefix, int ncon, int n,
                                                  int tab_doubles, int cls_max, int hh_split) {
  extern __shared__ double dyn[];  // ladder tables, a virtual list, W rows
  __shared__ double red[3 * 64 * 8];
  const int bid = blockIdx.x;
  if (bid < nhh_pad) {
    if (bid < nhh) hh_block<G, HS>(solves, hh_prefix, n, bid, hh_split, ladder_lds(dyn, tab_doubles, cls_max), red);
    return;
  }
  int rb = bid - nhh_pad;
  if (rb < nlad_pad) {
    const int t = xcd_block(rb, nlad_pad);
    if (t < nlad) ladder_block<G, NC>(solves, lad_prefix, n, t, ladder_lds(dyn, tab_doubles, cls_max), red);
    return;
  }
  rb -= nlad_pad;
  // the wave index is uniform over the wave: saying so lets the item, the
  // descriptor and its pointers live in scalar registers (scalar loads)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int item = xcd_block(rb, gridDim.x - nhh_pad - nlad_pad) * kItemsPerBlock + wave;
  if (item < ncon) contract_item<QS>(solves, con_prefix, n, item);
}

// Block partials of the residual pass: max dT, max |R|, sum E, the non-finite
// count and the nd DIIS dots e_new . e_k (the new error against every history
// slot; HE is the element's entry of slot 0, slots stride apart), padded with
// zeros to NQ, written slot-major for k_finalize.
template <int NQ>
__device__ __forceinline__ void main_partials(const SolveG& S, int lb, int nd, int slot, bool lead,
                                              double dt_part, double r_part, double e_part, double bad,
                                              double eab, double eba, const double2* HE, int stride,
                                              double* red) {
  double v[NQ];
  v[0] = dt_part;
  v[1] = r_part;
  v[2] = e_part;
  v[3] = bad;
#pragma unroll
  for (int k = 0; k < NQ - 4; ++k) {
    double dk = 0.0;
    if (k < nd && lead) {
      if (k == slot) {
        dk = eab * eab + eba * eba;
      } else {
        const double2 h = HE[(size_t)k * stride];
        dk = eab * h.x + eba * h.y;
      }
    }
    v[4 + k] = dk;
  }
  const double r = block_reduce_scatter<NQ, 2>(v, red);
  const int q = threadIdx.x;
  if (q < 4 + nd) *partial(S, q == 0 ? 1 : q == 1 ? 2 : q == 2 ? 0 : q, lb) = r;
}

// --------------------------------------------------------------------------
// Residual assembly + update, one thread per dense (i,j,a): G from drivers,
// mosaics, ladders (pp + hh from the ladder and contraction passes), rings and
// crossed rings (slice products and the quadratic sums), then R = D*T - G,
// the damped update, dT, the energy and the DIIS dot products.
// One 256-thread element block gb of the launch (block index after the XCD
// renumbering); red and joo_l are the block's LDS.
// NQ: the partials' width (16 serves up to 12 DIIS entries, 32 up to 28),
// or 0 to choose it per solve.
template <int NQ>
__device__ __forceinline__ void main_block(const SolveDev* __restrict__ solves, const int* __restrict__ prefix, int n,
                           int gb, int flags, double* red, double* joo_l) {
  const int s = find_solve(prefix, n, gb);
  const SolveG& S = as_global(solves)[s];
  // the state, the element (one valid element per thread, by compact index
  // cp, which is also its history position) and the Joo parts are requested
  // together; a finished solve leaves after they arrive
  const int phase = S.st2->phase, slot = S.st2->slot;
  const int lb = gb - prefix[s];
  if (lb == 0) {  // the pass's state becomes the state between passes
    const auto from = reinterpret_cast<GlobalPtr<const double>>(S.st2);
    const auto to = reinterpret_cast<GlobalPtr<double>>(S.st);
    for (int w = threadIdx.x; w < kStateWords; w += blockDim.x) to[w] = from[w];
  }
  const int o = S.o, o2 = o * o, n3 = S.n3;
  const int cp = lb * kElemPerBlock + threadIdx.x;
  int x = 0, b = -1, pos = 0;
  if (cp < S.nvalid) {
    const uint2 cm = ld(S.cmap + cp);
    x = elem_x(S.ecode, cm);
    b = elem_partner(S.ecode, cm);
    pos = elem_pos(S.ecode, cm);
  }
  if (S.do_mosaic)  // Joo from its slice-range parts, in part order
    for (int k = threadIdx.x; k < o; k += blockDim.x) {
      double a = 0.0;
#pragma unroll
      for (int p = 0; p < kJooParts; ++p) a += S.joo[p * o + k];
      joo_l[k] = a;
    }
  __syncthreads();  // publishes joo_l
  if (pass_idle(S, phase)) return;
  const int nh = (flags & MF_ITERATE) ? S.ndiis : 0;
  const auto T = S.tcur;
  const bool valid = b >= 0;
  const int i = x % o, j = (x / o) % o, av = x / o2, a = o + av, bv = b - o;

  double e_part = 0.0, dt_part = 0.0, r_part = 0.0, bad = 0.0;
  double eab = 0.0, eba = 0.0;
  const bool lead = valid;
  if (lead) {
    const double2 t0 = T[x];
    const double t_ab = t0.x, t_ba = t0.y;
    // the four Coulomb elements of the element, v(n_p - n_q) for p occupied
    // and q virtual, from the per-solve virtual-occupied block; with split
    // range flags each role reads its own table (drivers direct/exchange,
    // ring and crossed-ring linear directs, energy)
    const size_t vo = (size_t)S.v * o;
    const auto VE = S.vso + S.rslot[RO_ENERGY] * vo;
    const double v_ia = VE[av * o + i], v_ib = VE[bv * o + i];
    const double v_ja = VE[av * o + j], v_jb = VE[bv * o + j];
    double d_ia = v_ia, d_ib = v_ib, r_ia = v_ia, r_jb = v_jb, x_ja = v_ja, x_ib = v_ib;
    if (S.main_split) {
      const auto VD = S.vso + S.rslot[RO_DRVDIR] * vo, VX = S.vso + S.rslot[RO_DRVEX] * vo;
      const auto VR = S.vso + S.rslot[RO_RING_LD] * vo, VC = S.vso + S.rslot[RO_XR_LD] * vo;
      d_ia = VD[av * o + i];
      d_ib = VX[bv * o + i];
      r_ia = VR[av * o + i];
      r_jb = VR[bv * o + j];
      x_ja = VC[av * o + j];
      x_ib = VC[bv * o + i];
    }
    // drivers (source/CCD.f90:484-501)
    double gab = d_ia;
    double gba = -d_ib;
    // mosaics (source/CCD.f90:571-582)
    if (S.do_mosaic) {
      const auto jvv = S.jvv;
      const double* joo = joo_l;
      gab = gab - 0.5 * (jvv[av] * t_ab + jvv[bv] * t_ab);
      gba = gba - 0.5 * (jvv[av] * t_ba + jvv[bv] * t_ba);
      gab = gab - 0.5 * (joo[i] * t_ab + joo[j] * t_ab);
      gba = gba - 0.5 * (joo[i] * t_ba + joo[j] * t_ba);
    }
    // ladders: particle-particle part with the hole-hole sum (k_contract)
    if (S.do_ladder) {
      const double2 lp = S.lad[x], lh = S.hh[pos];
      const double2 lh2 = (flags & MF_HH2) ? S.hh2[pos] : make_double2(0.0, 0.0);
      const double2 l2 = make_double2(lp.x + lh.x + lh2.x, lp.y + lh.y + lh2.y);
      gab += 0.5 * (l2.x + l2.y);
      gba += 0.5 * (l2.x - l2.y);
    }
    // slice products: element (i,j) of slice a, (j,i) of slice b
    if (S.do_ring || S.do_xring) {
      // linear rings (source/CCD.f90:716-763): v(q) rowsum - (T V), and
      // crossed rings (source/CCD.f90:845-892): v(q) colsum - (V T), of the
      // a-slice at (i,j) and the b-slice at (j,i)
      const double2 ua = S.tv[x];
      const double2 ub = S.tv[(size_t)j + o * i + (size_t)o2 * bv];
      gab -= ua.x + ub.x;
      gba -= ua.y + ub.y;
      if (S.do_ring) gab += r_ia * S.rsum[av * o + i] + r_jb * S.rsum[bv * o + j];
      if (S.do_xring) gba += x_ja * S.csum[bv * o + i] + x_ib * S.csum[av * o + j];
      const double4 qd = S.quad[x];
      if (S.do_ring) { gab += qd.x; gba += qd.y; }
      if (S.do_xring) { gab += qd.z; gba += qd.w; }
    }
    // GetRes + SolveCC + CCEnergy on this element (source/CCD.f90:378-417, 225-227)
    const double den = S.eig[i] + S.eig[j] - S.eig[a] - S.eig[b];
    const double rab = -gab + den * t_ab;
    const double rba = -gba + den * t_ba;
    const double df = S.df;
    const double nab = (gab / den - (1.0 - df) * t_ab) / df;
    const double nba = (gba / den - (1.0 - df) * t_ba) / df;
    dt_part = fmax(fabs(nab - t_ab), fabs(nba - t_ba));
    r_part = fmax(fabs(rab), fabs(rba));
    const double naa = nab + nba;
    e_part = 0.5 * (v_ia * (naa + nab) - v_ja * (naa + nba));
    if (!isfinite(nab) || !isfinite(nba) || !isfinite(rab) || !isfinite(rba)) bad = 1.0;
    if (flags & MF_WRITE_GR) {
      if (S.gout) { S.gout[x] = gab; S.gout[n3 + x] = gba; }
      if (S.rout) { S.rout[x] = rab; S.rout[n3 + x] = rba; }
    }
    if (flags & MF_ITERATE) {
      S.histT[(size_t)slot * S.nvalid + cp] = make_double2(nab, nba);
      if (S.diis_err_is_dt) { eab = nab - t_ab; eba = nba - t_ba; }
      else { eab = rab; eba = rba; }
      if (S.mode == UEGCCD_MODE_DIIS) S.histE[(size_t)slot * S.nvalid + cp] = make_double2(eab, eba);
    }
  }
  // block partials: max dT, max |R|, sum E, non-finite count, DIIS dots
  const int nd = (S.mode == UEGCCD_MODE_DIIS) ? nh : 0;
  const double2* HE = S.histE + (lead ? cp : 0);
  if (NQ == 16 || (NQ == 0 && nd <= 12))
    main_partials<16>(S, lb, nd, slot, lead, dt_part, r_part, e_part, bad, eab, eba, HE, S.nvalid, red);
  else
    main_partials<32>(S, lb, nd, slot, lead, dt_part, r_part, e_part, bad, eab, eba, HE, S.nvalid, red);
}

// (NQ = 16, the usual DIIS length, is held to 128 VGPRs: four waves per SIMD
// to hide the element gathers)
template <int NQ>
__global__ __launch_bounds__(kElemThreads, NQ == 16 ? 4 : 1) void k_main(const SolveDev* __restrict__ solves,
                                                                        const int* __restrict__ prefix, int n,
                                                                        int flags) {
  __shared__ double red[(kElemThreads / 64) * 32];
  __shared__ double joo_l[kMaxOcc];
  main_block<NQ>(solves, prefix, n, xcd_block(blockIdx.x, gridDim.x), flags, red, joo_l);
}

// Sums and maxima of one solve's element-block partials by one block, in a
// fixed order.  Slot-major partials: quantity q of every element block is
// contiguous; q = 1 (max dT) and q = 2 (max |R|) are maxima, E (0), the
// non-finite count (3) and the nd DIIS dots (4..) sums.  Each of the first
// kFinThreads threads takes blocks t, t + kFinThreads, ... (further threads
// add zeros, so a block of any size of at least kFinThreads sums in the same
// order) and loads every quantity of a block at once, so the loads of a
// round are in flight together (the partials were written on other XCDs and
// miss this L2).  vals[0..3+nd] (E, max dT, max |R|, non-finite count, dots)
// are written by threads 0..3+nd: valid after a barrier.  red holds
// (blockDim/64)*32 doubles.
template <int NQ>
__device__ __forceinline__ void partials_load(const SolveG& S, int nd, int k, PartialLoads<NQ>& L) {
  constexpr int kStep = NQ <= 16 ? 2 : 1;
  const int nq = 4 + nd, nb = S.nblk_elem, k2 = k + kFinThreads;
  const bool one = k < nb && threadIdx.x < kFinThreads, two = one && kStep == 2 && k2 < nb;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int src = q == 0 ? 1 : q == 1 ? 2 : q == 2 ? 0 : q;
    L.p[q] = src < nq && one ? *partial(S, src, k) : 0.0;
    L.p2[q] = src < nq && two ? *partial(S, src, k2) : 0.0;
  }
}
// The reduction of the partials from the first round of loads (L, requested
// by partials_load(S, nd, threadIdx.x, L)) on: blocks k, k + kFinThreads, ...
// of thread k in that order, then the fixed tree of block_reduce_scatter.
template <int NQ>
__device__ __forceinline__ void partials_finish(const SolveG& S, int nd, PartialLoads<NQ>& L, double* vals,
                                                double* red) {
  constexpr int kStep = NQ <= 16 ? 2 : 1;
  const int nq = 4 + nd, nb = S.nblk_elem;
  double acc[NQ];  // dT, max|R|, E, non-finite, dots
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  for (int k = threadIdx.x; k < nb && threadIdx.x < kFinThreads; k += kStep * kFinThreads) {
    if (k != (int)threadIdx.x) partials_load<NQ>(S, nd, k, L);
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = q < 2 ? fmax(acc[q], L.p[q]) : acc[q] + L.p[q];
    if (kStep == 2 && k + kFinThreads < nb) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) acc[q] = q < 2 ? fmax(acc[q], L.p2[q]) : acc[q] + L.p2[q];
    }
  }
  const double r = block_reduce_scatter<NQ, 2>(acc, red);
  const int q = threadIdx.x;
  if (q < nq) vals[q == 0 ? 1 : q == 1 ? 2 : q == 2 ? 0 : q] = r;
}
template <int NQ>
__device__ __forceinline__ void reduce_partials_n(const SolveG& S, int nd, double* vals, double* red) {
  PartialLoads<NQ> L;
  partials_load<NQ>(S, nd, threadIdx.x, L);
  partials_finish<NQ>(S, nd, L, vals, red);
}
// MAXQ: the widest partials the caller reduces (16: DIIS histories up to 12,
// the only width the slice blocks' folded update instantiates).
template <int MAXQ>
__device__ __forceinline__ void reduce_partials(const SolveG& S, int nd, double* vals, double* red) {
  if (MAXQ == 16 || nd <= 12)
    reduce_partials_n<16>(S, nd, vals, red);
  else
    reduce_partials_n<MAXQ == 16 ? 16 : 32>(S, nd, vals, red);
}

// The iteration bookkeeping of source/CCD.f90:80-99 after a residual pass, on
// the state copy in LDS, by one thread: NIter, E, dT, convergence, the mixed
// precision switch, the DIIS history and its B matrix row, the trace line.
// Returns whether the next amplitudes are to be extrapolated.
__device__ __forceinline__ bool advance_state(const SolveG& S, State* st, const double* vals, int phase0,
                                              bool write_trace) {
  const double e = vals[0], dt = vals[1], rmax = vals[2];
  if (vals[3] > 0.0) st->nonfinite = 1;
  if (phase0 == PH_FINAL) {
    // the final residual check (reference mode), or in DIIS mode, whose final
    // pass computes nothing, the residual of the converged pass
    st->maxres = S.mode == UEGCCD_MODE_DIIS ? st->maxres_in : rmax;
    st->phase = PH_DONE;
    return false;
  }
  st->niter += 1;
  st->ecorr_prev = st->ecorr;
  st->ecorr = e;
  st->dt = dt;
  st->maxres_in = rmax;
  if (write_trace && S.trace && st->niter <= S.maxiter) {
    S.trace[2 * (st->niter - 1)] = e;
    S.trace[2 * (st->niter - 1) + 1] = dt;
  }
  const int slot = st->slot;
  st->slot_cur = slot;
  st->slot = (slot + 1) % S.ndiis;
  bool switched = false;
  if (!st->prec64) {
    st->iters_fp32 += 1;
    if (rmax < S.mixed_switch) { st->prec64 = 1; switched = true; }
  }
  bool conv;
  if (S.mode == UEGCCD_MODE_REFERENCE)
    conv = dt < S.tol_dt;
  else
    conv = st->prec64 && !switched && fabs(e - st->ecorr_prev) < S.tol_e && rmax < S.tol_r;
  const bool stop = conv || st->niter >= S.maxiter;
  bool want = false;
  if (S.mode == UEGCCD_MODE_DIIS) {
    // (the history carries over the precision switch: the FP32 phase's error
    // vectors only shape the extrapolation, the FP64 residuals decide
    // convergence; restarting it there cost a plain step and a rebuilt
    // history, 5% more iterations over the 256-request stream)
    {
      st->nhist = st->nhist < S.ndiis ? st->nhist + 1 : S.ndiis;
      for (int k = 0; k < S.ndiis; ++k) {
        st->bmat[slot * kMaxDiis + k] = vals[4 + k];
        st->bmat[k * kMaxDiis + slot] = vals[4 + k];
      }
      want = !stop && st->niter >= S.diis_start && st->nhist >= 2 && !st->nonfinite;
    }
  }
  if (st->nonfinite) {
    st->phase = PH_DONE;
  } else {
    if (conv) st->converged = 1;
    if (stop) st->phase = PH_FINAL;
  }
  return want;
}

// Reductions of a pass and the state that follows:
// FIN_ITER after a residual pass (bookkeeping and DIIS solve), FIN_INIT after
// the first pass (E and dT of the starting amplitudes, source/CCD.f90:62-64),
// FIN_SUM after a plain energy evaluation.  The iteration state is worked on
// in LDS: one parallel copy in, one out, instead of a chain of dependent
// global accesses by a single thread.
// finalize_state leaves the new state of solve S in sst (LDS) for a block of
// kFinThreads threads or more (vals and red are its LDS); it returns the
// phase the state had (a FIN_ITER on a finished solve changes nothing).
// write_trace: this block writes the trace line (one block per solve does).
__device__ __forceinline__ void state_load(const SolveG& S, StateLoads& L) {
  const auto src = reinterpret_cast<GlobalPtr<const double>>(S.st);
  const int w0 = threadIdx.x, w1 = threadIdx.x + kFinThreads;
  const bool two = threadIdx.x < kFinThreads;
  L.sw0 = w0 < kStateWords ? src[w0] : 0.0;
  L.sw1 = two && w1 < kStateWords ? src[w1] : 0.0;
}
// The DIIS dots a FIN_ITER reduction reads.
__device__ __forceinline__ int finalize_dots(const SolveG& S, int kind) {
  return (kind == FIN_ITER && S.mode == UEGCCD_MODE_DIIS) ? S.ndiis : 0;
}
__device__ __forceinline__ int finalize_apply(const SolveG& S, int kind, int prec64, const double* vals,
                                              State& sst, bool write_trace, const StateLoads& SL);

template <int MAXQ>
__device__ __forceinline__ int finalize_state(const SolveG& S, int kind, int prec64, double* vals, double* red,
                                              State& sst, bool write_trace) {
  // One round of loads: each thread's state words go out with its partials
  // (the DIIS dots are read whatever the phase; a finishing pass ignores them).
  StateLoads SL;
  state_load(S, SL);
  reduce_partials<MAXQ>(S, finalize_dots(S, kind), vals, red);
  return finalize_apply(S, kind, prec64, vals, sst, write_trace, SL);
}

// The state update from the reduced partials vals and the state words SL.
__device__ __forceinline__ int finalize_apply(const SolveG& S, int kind, int prec64, const double* vals,
                                              State& sst, bool write_trace, const StateLoads& SL) {
  const int w0 = threadIdx.x, w1 = threadIdx.x + kFinThreads;
  const bool two = threadIdx.x < kFinThreads;
  double* dst = reinterpret_cast<double*>(&sst);
  if (w0 < kStateWords) dst[w0] = SL.sw0;
  if (two && w1 < kStateWords) dst[w1] = SL.sw1;
  __syncthreads();
  const int phase0 = sst.phase;
  if (kind == FIN_ITER && phase0 == PH_DONE) return phase0;  // the whole block
  if (threadIdx.x == 0) {
    State* st = &sst;
    const double e = vals[0], dt = vals[1];
    if (kind == FIN_ITER) {
      st->next_src = advance_state(S, st, vals, phase0, write_trace) ? SRC_DIIS : SRC_PLAIN;
    } else {
      if (kind == FIN_INIT) {  // a fresh state: nothing carries over from an earlier solve
        st->nonfinite = 0;
        st->prec64 = prec64 < 0 ? S.prec0 : prec64;
        st->iters_fp32 = 0;
        st->maxres = st->maxres_in = 0.0;
      }
      if (vals[3] > 0.0) st->nonfinite = 1;
      st->ecorr = e;
      if (kind == FIN_INIT) {
        st->niter = 1;
        st->ecorr_prev = e;
        st->e_guess = e;
        st->dt = dt;
        st->slot = 0;
        st->slot_cur = 0;
        st->nhist = 0;
        st->next_src = SRC_TCUR;
        st->converged = 0;
        if (write_trace && S.trace && S.maxiter > 0) { S.trace[0] = e; S.trace[1] = dt; }
        const bool go = S.mode == UEGCCD_MODE_REFERENCE ? (dt >= S.tol_dt) : true;
        st->phase = (go && st->niter < S.maxiter && !st->nonfinite) ? PH_ITER : PH_FINAL;
        if (!go) st->converged = 1;
      }
    }
  }
  __syncthreads();
  return phase0;
}

// Writes an LDS state to a device state (each thread its words).
__device__ __forceinline__ void put_state(const State& sst, GlobalPtr<State> to) {
  const double* src = reinterpret_cast<const double*>(&sst);
  auto out = reinterpret_cast<GlobalPtr<double>>(to);
  for (int w = threadIdx.x; w < kStateWords; w += blockDim.x) out[w] = src[w];
}

// The reductions and state update of solve sidx by one kFinThreads block:
// the new state goes to st and to the pass copy st2.
__device__ __forceinline__ void finalize_block(const SolveDev* __restrict__ solves, int sidx, int kind, int prec64,
                               double* vals, double* red, State& sst) {
  const SolveG& S = as_global(solves)[sidx];
  const int phase0 = finalize_state<32>(S, kind, prec64, vals, red, sst, true);
  if (kind == FIN_ITER && phase0 == PH_DONE) return;
  put_state(sst, S.st);
  put_state(sst, S.st2);
}

__global__ __launch_bounds__(kFinThreads) void k_finalize(const SolveDev* __restrict__ solves, int kind,
                                                          int prec64) {
  __shared__ double vals[kPartialStride];
  __shared__ double red[(kFinThreads / 64) * 32];
  __shared__ State sst;
  finalize_block(solves, blockIdx.x, kind, prec64, vals, red, sst);
}

__global__ void k_state_reset(const SolveDev* __restrict__ solves, int phase, int prec64) {
  const SolveG& S = as_global(solves)[blockIdx.x];
  if (threadIdx.x > 1) return;
  const auto st = threadIdx.x == 0 ? S.st : S.st2;  // both copies
  st->niter = 0;
  st->phase = phase;
  st->slot = 0;
  st->slot_cur = 0;
  st->nhist = 0;
  st->next_src = SRC_TCUR;
  st->prec64 = prec64;
  st->iters_fp32 = 0;
  st->nonfinite = 0;
  st->converged = 0;
  st->ecorr = st->ecorr_prev = st->dt = st->maxres = st->maxres_in = st->e_guess = 0.0;
}

// Copies every solve's state into one contiguous buffer (one host copy per poll).
__global__ void k_gather_states(const SolveDev* __restrict__ solves, State* __restrict__ out) {
  const double* src = reinterpret_cast<const double*>(solves[blockIdx.x].st);
  double* dst = reinterpret_cast<double*>(out + blockIdx.x);
  for (int k = threadIdx.x; k < kStateWords; k += blockDim.x) dst[k] = src[k];
}

// --------------------------------------------------------------------------
// MP2 connectivity (DrvMBPT, source/MP2.f90:75-85): for every valid (i,j,a)
// with partner b, one count at |n_i - n_a|^2 and one at |n_i - n_b|^2.
// Per-block histograms in LDS, flushed with integer atomics (exact, any order).
constexpr int kMaxConnBins = 4096;
__global__ __launch_bounds__(256) void k_connectivity(const SolveDev* __restrict__ solves,
                                                      int nbins, int* __restrict__ hist,
                                                      int* __restrict__ overflow) {
  __shared__ int h[kMaxConnBins];
  const SolveG& S = as_global(solves)[0];
  for (int k = threadIdx.x; k < nbins; k += blockDim.x) h[k] = 0;
  __syncthreads();
  const int c = blockIdx.x * blockDim.x + threadIdx.x;  // compact index of a valid element
  if (c < S.nvalid) {
    const uint2 cm = ld(S.cmap + c);
    const int x = elem_x(S.ecode, cm), b = elem_partner(S.ecode, cm);
    const int o = S.o, i = x % o, a = o + x / (o * o);
    const int4 ni = S.nv4[i];
    const int d1 = d2(ni, S.nv4[a]), d2b = d2(ni, S.nv4[b]);
    if (d1 > nbins || d2b > nbins) {
      atomicOr(overflow, 1);
    } else {
      atomicAdd(&h[d1 - 1], 1);
      atomicAdd(&h[d2b - 1], 1);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}

int connectivity_on_device(const BatchDev& b, int nvalid, int nbins, int* hist, int* overflow,
                           hipStream_t s) {
  if (nbins > kMaxConnBins) return -1;
  hipLaunchKernelGGL(k_connectivity, dim3(std::max(1, (nvalid + 255) / 256)), dim3(256), 0, s, b.solves,
                     nbins, hist, overflow);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------------------
// dRPA / rCCD (source/dRPA.f90): the ring-only CCD on the spin-summed T2
// (held in the first component of tcur).  Every ring term of GetG2Rings
// (source/dRPA.f90:474-548) carries v(n_i - n_a) (the momentum of each
// Coulomb element there is n_i - n_a or its negative), so with the row sums
// R(i,a) = sum_k T(i,k,a) and the transfer sums Q(j,q) = sum over the pairs
// (l,d) of P(q) of T(l,j,d),
//   G(i,j,a) = v_ia [1 + 2 R(i,a) + 2 R(j,b) + 4 R(i,a) Q(j, n_a - n_i)]
//              - 1/2 (Jvv(a) + Jvv(b) + Joo(i) + Joo(j)) T(i,j,a)
// with the mosaic sums (source/dRPA.f90:420-436) Jvv(c) = sum_k 2 v(n_c - n_k)
// R(k,c) and Joo(k) = sum_c 2 v(n_c - n_k) R(k,c).  Work buffers: R in rsum,
// Joo in joo, Jvv in jvv, Q in lad (nqb*o doubles <= n3), and the transposed
// amplitudes tT(j,i,a) = T(i,j,a) in tv (n3 doubles), which make the sums
// over the first index contiguous reads.

__device__ inline GlobalPtr<double> drpa_tt(const SolveG& S) {
  return reinterpret_cast<GlobalPtr<double>>(S.tv);
}

// tT of the starting amplitudes (every position, zeros included), once per solve.
__global__ __launch_bounds__(256) void k_drpa_transpose(const SolveDev* __restrict__ solves) {
  const SolveG& S = as_global(solves)[0];
  const int o = S.o;
  const int64_t n3 = (int64_t)o * o * S.v, e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n3) return;
  const int j = (int)(e % o), i = (int)((e / o) % o);
  const int64_t slab = e / ((int64_t)o * o);
  drpa_tt(S)[e] = S.tcur[(size_t)i + (size_t)o * j + (size_t)o * o * slab].x;
}

// Block g*o + k (g < kDrpaJooParts, one wave): the part of Joo(k) from the
// slabs c of range g, sum over (c, j) of 2 v(n_c - n_k) tT(j,k,c), the lanes
// taking the pairs in turn.  Block P*o + a: R(:,a) and Jvv(a) of slab a.
// Block P*o + v + q: Q(:,q) of transfer q, lane j reading tT(j,l,d) of each
// pair (l,d) of P(q).
__global__ __launch_bounds__(64) void k_drpa_sums(const SolveDev* __restrict__ solves) {
  const SolveG& S = as_global(solves)[0];
  if (S.st->phase == PH_DONE) return;
  const int o = S.o, o2 = o * o, lane = threadIdx.x, nj = kDrpaJooParts * o;
  const auto T = S.tcur;
  const auto tT = drpa_tt(S);
  int blk = blockIdx.x;
  if (blk < nj) {
    if (!S.do_mosaic) return;
    const int k = blk % o, g = blk / o;
    const int c0 = (int)((int64_t)S.v * g / kDrpaJooParts), c1 = (int)((int64_t)S.v * (g + 1) / kDrpaJooParts);
    const int cnt = (c1 - c0) * o;
    double p = 0.0;
#pragma unroll 4
    for (int m = lane; m < cnt; m += 64) {
      const int c = c0 + m / o, j = m - (m / o) * o;
      p += S.vso[c * o + k] * tT[(size_t)j + (size_t)o * k + (size_t)o2 * c];
    }
    const double jo = wave_sum(p);
    if (lane == 0) S.joo[g * o + k] = 2.0 * jo;
    return;
  }
  blk -= nj;
  if (blk < S.v) {
    const int av = blk;
    double r = 0.0;
    if (lane < o) {
#pragma unroll 4
      for (int k = 0; k < o; ++k) r += T[(size_t)lane + o * k + (size_t)o2 * av].x;
    }
    if (lane < o) S.rsum[av * o + lane] = r;
    const double jv = wave_sum(lane < o ? 2.0 * S.vso[av * o + lane] * r : 0.0);
    if (lane == 0) S.jvv[av] = jv;
  } else {
    // the pairs of P(q) are loaded 64 at a time, one per lane, and each step
    // takes its row offset by readlane, so the row loads do not wait on the
    // pair list (the loop is uniform over the wave; lanes past o read row
    // o - 1 and store nothing)
    const int q = blk - S.v;
    const int r0 = S.qb_roff[q], r1 = S.qb_roff[q + 1], jl = min(lane, o - 1);
    double acc = 0.0;
    for (int c0 = r0; c0 < r1; c0 += 64) {
      const int nr = min(64, r1 - c0);
      const int pr = lane < nr ? S.qb_rows[c0 + lane] : 0;
      const int off = o * (pr % o) + o2 * (pr / o);  // o*l + o^2*d of pair (l, d)
#pragma unroll 8
      for (int r = 0; r < nr; ++r) acc += tT[(size_t)jl + __builtin_amdgcn_readlane(off, r)];
    }
    if (lane < o) reinterpret_cast<GlobalPtr<double>>(S.lad)[(size_t)q * o + lane] = acc;
  }
}

// One thread per valid element: G, R = D*T - G (GetRes, source/dRPA.f90:282-312),
// the damped update (SolveCC, 219-250), dT and the energy of the new
// amplitudes (CCEnergy with CSosex 0, 181-216).  In the final pass (the
// residual check, 100-103) the amplitudes stay.
__global__ __launch_bounds__(kElemThreads) void k_drpa_main(const SolveDev* __restrict__ solves) {
  __shared__ double red[(kElemThreads / 64) * 32];
  const SolveG& S = as_global(solves)[0];
  const int phase = S.st->phase;
  if (phase == PH_DONE) return;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int o = S.o, cp = lb * kElemPerBlock + threadIdx.x;
  int x = 0, b = -1;
  if (cp < S.nvalid) {
    const uint2 cm = ld(S.cmap + cp);
    x = elem_x(S.ecode, cm);
    b = elem_partner(S.ecode, cm);
  }
  __shared__ double joo_l[kMaxOcc];
  if (S.do_mosaic && (int)threadIdx.x < o) {
    double t = 0.0;
    for (int g = 0; g < kDrpaJooParts; ++g) t += S.joo[g * o + threadIdx.x];
    joo_l[threadIdx.x] = t;
  }
  __syncthreads();
  const bool lead = cp < S.nvalid;
  double e_part = 0.0, dt_part = 0.0, r_part = 0.0, bad = 0.0;
  if (lead) {
    const int i = x % o, j = (x / o) % o, av = x / (o * o), a = o + av, bv = b - o;
    const double t = S.tcur[x].x;
    const double vq = S.vso[av * o + i];  // v(n_i - n_a) = v(n_j - n_b)
    double g = vq;
    if (S.do_mosaic) {
      g = g - 0.5 * (S.jvv[av] * t + S.jvv[bv] * t);
      g = g - 0.5 * (joo_l[i] * t + joo_l[j] * t);
    }
    if (S.do_ring) {
      const double ria = S.rsum[av * o + i], rjb = S.rsum[bv * o + j];
      const double qj = reinterpret_cast<GlobalPtr<const double>>(S.lad)[(size_t)S.qb_of[av * o + i] * o + j];
      g += 2 * vq * ria + 2 * vq * rjb + 4 * vq * ria * qj;
    }
    const double den = S.eig[i] + S.eig[j] - S.eig[a] - S.eig[b];
    const double res = -g + den * t;
    r_part = fabs(res);
    double tn = t;
    if (phase == PH_ITER) {
      const double df = S.df;
      tn = (g / den - (1.0 - df) * t) / df;
      dt_part = fabs(tn - t);
      S.tcur[x] = make_double2(tn, 0.0);
      drpa_tt(S)[(size_t)j + (size_t)o * i + (size_t)o * o * av] = tn;
    }
    e_part = vq * tn * 2;
    if (!isfinite(tn) || !isfinite(res)) bad = 1.0;
  }
  main_partials<16>(S, lb, 0, 0, lead, dt_part, r_part, e_part, bad, 0.0, 0.0, S.histE, 0, red);
}

// --------------------------------------------------------------------------
// Structure factor (CCEnergyAndSF, source/CCD.f90:240-351).  Each valid
// element (i,j,a) adds to the cell q = n_a - n_i (direct) and subtracts at
// q = n_a - n_j (exchange); with T^aa = T^ab + T^ba the three sums are
//   Sum: +(2 T^ab + T^ba), -(T^ab + 2 T^ba)   DSum: +2 T^ab, -2 T^ba
//   XSum: +T^ba, -T^ab.
// So a pair (l,d) of the ring sets P(q) = {(l,d) : n_d - n_l = q} brings its
// row sums over j (direct) less its column sums over i (exchange), and a cell
// is the sum over its pairs.  Invalid amplitudes are zero.  Scratch: the pair
// sums in jring (3 o v doubles), the cell sums in jxring (3 nq doubles).

// Block a (one wave, lane l < o): the three sums of pair (l, a).
__global__ __launch_bounds__(64) void k_sf_pairs(const SolveDev* __restrict__ solves) {
  const SolveG& S = as_global(solves)[0];
  const int o = S.o, o2 = o * o, av = blockIdx.x, l = threadIdx.x;
  if (l >= o) return;
  const auto T = S.tcur;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = 0; k < o; ++k) {
    const double2 d = T[(size_t)l + o * k + (size_t)o2 * av];  // (l,k,a): the direct row of l
    const double2 x = T[(size_t)k + o * l + (size_t)o2 * av];  // (k,l,a): the exchange column of l
    s0 += (2.0 * d.x + d.y) - (x.x + 2.0 * x.y);
    s1 += 2.0 * d.x - 2.0 * x.y;
    s2 += d.y - x.x;
  }
  const size_t w = 3 * ((size_t)av * o + l);
  const auto JP = reinterpret_cast<GlobalPtr<double>>(S.jring);
  JP[w] = s0;
  JP[w + 1] = s1;
  JP[w + 2] = s2;
}

// Block q (one wave): the cell sums of transfer q over the pairs of P(q).
__global__ __launch_bounds__(64) void k_sf_cells(const SolveDev* __restrict__ solves) {
  const SolveG& S = as_global(solves)[0];
  const int q = blockIdx.x, r0 = S.qb_roff[q], r1 = S.qb_roff[q + 1];
  const auto JP = reinterpret_cast<GlobalPtr<const double>>(S.jring);
  const auto JC = reinterpret_cast<GlobalPtr<double>>(S.jxring);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int r = r0 + (int)threadIdx.x; r < r1; r += 64) {
    const size_t w = 3 * (size_t)S.qb_rows[r];  // row l + o*d of pair (l, d)
    a0 += JP[w];
    a1 += JP[w + 1];
    a2 += JP[w + 2];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  if (threadIdx.x == 0) {
    JC[3 * (size_t)q] = a0;
    JC[3 * (size_t)q + 1] = a1;
    JC[3 * (size_t)q + 2] = a2;
  }
}

// --------------------------------------------------------------------------
// Plane-wave basis on the device (init_HEG + SetupBasis + change_rs,
// source/HEG.f90:150-324, 494-533, 581-615).  Every floating-point step is an
// explicitly rounded operation in the host builder's order (no contraction
// into fused multiply-adds), so the ordering keys, k-vectors, eigenvalues and
// E_HF are the host's to the last bit.

// Keys |n + twist|^2 of the cube [-cell, cell]^3 in enumeration order (x
// slowest), +inf outside |n|^2 <= maxk and in the padding.
__global__ void k_basis_keys(int cell, int maxk, int twisted, double tx, double ty, double tz, int npad,
                             double* __restrict__ key, int* __restrict__ idx) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= npad) return;
  const int w = 2 * cell + 1;
  double kv = __longlong_as_double(0x7ff0000000000000LL);  // +inf
  if (e < w * w * w) {
    const int a = e / (w * w) - cell, b = (e / w) % w - cell, c = e % w - cell;
    const double x = a + tx, y = b + ty, z = c + tz;
    const double k2 = x * x + y * y + z * z;
    // the reference's sphere |n|^2 <= maxk, or the twisted one |n+t|^2 <= maxk
    if (twisted ? k2 <= (double)maxk : a * a + b * b + c * c <= maxk) kv = k2;
  }
  key[e] = kv;
  idx[e] = e;
}

// Bitonic sort of (key, enumeration index) pairs in LDS by one block: the
// index breaks ties, which is the stable order of the reference's sort.  Also
// counts the finite keys (the plane waves inside the cutoff).
__global__ __launch_bounds__(1024) void k_basis_sort(int npad, double* __restrict__ key,
                                                     int* __restrict__ idx, int* __restrict__ count) {
  extern __shared__ double sk[];
  int* si = reinterpret_cast<int*>(sk + npad);
  __shared__ int nfin;
  if (threadIdx.x == 0) nfin = 0;
  for (int t = threadIdx.x; t < npad; t += blockDim.x) {
    sk[t] = key[t];
    si[t] = idx[t];
  }
  __syncthreads();
  for (int k = 2; k <= npad; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < npad; t += blockDim.x) {
        const int u = t ^ j;
        if (u > t) {
          const bool up = (t & k) == 0;
          const double ka = sk[t], kb = sk[u];
          const int ia = si[t], ib = si[u];
          const bool a_gt_b = ka > kb || (ka == kb && ia > ib);
          if (a_gt_b == up) {
            sk[t] = kb; sk[u] = ka;
            si[t] = ib; si[u] = ia;
          }
        }
      }
      __syncthreads();
    }
  }
  int mine = 0;
  for (int t = threadIdx.x; t < npad; t += blockDim.x) {
    key[t] = sk[t];
    idx[t] = si[t];
    mine += isfinite(sk[t]) ? 1 : 0;
  }
  atomicAdd(&nfin, mine);
  __syncthreads();
  if (threadIdx.x == 0) *count = nfin;
}

struct BasisArgs {
  int cell, nb, no, mc, nel;
  double tx, ty, tz;
  double L1, rescale, omega, mad, gap, xf;
};

__device__ __forceinline__ void basis_kvec(const BasisArgs& A, const int* idx, int q, double* k,
                                           int* n) {
#pragma clang fp contract(off)
  const int src = q < A.no ? q : q + A.mc - A.no;
  const int e = idx[src], w = 2 * A.cell + 1;
  n[0] = e / (w * w) - A.cell;
  n[1] = (e / w) % w - A.cell;
  n[2] = e % w - A.cell;
  const double tw[3] = {A.tx, A.ty, A.tz};
  const double g = 2.0 * 3.14159265358979323846264338327950288 / A.L1;
  const double inv = 1.0 / A.rescale;
  for (int d = 0; d < 3; ++d) k[d] = ((n[d] + tw[d]) * g) * inv;
}

// Orbitals of the truncated basis: n-vectors (nVec(NAO,3) column-major), HF
// eigenvalues with the exchange over occupied orbitals and the Madelung term
// on the occupied diagonal; E_HF by thread 0 of block 0 in the host's order.
__global__ __launch_bounds__(256) void k_basis_orbitals(BasisArgs A, const int* __restrict__ idx,
                                                        int* __restrict__ nvec,
                                                        double* __restrict__ eigen,
                                                        double* __restrict__ ehf_out) {
#pragma clang fp contract(off)
  __shared__ double ko[kMaxOcc * 3];
  for (int l = threadIdx.x; l < A.no; l += blockDim.x) {
    int n[3];
    basis_kvec(A, idx, l, ko + 3 * l, n);
  }
  __syncthreads();
  const double pi4 = 4.0 * 3.14159265358979323846264338327950288;
  const double inv_omega = 1.0 / A.omega;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < A.nb) {
    double k1[3];
    int n[3];
    basis_kvec(A, idx, q, k1, n);
    for (int d = 0; d < 3; ++d) nvec[(size_t)d * A.nb + q] = n[d];
    double e = 0.5 * (k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
    for (int l = 0; l < A.no; ++l) {
      if (l != q) {
        const double d0 = k1[0] - ko[3 * l], d1 = k1[1] - ko[3 * l + 1], d2v = k1[2] - ko[3 * l + 2];
        e = e - inv_omega * (pi4 / (d0 * d0 + d1 * d1 + d2v * d2v)) * A.xf;
      } else {
        e = e + A.mad - A.gap;
      }
    }
    eigen[q] = e;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double ehf = 0.0;
    for (int a = 0; a < A.no; ++a) {
      const double* k1 = ko + 3 * a;
      ehf = ehf + 0.5 * (k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
      for (int l = 0; l < A.no; ++l) {
        if (l == a) continue;
        const double* k2 = ko + 3 * l;
        const double d0 = k1[0] - k2[0], d1 = k1[1] - k2[1], d2v = k1[2] - k2[2];
        ehf = ehf - 0.5 * inv_omega * (pi4 / (d0 * d0 + d1 * d1 + d2v * d2v));
      }
    }
    *ehf_out = ehf * 2.0 + A.mad * 0.5 * A.nel;
  }
}

int basis_on_device(int cell, int maxk, int twisted, const double* twist, int npad, double* key, int* idx,
                    int* count, hipStream_t s) {
  hipLaunchKernelGGL(k_basis_keys, dim3((npad + 255) / 256), dim3(256), 0, s, cell, maxk, twisted, twist[0],
                     twist[1], twist[2], npad, key, idx);
  const int lds = npad * (int)(sizeof(double) + sizeof(int));
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_basis_sort),
                      hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  hipLaunchKernelGGL(k_basis_sort, dim3(1), dim3(1024), lds, s, npad, key, idx, count);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int basis_orbitals_on_device(int cell, int nb, int no, int mc, int nel, const double* twist,
                             double L1, double rescale, double omega, double mad, double gap,
                             double xf, const int* idx, int* nvec, double* eigen, double* ehf,
                             hipStream_t s) {
  BasisArgs A{cell, nb, no, mc, nel, twist[0], twist[1], twist[2], L1, rescale, omega, mad, gap, xf};
  hipLaunchKernelGGL(k_basis_orbitals, dim3((nb + 255) / 256), dim3(256), 0, s, A, idx, nvec, eigen,
                     ehf);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------------------
// Persistent pass kernel: a chunk of passes of one solve in one cooperative
// launch.  Every block (256 threads) loops over each phase's work items
// (slice blocks, ladder tiles, contraction items four to a block, element
// blocks, the solve's finalize) and a grid-wide barrier separates the
// phases, in place of the five launches of a pass.  The LDS of the phases
// share one buffer; a block barrier opens every work item, since the waves
// of a block leave an item's epilogue at different times.
struct PersistLds {
  union {
    struct { Lds L; SliceDiis D; } sl;
    struct { double red[3 * 64 * 8]; } la;
    struct { double red[(kElemThreads / 64) * 32]; double joo[kMaxOcc]; } mn;
    struct { double vals[kPartialStride]; double red[(kFinThreads / 64) * 32]; State sst; } fi;
  } u;
};

__global__ __launch_bounds__(256) void k_persist(const SolveDev* __restrict__ solves,
                                                 const int* __restrict__ slice_prefix,
                                                 const int* __restrict__ hh_prefix,
                                                 const int* __restrict__ lad_prefix,
                                                 const int* __restrict__ con_prefix,
                                                 const int* __restrict__ elem_prefix, int nslice, int nhh,
                                                 int nlad, int ncon, int nelem, int passes, int tab_doubles,
                                                 int cls_max) {
  __shared__ PersistLds P;
  extern __shared__ double pdyn[];  // slice slabs, then the ladder tables, a virtual list, W rows
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ncb = (ncon + kItemsPerBlock - 1) / kItemsPerBlock;
  const LadderLds ll = ladder_lds(pdyn, tab_doubles, cls_max);
  for (int pass = 0; pass < passes; ++pass) {
    for (int it = blockIdx.x; it < nslice; it += gridDim.x) {
      __syncthreads();
      slice_block<256, false>(solves, slice_prefix, 1, it, P.u.sl.L, P.u.sl.D, pdyn, nullptr, 0, FoldArgs{}, nullptr,
                              nullptr, nullptr);
    }
    grid.sync();
    for (int it = blockIdx.x; it < nhh + nlad + ncb; it += gridDim.x) {
      __syncthreads();
      if (it < nhh) {
        hh_block<2, 8>(solves, hh_prefix, 1, it, 1, ll, P.u.la.red);
      } else if (it < nhh + nlad) {
        ladder_block<2, 1>(solves, lad_prefix, 1, it - nhh, ll, P.u.la.red);
      } else {
        const int item = (it - nhh - nlad) * kItemsPerBlock + wave;
        if (item < ncon) contract_item<8>(solves, con_prefix, 1, item);
      }
    }
    grid.sync();
    for (int it = blockIdx.x; it < nelem; it += gridDim.x) {
      __syncthreads();
      main_block<0>(solves, elem_prefix, 1, it, MF_ITERATE | MF_HH2, P.u.mn.red, P.u.mn.joo);
    }
    grid.sync();
    if (blockIdx.x == 0) finalize_block(solves, 0, FIN_ITER, 1, P.u.fi.vals, P.u.fi.red, P.u.fi.sst);
    grid.sync();
  }
}

int launch_persistent(const BatchDev& b, int dyn_lds, int passes, hipStream_t s) {
  static int grid = 0, grid_lds = -1;
  if (b.n != 1 || b.slice_split) return -1;
  if (grid_lds != dyn_lds) {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
    if (dyn_lds > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_persist), hipFuncAttributeMaxDynamicSharedMemorySize,
                            dyn_lds) != hipSuccess)
      return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persist, 256, dyn_lds) != hipSuccess ||
        per_cu < 1)
      return -1;
    grid = per_cu * prop.multiProcessorCount;
    grid_lds = dyn_lds;
  }
  const int nslice = b.slice_blocks;
  const int nhh = b.nhh, nlad = b.nlad, ncon = b.ncon, nelem = b.elem_blocks;
  const int tab_doubles = b.tab_doubles, cls_max = b.cls_max;
  // no more blocks than the largest phase has work items: a grid barrier's
  // cost grows with the blocks that take part
  const int need = std::max(std::max(nslice, nhh + nlad + (ncon + kItemsPerBlock - 1) / kItemsPerBlock), nelem);
  const int g = std::max(1, std::min(grid, need));
  int cls = 0;
  for (int q = 0; q < kSliceClasses; ++q)
    if (b.slice_cls_blocks[q] > 0) cls = q;
  const SolveDev* sd = b.solves;
  const int* sp = b.slice_prefix[cls];
  const int* hp = b.hh_prefix;
  const int* lp = b.lad_prefix;
  const int* cp = b.con_prefix;
  const int* ep = b.elem_prefix;
  void* args[] = {(void*)&sd,    (void*)&sp,     (void*)&hp,          (void*)&lp,
                  (void*)&cp,    (void*)&ep,     (void*)&nslice,      (void*)&nhh,
                  (void*)&nlad,  (void*)&ncon,   (void*)&nelem,       (void*)&passes,
                  (void*)&tab_doubles, (void*)&cls_max};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_persist), dim3(g), dim3(256), args,
                                    (unsigned)dyn_lds, s) == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------------------
// Dynamic LDS above the 64 KiB default needs an explicit per-kernel limit.
// The limit of each kernel is the LDS of a CU less the kernel's own static
// LDS (as the runtime reports it).
static bool allow_dyn_lds(const void* f) {
  hipFuncAttributes a;
  if (hipFuncGetAttributes(&a, f) != hipSuccess) return false;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxDynLds - (int)a.sharedSizeBytes) ==
         hipSuccess;
}

static void allow_large_lds() {
  static bool done = false;
  if (done) return;
  const void* fs[] = {
      reinterpret_cast<const void*>(&k_slice<256, false>), reinterpret_cast<const void*>(&k_slice<512, false>),
      reinterpret_cast<const void*>(&k_slice<768, false>), reinterpret_cast<const void*>(&k_slice<256, true>),
      reinterpret_cast<const void*>(&k_slice<512, true>),  reinterpret_cast<const void*>(&k_slice<768, true>),
#define UEGCCD_CONTRACT_NC(G, Q, H) reinterpret_cast<const void*>(&k_contract<G, Q, H, 1>), \
                                    reinterpret_cast<const void*>(&k_contract<G, Q, H, 2>)
      UEGCCD_CONTRACT_NC(1, 4, 4), UEGCCD_CONTRACT_NC(1, 4, 8), UEGCCD_CONTRACT_NC(1, 8, 4),
      UEGCCD_CONTRACT_NC(1, 8, 8), UEGCCD_CONTRACT_NC(2, 4, 4), UEGCCD_CONTRACT_NC(2, 4, 8),
      UEGCCD_CONTRACT_NC(2, 8, 4), UEGCCD_CONTRACT_NC(2, 8, 8), UEGCCD_CONTRACT_NC(4, 4, 4),
      UEGCCD_CONTRACT_NC(4, 4, 8), UEGCCD_CONTRACT_NC(4, 8, 4), UEGCCD_CONTRACT_NC(4, 8, 8)};
#undef UEGCCD_CONTRACT_NC
  bool refused = false;
  for (const void* f : fs) refused |= !allow_dyn_lds(f);
  if (refused) (void)hipGetLastError();  // a refused limit leaves the default, and no error behind
  done = true;
}

// Launch through the extended API with an optional stop event, which the
// dispatch itself writes when it completes.  A kernel is timed by the stop
// events of the kernel before it and of itself: kernels of one stream run
// back to back (a kernel trace shows no gap between them), so the difference
// is the kernel's trace duration.  (A start event would be a separate marker
// packet ahead of the dispatch and add its own time.)
template <typename F, typename... Args>
static void launch_k(F kernel, hipEvent_t stop, dim3 grid, dim3 block, uint32_t lds, hipStream_t s,
                     Args... args) {
  hipExtLaunchKernelGGL(kernel, grid, block, lds, s, nullptr, stop, 0, args...);
}

void prepare_launches() { allow_large_lds(); }

int contract_lds_limit() {
  static const int lim = [] {
    const void* fs[] = {
#define UEGCCD_CONTRACT_NC(G, Q, H) reinterpret_cast<const void*>(&k_contract<G, Q, H, 1>), \
                                    reinterpret_cast<const void*>(&k_contract<G, Q, H, 2>)
        UEGCCD_CONTRACT_NC(1, 4, 4), UEGCCD_CONTRACT_NC(1, 4, 8), UEGCCD_CONTRACT_NC(1, 8, 4),
        UEGCCD_CONTRACT_NC(1, 8, 8), UEGCCD_CONTRACT_NC(2, 4, 4), UEGCCD_CONTRACT_NC(2, 4, 8),
        UEGCCD_CONTRACT_NC(2, 8, 4), UEGCCD_CONTRACT_NC(2, 8, 8), UEGCCD_CONTRACT_NC(4, 4, 4),
        UEGCCD_CONTRACT_NC(4, 4, 8), UEGCCD_CONTRACT_NC(4, 8, 4), UEGCCD_CONTRACT_NC(4, 8, 8)};
#undef UEGCCD_CONTRACT_NC
    int m = kMaxDynLds;
    for (const void* f : fs) {
      hipFuncAttributes a;
      if (hipFuncGetAttributes(&a, f) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
      }
      m = std::min(m, kMaxDynLds - (int)a.sharedSizeBytes);
    }
    return m;
  }();
  return lim;
}

static void launch_slice(const BatchDev& b, hipStream_t s, hipEvent_t stop = nullptr, int fin_prev = 0) {
  allow_large_lds();
  const auto* sd = (const SolveDev*)b.solves;
  int last = -1;  // the stop event goes with the last launch
  for (int q = 0; q < kSliceClasses; ++q)
    if (b.slice_cls_blocks[q] > 0) last = q;
  for (int q = 0; q < kSliceClasses; ++q) {
    if (b.slice_cls_blocks[q] == 0) continue;
    const dim3 g(b.slice_cls_blocks[q]), t(256 * (q + 1));
    const auto* pf = (const int*)b.slice_prefix[q];
    const hipEvent_t ev = q == last ? stop : nullptr;
    if (b.slice_split) {
      if (q == 0) launch_k(k_slice<256, true>, ev, g, t, b.slice_lds[q], s, sd, pf, b.n, fin_prev, b.fold);
      else if (q == 1) launch_k(k_slice<512, true>, ev, g, t, b.slice_lds[q], s, sd, pf, b.n, fin_prev, b.fold);
      else launch_k(k_slice<768, true>, ev, g, t, b.slice_lds[q], s, sd, pf, b.n, fin_prev, b.fold);
    } else {
      if (q == 0) launch_k(k_slice<256, false>, ev, g, t, b.slice_lds[q], s, sd, pf, b.n, fin_prev, b.fold);
      else if (q == 1) launch_k(k_slice<512, false>, ev, g, t, b.slice_lds[q], s, sd, pf, b.n, fin_prev, b.fold);
      else launch_k(k_slice<768, false>, ev, g, t, b.slice_lds[q], s, sd, pf, b.n, fin_prev, b.fold);
    }
  }
}

template <int G, int QS, int HS>
static void launch_contract_t(const BatchDev& b, hipStream_t s, hipEvent_t stop) {
  // (two column tiles per ladder block, sharing each k-step's amplitude
  // loads: with the hole-hole units split, a solve alone gains from the
  // shorter launch too, C2's contraction 18.8 -> 17.9 us; the one-tile items
  // serve the persistent kernel)
  const int nc = 2, nlad = nc == 2 ? b.nlad2 : b.nlad, nlad_pad = (nlad + 7) & ~7;
  const int* lad_prefix = nc == 2 ? b.lad_prefix2 : b.lad_prefix;
  const int grid = b.nhh_pad + nlad_pad + (b.ncon + kItemsPerBlock - 1) / kItemsPerBlock;
  if (nc == 2)
    launch_k(k_contract<G, QS, HS, 2>, stop, dim3(grid), dim3(256), (uint32_t)b.con_lds, s,
             (const SolveDev*)b.solves, (const int*)b.hh_prefix, b.nhh, b.nhh_pad, lad_prefix,
             nlad, nlad_pad, (const int*)b.con_prefix, b.ncon, b.n, b.tab_doubles, b.cls_max, b.hh_split);
  else
    launch_k(k_contract<G, QS, HS, 1>, stop, dim3(grid), dim3(256), (uint32_t)b.con_lds, s,
             (const SolveDev*)b.solves, (const int*)b.hh_prefix, b.nhh, b.nhh_pad, lad_prefix,
             nlad, nlad_pad, (const int*)b.con_prefix, b.ncon, b.n, b.tab_doubles, b.cls_max, b.hh_split);
}

template <int G>
static void launch_contract_g(const BatchDev& b, hipStream_t s, hipEvent_t stop) {
  const int q8 = b.tiles.quad_steps == 4 ? 0 : 1, h8 = b.tiles.hh_steps == 4 ? 0 : 1;
  if (q8 && h8) launch_contract_t<G, 8, 8>(b, s, stop);
  else if (q8) launch_contract_t<G, 8, 4>(b, s, stop);
  else if (h8) launch_contract_t<G, 4, 8>(b, s, stop);
  else launch_contract_t<G, 4, 4>(b, s, stop);
}

static void launch_contract(const BatchDev& b, hipStream_t s, hipEvent_t stop = nullptr) {
  allow_large_lds();
  switch (b.tiles.ladder_group) {
    case 1: launch_contract_g<1>(b, s, stop); break;
    case 4: launch_contract_g<4>(b, s, stop); break;
    default: launch_contract_g<2>(b, s, stop); break;
  }
}

void launch_init(const BatchDev& b, hipStream_t s, int drpa) {
  hipLaunchKernelGGL(k_init, dim3(b.elem_blocks), dim3(kElemPerBlock), 0, s, b.solves,
                     b.elem_prefix, b.n, drpa);
}
void launch_mp2(const BatchDev& b, hipStream_t s, int drpa) {
  hipLaunchKernelGGL(k_mp2, dim3(b.elem_blocks), dim3(kElemPerBlock), 0, s, b.solves,
                     b.elem_prefix, b.n, drpa);
}

void launch_sf(const BatchDev& b, int v, int nq, hipStream_t s) {
  hipLaunchKernelGGL(k_sf_pairs, dim3(v), dim3(64), 0, s, b.solves);
  hipLaunchKernelGGL(k_sf_cells, dim3(nq), dim3(64), 0, s, b.solves);
}
// Amplitudes between a solve's dense (i,j,a) arrays, staged in device
// memory, and tcur.  k_place writes each valid element's (T^ab, T^ba) into
// tcur (whose other entries the caller has cleared; a NULL array reads as
// zeros); k_split unpacks every entry of tcur into T^ab, T^ba and
// T^aa = T^ab + T^ba (invalid entries hold zeros).
__global__ __launch_bounds__(256) void k_place_amplitudes(const SolveDev* __restrict__ solves,
                                                          const double* __restrict__ ab,
                                                          const double* __restrict__ ba) {
  const SolveG& S = as_global(solves)[0];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.nvalid) return;
  const int x = elem_x(S.ecode, ld(S.cmap + c));
  st2(S.tcur + x, make_double2(ab ? ab[x] : 0.0, ba ? ba[x] : 0.0));
}

__global__ __launch_bounds__(256) void k_split_amplitudes(const SolveDev* __restrict__ solves, double* __restrict__ ab,
                                                          double* __restrict__ ba, double* __restrict__ aa) {
  const SolveG& S = as_global(solves)[0];
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= S.n3) return;
  const double2 t = S.tcur[x];
  ab[x] = t.x;
  ba[x] = t.y;
  aa[x] = t.x + t.y;
}

void launch_place_amplitudes(const BatchDev& b, int64_t nvalid, const double* ab, const double* ba,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_place_amplitudes, dim3((unsigned)std::max<int64_t>(1, (nvalid + 255) / 256)), dim3(256), 0,
                     s, b.solves, ab, ba);
}

void launch_split_amplitudes(const BatchDev& b, int64_t n3, double* ab, double* ba, double* aa, hipStream_t s) {
  hipLaunchKernelGGL(k_split_amplitudes, dim3((unsigned)std::max<int64_t>(1, (n3 + 255) / 256)), dim3(256), 0, s,
                     b.solves, ab, ba, aa);
}

void launch_drpa_start(const BatchDev& b, int64_t n3, hipStream_t s) {
  hipLaunchKernelGGL(k_drpa_transpose, dim3((unsigned)std::max<int64_t>(1, (n3 + 255) / 256)), dim3(256), 0, s,
                     b.solves);
}
void launch_drpa_pass(const BatchDev& b, hipStream_t s) {
  // a batch of one: the sums over its slabs and transfers, then its elements
  hipLaunchKernelGGL(k_drpa_sums, dim3(b.drpa_sum_blocks), dim3(64), 0, s, b.solves);
  hipLaunchKernelGGL(k_drpa_main, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinThreads), 0, s, b.solves, (int)FIN_ITER, 1);
}
void launch_finalize(const BatchDev& b, int kind, hipStream_t s, int prec64) {
  hipLaunchKernelGGL(k_finalize, dim3(b.n), dim3(kFinThreads), 0, s, b.solves, kind, prec64);
}
void launch_state_reset(const BatchDev& b, int phase, int prec64, hipStream_t s) {
  hipLaunchKernelGGL(k_state_reset, dim3(b.n), dim3(64), 0, s, b.solves, phase, prec64);
}

// --------------------------------------------------------------------------
// Maps of a context on the device (launch_plan).  The box index of momentum n
// is center + n.x + w n.y + w^2 n.z; every lookup n_p + n_q - n_r of these
// kernels, n_i + n_j - n_a and n_l + n_p - n_i with i, j, l occupied, stays
// inside the box (its half-width is max|n| + 2 max|n_occupied| + 1).
__device__ __forceinline__ int box_lin(int4 n, int w) { return n.x + w * n.y + w * w * n.z; }

// Exclusive prefix of `flag` over a 256-thread block in thread order (wave
// ballots, then the four wave counts); *total receives the block's count.
__device__ __forceinline__ int block_rank(bool flag, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(flag);
  const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wave] = __popcll(m);
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < wave) before += wsum[k];
    all += wsum[k];
  }
  __syncthreads();  // wsum is written again by the next call
  *total = all;
  return before + in_wave;
}

// The box: orbital p + 1 at the box index of n_p.  Every planning kernel
// covers a batch: a block finds its solve by the launch's block prefix.
__global__ __launch_bounds__(256) void k_plan_box(const PlanArgs* __restrict__ pa, const int* __restrict__ prefix,
                                                  int n) {
  const int s = find_solve(prefix, n, blockIdx.x);
  const PlanArgs& P = pa[s];
  const SolveDev& S = *P.sd;
  const int p = (blockIdx.x - prefix[s]) * blockDim.x + threadIdx.x;
  if (p < S.nao) P.box[P.center + box_lin(S.nv4[p], P.w)] = p + 1;
}

// Blocks [0, nK) of a solve: block K's virtual list, the virtuals a whose
// partner b = K - n_a is a virtual orbital (source/CCD.f90: FindIndex(i,j,a)
// > NOcc), in a order, and civ[a][K].  The rest: qb_d[q][l], the virtual
// index of n_l + q (or -1), one thread per (q, l).
__global__ __launch_bounds__(256) void k_plan_blocks(const PlanArgs* __restrict__ pa,
                                                     const int* __restrict__ prefix, int n) {
  __shared__ int wsum[4];
  const int s = find_solve(prefix, n, blockIdx.x);
  const PlanArgs& P = pa[s];
  const SolveDev& S = *P.sd;
  const int o = S.o, v = S.v, nK = S.nK, w = P.w;
  const int blk = blockIdx.x - prefix[s];
  if (blk < nK) {
    const int K = blk;
    const int x0 = S.kb_pairs[S.kb_poff[K]];
    const int kb = P.center + box_lin(S.nv4[x0 % o], w) + box_lin(S.nv4[x0 / o], w);
    const int voff = S.kb_voff[K];
    int* const va = const_cast<int*>(S.kb_va);
    int* const vb = const_cast<int*>(S.kb_vb);
    int4* const nva = const_cast<int4*>(S.kb_nva);
    int run = 0;
    for (int a0 = 0; a0 < v; a0 += 256) {
      const int av = a0 + threadIdx.x;
      int b = -1;
      int4 na = make_int4(0, 0, 0, 0);
      if (av < v) {
        na = S.nv4[o + av];
        b = P.box[kb - box_lin(na, w)] - 1;
      }
      const bool valid = av < v && b >= o;
      int tot;
      const int rank = block_rank(valid, wsum, &tot);
      if (valid) {
        const int e = voff + run + rank;
        va[e] = o + av;
        vb[e] = b;
        nva[e] = make_int4(na.x, na.y, na.z, o + av);
      }
      if (av < v) P.civ[(size_t)av * nK + K] = valid ? run + rank : -1;
      run += tot;
    }
    return;
  }
  const int t = (blk - nK) * blockDim.x + threadIdx.x;
  if (t >= S.nqb * o) return;
  const int q = t / o, l = t - q * o;
  const int r = S.qb_rows[S.qb_roff[q]];  // a pair (i, p) of the transfer: q = n_p - n_i
  const int i = r % o, pv = r / o;
  const int ql = box_lin(S.nv4[o + pv], w) - box_lin(S.nv4[i], w);
  const int p = P.box[P.center + box_lin(S.nv4[l], w) + ql] - 1;
  const_cast<int*>(S.qb_d)[t] = p >= o ? p - o : -1;
}

// One block per virtual slice a of a solve: its valid elements (i, j, a) in
// element order, pair x = i + o j ascending, from slab_coff[a] on: the
// element, its K-blocked position and its partner.  Elements with a core
// index (i or j < coren, the CoreN masking) are left out.
__global__ __launch_bounds__(256) void k_plan_cmap(const PlanArgs* __restrict__ pa, const int* __restrict__ prefix,
                                                   int n) {
  __shared__ int wsum[4];
  const int s = find_solve(prefix, n, blockIdx.x);
  const PlanArgs& P = pa[s];
  const SolveDev& S = *P.sd;
  const int o = S.o, o2 = o * o, nK = S.nK, av = blockIdx.x - prefix[s];
  const int base = S.slab_coff[av];
  const int* civ = P.civ + (size_t)av * nK;
  uint2* const cmap = const_cast<uint2*>(S.cmap);
  int run = 0;
  for (int p0 = 0; p0 < o2; p0 += 256) {
    const int pr = p0 + threadIdx.x;
    int K = 0, ci = -1;
    if (pr < o2) {
      K = S.pair_kb[pr];
      ci = civ[K];
      if (pr % o < P.coren || pr / o < P.coren) ci = -1;
    }
    int tot;
    const int rank = block_rank(ci >= 0, wsum, &tot);
    if (ci >= 0) {
      const int noo = S.kb_poff[K + 1] - S.kb_poff[K];
      const int pos = S.kb_toff[K] + ci * noo + S.pair_pos[pr];
      cmap[base + run + rank] = elem_map(S.ecode, pr + o2 * av, pos, S.kb_vb[S.kb_voff[K] + ci]);
    }
    run += tot;
  }
}

void launch_plan(const BatchDev& b, hipStream_t s) {
  hipLaunchKernelGGL(k_plan_box, dim3(b.plan_blocks[0]), dim3(256), 0, s, b.plan, b.plan_prefix[0], b.n);
  hipLaunchKernelGGL(k_plan_blocks, dim3(b.plan_blocks[1]), dim3(256), 0, s, b.plan, b.plan_prefix[1], b.n);
  hipLaunchKernelGGL(k_plan_cmap, dim3(b.plan_blocks[2]), dim3(256), 0, s, b.plan, b.plan_prefix[2], b.n);
}

// Job blockIdx.y of a scatter launch, its 16-byte words strided over the
// launch's blocks in x.
__global__ __launch_bounds__(256) void k_scatter(const CopyJob* __restrict__ jobs) {
  const CopyJob J = jobs[blockIdx.y];
  uint4* const d = reinterpret_cast<uint4*>(J.dst);
  const uint4* const src = reinterpret_cast<const uint4*>(J.src);
  const size_t n16 = J.bytes / 16;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += (size_t)gridDim.x * blockDim.x)
    d[k] = src ? src[k] : make_uint4(0u, 0u, 0u, 0u);
}

void launch_scatter(const CopyJob* jobs, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_scatter, dim3(64, n), dim3(256), 0, s, jobs);
}

__global__ void k_zero(double* __restrict__ p, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    p[k] = 0.0;
}

void launch_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / sizeof(double);
  if (n == 0) return;
  const unsigned grid = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_zero, dim3(grid), dim3(256), 0, s, static_cast<double*>(p), n);
}

void launch_gather_states(const BatchDev& b, State* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_states, dim3(b.n), dim3(64), 0, s, b.solves, out);
}

// The kernels of one pass in launch order (numbering as KernelTimer::which).
static int pass_kernels(const BatchDev& b, int main_flags, int* order) {
  (void)main_flags;
  int n = 0;
  if (b.slice_blocks > 0) order[n++] = 0;
  if (b.nhh + b.nlad + b.ncon > 0) order[n++] = 1;
  order[n++] = 3;
  return n;
}

int first_pass_kernel(const BatchDev& b, int main_flags) {
  int order[5];
  pass_kernels(b, main_flags, order);
  return order[0];
}

void launch_one(const BatchDev& b, int which, hipStream_t s) {
  switch (which) {
    case 0:
      if (b.slice_blocks > 0) launch_slice(b, s);
      break;
    case 1:
      if (b.nhh + b.nlad + b.ncon > 0) launch_contract(b, s);
      break;
    case 3:
      if (b.main_nq == 16)
        hipLaunchKernelGGL(k_main<16>, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                           b.elem_prefix, b.n, (int)MF_ITERATE | (b.hh_split ? (int)MF_HH2 : 0));
      else
        hipLaunchKernelGGL(k_main<32>, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                           b.elem_prefix, b.n, (int)MF_ITERATE | (b.hh_split ? (int)MF_HH2 : 0));
      break;
    default:
      break;
  }
}

void launch_iteration(const BatchDev& b, int main_flags, hipStream_t s, const KernelTimer* timer,
                      hipEvent_t tail_stop, int fin_prev) {
  int order[5];
  const int n = pass_kernels(b, main_flags, order);
  hipEvent_t stop[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < n && timer; ++i)
    if (order[i] == timer->which) {
      stop[i] = timer->after;
      if (i > 0) stop[i - 1] = timer->before;
    }
  if (tail_stop) stop[n - 1] = tail_stop;
  for (int i = 0; i < n; ++i) {
    switch (order[i]) {
      case 0: launch_slice(b, s, stop[i], fin_prev); break;
      case 1: launch_contract(b, s, stop[i]); break;
      case 3:
        if (b.main_nq == 16)
          launch_k(k_main<16>, stop[i], dim3(b.elem_blocks), dim3(kElemThreads), 0, s,
                   (const SolveDev*)b.solves, (const int*)b.elem_prefix, b.n,
                   main_flags | (b.hh_split ? (int)MF_HH2 : 0));
        else
          launch_k(k_main<32>, stop[i], dim3(b.elem_blocks), dim3(kElemThreads), 0, s,
                   (const SolveDev*)b.solves, (const int*)b.elem_prefix, b.n,
                   main_flags | (b.hh_split ? (int)MF_HH2 : 0));
        break;
    }
  }
}

}  // namespace uegccd
