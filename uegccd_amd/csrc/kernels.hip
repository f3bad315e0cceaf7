// CDNA4 (gfx950) kernels of the CCD amplitude iteration.
//
// One residual pass (TermsWrapper, source/CCD.f90:146-181) plus the update
// (SolveCC/GetRes/CnvrgCC/CCEnergy, source/CCD.f90:186-462) is five launches:
//   k_mosaic_j   Jvv(c), Joo(k) of GetG2Mosaics (source/CCD.f90:548-567)
//   k_ring_j     quadratic ring / crossed-ring intermediates
//                (source/CCD.f90:771-797, 899-925)
//   k_ladder     both ladders as pair-momentum-blocked GEMMs on MFMA
//                (source/CCD.f90:611-681)
//   k_main       drivers + mosaics + rings + crossed rings + ladder contraction,
//                then R = D*T - G, the damped update, dT, energy, DIIS dots
//   k_finalize   deterministic reductions, convergence, DIIS solve
//   k_extrap     next amplitudes (DIIS extrapolation or the plain update)
// Every kernel resolves its solve from a block-prefix table, so one launch
// covers any number of independent solves (twists, r_s values, bases).
#include "uegccd_internal.h"
#include "../../include/uegccd.h"

namespace uegccd {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int find_solve(const int* __restrict__ prefix, int n, int blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= blk) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int d2(int4 p, int4 q) {
  int a = p.x - q.x, b = p.y - q.y, c = p.z - q.z;
  return a * a + b * b + c * c;
}

// Block-wide reductions (256 threads = 4 wave64s), fixed order.
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

// Coulomb table and n-vectors staged in LDS for the element kernels.
struct Lds {
  int4 nv[kMaxNao];
  double vt[kMaxN2];
  double red[16];
};

__device__ __forceinline__ void stage(const SolveDev& S, Lds& L) {
  for (int k = threadIdx.x; k < S.nao; k += blockDim.x) L.nv[k] = S.nv4[k];
  for (int k = threadIdx.x; k <= S.n2max; k += blockDim.x) L.vt[k] = S.vtab[k];
  __syncthreads();
}

#define VPQ(p, q) (L.vt[d2(L.nv[p], L.nv[q])])

// --------------------------------------------------------------------------
// Initial state: E(T0) and max|T0| (source/CCD.f90:62-64).
__global__ __launch_bounds__(kElemThreads) void k_init(const SolveDev* __restrict__ solves,
                                                        const int* __restrict__ prefix, int n) {
  __shared__ Lds L;
  const int s = find_solve(prefix, n, blockIdx.x);
  const SolveDev& S = solves[s];
  const int lb = blockIdx.x - prefix[s];
  stage(S, L);
  const int o = S.o, o2 = o * o;
  const int x = lb * blockDim.x + threadIdx.x;
  double e = 0.0, tmax = 0.0, bad = 0.0;
  if (x < S.n3) {
    const int b = S.pidx[x];
    if (b >= 0) {
      const int i = x % o, j = (x / o) % o, a = o + x / o2;
      const double tab = S.tcur[x], tba = S.tcur[S.n3 + x], taa = tab + tba;
      e = 0.5 * (VPQ(a, i) * (taa + tab) - VPQ(a, j) * (taa + tba));
      tmax = fmax(fabs(tab), fabs(tba));
      if (!isfinite(tab) || !isfinite(tba)) bad = 1.0;
    }
  }
  double* P = S.partials + (size_t)lb * kPartialStride;
  e = block_sum(e, L.red);
  tmax = block_max(tmax, L.red);
  bad = block_sum(bad, L.red);
  if (threadIdx.x == 0) { P[0] = e; P[1] = tmax; P[2] = 0.0; P[3] = bad; }
}

// DrvMBPT amplitudes and E(2) (source/MP2.f90:56-74); writes tcur.
__global__ __launch_bounds__(kElemThreads) void k_mp2(const SolveDev* __restrict__ solves,
                                                       const int* __restrict__ prefix, int n) {
  __shared__ Lds L;
  const int s = find_solve(prefix, n, blockIdx.x);
  const SolveDev& S = solves[s];
  const int lb = blockIdx.x - prefix[s];
  stage(S, L);
  const int o = S.o, o2 = o * o;
  const int x = lb * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (x < S.n3) {
    const int b = S.pidx[x];
    double tab = 0.0, tba = 0.0;
    if (b >= 0) {
      const int i = x % o, j = (x / o) % o, a = o + x / o2;
      const double vab = VPQ(i, a), vba = VPQ(i, b);
      const double den = S.eig[i] + S.eig[j] - S.eig[a] - S.eig[b];
      tab = vab / den;
      tba = -vba / den;
      const double taa = tab + tba;
      e = 0.5 * taa * (vab - vba) + 0.5 * tab * vab - 0.5 * tba * vba;
    }
    S.tcur[x] = tab;
    S.tcur[S.n3 + x] = tba;
  }
  e = block_sum(e, L.red);
  if (threadIdx.x == 0) {
    double* P = S.partials + (size_t)lb * kPartialStride;
    P[0] = e; P[1] = 0.0; P[2] = 0.0; P[3] = 0.0;
  }
}

// --------------------------------------------------------------------------
// Mosaic intermediates (source/CCD.f90:548-567): blocks [0,v) give Jvv(c),
// blocks [v, v+o) give Joo(k).
__global__ __launch_bounds__(kElemThreads) void k_mosaic_j(const SolveDev* __restrict__ solves,
                                                            const int* __restrict__ prefix, int n) {
  __shared__ Lds L;
  const int s = find_solve(prefix, n, blockIdx.x);
  const SolveDev& S = solves[s];
  if (S.st->phase == PH_DONE || !S.do_mosaic) return;
  const int lb = blockIdx.x - prefix[s];
  stage(S, L);
  const int o = S.o, o2 = o * o, v = S.v;
  const double* tab = S.tcur;
  const double* tba = S.tcur + S.n3;
  double acc = 0.0;
  if (lb < v) {
    const int cv = lb, c = o + cv;
    for (int r = threadIdx.x; r < o2; r += blockDim.x) {
      const int x = r + o2 * cv;
      if (S.pidx[x] < 0) continue;
      const int k = r % o, l = r / o;
      const double ta = tab[x], tb = tba[x], taa = ta + tb;
      acc += VPQ(c, k) * (taa + ta) - VPQ(c, l) * (taa + tb);
    }
  } else {
    const int k = lb - v;
    for (int r = threadIdx.x; r < o * v; r += blockDim.x) {
      const int l = r % o, cv = r / o, c = o + cv;
      const int x = k + o * l + o2 * cv;
      if (S.pidx[x] < 0) continue;
      const double ta = tab[x], tb = tba[x], taa = ta + tb;
      acc += VPQ(c, k) * (taa + ta) - VPQ(c, l) * (taa + tb);
    }
  }
  acc = block_sum(acc, L.red);
  if (threadIdx.x == 0) S.jvo[lb] = acc;
}

// Quadratic ring (J1,J2,J3 of source/CCD.f90:771-797) and crossed-ring
// (source/CCD.f90:899-925) intermediates, one thread per (i, a|b, l).
// Layout: jr[m*ovo + (av*o + i)*o + l], m = 0..2 rings, 3..5 crossed rings.
__global__ __launch_bounds__(kElemThreads) void k_ring_j(const SolveDev* __restrict__ solves,
                                                          const int* __restrict__ prefix, int n) {
  __shared__ Lds L;
  const int s = find_solve(prefix, n, blockIdx.x);
  const SolveDev& S = solves[s];
  if (S.st->phase == PH_DONE) return;
  const int lb = blockIdx.x - prefix[s];
  stage(S, L);
  const int o = S.o, o2 = o * o, v = S.v;
  const int ovo = o * v * o;
  const int t = lb * blockDim.x + threadIdx.x;
  if (t >= 2 * ovo) return;
  const int half = t / ovo, r = t - half * ovo;
  if (half == 0 && !S.do_ring) return;
  if (half == 1 && !S.do_xring) return;
  const int l = r % o, i = (r / o) % o, av = r / o2;
  const double* tab = S.tcur;
  const double* tba = S.tcur + S.n3;
  double j1 = 0.0, j2 = 0.0, j3 = 0.0;
  if (half == 0) {
    for (int k = 0; k < o; ++k) {
      const int y = i + o * k + o2 * av;
      const int c = S.pidx[y];
      if (c < 0) continue;
      const int d = S.pidx[k + o * l + o2 * (c - o)];
      if (d < 0) continue;
      const double vkl = VPQ(c, k), vlk = VPQ(c, l);
      const double ta = tab[y], tb = tba[y], taa = ta + tb;
      j1 += taa * (vkl - vlk) + ta * vkl;
      j2 += ta * (vkl - vlk) + taa * vkl;
      j3 -= tb * vlk;
    }
  } else {
    for (int k = 0; k < o; ++k) {
      const int c = S.pidx[i + o * k + o2 * av];
      if (c < 0) continue;
      const int d = S.pidx[k + o * l + o2 * (c - o)];
      if (d < 0) continue;
      const double vkl = VPQ(c, k), vlk = VPQ(c, l);
      const int y = k + o * i + o2 * av;
      const double ta = tab[y], tb = tba[y], taa = ta + tb;
      j1 += taa * (vkl - vlk) + tb * vkl;
      j2 += tb * (vkl - vlk) + taa * vkl;
      j3 += ta * vlk;
    }
  }
  double* J = S.jr + (size_t)(3 * half) * ovo + r;
  J[0] = j1;
  J[ovo] = j2;
  J[2 * ovo] = j3;
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
// so one tile kernel serves both: rows = pairs of K (16 per tile), columns =
// the block's a list followed by its pair list.  The B operand is never
// stored: each lane synthesises its element from two n-vectors and the
// Coulomb table.  One wave64 per 16x16 tile, v_mfma_f64_16x16x4f64 (or the
// f32 form while the mixed-precision schedule is in its FP32 phase).
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

template <typename Real>
__device__ void ladder_tile(const SolveDev& S, int K, int row0, int col0) {
  typedef typename MfmaTraits<Real>::acc_t acc_t;
  const int o = S.o, o2 = o * o;
  const int poff = S.kb_poff[K], noo = S.kb_poff[K + 1] - poff;
  const int voff = S.kb_voff[K], nvv = S.kb_voff[K + 1] - voff;
  const int ncol = nvv + noo;
  const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
  const Real* vt = nullptr;
  const double* vt64 = S.vtab;
  const float* vt32 = S.vtab32;
  (void)vt;
  const double* tab = S.tcur;
  const double* tba = S.tcur + S.n3;

  const int rowA = row0 + r16;
  const int xpA = rowA < noo ? S.kb_pairs[poff + rowA] : -1;
  const int col = col0 + r16;
  const bool colv = col < ncol;
  int cx = 0, cy = 0;
  if (colv) {
    if (col < nvv) {
      cx = S.kb_va[voff + col];
      cy = S.kb_vb[voff + col];
    } else {
      const int xq = S.kb_pairs[poff + col - nvv];
      cx = xq % o;
      cy = xq / o;
    }
  }
  const int4 nx = S.nv4[cx], ny = S.nv4[cy];
  acc_t accs = {0, 0, 0, 0}, accp = {0, 0, 0, 0};
  for (int c0 = 0; c0 < nvv; c0 += 4) {
    const int ci = c0 + kq;
    Real sv = 0, pv = 0, bs = 0, bp = 0;
    if (ci < nvv) {
      const int c = S.kb_va[voff + ci];
      const int4 nc = S.nv4[c];
      if (xpA >= 0) {
        const size_t y = (size_t)xpA + (size_t)o2 * (c - o);
        const double ta = tab[y], tb = tba[y];
        sv = (Real)(ta + tb);
        pv = (Real)(ta - tb);
      }
      if (colv) {
        Real vx, vy;
        if (sizeof(Real) == 8) {
          vx = (Real)vt64[d2(nc, nx)];
          vy = (Real)vt64[d2(nc, ny)];
        } else {
          vx = (Real)vt32[d2(nc, nx)];
          vy = (Real)vt32[d2(nc, ny)];
        }
        bs = vx - vy;
        bp = vx + vy;
      }
    }
    accs = MfmaTraits<Real>::mma(sv, bs, accs);
    accp = MfmaTraits<Real>::mma(pv, bp, accp);
  }
  if (!colv) return;
  for (int r = 0; r < 4; ++r) {
    const int rowo = row0 + MfmaTraits<Real>::row(kq, r);
    if (rowo >= noo) continue;
    const int xpo = S.kb_pairs[poff + rowo];
    const double as = (double)accs[r], ap = (double)accp[r];
    if (col < nvv) {
      const size_t y = (size_t)xpo + (size_t)o2 * (cx - o);
      S.lad[y] = 0.5 * as;
      S.lad[S.n3 + y] = 0.5 * ap;
    } else {
      const int q = col - nvv;
      const int i = xpo % o;
      const double vik = vt64[d2(S.nv4[i], nx)], vil = vt64[d2(S.nv4[i], ny)];
      const int wi = S.kb_woff[K] + rowo * noo + q;
      S.w[wi] = vik - vil + 0.5 * as;
      S.w[S.wtot + wi] = vik + vil + 0.5 * ap;
    }
  }
}

__global__ __launch_bounds__(64) void k_ladder(const SolveDev* __restrict__ solves,
                                               const int4* __restrict__ items, int nitems) {
  const int4 it = items[blockIdx.x];
  const SolveDev& S = solves[it.x];
  if (S.st->phase == PH_DONE || !S.do_ladder) return;
  if (S.st->prec64)
    ladder_tile<double>(S, it.y, it.z, it.w);
  else
    ladder_tile<float>(S, it.y, it.z, it.w);
}

// --------------------------------------------------------------------------
// Residual assembly + update, one thread per dense (i,j,a).
__global__ __launch_bounds__(kElemThreads) void k_main(const SolveDev* __restrict__ solves,
                                                        const int* __restrict__ prefix, int n,
                                                        int flags) {
  __shared__ Lds L;
  const int s = find_solve(prefix, n, blockIdx.x);
  const SolveDev& S = solves[s];
  if (S.st->phase == PH_DONE) return;
  const int lb = blockIdx.x - prefix[s];
  stage(S, L);
  const int o = S.o, o2 = o * o, n3 = S.n3;
  const int x = lb * blockDim.x + threadIdx.x;
  const int nh = (flags & MF_ITERATE) ? S.ndiis : 0;
  const int slot = S.st->slot;
  const double* tab = S.tcur;
  const double* tba = S.tcur + n3;

  double e_part = 0.0, dt_part = 0.0, r_part = 0.0, bad = 0.0;
  double eab = 0.0, eba = 0.0;
  bool valid = false;
  if (x < n3) {
    const int b = S.pidx[x];
    if (b >= 0) {
      valid = true;
      const int i = x % o, j = (x / o) % o, av = x / o2, a = o + av, bv = b - o;
      const double t_ab = tab[x], t_ba = tba[x];
      // drivers (source/CCD.f90:484-501)
      double gab = VPQ(i, a);
      double gba = -VPQ(i, b);
      // mosaics (source/CCD.f90:571-582)
      if (S.do_mosaic) {
        const double* jvv = S.jvo;
        const double* joo = S.jvo + S.v;
        gab = gab - 0.5 * (jvv[av] * t_ab + jvv[bv] * t_ab);
        gba = gba - 0.5 * (jvv[av] * t_ba + jvv[bv] * t_ba);
        gab = gab - 0.5 * (joo[i] * t_ab + joo[j] * t_ab);
        gba = gba - 0.5 * (joo[i] * t_ba + joo[j] * t_ba);
      }
      // ladders: pp part from k_ladder, hh contraction with W (source/CCD.f90:651-656)
      if (S.do_ladder) {
        const int xp = i + o * j;
        const int K = S.pair_kb[xp], pos = S.pair_pos[xp];
        const int poff = S.kb_poff[K], noo = S.kb_poff[K + 1] - poff;
        const double* ws = S.w + S.kb_woff[K] + pos * noo;
        const double* wp = ws + S.wtot;
        double hs = 0.0, hp = 0.0;
        for (int q = 0; q < noo; ++q) {
          const int y = S.kb_pairs[poff + q] + o2 * av;
          const double ta = tab[y], tb = tba[y];
          hs += ws[q] * (ta + tb);
          hp += wp[q] * (ta - tb);
        }
        const double ls = S.lad[x] + 0.5 * hs;
        const double lp = S.lad[n3 + x] + 0.5 * hp;
        gab += 0.5 * (ls + lp);
        gba += 0.5 * (ls - lp);
      }
      const int ovo = o * S.v * o;
      // rings (source/CCD.f90:716-811)
      if (S.do_ring) {
        double v1 = VPQ(i, a);
        for (int k = 0; k < o; ++k) {
          const int y = i + o * k + o2 * av;
          if (S.pidx[y] < 0) continue;
          const double v2 = VPQ(k, j);
          const double ta = tab[y], tb = tba[y];
          gab += (v1 - v2) * ta + v1 * (ta + tb);
          gba -= v2 * tb;
        }
        v1 = VPQ(j, b);
        for (int k = 0; k < o; ++k) {
          const int y = j + o * k + o2 * bv;
          if (S.pidx[y] < 0) continue;
          const double v2 = VPQ(k, i);
          const double ta = tab[y], tb = tba[y];
          gab += (v1 - v2) * ta + v1 * (ta + tb);
          gba -= v2 * tb;
        }
        const double* J = S.jr + (size_t)(av * o + i) * o;
        for (int l = 0; l < o; ++l) {
          const int d = S.pidx[j + o * l + o2 * bv];
          if (d < 0) continue;
          const int y = l + o * j + o2 * (d - o);
          const double ta = tab[y], tb = tba[y];
          gab += J[ovo + l] * (ta + tb) + J[l] * ta;
          gba += J[2 * ovo + l] * tb;
        }
      }
      // crossed rings (source/CCD.f90:845-937)
      if (S.do_xring) {
        const double vja = VPQ(j, a);
        for (int k = 0; k < o; ++k) {
          const int y = k + o * i + o2 * bv;
          if (S.pidx[y] < 0) continue;
          const double vjk = VPQ(j, k);
          const double ta = tab[y], tb = tba[y];
          gab -= vjk * ta;
          gba += (vja - vjk) * tb + vja * (ta + tb);
        }
        const double vib = VPQ(i, b);
        for (int k = 0; k < o; ++k) {
          const int y = k + o * j + o2 * av;
          if (S.pidx[y] < 0) continue;
          const double vik = VPQ(i, k);
          const double ta = tab[y], tb = tba[y];
          gab -= vik * ta;
          gba += (vib - vik) * tb + vib * (ta + tb);
        }
        const double* J = S.jr + (size_t)3 * ovo + (size_t)(bv * o + i) * o;
        for (int l = 0; l < o; ++l) {
          const int y = l + o * j + o2 * av;
          if (S.pidx[y] < 0) continue;
          const double ta = tab[y], tb = tba[y];
          gba = gba - J[l] * tb - J[ovo + l] * (ta + tb);
          gab += J[2 * ovo + l] * ta;
        }
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
      e_part = 0.5 * (VPQ(a, i) * (naa + nab) - VPQ(a, j) * (naa + nba));
      if (!isfinite(nab) || !isfinite(nba) || !isfinite(rab) || !isfinite(rba)) bad = 1.0;
      if (flags & MF_WRITE_GR) {
        if (S.gout) { S.gout[x] = gab; S.gout[n3 + x] = gba; }
        if (S.rout) { S.rout[x] = rab; S.rout[n3 + x] = rba; }
      }
      if (flags & MF_ITERATE) {
        double* HT = S.histT + (size_t)slot * 2 * n3;
        HT[x] = nab;
        HT[n3 + x] = nba;
        if (S.diis_err_is_dt) { eab = nab - t_ab; eba = nba - t_ba; }
        else { eab = rab; eba = rba; }
        if (S.mode == UEGCCD_MODE_DIIS) {
          double* HE = S.histE + (size_t)slot * 2 * n3;
          HE[x] = eab;
          HE[n3 + x] = eba;
        }
      }
    }
  }
  double* P = S.partials + (size_t)lb * kPartialStride;
  e_part = block_sum(e_part, L.red);
  dt_part = block_max(dt_part, L.red);
  r_part = block_max(r_part, L.red);
  bad = block_sum(bad, L.red);
  if (threadIdx.x == 0) { P[0] = e_part; P[1] = dt_part; P[2] = r_part; P[3] = bad; }
  if (S.mode == UEGCCD_MODE_DIIS && nh > 0) {
    for (int k = 0; k < nh; ++k) {
      double dk = 0.0;
      if (valid) {
        if (k == slot) {
          dk = eab * eab + eba * eba;
        } else {
          const double* HE = S.histE + (size_t)k * 2 * n3;
          dk = eab * HE[x] + eba * HE[n3 + x];
        }
      }
      dk = block_sum(dk, L.red);
      if (threadIdx.x == 0) P[4 + k] = dk;
    }
  }
}

// --------------------------------------------------------------------------
// Small dense solve for the DIIS coefficients (FP64, partial pivoting).
__device__ bool diis_coefficients(const State& st, int ndiis, double* coef) {
  const int nh = st.nhist;
  int sl[kMaxDiis];
  for (int r = 0; r < nh; ++r) sl[r] = ((st.slot_cur - (nh - 1 - r)) % ndiis + ndiis) % ndiis;
  const int m = nh + 1;
  double A[(kMaxDiis + 1) * (kMaxDiis + 1)];
  double rhs[kMaxDiis + 1];
  double scale = 0.0;
  for (int r = 0; r < nh; ++r) scale = fmax(scale, fabs(st.bmat[sl[r] * kMaxDiis + sl[r]]));
  if (!(scale > 0.0) || !isfinite(scale)) return false;
  for (int r = 0; r < nh; ++r) {
    for (int c = 0; c < nh; ++c) A[r * m + c] = st.bmat[sl[r] * kMaxDiis + sl[c]] / scale;
    A[r * m + nh] = -1.0;
    A[nh * m + r] = -1.0;
    rhs[r] = 0.0;
  }
  A[nh * m + nh] = 0.0;
  rhs[nh] = -1.0;
  for (int c = 0; c < m; ++c) {
    int p = c;
    for (int r = c + 1; r < m; ++r)
      if (fabs(A[r * m + c]) > fabs(A[p * m + c])) p = r;
    if (fabs(A[p * m + c]) < 1e-14) return false;
    if (p != c) {
      for (int k = 0; k < m; ++k) { double t = A[c * m + k]; A[c * m + k] = A[p * m + k]; A[p * m + k] = t; }
      double t = rhs[c]; rhs[c] = rhs[p]; rhs[p] = t;
    }
    for (int r = c + 1; r < m; ++r) {
      const double f = A[r * m + c] / A[c * m + c];
      for (int k = c; k < m; ++k) A[r * m + k] -= f * A[c * m + k];
      rhs[r] -= f * rhs[c];
    }
  }
  for (int r = m - 1; r >= 0; --r) {
    double acc = rhs[r];
    for (int k = r + 1; k < m; ++k) acc -= A[r * m + k] * rhs[k];
    rhs[r] = acc / A[r * m + r];
  }
  for (int k = 0; k < ndiis; ++k) coef[k] = 0.0;
  for (int r = 0; r < nh; ++r) {
    if (!isfinite(rhs[r])) return false;
    coef[sl[r]] = rhs[r];
  }
  return true;
}

// Reductions of the block partials and the iteration bookkeeping of
// source/CCD.f90:80-99 (one block per solve).
__global__ __launch_bounds__(256) void k_finalize(const SolveDev* __restrict__ solves, int kind) {
  __shared__ double red[16];
  __shared__ double vals[4 + kMaxDiis];
  const SolveDev& S = solves[blockIdx.x];
  State* st = S.st;
  if (kind == FIN_ITER && st->phase == PH_DONE) return;
  const int nb = S.nblk_elem;
  const int nd = (kind == FIN_ITER && S.mode == UEGCCD_MODE_DIIS && st->phase == PH_ITER) ? S.ndiis : 0;
  {
    double e = 0.0, dt = 0.0, r = 0.0, bad = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
      const double* P = S.partials + (size_t)k * kPartialStride;
      e += P[0];
      dt = fmax(dt, P[1]);
      r = fmax(r, P[2]);
      bad += P[3];
    }
    e = block_sum(e, red);
    dt = block_max(dt, red);
    r = block_max(r, red);
    bad = block_sum(bad, red);
    if (threadIdx.x == 0) { vals[0] = e; vals[1] = dt; vals[2] = r; vals[3] = bad; }
    for (int q = 0; q < nd; ++q) {
      double d = 0.0;
      for (int k = threadIdx.x; k < nb; k += blockDim.x) d += S.partials[(size_t)k * kPartialStride + 4 + q];
      d = block_sum(d, red);
      if (threadIdx.x == 0) vals[4 + q] = d;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double e = vals[0], dt = vals[1], rmax = vals[2];
  if (vals[3] > 0.0) st->nonfinite = 1;
  if (kind == FIN_SUM) { st->ecorr = e; return; }
  if (kind == FIN_INIT) {
    st->niter = 1;
    st->ecorr = e;
    st->ecorr_prev = e;
    st->dt = dt;
    st->slot = 0;
    st->slot_cur = 0;
    st->nhist = 0;
    st->use_coef = 0;
    st->converged = 0;
    if (S.trace && S.maxiter > 0) { S.trace[0] = e; S.trace[1] = dt; }
    const bool go = S.mode == UEGCCD_MODE_REFERENCE ? (dt >= S.tol_dt) : true;
    st->phase = (go && st->niter < S.maxiter && !st->nonfinite) ? PH_ITER : PH_FINAL;
    if (!go) st->converged = 1;
    return;
  }
  // FIN_ITER
  if (st->phase == PH_FINAL) {
    st->maxres = rmax;
    st->phase = PH_DONE;
    return;
  }
  st->niter += 1;
  st->ecorr_prev = st->ecorr;
  st->ecorr = e;
  st->dt = dt;
  st->maxres_in = rmax;
  if (S.trace && st->niter <= S.maxiter) {
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
  st->use_coef = 0;
  if (S.mode == UEGCCD_MODE_DIIS) {
    if (switched) {
      st->nhist = 0;  // residuals of the FP32 phase do not enter FP64 extrapolation
    } else {
      // B-matrix row/column of the new entry
      st->nhist = st->nhist < S.ndiis ? st->nhist + 1 : S.ndiis;
      for (int k = 0; k < S.ndiis; ++k) {
        st->bmat[slot * kMaxDiis + k] = vals[4 + k];
        st->bmat[k * kMaxDiis + slot] = vals[4 + k];
      }
      if (!stop && st->niter >= S.diis_start && st->nhist >= 2)
        st->use_coef = diis_coefficients(*st, S.ndiis, st->coef) ? 1 : 0;
    }
  }
  if (st->nonfinite) {
    st->phase = PH_DONE;
    return;
  }
  if (conv) st->converged = 1;
  if (stop) st->phase = PH_FINAL;
}

// Amplitudes for the next pass: DIIS extrapolation or the plain update.
__global__ __launch_bounds__(kElemThreads) void k_extrap(const SolveDev* __restrict__ solves,
                                                          const int* __restrict__ prefix, int n) {
  const int s = find_solve(prefix, n, blockIdx.x);
  const SolveDev& S = solves[s];
  const State* st = S.st;
  if (st->phase == PH_DONE) return;
  const int lb = blockIdx.x - prefix[s];
  const int x = lb * blockDim.x + threadIdx.x;
  const int n3 = S.n3;
  if (x >= n3 || S.pidx[x] < 0) return;
  if (st->use_coef) {
    double ab = 0.0, ba = 0.0;
    for (int k = 0; k < S.ndiis; ++k) {
      const double c = st->coef[k];
      if (c == 0.0) continue;
      const double* HT = S.histT + (size_t)k * 2 * n3;
      ab += c * HT[x];
      ba += c * HT[n3 + x];
    }
    S.tcur[x] = ab;
    S.tcur[n3 + x] = ba;
  } else {
    const double* HT = S.histT + (size_t)st->slot_cur * 2 * n3;
    S.tcur[x] = HT[x];
    S.tcur[n3 + x] = HT[n3 + x];
  }
}

__global__ void k_state_reset(const SolveDev* __restrict__ solves, int phase, int prec64) {
  const SolveDev& S = solves[blockIdx.x];
  if (threadIdx.x != 0) return;
  State* st = S.st;
  st->niter = 0;
  st->phase = phase;
  st->slot = 0;
  st->slot_cur = 0;
  st->nhist = 0;
  st->use_coef = 0;
  st->prec64 = prec64;
  st->iters_fp32 = 0;
  st->nonfinite = 0;
  st->converged = 0;
  st->ecorr = st->ecorr_prev = st->dt = st->maxres = st->maxres_in = 0.0;
}

// --------------------------------------------------------------------------
void launch_init(const BatchDev& b, hipStream_t s) {
  hipLaunchKernelGGL(k_init, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                     b.elem_prefix, b.n);
}
void launch_mp2(const BatchDev& b, hipStream_t s) {
  hipLaunchKernelGGL(k_mp2, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                     b.elem_prefix, b.n);
}
void launch_finalize(const BatchDev& b, int kind, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(b.n), dim3(256), 0, s, b.solves, kind);
}
void launch_state_reset(const BatchDev& b, int phase, int prec64, hipStream_t s) {
  hipLaunchKernelGGL(k_state_reset, dim3(b.n), dim3(64), 0, s, b.solves, phase, prec64);
}
void launch_one(const BatchDev& b, int which, hipStream_t s) {
  switch (which) {
    case 0:
      if (b.mos_blocks > 0)
        hipLaunchKernelGGL(k_mosaic_j, dim3(b.mos_blocks), dim3(kElemThreads), 0, s, b.solves,
                           b.mos_prefix, b.n);
      break;
    case 1:
      if (b.ring_blocks > 0)
        hipLaunchKernelGGL(k_ring_j, dim3(b.ring_blocks), dim3(kElemThreads), 0, s, b.solves,
                           b.ring_prefix, b.n);
      break;
    case 2:
      if (b.nlad > 0)
        hipLaunchKernelGGL(k_ladder, dim3(b.nlad), dim3(64), 0, s, b.solves, b.lad_items, b.nlad);
      break;
    case 3:
      hipLaunchKernelGGL(k_main, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                         b.elem_prefix, b.n, (int)MF_ITERATE);
      break;
    case 4:
      hipLaunchKernelGGL(k_extrap, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                         b.elem_prefix, b.n);
      break;
    default:
      break;
  }
}

void launch_iteration(const BatchDev& b, int main_flags, hipStream_t s) {
  if (b.mos_blocks > 0)
    hipLaunchKernelGGL(k_mosaic_j, dim3(b.mos_blocks), dim3(kElemThreads), 0, s, b.solves,
                       b.mos_prefix, b.n);
  if (b.ring_blocks > 0)
    hipLaunchKernelGGL(k_ring_j, dim3(b.ring_blocks), dim3(kElemThreads), 0, s, b.solves,
                       b.ring_prefix, b.n);
  if (b.nlad > 0)
    hipLaunchKernelGGL(k_ladder, dim3(b.nlad), dim3(64), 0, s, b.solves, b.lad_items, b.nlad);
  hipLaunchKernelGGL(k_main, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                     b.elem_prefix, b.n, main_flags);
  if (main_flags & MF_ITERATE) {
    hipLaunchKernelGGL(k_finalize, dim3(b.n), dim3(256), 0, s, b.solves, (int)FIN_ITER);
    hipLaunchKernelGGL(k_extrap, dim3(b.elem_blocks), dim3(kElemThreads), 0, s, b.solves,
                       b.elem_prefix, b.n);
  }
}

}  // namespace uegccd
