// CDNA4 (gfx950) kernels of the adaptive FAS V-cycle.
//
// One launch sweeps every patch of a level: work is described by per-level
// tables (256-cell tiles, ghost records, ring gathers, footprint cells,
// reflux targets) whose indices are absolute positions in the level's packed
// buffers.  All kernels are templated on the storage type T: double for the
// plain path (built with -ffp-contract=off and evaluated in the reference's
// operation order, so FP64 results equal the CPU oracle bit for bit) and float
// for the inner V-cycles of the mixed-precision path.  Reflux sums, the
// composite residual and the coarsest solve always accumulate in FP64.
//
// Formulas follow:
//   relax            EllipticBSAM/src/ebKernels.m:37-95
//   ghost records    EllipticBSAM/src/ebFillGhosts.m:18-92, ebInterpCf.m:41-153
//   restrict         EllipticBSAM/src/ebTransfer.m:10-33, ebKernels.m:98-102
//   coarse load      EllipticBSAM/src/ebCoarseLoad.m:17-32, ebKernels.m:13-25
//   reflux           EllipticBSAM/src/ebFaceCorrections.m:107-133, ebCoarseLoad.m:37-45
//   correct          EllipticBSAM/src/ebTransfer.m:57-71, ebKernels.m:105-123
//   coarsest         EllipticBSAM/src/ebCoarsest.m:79-91
//   residual         EllipticBSAM/src/ebCompResidual.m:302-332
//   composite mean   EllipticBSAM/src/ebCompositeMean.m:4-23
#include "bsam_kernels.h"

namespace bsam {

#define A1_ (-1.0 / 5.0)
#define A2_ (2.0 / 3.0)
#define A3_ (8.0 / 15.0)
#define B1_ (5.0 / 32.0)
#define B2_ (15.0 / 16.0)
#define B3_ (-3.0 / 32.0)
#define C83_ (8.0 / 3.0)
#define C13_ (1.0 / 3.0)
#define C23_ (2.0 / 3.0)

// ---------------------------------------------------------------- smoother
template <typename T>
__global__ void __launch_bounds__(kTile) k_relax(const PatchInfo* __restrict__ P,
                                                 const int2* __restrict__ tiles, T* __restrict__ Q,
                                                 const T* __restrict__ F, const T* __restrict__ ID,
                                                 const T* __restrict__ DeW, const T* __restrict__ DnS,
                                                 const T* __restrict__ G, T hx, T hy, int colour,
                                                 T omega, int homog) {
  const int2 t = tiles[blockIdx.x];
  const PatchInfo& pi = P[t.x];
  const int mx = pi.mx, my = pi.my;
  const int c = t.y + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = c / mx;
  const int i = c - j * mx;
  // (li + lj + par0) mod 2 == colour-1 with 1-based local indices
  if (((i + j + pi.par0) & 1) != colour - 1) return;
  const int ldq = mx + 2;
  const int qc = pi.qoff + (i + 1) + (j + 1) * ldq;
  T qe = Q[qc + 1], qw = Q[qc - 1], qn = Q[qc + ldq], qs = Q[qc - ldq];
  // partial physical ghosts: Dirichlet (8/3)g + (1/3)u_2, Neumann h g
  if (i == 0 && pi.bc[0]) {
    const T g = homog ? T(0) : G[pi.goff + j];
    qw = pi.bc[0] == 1 ? T(C83_) * g + T(C13_) * Q[qc + 1] : hx * g;
  }
  if (i == mx - 1 && pi.bc[1]) {
    const T g = homog ? T(0) : G[pi.goff + my + j];
    qe = pi.bc[1] == 1 ? T(C83_) * g + T(C13_) * Q[qc - 1] : hx * g;
  }
  if (j == 0 && pi.bc[2]) {
    const T g = homog ? T(0) : G[pi.goff + 2 * my + i];
    qs = pi.bc[2] == 1 ? T(C83_) * g + T(C13_) * Q[qc + ldq] : hy * g;
  }
  if (j == my - 1 && pi.bc[3]) {
    const T g = homog ? T(0) : G[pi.goff + 2 * my + mx + i];
    qn = pi.bc[3] == 1 ? T(C83_) * g + T(C13_) * Q[qc - ldq] : hy * g;
  }
  const int cc = pi.coff + c;
  const int fx = pi.xoff + i + j * (mx + 1);
  const int fy = pi.yoff + i + j * mx;
  const T hx2 = hx * hx, hy2 = hy * hy;
  T tmp = F[cc] + qe * DeW[fx + 1] / hx2;
  tmp = tmp + qw * DeW[fx] / hx2;
  tmp = tmp + qn * DnS[fy + mx] / hy2;
  tmp = tmp + qs * DnS[fy] / hy2;
  tmp = tmp * ID[cc];
  if (omega == T(1)) {
    Q[qc] = tmp;
  } else {
    Q[qc] = omega * tmp + (T(1) - omega) * Q[qc];
  }
}

// ---------------------------------------------------------------- ghosts
template <typename T>
__device__ __forceinline__ T ghost_value(const GhostRec& r, const T* __restrict__ Q,
                                         const T* __restrict__ Pq, int homog) {
  switch (r.kind) {
    case G_COPY:
      return Q[r.s[0]];
    case G_QIF:
      return (T(A1_) * Q[r.s[0]] + T(A2_) * Q[r.s[1]]) + T(A3_) * Pq[r.s[2]];
    case G_QIF_END:
      return (T(A1_) * Q[r.s[0]] + T(A2_) * Q[r.s[1]]) +
             T(A3_) * ((T(B1_) * Pq[r.s[2]] + T(B2_) * Pq[r.s[3]]) + T(B3_) * Pq[r.s[4]]);
    case G_LIF:
      return (T(C23_) * Pq[r.s[2]] + T(C23_) * Q[r.s[0]]) - T(C13_) * Q[r.s[1]];
    case G_LIF_CORNER:
      return T(C23_) * Pq[r.s[2]] + T(C13_) * Q[r.s[0]];
    case G_DIR: {
      const T g = homog ? T(0) : T(r.c);
      return (T(C83_) * g - T(2) * Q[r.s[0]]) + T(C13_) * Q[r.s[1]];
    }
    case G_NEU: {
      const T hg = homog ? T(0) : T(r.c);
      return Q[r.s[0]] + hg;
    }
    default:
      return T(0);
  }
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_fill(const GhostRec* __restrict__ recs, int nrec,
                                                const RingRec* __restrict__ ring, int nring,
                                                T* __restrict__ Q, const T* __restrict__ Pq,
                                                T* __restrict__ QC, int homog) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < nrec) {
    const GhostRec r = recs[t];
    T v;
    if (r.kind == G_PCORNER) {
      // physical corner from the two adjacent side ghosts (ebFillGhosts.m:85-89)
      const T a = ghost_value<T>(recs[r.s[0]], Q, Pq, homog);
      const T b = ghost_value<T>(recs[r.s[1]], Q, Pq, homog);
      v = (a + b) - Q[r.s[2]];
    } else {
      v = ghost_value<T>(r, Q, Pq, homog);
    }
    Q[r.dst] = v;
  } else if (t < nrec + nring) {
    const RingRec rr = ring[t - nrec];
    QC[rr.dst] = Pq[rr.src];
  }
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_ring(const RingRec* __restrict__ ring, int nring,
                                                const T* __restrict__ Pq, T* __restrict__ QC) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t < nring) {
    const RingRec rr = ring[t];
    QC[rr.dst] = Pq[rr.src];
  }
}

// ---------------------------------------------------------------- transfers
template <typename T>
__global__ void __launch_bounds__(kTile) k_restrict(const PatchInfo* __restrict__ P,
                                                    const FootRec* __restrict__ foot, int nfoot,
                                                    const T* __restrict__ Q, T* __restrict__ QC,
                                                    T* __restrict__ Pq, int qc_only) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t >= nfoot) return;
  const FootRec fr = foot[t];
  const PatchInfo& pi = P[fr.p];
  const int ldq = pi.mx + 2;
  const int q00 = pi.qoff + (2 * fr.ci + 1) + (2 * fr.cj + 1) * ldq;
  const T R = T(0.25) * (((Q[q00] + Q[q00 + 1]) + Q[q00 + ldq]) + Q[q00 + ldq + 1]);
  QC[pi.qcoff + (fr.ci + 1) + (fr.cj + 1) * (pi.mx / 2 + 2)] = R;
  if (!qc_only) Pq[fr.pq] = R;
}

// L(u) at one cell of a padded array (ebKernels.m:13-25)
template <typename T>
__device__ __forceinline__ T op_at(const T* __restrict__ Q, int qc, int ldq, T dw, T de, T ds, T dn,
                                   T cc, T hx2, T hy2) {
  const T u = Q[qc];
  const T fw = dw * (u - Q[qc - 1]);
  const T fe = de * (Q[qc + 1] - u);
  const T fs = ds * (u - Q[qc - ldq]);
  const T fn = dn * (Q[qc + ldq] - u);
  return (-(fe - fw) / hx2 - (fn - fs) / hy2) + cc * u;
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_coarse_load(
    const PatchInfo* __restrict__ P, const FootRec* __restrict__ foot, int nfoot,
    const T* __restrict__ Q, const T* __restrict__ F, const T* __restrict__ DeW,
    const T* __restrict__ DnS, const T* __restrict__ CC, const T* __restrict__ QC, T hx, T hy,
    const T* __restrict__ pDeW, const T* __restrict__ pDnS, const T* __restrict__ pCC,
    T* __restrict__ pF, T hcx, T hcy) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t >= nfoot) return;
  const FootRec fr = foot[t];
  const PatchInfo& pi = P[fr.p];
  const int mx = pi.mx;
  const int ldq = mx + 2;
  const T hx2 = hx * hx, hy2 = hy * hy;
  T rf[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int i = 2 * fr.ci + (s & 1);
    const int j = 2 * fr.cj + (s >> 1);
    const int qc = pi.qoff + (i + 1) + (j + 1) * ldq;
    const int cc = pi.coff + i + j * mx;
    const int fx = pi.xoff + i + j * (mx + 1);
    const int fy = pi.yoff + i + j * mx;
    rf[s] = F[cc] - op_at<T>(Q, qc, ldq, DeW[fx], DeW[fx + 1], DnS[fy], DnS[fy + mx], CC[cc], hx2, hy2);
  }
  const T R = T(0.25) * (((rf[0] + rf[1]) + rf[2]) + rf[3]);
  const int ldc = mx / 2 + 2;
  const int qcc = pi.qcoff + (fr.ci + 1) + (fr.cj + 1) * ldc;
  const T Lc = op_at<T>(QC, qcc, ldc, pDeW[fr.dew_w], pDeW[fr.dew_e], pDnS[fr.dns_s], pDnS[fr.dns_n],
                        pCC[fr.pc], hcx * hcx, hcy * hcy);
  pF[fr.pc] = R + Lc;
}

// one flux-balance contribution (ebFaceCorrections.m:111-133), in FP64
template <typename T>
__device__ __forceinline__ double corr_value(const CorrRec& e, const T* __restrict__ Q,
                                             const T* __restrict__ QC, const T* __restrict__ DeW,
                                             const T* __restrict__ DnS, double hcx2, double hcy2) {
  const T* D = e.face < 2 ? DeW : DnS;
  const double h2 = e.face < 2 ? hcx2 : hcy2;
  const double t1 = (double)D[e.d1] * ((double)Q[e.in1] - (double)Q[e.g1]);
  const double t2 = (double)D[e.d2] * ((double)Q[e.in2] - (double)Q[e.g2]);
  const double t3 = e.cf * ((double)QC[e.ring] - (double)QC[e.qcin]);
  return ((t1 + t2) + t3) / h2;
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_reflux(const int* __restrict__ tgt_cell,
                                                  const int* __restrict__ tgt_ptr,
                                                  const int* __restrict__ tgt_ent, int ntgt,
                                                  const CorrRec* __restrict__ corr,
                                                  const T* __restrict__ Q, const T* __restrict__ QC,
                                                  const T* __restrict__ DeW, const T* __restrict__ DnS,
                                                  double hcx2, double hcy2, T* __restrict__ pF) {
  const int t = blockIdx.x * kTile + threadIdx.x;
  if (t >= ntgt) return;
  const int cell = tgt_cell[t];
  double acc = (double)pF[cell];
  for (int e = tgt_ptr[t]; e < tgt_ptr[t + 1]; ++e)
    acc = acc + corr_value<T>(corr[tgt_ent[e]], Q, QC, DeW, DnS, hcx2, hcy2);
  pF[cell] = (T)acc;
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_correct(const PatchInfo* __restrict__ P,
                                                   const int2* __restrict__ tiles,
                                                   T* __restrict__ Q, const T* __restrict__ QC,
                                                   const int* __restrict__ wmap,
                                                   const T* __restrict__ Pq, int init) {
  const int2 t = tiles[blockIdx.x];
  const PatchInfo& pi = P[t.x];
  const int mx = pi.mx, my = pi.my;
  const int c = t.y + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = c / mx;
  const int i = c - j * mx;
  const int ldc = mx / 2 + 2;
  const int w0 = pi.qcoff + ((i >> 1) + 1) + ((j >> 1) + 1) * ldc;
  const int dx = (i & 1) ? 1 : -1;
  const int dy = (j & 1) ? ldc : -ldc;
  T e_c, e_x, e_y, e_d;
  if (init) {
    e_c = Pq[wmap[w0]];
    e_x = Pq[wmap[w0 + dx]];
    e_y = Pq[wmap[w0 + dy]];
    e_d = Pq[wmap[w0 + dx + dy]];
  } else {
    e_c = Pq[wmap[w0]] - QC[w0];
    e_x = Pq[wmap[w0 + dx]] - QC[w0 + dx];
    e_y = Pq[wmap[w0 + dy]] - QC[w0 + dy];
    e_d = Pq[wmap[w0 + dx + dy]] - QC[w0 + dx + dy];
  }
  const T p = (((T(9) * e_c + T(3) * e_x) + T(3) * e_y) + e_d) / T(16);
  const int qc = pi.qoff + (i + 1) + (j + 1) * (mx + 2);
  Q[qc] = init ? p : Q[qc] + p;
}

// ---------------------------------------------------------------- coarsest
template <typename T>
__global__ void __launch_bounds__(64) k_coarsest(const double* __restrict__ Ainv,
                                                 const double* __restrict__ bcRhs, int N, int M,
                                                 int n1, const T* __restrict__ F, T* __restrict__ Q,
                                                 int homog) {
  extern __shared__ double rhs[];
  for (int j = threadIdx.x; j < M; j += blockDim.x)
    rhs[j] = j < N ? (homog ? (double)F[j] : (double)F[j] + bcRhs[j]) : 0.0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double acc = 0.0;
  for (int j = 0; j < M; ++j) acc = acc + Ainv[i + (size_t)j * M] * rhs[j];
  const int ci = i % n1, cj = i / n1;
  Q[(ci + 1) + (cj + 1) * (n1 + 2)] = (T)acc;
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}

// block-level reduction of (max, sum, max, count) into out (thread 0 writes)
__device__ __forceinline__ void block_reduce4(double a, double b, double c, long long n,
                                              ResidualOut* out) {
  __shared__ double sa[kTile / 64], sb[kTile / 64], sc[kTile / 64];
  __shared__ long long sn[kTile / 64];
  a = wave_max(a);
  b = wave_sum(b);
  c = wave_max(c);
  n = wave_sum_ll(n);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sa[w] = a; sb[w] = b; sc[w] = c; sn[w] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ra = sa[0], rb = sb[0], rc = sc[0];
    long long rn = sn[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      ra = fmax(ra, sa[k]); rb = rb + sb[k]; rc = fmax(rc, sc[k]); rn += sn[k];
    }
    out->linf = ra; out->l2sq = rb; out->fnorm = rc; out->ncell = rn;
  }
}

// composite residual of one level: r = FT - L(Q) + reflux chunks; norms over
// visible cells; optional FP32/FP64 copy of r (defect-correction right side)
template <typename T, typename TO>
__global__ void __launch_bounds__(kTile) k_residual(
    const PatchInfo* __restrict__ P, const int2* __restrict__ tiles, const T* __restrict__ Q,
    const T* __restrict__ FT, const T* __restrict__ DeW, const T* __restrict__ DnS,
    const T* __restrict__ CC, T hx, T hy, const uint8_t* __restrict__ vis,
    const int* __restrict__ tgt_of_cell, const int* __restrict__ tgt_ptr,
    const int* __restrict__ tgt_ent, const CorrRec* __restrict__ fcorr, const T* __restrict__ fQ,
    const T* __restrict__ fQC, const T* __restrict__ fDeW, const T* __restrict__ fDnS,
    double hx2c, double hy2c, double area, ResidualOut* __restrict__ part, TO* __restrict__ rout) {
  const int2 t = tiles[blockIdx.x];
  const PatchInfo& pi = P[t.x];
  const int mx = pi.mx, my = pi.my;
  const int c = t.y + (int)threadIdx.x;
  double a = 0.0, b = 0.0, f = 0.0;
  long long n = 0;
  if (c < mx * my) {
    const int j = c / mx;
    const int i = c - j * mx;
    const int ldq = mx + 2;
    const int qc = pi.qoff + (i + 1) + (j + 1) * ldq;
    const int cc = pi.coff + c;
    const int fx = pi.xoff + i + j * (mx + 1);
    const int fy = pi.yoff + i + j * mx;
    T r = FT[cc] - op_at<T>(Q, qc, ldq, DeW[fx], DeW[fx + 1], DnS[fy], DnS[fy + mx], CC[cc], hx * hx, hy * hy);
    double rd = (double)r;
    if (tgt_of_cell) {
      const int row = tgt_of_cell[cc];
      if (row >= 0) {
        for (int e = tgt_ptr[row]; e < tgt_ptr[row + 1]; ++e)
          rd = rd + corr_value<T>(fcorr[tgt_ent[e]], fQ, fQC, fDeW, fDnS, hx2c, hy2c);
      }
    }
    if (rout) rout[cc] = (TO)rd;
    if (vis[cc]) {
      a = fabs(rd);
      b = rd * rd;
      f = fabs((double)FT[cc]);
      n = 1;
    }
  }
  block_reduce4(a, b * area, f, n, &part[blockIdx.x]);
}

__global__ void __launch_bounds__(kTile) k_reduce_parts(const ResidualOut* __restrict__ parts,
                                                        int nparts, ResidualOut* __restrict__ out) {
  double a = 0.0, b = 0.0, f = 0.0;
  long long n = 0;
  for (int k = threadIdx.x; k < nparts; k += kTile) {
    const ResidualOut p = parts[k];
    a = fmax(a, p.linf);
    b = b + p.l2sq;
    f = fmax(f, p.fnorm);
    n += p.ncell;
  }
  block_reduce4(a, b, f, n, out);
}

// composite mean: sum of visible u * da (ebCompositeMean.m:4-23)
template <typename T>
__global__ void __launch_bounds__(kTile) k_mass(const PatchInfo* __restrict__ P,
                                                const int2* __restrict__ tiles,
                                                const T* __restrict__ Q,
                                                const uint8_t* __restrict__ vis, double da,
                                                ResidualOut* __restrict__ part) {
  const int2 t = tiles[blockIdx.x];
  const PatchInfo& pi = P[t.x];
  const int mx = pi.mx, my = pi.my;
  const int c = t.y + (int)threadIdx.x;
  double s = 0.0;
  long long n = 0;
  if (c < mx * my && vis[pi.coff + c]) {
    const int j = c / mx;
    const int i = c - j * mx;
    s = (double)Q[pi.qoff + (i + 1) + (j + 1) * (mx + 2)];
    n = 1;
  }
  block_reduce4(0.0, s * da, 0.0, n, &part[blockIdx.x]);
}

// ---------------------------------------------------------------- elementwise
// Q64 interior += Q32 interior (defect-correction update)
__global__ void __launch_bounds__(kTile) k_add_correction(const PatchInfo* __restrict__ P,
                                                          const int2* __restrict__ tiles,
                                                          double* __restrict__ Q,
                                                          const float* __restrict__ E) {
  const int2 t = tiles[blockIdx.x];
  const PatchInfo& pi = P[t.x];
  const int mx = pi.mx, my = pi.my;
  const int c = t.y + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = c / mx;
  const int i = c - j * mx;
  const int qc = pi.qoff + (i + 1) + (j + 1) * (mx + 2);
  Q[qc] = Q[qc] + (double)E[qc];
}

template <typename T>
__global__ void __launch_bounds__(kTile) k_fill_value(T* __restrict__ a, long long n, T v) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t < n) a[t] = v;
}

template <typename TD, typename TS>
__global__ void __launch_bounds__(kTile) k_convert(TD* __restrict__ d, const TS* __restrict__ s,
                                                   long long n) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t < n) d[t] = (TD)s[t];
}

// BDF2 / BE right-hand side (ParabolicBSAM/src/ebpRhs.m:149-151)
__global__ void __launch_bounds__(kTile) k_bdf2_rhs(double* __restrict__ FT, const double* __restrict__ f,
                                                    const double* __restrict__ Uo,
                                                    const double* __restrict__ Uoo, double dt,
                                                    int first, long long n) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t >= n) return;
  const double fv = f ? f[t] : 0.0;
  FT[t] = first ? fv + Uo[t] / dt : fv + (2.0 * Uo[t] - 0.5 * Uoo[t]) / dt;
}

// history rotation: Uoo = Uo; Uo = interior(Q)
__global__ void __launch_bounds__(kTile) k_history(const PatchInfo* __restrict__ P,
                                                   const int2* __restrict__ tiles,
                                                   const double* __restrict__ Q,
                                                   double* __restrict__ Uo, double* __restrict__ Uoo,
                                                   int rotate) {
  const int2 t = tiles[blockIdx.x];
  const PatchInfo& pi = P[t.x];
  const int mx = pi.mx, my = pi.my;
  const int c = t.y + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = c / mx;
  const int i = c - j * mx;
  const double u = Q[pi.qoff + (i + 1) + (j + 1) * (mx + 2)];
  const int cc = pi.coff + c;
  Uoo[cc] = rotate ? Uo[cc] : u;
  Uo[cc] = u;
}

// InvDiagS refresh for a reaction shift: 1/(1/ID_phys + fold) with
// ID_phys = 1/(sum D/h^2 + C) (ebKernels.m:28-34, ebBcFold.m:12-42)
__global__ void __launch_bounds__(kTile) k_invdiag(const PatchInfo* __restrict__ P,
                                                   const int2* __restrict__ tiles,
                                                   const double* __restrict__ DeW,
                                                   const double* __restrict__ DnS,
                                                   const double* __restrict__ CC, double hx, double hy,
                                                   double* __restrict__ ID, float* __restrict__ ID32) {
  const int2 t = tiles[blockIdx.x];
  const PatchInfo& pi = P[t.x];
  const int mx = pi.mx, my = pi.my;
  const int c = t.y + (int)threadIdx.x;
  if (c >= mx * my) return;
  const int j = c / mx;
  const int i = c - j * mx;
  const int fx = pi.xoff + i + j * (mx + 1);
  const int fy = pi.yoff + i + j * mx;
  const double hx2 = hx * hx, hy2 = hy * hy;
  const double den = (DeW[fx + 1] + DeW[fx]) / hx2 + (DnS[fy + mx] + DnS[fy]) / hy2 + CC[pi.coff + c];
  const double idp = 1.0 / den;
  double add = 0.0;
  if (i == 0) {
    if (pi.bc[0] == 1) add = add + 2 * DeW[fx] / hx2;
    else if (pi.bc[0] == 2) add = add - DeW[fx] / hx2;
  }
  if (i == mx - 1) {
    if (pi.bc[1] == 1) add = add + 2 * DeW[fx + 1] / hx2;
    else if (pi.bc[1] == 2) add = add - DeW[fx + 1] / hx2;
  }
  if (j == 0) {
    if (pi.bc[2] == 1) add = add + 2 * DnS[fy] / hy2;
    else if (pi.bc[2] == 2) add = add - DnS[fy] / hy2;
  }
  if (j == my - 1) {
    if (pi.bc[3] == 1) add = add + 2 * DnS[fy + mx] / hy2;
    else if (pi.bc[3] == 2) add = add - DnS[fy + mx] / hy2;
  }
  const double v = 1.0 / (1.0 / idp + add);
  ID[pi.coff + c] = v;
  if (ID32) ID32[pi.coff + c] = (float)v;
}

__global__ void __launch_bounds__(kTile) k_add_scalar(double* __restrict__ a, long long n, double s,
                                                      float* __restrict__ a32) {
  const long long t = (long long)blockIdx.x * kTile + threadIdx.x;
  if (t < n) {
    const double v = a[t] + s;
    a[t] = v;
    if (a32) a32[t] = (float)v;
  }
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(long long n) { return (unsigned)((n + kTile - 1) / kTile); }

template <typename T>
void launch_relax(const LevelDev<T>& L, int colour, T omega, int homog, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_relax<T>, dim3(L.ntile), dim3(kTile), 0, s, L.P, L.tiles, L.Q, L.F, L.ID,
                     L.DeW, L.DnS, L.G, (T)L.h[0], (T)L.h[1], colour, omega, homog);
}

template <typename T>
void launch_fill(const LevelDev<T>& L, const LevelDev<T>* parent, int homog, hipStream_t s) {
  const int nring = parent ? L.nring_cf : 0;
  const long long n = (long long)L.nghost + nring;
  if (!n) return;
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(kTile), 0, s, L.ghost, L.nghost, L.ring_cf,
                     nring, L.Q, parent ? (const T*)parent->Q : (const T*)nullptr, L.QC, homog);
}

template <typename T>
void launch_ring(const LevelDev<T>& L, const LevelDev<T>& parent, hipStream_t s) {
  if (!L.nring_all) return;
  hipLaunchKernelGGL(k_ring<T>, dim3(nblk(L.nring_all)), dim3(kTile), 0, s, L.ring_all, L.nring_all,
                     (const T*)parent.Q, L.QC);
}

template <typename T>
void launch_restrict(const LevelDev<T>& L, const LevelDev<T>& parent, int qc_only, hipStream_t s) {
  if (!L.nfoot) return;
  hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(L.nfoot)), dim3(kTile), 0, s, L.P, L.foot, L.nfoot,
                     (const T*)L.Q, L.QC, parent.Q, qc_only);
}

template <typename T>
void launch_coarse_load(const LevelDev<T>& L, const LevelDev<T>& parent, int mass_corr,
                        hipStream_t s) {
  if (L.nfoot) {
    hipLaunchKernelGGL(k_coarse_load<T>, dim3(nblk(L.nfoot)), dim3(kTile), 0, s, L.P, L.foot, L.nfoot,
                       (const T*)L.Q, (const T*)L.F, (const T*)L.DeW, (const T*)L.DnS,
                       (const T*)L.CC, (const T*)L.QC, (T)L.h[0], (T)L.h[1], (const T*)parent.DeW,
                       (const T*)parent.DnS, (const T*)parent.CC, parent.F, (T)parent.h[0],
                       (T)parent.h[1]);
  }
  if (mass_corr && parent.ntgt) {
    hipLaunchKernelGGL(k_reflux<T>, dim3(nblk(parent.ntgt)), dim3(kTile), 0, s, parent.tgt_cell,
                       parent.tgt_ptr, parent.tgt_ent, parent.ntgt, L.corr, (const T*)L.Q,
                       (const T*)L.QC, (const T*)L.DeW, (const T*)L.DnS, parent.h[0] * parent.h[0],
                       parent.h[1] * parent.h[1], parent.F);
  }
}

template <typename T>
void launch_correct(const LevelDev<T>& L, const LevelDev<T>& parent, int init, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_correct<T>, dim3(L.ntile), dim3(kTile), 0, s, L.P, L.tiles, L.Q,
                     (const T*)L.QC, (const int*)L.wmap, (const T*)parent.Q, init);
}

template <typename T>
void launch_coarsest(const CoarsestDev& C, const LevelDev<T>& L1, int homog, hipStream_t s) {
  const unsigned nb = (unsigned)((C.N + 63) / 64);
  hipLaunchKernelGGL(k_coarsest<T>, dim3(nb), dim3(64), C.M * sizeof(double), s, C.Ainv, C.bcRhs,
                     C.N, C.M, L1.n[0], (const T*)L1.F, L1.Q, homog);
}

template <typename T, typename TO>
void launch_residual(const LevelDev<T>& L, const LevelDev<T>* fine, int mass_corr,
                     ResidualOut* parts, TO* rout, hipStream_t s) {
  if (!L.ntile) return;
  const bool corr = mass_corr && fine && L.ntgt;
  hipLaunchKernelGGL((k_residual<T, TO>), dim3(L.ntile), dim3(kTile), 0, s, L.P, L.tiles, (const T*)L.Q,
                     (const T*)L.FT, (const T*)L.DeW, (const T*)L.DnS, (const T*)L.CC, (T)L.h[0],
                     (T)L.h[1], (const uint8_t*)L.vis, corr ? (const int*)L.tgt_of_cell : nullptr,
                     (const int*)L.tgt_ptr, (const int*)L.tgt_ent,
                     corr ? (const CorrRec*)fine->corr : nullptr, corr ? (const T*)fine->Q : nullptr,
                     corr ? (const T*)fine->QC : nullptr, corr ? (const T*)fine->DeW : nullptr,
                     corr ? (const T*)fine->DnS : nullptr, L.h[0] * L.h[0], L.h[1] * L.h[1],
                     L.h[0] * L.h[1], parts, rout);
}

void launch_reduce(const ResidualOut* parts, int nparts, ResidualOut* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(kTile), 0, s, parts, nparts, out);
}

template <typename T>
void launch_mass(const LevelDev<T>& L, ResidualOut* parts, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_mass<T>, dim3(L.ntile), dim3(kTile), 0, s, L.P, L.tiles, (const T*)L.Q,
                     (const uint8_t*)L.vis, L.h[0] * L.h[1], parts);
}

void launch_add_correction(const LevelDev<double>& L, const LevelDev<float>& E, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_add_correction, dim3(L.ntile), dim3(kTile), 0, s, L.P, L.tiles, L.Q,
                     (const float*)E.Q);
}

template <typename T>
void launch_fill_value(T* a, long long n, T v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_value<T>, dim3(nblk(n)), dim3(kTile), 0, s, a, n, v);
}

template <typename TD, typename TS>
void launch_convert(TD* d, const TS* src, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_convert<TD, TS>), dim3(nblk(n)), dim3(kTile), 0, s, d, src, n);
}

void launch_bdf2_rhs(double* FT, const double* f, const double* Uo, const double* Uoo, double dt,
                     int first, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_bdf2_rhs, dim3(nblk(n)), dim3(kTile), 0, s, FT, f, Uo, Uoo, dt, first, n);
}

void launch_history(const LevelDev<double>& L, double* Uo, double* Uoo, int rotate, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_history, dim3(L.ntile), dim3(kTile), 0, s, L.P, L.tiles, (const double*)L.Q,
                     Uo, Uoo, rotate);
}

void launch_invdiag(const LevelDev<double>& L, float* ID32, hipStream_t s) {
  if (!L.ntile) return;
  hipLaunchKernelGGL(k_invdiag, dim3(L.ntile), dim3(kTile), 0, s, L.P, L.tiles, (const double*)L.DeW,
                     (const double*)L.DnS, (const double*)L.CC, L.h[0], L.h[1], L.ID, ID32);
}

void launch_add_scalar(double* a, long long n, double sft, float* a32, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add_scalar, dim3(nblk(n)), dim3(kTile), 0, s, a, n, sft, a32);
}

#define BSAM_INST(T)                                                                         \
  template void launch_relax<T>(const LevelDev<T>&, int, T, int, hipStream_t);               \
  template void launch_fill<T>(const LevelDev<T>&, const LevelDev<T>*, int, hipStream_t);    \
  template void launch_ring<T>(const LevelDev<T>&, const LevelDev<T>&, hipStream_t);         \
  template void launch_restrict<T>(const LevelDev<T>&, const LevelDev<T>&, int, hipStream_t);     \
  template void launch_coarse_load<T>(const LevelDev<T>&, const LevelDev<T>&, int, hipStream_t); \
  template void launch_correct<T>(const LevelDev<T>&, const LevelDev<T>&, int, hipStream_t); \
  template void launch_coarsest<T>(const CoarsestDev&, const LevelDev<T>&, int, hipStream_t); \
  template void launch_mass<T>(const LevelDev<T>&, ResidualOut*, hipStream_t);              \
  template void launch_fill_value<T>(T*, long long, T, hipStream_t);
BSAM_INST(double)
BSAM_INST(float)
template void launch_residual<double, float>(const LevelDev<double>&, const LevelDev<double>*, int,
                                             ResidualOut*, float*, hipStream_t);
template void launch_convert<float, double>(float*, const double*, long long, hipStream_t);
template void launch_convert<double, float>(double*, const float*, long long, hipStream_t);

}  // namespace bsam
