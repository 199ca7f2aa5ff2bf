// Kernel (b), part 2: divide-and-conquer eigensolver for the symmetric
// tridiagonal matrix produced by the Householder reduction (the role of the
// tridiagonal stage inside DSYEVR, imcom_matrix.f90:642).
//
// Cuppen's method with Gu-Eisenstat eigenvectors, fully on the GPU:
//  * the matrix is scaled to unit max-norm, every off-diagonal e_i becomes a
//    split point of a balanced binary tree whose leaves are 1x1, and the
//    leaves hold d_i - |e_{i-1}| - |e_i| with eigenvector 1;
//  * merges are processed level by level (all merges of a level in one launch
//    per step).  A merge of blocks (D1,Z1), (D2,Z2) coupled by beta solves
//    D + rho z z^T, rho = 2|beta|, z = [last row of Z1, sign(beta) first row
//    of Z2] / sqrt(2), after deflation:
//      - small rho|z_i| <= tol, tol = 8 eps max(|d|max, |z|max): the column
//        keeps its eigenvector;
//      - close pairs |t c s| <= tol: a Givens rotation zeroes one z entry;
//    (the deflation rules of LAPACK's dlaed2, kept step for step);
//  * the k non-deflated roots of 1 + rho sum z_i^2/(d_i - x) = 0 are found
//    one per thread, stored as (origin pole, offset tau) so d_i - lambda_j is
//    formed without cancellation; the root iteration models the two nearest
//    poles (value and slope matched) and falls back to bisection;
//  * z is recomputed from the roots (Loewner / Gu-Eisenstat) so the
//    eigenvectors u_j = (zhat_i / (d_i - lambda_j))_i are orthogonal to
//    working precision;
//  * new eigenvectors = Z(:, non-deflated) x U on the FP64 matrix cores
//    (grouped GEMM over the merges of the level);
//  * eigenvalues of each merge are left in ascending order.
// One host synchronisation per level reads the deflation counts.
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace imcom {

struct DcMerge {
  int off, n1, n2;   // off: offset in the concatenated vectors of all tiles
  double* Z;         // this merge's tile eigenvector matrix
  int64_t ldz;       // its leading dimension (the tile's n)
  int loff;          // offset of the merge inside its tile
  double rho;    // |beta|
  double sigma;  // sign(beta)
  int64_t u_off;   // offset of this merge's k x k U block (inside a K x K slot)
  int64_t zg_off;  // offset of this merge's K x K gather block (and K x k output)
};

struct DcCount {
  int k;     // non-deflated
  int nrot;  // deflation rotations
  int k1;    // non-deflated columns with entries in the first half's rows only
  int k2;    // ... in both halves' rows (mixed by a rotation); k - k1 - k2: second half only
};

struct DcPlan;

struct DcWork {
  int64_t n = 0;    // capacity of the concatenated vectors
  int64_t nn = 0;   // capacity of the matrix workspaces (sum of n_t^2)
  double *D = nullptr, *ds = nullptr, *zs = nullptr, *dl = nullptr, *wz = nullptr,
         *dfd = nullptr, *rc = nullptr, *rs = nullptr, *tau = nullptr, *lam = nullptr,
         *zhat = nullptr, *U = nullptr, *Zg = nullptr, *Zo = nullptr, *scale = nullptr,
         *dcat = nullptr, *ecat = nullptr;
  int *cs = nullptr, *ndcol = nullptr, *dfcol = nullptr, *rp = nullptr, *rn = nullptr,
      *org = nullptr, *d_prefix = nullptr, *ndpos = nullptr;
  DcMerge* d_merges = nullptr;
  DcCount* d_counts = nullptr;
  GemmDesc* d_descs = nullptr;
  DcPlan* plan = nullptr;  // launch plan for the last tile sizes (kept across dc_free)
};

static void dc_free(DcWork* w) {
  void* ptrs[] = {w->D,   w->ds,    w->zs,    w->dl,       w->wz,       w->dfd,    w->rc,
                  w->rs,  w->tau,   w->lam,   w->zhat,     w->U,        w->Zg,     w->Zo,
                  w->scale, w->dcat, w->ecat, w->cs,  w->ndcol, w->dfcol,    w->rp,
                  w->rn,     w->org, w->d_prefix, w->d_merges, w->d_counts, w->d_descs,
                  w->ndpos};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  DcPlan* plan = w->plan;
  *w = DcWork();
  w->plan = plan;
}

static void dc_plan_delete(DcPlan* p);

DcWork* dc_create() { return new DcWork(); }
void dc_destroy(DcWork* w) {
  if (!w) return;
  dc_free(w);
  dc_plan_delete(w->plan);
  delete w;
}

static hipError_t dc_reserve(DcWork* w, int64_t n, int64_t nn) {
  if (w->n >= n && w->nn >= nn) return hipSuccess;
  n = std::max(n, w->n);
  nn = std::max(nn, w->nn);
  dc_free(w);
  const size_t nd = (size_t)n * sizeof(double), ni = (size_t)n * sizeof(int);
  const size_t nnb = (size_t)nn * sizeof(double);
  double** dv[] = {&w->D, &w->ds, &w->zs, &w->dl, &w->wz, &w->dfd, &w->rc, &w->rs,
                   &w->tau, &w->lam, &w->zhat, &w->dcat, &w->ecat};
  for (double** q : dv) IMCOM_HIP_TRY(hipMalloc(q, nd));
  IMCOM_HIP_TRY(hipMalloc(&w->scale, sizeof(double) * 2));
  IMCOM_HIP_TRY(hipMalloc(&w->U, nnb));
  IMCOM_HIP_TRY(hipMalloc(&w->Zg, nnb));
  IMCOM_HIP_TRY(hipMalloc(&w->Zo, nnb));
  int** iv[] = {&w->cs, &w->ndcol, &w->dfcol, &w->rp, &w->rn, &w->org, &w->ndpos};
  for (int** q : iv) IMCOM_HIP_TRY(hipMalloc(q, ni));
  IMCOM_HIP_TRY(hipMalloc(&w->d_prefix, ni + sizeof(int)));
  IMCOM_HIP_TRY(hipMalloc(&w->d_merges, (size_t)n * sizeof(DcMerge)));
  IMCOM_HIP_TRY(hipMalloc(&w->d_counts, (size_t)n * sizeof(DcCount)));
  IMCOM_HIP_TRY(hipMalloc(&w->d_descs, (size_t)n * sizeof(GemmDesc)));
  w->n = n;
  w->nn = nn;
  return hipSuccess;
}

// ---- set-up ----------------------------------------------------------------
// scale = max(|d|, |e|) (one workgroup)
__global__ __launch_bounds__(1024) void dc_scale_kernel(int64_t n, const double* __restrict__ d,
                                                        const double* __restrict__ e,
                                                        double* __restrict__ scale) {
  __shared__ double red[16];
  double mx = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    mx = fmax(mx, fabs(d[i]));
    if (i < n - 1) mx = fmax(mx, fabs(e[i]));
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    scale[0] = m > 0.0 ? m : 1.0;
  }
}

// leaves: D_i = (d_i - |e_{i-1}| - |e_i|) / scale, Z = I
__global__ void dc_leaves_kernel(int64_t n, const double* __restrict__ d,
                                 const double* __restrict__ e, const double* __restrict__ scale,
                                 double* __restrict__ D) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = scale[0];
  double v = d[i] / s;
  if (i > 0) v -= fabs(e[i - 1] / s);
  if (i < n - 1) v -= fabs(e[i] / s);
  D[i] = v;
}

__global__ void dc_identity_kernel(int64_t n, double* __restrict__ Z) {
  const int64_t j = blockIdx.x;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) Z[i + j * n] = (i == j) ? 1.0 : 0.0;
}

__device__ __forceinline__ int find_merge(const DcMerge* __restrict__ mg, int nm, int64_t g) {
  // last merge with off <= g; -1 when g is not covered by a merge of this level
  int lo = 0, hi = nm - 1;
  if (nm <= 0 || g < mg[0].off) return -1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mg[mid].off <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  const DcMerge& m = mg[lo];
  return (g < (int64_t)m.off + m.n1 + m.n2) ? lo : -1;
}

// ---- step 1: merge-sort d of the two children, gather z ---------------------
__global__ void dc_sort_kernel(int64_t n, const DcMerge* __restrict__ mg, int nm,
                               const double* __restrict__ D, const double* __restrict__ Z,
                               double* __restrict__ ds, double* __restrict__ zs,
                               int* __restrict__ cs) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int mi = find_merge(mg, nm, g);
  if (mi < 0) return;
  const DcMerge m = mg[mi];
  const int q = (int)(g - m.off);
  const double* D1 = D + m.off;
  const double* D2 = D + m.off + m.n1;
  const double v = D[g];
  int rank;
  double z;
  const double r2 = 0.70710678118654752440;  // 1/sqrt(2)
  if (q < m.n1) {
    // rank in D2: count of D2 < v
    int lo = 0, hi = m.n2;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (D2[mid] < v)
        lo = mid + 1;
      else
        hi = mid;
    }
    rank = q + lo;
    z = m.Z[(m.loff + m.n1 - 1) + (int64_t)(m.loff + q) * m.ldz] * r2;
  } else {
    // rank in D1: count of D1 <= v
    int lo = 0, hi = m.n1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (D1[mid] <= v)
        lo = mid + 1;
      else
        hi = mid;
    }
    rank = (q - m.n1) + lo;
    z = m.sigma * m.Z[(m.loff + m.n1) + (int64_t)(m.loff + q) * m.ldz] * r2;
  }
  ds[m.off + rank] = v;
  zs[m.off + rank] = z;
  cs[m.off + rank] = q;
}

// ---- step 2: deflation scan (dlaed2 rules), one wavefront per merge -------
__device__ __forceinline__ double dlapy2_d(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// dlaed2's rotation test for the candidate (nd, nz) against the previous
// candidate (pd, pz): c, s of the rotation that zeroes one z entry, and
// whether the rotated matrix's off-diagonal entry t c s is negligible.
// Unfused, so that both code paths below take the same decision.
__device__ __forceinline__ bool dc_rot_test(double pd, double pz, double nd, double nz, double tol,
                                            double& c, double& s, double& tau) {
#pragma clang fp contract(off)
  s = pz;
  c = nz;
  tau = dlapy2_d(c, s);
  const double t = nd - pd;
  c = c / tau;
  s = -s / tau;
  return fabs(t * c * s) <= tol;
}

// One wavefront per merge, 64 entries at a time: each lane takes one entry of
// the sorted (d, z); entries with negligible z deflate, every other entry
// (a candidate) is tested against the previous candidate.  A chunk with no
// deflating rotation, whose deflated values keep ascending, is resolved at
// once (ballots and prefix counts give each output its place); a chunk with a
// rotation is walked by lane 0 with the scalar rules, since a rotation
// changes the previous candidate's d and z.  The outputs are those of the
// scalar scan over the whole merge.  Scan state: the pending candidate
// (pd, pz, pcol) once one exists, k non-deflated so far, ndf deflated, nrot
// rotations, lastv the largest deflated value appended, pmix whether the
// pending column mixes both halves (a rotation between a first-half and a
// second-half column, or a column that already did).
//
// Each non-deflated column's rows are nonzero in the first half, the second
// half or both (type 0 / 1 / 2: the first half's columns, the rotated mixed
// ones, the second half's).  ndpos[q] receives the column's place in the
// type-grouped order (all type 0, then 1, then 2, each in scan order), and
// counts the numbers k1 (type 0) and k2 (type 1): the eigenvector product
// then runs per half over the columns with entries in that half.
__device__ __forceinline__ int dc_col_type(int col, bool mix, int n1) {
  return mix ? 1 : (col < n1 ? 0 : 2);
}

__global__ __launch_bounds__(64) void dc_deflate_kernel(const DcMerge* __restrict__ mg, int nm,
                                  const double* __restrict__ ds, const double* __restrict__ zs,
                                  const int* __restrict__ cs, double* __restrict__ dl,
                                  double* __restrict__ wz, int* __restrict__ ndcol,
                                  int* __restrict__ ndpos,
                                  double* __restrict__ dfd, int* __restrict__ dfcol,
                                  int* __restrict__ rp, int* __restrict__ rn,
                                  double* __restrict__ rc, double* __restrict__ rs,
                                  DcCount* __restrict__ counts) {
  __shared__ double cd[64], cz[64];
  __shared__ int ccol[64];
  __shared__ double st_d[3];  // pd, pz, lastv after a walked chunk
  __shared__ int st_i[6];     // pcol, have_prev, k, ndf, nrot, pmix
  const int mi = blockIdx.x;
  if (mi >= nm) return;
  const int lane = threadIdx.x;
  const DcMerge m = mg[mi];
  const int K = m.n1 + m.n2;
  const int n1 = m.n1;
  const int o = m.off;
  const double rho = 2.0 * m.rho;
  const double eps = 1.1102230246251565e-16;
  double dmax = 0.0, zmax = 0.0;
  for (int j = lane; j < K; j += 64) {
    dmax = fmax(dmax, fabs(ds[o + j]));
    zmax = fmax(zmax, fabs(zs[o + j]));
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
    zmax = fmax(zmax, __shfl_xor(zmax, off, 64));
  }
  const double tol = 8.0 * eps * fmax(dmax, zmax);
  if (rho * zmax <= tol) {  // every entry deflates, in their (ascending) order
    for (int j = lane; j < K; j += 64) {
      dfd[o + j] = ds[o + j];
      dfcol[o + j] = cs[o + j];
    }
    if (lane == 0) {
      counts[mi].k = 0;
      counts[mi].nrot = 0;
      counts[mi].k1 = 0;
      counts[mi].k2 = 0;
    }
    return;
  }
  bool have_prev = false, pmix = false;
  double pd = 0.0, pz = 0.0, lastv = -HUGE_VAL;
  int pcol = 0, k = 0, ndf = 0, nrot = 0;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));  // lanes < lane
  for (int b = 0; b < K; b += 64) {
    const int j = b + lane;
    const bool in = j < K;
    const double nd = in ? ds[o + j] : 0.0, nz = in ? zs[o + j] : 0.0;
    const int ncol = in ? cs[o + j] : 0;
    const bool small = in && rho * fabs(nz) <= tol;
    const bool cand = in && !small;
    const uint64_t cm = __ballot(cand), sm = __ballot(small);
    // previous candidate (in this chunk, else the pending one) and previous
    // deflated value appended (in this chunk, else lastv)
    const uint64_t cb = cm & below, sb = sm & below;
    const int pl = cb ? 63 - __clzll(cb) : 0;
    const int sl = sb ? 63 - __clzll(sb) : 0;
    const double qd_in = __shfl(nd, pl, 64), qz_in = __shfl(nz, pl, 64);
    const double sv_in = __shfl(nd, sl, 64);
    bool rot = false;
    if (cand && (cb || have_prev)) {
      double c, s, tau;
      rot = dc_rot_test(cb ? qd_in : pd, cb ? qz_in : pz, nd, nz, tol, c, s, tau);
    }
    const bool unordered = small && nd < (sb ? sv_in : lastv);
    if (__ballot(rot) == 0 && __ballot(unordered) == 0) {
      const int r = __popcll(cm);
      if (r > 0) {
        int base = k;
        if (have_prev) {
          if (lane == 0) {
            dl[o + k] = pd;
            wz[o + k] = pz;
            ndcol[o + k] = pcol;
            ndpos[o + k] = dc_col_type(pcol, pmix, n1);
          }
          base = k + 1;
        }
        const int idx = __popcll(cb);
        if (cand && idx < r - 1) {
          dl[o + base + idx] = nd;
          wz[o + base + idx] = nz;
          ndcol[o + base + idx] = ncol;
          ndpos[o + base + idx] = dc_col_type(ncol, false, n1);
        }
        k = base + r - 1;
        const int L = 63 - __clzll(cm);
        pd = __shfl(nd, L, 64);
        pz = __shfl(nz, L, 64);
        pcol = __shfl(ncol, L, 64);
        have_prev = true;
        pmix = false;
      }
      const int ns = __popcll(sm);
      if (small) {
        const int si = ndf + __popcll(sb);
        dfd[o + si] = nd;
        dfcol[o + si] = ncol;
      }
      if (ns > 0) lastv = __shfl(nd, 63 - __clzll(sm), 64);
      ndf += ns;
    } else {
      cd[lane] = nd;
      cz[lane] = nz;
      ccol[lane] = ncol;
      __syncthreads();
      if (lane == 0) {
        const int cnt = min(64, K - b);
        for (int q = 0; q < cnt; ++q) {
          const double vd = cd[q], vz = cz[q];
          const int vc = ccol[q];
          const bool vsmall = rho * fabs(vz) <= tol;
          auto push_df = [&](double v, int col) {
            // ascending insertion (values arrive almost sorted)
            int p = ndf;
            if (v < lastv) {
              while (p > 0 && dfd[o + p - 1] > v) {
                dfd[o + p] = dfd[o + p - 1];
                dfcol[o + p] = dfcol[o + p - 1];
                --p;
              }
            } else {
              lastv = v;
            }
            dfd[o + p] = v;
            dfcol[o + p] = col;
            ++ndf;
          };
          if (vsmall) {
            push_df(vd, vc);
          } else if (!have_prev) {
            pd = vd;
            pz = vz;
            pcol = vc;
            have_prev = true;
            pmix = false;
          } else {
            double c, s, tau;
            if (dc_rot_test(pd, pz, vd, vz, tol, c, s, tau)) {
              // rotate columns (pcol, vc): x' = c x + s y, y' = c y - s x
              rp[o + nrot] = pcol;
              rn[o + nrot] = vc;
              rc[o + nrot] = c;
              rs[o + nrot] = s;
              ++nrot;
              double tp, tn;
              {
#pragma clang fp contract(off)
                tp = pd * c * c + vd * s * s;
                tn = pd * s * s + vd * c * c;
              }
              push_df(tp, pcol);
              pmix = pmix || ((pcol < n1) != (vc < n1));
              pd = tn;
              pz = tau;
              pcol = vc;
            } else {
              dl[o + k] = pd;
              wz[o + k] = pz;
              ndcol[o + k] = pcol;
              ndpos[o + k] = dc_col_type(pcol, pmix, n1);
              ++k;
              pd = vd;
              pz = vz;
              pcol = vc;
              pmix = false;
            }
          }
        }
        st_d[0] = pd;
        st_d[1] = pz;
        st_d[2] = lastv;
        st_i[0] = pcol;
        st_i[1] = have_prev ? 1 : 0;
        st_i[2] = k;
        st_i[3] = ndf;
        st_i[4] = nrot;
        st_i[5] = pmix ? 1 : 0;
      }
      __syncthreads();
      pd = st_d[0];
      pz = st_d[1];
      lastv = st_d[2];
      pcol = st_i[0];
      have_prev = st_i[1] != 0;
      k = st_i[2];
      ndf = st_i[3];
      nrot = st_i[4];
      pmix = st_i[5] != 0;
      __syncthreads();  // before the next walked chunk overwrites the staging arrays
    }
  }
  if (lane == 0) {
    dl[o + k] = pd;
    wz[o + k] = pz;
    ndcol[o + k] = pcol;
    ndpos[o + k] = dc_col_type(pcol, pmix, n1);
  }
  ++k;
  __syncthreads();  // the types, written by any lane, are read back below
  // type-grouped places: counts of types 0 and 1, then each column's place
  int c0 = 0, c1 = 0;
  for (int b = 0; b < k; b += 64) {
    const int q = b + lane;
    const int t = q < k ? ndpos[o + q] : -1;
    c0 += __popcll(__ballot(t == 0));
    c1 += __popcll(__ballot(t == 1));
  }
  int r0 = 0, r1 = c0, r2 = c0 + c1;  // next place of each type
  for (int b = 0; b < k; b += 64) {
    const int q = b + lane;
    const int t = q < k ? ndpos[o + q] : -1;
    const uint64_t m0 = __ballot(t == 0), m1 = __ballot(t == 1), m2 = __ballot(t == 2);
    if (q < k) {
      const uint64_t mt = t == 0 ? m0 : (t == 1 ? m1 : m2);
      const int rt = t == 0 ? r0 : (t == 1 ? r1 : r2);
      ndpos[o + q] = rt + __popcll(mt & below);
    }
    r0 += __popcll(m0);
    r1 += __popcll(m1);
    r2 += __popcll(m2);
  }
  if (lane == 0) {
    counts[mi].k = k;
    counts[mi].nrot = nrot;
    counts[mi].k1 = c0;
    counts[mi].k2 = c1;
  }
}

// ---- step 3: apply the deflation rotations to Z's columns ------------------
__global__ void dc_rotate_kernel(int64_t n, const DcMerge* __restrict__ mg, int nm,
                                 const DcCount* __restrict__ counts, const int* __restrict__ rp,
                                 const int* __restrict__ rn, const double* __restrict__ rc,
                                 const double* __restrict__ rs, double* __restrict__ Z) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // row
  if (g >= n) return;
  const int mi = find_merge(mg, nm, g);
  if (mi < 0) return;
  const DcMerge m = mg[mi];
  const int nr = counts[mi].nrot;
  const int o = m.off;
  for (int r = 0; r < nr; ++r) {
    const int64_t row = m.loff + (g - m.off);
    double* x = m.Z + row + (int64_t)(m.loff + rp[o + r]) * m.ldz;
    double* y = m.Z + row + (int64_t)(m.loff + rn[o + r]) * m.ldz;
    const double c = rc[o + r], s = rs[o + r];
    const double xv = *x, yv = *y;
    *x = c * xv + s * yv;
    *y = c * yv - s * xv;
  }
}

// ---- step 4: secular equation roots ----------------------------------------
// f(x) = 1 + rho sum_i w_i^2 / (d_i - x); root j lies in (d_j, d_{j+1}) (last
// root in (d_{k-1}, d_{k-1} + rho |w|^2]).  Stored as lambda_j = d_org + tau.
// f = 1 + psi + phi at tau from d_org, with psi, dpsi over the poles
// i <= j and phi, dphi over i > j.  The k-term sums are split over the 64
// lanes of a wavefront (lane l takes terms l, l + 64, ...) and combined by a
// butterfly, which leaves bitwise the same sums in every lane.
__device__ void secular_eval_wave(int k, const double* __restrict__ dl,
                                  const double* __restrict__ w, double rho, int org, double tau,
                                  int j, double& f, double& psi, double& dpsi, double& phi,
                                  double& dphi) {
  const int lane = threadIdx.x & 63;
  const double dorg = dl[org];
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  for (int i = lane; i < k; i += 64) {
    const double del = (dl[i] - dorg) - tau;
    const double t = w[i] / del;
    if (i <= j) {
      a += w[i] * t;
      b += t * t;
    } else {
      c += w[i] * t;
      d += t * t;
    }
  }
  psi = wave_sum(a) * rho;
  dpsi = wave_sum(b) * rho;
  phi = wave_sum(c) * rho;
  dphi = wave_sum(d) * rho;
  f = 1.0 + psi + phi;
}

__global__ __launch_bounds__(256) void dc_secular_kernel(const DcMerge* __restrict__ mg, const DcCount* __restrict__ counts,
                                  const int* __restrict__ rprefix, int nm, int total,
                                  const double* __restrict__ dlall, const double* __restrict__ wall,
                                  int* __restrict__ orgall, double* __restrict__ tauall,
                                  double* __restrict__ lamall) {
  // one wavefront per root (uniform control flow: every lane holds the
  // same sums)
  const int gr = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (gr >= total) return;
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rprefix[mid] <= gr)
      lo = mid;
    else
      hi = mid - 1;
  }
  const DcMerge m = mg[lo];
  const int k = counts[lo].k;
  const int j = gr - rprefix[lo];
  if (j >= k) return;
  const double* dl = dlall + m.off;
  const double* w = wall + m.off;
  const double rho = 2.0 * m.rho;
  const double eps = 1.1102230246251565e-16;
  int org;
  double tlo, thi, tau;
  if (k == 1) {
    if (lane == 0) {
      orgall[m.off] = 0;
      tauall[m.off] = rho * w[0] * w[0];
      lamall[m.off] = dl[0] + rho * w[0] * w[0];
    }
    return;
  }
  if (j < k - 1) {
    const double gap = dl[j + 1] - dl[j];
    const double mid = 0.5 * gap;
    double f, psi, dpsi, phi, dphi;
    secular_eval_wave(k, dl, w, rho, j, mid, j, f, psi, dpsi, phi, dphi);
    if (f >= 0.0) {
      org = j;  // root in (d_j, mid]
      tlo = 0.0;
      thi = mid;
    } else {
      org = j + 1;  // root in (mid, d_{j+1}) -> tau in (-(gap - mid), 0)
      tlo = -(gap - mid);
      thi = 0.0;
    }
    tau = 0.5 * (tlo + thi);
  } else {
    double wn = 0.0;
    for (int i = lane; i < k; i += 64) wn += w[i] * w[i];
    wn = wave_sum(wn);
    org = k - 1;
    tlo = 0.0;
    thi = rho * wn;
    tau = 0.5 * thi;
  }
  const double dorg = dl[org];
  for (int it = 0; it < 200; ++it) {
    double f, psi, dpsi, phi, dphi;
    secular_eval_wave(k, dl, w, rho, org, tau, j, f, psi, dpsi, phi, dphi);
    if (f == 0.0) break;
    const double erretm = 8.0 * (phi - psi) + 1.0 + fabs(tau) * (dpsi + dphi);
    if (fabs(f) <= eps * erretm) break;
    if (f < 0.0)
      tlo = tau;
    else
      thi = tau;
    // two nearest poles around the root, relative to the current iterate
    double tnew;
    if (j < k - 1) {
      const double Dj = (dl[j] - dorg) - tau;          // < 0
      const double Dj1 = (dl[j + 1] - dorg) - tau;     // > 0
      const double s = Dj * Dj * dpsi, S = Dj1 * Dj1 * dphi;
      const double c = f - Dj * dpsi - Dj1 * dphi;
      const double a = c * (Dj + Dj1) + s + S;
      const double b = c * Dj * Dj1 + s * Dj1 + S * Dj;  // = f Dj Dj1
      // root of c eta^2 - a eta + b = 0 nearest to 0: eta = 2b / (a + sign(a) sqrt(disc))
      const double disc = sqrt(fabs(a * a - 4.0 * b * c));
      const double den = a + copysign(disc, a);
      tnew = (den != 0.0) ? tau + 2.0 * b / den : 0.5 * (tlo + thi);
    } else {
      const double Dk = (dl[k - 1] - dorg) - tau;  // < 0
      const double s = Dk * Dk * dpsi;
      const double c = f - Dk * dpsi;
      // c + s / (Dk - eta) = 0  ->  eta = Dk + s / c
      tnew = (c != 0.0) ? tau + (Dk + s / c) : 0.5 * (tlo + thi);
    }
    if (!(tnew > tlo && tnew < thi)) tnew = 0.5 * (tlo + thi);
    if (tnew == tau || thi - tlo <= 4.0 * eps * fmax(fabs(tlo), fabs(thi))) {
      tau = tnew;
      break;
    }
    tau = tnew;
  }
  if (lane == 0) {
    orgall[m.off + j] = org;
    tauall[m.off + j] = tau;
    lamall[m.off + j] = dorg + tau;
  }
}

// ---- step 5: Gu-Eisenstat zhat ----------------------------------------------
__global__ void dc_zhat_kernel(const DcMerge* __restrict__ mg, const DcCount* __restrict__ counts,
                               const int* __restrict__ rprefix, int nm, int total,
                               const double* __restrict__ dlall, const double* __restrict__ wall,
                               const int* __restrict__ orgall, const double* __restrict__ tauall,
                               double* __restrict__ zhat) {
  const int gr = blockIdx.x * blockDim.x + threadIdx.x;
  if (gr >= total) return;
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rprefix[mid] <= gr)
      lo = mid;
    else
      hi = mid - 1;
  }
  const DcMerge m = mg[lo];
  const int k = counts[lo].k;
  const int i = gr - rprefix[lo];
  if (i >= k) return;
  const double* dl = dlall + m.off;
  const int* org = orgall + m.off;
  const double* tau = tauall + m.off;
  const double di = dl[i];
  // W(i) = delta_ii prod_{j != i} delta_ij / (d_i - d_j),  delta_ij = d_i - lambda_j
  double Wi = (di - dl[org[i]]) - tau[i];
  for (int j = 0; j < k; ++j) {
    if (j == i) continue;
    const double dij = (di - dl[org[j]]) - tau[j];
    Wi *= dij / (di - dl[j]);
  }
  const double zi = wall[m.off + i];
  zhat[m.off + i] = copysign(sqrt(fmax(-Wi, 0.0)), zi);
}

// ---- step 6: eigenvectors of D + rho z z^T, normalised ----------------------
// one workgroup per (merge, column j); U_m is k x k, ld k, at u_off.
__global__ __launch_bounds__(256) void dc_vec_kernel(const DcMerge* __restrict__ mg,
                                                     const DcCount* __restrict__ counts,
                                                     const int* __restrict__ rprefix, int nm,
                                                     const double* __restrict__ dlall,
                                                     const int* __restrict__ orgall,
                                                     const double* __restrict__ tauall,
                                                     const double* __restrict__ zhat,
                                                     const int* __restrict__ ndpos,
                                                     double* __restrict__ Uall) {
  __shared__ double red[4];
  const int gr = blockIdx.x;
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (rprefix[mid] <= gr)
      lo = mid;
    else
      hi = mid - 1;
  }
  const DcMerge m = mg[lo];
  const int k = counts[lo].k;
  const int j = gr - rprefix[lo];
  if (j >= k) return;
  const double* dl = dlall + m.off;
  const double dorg = dl[orgall[m.off + j]];
  const double tj = tauall[m.off + j];
  double* U = Uall + m.u_off + (int64_t)j * k;
  double ss = 0.0;
  for (int i = threadIdx.x; i < k; i += blockDim.x) {  // row i goes to its grouped place
    const double del = (dl[i] - dorg) - tj;
    const double u = zhat[m.off + i] / del;
    U[ndpos[m.off + i]] = u;
    ss += u * u;
  }
  ss = block_sum(ss, red);
  const double inv = 1.0 / sqrt(ss);
  for (int i = threadIdx.x; i < k; i += blockDim.x) U[i] *= inv;
}

// ---- step 7: gather Z columns: non-deflated first (in their type-grouped
// places), then deflated; grid.x = global column g (merge m, local column q);
// Zg_m is K x K (ld K).
__global__ void dc_gather_kernel(int64_t n, const DcMerge* __restrict__ mg, int nm,
                                 const DcCount* __restrict__ counts, const int* __restrict__ ndcol,
                                 const int* __restrict__ ndpos, const int* __restrict__ dfcol,
                                 const double* __restrict__ Z, double* __restrict__ Zg) {
  const int64_t g = blockIdx.x;
  const int mi = find_merge(mg, nm, g);
  if (mi < 0) return;
  const DcMerge m = mg[mi];
  const int K = m.n1 + m.n2;
  const int k = counts[mi].k;
  const int q = (int)(g - m.off);
  const int src = (q < k) ? ndcol[m.off + q] : dfcol[m.off + (q - k)];
  const int dst = (q < k) ? ndpos[m.off + q] : q;
  const double* zc = m.Z + m.loff + (int64_t)(m.loff + src) * m.ldz;
  double* out = Zg + m.zg_off + (int64_t)dst * K;
  for (int r = threadIdx.x; r < K; r += blockDim.x) out[r] = zc[r];
}

// ---- step 8: place eigenpairs in ascending order ----------------------------
// Non-deflated column j (new vectors in Zo_m, K x k) and deflated column jd
// (in Zg_m at column k + jd) are merged by value.
__global__ void dc_place_kernel(int64_t n, const DcMerge* __restrict__ mg, int nm,
                                const DcCount* __restrict__ counts, const double* __restrict__ lam,
                                const double* __restrict__ dfd, const double* __restrict__ Zg,
                                const double* __restrict__ Zo, double* __restrict__ D,
                                double* __restrict__ Z) {
  const int64_t g = blockIdx.x;
  const int mi = find_merge(mg, nm, g);
  if (mi < 0) return;
  const DcMerge m = mg[mi];
  const int K = m.n1 + m.n2;
  const int k = counts[mi].k;
  const int ndf = K - k;
  const int q = (int)(g - m.off);
  const double* L = lam + m.off;
  const double* F = dfd + m.off;
  int pos;
  double val;
  const double* src;
  if (q < k) {
    val = L[q];
    int lo = 0, hi = ndf;  // count deflated < val
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (F[mid] < val)
        lo = mid + 1;
      else
        hi = mid;
    }
    pos = q + lo;
    src = Zo + m.zg_off + (int64_t)q * K;
  } else {
    const int jd = q - k;
    val = F[jd];
    int lo = 0, hi = k;  // count non-deflated <= val
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (L[mid] <= val)
        lo = mid + 1;
      else
        hi = mid;
    }
    pos = jd + lo;
    src = Zg + m.zg_off + (int64_t)q * K;
  }
  double* dst = m.Z + m.loff + (int64_t)(m.loff + pos) * m.ldz;
  for (int r = threadIdx.x; r < K; r += blockDim.x) dst[r] = src[r];
  if (threadIdx.x == 0) D[m.off + pos] = val;
}

__global__ void dc_unscale_kernel(int64_t n, const double* __restrict__ D,
                                  const double* __restrict__ scale, double* __restrict__ lam) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lam[i] = D[i] * scale[0];
}

// ---- plan-time set-up kernels --------------------------------------------------
// fill (in place of a memset, which a captured graph must not hold)
__global__ void dc_fill_kernel(int64_t n, double v, double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// coupling of every merge of every level: beta = e(off + n1 - 1) / scale
__global__ void dc_rho_kernel(DcMerge* __restrict__ mg, int nm, const double* __restrict__ e,
                              const double* __restrict__ scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  const double beta = e[mg[i].off + mg[i].n1 - 1] / scale[0];
  mg[i].rho = fabs(beta);
  mg[i].sigma = beta < 0.0 ? -1.0 : 1.0;
}

// the level's eigenvector products Zo_m (K x k) = Zg_m(:, 0:k) U_m (k x k,
// ld k) take their k from the deflation counts on the device: descriptors
// built for k = K are cut down here; blocks beyond a product's tiles exit
// Each merge's eigenvector product per half: rows [0, n1) take the grouped
// columns [0, k1 + k2) (types 0 and 1), rows [n1, K) the columns [k1, k)
// (types 1 and 2), each against the same rows of U.  dmerge[i] = 2 x (merge
// index in the level) + half.  Formed from the plan's fixed parts on every
// solve, so replays of a captured solve find them as set here.
__global__ void dc_fix_descs_kernel(GemmDesc* __restrict__ d, const int* __restrict__ dmerge,
                                    int nd, const DcCount* __restrict__ counts,
                                    const DcMerge* __restrict__ mg, double* __restrict__ Zg,
                                    double* __restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd) return;
  const int mi = dmerge[i] >> 1, half = dmerge[i] & 1;
  const DcCount c = counts[mi];
  const DcMerge m = mg[mi];
  const int K = m.n1 + m.n2;
  d[i].N = c.k;
  d[i].ldb = c.k > 0 ? c.k : 1;
  if (half == 0) {
    d[i].K = c.k1 + c.k2;
  } else {
    d[i].K = c.k - c.k1;
    d[i].A = Zg + m.zg_off + m.n1 + (int64_t)c.k1 * K;
    d[i].B = U + m.u_off + c.k1;
  }
}

// ---- plan -------------------------------------------------------------------
// Everything the levels need that depends only on the tile sizes (the merge
// trees, offsets, launch grids) is fixed on the host once, before any stream
// capture; what depends on the matrix (couplings, deflation counts, product
// sizes) is formed on the device.  A level's roots, U blocks and products are
// laid out for the worst case (no deflation: k = K), and the kernels skip the
// roots j >= k, so the whole solve is a fixed launch sequence with no host
// synchronisation: a HIP graph can hold it.
struct Node {
  int tile, off, size, height;
};

static int build_tree(int tile, int off, int size, std::vector<std::vector<Node>>& levels) {
  if (size <= 1) return 0;
  const int n1 = (size + 1) / 2;
  const int h1 = build_tree(tile, off, n1, levels);
  const int h2 = build_tree(tile, off + n1, size - n1, levels);
  const int h = std::max(h1, h2) + 1;
  if ((int)levels.size() < h) levels.resize(h);
  levels[h - 1].push_back(Node{tile, off, size, h});
  return h;
}

// merges whose worst-case product has at least this many 128x128 tiles go to
// the level's 128-tile grouped launch, the rest to its 64-tile launch
constexpr int kDc128Min = 512;

struct DcLevel {
  int merge_off, nm, maxK, roots;
  int desc_off[2], ndesc[2], blocks[2];  // [0]: 64x64 tiles, [1]: 128x128
};

struct DcPlan {
  std::vector<int64_t> ns;
  std::vector<double*> Z;
  std::vector<int64_t> base;
  int64_t N = 0;
  std::vector<DcLevel> levels;
  int nmerges = 0, ndescs = 0;
  DcMerge* d_merges = nullptr;  // all levels
  int* d_rprefix = nullptr;     // per level: nm + 1 root prefixes over K
  GemmDesc* d_descs = nullptr;  // per level and tile size
  int* d_dmerge = nullptr;      // descriptor -> merge index within its level
  int* d_gprefix = nullptr;     // grouped-launch block prefixes
  std::vector<int> rprefix_off, gprefix_off[2];
  void release() {
    void* p[] = {d_merges, d_rprefix, d_descs, d_dmerge, d_gprefix};
    for (void* q : p)
      if (q) (void)hipFree(q);
    *this = DcPlan();
  }
};

DcPlan* dc_plan_of(DcWork* w) {
  if (!w->plan) w->plan = new DcPlan();
  return w->plan;
}

static void dc_plan_delete(DcPlan* p) {
  if (!p) return;
  p->release();
  delete p;
}

static hipError_t dc_build_plan(hipStream_t st, DcWork* w, DcPlan& P, int ntiles,
                                const int64_t* ns, double* const* Z_t) {
  P.release();
  P.ns.assign(ns, ns + ntiles);
  P.Z.assign(Z_t, Z_t + ntiles);
  P.base.assign(ntiles + 1, 0);
  int64_t nn = 0;
  for (int t = 0; t < ntiles; ++t) {
    P.base[t + 1] = P.base[t] + ns[t];
    nn += ns[t] * ns[t];
  }
  P.N = P.base[ntiles];
  IMCOM_HIP_TRY(dc_reserve(w, P.N, nn));
  std::vector<std::vector<Node>> levels;
  for (int t = 0; t < ntiles; ++t) build_tree(t, 0, (int)ns[t], levels);
  std::vector<DcMerge> hm;
  std::vector<int> rpre, gpre, dmerge;
  std::vector<GemmDesc> descs;
  for (std::vector<Node>& lv : levels) {
    std::sort(lv.begin(), lv.end(), [&](const Node& a, const Node& b) {
      return P.base[a.tile] + a.off < P.base[b.tile] + b.off;
    });
    DcLevel L{};
    L.merge_off = (int)hm.size();
    L.nm = (int)lv.size();
    P.rprefix_off.push_back((int)rpre.size());
    rpre.push_back(0);
    int64_t zoff = 0;
    std::vector<GemmDesc> dl[2];
    std::vector<int> ml[2];
    for (int i = 0; i < L.nm; ++i) {
      const Node& nd = lv[i];
      DcMerge m{};
      m.n1 = (nd.size + 1) / 2;
      m.n2 = nd.size - m.n1;
      m.off = (int)(P.base[nd.tile] + nd.off);
      m.Z = Z_t[nd.tile];
      m.ldz = ns[nd.tile];
      m.loff = nd.off;
      m.zg_off = zoff;
      m.u_off = zoff;  // a K x K block holds the k x k U
      hm.push_back(m);
      rpre.push_back(rpre.back() + nd.size);
      const int K = nd.size;
      zoff += (int64_t)K * K;
      L.maxK = std::max(L.maxK, K);
      const int big = (int64_t)((K + 127) / 128) * ((K + 127) / 128) >= kDc128Min ? 1 : 0;
      // per half (sizes and the second half's operands set on the device)
      dl[big].push_back(GemmDesc{m.n1, K, K, 1.0, 0.0, w->Zg + m.zg_off, K, w->U + m.u_off, K,
                                 w->Zo + m.zg_off, K});
      ml[big].push_back(2 * i);
      dl[big].push_back(GemmDesc{m.n2, K, K, 1.0, 0.0, w->Zg + m.zg_off + m.n1, K, w->U + m.u_off,
                                 K, w->Zo + m.zg_off + m.n1, K});
      ml[big].push_back(2 * i + 1);
    }
    L.roots = rpre.back() - rpre[P.rprefix_off.back()];
    for (int b = 0; b < 2; ++b) {
      const int tile = b ? 128 : 64;
      L.desc_off[b] = (int)descs.size();
      L.ndesc[b] = (int)dl[b].size();
      P.gprefix_off[b].push_back((int)gpre.size());
      gpre.push_back(0);
      for (size_t q = 0; q < dl[b].size(); ++q) {
        descs.push_back(dl[b][q]);
        dmerge.push_back(ml[b][q]);
        const int M = dl[b][q].M, Nw = dl[b][q].N;  // worst case: N = K
        gpre.push_back(gpre.back() + ((M + tile - 1) / tile) * ((Nw + tile - 1) / tile));
      }
      L.blocks[b] = gpre.back();
    }
    P.levels.push_back(L);
  }
  // root prefixes are per level (relative to the level's first entry)
  for (size_t l = 0; l < P.levels.size(); ++l) {
    const int o = P.rprefix_off[l];
    const int first = rpre[o];
    for (int i = 0; i <= P.levels[l].nm; ++i) rpre[o + i] -= first;
  }
  P.nmerges = (int)hm.size();
  P.ndescs = (int)descs.size();
  auto up = [&](void** dst, const void* src, size_t bytes) -> hipError_t {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return e;
    return src ? hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, st) : hipSuccess;
  };
  IMCOM_HIP_TRY(up((void**)&P.d_merges, hm.data(), hm.size() * sizeof(DcMerge)));
  IMCOM_HIP_TRY(up((void**)&P.d_rprefix, rpre.data(), rpre.size() * sizeof(int)));
  IMCOM_HIP_TRY(up((void**)&P.d_descs, descs.empty() ? nullptr : descs.data(),
                   descs.size() * sizeof(GemmDesc)));
  IMCOM_HIP_TRY(up((void**)&P.d_dmerge, dmerge.empty() ? nullptr : dmerge.data(),
                   dmerge.size() * sizeof(int)));
  IMCOM_HIP_TRY(up((void**)&P.d_gprefix, gpre.data(), gpre.size() * sizeof(int)));
  // the host vectors must outlive the copies; this runs before any capture
  return hipStreamSynchronize(st);
}

static bool dc_plan_matches(const DcPlan& P, int ntiles, const int64_t* ns, double* const* Z_t) {
  if ((int)P.ns.size() != ntiles || P.d_merges == nullptr) return false;
  for (int t = 0; t < ntiles; ++t)
    if (P.ns[t] != ns[t] || P.Z[t] != Z_t[t]) return false;
  return true;
}

hipError_t dc_prepare(hipStream_t st, DcWork* w, int ntiles, const int64_t* ns, double* const* Z_t) {
  DcPlan& P = *dc_plan_of(w);
  if (dc_plan_matches(P, ntiles, ns, Z_t)) return hipSuccess;
  return dc_build_plan(st, w, P, ntiles, ns, Z_t);
}

// Tiles' tridiagonals (d_t: n_t, e_t: n_t - 1, device) solved together: their
// vectors are concatenated (with zero coupling between tiles), each tile keeps
// its own balanced merge tree, and merges of equal height across all tiles
// share the launches of a level.  lam_t (n_t) ascending, Z_t (n_t x n_t).
// The plan for these sizes and eigenvector buffers is built on first use
// (dc_prepare); inside a stream capture it must exist already.
hipError_t dc_solve_batched(hipStream_t st, DcWork* w, int ntiles, const int64_t* ns,
                            const double* const* d_t, const double* const* e_t,
                            double* const* lam_t, double* const* Z_t) {
  DcPlan& P = *dc_plan_of(w);
  if (!dc_plan_matches(P, ntiles, ns, Z_t)) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    IMCOM_HIP_TRY(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return hipErrorStreamCaptureUnsupported;
    IMCOM_HIP_TRY(dc_build_plan(st, w, P, ntiles, ns, Z_t));
  }
  const int64_t N = P.N;
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(dc_fill_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, 0.0, w->ecat);
  for (int t = 0; t < ntiles; ++t) {
    IMCOM_HIP_TRY(hipMemcpyAsync(w->dcat + P.base[t], d_t[t], ns[t] * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
    if (ns[t] > 1)
      IMCOM_HIP_TRY(hipMemcpyAsync(w->ecat + P.base[t], e_t[t], (ns[t] - 1) * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(dc_identity_kernel, dim3((unsigned)ns[t]), dim3(256), 0, st, ns[t], Z_t[t]);
  }
  hipLaunchKernelGGL(dc_scale_kernel, dim3(1), dim3(1024), 0, st, N, w->dcat, w->ecat, w->scale);
  hipLaunchKernelGGL(dc_leaves_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, w->dcat, w->ecat,
                     w->scale, w->D);
  if (P.nmerges > 0)
    hipLaunchKernelGGL(dc_rho_kernel, dim3(ceil_div(P.nmerges, 256)), dim3(256), 0, st, P.d_merges,
                       P.nmerges, w->ecat, w->scale);
  for (size_t l = 0; l < P.levels.size(); ++l) {
    const DcLevel& L = P.levels[l];
    const int nm = L.nm;
    const DcMerge* mg = P.d_merges + L.merge_off;
    DcCount* counts = w->d_counts + L.merge_off;
    const int* rpre = P.d_rprefix + P.rprefix_off[l];
    hipLaunchKernelGGL(dc_sort_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, mg, nm, w->D,
                       (const double*)nullptr, w->ds, w->zs, w->cs);
    if (nm > 0)
      hipLaunchKernelGGL(dc_deflate_kernel, dim3(nm), dim3(64), 0, st, mg, nm, w->ds,
                       w->zs, w->cs, w->dl, w->wz, w->ndcol, w->ndpos, w->dfd, w->dfcol, w->rp,
                       w->rn, w->rc, w->rs, counts);
    hipLaunchKernelGGL(dc_rotate_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, mg, nm,
                       (const DcCount*)counts, w->rp, w->rn, w->rc, w->rs, (double*)nullptr);
    hipLaunchKernelGGL(dc_secular_kernel, dim3(ceil_div(L.roots, 4)), dim3(256), 0, st, mg,
                       (const DcCount*)counts, rpre, nm, L.roots, w->dl, w->wz, w->org, w->tau,
                       w->lam);
    hipLaunchKernelGGL(dc_zhat_kernel, dim3(ceil_div(L.roots, 64)), dim3(64), 0, st, mg,
                       (const DcCount*)counts, rpre, nm, L.roots, w->dl, w->wz, w->org, w->tau,
                       w->zhat);
    hipLaunchKernelGGL(dc_vec_kernel, dim3(L.roots), dim3(256), 0, st, mg, (const DcCount*)counts,
                       rpre, nm, w->dl, w->org, w->tau, w->zhat, (const int*)w->ndpos, w->U);
    const int tpb = std::min(256, std::max(64, (L.maxK + 63) / 64 * 64));
    hipLaunchKernelGGL(dc_gather_kernel, dim3((unsigned)N), dim3(tpb), 0, st, N, mg, nm,
                       (const DcCount*)counts, w->ndcol, w->ndpos, w->dfcol, (const double*)nullptr,
                       w->Zg);
    // Zo_m = Zg_m(:, 0:k) U_m, with k from the device counts
    for (int b = 0; b < 2; ++b) {
      if (L.ndesc[b] == 0) continue;
      GemmDesc* d = P.d_descs + L.desc_off[b];
      hipLaunchKernelGGL(dc_fix_descs_kernel, dim3(ceil_div(L.ndesc[b], 64)), dim3(64), 0, st, d,
                         P.d_dmerge + L.desc_off[b], L.ndesc[b], (const DcCount*)counts, mg, w->Zg,
                         w->U);
      IMCOM_HIP_TRY(dgemm_grouped(st, false, false, d, P.d_gprefix + P.gprefix_off[b][l],
                                  L.ndesc[b], L.blocks[b], b ? 128 : 64));
    }
    hipLaunchKernelGGL(dc_place_kernel, dim3((unsigned)N), dim3(tpb), 0, st, N, mg, nm,
                       (const DcCount*)counts, w->lam, w->dfd, w->Zg, w->Zo, w->D, (double*)nullptr);
  }
  hipLaunchKernelGGL(dc_unscale_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, st, N, w->D, w->scale,
                     w->lam);
  for (int t = 0; t < ntiles; ++t)
    IMCOM_HIP_TRY(hipMemcpyAsync(lam_t[t], w->lam + P.base[t], ns[t] * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
  return hipGetLastError();
}

hipError_t dc_solve(hipStream_t st, DcWork* w, int64_t n, const double* d, const double* e,
                    double* lam, double* Z) {
  return dc_solve_batched(st, w, 1, &n, &d, &e, &lam, &Z);
}

}  // namespace imcom
