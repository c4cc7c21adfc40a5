// Device-resident setup of the smoothed-aggregation hierarchy
// (src/solver/iterative/amg.rs:108-341, the CSR products of
// src/solver/iterative/csr.rs:7-189).  Part of the PCG translation unit:
// included by pcg_gpu.hip inside its anonymous namespace, after DBuf, DevMat
// and make_mat.
//
// Every level matrix, prolongator and product stays in HBM as scalar CSR
// (int64 row pointers, int32 columns sorted per row, FP64 values).  The GPU
// does the full symmetric CSR of the caller's lower triangle, the diagonal,
// the block strengths (squared Frobenius norms), the tentative prolongator
// (a modified Gram-Schmidt QR per aggregate), the smoothing
// P = P~ - (omega / d) A P~, the transpose and the Galerkin products, and
// the BSR operators of the V-cycle.  The host does what is sequential and
// small: the three aggregation passes over the strong lists (integers
// downloaded per level), the spectral radius bookkeeping, and the dense
// LDL^T of the coarsest matrix.
//
// Every floating-point value is formed by the host setup's operations in
// their order (amg_setup.cpp, itself the reference's order): sums run
// sequentially from zero in the host loops' order and the kernels are
// compiled without contraction, so the hierarchy is bit for bit the host
// setup's.  A product row is one wave: its column set is found in an LDS hash
// table and sorted in LDS (bitonic), and its values accumulate in LDS in the
// order of the left factor's entries, one entry at a time (lanes over the
// right factor's row, a barrier between entries).  Rows too long for the LDS
// tables make the setup throw DeviceSetupUnsupported, and the caller then runs
// the host setup.

struct DeviceSetupUnsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct DCsr {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  DBuf ptr, idx, val;  // int64 [nrows + 1], int32 [nnz], double [nnz]
};

constexpr int SCAN_T = 256, SCAN_E = 8;
constexpr int WAVE = 64;

// ---------------------------------------------------------------- scans

__global__ __launch_bounds__(SCAN_T) void k_scan_local(const int64_t* in, int64_t* out, int64_t n, int64_t* bsum) {
  __shared__ int64_t sh[SCAN_T];
  const int t = threadIdx.x;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * (SCAN_T * SCAN_E) + static_cast<int64_t>(t) * SCAN_E;
  int64_t v[SCAN_E];
  int64_t s = 0;
  for (int e = 0; e < SCAN_E; ++e) {
    v[e] = base + e < n ? in[base + e] : 0;
    s += v[e];
  }
  sh[t] = s;
  __syncthreads();
  for (int o = 1; o < SCAN_T; o <<= 1) {
    const int64_t add = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  int64_t run = sh[t] - s;
  for (int e = 0; e < SCAN_E; ++e)
    if (base + e < n) {
      out[base + e] = run;
      run += v[e];
    }
  if (t == SCAN_T - 1) bsum[blockIdx.x] = sh[t];
}

__global__ void k_scan_add(int64_t* out, int64_t n, const int64_t* boff) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] += boff[i / (SCAN_T * SCAN_E)];
}

// exclusive prefix sum of n values (in place allowed); returns the total
int64_t scan_exclusive(const int64_t* in, int64_t* out, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  int64_t last_in = 0;
  PHIPCHK(hipMemcpyAsync(&last_in, in + n - 1, 8, hipMemcpyDeviceToHost, s));
  PHIPCHK(hipStreamSynchronize(s));
  const int64_t per = SCAN_T * SCAN_E;
  const int64_t nb = (n + per - 1) / per;
  DBuf bsum, boff;
  bsum.alloc(nb * 8);
  boff.alloc(nb * 8);
  hipLaunchKernelGGL(k_scan_local, dim3(static_cast<unsigned>(nb)), dim3(SCAN_T), 0, s, in, out, n, bsum.as<int64_t>());
  PHIPCHK(hipGetLastError());
  if (nb > 1) {
    scan_exclusive(bsum.as<int64_t>(), boff.as<int64_t>(), nb, s);
    hipLaunchKernelGGL(k_scan_add, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, out, n,
                       boff.as<int64_t>());
    PHIPCHK(hipGetLastError());
  }
  int64_t last_out = 0;
  PHIPCHK(hipMemcpyAsync(&last_out, out + n - 1, 8, hipMemcpyDeviceToHost, s));
  PHIPCHK(hipStreamSynchronize(s));
  return last_out + last_in;
}

// counts [n] -> ptr [n + 1]; returns the total
int64_t counts_to_ptr(const DBuf& counts, DBuf& ptr, int64_t n, hipStream_t s) {
  ptr.alloc((n + 1) * 8);
  const int64_t tot = n > 0 ? scan_exclusive(counts.as<int64_t>(), ptr.as<int64_t>(), n, s) : 0;
  PHIPCHK(hipMemcpyAsync(ptr.as<int64_t>() + n, &tot, 8, hipMemcpyHostToDevice, s));
  PHIPCHK(hipStreamSynchronize(s));
  return tot;
}

unsigned nblocks(int64_t n, int t = 256) { return static_cast<unsigned>(std::max<int64_t>(1, (n + t - 1) / t)); }

template <typename T>
std::vector<T> download_vec(const DBuf& b, int64_t n, hipStream_t s) {
  std::vector<T> v(n);
  if (n) PHIPCHK(hipMemcpyAsync(v.data(), b.p, n * sizeof(T), hipMemcpyDeviceToHost, s));
  PHIPCHK(hipStreamSynchronize(s));
  return v;
}

// ------------------------------------------------- sorting row segments

// Sorts the entries of each row by column (keys int32, values double), one
// wave per row: bitonic sort of the row padded to a power of two in LDS.
__global__ __launch_bounds__(WAVE) void k_sort_rows(int64_t nrows, const int64_t* __restrict__ ptr, int32_t* idx,
                                                    double* val, int cap) {
  extern __shared__ unsigned char smem[];
  int32_t* key = reinterpret_cast<int32_t*>(smem);
  double* v = reinterpret_cast<double*>(smem + static_cast<size_t>(cap) * 4);
  const int64_t r = blockIdx.x;
  if (r >= nrows) return;
  const int64_t b = ptr[r];
  const int m = static_cast<int>(ptr[r + 1] - b);
  if (m > cap) return;  // a long row: k_sort_long sorts it
  int np = 1;
  while (np < m) np <<= 1;
  for (int i = threadIdx.x; i < np; i += WAVE) {
    key[i] = i < m ? idx[b + i] : INT32_MAX;
    v[i] = i < m ? val[b + i] : 0.0;
  }
  __syncthreads();
  for (int k = 2; k <= np; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np; i += WAVE) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          if ((key[i] > key[l]) == up) {
            const int32_t tk = key[i];
            key[i] = key[l];
            key[l] = tk;
            const double tv = v[i];
            v[i] = v[l];
            v[l] = tv;
          }
        }
      }
      __syncthreads();
    }
  for (int i = threadIdx.x; i < m; i += WAVE) {
    idx[b + i] = key[i];
    val[b + i] = v[i];
  }
}

// The rows longer than the LDS sort takes, one workgroup each: the same
// bitonic sort in global scratch (np entries per row at soff[j]).
__global__ __launch_bounds__(256) void k_sort_long(const int64_t* __restrict__ rows, const int64_t* __restrict__ soff,
                                                   const int64_t* __restrict__ ptr, int32_t* idx, double* val,
                                                   int32_t* skey, double* sval) {
  const int64_t r = rows[blockIdx.x];
  const int64_t b = ptr[r];
  const int64_t m = ptr[r + 1] - b;
  int64_t np = 1;
  while (np < m) np <<= 1;
  int32_t* key = skey + soff[blockIdx.x];
  double* v = sval + soff[blockIdx.x];
  for (int64_t i = threadIdx.x; i < np; i += 256) {
    key[i] = i < m ? idx[b + i] : INT32_MAX;
    v[i] = i < m ? val[b + i] : 0.0;
  }
  __syncthreads();
  for (int64_t k = 2; k <= np; k <<= 1)
    for (int64_t j = k >> 1; j > 0; j >>= 1) {
      for (int64_t i = threadIdx.x; i < np; i += 256) {
        const int64_t l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          if ((key[i] > key[l]) == up) {
            const int32_t tk = key[i];
            key[i] = key[l];
            key[l] = tk;
            const double tv = v[i];
            v[i] = v[l];
            v[l] = tv;
          }
        }
      }
      __threadfence_block();
      __syncthreads();
    }
  for (int64_t i = threadIdx.x; i < m; i += 256) {
    idx[b + i] = key[i];
    val[b + i] = v[i];
  }
}

__global__ void k_row_len_max(int64_t nrows, const int64_t* ptr, unsigned long long* mx) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < nrows) atomicMax(mx, static_cast<unsigned long long>(ptr[r + 1] - ptr[r]));
}

int next_pow2(int64_t m) {
  int64_t p = 1;
  while (p < m) p <<= 1;
  return static_cast<int>(p);
}

constexpr int LDS_SORT_MAX = 4096;  // entries per row the LDS sort takes

void sort_rows(DCsr& a, hipStream_t s) {
  if (a.nrows == 0) return;
  DBuf mx;
  mx.alloc(8);
  PHIPCHK(hipMemsetAsync(mx.p, 0, 8, s));
  hipLaunchKernelGGL(k_row_len_max, dim3(nblocks(a.nrows)), dim3(256), 0, s, a.nrows, a.ptr.as<int64_t>(),
                     mx.as<unsigned long long>());
  unsigned long long m = 0;
  PHIPCHK(hipMemcpyAsync(&m, mx.p, 8, hipMemcpyDeviceToHost, s));
  PHIPCHK(hipStreamSynchronize(s));
  const int cap = std::max(1, next_pow2(static_cast<int64_t>(std::min<unsigned long long>(m, LDS_SORT_MAX))));
  hipLaunchKernelGGL(k_sort_rows, dim3(static_cast<unsigned>(a.nrows)), dim3(WAVE), static_cast<size_t>(cap) * 12, s,
                     a.nrows, a.ptr.as<int64_t>(), a.idx.as<int32_t>(), a.val.as<double>(), cap);
  PHIPCHK(hipGetLastError());
  if (m <= static_cast<unsigned long long>(cap)) return;
  // rows longer than the LDS sort takes (the dense coarse levels): sorted in
  // global scratch, one workgroup per row
  const std::vector<int64_t> hp = download_vec<int64_t>(a.ptr, a.nrows + 1, s);
  std::vector<int64_t> rows, soff(1, 0);
  for (int64_t r = 0; r < a.nrows; ++r)
    if (hp[r + 1] - hp[r] > cap) {
      rows.push_back(r);
      soff.push_back(soff.back() + next_pow2(hp[r + 1] - hp[r]));
    }
  DBuf drows = upload(rows), dsoff = upload(soff), skey, sval;
  skey.alloc(soff.back() * 4);
  sval.alloc(soff.back() * 8);
  hipLaunchKernelGGL(k_sort_long, dim3(static_cast<unsigned>(rows.size())), dim3(256), 0, s, drows.as<int64_t>(),
                     dsoff.as<int64_t>(), a.ptr.as<int64_t>(), a.idx.as<int32_t>(), a.val.as<double>(),
                     skey.as<int32_t>(), sval.as<double>());
  PHIPCHK(hipGetLastError());
  PHIPCHK(hipStreamSynchronize(s));
}

// --------------------------------------- full CSR from the lower triangle

__global__ void k_lower_count(int64_t n, const int64_t* colptr, const int64_t* rowidx, unsigned long long* cnt) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
    const int64_t r = rowidx[p];
    atomicAdd(&cnt[r], 1ull);
    if (r != c) atomicAdd(&cnt[c], 1ull);
  }
}

__global__ void k_lower_fill(int64_t n, const int64_t* colptr, const int64_t* rowidx, const double* lv,
                             const int64_t* ptr, unsigned long long* fill, int32_t* idx, double* val) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int64_t p = colptr[c]; p < colptr[c + 1]; ++p) {
    const int64_t r = rowidx[p];
    const int64_t a = ptr[r] + static_cast<int64_t>(atomicAdd(&fill[r], 1ull));
    idx[a] = static_cast<int32_t>(c);
    val[a] = lv[p];
    if (r != c) {
      const int64_t b = ptr[c] + static_cast<int64_t>(atomicAdd(&fill[c], 1ull));
      idx[b] = static_cast<int32_t>(r);
      val[b] = lv[p];
    }
  }
}

// csr_from_lower (amg_setup.cpp; src/solver/iterative/csr.rs:21-60) on the device
DCsr dev_csr_from_lower(int64_t n, const int64_t* colptr, const int64_t* rowidx, const double* val, hipStream_t s) {
  const int64_t nl = colptr[n];
  DBuf dcp, dri, dv;
  dcp.alloc((n + 1) * 8);
  dri.alloc(std::max<int64_t>(nl, 1) * 8);
  dv.alloc(std::max<int64_t>(nl, 1) * 8);
  PHIPCHK(hipMemcpy(dcp.p, colptr, (n + 1) * 8, hipMemcpyHostToDevice));
  if (nl) {
    PHIPCHK(hipMemcpy(dri.p, rowidx, nl * 8, hipMemcpyHostToDevice));
    PHIPCHK(hipMemcpy(dv.p, val, nl * 8, hipMemcpyHostToDevice));
  }
  DCsr a;
  a.nrows = a.ncols = n;
  DBuf cnt;
  cnt.alloc(std::max<int64_t>(n, 1) * 8);
  PHIPCHK(hipMemsetAsync(cnt.p, 0, std::max<int64_t>(n, 1) * 8, s));
  hipLaunchKernelGGL(k_lower_count, dim3(nblocks(n)), dim3(256), 0, s, n, dcp.as<int64_t>(), dri.as<int64_t>(),
                     cnt.as<unsigned long long>());
  PHIPCHK(hipGetLastError());
  a.nnz = counts_to_ptr(cnt, a.ptr, n, s);
  a.idx.alloc(std::max<int64_t>(a.nnz, 1) * 4);
  a.val.alloc(std::max<int64_t>(a.nnz, 1) * 8);
  PHIPCHK(hipMemsetAsync(cnt.p, 0, std::max<int64_t>(n, 1) * 8, s));
  hipLaunchKernelGGL(k_lower_fill, dim3(nblocks(n)), dim3(256), 0, s, n, dcp.as<int64_t>(), dri.as<int64_t>(),
                     dv.as<double>(), a.ptr.as<int64_t>(), cnt.as<unsigned long long>(), a.idx.as<int32_t>(),
                     a.val.as<double>());
  PHIPCHK(hipGetLastError());
  sort_rows(a, s);
  PHIPCHK(hipStreamSynchronize(s));
  return a;
}

Csr download(const DCsr& a, hipStream_t s) {
  Csr m;
  m.nrows = a.nrows;
  m.ncols = a.ncols;
  m.ptr.resize(a.nrows + 1);
  m.idx.resize(a.nnz);
  m.val.resize(a.nnz);
  PHIPCHK(hipMemcpyAsync(m.ptr.data(), a.ptr.p, (a.nrows + 1) * 8, hipMemcpyDeviceToHost, s));
  if (a.nnz) {
    PHIPCHK(hipMemcpyAsync(m.idx.data(), a.idx.p, a.nnz * 4, hipMemcpyDeviceToHost, s));
    PHIPCHK(hipMemcpyAsync(m.val.data(), a.val.p, a.nnz * 8, hipMemcpyDeviceToHost, s));
  }
  PHIPCHK(hipStreamSynchronize(s));
  return m;
}


// ----------------------------------------------------------- diagonal

__global__ void k_diag_check(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, double* diag,
                             int32_t* bad) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t lo = ptr[i], hi = ptr[i + 1];
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (idx[mid] < i)
      lo = mid + 1;
    else
      hi = mid;
  }
  const double d = (lo < ptr[i + 1] && idx[lo] == i) ? val[lo] : 0.0;
  diag[i] = d;
  if (!(d > 0.0)) atomicOr(bad, 1);
}

// ----------------------------------------------------------- strength

__global__ void k_blkof(int64_t nb, const int64_t* bptr, int32_t* blkof) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  for (int64_t i = bptr[b]; i < bptr[b + 1]; ++i) blkof[i] = static_cast<int32_t>(b);
}

// The neighbour blocks of block b in the order of first appearance in its
// rows, with the squared Frobenius norm of each coupling block and of the
// diagonal block (amg_aggregate, amg_setup.cpp), one thread per block.  The
// lists live in scratch at the offset of the block's first entry.
__global__ void k_strength_lists(int64_t nb, const int64_t* __restrict__ bptr, const int64_t* __restrict__ ptr,
                                 const int32_t* __restrict__ idx, const double* __restrict__ val,
                                 const int32_t* __restrict__ blkof, int32_t* lid, double* lacc, int32_t* nlist,
                                 double* dn) {
#pragma clang fp contract(off)
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int64_t base = ptr[bptr[b]];
  int nl = 0, lastc = -1, lastj = 0;
  double dnb = 0.0;
  for (int64_t i = bptr[b]; i < bptr[b + 1]; ++i)
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      const int c = blkof[idx[p]];
      const double v = val[p];
      if (c == b) {
        dnb += v * v;
        continue;
      }
      int j = -1;
      if (c == lastc) {
        j = lastj;
      } else {
        for (int q = 0; q < nl; ++q)
          if (lid[base + q] == c) {
            j = q;
            break;
          }
        if (j < 0) {
          j = nl++;
          lid[base + j] = c;
          lacc[base + j] = 0.0;
        }
        lastc = c;
        lastj = j;
      }
      lacc[base + j] += v * v;
    }
  nlist[b] = nl;
  dn[b] = dnb;
}

// strong neighbours: acc > theta^2 sqrt(dn_b dn_c); pass 0 counts, pass 1 writes
__global__ void k_strong(int64_t nb, const int64_t* __restrict__ bptr, const int64_t* __restrict__ ptr,
                         const int32_t* __restrict__ lid, const double* __restrict__ lacc,
                         const int32_t* __restrict__ nlist, const double* __restrict__ dn, double theta,
                         int64_t* cnt, const int64_t* sptr, int32_t* sidx) {
#pragma clang fp contract(off)
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int64_t base = ptr[bptr[b]];
  int64_t k = 0;
  for (int q = 0; q < nlist[b]; ++q) {
    const int c = lid[base + q];
    if (lacc[base + q] > theta * theta * sqrt(dn[b] * dn[c])) {
      if (sidx) sidx[sptr[b] + k] = c;
      ++k;
    }
  }
  if (cnt) cnt[b] = k;
}

// ---------------------------------------------------------- tentative

// One aggregate per thread: the modified Gram-Schmidt QR (two passes) of the
// near-null space rows of its DOFs, columns below 1e-10 x the largest norm
// dropped (tentative(), amg_setup.cpp; src/solver/iterative/amg.rs:184-257).
// q and kept live in scratch at the aggregate's DOF offset.
__global__ void k_tentative(int64_t nagg, const int64_t* __restrict__ gptr, const int32_t* __restrict__ gblk,
                            const int64_t* __restrict__ bptr, const int64_t* __restrict__ doff,
                            const double* __restrict__ b, int nb, double* qs, double* ks, int32_t* nk_out,
                            double* rmat, double* ptv) {
#pragma clang fp contract(off)
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nagg) return;
  const int64_t nd = doff[g + 1] - doff[g];
  double* q = qs + static_cast<int64_t>(nb) * doff[g];
  double* kept = ks + static_cast<int64_t>(nb + 1) * doff[g];
  {
    int64_t li = 0;
    for (int64_t m = gptr[g]; m < gptr[g + 1]; ++m)
      for (int64_t i = bptr[gblk[m]]; i < bptr[gblk[m] + 1]; ++i, ++li)
        for (int c = 0; c < nb; ++c) q[c * nd + li] = b[i * nb + c];
  }
  const int64_t rr = nb < nd ? nb : nd;
  double r[36];
  for (int i = 0; i < 36; ++i) r[i] = 0.0;
  for (int64_t i = 0; i < static_cast<int64_t>(nb) * nd + nd; ++i) kept[i] = 0.0;
  int64_t nk = 0;
  double scale = 0.0;
  for (int c = 0; c < nb; ++c) {
    double s = 0.0;
    for (int64_t i = 0; i < nd; ++i) s += q[c * nd + i] * q[c * nd + i];
    scale = fmax(scale, s);
  }
  scale = sqrt(scale);
  for (int c = 0; c < nb; ++c) {
    double* v = q + c * nd;
    for (int pass = 0; pass < 2; ++pass)
      for (int64_t k = 0; k < nk; ++k) {
        const double* qk = kept + k * nd;
        double s = 0.0;
        for (int64_t i = 0; i < nd; ++i) s += qk[i] * v[i];
        if (pass == 0)
          r[k * nb + c] = s;
        else
          r[k * nb + c] += s;
        for (int64_t i = 0; i < nd; ++i) v[i] -= s * qk[i];
      }
    double norm = 0.0;
    for (int64_t i = 0; i < nd; ++i) norm += v[i] * v[i];
    norm = sqrt(norm);
    if (nk < nd && norm > 1e-10 * fmax(scale, DBL_MIN)) {
      r[nk * nb + c] = norm;
      for (int64_t i = 0; i < nd; ++i) kept[nk * nd + i] = v[i] / norm;
      ++nk;
    }
  }
  if (nk == 0) {
    const double s = 1.0 / sqrt(static_cast<double>(nd));
    for (int64_t i = 0; i < nd; ++i) kept[i] = s;
    nk = 1;
  }
  nk_out[g] = static_cast<int32_t>(nk);
  for (int64_t k = 0; k < nk; ++k)
    for (int c = 0; c < nb; ++c) rmat[(g * nb + k) * nb + c] = k < rr ? r[k * nb + c] : 0.0;
  int64_t li = 0;
  for (int64_t m = gptr[g]; m < gptr[g + 1]; ++m)
    for (int64_t i = bptr[gblk[m]]; i < bptr[gblk[m] + 1]; ++i, ++li)
      for (int64_t k = 0; k < nb; ++k) ptv[i * nb + k] = k < nk ? kept[k * nd + li] : 0.0;
}

// P~ rows: the nonzero kept values of each DOF's aggregate, columns cptr[g] + k
__global__ void k_ptent(int64_t n, const int32_t* __restrict__ blkof, const int32_t* __restrict__ agg,
                        const int32_t* __restrict__ nk, const int64_t* __restrict__ cptr,
                        const double* __restrict__ ptv, int nb, int64_t* cnt, const int64_t* ptr, int32_t* idx,
                        double* val) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = agg[blkof[i]];
  int64_t k = 0;
  for (int j = 0; j < nk[g]; ++j) {
    const double x = ptv[i * nb + j];
    if (x != 0.0) {
      if (idx) {
        idx[ptr[i] + k] = static_cast<int32_t>(cptr[g] + j);
        val[ptr[i] + k] = x;
      }
      ++k;
    }
  }
  if (cnt) cnt[i] = k;
}

// the coarse near-null space: row cptr[g] + k of the aggregate's R factor
__global__ void k_coarse_b(int64_t nagg, const int32_t* nk, const int64_t* cptr, const double* rmat, int nb,
                           double* bc) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= nagg) return;
  for (int k = 0; k < nk[g]; ++k)
    for (int c = 0; c < nb; ++c) bc[(cptr[g] + k) * nb + c] = rmat[(g * nb + k) * nb + c];
}

// --------------------------------------------------------- products

__device__ __forceinline__ uint32_t hash_key(int32_t k, int mask) {
  return (static_cast<uint32_t>(k) * 0x9E3779B1u) & static_cast<uint32_t>(mask);
}

// candidate columns of each row of A B (an upper bound of its length)
__global__ void k_spgemm_cand(int64_t nrows, const int64_t* __restrict__ ap, const int32_t* __restrict__ ai,
                              const int64_t* __restrict__ bp, unsigned long long* mx) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  int64_t s = 0;
  for (int64_t p = ap[i]; p < ap[i + 1]; ++p) s += bp[ai[p] + 1] - bp[ai[p]];
  atomicMax(mx, static_cast<unsigned long long>(s));
}

// Column set of each row of A B: inserted into an LDS hash table of hs
// slots (lanes over A's entries), counted (pass 0) or compacted, sorted
// (bitonic) and written (pass 1).  A row with more than hs / 2 columns sets
// *overflow and stops inserting (the table always keeps free slots, so every
// probe ends).
__global__ __launch_bounds__(WAVE) void k_spgemm_sym(int64_t nrows, const int64_t* __restrict__ ap,
                                                     const int32_t* __restrict__ ai, const int64_t* __restrict__ bp,
                                                     const int32_t* __restrict__ bi, int hs, int64_t* cnt,
                                                     const int64_t* cptr, int32_t* ci, int32_t* overflow) {
  extern __shared__ unsigned char smem[];
  int32_t* tab = reinterpret_cast<int32_t*>(smem);
  int32_t* dense = tab + hs;
  __shared__ int32_t n_unique, n_dense;
  const int64_t i = blockIdx.x;
  if (i >= nrows) return;
  const int mask = hs - 1;
  const int limit = hs / 2;
  for (int t = threadIdx.x; t < hs; t += WAVE) tab[t] = -1;
  if (threadIdx.x == 0) {
    n_unique = 0;
    n_dense = 0;
  }
  __syncthreads();
  bool full = false;
  for (int64_t p = ap[i] + threadIdx.x; p < ap[i + 1] && !full; p += WAVE) {
    const int32_t k = ai[p];
    for (int64_t q = bp[k]; q < bp[k + 1]; ++q) {
      if (*reinterpret_cast<volatile int32_t*>(&n_unique) >= limit) {
        full = true;
        break;
      }
      const int32_t key = bi[q];
      uint32_t h = hash_key(key, mask);
      while (true) {
        const int32_t old = atomicCAS(&tab[h], -1, key);
        if (old == -1) {
          atomicAdd(&n_unique, 1);
          break;
        }
        if (old == key) break;
        h = (h + 1) & static_cast<uint32_t>(mask);
      }
    }
  }
  if (full) atomicOr(overflow, 1);
  __syncthreads();
  if (!ci) {
    if (threadIdx.x == 0) cnt[i] = n_unique;
    return;
  }
  const int m = n_unique;
  for (int t = threadIdx.x; t < hs; t += WAVE)
    if (tab[t] >= 0) dense[atomicAdd(&n_dense, 1)] = tab[t];
  __syncthreads();
  int np = 1;
  while (np < m) np <<= 1;
  for (int t = m + threadIdx.x; t < np; t += WAVE) dense[t] = INT32_MAX;
  __syncthreads();
  for (int k = 2; k <= np; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < np; t += WAVE) {
        const int l = t ^ j;
        if (l > t) {
          const bool up = (t & k) == 0;
          if ((dense[t] > dense[l]) == up) {
            const int32_t x = dense[t];
            dense[t] = dense[l];
            dense[l] = x;
          }
        }
      }
      __syncthreads();
    }
  const int64_t b = cptr[i];
  for (int t = threadIdx.x; t < m; t += WAVE) ci[b + t] = dense[t];
}

// Values of A B on its pattern: acc[c] += a_ip b_kc over A's entries p in
// order (csr_mul, amg_setup.cpp), lanes over B's row k, a barrier between
// entries so that each slot sums in that order.  The right factor's rows are
// staged in LDS for up to 64 consecutive entries p at a time (sb entries of
// room), so their global loads are in flight together; a row of B longer
// than the stage is read from global memory directly.
__global__ __launch_bounds__(WAVE) void k_spgemm_num(int64_t nrows, const int64_t* __restrict__ ap,
                                                     const int32_t* __restrict__ ai, const double* __restrict__ av,
                                                     const int64_t* __restrict__ bp, const int32_t* __restrict__ bi,
                                                     const double* __restrict__ bv, const int64_t* __restrict__ cptr,
                                                     const int32_t* __restrict__ ci, double* cv, int cap, int sb) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem[];
  double* acc = reinterpret_cast<double*>(smem);
  double* sval = acc + cap;
  double* sa = sval + sb;
  int32_t* cols = reinterpret_cast<int32_t*>(sa + WAVE);
  int32_t* skey = cols + cap;
  int32_t* soff = skey + sb;  // WAVE + 1 entries
  const int64_t i = blockIdx.x;
  if (i >= nrows) return;
  const int lane = threadIdx.x;
  const int64_t cb = cptr[i];
  const int m = static_cast<int>(cptr[i + 1] - cb);
  for (int t = lane; t < m; t += WAVE) {
    cols[t] = ci[cb + t];
    acc[t] = 0.0;
  }
  __syncthreads();
  auto update = [&](int32_t key, double a, double b) {
    int lo = 0, hi = m;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cols[mid] < key)
        lo = mid + 1;
      else
        hi = mid;
    }
    acc[lo] += a * b;
  };
  const int64_t pe = ap[i + 1];
  if (sb == 0) {
    // long rows of B: every entry p straight from global memory
    for (int64_t p = ap[i]; p < pe; ++p) {
      const int32_t k = ai[p];
      const double a = av[p];
      for (int64_t q = bp[k] + lane; q < bp[k + 1]; q += WAVE) update(bi[q], a, bv[q]);
      __syncthreads();
    }
  }
  for (int64_t p0 = sb == 0 ? pe : ap[i]; p0 < pe;) {
    // the B rows of entries p0, p0 + 1, ...: lengths, an inclusive scan over the wave
    const int64_t p = p0 + lane;
    int32_t k = 0;
    int64_t len = 0, start = 0;
    double a = 0.0;
    if (p < pe) {
      k = ai[p];
      a = av[p];
      start = bp[k];
      len = bp[k + 1] - start;
    }
    int64_t incl = len;
    for (int o = 1; o < WAVE; o <<= 1) {
      const int64_t t = __shfl_up(incl, o, WAVE);
      if (lane >= o) incl += t;
    }
    const bool fits = p < pe && incl <= sb;
    const int np = __popcll(__ballot(fits));
    if (np == 0) {
      // B's row of p0 alone exceeds the stage: straight from global memory
      const int64_t s0 = __shfl(start, 0, WAVE), l0 = __shfl(len, 0, WAVE);
      const double a0 = __shfl(a, 0, WAVE);
      for (int64_t q = s0 + lane; q < s0 + l0; q += WAVE) update(bi[q], a0, bv[q]);
      __syncthreads();
      p0 += 1;
      continue;
    }
    if (lane < np) {
      soff[lane] = static_cast<int32_t>(incl - len);
      sa[lane] = a;
    }
    if (lane == np - 1) soff[np] = static_cast<int32_t>(incl);
    __syncthreads();
    const int tot = soff[np];
    // copy the staged rows: entry t belongs to the last j with soff[j] <= t
    for (int t = lane; t < tot; t += WAVE) {
      int lo = 0, hi = np;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (soff[mid] <= t)
          lo = mid;
        else
          hi = mid;
      }
      const int64_t src = bp[ai[p0 + lo]] + (t - soff[lo]);
      skey[t] = bi[src];
      sval[t] = bv[src];
    }
    __syncthreads();
    for (int j = 0; j < np; ++j) {
      const double aj = sa[j];
      for (int t = soff[j] + lane; t < soff[j + 1]; t += WAVE) update(skey[t], aj, sval[t]);
      __syncthreads();
    }
    p0 += np;
  }
  for (int t = lane; t < m; t += WAVE) cv[cb + t] = acc[t];
}

constexpr int LDS_HASH_MAX = 8192;  // hash slots of the symbolic product (32 KB of LDS)

// C = A B (csr_mul on the device)
DCsr dev_spgemm(const DCsr& a, const DCsr& b, hipStream_t s) {
  if (a.ncols != b.nrows) throw std::runtime_error("internal: spgemm shape mismatch");
  DCsr c;
  c.nrows = a.nrows;
  c.ncols = b.ncols;
  DBuf mx, ovf;
  mx.alloc(8);
  ovf.alloc(4);
  PHIPCHK(hipMemsetAsync(mx.p, 0, 8, s));
  PHIPCHK(hipMemsetAsync(ovf.p, 0, 4, s));
  if (a.nrows)
    hipLaunchKernelGGL(k_spgemm_cand, dim3(nblocks(a.nrows)), dim3(256), 0, s, a.nrows, a.ptr.as<int64_t>(),
                       a.idx.as<int32_t>(), b.ptr.as<int64_t>(), mx.as<unsigned long long>());
  unsigned long long cand = 0;
  PHIPCHK(hipMemcpyAsync(&cand, mx.p, 8, hipMemcpyDeviceToHost, s));
  PHIPCHK(hipStreamSynchronize(s));
  // the table: twice the largest possible row, at most LDS_HASH_MAX slots
  const int64_t bound = std::min<int64_t>(static_cast<int64_t>(cand), b.ncols);
  // (at least 256 slots: a full wave can insert 63 keys past the half-full
  // mark, and a probe needs a free slot)
  const int hs_full = std::min(LDS_HASH_MAX, std::max(256, next_pow2(2 * std::max<int64_t>(bound, 1))));
  // a small table first (the candidate bound counts repeated columns many
  // times over); the full one only when some row overflows it
  int hs = std::min(hs_full, 2048);
  DBuf cnt;
  cnt.alloc(std::max<int64_t>(a.nrows, 1) * 8);
  for (;;) {
    const size_t sym_lds = static_cast<size_t>(hs) * 4 + static_cast<size_t>(hs / 2) * 4;
    if (a.nrows)
      hipLaunchKernelGGL(k_spgemm_sym, dim3(static_cast<unsigned>(a.nrows)), dim3(WAVE), sym_lds, s, a.nrows,
                         a.ptr.as<int64_t>(), a.idx.as<int32_t>(), b.ptr.as<int64_t>(), b.idx.as<int32_t>(), hs,
                         cnt.as<int64_t>(), static_cast<const int64_t*>(nullptr), static_cast<int32_t*>(nullptr),
                         ovf.as<int32_t>());
    PHIPCHK(hipGetLastError());
    int32_t over = 0;
    PHIPCHK(hipMemcpyAsync(&over, ovf.p, 4, hipMemcpyDeviceToHost, s));
    PHIPCHK(hipStreamSynchronize(s));
    if (!over) break;
    if (hs == hs_full) throw DeviceSetupUnsupported("AMG device setup: a product row too long for the LDS table");
    hs = hs_full;
    PHIPCHK(hipMemsetAsync(ovf.p, 0, 4, s));
  }
  const size_t sym_lds = static_cast<size_t>(hs) * 4 + static_cast<size_t>(hs / 2) * 4;
  c.nnz = counts_to_ptr(cnt, c.ptr, a.nrows, s);
  c.idx.alloc(std::max<int64_t>(c.nnz, 1) * 4);
  c.val.alloc(std::max<int64_t>(c.nnz, 1) * 8);
  if (a.nrows) {
    PHIPCHK(hipMemsetAsync(mx.p, 0, 8, s));
    hipLaunchKernelGGL(k_row_len_max, dim3(nblocks(c.nrows)), dim3(256), 0, s, c.nrows, c.ptr.as<int64_t>(),
                       mx.as<unsigned long long>());
    unsigned long long longest = 0;
    PHIPCHK(hipMemcpyAsync(&longest, mx.p, 8, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_spgemm_sym, dim3(static_cast<unsigned>(a.nrows)), dim3(WAVE), sym_lds, s, a.nrows,
                       a.ptr.as<int64_t>(), a.idx.as<int32_t>(), b.ptr.as<int64_t>(), b.idx.as<int32_t>(), hs,
                       static_cast<int64_t*>(nullptr), c.ptr.as<int64_t>(), c.idx.as<int32_t>(), ovf.as<int32_t>());
    PHIPCHK(hipGetLastError());
    PHIPCHK(hipStreamSynchronize(s));
    const int ccap = std::max<int>(1, static_cast<int>(longest));
    // the stage: room for the candidates of a row, at most 1,024 entries
    // (12 KB) so that many rows stay resident per compute unit, and within
    // 56 KB of LDS in all
    // (rows of B of 256 entries and more on average keep the lanes busy
    // without it: sb = 0 reads them straight from global memory)
    const bool direct = b.nrows > 0 && b.nnz >= 256 * b.nrows;
    const int sb = direct ? 0
                          : std::max(64, std::min({1024,
                                                   next_pow2(static_cast<int64_t>(std::max<unsigned long long>(cand, 1))),
                                                   (56 * 1024 - ccap * 12 - WAVE * 12 - 8) / 12}));
    const size_t lds = static_cast<size_t>(ccap) * 12 + static_cast<size_t>(sb) * 12 + WAVE * 8 + (WAVE + 1) * 4;
    hipLaunchKernelGGL(k_spgemm_num, dim3(static_cast<unsigned>(a.nrows)), dim3(WAVE), lds, s, a.nrows,
                       a.ptr.as<int64_t>(), a.idx.as<int32_t>(), a.val.as<double>(), b.ptr.as<int64_t>(),
                       b.idx.as<int32_t>(), b.val.as<double>(), c.ptr.as<int64_t>(), c.idx.as<int32_t>(),
                       c.val.as<double>(), ccap, sb);
    PHIPCHK(hipGetLastError());
  }
  PHIPCHK(hipStreamSynchronize(s));
  return c;
}

// P = P~ - (omega / d) A P~ merged row by row, zeros dropped (sub_scaled,
// amg_setup.cpp); pass 0 counts, pass 1 writes
__global__ void k_sub_scaled(int64_t n, const int64_t* __restrict__ pp, const int32_t* __restrict__ pi,
                             const double* __restrict__ pv, const int64_t* __restrict__ qp,
                             const int32_t* __restrict__ qi, const double* __restrict__ qv,
                             const double* __restrict__ diag, double omega, int64_t* cnt, const int64_t* mp,
                             int32_t* mi, double* mv) {
#pragma clang fp contract(off)
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = omega / diag[i];
  int64_t p = pp[i], pe = pp[i + 1], q = qp[i], qe = qp[i + 1];
  int64_t k = 0;
  while (p < pe || q < qe) {
    const int64_t cp = p < pe ? pi[p] : INT64_MAX;
    const int64_t cq = q < qe ? qi[q] : INT64_MAX;
    int64_t c;
    double v;
    if (cp == cq) {
      v = pv[p] - s * qv[q];
      ++p;
      ++q;
      c = cp;
    } else if (cp < cq) {
      v = pv[p];
      ++p;
      c = cp;
    } else {
      v = -s * qv[q];
      ++q;
      c = cq;
    }
    if (v != 0.0) {
      if (mi) {
        mi[mp[i] + k] = static_cast<int32_t>(c);
        mv[mp[i] + k] = v;
      }
      ++k;
    }
  }
  if (cnt) cnt[i] = k;
}

// transpose: counts per column, then a fill (sorted per row afterwards)
__global__ void k_tr_count(int64_t n, const int64_t* ptr, const int32_t* idx, unsigned long long* cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) atomicAdd(&cnt[idx[p]], 1ull);
}

__global__ void k_tr_fill(int64_t n, const int64_t* ptr, const int32_t* idx, const double* val, const int64_t* tptr,
                          unsigned long long* fill, int32_t* tidx, double* tval) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
    const int32_t c = idx[p];
    const int64_t a = tptr[c] + static_cast<int64_t>(atomicAdd(&fill[c], 1ull));
    tidx[a] = static_cast<int32_t>(i);
    tval[a] = val[p];
  }
}

DCsr dev_transpose(const DCsr& a, hipStream_t s) {
  DCsr t;
  t.nrows = a.ncols;
  t.ncols = a.nrows;
  DBuf cnt;
  cnt.alloc(std::max<int64_t>(t.nrows, 1) * 8);
  PHIPCHK(hipMemsetAsync(cnt.p, 0, std::max<int64_t>(t.nrows, 1) * 8, s));
  if (a.nrows)
    hipLaunchKernelGGL(k_tr_count, dim3(nblocks(a.nrows)), dim3(256), 0, s, a.nrows, a.ptr.as<int64_t>(),
                       a.idx.as<int32_t>(), cnt.as<unsigned long long>());
  t.nnz = counts_to_ptr(cnt, t.ptr, t.nrows, s);
  t.idx.alloc(std::max<int64_t>(t.nnz, 1) * 4);
  t.val.alloc(std::max<int64_t>(t.nnz, 1) * 8);
  PHIPCHK(hipMemsetAsync(cnt.p, 0, std::max<int64_t>(t.nrows, 1) * 8, s));
  if (a.nrows)
    hipLaunchKernelGGL(k_tr_fill, dim3(nblocks(a.nrows)), dim3(256), 0, s, a.nrows, a.ptr.as<int64_t>(),
                       a.idx.as<int32_t>(), a.val.as<double>(), t.ptr.as<int64_t>(), cnt.as<unsigned long long>(),
                       t.idx.as<int32_t>(), t.val.as<double>());
  PHIPCHK(hipGetLastError());
  sort_rows(t, s);
  PHIPCHK(hipStreamSynchronize(s));
  return t;
}

// ----------------------------------------- BSR operators of the V-cycle

// Block pattern of each block row: the sorted union of the column blocks of
// its bs rows (a merge of the rows' sorted columns); pass 0 counts, pass 1
// writes.
__global__ void k_bsr_pattern(int64_t nbr, int bs, int bsc, const int64_t* __restrict__ ptr,
                              const int32_t* __restrict__ idx, int64_t* cnt, const int64_t* bptr, int32_t* bidx) {
  const int64_t br = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (br >= nbr) return;
  int64_t h[6], e[6];
  for (int r = 0; r < bs; ++r) {
    h[r] = ptr[br * bs + r];
    e[r] = ptr[br * bs + r + 1];
  }
  int64_t k = 0;
  while (true) {
    int32_t mb = INT32_MAX;
    for (int r = 0; r < bs; ++r)
      if (h[r] < e[r]) mb = min(mb, idx[h[r]] / bsc);
    if (mb == INT32_MAX) break;
    if (bidx) bidx[bptr[br] + k] = mb;
    ++k;
    for (int r = 0; r < bs; ++r)
      while (h[r] < e[r] && idx[h[r]] / bsc == mb) ++h[r];
  }
  if (cnt) cnt[br] = k;
}

__device__ __forceinline__ uint16_t to_bf16_dev(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// values into the block slots (make_mat's placement and layouts)
__global__ void k_bsr_values(int64_t nrows, int bs, int bsc, const int64_t* __restrict__ ptr,
                             const int32_t* __restrict__ idx, const double* __restrict__ val,
                             const int64_t* __restrict__ bptr, const int32_t* __restrict__ bidx, int64_t nnzb,
                             int soa, double* v64, float* v32, uint16_t* v16) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int64_t br = i / bs;
  const int ri = static_cast<int>(i % bs);
  const int bb = bs * bsc;
  const int g8 = granule(bb, 8), g4 = granule(bb, 4), g2 = granule(bb, 2);
  for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
    const int32_t c = idx[p];
    const int32_t bc = c / bsc;
    int64_t lo = bptr[br], hi = bptr[br + 1];
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (bidx[mid] < bc)
        lo = mid + 1;
      else
        hi = mid;
    }
    const int w = ri * bsc + c % bsc;
    const double x = val[p];
    auto at = [&](int gr) {
      return soa ? (static_cast<int64_t>(w / gr) * nnzb + lo) * gr + w % gr : lo * bb + w;
    };
    if (v64) v64[at(g8)] = x;
    if (v32) v32[at(g4)] = static_cast<float>(x);
    if (v16) v16[at(g2)] = to_bf16_dev(static_cast<float>(x));
  }
}

// make_mat (scalar CSR -> bs x bsc blocks) from a device CSR
DevMat make_mat_dev(const DCsr& a, int bs, int bsc, bool keep64, bool keep32, bool soa, bool keep16, hipStream_t s) {
  if (a.nrows % bs || a.ncols % bsc || bs > 6) throw std::runtime_error("internal: block sizes do not divide the matrix");
  DevMat M;
  M.bs = bs;
  M.bsc = bsc;
  M.soa = soa && bs * bsc > 1;
  M.nbr = a.nrows / bs;
  M.nbc = a.ncols / bsc;
  DBuf cnt;
  cnt.alloc(std::max<int64_t>(M.nbr, 1) * 8);
  if (M.nbr)
    hipLaunchKernelGGL(k_bsr_pattern, dim3(nblocks(M.nbr)), dim3(256), 0, s, M.nbr, bs, bsc, a.ptr.as<int64_t>(),
                       a.idx.as<int32_t>(), cnt.as<int64_t>(), static_cast<const int64_t*>(nullptr),
                       static_cast<int32_t*>(nullptr));
  M.nnzb = counts_to_ptr(cnt, M.ptr, M.nbr, s);
  M.idx.alloc(std::max<int64_t>(M.nnzb, 1) * 4);
  if (M.nbr)
    hipLaunchKernelGGL(k_bsr_pattern, dim3(nblocks(M.nbr)), dim3(256), 0, s, M.nbr, bs, bsc, a.ptr.as<int64_t>(),
                       a.idx.as<int32_t>(), static_cast<int64_t*>(nullptr), M.ptr.as<int64_t>(), M.idx.as<int32_t>());
  M.group = pick_group(M.nbr ? static_cast<double>(M.nnzb) / M.nbr : 1.0);
  const size_t nv = static_cast<size_t>(std::max<int64_t>(M.nnzb, 1)) * bs * bsc;
  if (keep64) {
    M.v64.alloc(nv * 8);
    PHIPCHK(hipMemsetAsync(M.v64.p, 0, nv * 8, s));
  }
  if (keep32) {
    M.v32.alloc(nv * 4);
    PHIPCHK(hipMemsetAsync(M.v32.p, 0, nv * 4, s));
  }
  if (keep16) {
    M.v16.alloc(nv * 2);
    PHIPCHK(hipMemsetAsync(M.v16.p, 0, nv * 2, s));
  }
  if (a.nrows)
    hipLaunchKernelGGL(k_bsr_values, dim3(nblocks(a.nrows)), dim3(256), 0, s, a.nrows, bs, bsc, a.ptr.as<int64_t>(),
                       a.idx.as<int32_t>(), a.val.as<double>(), M.ptr.as<int64_t>(), M.idx.as<int32_t>(), M.nnzb,
                       M.soa ? 1 : 0, M.v64.as<double>(), M.v32.as<float>(), M.v16.as<uint16_t>());
  PHIPCHK(hipGetLastError());
  PHIPCHK(hipStreamSynchronize(s));
  return M;
}

DCsr copy_csr(const DCsr& a, hipStream_t s) {
  DCsr c;
  c.nrows = a.nrows;
  c.ncols = a.ncols;
  c.nnz = a.nnz;
  c.ptr.alloc((a.nrows + 1) * 8);
  c.idx.alloc(std::max<int64_t>(a.nnz, 1) * 4);
  c.val.alloc(std::max<int64_t>(a.nnz, 1) * 8);
  PHIPCHK(hipMemcpyAsync(c.ptr.p, a.ptr.p, (a.nrows + 1) * 8, hipMemcpyDeviceToDevice, s));
  if (a.nnz) {
    PHIPCHK(hipMemcpyAsync(c.idx.p, a.idx.p, a.nnz * 4, hipMemcpyDeviceToDevice, s));
    PHIPCHK(hipMemcpyAsync(c.val.p, a.val.p, a.nnz * 8, hipMemcpyDeviceToDevice, s));
  }
  return c;
}

// ----------------------------------------------------------- hierarchy

struct DevLevel {
  DCsr A, P, R;
  std::vector<double> diag;
  double rho = 0.0, omega = 0.0;
  std::vector<int64_t> bptr;
};

struct DevHierarchy {
  std::vector<DevLevel> levels;
  Csr coarse;
  std::vector<int64_t> coarse_bptr;
  std::vector<double> coarse_l, coarse_d;
  std::vector<int64_t> sizes;
  double complexity = 0.0;
  std::unique_ptr<AmgHost> kept;  // the host copy, when asked for
};

// The setup loop of amg_setup (amg_setup.cpp; src/solver/iterative/amg.rs:282-341)
// on device matrices.  radius(level matrix, diag) returns 1.05 x the power
// method's estimate (level 0 may use the solver's own BSR copy).
DevHierarchy amg_setup_device(DCsr a, std::vector<int64_t> bptr, const std::vector<double>& b_host, int nb,
                              const std::function<double(const DCsr&, int, const std::vector<double>&)>& radius,
                              bool keep, hipStream_t s) {
  DevHierarchy H;
  if (keep) H.kept = std::make_unique<AmgHost>();
  H.sizes.push_back(a.nrows);
  const double fine_nnz = static_cast<double>(a.nnz);
  double total = fine_nnz;
  double theta = 0.08;
  const char* tenv = std::getenv("RUSTYFEM_AMG_TIMING");
  const bool timing = tenv && *tenv && *tenv != '0';
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto secs = [](std::chrono::steady_clock::time_point a0, std::chrono::steady_clock::time_point a1) {
    return std::chrono::duration<double>(a1 - a0).count();
  };
  DBuf bdev = upload(b_host);
  std::vector<double> b_keep = keep ? b_host : std::vector<double>();
  DBuf flag;
  flag.alloc(4);
  int level = 0;
  while (a.nrows > 1500 && H.levels.size() < 12) {
    const auto t0 = now();
    const int64_t n = a.nrows;
    const int64_t nblk = static_cast<int64_t>(bptr.size()) - 1;
    // diagonal
    DBuf diag;
    diag.alloc(n * 8);
    PHIPCHK(hipMemsetAsync(flag.p, 0, 4, s));
    hipLaunchKernelGGL(k_diag_check, dim3(nblocks(n)), dim3(256), 0, s, n, a.ptr.as<int64_t>(), a.idx.as<int32_t>(),
                       a.val.as<double>(), diag.as<double>(), flag.as<int32_t>());
    int32_t bad = 0;
    PHIPCHK(hipMemcpyAsync(&bad, flag.p, 4, hipMemcpyDeviceToHost, s));
    PHIPCHK(hipStreamSynchronize(s));
    if (bad) throw AmgBreakdown("AMG setup: a diagonal entry is not positive");
    // strength and the strong lists
    DBuf dbptr = upload(bptr), blkof, lid, lacc, nlist, dn, scnt, sptr_d, sidx_d;
    blkof.alloc(n * 4);
    hipLaunchKernelGGL(k_blkof, dim3(nblocks(nblk)), dim3(256), 0, s, nblk, dbptr.as<int64_t>(), blkof.as<int32_t>());
    lid.alloc(std::max<int64_t>(a.nnz, 1) * 4);
    lacc.alloc(std::max<int64_t>(a.nnz, 1) * 8);
    nlist.alloc(nblk * 4);
    dn.alloc(nblk * 8);
    hipLaunchKernelGGL(k_strength_lists, dim3(nblocks(nblk)), dim3(256), 0, s, nblk, dbptr.as<int64_t>(),
                       a.ptr.as<int64_t>(), a.idx.as<int32_t>(), a.val.as<double>(), blkof.as<int32_t>(),
                       lid.as<int32_t>(), lacc.as<double>(), nlist.as<int32_t>(), dn.as<double>());
    scnt.alloc(nblk * 8);
    hipLaunchKernelGGL(k_strong, dim3(nblocks(nblk)), dim3(256), 0, s, nblk, dbptr.as<int64_t>(), a.ptr.as<int64_t>(),
                       lid.as<int32_t>(), lacc.as<double>(), nlist.as<int32_t>(), dn.as<double>(), theta,
                       scnt.as<int64_t>(), static_cast<const int64_t*>(nullptr), static_cast<int32_t*>(nullptr));
    const int64_t nstrong = counts_to_ptr(scnt, sptr_d, nblk, s);
    sidx_d.alloc(std::max<int64_t>(nstrong, 1) * 4);
    hipLaunchKernelGGL(k_strong, dim3(nblocks(nblk)), dim3(256), 0, s, nblk, dbptr.as<int64_t>(), a.ptr.as<int64_t>(),
                       lid.as<int32_t>(), lacc.as<double>(), nlist.as<int32_t>(), dn.as<double>(), theta,
                       static_cast<int64_t*>(nullptr), sptr_d.as<int64_t>(), sidx_d.as<int32_t>());
    PHIPCHK(hipGetLastError());
    const std::vector<int64_t> sptr = download_vec<int64_t>(sptr_d, nblk + 1, s);
    const std::vector<int32_t> sidx = download_vec<int32_t>(sidx_d, nstrong, s);
    lid.release();
    lacc.release();
    int64_t nagg = 0;
    const std::vector<int64_t> agg = aggregate_passes(nblk, sptr, sidx, nagg);
    const auto t1 = now();
    // tentative prolongator
    std::vector<int64_t> gptr(nagg + 1, 0), doff(nagg + 1, 0);
    std::vector<int32_t> gblk(nblk), agg32(nblk);
    for (int64_t k = 0; k < nblk; ++k) {
      ++gptr[agg[k] + 1];
      doff[agg[k] + 1] += bptr[k + 1] - bptr[k];
      agg32[k] = static_cast<int32_t>(agg[k]);
    }
    for (int64_t g = 0; g < nagg; ++g) {
      gptr[g + 1] += gptr[g];
      doff[g + 1] += doff[g];
    }
    {
      std::vector<int64_t> pos(gptr.begin(), gptr.end() - 1);
      for (int64_t k = 0; k < nblk; ++k) gblk[pos[agg[k]]++] = static_cast<int32_t>(k);
    }
    DBuf dgptr = upload(gptr), dgblk = upload(gblk), ddoff = upload(doff), dagg = upload(agg32), qs, ks, dnk, rmat, ptv;
    qs.alloc(std::max<int64_t>(n, 1) * nb * 8);
    ks.alloc(std::max<int64_t>(n, 1) * (nb + 1) * 8);
    dnk.alloc(std::max<int64_t>(nagg, 1) * 4);
    rmat.alloc(std::max<int64_t>(nagg, 1) * nb * nb * 8);
    ptv.alloc(std::max<int64_t>(n, 1) * nb * 8);
    hipLaunchKernelGGL(k_tentative, dim3(nblocks(nagg, 64)), dim3(64), 0, s, nagg, dgptr.as<int64_t>(),
                       dgblk.as<int32_t>(), dbptr.as<int64_t>(), ddoff.as<int64_t>(), bdev.as<double>(), nb,
                       qs.as<double>(), ks.as<double>(), dnk.as<int32_t>(), rmat.as<double>(), ptv.as<double>());
    PHIPCHK(hipGetLastError());
    const std::vector<int32_t> nk = download_vec<int32_t>(dnk, nagg, s);
    qs.release();
    ks.release();
    std::vector<int64_t> cptr(nagg + 1, 0);
    for (int64_t g = 0; g < nagg; ++g) cptr[g + 1] = cptr[g] + nk[g];
    const int64_t nc = cptr[nagg];
    DBuf dcptr = upload(cptr), pcnt;
    DCsr pt;
    pt.nrows = n;
    pt.ncols = nc;
    pcnt.alloc(n * 8);
    hipLaunchKernelGGL(k_ptent, dim3(nblocks(n)), dim3(256), 0, s, n, blkof.as<int32_t>(), dagg.as<int32_t>(),
                       dnk.as<int32_t>(), dcptr.as<int64_t>(), ptv.as<double>(), nb, pcnt.as<int64_t>(),
                       static_cast<const int64_t*>(nullptr), static_cast<int32_t*>(nullptr),
                       static_cast<double*>(nullptr));
    pt.nnz = counts_to_ptr(pcnt, pt.ptr, n, s);
    pt.idx.alloc(std::max<int64_t>(pt.nnz, 1) * 4);
    pt.val.alloc(std::max<int64_t>(pt.nnz, 1) * 8);
    hipLaunchKernelGGL(k_ptent, dim3(nblocks(n)), dim3(256), 0, s, n, blkof.as<int32_t>(), dagg.as<int32_t>(),
                       dnk.as<int32_t>(), dcptr.as<int64_t>(), ptv.as<double>(), nb, static_cast<int64_t*>(nullptr),
                       pt.ptr.as<int64_t>(), pt.idx.as<int32_t>(), pt.val.as<double>());
    DBuf bc;
    bc.alloc(std::max<int64_t>(nc, 1) * nb * 8);
    hipLaunchKernelGGL(k_coarse_b, dim3(nblocks(nagg)), dim3(256), 0, s, nagg, dnk.as<int32_t>(), dcptr.as<int64_t>(),
                       rmat.as<double>(), nb, bc.as<double>());
    PHIPCHK(hipGetLastError());
    PHIPCHK(hipStreamSynchronize(s));
    const auto t2 = now();
    if (static_cast<double>(nc) > 0.85 * static_cast<double>(n)) break;
    // damping from the spectral radius of D^-1 A
    std::vector<double> diag_h = download_vec<double>(diag, n, s);
    const double rho = radius(a, level, diag_h);
    const double omega = 4.0 / 3.0 / rho;
    const auto t3 = now();
    DCsr apt = dev_spgemm(a, pt, s);
    DCsr p;
    p.nrows = n;
    p.ncols = nc;
    {
      DBuf mc;
      mc.alloc(n * 8);
      hipLaunchKernelGGL(k_sub_scaled, dim3(nblocks(n)), dim3(256), 0, s, n, pt.ptr.as<int64_t>(), pt.idx.as<int32_t>(),
                         pt.val.as<double>(), apt.ptr.as<int64_t>(), apt.idx.as<int32_t>(), apt.val.as<double>(),
                         diag.as<double>(), omega, mc.as<int64_t>(), static_cast<const int64_t*>(nullptr),
                         static_cast<int32_t*>(nullptr), static_cast<double*>(nullptr));
      p.nnz = counts_to_ptr(mc, p.ptr, n, s);
      p.idx.alloc(std::max<int64_t>(p.nnz, 1) * 4);
      p.val.alloc(std::max<int64_t>(p.nnz, 1) * 8);
      hipLaunchKernelGGL(k_sub_scaled, dim3(nblocks(n)), dim3(256), 0, s, n, pt.ptr.as<int64_t>(), pt.idx.as<int32_t>(),
                         pt.val.as<double>(), apt.ptr.as<int64_t>(), apt.idx.as<int32_t>(), apt.val.as<double>(),
                         diag.as<double>(), omega, static_cast<int64_t*>(nullptr), p.ptr.as<int64_t>(),
                         p.idx.as<int32_t>(), p.val.as<double>());
      PHIPCHK(hipGetLastError());
      PHIPCHK(hipStreamSynchronize(s));
    }
    apt = DCsr();
    const auto t4 = now();
    DCsr r = dev_transpose(p, s);
    const auto t5 = now();
    DCsr apm = dev_spgemm(a, p, s);
    const auto t6 = now();
    DCsr ac = dev_spgemm(r, apm, s);
    apm = DCsr();
    const auto t7 = now();
    if (timing)
      std::fprintf(stderr,
                   "[amg-device] level %d: n %lld nnz %lld | aggregate %.3f s, tentative %.3f, spectral radius %.3f, "
                   "A P~ and smoothing %.3f, transpose %.3f, A P %.3f, R (A P) %.3f\n",
                   level, static_cast<long long>(n), static_cast<long long>(a.nnz), secs(t0, t1), secs(t1, t2),
                   secs(t2, t3), secs(t3, t4), secs(t4, t5), secs(t5, t6), secs(t6, t7));
    if (keep) {
      AmgHostLevel L;
      L.A = download(a, s);
      L.P = download(p, s);
      L.R = download(r, s);
      L.Pt = download(pt, s);
      L.diag = diag_h;
      L.rho = rho;
      L.omega = omega;
      L.nblk = nblk;
      L.bptr = bptr;
      L.agg = agg;
      L.B = b_keep;
      L.nbm = nb;
      H.kept->levels.push_back(std::move(L));
      b_keep = download_vec<double>(bc, nc * nb, s);
    }
    total += static_cast<double>(ac.nnz);
    H.sizes.push_back(ac.nrows);
    DevLevel L;
    L.A = std::move(a);
    L.P = std::move(p);
    L.R = std::move(r);
    L.diag = std::move(diag_h);
    L.rho = rho;
    L.omega = omega;
    L.bptr = bptr;
    H.levels.push_back(std::move(L));
    a = std::move(ac);
    bptr = std::move(cptr);
    bdev = std::move(bc);
    theta *= 0.5;
    ++level;
  }
  H.coarse = download(a, s);
  H.coarse_bptr = bptr;
  coarse_ldlt(H.coarse, H.coarse_l, H.coarse_d);
  H.complexity = total / fine_nnz;
  if (keep) {
    H.kept->coarse = H.coarse;
    H.kept->coarse_bptr = bptr;
    H.kept->coarse_l = H.coarse_l;
    H.kept->coarse_d = H.coarse_d;
    H.kept->sizes = H.sizes;
    H.kept->complexity = H.complexity;
    H.kept->coarse_B = b_keep;
  }
  return H;
}
