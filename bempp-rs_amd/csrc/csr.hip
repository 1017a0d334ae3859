// csr.hip -- the sparse singular part in CSR form, built on the device from the adjacent-pair
// kernel's COO values.
//
// Follows assemble_singular (src/boundary_assemblers.rs:122-155): the COO triplets of
// assemble_singular_part become CsrMatrix::from_aij(shape, rows, cols, data), i.e. rows in
// order, columns ascending within a row and duplicate (row, column) entries summed; when no
// |entry| exceeds 1e-10 the result is the empty matrix (all row pointers 0).
//
// Steps: (row * ncols + col) keys from the pair list and the cells' DOF maps; a stable radix
// sort of (key, COO position); each run of equal keys summed sequentially in COO order, in
// the output value type (so the sums equal a stable sort on the host followed by in-order
// accumulation, bit for bit); row pointers from the rows of the unique entries.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>
#include <string>

#include "core.hpp"
#include "kernels.hpp"

namespace bempp_amd {

#define HIP_CHECK_C(expr)                                                                         \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) throw Error(std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

namespace {

using u64 = unsigned long long;

// COO position e = pair * nbt * nbs + t + nbt * b (t: test basis function, b: trial basis
// function), the layout of the adjacent-pair kernels' COO mode. DOF tables: `tstride` /
// `sstride` entries per cell, -1 for a basis function the cell does not have (a triangle's
// fourth on a grid with quadrilaterals); such entries take the key nrows * ncols, which
// sorts after every real one.
__global__ void __launch_bounds__(256) csr_keys_kernel(const int4* __restrict__ pairs, long long n, int nbt, int nbs,
                                                       const int* __restrict__ test_dofs, int tstride,
                                                       const int* __restrict__ trial_dofs, int sstride, u64 nrows,
                                                       u64 ncols, const double* __restrict__ vals, int dtype,
                                                       u64* __restrict__ keys, unsigned* __restrict__ pos,
                                                       int* __restrict__ any_nonzero) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  bool nz = false;
  if (e < n) {
    const int nbl = nbt * nbs;
    const long long p = e / nbl;
    const int r = (int)(e - p * nbl);
    const int t = r % nbt, b = r / nbt;
    const int4 pr = pairs[p];
    const int row = test_dofs[(long long)pr.x * tstride + t];
    const int col = trial_dofs[(long long)pr.y * sstride + b];
    const bool valid = row >= 0 && col >= 0;
    keys[e] = valid ? (u64)row * ncols + (u64)col : nrows * ncols;
    pos[e] = (unsigned)e;
    const double re = valid ? vals[2 * e] : 0.0, im = valid ? vals[2 * e + 1] : 0.0;
    double mag;
    switch (dtype) {
      case kF64: mag = fabs(re); break;
      case kF32: mag = (double)fabsf((float)re); break;
      case kC64: mag = hypot(re, im); break;
      default: mag = (double)hypotf((float)re, (float)im); break;
    }
    nz = mag > 1e-10;
  }
  if (__any(nz) && (threadIdx.x & 63) == 0) atomicOr(any_nonzero, 1);
}

__global__ void __launch_bounds__(256) csr_heads_kernel(const u64* __restrict__ keys, long long n,
                                                        unsigned* __restrict__ head) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e < n) head[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
}

template <typename T>
struct Value;
template <>
struct Value<double> {
  static __device__ double load(const double* v, unsigned e) { return v[2 * (size_t)e]; }
  static __device__ double add(double a, double b) { return a + b; }
};
template <>
struct Value<float> {
  static __device__ float load(const double* v, unsigned e) { return (float)v[2 * (size_t)e]; }
  static __device__ float add(float a, float b) { return a + b; }
};
template <>
struct Value<double2> {
  static __device__ double2 load(const double* v, unsigned e) {
    return make_double2(v[2 * (size_t)e], v[2 * (size_t)e + 1]);
  }
  static __device__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
};
template <>
struct Value<float2> {
  static __device__ float2 load(const double* v, unsigned e) {
    return make_float2((float)v[2 * (size_t)e], (float)v[2 * (size_t)e + 1]);
  }
  static __device__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
};

// One thread per run of equal keys (started by its head): the run's values summed in COO
// order into unique entry scan[e] - 1.
template <typename T>
__global__ void __launch_bounds__(256) csr_segments_kernel(const u64* __restrict__ keys,
                                                           const unsigned* __restrict__ order,
                                                           const unsigned* __restrict__ scan, long long n,
                                                           const double* __restrict__ vals, u64 ncols,
                                                           T* __restrict__ data, long long* __restrict__ indices,
                                                           long long* __restrict__ urow) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const u64 key = keys[e];
  if (e > 0 && keys[e - 1] == key) return;
  T acc = Value<T>::load(vals, order[e]);
  for (long long k = e + 1; k < n && keys[k] == key; ++k) acc = Value<T>::add(acc, Value<T>::load(vals, order[k]));
  const long long j = (long long)scan[e] - 1;
  data[j] = acc;
  indices[j] = (long long)(key % ncols);
  urow[j] = (long long)(key / ncols);
}

// indptr[r] = first unique entry whose row is >= r, for r = 0..nrows.
__global__ void __launch_bounds__(256) csr_indptr_kernel(const long long* __restrict__ urow, long long nnz,
                                                         long long nrows, long long* __restrict__ indptr) {
  const long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (j > nnz) return;
  const long long lo = j == 0 ? -1 : urow[j - 1];
  const long long hi = j < nnz ? urow[j] : nrows;
  for (long long r = lo + 1; r <= hi; ++r) indptr[r] = j;
}

struct DeviceBuffers {
  void* p = nullptr;
  DeviceBuffers() = default;
  DeviceBuffers(const DeviceBuffers&) = delete;
  DeviceBuffers& operator=(const DeviceBuffers&) = delete;
  ~DeviceBuffers() {
    if (p) (void)hipFree(p);
  }
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <typename T>
void launch_segments(dim3 g, hipStream_t s, const u64* keys, const unsigned* order, const unsigned* scan, long long n,
                     const double* vals, u64 ncols, void* data, long long* indices, long long* urow) {
  csr_segments_kernel<T><<<g, 256, 0, s>>>(keys, order, scan, n, vals, ncols, (T*)data, indices, urow);
}

}  // namespace

void coo_to_csr_device(const double* vals, const int4* pairs, int npairs, const int* test_dofs, int nbt, int tstride,
                       const int* trial_dofs, int nbs, int sstride, int64_t nrows, int64_t ncols, DType dtype,
                       hipStream_t s, CsrHost& out) {
  out.indptr.assign((size_t)nrows + 1, 0);
  out.indices.clear();
  out.data.clear();
  const long long n = (long long)npairs * nbt * nbs;
  if (n == 0 || nrows == 0 || ncols == 0) return;
  if (n >= (1LL << 32)) throw Error("too many singular COO entries for the device CSR build (2^32)");
  const u64 maxkey = (u64)nrows * (u64)ncols;  // the key of absent entries
  unsigned bits = 1;
  while (bits < 64 && (maxkey >> bits) != 0) ++bits;
  const size_t vb = dtype_size(dtype);
  // device layout: keys[2][n] u64, order[2][n] u32, head/scan[2][n] u32, data[n], indices[n],
  // urow[n], indptr[nrows + 1], flags, then the rocPRIM temporary storage
  size_t sort_tmp = 0, scan_tmp = 0;
  HIP_CHECK_C(rocprim::radix_sort_pairs(nullptr, sort_tmp, (const u64*)nullptr, (u64*)nullptr,
                                        (const unsigned*)nullptr, (unsigned*)nullptr, (unsigned)n, 0u, bits, s));
  HIP_CHECK_C(rocprim::inclusive_scan(nullptr, scan_tmp, (const unsigned*)nullptr, (unsigned*)nullptr, (unsigned)n,
                                      rocprim::plus<unsigned>(), s));
  const size_t o_keys = 0, o_keys2 = o_keys + align256(8 * n), o_ord = o_keys2 + align256(8 * n),
               o_ord2 = o_ord + align256(4 * n), o_head = o_ord2 + align256(4 * n), o_scan = o_head + align256(4 * n),
               o_data = o_scan + align256(4 * n), o_ind = o_data + align256(vb * n), o_urow = o_ind + align256(8 * n),
               o_ptr = o_urow + align256(8 * n), o_flag = o_ptr + align256(8 * ((size_t)nrows + 1)),
               o_tmp = o_flag + 256, total = o_tmp + align256(std::max(sort_tmp, scan_tmp));
  DeviceBuffers buf;
  HIP_CHECK_C(hipMalloc(&buf.p, total));
  char* base = (char*)buf.p;
  u64* keys = (u64*)(base + o_keys);
  u64* keys_sorted = (u64*)(base + o_keys2);
  unsigned* ord = (unsigned*)(base + o_ord);
  unsigned* ord_sorted = (unsigned*)(base + o_ord2);
  unsigned* head = (unsigned*)(base + o_head);
  unsigned* scan = (unsigned*)(base + o_scan);
  long long* indices = (long long*)(base + o_ind);
  long long* urow = (long long*)(base + o_urow);
  long long* indptr = (long long*)(base + o_ptr);
  int* flag = (int*)(base + o_flag);
  void* tmp = base + o_tmp;

  const dim3 g((unsigned)((n + 255) / 256));
  dev::fill_zero_kernel<int><<<1, 64, 0, s>>>(flag, 1);
  csr_keys_kernel<<<g, 256, 0, s>>>(pairs, n, nbt, nbs, test_dofs, tstride, trial_dofs, sstride, (u64)nrows,
                                    (u64)ncols, vals, (int)dtype, keys, ord, flag);
  HIP_CHECK_C(hipGetLastError());
  size_t b = sort_tmp;
  HIP_CHECK_C(rocprim::radix_sort_pairs(tmp, b, (const u64*)keys, keys_sorted, (const unsigned*)ord, ord_sorted,
                                        (unsigned)n, 0u, bits, s));
  csr_heads_kernel<<<g, 256, 0, s>>>(keys_sorted, n, head);
  HIP_CHECK_C(hipGetLastError());
  b = scan_tmp;
  HIP_CHECK_C(rocprim::inclusive_scan(tmp, b, (const unsigned*)head, scan, (unsigned)n, rocprim::plus<unsigned>(), s));
  switch (dtype) {
    case kF64: launch_segments<double>(g, s, keys_sorted, ord_sorted, scan, n, vals, (u64)ncols, base + o_data, indices, urow); break;
    case kF32: launch_segments<float>(g, s, keys_sorted, ord_sorted, scan, n, vals, (u64)ncols, base + o_data, indices, urow); break;
    case kC64: launch_segments<double2>(g, s, keys_sorted, ord_sorted, scan, n, vals, (u64)ncols, base + o_data, indices, urow); break;
    default: launch_segments<float2>(g, s, keys_sorted, ord_sorted, scan, n, vals, (u64)ncols, base + o_data, indices, urow); break;
  }
  HIP_CHECK_C(hipGetLastError());
  int any = 0;
  unsigned nnz = 0;
  u64 last = 0;
  HIP_CHECK_C(hipMemcpyAsync(&any, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK_C(hipMemcpyAsync(&nnz, scan + (n - 1), sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIP_CHECK_C(hipMemcpyAsync(&last, keys_sorted + (n - 1), sizeof(u64), hipMemcpyDeviceToHost, s));
  HIP_CHECK_C(hipStreamSynchronize(s));
  if (!any) return;
  if (last == maxkey) --nnz;  // the run of absent entries, last in key order
  csr_indptr_kernel<<<dim3((unsigned)((nnz + 1 + 255) / 256)), 256, 0, s>>>(urow, (long long)nnz, (long long)nrows,
                                                                            indptr);
  HIP_CHECK_C(hipGetLastError());
  out.indices.resize(nnz);
  out.data.resize((size_t)nnz * vb);
  HIP_CHECK_C(hipMemcpyAsync(out.indices.data(), indices, (size_t)nnz * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK_C(hipMemcpyAsync(out.data.data(), base + o_data, (size_t)nnz * vb, hipMemcpyDeviceToHost, s));
  HIP_CHECK_C(hipMemcpyAsync(out.indptr.data(), indptr, ((size_t)nrows + 1) * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK_C(hipStreamSynchronize(s));
}

}  // namespace bempp_amd
