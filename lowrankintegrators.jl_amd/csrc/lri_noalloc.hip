// lri_product_noalloc: the product pass of lri_debug_product (dA * B or
// dA' * B) on caller-owned device arrays, for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops).  Every
// intermediate -- the packed B operand of the pass and its split partials --
// lives in one file-scope device array, `scratch`, of kNoAllocScratchBytes per
// GPU (held from the library's first use on that GPU until the process ends).
//
// Host code never passes the array's name to a kernel: on the first call on a
// device (with that device current) hipGetSymbolAddress yields that device's
// address of its copy, which is kept per device and used for every launch.
// All calls on one GPU share its array; calls that overlap on one GPU (two
// streams, or two host threads) race on it.
#include <mutex>
#include <string>

#include "../../include/lri.h"
#include "lri_common.h"
#include "lri_kernels.h"

namespace lri {
namespace {

constexpr size_t kNoAllocScratchBytes = LRI_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};
thread_local std::string g_noalloc_error;

// zero n16 16-byte words (the packed operand's padding rows must read as zero:
// the passes multiply them with masked, zero A values, and a NaN left there by
// an earlier call would survive 0 * NaN)
__global__ __launch_bounds__(256) void k_zero16(u32x4_t* __restrict__ p, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256)
    p[i] = u32x4_t{0u, 0u, 0u, 0u};
}

int fail(int status, const std::string& msg) {
  g_noalloc_error = msg;
  return status;
}

// this device's address of `scratch` (looked up once per device, the device current)
hipError_t scratch_for(int dev, void** out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int rb_for(int mode, int64_t cols) {
  if (mode == PASS_F64) return cols <= 16 ? 16 : (cols <= 32 ? 32 : (cols <= 48 ? 48 : 64));
  return cols <= 32 ? 32 : 64;
}

}  // namespace

// Scratch bytes a call of this shape needs (the packed operand, then the partials).
static size_t noalloc_need(int mode, bool transposed, int64_t n, int64_t m, int rb, size_t* pk_bytes, int* splits,
                           int64_t* chunk) {
  const int64_t krows = transposed ? n : m;  // rows of B
  const int64_t kpad = (krows + 63) / 64 * 64;
  size_t pk = 0;
  if (mode == PASS_F64) pk = sizeof(double) * (size_t)kpad * rb;
  else if (mode == PASS_F32) pk = sizeof(float) * (size_t)kpad * rb;
  else pk = 2 * 2 * (size_t)kpad * rb;  // hi and lo panels
  const int64_t parts = plan_pass(mode, transposed, transposed ? m : n, transposed ? n : m, rb, splits, chunk);
  *pk_bytes = align256(pk);
  return *pk_bytes + align256((size_t)parts * (mode == PASS_F64 ? 8 : 4));
}

}  // namespace lri

using namespace lri;

extern "C" int lri_product_noalloc(int which, int dtype, int precision, int64_t n, int64_t m, const void* a_prev,
                                   const void* a_next, int64_t lda, const double* B, int64_t ldb, int64_t r,
                                   double* out, int64_t ldo, void* stream) {
  if (which != 0 && which != 1) return fail(LRI_EINVAL, "which must be 0 (dA B) or 1 (dA' B)");
  if (dtype != LRI_F64 && dtype != LRI_F32 && dtype != LRI_BF16) return fail(LRI_EINVAL, "unknown dtype");
  if (precision != LRI_PREC_F64 && precision != LRI_PREC_MIXED) return fail(LRI_EINVAL, "unknown precision");
  if (n < 1 || m < 1 || r < 1 || r > LRI_MAX_RANK) return fail(LRI_EINVAL, "need n, m >= 1 and 1 <= r <= LRI_MAX_RANK");
  if (!a_prev || !a_next || !B || !out) return fail(LRI_EINVAL, "a NULL array");
  const bool tr = which == 1;
  const int64_t brows = tr ? n : m, orows = tr ? m : n;
  if (lda < n || ldb < brows || ldo < orows) return fail(LRI_EINVAL, "leading dimensions too small");
  const int mode = (precision == LRI_PREC_F64 || dtype == LRI_F64) ? PASS_F64
                                                                     : (dtype == LRI_BF16 ? PASS_BF16X3 : PASS_F32);
  const int rb = rb_for(mode, r);
  size_t pk_bytes = 0;
  int splits = 0;
  int64_t chunk = 0;
  const size_t need = noalloc_need(mode, tr, n, m, rb, &pk_bytes, &splits, &chunk);
  if (need > kNoAllocScratchBytes)
    return fail(LRI_EINVAL, "n = " + std::to_string(n) + ", m = " + std::to_string(m) + ", r = " + std::to_string(r) +
                                " needs " + std::to_string(need) + " bytes of scratch, above LRI_NOALLOC_SCRATCH_BYTES = " +
                                std::to_string(kNoAllocScratchBytes));
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(LRI_EHIP, "no current HIP device");
  void* base = nullptr;
  hipError_t e = scratch_for(dev, &base);
  if (e != hipSuccess) return fail(LRI_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  hipStream_t st = (hipStream_t)stream;
  unsigned char* pk = static_cast<unsigned char*>(base);
  void* part = pk + pk_bytes;

  // 1. clear the packed operand, then pack B into it (the row kernel zeroes columns r..rb)
  const int64_t n16 = (int64_t)(pk_bytes / 16);
  hipLaunchKernelGGL(k_zero16, dim3((unsigned)std::min<int64_t>((n16 + 255) / 256, 1024)), dim3(256), 0, st,
                     reinterpret_cast<u32x4_t*>(pk), n16);
  if ((e = hipGetLastError()) != hipSuccess) return fail(LRI_EHIP, std::string("zero fill: ") + hipGetErrorString(e));
  const int64_t kpad = (brows + 63) / 64 * 64;
  RowArgs<double> P;
  P.rows = brows;
  P.r_out = (int)r;
  P.base = B;
  P.ld_base = ldb;
  P.pk_cols = rb;
  PassArgs a;
  if (mode == PASS_F64) {
    P.pk_f64 = reinterpret_cast<double*>(pk);
    a.B = pk;
  } else if (mode == PASS_F32) {
    P.pk_f32 = reinterpret_cast<float*>(pk);
    P.pk_layout = tr ? 1 : 0;
    a.B = pk;
  } else {
    P.pk_hi = reinterpret_cast<uint16_t*>(pk);
    P.pk_lo = reinterpret_cast<uint16_t*>(pk) + kpad * rb;
    P.pk_layout = tr ? 1 : 0;
    a.B = P.pk_hi;
    a.B2 = P.pk_lo;
    a.ldb = kpad;
  }
  if ((e = launch_rows<double>(P, st)) != hipSuccess) return fail(LRI_EHIP, std::string("pack: ") + hipGetErrorString(e));

  // 2. the streaming pass, split partials into the scratch
  a.anext = a_next;
  a.aprev = a_prev;
  a.dtype = dtype;
  a.lda = lda;
  a.n = n;
  a.m = m;
  a.rb = rb;
  a.part = part;
  a.splits = splits;
  a.chunk = chunk;
  e = tr ? launch_passT(mode, a, st) : launch_passN(mode, a, st);
  if (e != hipSuccess) return fail(LRI_EHIP, std::string("pass: ") + hipGetErrorString(e));

  // 3. the split partials summed in order into out
  if (mode == PASS_F64) {
    RowArgs<double> S;
    S.rows = orows;
    S.r_out = (int)r;
    S.parts = static_cast<const double*>(part);
    S.nsplit = splits;
    S.rb = rb;
    S.out = out;
    S.ldo = ldo;
    e = launch_rows<double>(S, st);
  } else {
    RowArgs<float> S;
    S.rows = orows;
    S.r_out = (int)r;
    S.parts = static_cast<const float*>(part);
    S.nsplit = splits;
    S.rb = rb;
    S.out = out;
    S.ldo = ldo;
    e = launch_rows<float>(S, st);
  }
  if (e != hipSuccess) return fail(LRI_EHIP, std::string("sum: ") + hipGetErrorString(e));
  return LRI_OK;
}

extern "C" const char* lri_noalloc_last_error(void) { return g_noalloc_error.c_str(); }
