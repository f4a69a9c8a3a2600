// madam_noalloc.hip -- triangle separation that allocates nothing, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// It runs the kernels of madam_tri_separate (madam_sep.hip) on the
// caller's device buffers, with their working memory -- the separation
// state, counters, histogram and key/value staging -- in one file-scope
// device array of fixed size instead of library allocations.  Every call on
// a GPU shares that GPU's array: calls that overlap on one GPU (two
// streams, two host threads) race and must be serialised by the caller.
// The array keeps the last call's contents; each call clears the parts that
// must start at zero with a fill kernel (no hipMemsetAsync, which graph
// capture records badly).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "madam_engine.h"
#include "madam_sep.h"

namespace madam {

// Largest K (the reference's cap on added triangles, MaxTriIneqAdded,
// admm.h:34) and the layout of the array, in bytes from its start.
constexpr int NA_KMAX = 50000;
constexpr size_t NA_VIEW = 0;
constexpr size_t NA_CTL = 256;
constexpr size_t NA_MAXB = NA_CTL + 256;
constexpr size_t NA_HIST = NA_MAXB + 256;
constexpr size_t NA_OKEY = NA_HIST + SEP_NBIN * 4;
constexpr size_t NA_BKEY = NA_OKEY + (size_t)NA_KMAX * 4;
constexpr size_t NA_BVAL = NA_BKEY + (size_t)SEP_BCAP * 4;
constexpr size_t NA_FKEY = NA_BVAL + (size_t)SEP_BCAP * 8;
constexpr size_t NA_FVAL = NA_FKEY + (size_t)NA_KMAX * 4;
constexpr size_t NA_BYTES = (NA_FVAL + (size_t)NA_KMAX * 8 + 255) & ~(size_t)255;
static_assert(sizeof(SepView) <= NA_CTL, "SepView fits its slot");
static_assert(CTL_N * 4 <= NA_MAXB - NA_CTL, "control words fit their slot");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Per-device address of `scratch`, resolved once per device at set-up.  In
// host code the name `scratch` is a host-side stand-in, never a device
// address: only these pointers are passed to kernels.
constexpr int NA_MAXDEV = 64;
static unsigned char* g_scratch_dev[NA_MAXDEV] = {nullptr};

// The separation view of this call into the array, its counters,
// histogram and max cleared.
__global__ __launch_bounds__(256) void k_na_init(SepView v, unsigned char* base) {
  const int tid = threadIdx.x;
  if (tid == 0) *(SepView*)(base + NA_VIEW) = v;
  uint32_t* ctl = (uint32_t*)(base + NA_CTL);
  uint32_t* hist = (uint32_t*)(base + NA_HIST);
  for (int i = tid; i < CTL_N; i += 256) ctl[i] = 0u;
  for (int i = tid; i < SEP_NBIN; i += 256) hist[i] = 0u;
  if (tid == 0) *(unsigned long long*)(base + NA_MAXB) = 0ull;
}

// Final keys -> the caller's madam_tri list (the decoding madam_tri_separate
// does on the host), the list length and the largest row value.
__global__ __launch_bounds__(256) void k_na_emit(const unsigned char* base, madam_tri* list,
                                                 int* size, double* maxval) {
  const SepView& s = *(const SepView*)(base + NA_VIEW);
  const uint32_t nf = s.ctl[CTL_NFINAL];
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q < nf) {
    const uint32_t key = s.fkey[q];
    madam_tri t;
    t.type = (int)(key >> 30) + 1;
    t.i = (key >> 20) & 1023;
    t.j = (key >> 10) & 1023;
    t.k = key & 1023;
    t.value = s.fval[q];
    t.y = 0.0;
    list[q] = t;
  }
  if (q == 0) {
    *size = (int)nf;
    const unsigned long long mb = s.maxbits[0];
    const unsigned long long bits =
        (mb & 0x8000000000000000ull) ? (mb & 0x7fffffffffffffffull) : ~mb;
    *maxval = __longlong_as_double((long long)bits);
  }
}

}  // namespace madam

namespace mh {

// Called by madam_open with `device` current.
int noalloc_setup(int device) {
  if (device < 0 || device >= madam::NA_MAXDEV) {
    set_err("device %d beyond the %d the no-allocation path supports", device, madam::NA_MAXDEV);
    return MADAM_E_ARG;
  }
  if (madam::g_scratch_dev[device]) return MADAM_OK;
  void* p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(madam::scratch)));
  madam::g_scratch_dev[device] = (unsigned char*)p;
  return MADAM_OK;
}

}  // namespace mh

using namespace mh;

extern "C" {

int madam_tri_separate_noalloc(int n, const double* X, int K, double thresh, madam_tri* list,
                               int* size, double* maxval, void* stream) {
  if (n < 3 || n > MADAM_NMAX || K < 0 || K > madam::NA_KMAX || !X || !size || !maxval ||
      (K > 0 && !list)) {
    set_err("madam_tri_separate_noalloc: n=%d (3..%d), K=%d (0..%d) or a null pointer", n,
            MADAM_NMAX, K, madam::NA_KMAX);
    return MADAM_E_ARG;
  }
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  unsigned char* base = (dev >= 0 && dev < madam::NA_MAXDEV) ? madam::g_scratch_dev[dev] : nullptr;
  if (!base) {
    set_err("madam_tri_separate_noalloc: call madam_open(%d) first (it resolves the scratch "
            "array of the device)", dev);
    return MADAM_E_ARG;
  }
  const hipStream_t st = (hipStream_t)stream;
  madam::SepView v;
  memset(&v, 0, sizeof v);
  v.n = n;
  v.K = K;
  v.thresh = thresh;
  v.X = X;
  v.ctl = (uint32_t*)(base + madam::NA_CTL);
  v.maxbits = (unsigned long long*)(base + madam::NA_MAXB);
  v.hist = (uint32_t*)(base + madam::NA_HIST);
  v.okey = (uint32_t*)(base + madam::NA_OKEY);
  v.bkey = (uint32_t*)(base + madam::NA_BKEY);
  v.bval = (double*)(base + madam::NA_BVAL);
  v.fkey = (uint32_t*)(base + madam::NA_FKEY);
  v.fval = (double*)(base + madam::NA_FVAL);
  madam::SepView* sv = (madam::SepView*)(base + madam::NA_VIEW);
  madam::k_na_init<<<1, 256, 0, st>>>(v, base);
  madam::launch_tri_separate(sv, 1, n, st);
  // the boundary refinement is launched every time (it returns at once
  // unless the boundary bin overflowed): no host round trip to decide
  madam::launch_tri_refine(sv, 1, n, st);
  madam::k_na_emit<<<(K + 255) / 256 > 0 ? (K + 255) / 256 : 1, 256, 0, st>>>(base, list, size,
                                                                             maxval);
  HIP_TRY(hipGetLastError());
  return MADAM_OK;
}

}  // extern "C"
