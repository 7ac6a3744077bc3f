// Runtime entry points: errors, device selection, memory, and the host-only
// numerics of the reference (aliased index ranges, ETDRK4 contour coefficients).
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "prof.h"
#include "reduce.h"

namespace ffhip {

// ----------------------------------------------------------------- profiler
bool g_prof_on = false;
struct ProfRec {
  const char* name;
  double bytes;
  hipEvent_t a, b;
};
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_evpool;
static size_t g_evnext = 0;

static hipEvent_t prof_event() {
  if (g_evnext == g_evpool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_evpool.push_back(e);
  }
  return g_evpool[g_evnext++];
}

void prof_begin_impl(const char* name, double bytes, hipStream_t st) {
  ProfRec r{name, bytes, prof_event(), prof_event()};
  if (r.a && r.b) {
    (void)hipEventRecord(r.a, st);
    g_prof.push_back(r);
  }
}

void prof_end_impl(hipStream_t st) {
  if (!g_prof.empty()) (void)hipEventRecord(g_prof.back().b, st);
}

bool prof_pair(const char* name, double bytes, hipEvent_t* a, hipEvent_t* b) {
  ProfRec r{name, bytes, prof_event(), prof_event()};
  if (!r.a || !r.b) return false;
  g_prof.push_back(r);
  *a = r.a;
  *b = r.b;
  return true;
}

static thread_local std::string g_err;

void clear_error() { g_err.clear(); }

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? FFHIP_ERR_OOM : FFHIP_ERR_HIP, "hipMalloc(%zu) failed: %s", bytes,
                hipGetErrorString(e));
  }
  e = hipMemset(*p, 0, bytes);
  if (e != hipSuccess) {
    (void)hipFree(*p);
    *p = nullptr;
    return fail(FFHIP_ERR_HIP, "hipMemset failed: %s", hipGetErrorString(e));
  }
  return FFHIP_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

static int g_lds = 0;

int max_lds_per_block() {
  if (g_lds == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      g_lds = v;
    else
      g_lds = 65536;
  }
  return g_lds;
}

// getaliasedwavenumbers (src/domains.jl:408-421)
static void aliased_range(int64_t nk, double f, int64_t* lo, int64_t* hi) {
  const double L = (1 - f) / 2;
  const double R = (1 + f) / 2;
  *lo = (int64_t)std::floor(L * (double)nk) + 1;
  *hi = (int64_t)std::ceil(R * (double)nk);
}

}  // namespace ffhip

using namespace ffhip;

extern "C" {

const char* ffhip_last_error(void) { return g_err.c_str(); }

int ffhip_version(int* major, int* minor, int* patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return FFHIP_OK;
}

int ffhip_init(int device) {
  clear_error();
  int n = 0;
  FF_HIP(hipGetDeviceCount(&n));
  FF_CHECK(device >= 0 && device < n, FFHIP_ERR_ARG, "device %d out of range (%d devices)", device, n);
  FF_HIP(hipSetDevice(device));
  g_lds = 0;
  (void)max_lds_per_block();
  return red_scratch_setup(device);
}

int ffhip_malloc(void** ptr, size_t bytes) {
  clear_error();
  FF_CHECK(ptr, FFHIP_ERR_ARG, "null output pointer");
  return dev_alloc(ptr, bytes);
}

int ffhip_free(void* ptr) {
  clear_error();
  if (ptr) FF_HIP(hipFree(ptr));
  return FFHIP_OK;
}

int ffhip_memcpy_htod(void* dst, const void* src, size_t bytes) {
  clear_error();
  FF_CHECK(dst && src, FFHIP_ERR_ARG, "null pointer");
  FF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FFHIP_OK;
}

int ffhip_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  clear_error();
  FF_CHECK(dst && src, FFHIP_ERR_ARG, "null pointer");
  FF_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FFHIP_OK;
}

int ffhip_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  clear_error();
  FF_CHECK(dst && src, FFHIP_ERR_ARG, "null pointer");
  FF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return FFHIP_OK;
}

int ffhip_memset_zero(void* dst, size_t bytes, void* stream) {
  clear_error();
  FF_CHECK(dst, FFHIP_ERR_ARG, "null pointer");
  FF_HIP(hipMemsetAsync(dst, 0, bytes, as_stream(stream)));
  return FFHIP_OK;
}

int ffhip_synchronize(void) {
  clear_error();
  FF_HIP(hipDeviceSynchronize());
  return FFHIP_OK;
}

int ffhip_profile_start(void) {
  clear_error();
  g_prof.clear();
  g_evnext = 0;
  g_prof_on = true;
  return FFHIP_OK;
}

int ffhip_profile_stop(char* buf, size_t len) {
  clear_error();
  g_prof_on = false;
  FF_HIP(hipDeviceSynchronize());
  struct Agg {
    std::string name;
    long count = 0;
    double ms = 0, bytes = 0;
  };
  std::vector<Agg> agg;
  for (const ProfRec& r : g_prof) {
    float ms = 0;
    FF_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    Agg* a = nullptr;
    for (auto& x : agg)
      if (x.name == r.name) a = &x;
    if (!a) {
      agg.push_back(Agg());
      a = &agg.back();
      a->name = r.name;
    }
    a->count += 1;
    a->ms += ms;
    a->bytes += r.bytes;
  }
  std::string js = "{\"kernels\": [";
  for (size_t i = 0; i < agg.size(); ++i) {
    char tmp[512];
    snprintf(tmp, sizeof tmp, "%s{\"name\": \"%s\", \"count\": %ld, \"total_ms\": %.6f, \"bytes\": %.1f}",
             i ? ", " : "", agg[i].name.c_str(), agg[i].count, agg[i].ms, agg[i].bytes);
    js += tmp;
  }
  js += "]}";
  g_prof.clear();
  g_evnext = 0;
  if (buf && len) {
    snprintf(buf, len, "%s", js.c_str());
    FF_CHECK(js.size() < len, FFHIP_ERR_ARG, "profile buffer too small (%zu bytes needed)", js.size() + 1);
  }
  return FFHIP_OK;
}

int ffhip_aliased_wavenumbers(int64_t nk, int64_t nkr, double f, int64_t* klo, int64_t* khi, int64_t* krlo,
                              int64_t* krhi) {
  clear_error();
  FF_CHECK(f < 1, FFHIP_ERR_ARG, "`aliased_fraction` must be less than 1");
  FF_CHECK(nk >= 1 && nkr >= 1, FFHIP_ERR_ARG, "sizes must be positive");
  int64_t lo = 1, hi = 0, rlo = 1, rhi = 0;
  if (f > 0) {
    aliased_range(nk, f, &lo, &hi);
    rlo = lo;
    rhi = nkr;
  }
  if (klo) *klo = lo;
  if (khi) *khi = hi;
  if (krlo) *krlo = rlo;
  if (krhi) *krhi = rhi;
  return FFHIP_OK;
}

// getetdcoeffs (src/timesteppers.jl:689-718), FP64 complex arithmetic.
int ffhip_etdcoeffs(double dt, const double* Lre, const double* Lim, int64_t count, int ncirc, double rcirc,
                    double* zr, double* ar, double* br, double* gr, double* zi, double* ai, double* bi,
                    double* gi) {
  clear_error();
  FF_CHECK(Lre && zr && ar && br && gr, FFHIP_ERR_ARG, "null pointer");
  FF_CHECK(ncirc >= 1, FFHIP_ERR_ARG, "ncirc must be >= 1");
  FF_CHECK(!Lim || (zi && ai && bi && gi), FFHIP_ERR_ARG, "complex L needs imaginary outputs");
  using cd = std::complex<double>;
  std::vector<cd> circ(ncirc);
  const double pi = 3.14159265358979323846264338327950288;
  for (int j = 0; j < ncirc; ++j) circ[j] = rcirc * std::exp(cd(0.0, 2 * pi / ncirc * (j + 0.5)));
  for (int64_t i = 0; i < count; ++i) {
    const cd hL = cd(dt * Lre[i], Lim ? dt * Lim[i] : 0.0);
    cd sz = 0, sa = 0, sb = 0, sg = 0;
    for (int j = 0; j < ncirc; ++j) {
      const cd z = hL + circ[j];
      const cd ez = std::exp(z);
      const cd z3 = z * z * z;
      sz += (std::exp(z / 2.0) - 1.0) / z;
      sa += (-4.0 - z + ez * (4.0 - 3.0 * z + z * z)) / z3;
      sb += (2.0 + z + ez * (-2.0 + z)) / z3;
      sg += (-4.0 - 3.0 * z - z * z + ez * (4.0 - z)) / z3;
    }
    const double nc = (double)ncirc;  // mean = sum / ncirc
    const cd Z = dt * (sz / nc), A = dt * (sa / nc), B = dt * (sb / nc), G = dt * (sg / nc);
    zr[i] = Z.real();
    ar[i] = A.real();
    br[i] = B.real();
    gr[i] = G.real();
    if (Lim) {
      zi[i] = Z.imag();
      ai[i] = A.imag();
      bi[i] = B.imag();
      gi[i] = G.imag();
    }
  }
  return FFHIP_OK;
}

}  // extern "C"
