/* runtime.cpp -- device initialisation, memory, copies, events and error reporting behind
 * the C-ABI.  Replaces the reference's CUDA runtime utilities (util/cuda/allocator.h,
 * cudamemcpy.h, pinner.h, check.h, timer.h, initialize.cu) with HIP runtime calls on one
 * library-owned stream. */
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include "internal.h"

namespace ispm {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int set_error_message(int code, const char *msg) { return fail(code, "%s", msg); }

int hip_status(hipError_t e, const char *what)
{
  if (e == hipSuccess) return ISPM_OK;
  int code = (e == hipErrorOutOfMemory) ? ISPM_ERR_NOMEM : ISPM_ERR_RUNTIME;
  return fail(code, "%s: %s", what, hipGetErrorString(e));
}

namespace {
struct device_state {
  bool ready = false;
  int init_status = ISPM_OK;
  hipStream_t own = nullptr;
  hipStream_t user = nullptr;
  reduce_scratch red{nullptr};
  int cus = 0;
};
std::mutex g_mu;
device_state g_dev[kMaxDevices];

device_state &cur()
{
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices) d = 0;
  return g_dev[d];
}

int init_locked(device_state &s)
{
  if (s.ready) return s.init_status;
  s.ready = true;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return s.init_status = hip_status(e, "hipGetDevice");
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) return s.init_status = hip_status(e, "hipGetDeviceProperties");
  /* The library carries gfx950 code objects only (reference: arch-mismatch abort,
   * initialize.cu:21-27). */
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return s.init_status = fail(ISPM_ERR_DEVICE, "device %d is %s; this library is built for gfx950",
                                dev, p.gcnArchName);
  s.cus = p.multiProcessorCount;
  e = hipStreamCreateWithFlags(&s.own, hipStreamNonBlocking);
  if (e != hipSuccess) return s.init_status = hip_status(e, "hipStreamCreate");
  e = hipMalloc((void **)&s.red.partials, sizeof(double) * kMaxBlocks * kMaxSums);
  if (e != hipSuccess) return s.init_status = hip_status(e, "hipMalloc(scratch)");
  int st = noalloc_scratch_init(dev);
  if (st) return s.init_status = st;
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return s.init_status = hip_status(e, "hipDeviceSynchronize");
  return s.init_status = ISPM_OK;
}
}  // namespace

int ensure_init()
{
  device_state &s = cur();
  if (s.ready) return s.init_status;
  std::lock_guard<std::mutex> g(g_mu);
  return init_locked(s);
}

hipStream_t stream()
{
  device_state &s = cur();
  return s.user ? s.user : s.own;
}

reduce_scratch &scratch() { return cur().red; }

bool initialized()
{
  const device_state &s = cur();
  return s.ready && s.init_status == ISPM_OK;
}

int num_cus() { return cur().cus; }

size_t type_size(int t)
{
  switch (t) {
    case ISPM_F32: return 4;
    case ISPM_F64: return 8;
    case ISPM_BF16: return 2;
    default: return 0;
  }
}

}  // namespace ispm

using namespace ispm;

extern "C" {

const char *ispm_last_error(void) { return g_err; }

int ispm_init(void) { return ensure_init(); }

int ispm_device_string(char *buf, int len)
{
  if (!buf || len <= 0) return fail(ISPM_ERR_INVALID, "ispm_device_string: null buffer");
  int dev = 0;
  ISPM_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  ISPM_HIP(hipGetDeviceProperties(&p, dev));
  /* the marketing name can be empty on this part: fall back to the ISA name without its
   * feature suffixes (":sramecc+:xnack-"), so that the string fits the reference's 32 bytes */
  char arch[sizeof p.gcnArchName];
  std::snprintf(arch, sizeof arch, "%s", p.gcnArchName);
  if (char *c = std::strchr(arch, ':')) *c = 0;
  const char *name = p.name[0] ? p.name : arch;
  snprintf(buf, len, "%04x:%02x:%02x (%s)", p.pciDomainID, p.pciBusID, p.pciDeviceID, name);
  return ISPM_OK;
}

int ispm_malloc(void **p, size_t bytes)
{
  if (!p) return fail(ISPM_ERR_INVALID, "ispm_malloc: null out pointer");
  *p = nullptr;
  int st = ensure_init();
  if (st) return st;
  if (bytes == 0) return ISPM_OK;
  ISPM_HIP(hipMalloc(p, bytes));
  /* zero-filled like cuda_device_allocator (util/cuda/allocator.h:17-18) */
  ISPM_HIP(hipMemsetAsync(*p, 0, bytes, stream()));
  ISPM_HIP(hipStreamSynchronize(stream()));
  return ISPM_OK;
}

int ispm_free(void *p)
{
  if (!p) return ISPM_OK;
  ISPM_HIP(hipStreamSynchronize(stream()));
  ISPM_HIP(hipFree(p));
  return ISPM_OK;
}

static hipMemcpyKind kind_of(int kind)
{
  switch (kind) {
    case ISPM_H2D: return hipMemcpyHostToDevice;
    case ISPM_D2H: return hipMemcpyDeviceToHost;
    case ISPM_D2D: return hipMemcpyDeviceToDevice;
    default: return hipMemcpyDefault;
  }
}

int ispm_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
  if (bytes == 0) return ISPM_OK;
  if (!dst || !src) return fail(ISPM_ERR_INVALID, "ispm_memcpy: null pointer");
  int st = ensure_init();
  if (st) return st;
  ISPM_HIP(hipMemcpyAsync(dst, src, bytes, kind_of(kind), stream()));
  ISPM_HIP(hipStreamSynchronize(stream()));
  return ISPM_OK;
}

int ispm_memcpy_async(void *dst, const void *src, size_t bytes, int kind)
{
  if (bytes == 0) return ISPM_OK;
  if (!dst || !src) return fail(ISPM_ERR_INVALID, "ispm_memcpy_async: null pointer");
  int st = ensure_init();
  if (st) return st;
  ISPM_HIP(hipMemcpyAsync(dst, src, bytes, kind_of(kind), stream()));
  return ISPM_OK;
}

int ispm_host_alloc_mapped(void **hp, size_t bytes)
{
  if (!hp) return fail(ISPM_ERR_INVALID, "ispm_host_alloc_mapped: null out pointer");
  int st = ensure_init();
  if (st) return st;
  ISPM_HIP(hipHostMalloc(hp, bytes ? bytes : 1, hipHostMallocMapped));
  std::memset(*hp, 0, bytes);
  return ISPM_OK;
}

int ispm_host_free(void *hp)
{
  if (!hp) return ISPM_OK;
  ISPM_HIP(hipHostFree(hp));
  return ISPM_OK;
}

int ispm_host_device_pointer(void **dp, void *hp)
{
  if (!dp || !hp) return fail(ISPM_ERR_INVALID, "ispm_host_device_pointer: null pointer");
  ISPM_HIP(hipHostGetDevicePointer(dp, hp, 0));
  return ISPM_OK;
}

int ispm_synchronize(void)
{
  int st = ensure_init();
  if (st) return st;
  ISPM_HIP(hipStreamSynchronize(stream()));
  return ISPM_OK;
}

void *ispm_get_stream(void)
{
  if (ensure_init()) return nullptr;
  return (void *)stream();
}

int ispm_set_stream(void *s)
{
  int st = ensure_init();
  if (st) return st;
  cur().user = (hipStream_t)s;
  return ISPM_OK;
}

int ispm_event_create(void **ev)
{
  if (!ev) return fail(ISPM_ERR_INVALID, "ispm_event_create: null out pointer");
  int st = ensure_init();
  if (st) return st;
  hipEvent_t e;
  ISPM_HIP(hipEventCreate(&e));
  *ev = (void *)e;
  return ISPM_OK;
}

int ispm_event_record(void *ev)
{
  ISPM_HIP(hipEventRecord((hipEvent_t)ev, stream()));
  return ISPM_OK;
}

int ispm_event_synchronize(void *ev)
{
  ISPM_HIP(hipEventSynchronize((hipEvent_t)ev));
  return ISPM_OK;
}

int ispm_event_elapsed_ms(void *a, void *b, float *ms)
{
  if (!ms) return fail(ISPM_ERR_INVALID, "ispm_event_elapsed_ms: null out pointer");
  ISPM_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return ISPM_OK;
}

int ispm_event_destroy(void *ev)
{
  if (!ev) return ISPM_OK;
  ISPM_HIP(hipEventDestroy((hipEvent_t)ev));
  return ISPM_OK;
}

}  // extern "C"
