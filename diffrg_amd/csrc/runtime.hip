// runtime.hip -- device context, memory, streams, events and the scratch pools of libdfrg.so.
#include "internal.hh"

#include <map>
#include <vector>

namespace dfrg
{
  namespace
  {
    thread_local std::string g_error;
    thread_local hipStream_t g_thread_stream = nullptr;
    thread_local int g_thread_stream_device = -1;

    // Size-class caching pools: blocks are never returned to the driver before dfrg_finalize,
    // so per-call scratch costs no hipMalloc after warm-up.
    struct Pool {
      std::mutex mutex;
      std::map<size_t, std::vector<void *>> free_blocks;
      std::map<void *, size_t> sizes;
    };
    Pool g_dev_pool, g_host_pool;

    // Scratch of the entry point that allocates nothing (dfrg_se_batch_run_noalloc): partial
    // slots, ticket counters and the work queue of se_kernel, DFRG_NOALLOC_SCRATCH_BYTES per
    // device, held until the process ends.  In host code `scratch` is only a handle: its device
    // address on each GPU is resolved once per device in dfrg_init (hipGetSymbolAddress) and kept in
    // g_scratch_ptr, which is what every host-side use takes.
    static __device__ __align__(256) unsigned char scratch[DFRG_NOALLOC_SCRATCH_BYTES];
    constexpr int kMaxDevices = 64;
    void *g_scratch_ptr[kMaxDevices] = {};

    size_t size_class(size_t bytes)
    {
      size_t c = 256;
      while (c < bytes) c <<= 1;
      return c;
    }
  } // namespace

  void set_error(const char *fmt, ...)
  {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
  }
  void clear_error() { g_error.clear(); }

  Context &context()
  {
    static Context ctx;
    return ctx;
  }

  int require_init()
  {
    Context &c = context();
    if (!c.initialised) {
      set_error("dfrg_init has not been called");
      return DFRG_ERR_NOT_INIT;
    }
    int cur = -1;
    DFRG_HIP(hipGetDevice(&cur));
    if (cur != c.device) DFRG_HIP(hipSetDevice(c.device));
    return DFRG_OK;
  }

  hipStream_t resolve_stream(void *stream)
  {
    if (stream) return (hipStream_t)stream;
    Context &c = context();
    if (g_thread_stream == nullptr || g_thread_stream_device != c.device) {
      hipStream_t s = nullptr;
      // default priority only
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
      g_thread_stream = s;
      g_thread_stream_device = c.device;
    }
    return g_thread_stream;
  }

  void *noalloc_scratch()
  {
    const int d = context().device;
    return d >= 0 && d < kMaxDevices ? g_scratch_ptr[d] : nullptr;
  }
} // namespace dfrg

using namespace dfrg;

extern "C" {

const char *dfrg_last_error(void) { return g_error.c_str(); }

int dfrg_device_count(int *count)
{
  DFRG_CHECK_ARG(count != nullptr, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    *count = 0;
    set_error("No HIP devices found");
    return DFRG_ERR_NO_DEVICE;
  }
  *count = n;
  return DFRG_OK;
}

int dfrg_init(int device)
{
  Context &c = context();
  std::lock_guard<std::mutex> lock(c.mutex);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    set_error("No HIP devices found");
    return DFRG_ERR_NO_DEVICE;
  }
  DFRG_CHECK_ARG(device >= 0 && device < n, "device %d out of range [0,%d)", device, n);
  if (c.initialised && c.device == device) {
    DFRG_HIP(hipSetDevice(device));
    return DFRG_OK;
  }
  DFRG_HIP(hipSetDevice(device));
  // this device's copy of the static scratch: resolve its address once, with the device
  // current, and zero it (its counters must start at zero)
  DFRG_CHECK_ARG(device < kMaxDevices, "device %d beyond the %d devices the library supports", device,
                 kMaxDevices);
  if (g_scratch_ptr[device] == nullptr) {
    void *p = nullptr;
    DFRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DFRG_HIP(hipMemset(p, 0, DFRG_NOALLOC_SCRATCH_BYTES));
    DFRG_HIP(hipDeviceSynchronize());
    g_scratch_ptr[device] = p;
  }
  hipDeviceProp_t prop;
  DFRG_HIP(hipGetDeviceProperties(&prop, device));
  c.device = device;
  c.cu_count = prop.multiProcessorCount;
  c.clock_khz = prop.clockRate;
  c.hbm_bytes = prop.totalGlobalMem;
  snprintf(c.name, sizeof(c.name), "%s", prop.gcnArchName);
  c.initialised = true;
  return DFRG_OK;
}

int dfrg_finalize(void)
{
  Context &c = context();
  std::lock_guard<std::mutex> lock(c.mutex);
  if (!c.initialised) return DFRG_OK;
  hipSetDevice(c.device);
  hipDeviceSynchronize();
  {
    std::lock_guard<std::mutex> l(g_dev_pool.mutex);
    for (auto &kv : g_dev_pool.sizes) hipFree(kv.first);
    g_dev_pool.sizes.clear();
    g_dev_pool.free_blocks.clear();
  }
  {
    std::lock_guard<std::mutex> l(g_host_pool.mutex);
    for (auto &kv : g_host_pool.sizes) hipHostFree(kv.first);
    g_host_pool.sizes.clear();
    g_host_pool.free_blocks.clear();
  }
  if (g_thread_stream) {
    hipStreamDestroy(g_thread_stream);
    g_thread_stream = nullptr;
  }
  c.initialised = false;
  return DFRG_OK;
}

int dfrg_device_props(char *name, int len, int *cu_count, int *clock_khz, unsigned long long *hbm_bytes)
{
  DFRG_TRY(require_init());
  Context &c = context();
  if (name && len > 0) snprintf(name, len, "%s", c.name);
  if (cu_count) *cu_count = c.cu_count;
  if (clock_khz) *clock_khz = c.clock_khz;
  if (hbm_bytes) *hbm_bytes = c.hbm_bytes;
  return DFRG_OK;
}

int dfrg_malloc(void **d_ptr, size_t bytes)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(d_ptr != nullptr, "d_ptr is NULL");
  DFRG_HIP(hipMalloc(d_ptr, bytes ? bytes : 1));
  return DFRG_OK;
}

int dfrg_free(void *d_ptr)
{
  DFRG_TRY(require_init());
  if (d_ptr) DFRG_HIP(hipFree(d_ptr));
  return DFRG_OK;
}

int dfrg_host_alloc(void **h_ptr, size_t bytes)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(h_ptr != nullptr, "h_ptr is NULL");
  DFRG_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return DFRG_OK;
}

int dfrg_host_free(void *h_ptr)
{
  DFRG_TRY(require_init());
  if (h_ptr) DFRG_HIP(hipHostFree(h_ptr));
  return DFRG_OK;
}

int dfrg_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, resolve_stream(stream)));
  return DFRG_OK;
}

int dfrg_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, resolve_stream(stream)));
  return DFRG_OK;
}

int dfrg_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, resolve_stream(stream)));
  return DFRG_OK;
}

int dfrg_thread_stream(void **stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(stream != nullptr, "stream is NULL");
  hipStream_t s = resolve_stream(nullptr);
  if (!s) {
    set_error("could not create the thread stream");
    return DFRG_ERR_HIP;
  }
  *stream = (void *)s;
  return DFRG_OK;
}

int dfrg_stream_create(void **stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(stream != nullptr, "stream is NULL");
  hipStream_t s;
  DFRG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return DFRG_OK;
}

int dfrg_stream_destroy(void *stream)
{
  DFRG_TRY(require_init());
  if (stream) DFRG_HIP(hipStreamDestroy((hipStream_t)stream));
  return DFRG_OK;
}

int dfrg_stream_sync(void *stream)
{
  DFRG_TRY(require_init());
  DFRG_HIP(hipStreamSynchronize(resolve_stream(stream)));
  return DFRG_OK;
}

int dfrg_event_create(void **event)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(event != nullptr, "event is NULL");
  hipEvent_t e;
  DFRG_HIP(hipEventCreate(&e));
  *event = (void *)e;
  return DFRG_OK;
}

int dfrg_event_destroy(void *event)
{
  DFRG_TRY(require_init());
  if (event) DFRG_HIP(hipEventDestroy((hipEvent_t)event));
  return DFRG_OK;
}

int dfrg_event_record(void *event, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_HIP(hipEventRecord((hipEvent_t)event, resolve_stream(stream)));
  return DFRG_OK;
}

int dfrg_event_sync(void *event)
{
  DFRG_TRY(require_init());
  DFRG_HIP(hipEventSynchronize((hipEvent_t)event));
  return DFRG_OK;
}

int dfrg_event_elapsed_ms(void *start, void *stop, float *ms)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(ms != nullptr, "ms is NULL");
  DFRG_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return DFRG_OK;
}

struct dfrg_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int dfrg_graph_begin(void *stream)
{
  DFRG_TRY(require_init());
  DFRG_HIP(hipStreamBeginCapture(resolve_stream(stream), hipStreamCaptureModeThreadLocal));
  return DFRG_OK;
}

int dfrg_graph_end(void *stream, dfrg_graph **graph)
{
  DFRG_TRY(require_init());
  hipGraph_t g = nullptr;
  // the capture ends in every case, so a failed argument check does not leave the stream capturing
  const hipError_t e = hipStreamEndCapture(resolve_stream(stream), &g);
  if (graph == nullptr) {
    if (g) hipGraphDestroy(g);
    DFRG_CHECK_ARG(false, "graph is NULL");
  }
  *graph = nullptr;
  if (e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
    return DFRG_ERR_HIP;
  }
  auto *out = new dfrg_graph;
  out->graph = g;
  const hipError_t ei = hipGraphInstantiate(&out->exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    hipGraphDestroy(g);
    delete out;
    set_error("hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    return DFRG_ERR_HIP;
  }
  *graph = out;
  return DFRG_OK;
}

int dfrg_graph_launch(dfrg_graph *graph, void *stream)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(graph != nullptr && graph->exec != nullptr, "NULL graph");
  DFRG_HIP(hipGraphLaunch(graph->exec, resolve_stream(stream)));
  return DFRG_OK;
}

int dfrg_graph_destroy(dfrg_graph *graph)
{
  if (!graph) return DFRG_OK;
  if (graph->exec) hipGraphExecDestroy(graph->exec);
  if (graph->graph) hipGraphDestroy(graph->graph);
  delete graph;
  return DFRG_OK;
}

int dfrg_pool_alloc(void **d_ptr, size_t bytes)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(d_ptr != nullptr, "d_ptr is NULL");
  const size_t c = size_class(bytes);
  {
    std::lock_guard<std::mutex> l(g_dev_pool.mutex);
    auto &v = g_dev_pool.free_blocks[c];
    if (!v.empty()) {
      *d_ptr = v.back();
      v.pop_back();
      return DFRG_OK;
    }
  }
  void *p = nullptr;
  DFRG_HIP(hipMalloc(&p, c));
  DFRG_HIP(hipMemset(p, 0, c));
  DFRG_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> l(g_dev_pool.mutex);
  g_dev_pool.sizes[p] = c;
  *d_ptr = p;
  return DFRG_OK;
}

int dfrg_pool_free(void *d_ptr)
{
  if (!d_ptr) return DFRG_OK;
  std::lock_guard<std::mutex> l(g_dev_pool.mutex);
  auto it = g_dev_pool.sizes.find(d_ptr);
  DFRG_CHECK_ARG(it != g_dev_pool.sizes.end(), "pointer %p was not allocated by dfrg_pool_alloc", d_ptr);
  g_dev_pool.free_blocks[it->second].push_back(d_ptr);
  return DFRG_OK;
}

int dfrg_pool_host_alloc(void **h_ptr, size_t bytes)
{
  DFRG_TRY(require_init());
  DFRG_CHECK_ARG(h_ptr != nullptr, "h_ptr is NULL");
  const size_t c = size_class(bytes);
  {
    std::lock_guard<std::mutex> l(g_host_pool.mutex);
    auto &v = g_host_pool.free_blocks[c];
    if (!v.empty()) {
      *h_ptr = v.back();
      v.pop_back();
      return DFRG_OK;
    }
  }
  void *p = nullptr;
  DFRG_HIP(hipHostMalloc(&p, c, hipHostMallocDefault));
  std::lock_guard<std::mutex> l(g_host_pool.mutex);
  g_host_pool.sizes[p] = c;
  *h_ptr = p;
  return DFRG_OK;
}

int dfrg_pool_host_free(void *h_ptr)
{
  if (!h_ptr) return DFRG_OK;
  std::lock_guard<std::mutex> l(g_host_pool.mutex);
  auto it = g_host_pool.sizes.find(h_ptr);
  DFRG_CHECK_ARG(it != g_host_pool.sizes.end(), "pointer %p was not allocated by dfrg_pool_host_alloc", h_ptr);
  g_host_pool.free_blocks[it->second].push_back(h_ptr);
  return DFRG_OK;
}

} // extern "C"
