// Batched sectors: every sector of a batch gets one apply H C0 and a
// fixed-length Taylor series exp(-i t H) C0 = sum_{k<=nterms} (-i t H)^k
// C0 / k!, all recorded once into a HIP graph and replayed with a single
// hipGraphLaunch.  The sectors are spread over several capture streams
// (longest-processing-time order, each to the least-loaded stream) so the
// graph has parallel branches: small sectors, which cannot fill the GPU
// alone, run side by side.  Each sector keeps its own kernel variant (the
// tile choice of its descriptor, picked by the host autotuner).
#include "fqehip_common.h"

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

struct fqehip_batch {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  int nodes = 0;
};

namespace fqehip {
namespace {

size_t esize(int precision) {
  return precision == FQEHIP_C128 ? 2 * sizeof(double) : 2 * sizeof(float);
}

// Work estimate of one apply of a sector (for the stream assignment).
double item_work(const fqehip_batch_item& it) {
  const double n = static_cast<double>(it.sector.lena) * it.sector.lenb;
  return n * (it.sector.npair + 8.0);
}

void release(fqehip_batch* b) {
  if (!b) return;
  // teardown: failures here leave nothing to recover, so they are ignored
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  delete b;
}

// The work of one item on stream st (see the file comment).
int enqueue_item(int precision, const fqehip_batch_item& it, double time,
                 int nterms, hipStream_t st, int* nodes) {
  const int64_t n = it.sector.lena * it.sector.lenb;
  if (n == 0) return FQEHIP_OK;
  void* sv = reinterpret_cast<void*>(st);
  if (hipMemcpyAsync(it.d_term, it.d_coeff0, n * esize(precision),
                     hipMemcpyDeviceToDevice, st) != hipSuccess) {
    set_error("batch: copy of the initial term failed");
    return FQEHIP_ERR_HIP;
  }
  int rc = fqehip_convert_to_c128(precision, n, it.d_coeff0, 1.0, 0.0,
                                  it.d_result, sv);
  if (rc) return rc;
  *nodes += 2;
  if (it.d_apply_out) {
    rc = fqehip_apply(precision, &it.sector, &it.op, it.d_coeff0, 1.0, 0.0,
                      0.0, 0.0, nullptr, it.d_apply_out, nullptr, nullptr,
                      nullptr, it.d_workspace, sv);
    if (rc) return rc;
    *nodes += 3;
  }
  for (int k = 1; k <= nterms; ++k) {
    rc = fqehip_apply(precision, &it.sector, &it.op, it.d_term, 0.0,
                      -time / static_cast<double>(k), 0.0, 0.0, nullptr,
                      it.d_term, it.d_result, nullptr, nullptr,
                      it.d_workspace, sv);
    if (rc) return rc;
    *nodes += 3;
  }
  return FQEHIP_OK;
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_batch_create(int precision, int nitems,
                                   const fqehip_batch_item* items,
                                   double time, int nterms, int nstreams,
                                   void* stream, fqehip_batch** out) {
  clear_error();
  FQEHIP_CHECK_ARG(out != nullptr, "NULL output handle");
  *out = nullptr;
  FQEHIP_CHECK_ARG(precision == FQEHIP_C128 || precision == FQEHIP_C64,
                   "precision=%d", precision);
  FQEHIP_CHECK_ARG(nitems >= 0 && (nitems == 0 || items), "nitems=%d", nitems);
  FQEHIP_CHECK_ARG(nterms >= 0 && nstreams >= 1 && nstreams <= 32,
                   "nterms=%d nstreams=%d", nterms, nstreams);
  for (int i = 0; i < nitems; ++i) {
    const fqehip_batch_item& it = items[i];
    const int64_t n = it.sector.lena * it.sector.lenb;
    FQEHIP_CHECK_ARG(it.op.a && !it.op.h1, "item %d: two-body operator needed", i);
    FQEHIP_CHECK_ARG(n == 0 || (it.d_coeff0 && it.d_term && it.d_result &&
                                it.d_workspace),
                     "item %d: NULL buffer", i);
    FQEHIP_CHECK_ARG(it.d_term != it.d_coeff0, "item %d: term aliases coeff0", i);
  }
  fqehip_batch* b = new (std::nothrow) fqehip_batch;
  if (!b) {
    set_error("batch: out of host memory");
    return FQEHIP_ERR_OOM;
  }
  // The graph is recorded on a stream of its own (a caller's stream may be
  // the legacy default stream, which cannot be captured); `stream` only
  // receives the launches.
  (void)stream;
  const int nb = std::min(nstreams, std::max(nitems, 1));
  for (int s = 0; s < nb + 1; ++s) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      release(b);
      set_error("batch: stream creation failed");
      return FQEHIP_ERR_HIP;
    }
    b->streams.push_back(st);
  }
  for (int e = 0; e < nb + 1; ++e) {
    hipEvent_t ev;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      release(b);
      set_error("batch: event creation failed");
      return FQEHIP_ERR_HIP;
    }
    b->events.push_back(ev);
  }
  hipStream_t origin = b->streams[nb];
  // longest processing time first, each to the least-loaded stream
  std::vector<int> order(nitems);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    return item_work(items[x]) > item_work(items[y]);
  });
  std::vector<double> load(nb, 0.0);
  std::vector<std::vector<int>> plan(nb);
  for (int i : order) {
    const int s = static_cast<int>(
        std::min_element(load.begin(), load.end()) - load.begin());
    plan[s].push_back(i);
    load[s] += item_work(items[i]);
  }
  if (hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal) !=
      hipSuccess) {
    release(b);
    set_error("batch: stream capture could not begin");
    return FQEHIP_ERR_HIP;
  }
  int rc = FQEHIP_OK;
  if (hipEventRecord(b->events[nb], origin) != hipSuccess) rc = FQEHIP_ERR_HIP;
  for (int s = 0; s < nb && rc == FQEHIP_OK; ++s)
    if (hipStreamWaitEvent(b->streams[s], b->events[nb], 0) != hipSuccess)
      rc = FQEHIP_ERR_HIP;
  for (int s = 0; s < nb && rc == FQEHIP_OK; ++s)
    for (int i : plan[s]) {
      rc = enqueue_item(precision, items[i], time, nterms, b->streams[s],
                        &b->nodes);
      if (rc) break;
    }
  for (int s = 0; s < nb && rc == FQEHIP_OK; ++s) {
    if (hipEventRecord(b->events[s], b->streams[s]) != hipSuccess ||
        hipStreamWaitEvent(origin, b->events[s], 0) != hipSuccess)
      rc = FQEHIP_ERR_HIP;
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(origin, &graph);
  if (rc != FQEHIP_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    release(b);
    if (rc == FQEHIP_OK) {
      set_error("batch: stream capture failed: %s", hipGetErrorString(ec));
      rc = FQEHIP_ERR_HIP;
    }
    return rc;
  }
  b->graph = graph;
  if (hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    release(b);
    set_error("batch: graph instantiation failed");
    return FQEHIP_ERR_HIP;
  }
  *out = b;
  return FQEHIP_OK;
}

extern "C" int fqehip_batch_launch(fqehip_batch* b, void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(b && b->exec, "NULL batch");
  const hipError_t e = hipGraphLaunch(b->exec, as_stream(stream));
  if (e != hipSuccess) {
    set_error("batch: graph launch failed: %s", hipGetErrorString(e));
    return FQEHIP_ERR_HIP;
  }
  return FQEHIP_OK;
}

extern "C" int fqehip_batch_nodes(const fqehip_batch* b) {
  return b ? b->nodes : 0;
}

extern "C" void fqehip_batch_destroy(fqehip_batch* b) { release(b); }
