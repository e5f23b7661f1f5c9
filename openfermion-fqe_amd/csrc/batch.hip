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
  fqehip::GroupSector* d_group = nullptr;   // grouped sectors (device)
  int ngroup = 0;
  int nodes = 0;
};

namespace fqehip {

// Defined in fused.hip.
int launch_group_rows(int precision, const GroupSector* d_secs, int nsec,
                      int64_t rows, int64_t max_lenb, int step,
                      hipStream_t st);
int launch_group_reduce(int precision, const GroupSector* d_secs, int nsec,
                        int64_t blocks, int step, double ar, double ai,
                        hipStream_t st);
int group_reduce_threads();

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
  if (b->d_group) (void)hipFree(b->d_group);
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
    // after the sector's first apply its workspace holds the operator's
    // tables
    rc = apply_reusing_tables(precision, &it.sector, &it.op, it.d_term, 0.0,
                              -time / static_cast<double>(k), 0.0, 0.0,
                              nullptr, it.d_term, it.d_result, nullptr,
                              nullptr, it.d_workspace, sv,
                              it.d_apply_out != nullptr || k > 1);
    if (rc) return rc;
    *nodes += 3;
  }
  return FQEHIP_OK;
}

// The grouped sectors on stream st: per sector the copy of C0 into the
// term buffer and the FP64 initial sum, then per step one grouped row
// launch and one grouped reduction over all of them.
int enqueue_group(int precision, const fqehip_batch_item* items,
                  const std::vector<int>& grouped, const GroupSector* d_secs,
                  int64_t rows, int64_t blocks, int64_t max_lenb,
                  bool any_apply, double time, int nterms, hipStream_t st,
                  int* nodes) {
  void* sv = reinterpret_cast<void*>(st);
  for (int i : grouped) {
    const fqehip_batch_item& it = items[i];
    const int64_t n = it.sector.lena * it.sector.lenb;
    if (hipMemcpyAsync(it.d_term, it.d_coeff0, n * esize(precision),
                       hipMemcpyDeviceToDevice, st) != hipSuccess) {
      set_error("batch: copy of the initial term failed");
      return FQEHIP_ERR_HIP;
    }
    int rc = fqehip_convert_to_c128(precision, n, it.d_coeff0, 1.0, 0.0,
                                    it.d_result, sv);
    if (rc) return rc;
    *nodes += 2;
  }
  const int nsec = static_cast<int>(grouped.size());
  for (int step = any_apply ? 0 : 1; step <= nterms; ++step) {
    int rc = launch_group_rows(precision, d_secs, nsec, rows, max_lenb, step,
                               st);
    if (rc) return rc;
    rc = launch_group_reduce(precision, d_secs, nsec, blocks, step, 0.0,
                             step ? -time / static_cast<double>(step) : 0.0, st);
    if (rc) return rc;
    *nodes += 2;
  }
  return FQEHIP_OK;
}

}  // namespace
}  // namespace fqehip

using namespace fqehip;

extern "C" int fqehip_batch_create(int precision, int nitems,
                                   const fqehip_batch_item* items,
                                   double time, int nterms, int nstreams,
                                   int group, void* stream,
                                   fqehip_batch** out) {
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
  // sectors of the fused row kinds (no split: complex64 rows4/rows8,
  // complex128 k_fused_f64) run as one group: one row launch and one
  // reduction per step for all of them
  std::vector<int> grouped, single;
  for (int i = 0; i < nitems; ++i) {
    const fqehip_batch_item& it = items[i];
    const bool g = group && it.sector.lena * it.sector.lenb > 0 &&
                   !it.sector.hbb_hi && !it.sector.hbb_f64 &&
                   it.sector.tile != FQEHIP_TILE_OZAKI &&
                   it.sector.nslot == it.sector.nalpha * (it.sector.norb - it.sector.nalpha) &&
                   fqehip_fused_supported(precision, &it.sector, &it.op);
    (g ? grouped : single).push_back(i);
  }
  int64_t grows = 0, gblocks = 0, gmax_lenb = 0;
  bool gapply = false;
  double gwork = 0.0;
  if (!grouped.empty()) {
    std::vector<GroupSector> hs(grouped.size());
    const int nt = group_reduce_threads();
    for (size_t j = 0; j < grouped.size(); ++j) {
      const fqehip_batch_item& it = items[grouped[j]];
      const fqehip_sector& sc = it.sector;
      const int64_t n = sc.lena * sc.lenb;
      GroupSector& g = hs[j];
      g.norb = sc.norb;
      g.compressed = sc.npair == sc.norb * (sc.norb + 1) / 2 ? 1 : 0;
      g.npair = sc.npair;
      g.nslot = sc.nslot;
      g.lena = sc.lena;
      g.lenb = sc.lenb;
      g.row0 = grows;
      g.blk0 = gblocks;
      g.amat = it.op.a;
      g.pta = sc.ptab_a;
      g.ptb = sc.ptab_b;
      g.fwa = sc.fwd_a;
      g.fwb = sc.fwd_b;
      g.sla = sc.slot_a;
      g.coeff0 = it.d_coeff0;
      g.apply_out = it.d_apply_out;
      g.term = it.d_term;
      g.result = it.d_result;
      g.slots = it.d_workspace;
      g.rowacc = static_cast<char*>(it.d_workspace) +
                 static_cast<int64_t>(sc.nslot) * n * esize(precision);
      grows += sc.lena;
      gblocks += (n + nt - 1) / nt;
      gmax_lenb = std::max(gmax_lenb, sc.lenb);
      gapply = gapply || it.d_apply_out != nullptr;
      gwork += item_work(it);
    }
    const size_t bytes = hs.size() * sizeof(GroupSector);
    void* dg = nullptr;
    if (hipMalloc(&dg, bytes) != hipSuccess) {
      release(b);
      set_error("batch: group descriptor allocation failed");
      return FQEHIP_ERR_OOM;
    }
    b->d_group = static_cast<GroupSector*>(dg);
    if (hipMemcpy(b->d_group, hs.data(), bytes, hipMemcpyHostToDevice) !=
        hipSuccess) {
      release(b);
      set_error("batch: group descriptor upload failed");
      return FQEHIP_ERR_HIP;
    }
    b->ngroup = static_cast<int>(grouped.size());
  }
  // longest processing time first, each to the least-loaded stream (the
  // group occupies stream 0)
  std::vector<int> order(single);
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    return item_work(items[x]) > item_work(items[y]);
  });
  std::vector<double> load(nb, 0.0);
  load[0] = gwork;
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
  if (rc == FQEHIP_OK && !grouped.empty())
    rc = enqueue_group(precision, items, grouped, b->d_group, grows, gblocks,
                       gmax_lenb, gapply, time, nterms, b->streams[0],
                       &b->nodes);
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

extern "C" int fqehip_batch_grouped(const fqehip_batch* b) {
  return b ? b->ngroup : 0;
}

extern "C" void fqehip_batch_destroy(fqehip_batch* b) { release(b); }
