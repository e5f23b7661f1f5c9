// A whole Taylor series (fqehip_taylor) recorded once into a HIP graph and
// replayed with one launch per time_evolve: every term's apply, norm
// reduction and device-flag test become graph nodes, so a series costs one
// host call however many terms it has.  Terms after convergence still run
// as nodes but their kernels return at once on the device flag, as in the
// eager series.
#include "fqehip_common.h"

struct fqehip_taylor_graph {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void release(fqehip_taylor_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

}  // namespace

using namespace fqehip;

extern "C" int fqehip_taylor_graph_create(
    int precision, const fqehip_sector* s, const fqehip_operator* op,
    double time, double accuracy, int expansion, void* d_term,
    double* d_result, int32_t* d_status, double* d_partials, double* d_norm,
    void* d_workspace, fqehip_taylor_graph** out) {
  clear_error();
  FQEHIP_CHECK_ARG(out != nullptr, "NULL output handle");
  *out = nullptr;
  fqehip_taylor_graph* g = new (std::nothrow) fqehip_taylor_graph();
  if (!g) {
    set_error("fqehip_taylor_graph_create: out of host memory");
    return FQEHIP_ERR_OOM;
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    release(g);
    set_error("fqehip_taylor_graph_create: hipStreamCreateWithFlags failed");
    return FQEHIP_ERR_HIP;
  }
  if (hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal) !=
      hipSuccess) {
    release(g);
    set_error("fqehip_taylor_graph_create: hipStreamBeginCapture failed");
    return FQEHIP_ERR_HIP;
  }
  const int rc = fqehip_taylor(precision, s, op, time, accuracy, expansion,
                               d_term, d_result, d_status, d_partials, d_norm,
                               d_workspace, g->stream);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
  g->graph = graph;
  if (rc != FQEHIP_OK) {
    release(g);
    return rc;                             // message set by fqehip_taylor
  }
  if (ec != hipSuccess || !graph) {
    release(g);
    set_error("fqehip_taylor_graph_create: capture failed: %s",
              hipGetErrorString(ec));
    return FQEHIP_ERR_HIP;
  }
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) !=
      hipSuccess) {
    release(g);
    set_error("fqehip_taylor_graph_create: hipGraphInstantiate failed");
    return FQEHIP_ERR_HIP;
  }
  *out = g;
  return FQEHIP_OK;
}

extern "C" int fqehip_taylor_graph_launch(fqehip_taylor_graph* g,
                                          void* stream) {
  clear_error();
  FQEHIP_CHECK_ARG(g && g->exec, "NULL graph");
  const hipError_t e = hipGraphLaunch(g->exec, as_stream(stream));
  if (e != hipSuccess) {
    set_error("fqehip_taylor_graph_launch: %s", hipGetErrorString(e));
    return FQEHIP_ERR_HIP;
  }
  return FQEHIP_OK;
}

extern "C" void fqehip_taylor_graph_destroy(fqehip_taylor_graph* g) {
  release(g);
}
