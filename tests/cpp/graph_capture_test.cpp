// Captures gary_tc_count (learned, heuristic and bitmap strategies) in a HIP graph on a caller's
// stream and replays it: every replay must add the eager count to the device counter, with
// eager counts of the program between replays. Plain g++ host code against libgary and the HIP
// runtime API, as a caller that captures its own work would write it.
//
//   graph_capture_test [scale]      (R-MAT, edgefactor 16; default 14)
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gary.h"

#define CHECK(x)                                                                 \
  do {                                                                           \
    const int rc_ = (x);                                                         \
    if (rc_ != 0) {                                                              \
      std::printf("FAILED %s: %d %s\n", #x, rc_, gary_last_error());             \
      return 1;                                                                  \
    }                                                                            \
  } while (0)
#define HCHECK(x)                                                                \
  do {                                                                           \
    const hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                      \
      std::printf("FAILED %s: %s\n", #x, hipGetErrorString(e_));                 \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

int main(int argc, char **argv) {
  const int scale = argc > 1 ? std::atoi(argv[1]) : 14;
  const int32_t n = 1 << scale;
  const int64_t nt = 16ll * n;
  std::vector<int32_t> src(nt), dst(nt);
  CHECK(gary_gen_rmat(scale, nt, 0.57, 0.19, 0.19, 41, src.data(), dst.data()));
  std::vector<int64_t> indPtr(n + 1);
  std::vector<int32_t> colIdx(nt);
  int64_t nnz = 0;
  CHECK(gary_csr_from_edges(n, nt, src.data(), dst.data(), GARY_CSR_ORIENT_DEGREE, indPtr.data(),
                            colIdx.data(), nt, &nnz));
  gary_ctx *ctx = nullptr;
  CHECK(gary_ctx_create(0, &ctx));
  CHECK(gary_ctx_load_csr(ctx, n, nnz, indPtr.data(), colIdx.data()));
  CHECK(gary_ctx_populate(ctx));
  gary_device_graph g;
  CHECK(gary_ctx_device_graph(ctx, &g));
  uint64_t want = 0;
  CHECK(gary_tc_count_sync(&g, GARY_TC_BITMAP, nullptr, &want));
  std::printf("rmat%d: %lld arcs, %llu triangles\n", scale, (long long)nnz,
              (unsigned long long)want);

  // a trained-looking model (any parameters give exact counts; the route table may be mixed)
  std::vector<float> params(GARY_MLP_PARAMS);
  uint64_t z = 0x5EEDull;
  for (float &p : params) {
    z = z * 6364136223846793005ull + 1442695040888963407ull;
    p = (float)((int64_t)(z >> 40) - (1 << 23)) / (float)(1 << 24);
  }
  CHECK(gary_selector_set_params(params.data(), 1, 0.f, 7, 0.f));

  hipStream_t st;
  HCHECK(hipStreamCreate(&st));
  uint64_t *d_count = nullptr;
  CHECK(gary_malloc(0, sizeof(uint64_t), (void **)&d_count));
  const int strategies[] = {GARY_TC_AUTO_LEARNED, GARY_TC_HEURISTIC, GARY_TC_BITMAP};
  for (int strat : strategies) {
    CHECK(gary_memset(0, d_count, 0, sizeof(uint64_t)));
    CHECK(gary_tc_count(&g, strat, nullptr, d_count, st));  // sizes the workspace
    HCHECK(hipStreamSynchronize(st));
    hipGraph_t graph;
    hipGraphExec_t exec;
    HCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
    const int rc = gary_tc_count(&g, strat, nullptr, d_count, st);
    HCHECK(hipStreamEndCapture(st, &graph));
    CHECK(rc);
    HCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    for (int rep = 1; rep <= 4; ++rep) {
      HCHECK(hipGraphLaunch(exec, st));
      HCHECK(hipStreamSynchronize(st));
      uint64_t got = 0;
      CHECK(gary_memcpy_d2h(0, &got, d_count, sizeof got));
      if (got != want * (uint64_t)(rep + 1)) {
        std::printf("FAILED strategy %d replay %d: %llu != %llu\n", strat, rep,
                    (unsigned long long)got, (unsigned long long)(want * (rep + 1)));
        return 1;
      }
      uint64_t eager = 0;  // eager work of the program between replays
      CHECK(gary_tc_count_sync(&g, strat, nullptr, &eager));
      if (eager != want) {
        std::printf("FAILED strategy %d eager count %llu != %llu\n", strat,
                    (unsigned long long)eager, (unsigned long long)want);
        return 1;
      }
    }
    HCHECK(hipGraphExecDestroy(exec));
    HCHECK(hipGraphDestroy(graph));
    std::printf("strategy %d: 4 replays ok\n", strat);
  }
  CHECK(gary_free(0, d_count));
  HCHECK(hipStreamDestroy(st));
  CHECK(gary_ctx_destroy(ctx));
  std::printf("ALL PASSED\n");
  return 0;
}
