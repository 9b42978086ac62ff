// tri_query_test.cpp — the reference's TriQueryTest (test/TriQueryTest.cu:39-153) restated over
// the C++ façade: plain g++ host code, no HIP compiler, linked against libgary.so and the CPU
// oracle. Usage: tri_query_test [rmat_scale [selector_model.json]]   (scale 0 or absent: ER
// G(4096, 65536); the model defaults to gary_amd/selector_model.json beside this binary's tree)
// Prints one line per case and exits non-zero on the first mismatch.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../oracle/oracle.h"
#include "DeviceGraph.h"
#include "kernels/GraphQueryHom.h"
#include "ops/Intersection.h"

namespace {

int failures = 0;

void expect(const char *name, uint64_t got, uint64_t want, double ms) {
  const bool ok = got == want;
  std::printf("%s %-28s count=%llu ref=%llu %.3f ms\n", ok ? "PASS" : "FAIL", name,
              (unsigned long long)got, (unsigned long long)want, ms);
  std::fflush(stdout);
  if (!ok) ++failures;
}

struct Timer {
  gary_timer *t = nullptr;
  Timer() { gary::check(gary_timer_create(&t)); }
  ~Timer() { gary_timer_destroy(t); }
  void start() { gary::check(gary_timer_start(t, nullptr)); }
  float stop() {
    float ms = 0;
    gary::check(gary_timer_stop(t, nullptr, &ms));
    return ms;
  }
};

}  // namespace

int main(int argc, char **argv) {
  const int scale = argc > 1 ? std::atoi(argv[1]) : 0;
  int ndev = 0;
  gary::check(gary_device_count(&ndev));
  if (ndev == 0) {
    std::printf("SKIP no GPU\n");
    return 0;
  }
  vidT n;
  std::vector<vidT> src, dst;
  if (scale == 0) {
    n = 4096;
    src.resize(65536);
    dst.resize(65536);
    gary::check(gary_gen_er(n, 65536, 0x6A7E0001ull, src.data(), dst.data()));
  } else {
    n = 1 << scale;
    src.resize((size_t)16 << scale);
    dst.resize((size_t)16 << scale);
    gary::check(gary_gen_rmat(scale, (int64_t)src.size(), 0.57, 0.19, 0.19, 0x6A7E0002ull,
                              src.data(), dst.data()));
  }
  auto csr = std::make_shared<GraphCSR>(GraphCSR::fromEdges(n, src, dst));
  std::printf("m=%d nnz=%lld\n", csr->numV(), (long long)csr->numE());

  DeviceGraphCtx ctx(csr);
  ctx.populateDevice();

  // RunCPU: cpuTriCount (restated in the oracle, 64-bit)
  auto c0 = std::chrono::steady_clock::now();
  const uint64_t countRef = or_tri_count(csr->numV(), csr->getPtr().data(), csr->getIdx().data(), 0);
  const double cpu_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
  expect("RunCPU", countRef, countRef, cpu_ms);

  Timer timer;
  DeviceGraph g{ctx.getDeviceCSR(), ctx.getDeviceCOO()};
  {  // RunGPULegacy
    DeviceArray<unsigned> c(1);
    c.zero();
    timer.start();
    graph_query::triangleCountLegacy(g, c.data());
    expect("RunGPULegacy", c.download()[0], countRef, timer.stop());
  }
  DeviceArray<uint64_t> counts(1);
  auto run = [&](const char *name, auto launch) {
    counts.zero();
    timer.start();
    launch();
    const float ms = timer.stop();
    expect(name, counts.download()[0], countRef, ms);
  };
  run("RunGPU", [&] { graph_query::triangleCountHom<false>(g, counts.data()); });
  run("RunGPUWithCache", [&] {
    graph_query::triangleCountHom<true, 32>(g, counts.data(), 1536, 256);
  });
  run("RunGPUWithCacheChunked", [&] {
    graph_query::triangleCountHom<true, 32, chunkedPolicy>(g, counts.data(), 1536, 256);
  });
  run("RunGPUWithCache64", [&] { graph_query::triangleCountHom<true, 64>(g, counts.data()); });
  run("RunGPUWithCacheDynamic", [&] {
    graph_query::triangleCountHom<true, 32, dynamicPolicy>(g, counts.data(), 1536, 256);
  });
  {  // Strategy::Learned runs the shipped model, not the heuristic fallback
    std::string model = argc > 2 ? argv[2] : std::string(argv[0]);
    if (argc <= 2) model = model.substr(0, model.find_last_of('/') + 1) +
                           "../../gary_amd/selector_model.json";
    graph_query::loadSelector(model);
    expect("loadSelector (trained)", graph_query::selectorTrained() ? 1 : 0, 1, 0.0);
  }
  const std::pair<const char *, graph_query::Strategy> strategies[] = {
      {"Strategy::Learned", graph_query::Strategy::Learned},
      {"Strategy::Heuristic", graph_query::Strategy::Heuristic},
      {"Strategy::Bitmap", graph_query::Strategy::Bitmap},
      {"Strategy::Merge", graph_query::Strategy::Merge},
  };
  for (auto &s : strategies)
    run(s.first, [&] { graph_query::triangleCount(g, s.second, counts.data()); });

  // SetOpTest-style check on two adjacency lists of the graph
  vidT a = 0;
  for (vidT v = 1; v < csr->numV(); ++v)
    if (csr->getDegree(v) > csr->getDegree(a)) a = v;
  const vidT b = csr->getDegree(a) > 0 ? csr->N(a)[0] : 0;
  const uint64_t want = or_intersect_count(csr->N(a), csr->getDegree(a), csr->N(b),
                                           csr->getDegree(b));
  for (auto m : {ops::IntersectMethod::BinarySearch, ops::IntersectMethod::CachedBinarySearch,
                 ops::IntersectMethod::Merge, ops::IntersectMethod::Bitmap}) {
    counts.zero();
    timer.start();
    ops::intersect_count(g.csr.colIdx + csr->getPtr()[a], csr->getDegree(a),
                         g.csr.colIdx + csr->getPtr()[b], csr->getDegree(b), counts.data(), m);
    expect("SetOp intersect_count", counts.download()[0], want, timer.stop());
  }
  {  // the same CSR built on the device from the uploaded tuples
    DeviceArray<vidT> ds(src.size()), dd(dst.size());
    ds.upload(src.data());
    dd.upload(dst.data());
    DeviceGraphCtx built;
    timer.start();
    built.buildFromEdges(n, ds.data(), dd.data(), (int64_t)src.size());
    const float ms = timer.stop();
    auto h = built.getHostCSR();
    const bool same = h->numE() == csr->numE() && h->getPtr() == csr->getPtr() &&
                      h->getIdx() == csr->getIdx();
    expect("DeviceGraphCtx::buildFromEdges", same ? 1 : 0, 1, ms);
  }
  {  // saveToBinary / loadFromBinary<GraphCSR> round trip of the reference's file format
    const std::string path = "/tmp/gary_tri_query_test_csr.bin";
    saveToBinary(*csr, path);
    GraphCSR back = loadCSRFromBinary(path);
    std::remove(path.c_str());
    const bool same = back.numV() == csr->numV() && back.getPtr() == csr->getPtr() &&
                      back.getIdx() == csr->getIdx() && back.getDegreeAll() == csr->getDegreeAll();
    expect("saveToBinary/loadFromBinary", same ? 1 : 0, 1, 0.0);
  }
  // the façade's exceptions: an out-of-range id is rejected before any device work
  try {
    std::vector<vidT> bad_s = {0}, bad_d = {n + 5};
    GraphCSR::fromEdges(n, bad_s, bad_d);
    std::printf("FAIL invalid edge accepted\n");
    ++failures;
  } catch (const gary::Error &e) {
    std::printf("PASS invalid edge rejected (%d): %s\n", e.code, e.what());
  }
  std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
  return failures ? 1 : 0;
}
