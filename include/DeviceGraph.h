// DeviceGraph.h — C++ façade over the C ABI (gary.h) with the reference's names:
// DeviceCSR / DeviceCOO / DeviceGraph views and the owning DeviceGraphCtx
// (include/DeviceGraph.cuh:15-133, src/DeviceGraph.cu:14-96), plus a minimal host GraphCSR
// (include/Graph.h:157-292). Header-only; links against libgary.so, needs no HIP compiler.
// Errors surface as gary::Error exceptions on this side of the ABI (the reference throws
// std::runtime_error); nothing throws across the ABI itself.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Defines.h"
#include "gary.h"

namespace gary {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

inline void check(int rc) {
  if (rc != GARY_OK) throw Error(rc, gary_last_error());
}

}  // namespace gary

using DeviceCSR = gary_device_csr;
using DeviceCOO = gary_device_coo;
using DeviceGraph = gary_device_graph;

static_assert(sizeof(DeviceCSR) == 40 && offsetof(DeviceCSR, nnz) == 8 &&
                  offsetof(DeviceCSR, indPtr) == 16 && offsetof(DeviceCSR, colIdx) == 24 &&
                  offsetof(DeviceCSR, deg) == 32,
              "DeviceCSR layout must match {vidT m; eidT nnz; eidT* indPtr; vidT* colIdx; vidT* deg}");
static_assert(sizeof(DeviceGraph) == 80 && offsetof(DeviceGraph, coo) == 40,
              "DeviceGraph layout must match {DeviceCSR csr; DeviceCOO coo}");

// Host CSR: row pointers (m+1), sorted adjacency lists, degrees derived from the row pointers.
class GraphCSR {
 public:
  GraphCSR() = default;
  GraphCSR(vidT m, eidT nnz, std::vector<eidT> ptr, std::vector<vidT> idx)
      : m_(m), nnz_(nnz), ptr_(std::move(ptr)), idx_(std::move(idx)), deg_(m) {
    for (vidT v = 0; v < m_; ++v) deg_[v] = (vidT)(ptr_[v + 1] - ptr_[v]);
  }
  vidT numV() const { return m_; }
  eidT numE() const { return nnz_; }
  vidT getDegree(vidT v) const { return deg_.at(v); }
  const std::vector<vidT> &getDegreeAll() const { return deg_; }
  const std::vector<eidT> &getPtr() const { return ptr_; }
  const std::vector<vidT> &getIdx() const { return idx_; }
  const vidT *N(vidT v) const { return idx_.data() + ptr_[v]; }

  // COO -> CSR from raw tuples: symmetrize, drop loops, dedup, then orient / keep both arcs.
  static GraphCSR fromEdges(vidT n, const std::vector<vidT> &src, const std::vector<vidT> &dst,
                            int mode = GARY_CSR_ORIENT_DEGREE) {
    if (src.size() != dst.size()) throw gary::Error(GARY_ERR_INVALID, "src/dst size mismatch");
    const int64_t nt = (int64_t)src.size();
    std::vector<eidT> ptr((size_t)n + 1);
    const bool sym = mode == GARY_CSR_SYMMETRIC || mode == GARY_CSR_SYMMETRIC_RANKED;
    std::vector<vidT> idx((size_t)(sym ? 2 * nt : nt) + 1);
    int64_t nnz = 0;
    gary::check(gary_csr_from_edges(n, nt, src.data(), dst.data(), mode, ptr.data(), idx.data(),
                                    (int64_t)idx.size() - 1, &nnz));
    idx.resize((size_t)nnz);
    return GraphCSR(n, nnz, std::move(ptr), std::move(idx));
  }

 private:
  vidT m_ = 0;
  eidT nnz_ = 0;
  std::vector<eidT> ptr_;
  std::vector<vidT> idx_;
  std::vector<vidT> deg_;
};

// loadFromBinary<GraphCSR> (include/Graph.h:301-309): a CSR file written by the reference's
// saveToBinary (cereal binary archive, SerDesVersion V1) or by saveToBinary below.
inline GraphCSR loadCSRFromBinary(const std::string &path) {
  int format = 0;
  vidT m = 0;
  eidT nnz = 0;
  gary::check(gary_graph_file_info(path.c_str(), &format, &m, &nnz));
  std::vector<eidT> ptr((size_t)m + 1);
  std::vector<vidT> idx((size_t)nnz);
  gary::check(gary_graph_file_read_csr(path.c_str(), ptr.data(), idx.data(), nullptr));
  return GraphCSR(m, nnz, std::move(ptr), std::move(idx));
}

// saveToBinary (include/Graph.h:311-316) of a CSR, readable by the reference's loadFromBinary.
inline void saveToBinary(const GraphCSR &g, const std::string &path) {
  gary::check(gary_graph_file_write_csr(path.c_str(), g.numV(), g.numE(), g.getPtr().data(),
                                        g.getIdx().data(), g.getDegreeAll().data()));
}

// loadFromSNAP (src/Graph.cpp:157-196): tuples of a SNAP edge list, n = max id + 1.
struct EdgeList {
  vidT n = 0;
  std::vector<vidT> src, dst;
};
inline EdgeList loadFromSNAP(const std::string &path, bool removeSelfLoops = true) {
  EdgeList e;
  int64_t nt = 0;
  gary::check(gary_snap_read(path.c_str(), removeSelfLoops, nullptr, nullptr, 0, &nt, &e.n));
  e.src.resize((size_t)nt);
  e.dst.resize((size_t)nt);
  gary::check(gary_snap_read(path.c_str(), removeSelfLoops, e.src.data(), e.dst.data(), nt, &nt,
                             &e.n));
  return e;
}

// Owner of one graph's device buffers.
class DeviceGraphCtx {
 public:
  explicit DeviceGraphCtx(int device = 0) { gary::check(gary_ctx_create(device, &h_)); }
  explicit DeviceGraphCtx(std::shared_ptr<GraphCSR> graph, int device = 0) : DeviceGraphCtx(device) {
    loadFromHost(std::move(graph));
  }
  DeviceGraphCtx(const DeviceGraphCtx &) = delete;
  DeviceGraphCtx &operator=(const DeviceGraphCtx &) = delete;
  ~DeviceGraphCtx() { gary_ctx_destroy(h_); }

  // H2D of indPtr and colIdx (replaces the current graph); deg is derived on the device.
  void loadFromHost(std::shared_ptr<GraphCSR> graph) {
    gary::check(gary_ctx_load_csr(h_, graph->numV(), graph->numE(), graph->getPtr().data(),
                                  graph->getIdx().data()));
    host_ = std::move(graph);
  }
  // Build the CSR on the device from device tuple arrays (symmetrize, drop loops, dedup, orient
  // by `mode`), replacing the host GraphCOO -> GraphCSR build followed by loadFromHost.
  void buildFromEdges(vidT n, const vidT *d_src, const vidT *d_dst, int64_t ntuples,
                      int mode = GARY_CSR_ORIENT_DEGREE, void *stream = nullptr) {
    gary::check(gary_ctx_build_from_edges(h_, n, ntuples, d_src, d_dst, mode, stream));
    host_.reset();
  }
  // Create the COO rowIdx on the device.
  void populateDevice() { gary::check(gary_ctx_populate(h_)); }
  void detachFromHost() { host_.reset(); }
  void clear() {
    gary::check(gary_ctx_clear(h_));
    host_.reset();
  }
  // The host graph; after detachFromHost it is copied back from the device (m+1 row pointers).
  std::shared_ptr<GraphCSR> getHostCSR() {
    if (!host_) {
      DeviceGraph g = view();
      std::vector<eidT> ptr((size_t)g.csr.m + 1);
      std::vector<vidT> idx((size_t)g.csr.nnz);
      gary::check(gary_ctx_copy_to_host(h_, ptr.data(), idx.data(), nullptr));
      host_ = std::make_shared<GraphCSR>(g.csr.m, g.csr.nnz, std::move(ptr), std::move(idx));
    }
    return host_;
  }
  DeviceCSR getDeviceCSR() { return view().csr; }
  DeviceCOO getDeviceCOO() {
    populateDevice();
    return view().coo;
  }
  DeviceGraph getDeviceGraph() { return {getDeviceCSR(), getDeviceCOO()}; }

 private:
  DeviceGraph view() {
    DeviceGraph g;
    gary::check(gary_ctx_device_graph(h_, &g));
    return g;
  }
  gary_ctx *h_ = nullptr;
  std::shared_ptr<GraphCSR> host_;
};

// A device allocation owned by the library (helper for callers without a HIP toolchain).
template <typename T>
class DeviceArray {
 public:
  explicit DeviceArray(size_t n, int device = 0) : n_(n), dev_(device) {
    void *p = nullptr;
    gary::check(gary_malloc(device, n * sizeof(T) + (n == 0), &p));
    p_ = (T *)p;
  }
  DeviceArray(const DeviceArray &) = delete;
  ~DeviceArray() { gary_free(dev_, p_); }
  T *data() { return p_; }
  size_t size() const { return n_; }
  void zero() { gary::check(gary_memset(dev_, p_, 0, n_ * sizeof(T))); }
  void upload(const T *src) { gary::check(gary_memcpy_h2d(dev_, p_, src, n_ * sizeof(T))); }
  std::vector<T> download() const {
    std::vector<T> out(n_);
    gary::check(gary_memcpy_d2h(dev_, out.data(), p_, n_ * sizeof(T)));
    return out;
  }

 private:
  size_t n_;
  int dev_;
  T *p_ = nullptr;
};
