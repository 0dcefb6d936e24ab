// starplat_graph.hpp -- the `graph` class that code generated by the mi355x
// backend compiles against. It keeps the public members and methods generated
// code and drivers use from StarPlat's graphcode/graph.hpp (class graph,
// :31-1215) -- indexofNodes, edgeList, rev_indexofNodes, srcList (:44-48),
// getEdgeLen() (:134-137), num_nodes() (:139-142), num_edges() (:145-152),
// parseGraph() (:861-865), check_if_nbr() (:233-287), getEdge() (:168-186),
// updateCSRDel() / updateCSRAdd() (:869-975) -- without graph.hpp's
// MKL/GNN/geometric includes. Parsing and CSR construction go through the
// host data layer (libspm_host.so, same semantics); device() returns the HBM
// mirror used by the Compute_* bodies. The update calls apply a batch to the
// HBM mirror and then refresh the public arrays from it (rows ascending).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "starplat_amd.h"
#include "starplat_host.h"

// graphcode/abstractGraph.hpp:16-25
class edge {
 public:
  int32_t source = 0;
  int32_t destination = 0;
  int32_t weight = 0;
  int32_t id = 0;
  int8_t dir = 1;
};

// graphcode/update.hpp:8-15
class update {
 public:
  char type = 'a';
  int32_t source = 0;
  int32_t destination = 0;
  int32_t weight = 1;
};

class graph {
 public:
  int32_t* indexofNodes = nullptr;
  int32_t* rev_indexofNodes = nullptr;
  int32_t* edgeList = nullptr;
  int32_t* srcList = nullptr;
  int32_t* edgeMap = nullptr;  // reverse-CSR slot -> CSR position (graph.hpp:57)

  explicit graph(const char* file) : filePath_(file ? file : "") {}
  graph(const graph&) = delete;
  graph& operator=(const graph&) = delete;
  ~graph() {
    if (dev_) spm_graph_destroy(dev_);
    if (h_) spmh_free(h_);
  }

  void parseGraph() {
    spmh_graph* h = spmh_parse_text(filePath_.c_str());
    if (!h) throw std::runtime_error("parseGraph: cannot read " + filePath_);
    adopt(h);
  }
  // Takes ownership of a CSR built by the host data layer (generators).
  void adopt(spmh_graph* h) {
    if (dev_) { spm_graph_destroy(dev_); dev_ = nullptr; }
    if (h_) spmh_free(h_);
    h_ = h;
    updated_ = false;
    indexofNodes = h->indexofNodes;
    rev_indexofNodes = h->rev_indexofNodes;
    edgeList = h->edgeList;
    srcList = h->srcList;
    build_edge_map(h->V, h->E);
  }

  int getOutDegree(int v) { return indexofNodes[v + 1] - indexofNodes[v]; }  // graph.hpp:316-320
  int getInDegree(int v) { return rev_indexofNodes[v + 1] - rev_indexofNodes[v]; }  // :322-326

  int num_nodes() { return h_ ? h_->V : 1; }
  int num_edges() { return updated_ ? int(adj_.size()) : h_ ? int(h_->E) : 0; }
  int* getEdgeLen() { return updated_ ? w_.data() : h_ ? h_->edgeLen : nullptr; }

  // Public arrays := the HBM mirror's current CSR (after updates applied on
  // the device, e.g. by spm_dyn_sssp).
  void sync_from_device() {
    if (dev_) refresh();
  }

  // graph::updateCSRDel(batchUpdate, k, size) (graph.hpp:869-930): the
  // deletions among batchUpdate[k, k+size). Every copy of an edge goes.
  void updateCSRDel(std::vector<update>& batchUpdate, int k, int size) { apply(batchUpdate, k, size, 'd'); }
  // graph::updateCSRAdd(batchUpdate, k, size) (graph.hpp:932-975): the
  // additions among batchUpdate[k, k+size), with their weights.
  void updateCSRAdd(std::vector<update>& batchUpdate, int k, int size) { apply(batchUpdate, k, size, 'a'); }

  bool check_if_nbr(int s, int d) {
    for (int32_t e = indexofNodes[s]; e < indexofNodes[s + 1]; ++e)
      if (edgeList[e] == d) return true;
    return false;
  }
  edge getEdge(int s, int d) {
    edge found;
    for (int32_t e = indexofNodes[s]; e < indexofNodes[s + 1]; ++e)
      if (edgeList[e] == d) {
        found.source = s;
        found.destination = d;
        found.weight = getEdgeLen()[e];
        found.id = e;
        return found;
      }
    return found;
  }

  // HBM mirror of the CSR, created on first use and reused by every
  // generated Compute_* call on this graph.
  spm_graph* device() {
    if (!dev_) {
      if (!h_) throw std::runtime_error("graph::device: parseGraph() first");
      int rc = spm_graph_create(h_->V, h_->E, h_->indexofNodes, h_->edgeList, h_->edgeLen,
                                h_->rev_indexofNodes, h_->srcList, &dev_);
      if (rc != SPM_OK) throw std::runtime_error(std::string("spm_graph_create: ") + spm_last_error());
    }
    return dev_;
  }

 private:
  void apply(std::vector<update>& u, int k, int size, char kind) {
    std::vector<int32_t> s, d, w;
    for (int i = k; i < k + size && i < int(u.size()); ++i)
      if (u[i].type == kind) {
        s.push_back(u[i].source);
        d.push_back(u[i].destination);
        w.push_back(u[i].weight);
      }
    int rc = kind == 'd' ? spm_graph_delete_edges(device(), s.data(), d.data(), int32_t(s.size()))
                         : spm_graph_add_edges(device(), s.data(), d.data(), w.data(), int32_t(s.size()));
    if (rc != SPM_OK) throw std::runtime_error(std::string("graph update: ") + spm_last_error());
    refresh();
  }
  // Public arrays := the HBM mirror's current (folded) CSR.
  void refresh() {
    int64_t E = 0;
    int rc = spm_graph_num_edges(dev_, &E);
    const size_t V = size_t(h_->V);
    off_.assign(V + 1, 0);
    roff_.assign(V + 1, 0);
    adj_.assign(size_t(E), 0);
    w_.assign(size_t(E), 0);
    radj_.assign(size_t(E), 0);
    if (rc == SPM_OK)
      rc = spm_graph_export(dev_, off_.data(), adj_.data(), h_->edgeLen ? w_.data() : nullptr, roff_.data(),
                            radj_.data());
    if (rc != SPM_OK) throw std::runtime_error(std::string("graph update: ") + spm_last_error());
    indexofNodes = off_.data();
    edgeList = adj_.data();
    rev_indexofNodes = roff_.data();
    srcList = radj_.data();
    updated_ = true;
    build_edge_map(h_->V, E);
  }
  void build_edge_map(int32_t V, int64_t E) {
    emap_.assign(size_t(E > 0 ? E : 1), 0);
    if (spmh_edge_map(V, indexofNodes, edgeList, rev_indexofNodes, srcList, emap_.data()) != 0)
      throw std::runtime_error("graph: the CSR and the reverse CSR disagree");
    edgeMap = emap_.data();
  }

  std::string filePath_;
  spmh_graph* h_ = nullptr;
  spm_graph* dev_ = nullptr;
  bool updated_ = false;
  std::vector<int32_t> off_, adj_, w_, roff_, radj_, emap_;
};

// update.hpp:18-72: one "<type> <src> <dst>" per line; every line is pushed
// twice, once reversed, as the reference does; additions weigh 1.
inline std::vector<update> parseUpdateFile(const char* updateFile) {
  std::vector<update> out;
  FILE* f = std::fopen(updateFile, "r");
  if (!f) return out;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    char t = 0;
    int s = 0, d = 0;
    if (std::sscanf(line, " %c %d %d", &t, &s, &d) != 3 || (t != 'a' && t != 'd')) continue;
    update u;
    u.type = t;
    u.source = s;
    u.destination = d;
    u.weight = 1;
    update r = u;
    r.source = d;
    r.destination = s;
    out.push_back(u);
    out.push_back(r);
  }
  std::fclose(f);
  return out;
}

// Status check used by generated code: a failing library call is fatal, as a
// failing CUDA call is in the reference's generated code (libcuda.cuh:215-222).
inline void spm_require(int rc, const char* what) {
  if (rc != SPM_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, spm_last_error());
    std::exit(1);
  }
}
