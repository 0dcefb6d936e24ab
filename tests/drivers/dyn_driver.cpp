// Driver for the dynamic-graph path through starplat_graph.hpp, in the role
// of a main for StarPlat's Dynamic programs: parse a graph.hpp-format text
// file and an update.hpp-format update file, then
//   dyn_driver updates <graph.txt> <updates.txt> <batch> <out.bin>
//     applies the updates batch by batch with graph::updateCSRDel /
//     graph::updateCSRAdd and writes the final CSR: V+1 offsets, E
//     neighbours, E weights (int32, E = num_edges());
//   dyn_driver sssp <graph.txt> <updates.txt> <batch> <out.bin> <src>
//     runs DynSSSP (spm_dyn_sssp) and writes dist then parent (V int32 each).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "starplat_graph.hpp"

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s updates|sssp <graph.txt> <updates.txt> <batch> <out.bin> [src]\n", argv[0]);
    return 2;
  }
  graph g(argv[2]);
  g.parseGraph();
  std::vector<update> updates = parseUpdateFile(argv[3]);
  const int batch = std::atoi(argv[4]);
  FILE* out = std::fopen(argv[5], "wb");
  if (!out || batch < 1) return 3;
  const int V = g.num_nodes();
  if (std::strcmp(argv[1], "updates") == 0) {
    for (int idx = 0; idx < int(updates.size()); idx += batch) {
      const int n = std::min(batch, int(updates.size()) - idx);
      g.updateCSRDel(updates, idx, n);
      g.updateCSRAdd(updates, idx, n);
    }
    const int E = g.num_edges();
    std::fwrite(g.indexofNodes, sizeof(int32_t), V + 1, out);
    std::fwrite(g.edgeList, sizeof(int32_t), E, out);
    std::fwrite(g.getEdgeLen(), sizeof(int32_t), E, out);
  } else {
    if (argc < 7) return 2;
    std::vector<char> types;
    std::vector<int32_t> s, d;
    for (const update& u : updates) {
      types.push_back(u.type);
      s.push_back(u.source);
      d.push_back(u.destination);
    }
    std::vector<int32_t> dist(V), parent(V);
    spm_require(spm_dyn_sssp(g.device(), std::atoi(argv[6]), types.data(), s.data(), d.data(), int32_t(s.size()),
                             batch, dist.data(), parent.data()),
                "spm_dyn_sssp");
    std::fwrite(dist.data(), sizeof(int32_t), V, out);
    std::fwrite(parent.data(), sizeof(int32_t), V, out);
  }
  std::fclose(out);
  return 0;
}
