// Driver for code generated by the mi355x backend, in the role of StarPlat's
// graphcode/main_functions/*main.cpp: parse a graph.hpp-format text file, call
// one generated Compute_* entry point, write the result as raw binary so the
// test can compare it with the oracle. Built with -include <generated .h> and
// one of -DRUN_SSSP / -DRUN_PR / -DRUN_TC / -DRUN_BC.
//
//   gen_sssp <graph.txt> <out.bin> <src>
//   gen_pr   <graph.txt> <out.bin> <beta> <delta> <maxIter>
//   gen_tc   <graph.txt> <out.bin>
//   gen_bc   <graph.txt> <out.bin> <src> [<src> ...]
//   gen_dynsssp <graph.txt> <out.bin> <updates.txt> <batch> <src>
//     (-DRUN_DYNSSSP -DDYN_FUNC=<the Dynamic function>): dist, parent, then
//     the updated graph's edge count)
//   gen_dyntc <graph.txt> <out.bin> <updates.txt> <batch>
//     (-DRUN_DYNTC -DDYN_FUNC=<the Dynamic function>): the triangle count
//     (int64), then the updated graph's edge count)
//   gen_dynpr <graph.txt> <out.bin> <updates.txt> <batch> <beta> <delta> <maxIter>
//     (-DRUN_DYNPR -DDYN_FUNC=<the Dynamic function>): ranks, then the
//     updated graph's edge count)
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <graph.txt> <out.bin> [args]\n", argv[0]);
    return 2;
  }
  graph g(argv[1]);
  g.parseGraph();
  const int V = g.num_nodes();
  (void)V;
  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 3;
#if defined(RUN_SSSP)
  if (argc < 4) return 2;
  std::vector<int> dist(V, INT_MAX);
  std::vector<int> weight(g.num_edges());
  int k = 0;  // weight by CSR position, from getEdge, as the reference driver builds it
  for (int v = 0; v < V; v++)
    for (int e = g.indexofNodes[v]; e < g.indexofNodes[v + 1]; e++) weight[k++] = g.getEdge(v, g.edgeList[e]).weight;
  Compute_SSSP(g, dist.data(), weight.data(), std::atoi(argv[3]));
  std::fwrite(dist.data(), sizeof(int), V, out);
#elif defined(RUN_PR)
  if (argc < 6) return 2;
  std::vector<float> pr(V);
  Compute_PR(g, (float)std::atof(argv[3]), (float)std::atof(argv[4]), std::atoi(argv[5]), pr.data());
  std::fwrite(pr.data(), sizeof(float), V, out);
#elif defined(RUN_TC)
  long n = Compute_TC(g);
  std::fwrite(&n, sizeof n, 1, out);
#elif defined(RUN_BC)
  std::set<int> sourceSet;
  for (int i = 3; i < argc; i++) sourceSet.insert(std::atoi(argv[i]));
  std::vector<float> BC(V);
  Compute_BC(g, BC.data(), sourceSet);
  std::fwrite(BC.data(), sizeof(float), V, out);
#elif defined(RUN_DYNSSSP)
  if (argc < 6) return 2;
  std::vector<int> dist(V), parent(V);
  std::vector<int> weight(g.num_edges());
  int k = 0;
  for (int v = 0; v < V; v++)
    for (int e = g.indexofNodes[v]; e < g.indexofNodes[v + 1]; e++) weight[k++] = g.getEdge(v, g.edgeList[e]).weight;
  std::vector<update> updateBatch = parseUpdateFile(argv[3]);
  DYN_FUNC(g, dist.data(), parent.data(), weight.data(), updateBatch, std::atoi(argv[4]), std::atoi(argv[5]));
  std::fwrite(dist.data(), sizeof(int), V, out);
  std::fwrite(parent.data(), sizeof(int), V, out);
  const int E = g.num_edges();
  std::fwrite(&E, sizeof E, 1, out);
#elif defined(RUN_DYNTC)
  if (argc < 5) return 2;
  std::vector<update> updateBatch = parseUpdateFile(argv[3]);
  DYN_FUNC(g, updateBatch, std::atoi(argv[4]));
  // the Dynamic function returns nothing; its count stays with the graph's device mirror
  int64_t count = 0;
  spm_require(spm_dyntc_count(g.device(), &count), "spm_dyntc_count");
  std::fwrite(&count, sizeof count, 1, out);
  const int E = g.num_edges();
  std::fwrite(&E, sizeof E, 1, out);
#elif defined(RUN_DYNPR)
  if (argc < 8) return 2;
  std::vector<float> pr(V);
  std::vector<update> updateBatch = parseUpdateFile(argv[3]);
  DYN_FUNC(g, (float)std::atof(argv[5]), (float)std::atof(argv[6]), std::atoi(argv[7]), pr.data(), updateBatch,
           std::atoi(argv[4]));
  std::fwrite(pr.data(), sizeof(float), V, out);
  const int E = g.num_edges();
  std::fwrite(&E, sizeof E, 1, out);
#else
#error "define one of RUN_SSSP, RUN_PR, RUN_TC, RUN_BC, RUN_DYNSSSP, RUN_DYNTC, RUN_DYNPR"
#endif
  std::fclose(out);
  return 0;
}
