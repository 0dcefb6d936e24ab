// Pull PageRank; stops when the L1 change of the ranks drops to beta or
// after maxIter sweeps.
function Compute_PR(Graph g, float beta, float delta, int maxIter, propNode<float> pageRank) {
  float n = g.num_nodes();
  propNode<float> nextRank;
  g.attachNodeProperty(pageRank = 1 / n, nextRank = 0);
  int it = 0;
  float change = 0.0;
  do {
    change = 0.0;
    forall (v in g.nodes()) {
      float acc = 0.0;
      for (u in g.nodes_to(v)) {
        acc = acc + u.pageRank / g.count_outNbrs(u);
      }
      float r = (1 - delta) / n + delta * acc;
      if (r - v.pageRank >= 0) {
        change += r - v.pageRank;
      } else {
        change += v.pageRank - r;
      }
      v.nextRank = r;
    }
    pageRank = nextRank;
    it++;
  } while ((change > beta) && (it < maxIter));
}
