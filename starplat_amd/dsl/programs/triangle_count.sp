// Each triangle a < b < c is counted once, at its middle vertex b.
function Compute_TC(Graph g) {
  long count = 0;
  forall (v in g.nodes()) {
    forall (a in g.neighbors(v).filter(a < v)) {
      forall (c in g.neighbors(v).filter(c > v)) {
        if (g.is_an_edge(a, c)) {
          count += 1;
        }
      }
    }
  }
  return count;
}
