// Bellman-Ford as a fixed point over the set of improved vertices.
function Compute_SSSP(Graph g, propNode<int> dist, propEdge<int> weight, node src) {
  propNode<bool> improved;
  g.attachNodeProperty(dist = INF, improved = False);
  src.improved = True;
  src.dist = 0;
  bool done = False;
  fixedPoint until (done : !improved) {
    forall (v in g.nodes().filter(improved == True)) {
      forall (t in g.neighbors(v)) {
        edge e = g.get_edge(v, t);
        <t.dist, t.improved> = <Min(t.dist, v.dist + e.weight), True>;
      }
    }
  }
}
