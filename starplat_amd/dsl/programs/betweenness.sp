// Brandes betweenness over a set of sources (no 1/2 factor).
function Compute_BC(Graph g, propNode<float> BC, SetN<g> sourceSet) {
  g.attachNodeProperty(BC = 0);
  propNode<float> paths;
  propNode<float> dep;
  for (s in sourceSet) {
    g.attachNodeProperty(dep = 0);
    g.attachNodeProperty(paths = 0);
    s.paths = 1;
    iterateInBFS(v in g.nodes() from s) {
      for (w in g.neighbors(v)) {
        w.paths += v.paths;
      }
    }
    iterateInReverse(v != s) {
      for (w in g.neighbors(v)) {
        v.dep = v.dep + (v.paths / w.paths) * (1 + w.dep);
      }
      v.BC = v.BC + v.dep;
    }
  }
}
