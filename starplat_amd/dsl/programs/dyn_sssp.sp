// Single-source shortest paths kept up to date under batches of edge
// deletions and additions: a Bellman-Ford fixed point that records each
// vertex's predecessor, then per batch: cut off what a deleted tree edge
// supported, re-derive it from in-neighbours, and push from the endpoints of
// useful added edges.
Static baseDistances(Graph g, propNode<int> d, propNode<int> pred, propEdge<int> w, int root) {
  propNode<bool> changed;
  propNode<bool> changed_next;
  g.attachNodeProperty(d = INF, changed = False, changed_next = False, pred = -1);
  root.changed = True;
  root.d = 0;
  bool stable = False;
  fixedPoint until (stable : !changed) {
    forall (x in g.nodes().filter(changed == True)) {
      forall (y in g.neighbors(x)) {
        edge xy = g.get_edge(x, y);
        <y.d, y.changed_next, y.pred> = <Min(y.d, x.d + xy.w), True, x>;
      }
    }
    changed = changed_next;
    g.attachNodeProperty(changed_next = False);
  }
}

Incremental(Graph g, propNode<int> d, propNode<int> pred, propEdge<int> w, propNode<bool> touched) {
  propNode<bool> touched_next;
  g.attachNodeProperty(touched_next = False);
  bool stable = False;
  fixedPoint until (stable : !touched) {
    forall (x in g.nodes().filter(touched == True)) {
      forall (y in g.neighbors(x)) {
        edge xy = g.get_edge(x, y);
        <y.d, y.touched_next, y.pred> = <Min(y.d, x.d + xy.w), True, x>;
      }
    }
    touched = touched_next;
    g.attachNodeProperty(touched_next = False);
  }
}

Decremental(Graph g, propNode<int> d, propNode<int> pred, propEdge<int> w, propNode<bool> cut) {
  bool stable = False;
  while (!stable) {
    stable = True;
    forall (x in g.nodes().filter(cut == False)) {
      node up = x.pred;
      if (up > -1 && up.cut) {
        x.d = INT_MAX/2;
        x.cut = True;
        x.pred = -1;
        stable = False;
      }
    }
  }
  stable = False;
  while (!stable) {
    stable = True;
    forall (x in g.nodes().filter(cut == True)) {
      for (y in g.nodes_to(x)) {
        edge yx = g.get_edge(y, x);
        if (x.d > y.d + yx.w) {
          x.d = y.d + yx.w;
          x.pred = y;
          stable = False;
        }
      }
    }
  }
}

Dynamic maintainDistances(Graph g, propNode<int> d, propNode<int> pred, propEdge<int> w, updates<g> stream, int perBatch, int root) {
  baseDistances(g, d, pred, w, root);
  Batch(stream : perBatch) {
    propNode<bool> cut;
    propNode<bool> grown;
    g.attachNodeProperty(cut = False, grown = False);
    OnDelete(e in stream.currentBatch()): {
      int a = e.source;
      int b = e.destination;
      if (b.pred == a) {
        b.d = INT_MAX/2;
        b.cut = True;
        b.pred = -1;
      }
    }
    g.updateCSRDel(stream);
    Decremental(g, d, pred, w, cut);
    OnAdd(e in stream.currentBatch()): {
      int a = e.source;
      int b = e.destination;
      if (b.d > a.d + 1) {
        b.grown = True;
        a.grown = True;
      }
    }
    g.updateCSRAdd(stream);
    Incremental(g, d, pred, w, grown);
  }
}
