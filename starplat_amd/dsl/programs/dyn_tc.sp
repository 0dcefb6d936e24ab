// Triangle count kept up to date under batches of edge deletions and
// additions. Each batch tags its own edges, then counts, for every update
// (a, b) and every common neighbour c, how many of the triangle's edges the
// batch holds: a triangle whose k edges are all in the batch is met from 2k
// directed updates, so class k is divided by 2k.
Static countAll(Graph g) {
  long found = 0;
  forall (x in g.nodes()) {
    forall (lo in g.neighbors(x).filter(lo < x)) {
      forall (hi in g.neighbors(x).filter(hi > x)) {
        if (g.is_an_edge(lo, hi)) {
          found += 1;
        }
      }
    }
  }
  return found;
}

Incremental(Graph g, int before, propEdge<bool> tagged, updates<g> added) {
  long total = before;
  int once = 0;
  int twice = 0;
  int thrice = 0;
  forall (up in added) {
    int a = up.source;
    int b = up.destination;
    forall (c in g.neighbors(a).filter(a != b && c != a && c != b)) {
      edge ac = g.get_edge(a, c);
      int mine = 1;
      bool closed = False;
      if (ac.tagged)
        mine = mine + 1;
      if (g.is_an_edge(b, c)) {
        edge bc = g.get_edge(b, c);
        closed = True;
        if (bc.tagged)
          mine = mine + 1;
      }
      if (closed) {
        if (mine == 1) {
          once += 1;
        } else if (mine == 2) {
          twice += 1;
        } else {
          thrice += 1;
        }
      }
    }
  }
  total = total + (once / 2 + twice / 4 + thrice / 6);
  return total;
}

Decremental(Graph g, int before, propEdge<bool> tagged, updates<g> removed) {
  long total = before;
  int once = 0;
  int twice = 0;
  int thrice = 0;
  forall (up in removed) {
    int a = up.source;
    int b = up.destination;
    forall (c in g.neighbors(a).filter(a != b && c != a && c != b)) {
      edge ac = g.get_edge(a, c);
      int mine = 1;
      bool closed = False;
      if (ac.tagged)
        mine = mine + 1;
      if (g.is_an_edge(b, c)) {
        edge bc = g.get_edge(b, c);
        closed = True;
        if (bc.tagged)
          mine = mine + 1;
      }
      if (closed) {
        if (mine == 1) {
          once += 1;
        } else if (mine == 2) {
          twice += 1;
        } else {
          thrice += 1;
        }
      }
    }
  }
  total = total - (once / 2 + twice / 4 + thrice / 6);
  return total;
}

Dynamic keepCount(Graph g, updates<g> stream, int chunk) {
  int tri = countAll(g);
  Batch (stream : chunk) {
    propEdge<bool> gone;
    g.attachEdgeProperty(gone = false);
    updates<g> outgoing = stream.currentBatch(0);
    updates<g> incoming = stream.currentBatch(1);
    OnDelete (e in stream.currentBatch()): {
      int p = e.source;
      int q = e.destination;
      for (r in g.neighbors(p)) {
        edge pr = g.get_edge(p, r);
        if (r == q)
          pr.gone = True;
      }
    }
    tri = Decremental(g, tri, gone, outgoing);
    g.updateCSRDel(stream);
    g.updateCSRAdd(stream);
    propEdge<bool> fresh;
    g.attachEdgeProperty(fresh = false);
    OnAdd (e in stream.currentBatch()): {
      int p = e.source;
      int q = e.destination;
      for (r in g.neighbors(p)) {
        edge pr = g.get_edge(p, r);
        if (r == q)
          pr.fresh = True;
      }
    }
    tri = Incremental(g, tri, fresh, incoming);
  }
}
