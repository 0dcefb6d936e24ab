// PageRank refreshed after batches of edge deletions and additions: a pull
// sweep from uniform ranks, then per batch the vertices touched by the batch
// (and, through propagateNodeFlags, every vertex with an in-edge) are swept
// again after the deletions and again after the additions.
Static firstRanks(Graph g, float tol, float damp, int rounds, propNode<float> rank) {
  float total = g.num_nodes();
  propNode<float> rank_next;
  g.attachNodeProperty(rank = 1 / total, rank_next = 0);
  int done = 0;
  float change;
  do {
    change = 0.0;
    forall (x in g.nodes()) {
      float acc = 0.0;
      for (y in g.nodes_to(x)) {
        acc = acc + y.rank / g.count_outNbrs(y);
      }
      float fresh = (1 - damp) / total + damp * acc;
      change += fresh - x.rank;
      x.rank_next = fresh;
    }
    rank = rank_next;
    done++;
  } while ((change > tol) && (done < rounds));
}

Incremental(Graph g, float tol, float damp, int rounds, propNode<float> rank, propNode<bool> hit) {
  propNode<float> rank_next;
  g.attachNodeProperty(rank_next = 0);
  int done = 0;
  float change;
  do {
    change = 0.0;
    forall (x in g.nodes().filter(hit == True)) {
      float acc = 0.0;
      for (y in g.nodes_to(x)) {
        acc = acc + y.rank / g.count_outNbrs(y);
      }
      float fresh = (1 - damp) / g.num_nodes() + damp * acc;
      change += fresh - x.rank;
      x.rank_next = fresh;
    }
    forall (x in g.nodes().filter(hit == True)) {
      x.rank = x.rank_next;
    }
    done++;
  } while ((change > tol) && (done < rounds));
}

Decremental(Graph g, float tol, float damp, int rounds, propNode<float> rank, propNode<bool> hit) {
  propNode<float> rank_next;
  g.attachNodeProperty(rank_next = 0);
  int done = 0;
  float change;
  do {
    change = 0.0;
    forall (x in g.nodes().filter(hit == True)) {
      float acc = 0.0;
      for (y in g.nodes_to(x)) {
        acc = acc + y.rank / g.count_outNbrs(y);
      }
      float fresh = (1 - damp) / g.num_nodes() + damp * acc;
      change += fresh - x.rank;
      x.rank_next = fresh;
    }
    forall (x in g.nodes().filter(hit == True)) {
      x.rank = x.rank_next;
    }
    done++;
  } while ((change > tol) && (done < rounds));
}

Dynamic refreshRanks(Graph g, float tol, float damp, int rounds, propNode<float> rank, updates<g> stream, int chunk) {
  firstRanks(g, tol, damp, rounds, rank);
  Batch (stream : chunk) {
    propNode<bool> lost;
    propNode<bool> gained;
    g.attachNodeProperty(lost = False, gained = False);
    updates<g> outgoing = stream.currentBatch(0);
    updates<g> incoming = stream.currentBatch(1);
    OnDelete (e in stream.currentBatch()): {
      node p = e.source;
      node q = e.destination;
      q.lost = True;
    }
    g.propagateNodeFlags(lost);
    g.updateCSRDel(stream);
    Decremental(g, tol, damp, rounds, rank, lost);
    OnAdd (e in stream.currentBatch()): {
      node p = e.source;
      node q = e.destination;
      q.gained = True;
    }
    g.propagateNodeFlags(gained);
    g.updateCSRAdd(stream);
    Incremental(g, tol, damp, rounds, rank, gained);
  }
}
