"""Edge-update batches, as graphcode/update.hpp defines them.

``update`` mirrors the reference's class (type 'a' or 'd', source,
destination, weight); ``parseUpdateFile`` follows update.hpp:18-72 -- one
``<type> <src> <dst>`` per line, each line pushed twice, once reversed (the
reference does so unconditionally, "only required for undirected graph
case"), additions weighing 1; ``getDeletions`` / ``getAdditions`` follow
:74-102. ``as_arrays`` turns a slice into the plain arrays the C ABI takes.
"""
from __future__ import annotations

import numpy as np


class update:
    __slots__ = ("type", "source", "destination", "weight")

    def __init__(self, type, source, destination, weight=1):
        self.type = type
        self.source = int(source)
        self.destination = int(destination)
        self.weight = int(weight)

    def __repr__(self):
        return f"update({self.type!r}, {self.source}, {self.destination}, {self.weight})"


def parseUpdateFile(updateFile):
    """update.hpp:18-72. Lines that do not hold a type and two ids are skipped
    (the reference reads garbage for them)."""
    updates = []
    with open(updateFile) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 3 or parts[0] not in ("a", "d"):
                continue
            t, s, d = parts[0], int(parts[1]), int(parts[2])
            updates.append(update(t, s, d, 1))
            updates.append(update(t, d, s, 1))
    return updates


def getDeletions(updateIndex, batchSize, updateVec):
    """update.hpp:74-87."""
    return [u for u in updateVec[updateIndex:updateIndex + batchSize] if u.type == "d"]


def getAdditions(updateIndex, batchSize, updateVec):
    """update.hpp:89-102."""
    return [u for u in updateVec[updateIndex:updateIndex + batchSize] if u.type == "a"]


def as_arrays(updateVec):
    """(types as bytes, sources, destinations, weights) of a list of updates."""
    n = len(updateVec)
    types = "".join(u.type for u in updateVec).encode()
    src = np.fromiter((u.source for u in updateVec), np.int32, n)
    dst = np.fromiter((u.destination for u in updateVec), np.int32, n)
    w = np.fromiter((u.weight for u in updateVec), np.int32, n)
    return types, src, dst, w


def write_update_file(path, types, src, dst):
    """Writes ``<type> <src> <dst>`` lines (one per pair, not doubled)."""
    with open(path, "w") as f:
        for t, s, d in zip(types, src, dst):
            f.write(f"{chr(t) if isinstance(t, int) else t} {int(s)} {int(d)}\n")
