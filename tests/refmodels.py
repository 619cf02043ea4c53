"""Independent formulations used to pin the oracle (pure Python, tiny inputs).

These are not restatements of the oracle's row loop:
  * best_path_score    -- shortest path over the alignment state graph
                          (row, char vertex) with the three ConSGA ops
                          (src/traceback.hpp:34-36), by dynamic programming in
                          (row, topological vertex) order;
  * chain_score        -- for a graph that is one chain of chars, the textbook
                          form: sigma(G[a], r[0]) + Levenshtein(r[1:], G[a+1:b+1])
                          minimised over the window a <= b;
  * levenshtein        -- the textbook unit-cost edit distance.
"""
from __future__ import annotations


def sigma(alphabet: int, a: int, b: int, m: int = 0, s: int = 1) -> int:
    if alphabet == 5:
        return m if a == b else s
    N = ord("N")
    return m if (a == N or b == N or a == b) else s


def levenshtein(x: bytes, y: bytes) -> int:
    prev = list(range(len(y) + 1))
    for i in range(1, len(x) + 1):
        cur = [i] + [0] * len(y)
        for j in range(1, len(y) + 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (0 if x[i - 1] == y[j - 1] else 1))
        prev = cur
    return prev[len(y)]


def chain_score(labels: bytes, read: bytes, alphabet: int = 5) -> int:
    """Score of `read` against a single chain of chars (only valid for alphabet 5
    with N-free inputs, where sigma is plain equality)."""
    assert alphabet == 5
    best = None
    n = len(labels)
    for a in range(n):
        head = 0 if labels[a] == read[0] else 1
        for b in range(a, n):
            c = head + levenshtein(read[1:], labels[a + 1:b + 1])
            best = c if best is None else min(best, c)
    return best


def best_path_score(V: int, labels: bytes, in_edges: list, read: bytes, alphabet: int = 5,
                    costs=(0, 1, 1, 1)) -> int:
    """min over all alignment paths; in_edges[v] = list of char in-neighbours.
    The char graph must be a DAG; vertices are relaxed in topological order."""
    m, s, d, ins = costs
    outs = [[] for _ in range(V)]
    indeg = [0] * V
    for v in range(V):
        for u in in_edges[v]:
            outs[u].append(v)
            indeg[v] += 1
    order = [v for v in range(V) if indeg[v] == 0]
    k = 0
    while k < len(order):
        for w in outs[order[k]]:
            indeg[w] -= 1
            if indeg[w] == 0:
                order.append(w)
        k += 1
    assert len(order) == V, "not a DAG"
    INF = 1 << 40
    cost = [sigma(alphabet, labels[v], read[0], m, s) for v in range(V)]  # row 0: start states
    for i in range(1, len(read)):
        new = [INF] * V
        for v in order:
            best = cost[v] + d  # DELETE from (i-1, v)
            g = sigma(alphabet, labels[v], read[i], m, s)
            for u in in_edges[v]:
                best = min(best, cost[u] + g)  # MATCHSUB from (i-1, u)
                best = min(best, new[u] + ins)  # INSERT from (i, u); u precedes v in `order`
            new[v] = best
        cost = new
    return min(cost)


def enumerate_paths_score(V: int, labels: bytes, in_edges: list, read: bytes, alphabet: int = 5) -> int:
    """Exhaustive search over every op sequence (tiny graphs only)."""
    outs = [[] for _ in range(V)]
    for v in range(V):
        for u in in_edges[v]:
            outs[u].append(v)
    n = len(read)
    best = [1 << 40]

    def walk(i: int, v: int, c: int, ins_allowed: bool):
        if c >= best[0]:
            return
        if i == n - 1:
            best[0] = min(best[0], c)
        if i + 1 < n:
            walk(i + 1, v, c + 1, True)  # DELETE
            for w in outs[v]:
                walk(i + 1, w, c + sigma(alphabet, labels[w], read[i + 1]), True)  # MATCHSUB
        if ins_allowed:
            for w in outs[v]:
                walk(i, w, c + 1, True)  # INSERT (not on row 0)

    for v in range(V):
        walk(0, v, sigma(alphabet, labels[v], read[0]), False)
    return best[0]


def expand(strings: list, edges: list):
    """Char graph of a string graph (src/handlegraph_to_csr.cpp:165-247):
    returns (labels, str_off, char in-edge lists)."""
    off = [0]
    for s in strings:
        off.append(off[-1] + len(s))
    V = off[-1]
    ins = [[] for _ in range(V)]
    for si, s in enumerate(strings):
        for j in range(1, len(s)):
            ins[off[si] + j].append(off[si] + j - 1)
    for (a, b) in edges:
        ins[off[b]].append(off[a + 1] - 1)
    for l in ins:
        l.sort()
    return b"".join(s.encode() if isinstance(s, str) else s for s in strings), off, ins


def write_dataset(tmp, strings: list, edges: list, name: str = "g"):
    """Writes the five drop-in input files for a string graph; returns their paths."""
    import os
    labels, off, ins = expand(strings, edges)
    V = len(labels)
    outs = [[] for _ in range(V)]
    for v in range(V):
        for u in ins[v]:
            outs[u].append(v)

    def adj(path, n, out_lists):
        m = sum(len(o) for o in out_lists)
        with open(path, "w") as f:
            f.write(f"AdjacencyGraph\n{n}\n{m}\n")
            acc = 0
            for o in out_lists:
                f.write(f"{acc}\n")
                acc += len(o)
            for o in out_lists:
                for v in sorted(o):
                    f.write(f"{v}\n")
            for _ in range(m):
                f.write("1.0\n")

    sout = [[] for _ in strings]
    for a, b in edges:
        sout[a].append(b)
    p = {k: os.path.join(str(tmp), f"{name}.{k}") for k in ("char.adj", "str.adj", "labels", "label_offsets")}
    adj(p["char.adj"], V, outs)
    adj(p["str.adj"], len(strings), sout)
    with open(p["labels"], "wb") as f:
        f.write(labels + b"\n")
    with open(p["label_offsets"], "w") as f:
        f.write(" ".join(str(x) for x in off) + "\n")
    return p


def random_string_dag(rng, n_strings: int, max_len: int, p_edge: float = 0.3, alphabet=b"ACGT"):
    strings = [rbytes(rng, int(rng.integers(1, max_len + 1)), alphabet) for _ in range(n_strings)]
    edges = []
    for a in range(n_strings):
        for b in range(a + 1, n_strings):
            if b == a + 1 or rng.random() < p_edge:
                edges.append((a, b))
    return strings, edges


def rbytes(rng, n: int, alphabet=b"ACGT") -> bytes:
    """n random symbols of `alphabet` as bytes."""
    import numpy as np
    return bytes(np.asarray(rng.choice(list(alphabet), size=n), dtype=np.uint8))
