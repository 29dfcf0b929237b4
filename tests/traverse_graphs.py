"""Small CSR graphs with known BFS distances and components, shared by the
oracle tests and the GPU parity tests.  The shapes follow the reference's own
analytics tests (crates/graph-compute/tests/gpu_bfs.rs and
gpu_connected_components.rs): path, star, grid, disjoint cliques, isolated
nodes, two cycles and random symmetric graphs of up to 40 nodes."""
import numpy as np

from oracle import oracle as O


def csr_from_adj(adj):
    off = [0]
    nb = []
    for s in adj:
        nb.extend(sorted(s))
        off.append(len(nb))
    return np.array(off, np.uint32), np.array(nb, np.uint32)


def path(n):
    adj = [set() for _ in range(n)]
    for i in range(n - 1):
        adj[i].add(i + 1)
        adj[i + 1].add(i)
    return csr_from_adj(adj)


def star(leaves):
    adj = [set(range(1, leaves + 1))] + [{0} for _ in range(leaves)]
    return csr_from_adj(adj)


def grid(rows, cols):
    adj = [set() for _ in range(rows * cols)]
    for r in range(rows):
        for c in range(cols):
            v = r * cols + c
            if c + 1 < cols:
                adj[v].add(v + 1)
                adj[v + 1].add(v)
            if r + 1 < rows:
                adj[v].add(v + cols)
                adj[v + cols].add(v)
    return csr_from_adj(adj)


def cliques(k, size):
    n = k * size
    adj = [set(range((v // size) * size, (v // size + 1) * size)) - {v} for v in range(n)]
    return csr_from_adj(adj)


def cycles(*lengths):
    adj = []
    base = 0
    for L in lengths:
        for i in range(L):
            adj.append({base + (i + 1) % L, base + (i - 1) % L})
        base += L
    return csr_from_adj(adj)


def random_graph(seed, max_n=40, max_edges=80):
    """The reference proptest's shape: n in 1..40, up to 80 raw pairs mod n,
    symmetric, self loops dropped."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, max_n))
    m = int(rng.integers(0, max_edges))
    raw = rng.integers(0, 40, size=(m, 2)) % n
    adj = [set() for _ in range(n)]
    for a, b in raw:
        if a != b:
            adj[a].add(int(b))
            adj[b].add(int(a))
    return csr_from_adj(adj), int(rng.integers(0, 40)) % n


def broom(hub_leaves=6000, tail=100):
    """Hub 0 with `hub_leaves` leaves, every leaf also joined to one collector
    node, then a path of `tail` nodes from the collector.  From the hub the
    frontier jumps to thousands of nodes and back to one; from the tail's end it
    grows to thousands at the leaves."""
    n = 1 + hub_leaves + 1 + tail
    col = hub_leaves + 1
    pairs = []
    for l in range(1, hub_leaves + 1):
        pairs += [0, l, l, col]
    prev = col
    for t in range(tail):
        pairs += [prev, col + 1 + t]
        prev = col + 1 + t
    off, nb = O.csr_from_pairs(n, np.array(pairs, np.uint32))
    return off, nb


def directed_only(n=30, seed=5):
    """A CSR whose rows are not symmetric (entry v->u without u->v): components
    must still treat every entry as an undirected edge."""
    rng = np.random.default_rng(seed)
    adj = [set() for _ in range(n)]
    for _ in range(n // 2):
        a, b = rng.integers(0, n, 2)
        if a != b:
            adj[int(a)].add(int(b))
    return csr_from_adj(adj)


def directed_hub(leaves=10000):
    """A CSR that is not symmetric, with a frontier too large for the LDS queue:
    0 -> 1..leaves (level 1); leaf i -> leaves+i for i <= leaves/2 (level 2,
    reached only along out-edges); leaves/2 more nodes that point INTO the
    leaves and are never reached from 0.  A bottom-up level that scanned a
    node's own row would get both of the last two groups wrong."""
    half = leaves // 2
    n = 1 + leaves + half + half
    adj = [set() for _ in range(n)]
    adj[0] = set(range(1, leaves + 1))
    for i in range(1, half + 1):
        adj[i].add(leaves + i)
    for k in range(half):
        adj[1 + leaves + half + k].add(1 + k)
    return csr_from_adj(adj)


def shuffled_rows(off, nb, seed=1):
    """The same graph with every row in a random order (not searchable by
    bisection: such a CSR is treated as directed)."""
    rng = np.random.default_rng(seed)
    nb = nb.copy()
    for v in range(len(off) - 1):
        seg = nb[off[v]:off[v + 1]]
        rng.shuffle(seg)
    return off, nb
