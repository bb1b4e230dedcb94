"""Tree helpers shared by the tests: random topologies with AABBs, subtree
walks, degenerate meshes.  Plain numpy, no GPU."""
import numpy as np

from obvh_amd import synth
from obvh_amd.layouts import NODE_DTYPE

INV = 0xFFFFFFFF


def random_tree(n_leaves: int, seed: int):
    """Random binary tree over shuffled leaves: internal nodes numbered in
    breadth-first order (root 0), leaves I..2N-2 holding a random face
    permutation.  Sibling leaves are rarely adjacent."""
    rng = np.random.default_rng(seed)
    I = n_leaves - 1
    nodes = np.zeros(2 * n_leaves - 1, dtype=NODE_DTYPE)
    parents = np.zeros(2 * n_leaves - 1, dtype=np.uint32)
    parents[0] = INV
    leaf_slots = rng.permutation(n_leaves)  # leaf slot of each face
    queue = [(0, np.arange(n_leaves))]
    next_internal = 1
    head = 0
    while head < len(queue):
        idx, faces = queue[head]
        head += 1
        cut = int(rng.integers(1, len(faces)))
        kids = []
        for part in (faces[:cut], faces[cut:]):
            if len(part) == 1:
                kids.append(I + int(leaf_slots[part[0]]))
            else:
                kids.append(next_internal)
                queue.append((next_internal, part))
                next_internal += 1
        nodes[idx]["a"], nodes[idx]["b"] = kids
        for k in kids:
            parents[k] = idx
    for f in range(n_leaves):
        L = nodes[I + leaf_slots[f]]
        L["a"] = f
        L["nprim"] = 1
    return nodes, parents


def contiguous_tree(n_leaves: int, seed: int, spine_block: int = 0):
    """Random binary tree in which every subtree holds a contiguous run of
    leaf positions (leaf node I + f holds face f), internal nodes numbered
    breadth-first rather than as a Karras tree.  spine_block > 0: a node with
    more than 2 * spine_block leaves peels a block of spine_block to
    1.5 * spine_block leaves off its right end, so the leftmost leaves lie
    below a chain of about n_leaves / spine_block ancestors."""
    rng = np.random.default_rng(seed)
    I = n_leaves - 1
    nodes = np.zeros(2 * n_leaves - 1, dtype=NODE_DTYPE)
    parents = np.zeros(2 * n_leaves - 1, dtype=np.uint32)
    parents[0] = INV
    queue = [(0, 0, n_leaves)]  # (node, first leaf, leaf count)
    next_internal = 1
    head = 0
    while head < len(queue):
        idx, first, cnt = queue[head]
        head += 1
        if spine_block and cnt > 2 * spine_block:
            cut = cnt - spine_block - int(rng.integers(0, spine_block // 2 + 1))
        else:
            cut = int(rng.integers(1, cnt))
        kids = []
        for f0, c in ((first, cut), (first + cut, cnt - cut)):
            if c == 1:
                kids.append(I + f0)
            else:
                kids.append(next_internal)
                queue.append((next_internal, f0, c))
                next_internal += 1
        nodes[idx]["a"], nodes[idx]["b"] = kids
        for k in kids:
            parents[k] = idx
    leaf = nodes[I:]
    leaf["a"] = np.arange(n_leaves)
    leaf["nprim"] = 1
    return nodes, parents


def fill_bounds(nodes, V, F):
    """Leaf AABBs from their triangles, internal AABBs as unions (in place)."""
    total = nodes.shape[0]
    I = (total - 1) // 2
    leaf = nodes[I:]
    tri = V[F[leaf["a"], :3].astype(np.int64)]  # (N, 3, 3)
    leaf["bmin"] = tri.min(axis=1)
    leaf["bmax"] = tri.max(axis=1)
    order = bfs_order(nodes)
    for v in order[::-1]:
        if v < I:
            a, b = int(nodes[v]["a"]), int(nodes[v]["b"])
            nodes[v]["bmin"] = np.minimum(nodes[a]["bmin"], nodes[b]["bmin"])
            nodes[v]["bmax"] = np.maximum(nodes[a]["bmax"], nodes[b]["bmax"])
    return nodes


def bfs_order(nodes, root: int = 0):
    total = nodes.shape[0]
    I = (total - 1) // 2
    order, head = [root], 0
    while head < len(order):
        v = order[head]
        head += 1
        if v < I:
            order.append(int(nodes[v]["a"]))
            order.append(int(nodes[v]["b"]))
    return order


def leaves_right_first(nodes, v, I):
    """Leaf nodes below v visited right child first (recursive definition)."""
    if v >= I:
        return [v]
    return leaves_right_first(nodes, int(nodes[v]["b"]), I) + leaves_right_first(nodes, int(nodes[v]["a"]), I)


def flat_grid(n: int, seed: int):
    """An n x n vertex grid in the plane z = 0.25 (jittered in x and y), two
    triangles per cell: every node's points are coplanar, so the candidate
    search finds no apex off the base plane (its third degenerate exit)."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 2).astype(np.float64)
    g += rng.uniform(-0.3, 0.3, g.shape)
    V = np.concatenate([g / n, np.full((n * n, 1), 0.25)], 1).astype(np.float32)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a, b, c, d = (i * n + j).ravel(), (i * n + j + 1).ravel(), ((i + 1) * n + j).ravel(), ((i + 1) * n + j + 1).ravel()
    F = np.concatenate([np.stack([a, b, c], 1), np.stack([b, d, c], 1)])
    return V, np.concatenate([F, np.zeros((len(F), 1), np.int64)], 1).astype(np.uint32)


def doubled_vertices(seed: int):
    """An icosphere whose every vertex is stored twice, the faces using either
    copy at random: equal positions under different ids, so slab ties are
    decided by the tie rule and distinct ids need not be distinct points."""
    V, F = synth.icosphere(10, seed=seed)
    rng = np.random.default_rng(seed)
    V2 = np.concatenate([V, V])
    F2 = F.copy()
    F2[:, :3] += (rng.integers(0, 2, F[:, :3].shape) * len(V)).astype(np.uint32)
    return V2, F2


def slivers(seed: int):
    """Hairball tubes where every third triangle is collapsed onto a segment
    (its third vertex moved to the midpoint of the first two): zero-area
    triangles among ordinary ones."""
    V, F = synth.hairball(tubes=60, seed=seed)
    V = V.copy()
    for f in F[::3]:
        V[f[2]] = (V[f[0]] + V[f[1]]) * np.float32(0.5)
    return V, F
