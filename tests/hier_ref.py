"""Independent restatement of the MAS32 hierarchy from its definition.

Used only to check the oracle (and, through it, the library).  Where the
reference builds components with 32-bit masks and bit-BFS inside warps
(TraditionalMAS32Preconditioner.cu:463-1057), this file uses union-find over
explicit edge lists, so a shared misunderstanding of the bit tricks cannot
hide in both.

Definition:
  level 0 nodes are slots (bank = slot // B); two nodes are adjacent if a mesh
  edge or a contact pair joins them and they lie in the same bank.  The
  connected components within each bank become the level-1 nodes, numbered in
  order of (bank, lowest lane of the component).  Level l+1 is built from
  level l the same way, where level-l nodes adjacent through any mesh edge or
  contact pair of their fine vertices, and lying in one level-l bank, are
  joined.  Slots: level 0 occupies [0, n_slots); level l occupies
  [begin_l, begin_l + count_l) with begin_{l+1} = begin_l + pad_B(count_l).
"""
import numpy as np


def _pad(v, b):
    return (v + b - 1) // b * b


def decode_pair(p):
    x, y, z, w = (int(t) for t in p)
    if x >= 0:
        if w < 0:
            w = -w - 1
        return [x, y, z, w]
    x = -x - 1
    if z < 0:
        if y < 0:
            return [x, -y - 1, -z - 1, -w - 1]
        return [x, y]
    if w < 0:
        if y < 0:
            return [x, -y - 1, z, -w - 1]
        return [x, y, z]
    return [x, y, z, w]


class _UF:
    def __init__(self, n):
        self.p = list(range(n))

    def find(self, a):
        while self.p[a] != a:
            self.p[a] = self.p[self.p[a]]
            a = self.p[a]
        return a

    def union(self, a, b):
        ra, rb = self.find(a), self.find(b)
        if ra != rb:
            self.p[max(ra, rb)] = min(ra, rb)


def build(n_fem, n_slots, edges, part_to_real, real_to_part, levels, pairs=(), node_offset=0, bank=32):
    """edges: iterable of (u, v) real vertex ids.  Returns a dict like the oracle's."""
    groups = []   # per contact: list of fine vertex ids (in range)
    for p in pairs:
        vs = [v - node_offset for v in decode_pair(p)]
        groups.append([v for v in vs if 0 <= v < n_fem])
    links = [(int(u), int(v)) for u, v in edges if u != v]
    for g in groups:
        for i in range(len(g)):
            for j in range(i + 1, len(g)):
                if g[i] != g[j]:
                    links.append((g[i], g[j]))

    # level 0 ids are slots
    ident = np.array(real_to_part, dtype=np.int64)          # fine vertex -> node id at current level
    count = n_slots
    valid = [part_to_real[s] >= 0 for s in range(n_slots)]
    level_count = [0] * (levels + 1)
    level_begin = [0] * (levels + 1)
    level_begin[1] = _pad(n_slots, bank)
    going_next = {}
    coarse_space = []
    fine_mask = np.zeros(n_fem, np.uint32)
    prefix = np.zeros(n_slots // bank, np.int64)
    begin = 0
    for level in range(levels):
        uf = _UF(count)
        for u, v in links:
            a, b = int(ident[u]), int(ident[v])
            if a != b and a // bank == b // bank:
                uf.union(a, b)
        # components, ordered by (bank, lowest lane)
        newid = {}
        nxt = 0
        comps_per_bank = {}
        for c in range(count):
            if level == 0 and not valid[c]:
                continue
            r = uf.find(c)
            if r == c:                                      # c is the lowest member
                newid[c] = nxt
                nxt += 1
                comps_per_bank[c // bank] = comps_per_bank.get(c // bank, 0) + 1
        for c in range(count):
            if level == 0 and not valid[c]:
                continue
            newid[c] = newid[uf.find(c)]
        if level == 0:
            for b in range(n_slots // bank):
                prefix[b] = comps_per_bank.get(b, 0)
            for v in range(n_fem):
                s = int(real_to_part[v])
                root = uf.find(s)
                mask = 0
                for lane in range(bank):
                    t = (s // bank) * bank + lane
                    if t < n_slots and valid[t] and uf.find(t) == root:
                        mask |= 1 << lane
                fine_mask[v] = mask
        if level + 1 <= levels:
            level_count[level + 1] = nxt
            if level >= 1:
                level_begin[level + 1] = begin + _pad(count, bank)
        nb = level_begin[level + 1]
        if level == 0:
            for v in range(n_fem):
                going_next[v] = newid[int(real_to_part[v])] + nb
        else:
            for c in range(count):
                going_next[begin + c] = newid[c] + nb
        ident = np.array([newid[int(i)] for i in ident], dtype=np.int64)
        coarse_space.append(ident.copy())
        begin = level_begin[level + 1]
        count = nxt
    total = level_begin[levels]
    coarse_table = np.zeros((n_fem, 6), np.int64)
    for v in range(n_fem):
        cur = v
        for l in range(levels - 1):
            cur = going_next[cur]
            coarse_table[v, l] = cur
    return dict(level_count=np.array(level_count), level_begin=np.array(level_begin), fine_mask=fine_mask,
                prefix=prefix, going_next=going_next, coarse_table=coarse_table, total=total)


def components_from_hierarchy(h, r2p, bank):
    """Per-bank component lists (comp_ptr, comp_mask, comp_parent) implied by
    a hierarchy's going_next, fine masks and level layout: the members of a
    component share one parent, so grouping a bank's nodes by parent gives its
    components, and parents are numbered in (bank, leader lane) order, so
    sorting by parent gives the leader order.  The top level lists none."""
    L, T = int(h["levels"]), int(h["total"])
    lb, lc = h["level_begin"], h["level_count"]
    n_slots = int(lb[1])
    per_bank = [[] for _ in range(T // bank)]
    if L >= 2:
        seen = {}
        for v, s in enumerate(r2p):
            p = int(h["going_next"][v])
            seen.setdefault((int(s) // bank, p), int(h["fine_mask"][v]))
        for (b, p), m in seen.items():
            per_bank[b].append((p, m))
    for lvl in range(1, L - 1):
        groups = {}
        for s in range(int(lb[lvl]), int(lb[lvl]) + int(lc[lvl])):
            p = int(h["going_next"][s])
            groups[(s // bank, p)] = groups.get((s // bank, p), 0) | (1 << (s % bank))
        for (b, p), m in groups.items():
            per_bank[b].append((p, m))
    ptr = np.zeros(T // bank + 1, np.int32)
    masks, parents = [], []
    for b, lst in enumerate(per_bank):
        lst.sort()
        ptr[b + 1] = ptr[b] + len(lst)
        parents.extend(p for p, _ in lst)
        masks.extend(m for _, m in lst)
    assert n_slots % bank == 0
    return ptr, np.array(masks, np.uint32), np.array(parents, np.int32)
