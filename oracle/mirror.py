"""CPU restatement of the reference's supernode compressor with mirror vertices
and of its sketch iteration, for small graphs (pure Python over sets, the way the
reference's loops read).

TEST INFRASTRUCTURE ONLY: the checker of suminc_compress_mirror and of the
sketch phase driven through suminc_pr_step_phase / suminc_pr_correct; nothing in
the product imports it.

Every function names the reference lines it follows.  Where the reference leaves
an order open (std::unordered_set iteration), the restatement takes ids in
ascending order, clusters in cluster-id order, and numbers mirror ids cluster by
cluster, in-mirrors before out-mirrors, by master id.

Pinned by: the hand-worked example in tests/test_gpu_mirror.py (every mirror
rule: selection by mirror_k, the four benefits, absorption of unclustered
candidates, entry priority, the master types), the closed-form rows (a dense
solve per supernode source), and the end-to-end fixed point (the sketch query
equals scipy's solve of the plain system).
"""
from __future__ import annotations

import numpy as np

INF = 0xFFFFFFFF


def _adj(off, col):
    n = len(off) - 1
    return [list(map(int, col[off[v]:off[v + 1]])) for v in range(n)]


def _transpose(n, out):
    inn = [[] for _ in range(n)]
    for u in range(n):
        for v in out[u]:
            inn[v].append(u)
    return inn


def compress(off, col, cluster_of, min_node=2, max_node=1024, threshold=1.0, mirror_k=4.0):
    """get_init_supernode_by_clusters (compressor_base.h:3253-3611),
    final_build_supernode (:3616-3783), build_subgraph_mirror (:3786-3856),
    judge_out_bound_node (:3924-4007), get_nodetype_mirror (:4144-4182).
    Returns the structure (no rows)."""
    n = len(off) - 1
    out = _adj(off, col)
    inn = _transpose(n, out)
    use_mirror = mirror_k != 1e8
    k = int(mirror_k)
    obj = np.float32(threshold)
    # kept clusters in id order (compress_by_cluster, :3160-3250)
    raw = {}
    for v in range(n):
        c = int(cluster_of[v])
        if c != INF:
            raw.setdefault(c, []).append(v)
    clusters = [sorted(raw[c]) for c in sorted(raw) if min_node <= len(raw[c]) <= max_node]
    id2cl = [None] * n
    for j, P in enumerate(clusters):
        for v in P:
            id2cl[v] = j
    id2sp = [None] * n
    fc = [0] * n  # 0: FC_default_value, -1: member, 1: supernode source (Fc >= 0)
    members, sources, in_m, out_m = [], [], [], []
    for j, P in enumerate(clusters):
        in_freq, out_freq = {}, {}
        old_entry, old_exit = set(), set()
        inner = 0
        for u in P:
            for w in inn[u]:
                if id2cl[w] != j:
                    if use_mirror:
                        sp = id2sp[w]
                        if sp is None or u not in out_m[sp]:
                            in_freq[w] = in_freq.get(w, 0) + 1
                    old_entry.add(u)
                else:
                    inner += 1
            for w in out[u]:
                if id2cl[w] != j:
                    if use_mirror:
                        sp = id2sp[w]
                        if sp is None or u not in in_m[sp]:
                            out_freq[w] = out_freq.get(w, 0) + 1
                    old_exit.add(u)
        in_mirror = {w for w, f in in_freq.items() if f > k}
        out_mirror = {w for w, f in out_freq.items() if f > k}
        in_edges = sum(f for f in in_freq.values() if f > k)
        out_edges = sum(f for f in out_freq.values() if f > k)
        Pset = set(P)
        in_P, out_P = Pset | in_mirror, Pset | out_mirror
        B = {v for v in P if any(w not in out_P for w in out[v])}
        S = {v for v in P if any(w not in in_P for w in inn[v])}
        oX, oE = len(old_exit), len(old_entry)
        nX, nE = len(B) + len(out_mirror), len(S) + len(in_mirror)
        benefit = [np.float32(float(inner) - float(oX * oE)),
                   np.float32(float(inner + in_edges) - float(oX * nE + len(in_mirror))),
                   np.float32(float(inner + out_edges) - float(nX * oE + len(out_mirror))),
                   np.float32(float(inner + in_edges + out_edges) - float(nX * nE + len(in_mirror) + len(out_mirror)))]
        mi = 0
        for i in range(1, 4):
            if benefit[mi] < benefit[i]:
                mi = i
        if not benefit[mi] > obj:
            continue
        if mi == 0:
            S, in_mirror, out_mirror = set(old_entry), set(), set()
        elif mi == 1:
            out_mirror = set()
        elif mi == 2:
            S, in_mirror = set(old_entry), set()
        cid = len(members)
        members.append(list(P))
        sources.append(sorted(S))
        in_m.append(in_mirror)
        out_m.append(out_mirror)
        for u in P:
            fc[u] = -1
            id2sp[u] = cid
    K = len(members)
    # the accepted clusters before the final build (the cluster cache's content)
    acc = {"members": [list(P) for P in members], "sources": [list(S) for S in sources],
           "in": [sorted(x) for x in in_m], "out": [sorted(x) for x in out_m]}
    # final_build_supernode: absorption, mirror ids, supernode sources
    all_n = n
    master = []
    ids = []
    in_ids, out_ids = [], []
    src_list, src_cluster = [], []
    for i in range(K):
        P, S = members[i], sources[i]
        for which in (0, 1):
            cand = sorted(in_m[i] if which == 0 else out_m[i])
            dele = []
            for v in cand:
                if fc[v] != 0:
                    continue
                P.append(v)
                if any(id2sp[w] != i and w not in in_m[i] for w in inn[v]):
                    S.append(v)
                fc[v] = -1
                id2sp[v] = i
                dele.append(v)
            for v in dele:
                in_m[i].discard(v)
                out_m[i].discard(v)
        lst = list(P)
        ii, oo = [], []
        for v in sorted(in_m[i]):
            master.append(v)
            lst.append(all_n)
            ii.append(all_n)
            all_n += 1
        for v in sorted(out_m[i]):
            master.append(v)
            lst.append(all_n)
            oo.append(all_n)
            all_n += 1
        ids.append(lst)
        in_ids.append(ii)
        out_ids.append(oo)
        for s in S:
            fc[s] = 1
            src_list.append(s)
            src_cluster.append(i)
        for m in ii:
            src_list.append(m)
            src_cluster.append(i)
    cl_of = [INF] * all_n
    for v in range(n):
        if id2sp[v] is not None:
            cl_of[v] = id2sp[v]
    for i in range(K):
        for m in in_ids[i] + out_ids[i]:
            cl_of[m] = i
    # build_subgraph_mirror
    sub = [[] for _ in range(all_n)]
    for i in range(K):
        for v in members[i]:
            sub[v] = [w for w in out[v] if cl_of[w] == i]
        for m in in_ids[i]:
            sub[m] = [w for w in out[master[m - n]] if cl_of[w] == i]
    for i in range(K):
        for m in out_ids[i]:
            v = master[m - n]
            v_in = in_m[cl_of[v]]
            for w in inn[v]:
                if cl_of[w] == i and w not in v_in:
                    sub[w].append(m)
    # judge_out_bound_node
    out_bound = [False] * n
    for i in range(K):
        for v in members[i]:
            if any(cl_of[w] != i and w not in out_m[i] for w in out[v]):
                out_bound[v] = True
    for i in range(K):
        for v in in_m[i]:
            out_bound[v] = True
    # get_nodetype_mirror
    typ = [0] * n
    for v in range(n):
        t = 0
        if fc[v] == 0:
            t = 0
        elif fc[v] > 0:
            t = 2
        elif not out_bound[v]:
            t = 4
        if out_bound[v]:
            t = 3 if fc[v] > 0 else 1
        typ[v] = t
    for i in range(K):
        for v in in_m[i]:
            if typ[v] == 1:
                typ[v] = 5
            elif typ[v] == 3:
                typ[v] = 6
    return {"n": n, "all_node_num": all_n, "K": K, "members": members, "ids": ids, "in_m": in_m, "out_m": out_m,
            "in_ids": in_ids, "out_ids": out_ids, "master": master, "cluster_of": cl_of, "sub": sub,
            "out_bound": out_bound, "node_type": typ, "sources": src_list, "source_cluster": src_cluster,
            "out": out, "acceptance": acc}


def cluster_cache_bytes(st):
    """serialize_cluster (compressor_base.h:3081-3150) of the accepted clusters,
    little-endian, lists ascending: u32 init_mirror_num, u32 K, per cluster five
    lists (supernode_ids, cluster_ids, supernode_source, in-mirror masters, out-mirror
    masters) as u32 count + ids, i32 Fc[n] (-(cluster + 1) for members, INT32_MAX
    otherwise), u32 id2spids[n] (cluster or 0xFFFFFFFF), i32 -1."""
    a = st["acceptance"]
    n = st["n"]
    K = len(a["members"])
    words = [sum(len(x) for x in a["in"]) + sum(len(x) for x in a["out"]), K]
    fc = np.full(n, 2**31 - 1, np.int64)
    sp = np.full(n, INF, np.int64)
    for i in range(K):
        for lst in (a["members"][i], a["members"][i], a["sources"][i], a["in"][i], a["out"][i]):
            words.append(len(lst))
            words.extend(lst)
        for u in a["members"][i]:
            fc[u] = -(i + 1)
            sp[u] = i
    out = np.array(words, "<u4").tobytes() + fc.astype("<i4").tobytes() + sp.astype("<u4").tobytes()
    return out + np.array([-1], "<i4").tobytes()


def shortcut_rows(st, d):
    """build_iter_index_mirror (iter_compressor.h:318-491) in closed form: for each
    supernode source, reach = (I - A_mm)^-1 x0 over the cluster's members, where
    A_mm[x, v] = d / outdeg(v) per internal edge v -> x and x0 is the unit at an
    entry, or an in-mirror's first push d / outdeg(master) along its master's
    in-cluster edges; out-mirrors collect A_om reach.  Returns per source the
    bound_delta (exits, out-mirrors), inner_value (other members) and an empty
    inner_delta (the exact solve leaves no residual)."""
    n, out = st["n"], st["out"]
    deg = [len(out[v]) for v in range(n)]
    rows = {}
    for s, i in zip(st["sources"], st["source_cluster"]):
        mem = st["members"][i]
        midx = {v: a for a, v in enumerate(mem)}
        oidx = {m: a for a, m in enumerate(st["out_ids"][i])}
        A = np.zeros((len(mem), len(mem)))
        Ao = np.zeros((len(oidx), len(mem)))
        for a, v in enumerate(mem):
            for x in st["sub"][v]:
                if x in midx:
                    A[midx[x], a] += d / deg[v]
                else:
                    Ao[oidx[x], a] += d / deg[v]
        x0 = np.zeros(len(mem))
        if s < n:
            x0[midx[s]] = 1.0
        else:
            dm = deg[st["master"][s - n]]
            for x in st["sub"][s]:
                x0[midx[x]] += d / dm
        r = np.linalg.solve(np.eye(len(mem)) - A, x0)
        om = Ao @ r
        bd, iv = [], []
        for a, v in enumerate(mem):
            if r[a] != 0.0:
                (bd if st["out_bound"][v] else iv).append((v, r[a]))
        for m, a in oidx.items():
            if om[a] != 0.0:
                bd.append((m, om[a]))
        rows[s] = {"bound_delta": sorted(bd), "inner_value": iv, "inner_delta": []}
    return rows


def sketch_csr(st, rows):
    """sketch2csr_mirror (compressor_base.h:4636-4899): is_e_ (bound_delta of
    in-mirrors and of OnlyIn / BothOutIn / BothOutInMaster vertices), ib_e_
    (Singles: every out-edge; OnlyOut / BothOutIn / OutMaster / BothOutInMaster:
    the out-edges that leave the cluster, go to no out-mirror of it and leave no
    in-mirror of the vertex in the target's cluster), sync_e_ (master -> in-mirror,
    out-mirror -> master).  Rows as lists keyed by id."""
    n, all_n, out = st["n"], st["all_node_num"], st["out"]
    typ, cl = st["node_type"], st["cluster_of"]
    is_rows = [[] for _ in range(all_n)]
    for m in range(n, all_n):
        if m in rows:
            is_rows[m] = rows[m]["bound_delta"]
    for v in range(n):
        if typ[v] in (2, 3, 6):
            is_rows[v] = rows[v]["bound_delta"]
    ib = [[] for _ in range(n)]
    for u in range(n):
        if typ[u] == 0:
            ib[u] = list(out[u])
        elif typ[u] in (1, 3, 5, 6):
            i = cl[u]
            for w in out[u]:
                ow = cl[w]
                if ow == i or w in st["out_m"][i]:
                    continue
                if ow != INF and u in st["in_m"][ow]:
                    continue
                ib[u].append(w)
    sync = [[] for _ in range(all_n)]
    for i in range(st["K"]):
        for m in st["in_ids"][i]:
            sync[st["master"][m - n]].append(m)
        for m in st["out_ids"][i]:
            sync[m].append(st["master"][m - n])
    iv = [[] for _ in range(all_n)]
    idl = [[] for _ in range(all_n)]
    for s, r in rows.items():
        iv[s] = r["inner_value"]
        idl[s] = r["inner_delta"]
    return {"is": is_rows, "ib": ib, "sync": sync, "iv": iv, "id": idl}


def to_arrays(st, csr):
    """The suminc_sketch_csr layout (row pointers + columns + float weights)."""
    def rows(r, weighted):
        off = np.zeros(len(r) + 1, np.uint64)
        cols, ws = [], []
        for a, row in enumerate(r):
            off[a + 1] = off[a] + len(row)
            for e in row:
                if weighted:
                    cols.append(e[0])
                    ws.append(e[1])
                else:
                    cols.append(e)
        return off, np.array(cols, np.uint32), np.array(ws, np.float32)
    out = {"n": st["n"], "all_node_num": st["all_node_num"],
           "node_type": np.array(st["node_type"], np.uint8),
           "mirror_master": np.array(st["master"], np.uint32),
           "cluster_of": np.array(st["cluster_of"], np.uint32)}
    out["is_off"], out["is_col"], out["is_w"] = rows(csr["is"], True)
    out["ib_off"], out["ib_col"], _ = rows(csr["ib"], False)
    out["sync_off"], out["sync_col"], _ = rows(csr["sync"], False)
    out["iv_off"], out["iv_col"], out["iv_w"] = rows(csr["iv"], True)
    out["id_off"], out["id_col"], out["id_w"] = rows(csr["id"], True)
    out["sub_off"], out["sub_col"], _ = rows(st["sub"], False)
    return out


def _row(sk, name, u):
    o = sk[name + "_off"]
    cols = sk[name + "_col"][o[u]:o[u + 1]].astype(np.int64)
    w = sk[name + "_w"][o[u]:o[u + 1]].astype(np.float64) if name + "_w" in sk else None
    return cols, w


def _exit_pass(sk, off, src, rank, delta, bound, spnode, d, eps, flush):
    """Exits (OnlyOut, BothOutIn, OutMaster, BothOutInMaster) send their pending
    sum src[v] (|src| > eps, or any non-zero with flush): into rank (unless rank is
    None: the sum is in rank already), d/outdeg of it along the ib_e_ row, and a
    master into each in-mirror m of its sync_e_ row (spnode[m], then bound over m's
    is_e_ row); then the pushes land in delta and the out-mirrors flush their bound
    into their masters' delta (sum_sync_iter_worker.h:1042-1127;
    tjn::OutMirrorSyncToMaster)."""
    n, all_n = int(sk["n"]), int(sk["all_node_num"])
    typ = sk["node_type"]
    deg = np.diff(off)
    pushed = np.zeros(n)
    items = []
    for v in range(n):
        t = typ[v]
        if t not in (1, 3, 5, 6):
            continue
        b = src[v]
        if not (b != 0.0 if flush else abs(b) > eps):
            continue
        src[v] = 0.0
        if rank is not None:
            rank[v] += b
        cols, _ = _row(sk, "ib", v)
        if deg[v] and len(cols):
            np.add.at(pushed, cols, d * b / deg[v])
        if t in (5, 6):
            for m in _row(sk, "sync", v)[0]:
                spnode[m] += b
                items.append((int(m), b))
    for m, b in items:
        cols, w = _row(sk, "is", m)
        np.add.at(bound, cols, b * w)
    delta += pushed
    for o in range(n, all_n):
        cols, _ = _row(sk, "sync", o)
        if len(cols) and bound[o] != 0.0:
            delta[int(cols[0])] += bound[o]
            bound[o] = 0.0


def precompute(sk, off, delta, rank, bound, spnode, d, T, maxit=100):
    """The precompute that opens the compressed phase (precompute_spnode_one /
    run_to_convergence_for_precpt, iter_compressor.h:534-560, 725-765; the exits'
    send, sum_sync_iter_worker.h:387-431; precompute_spnode_two, :714-722) as the
    device runs it: synchronous Jacobi rounds of all clusters at once over the
    subgraph rows (every member with delta takes it into rank and its pass-through
    sum, then sends d/outdeg along the row: members' delta, out-mirrors' bound),
    until the mass moved in a round is at most T/n per member or maxit rounds;
    then the exits send their pass-through sums."""
    n = int(sk["n"])
    typ = sk["node_type"]
    deg = np.diff(off)
    members = [v for v in range(n) if typ[v] != 0]
    pre = np.zeros(n)
    thr = T / n * len(members)
    for _ in range(maxit if members else 0):
        moved = 0.0
        items = []
        for v in members:
            x = delta[v]
            if x == 0.0:
                continue
            delta[v] = 0.0
            rank[v] += x
            pre[v] += x
            moved += abs(x)
            cols, _ = _row(sk, "sub", v)
            if deg[v] and len(cols):
                items.append((cols, d * x / deg[v]))
        for cols, c in items:
            mem = cols[cols < n]
            np.add.at(delta, mem, c)
            np.add.at(bound, cols[cols >= n], c)
        if moved <= thr:
            break
    _exit_pass(sk, off, pre, None, delta, bound, spnode, d, 0.0, True)


def sketch_round(sk, off, delta, rank, bound, spnode, d, eps):
    """One synchronous round of the sketch phase by node type
    (sum_sync_iter_worker.h:949-1127) in the order the device runs it: Singles
    push and entries inject (delta -> spnode, bound += delta * is_e_); exits turn
    their inflow into rank and push d/outdeg of it along ib_e_, masters feed their
    in-mirrors' rows; then every push lands in delta and the out-mirrors flush into
    their masters' delta.  Arrays (FP64) are updated in place."""
    n = int(sk["n"])
    typ = sk["node_type"]
    deg = np.diff(off)
    pushed = np.zeros(n)
    for v in range(n):  # Singles and entries
        t = typ[v]
        if t == 0 and abs(delta[v]) > eps:
            x = delta[v]
            delta[v] = 0.0
            rank[v] += x
            cols, _ = _row(sk, "ib", v)
            if deg[v] and len(cols):
                np.add.at(pushed, cols, d * x / deg[v])
        elif t in (2, 3, 6) and abs(delta[v]) > eps:
            x = delta[v]
            delta[v] = 0.0
            spnode[v] += x
            cols, w = _row(sk, "is", v)
            np.add.at(bound, cols, x * w)
    # the Singles' pushes land with the exits' (nothing in the exit pass reads delta)
    delta += pushed
    _exit_pass(sk, off, bound, rank, delta, bound, spnode, d, eps, False)


def correct(sk, off, delta, rank, bound, spnode, d, T):
    """The correction (sum_sync_iter_worker.h:1385-1491) as the device runs it: the
    exits' pending inflow is flushed (exit passes until sum|bound| <= T / 1000, at
    most 256; the rest is dropped), then every non-zero spnode is spread through
    inner_value (rank) and inner_delta (delta)."""
    all_n = int(sk["all_node_num"])
    for _ in range(256):
        if np.abs(bound).sum() <= 1e-3 * T:
            break
        _exit_pass(sk, off, bound, rank, delta, bound, spnode, d, 0.0, True)
    bound[:] = 0.0
    for u in range(all_n):
        s = spnode[u]
        if s == 0.0:
            continue
        spnode[u] = 0.0
        cols, w = _row(sk, "iv", u)
        np.add.at(rank, cols, s * w)
        cols, w = _row(sk, "id", u)
        np.add.at(delta, cols, s * w)
