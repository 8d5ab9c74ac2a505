/*
 * Per-precision body of the CPU oracle.  TEST INFRASTRUCTURE ONLY: nothing in
 * the product path links or calls this file.  It is included twice by
 * hdb_oracle.c with REAL / SFX / REAL_MAXV defined for float and double.
 *
 * Every stage restates the reference's per-window algorithm, one window at a
 * time, serially, in the reference's Real precision, with no FMA contraction
 * (built with -ffp-contract=off):
 *   distances  reference hdbscan_gpu/_kernels.py:45-73
 *   core       reference hdbscan_gpu/_cuda.py:114-118 (self-inclusive order statistic)
 *   prim       reference hdbscan_gpu/_kernels.py:324-355 (serial form; same as
 *              sklearn _linkage.pyx:111-224)
 *   argsort    reference hdbscan_gpu/_kernels.py:358-516 (NumPy scalar aquicksort)
 *   linkage    reference hdbscan_gpu/_kernels.py:518-556
 *   condense   reference hdbscan_gpu/_kernels.py:558-643
 *   eom        reference hdbscan_gpu/_kernels.py:671-709
 *   labels     reference hdbscan_gpu/_kernels.py:712-760
 * The stages follow scikit-learn's _linkage.pyx / _tree.pyx and NumPy's
 * scalar aquicksort / aheapsort step for step (BSD-3; notices in
 * THIRD_PARTY_NOTICES.md at the repository root).
 */

#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define F(name) ORC_CAT(name, SFX)

/* d[i][j] = sqrt(sum_k (x_ik - x_jk)^2), k ascending from total = 0. */
static REAL F(orc_pair_)(const REAL *xi, const REAL *xj, int d)
{
    REAL total = 0;
    for (int k = 0; k < d; ++k) {
        REAL delta = xi[k] - xj[k];
        total = total + delta * delta;
    }
    return SQRTF(total);
}

void F(orc_distances_)(const REAL *x, int n, int d, REAL *out)
{
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            REAL v = F(orc_pair_)(x + (size_t)i * d, x + (size_t)j * d, d);
            out[(size_t)i * n + j] = v;
            out[(size_t)j * n + i] = v;
        }
}

static int F(orc_cmp_)(const void *a, const void *b)
{
    REAL x = *(const REAL *)a, y = *(const REAL *)b;
    return (x > y) - (x < y);
}

/* core[i] = (ms-1)-th smallest entry of row i, diagonal included. */
void F(orc_core_)(const REAL *dist, int n, int ms, REAL *core)
{
    REAL *row = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        memcpy(row, dist + (size_t)i * n, sizeof(REAL) * (size_t)n);
        qsort(row, (size_t)n, sizeof(REAL), F(orc_cmp_));
        core[i] = row[ms - 1];
    }
    free(row);
}

/* Serial Prim from vertex 0 over w = max(core_u, core_v, d_uv) (or d_uv).
 * Strict-less updates keep the earlier source on equal weights; the first
 * (lowest index) minimum becomes the next vertex. */
void F(orc_prim_)(const REAL *dist, const REAL *core, int n, int unit_core,
                  int *edge_u, int *edge_v, REAL *edge_w)
{
    unsigned char *in_tree = (unsigned char *)calloc((size_t)n, 1);
    REAL *best = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    int *sources = (int *)malloc(sizeof(int) * (size_t)n);
    for (int j = 0; j < n; ++j) { best[j] = (REAL)INFINITY; sources[j] = 1; }
    int current = 0;
    for (int i = 0; i < n - 1; ++i) {
        in_tree[current] = 1;
        REAL candidate = REAL_MAXV;
        int next = 0, source = 0;
        const REAL *row = dist + (size_t)current * n;
        for (int j = 0; j < n; ++j) {
            if (in_tree[j]) continue;
            REAL old = best[j];
            REAL w = row[j];
            if (!unit_core) {
                REAL c = FMAXF(core[current], core[j]);
                w = FMAXF(c, w);
            }
            if (w < old) {
                best[j] = w;
                sources[j] = current;
                if (w < candidate) { candidate = w; source = current; next = j; }
            } else if (old < candidate) {
                candidate = old;
                source = sources[j];
                next = j;
            }
        }
        edge_u[i] = source;
        edge_v[i] = next;
        edge_w[i] = candidate;
        current = next;
    }
    free(in_tree); free(best); free(sources);
}

/* NumPy scalar indirect heapsort on order[lo..hi]. */
static void F(orc_aheapsort_)(const REAL *v, int *order, int lo, int hi)
{
    int *a = order + lo - 1; /* one-based */
    int n = hi - lo + 1;
    for (int l = n >> 1; l > 0; --l) {
        int tmp = a[l], i = l, j = l << 1;
        while (j <= n) {
            if (j < n && v[a[j]] < v[a[j + 1]]) j += 1;
            if (v[tmp] < v[a[j]]) { a[i] = a[j]; i = j; j += j; }
            else break;
        }
        a[i] = tmp;
    }
    while (n > 1) {
        int tmp = a[n];
        a[n] = a[1];
        n -= 1;
        int i = 1, j = 2;
        while (j <= n) {
            if (j < n && v[a[j]] < v[a[j + 1]]) j++;
            if (v[tmp] < v[a[j]]) { a[i] = a[j]; i = j; j += j; }
            else break;
        }
        a[i] = tmp;
    }
}

#define ORC_SWAP(p, q) do { int t_ = (p); (p) = (q); (q) = t_; } while (0)

/* order = argsort(v) in NumPy's scalar introsort order.  depth < -1000 means
 * "use NumPy's initial depth 2*floor(log2(m))". */
void F(orc_argsort_)(const REAL *v, int m, int depth_override, int *order)
{
    for (int i = 0; i < m; ++i) order[i] = i;
    if (m < 2) return;
    int stack_lo[128], stack_hi[128], stack_depth[128];
    int top = 0;
    int lo = 0, hi = m - 1;
    int depth = 0;
    for (int c = m; c > 1; c >>= 1) depth += 2;
    if (depth_override > -1000) depth = depth_override;
    for (;;) {
        if (depth < 0) {
            F(orc_aheapsort_)(v, order, lo, hi);
        } else {
            while (hi - lo > ORC_SMALL_QUICKSORT) {
                int mid = lo + ((hi - lo) >> 1);
                if (v[order[mid]] < v[order[lo]]) ORC_SWAP(order[mid], order[lo]);
                if (v[order[hi]] < v[order[mid]]) ORC_SWAP(order[hi], order[mid]);
                if (v[order[mid]] < v[order[lo]]) ORC_SWAP(order[mid], order[lo]);
                REAL pivot = v[order[mid]];
                int i = lo, j = hi - 1;
                ORC_SWAP(order[mid], order[j]);
                for (;;) {
                    do { ++i; } while (v[order[i]] < pivot);
                    do { --j; } while (pivot < v[order[j]]);
                    if (i >= j) break;
                    ORC_SWAP(order[i], order[j]);
                }
                ORC_SWAP(order[i], order[hi - 1]);
                --depth;
                if (i - lo < hi - i) {
                    stack_lo[top] = i + 1; stack_hi[top] = hi; hi = i - 1;
                } else {
                    stack_lo[top] = lo; stack_hi[top] = i - 1; lo = i + 1;
                }
                stack_depth[top++] = depth;
            }
            for (int i = lo + 1; i <= hi; ++i) {
                int value = order[i], j = i;
                while (j > lo && v[value] < v[order[j - 1]]) { order[j] = order[j - 1]; --j; }
                order[j] = value;
            }
        }
        if (top == 0) break;
        --top;
        lo = stack_lo[top]; hi = stack_hi[top]; depth = stack_depth[top];
    }
}

/* Kruskal over the sorted MST edges.  Node n+i joins the newest nodes that
 * contain edge_u and edge_v; union by size with path halving. */
void F(orc_linkage_)(const int *edge_u, const int *edge_v, const REAL *edge_w,
                     const int *order, int n, int *left, int *right, int *sizes,
                     REAL *distance)
{
    int *uf = (int *)malloc(sizeof(int) * (size_t)n);
    int *node_of = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) { uf[i] = i; node_of[i] = i; }
    for (int i = 0; i < n - 1; ++i) {
        int e = order[i], a = edge_u[e], c = edge_v[e];
        while (uf[a] != a) { uf[a] = uf[uf[a]]; a = uf[a]; }
        while (uf[c] != c) { uf[c] = uf[uf[c]]; c = uf[c]; }
        int ca = node_of[a], cc = node_of[c];
        int sa = ca < n ? 1 : sizes[ca - n];
        int sc = cc < n ? 1 : sizes[cc - n];
        left[i] = ca; right[i] = cc; sizes[i] = sa + sc;
        distance[i] = edge_w[e];
        if (sa < sc) { uf[a] = c; node_of[c] = n + i; }
        else { uf[c] = a; node_of[a] = n + i; }
    }
    free(uf); free(node_of);
}

/* Breadth-first condensation from the root 2n-2.  Cluster n is the root;
 * children of a split with both sides >= mcs get consecutive new IDs.
 * Outputs are indexed by cluster - n (parent, birth, stability, selected)
 * and by point (point_parent).  Returns next_label. */
int F(orc_condense_)(const int *left, const int *right, const int *sizes,
                     const REAL *distance, int n, long long mcs,
                     int *parent, REAL *birth, REAL *stability,
                     unsigned char *selected, int *point_parent)
{
    int *relabel = (int *)malloc(sizeof(int) * (size_t)n);
    int *queue = (int *)malloc(sizeof(int) * (size_t)n);
    int *prune = (int *)malloc(sizeof(int) * (size_t)(2 * n));
    for (int i = 0; i < n; ++i) {
        relabel[i] = -1; parent[i] = -1; selected[i] = 0;
        birth[i] = 0; stability[i] = 0; point_parent[i] = -1;
    }
    int root = 2 * n - 2, head = 0, tail = 1, next_label = n + 1;
    queue[0] = root;
    relabel[root - n] = n;
    while (head < tail) {
        int node = queue[head++];
        int a = left[node - n], c = right[node - n];
        int cluster = relabel[node - n];
        REAL cluster_birth = birth[cluster - n];
        REAL cluster_stability = stability[cluster - n];
        REAL dist = distance[node - n];
        REAL lambda = dist > 0 ? (REAL)1 / dist : (REAL)INFINITY;
        long long size_a = a < n ? 1 : sizes[a - n];
        long long size_c = c < n ? 1 : sizes[c - n];
        int large_a = size_a >= mcs, large_c = size_c >= mcs;
        int children[2] = {a, c};
        if (large_a && large_c) {
            for (int k = 0; k < 2; ++k) {
                int child = children[k], label = next_label++;
                relabel[child - n] = label;
                parent[label - n] = cluster;
                birth[label - n] = lambda;
                selected[label - n] = 1;
                long long sz = k == 0 ? size_a : size_c;
                cluster_stability += (lambda - cluster_birth) * (REAL)sz;
                queue[tail++] = child;
            }
        } else {
            for (int k = 0; k < 2; ++k) {
                int child = children[k];
                if (k == 0 ? large_a : large_c) {
                    relabel[child - n] = cluster;
                    queue[tail++] = child;
                } else {
                    int ph = 0, pt = 0, leaf = child;
                    for (;;) {
                        if (leaf < n) {
                            point_parent[leaf] = cluster;
                            cluster_stability += lambda - cluster_birth;
                        } else {
                            prune[pt++] = left[leaf - n];
                            prune[pt++] = right[leaf - n];
                        }
                        if (ph == pt) break;
                        leaf = prune[ph++];
                    }
                }
            }
        }
        stability[cluster - n] = cluster_stability;
    }
    free(relabel); free(queue); free(prune);
    return next_label;
}

/* Excess-of-mass selection, optional epsilon merge, ascending-ID label
 * numbering and point labels.  selected/stability are updated in place. */
void F(orc_select_)(const int *parent, const REAL *birth, REAL *stability,
                    unsigned char *selected, const int *point_parent, int n,
                    int next_label, REAL epsilon, int *labels)
{
    int count = next_label - n; /* clusters n .. next_label-1 */
    REAL *subtree = (REAL *)calloc((size_t)count, sizeof(REAL));
    int *first_child = (int *)malloc(sizeof(int) * (size_t)count);
    int *next_sibling = (int *)malloc(sizeof(int) * (size_t)count);
    int *blocked = (int *)calloc((size_t)count, sizeof(int));
    int *stop = (int *)malloc(sizeof(int) * (size_t)count);
    int *mark = (int *)calloc((size_t)count, sizeof(int));
    int *label_map = (int *)malloc(sizeof(int) * (size_t)count);
    int *resolved = (int *)malloc(sizeof(int) * (size_t)count);
    (void)subtree;
    for (int c = 0; c < count; ++c) { first_child[c] = -1; next_sibling[c] = -1; label_map[c] = -1; }
    /* Children in ascending ID order. */
    for (int child = next_label - 1; child > n; --child) {
        int p = parent[child - n];
        next_sibling[child - n] = first_child[p - n];
        first_child[p - n] = child;
    }
    for (int node = next_label - 1; node > n; --node) {
        REAL s = 0;
        for (int ch = first_child[node - n]; ch != -1; ch = next_sibling[ch - n])
            s += stability[ch - n];
        if (s > stability[node - n]) {
            selected[node - n] = 0;
            stability[node - n] = s;
        }
    }
    blocked[0] = 0;
    for (int node = n + 1; node < next_label; ++node) {
        int inherited = blocked[parent[node - n] - n];
        blocked[node - n] = inherited || selected[node - n];
        if (inherited) selected[node - n] = 0;
    }
    if (epsilon > 0) {
        for (int node = n + 1; node < next_label; ++node) {
            int anc = parent[node - n];
            stop[node - n] = anc == n ? node
                : ((REAL)1 / birth[anc - n] > epsilon ? anc : stop[anc - n]);
            if (!selected[node - n]) continue;
            int chosen = (REAL)1 / birth[node - n] < epsilon ? stop[node - n] : node;
            mark[chosen - n] = 1;
        }
        for (int node = n + 1; node < next_label; ++node)
            selected[node - n] = (unsigned char)(mark[node - n] == 1);
    }
    int label = 0;
    for (int node = n + 1; node < next_label; ++node)
        if (selected[node - n]) label_map[node - n] = label++;
    resolved[0] = -1;
    for (int node = n + 1; node < next_label; ++node)
        resolved[node - n] = selected[node - n] ? label_map[node - n] : resolved[parent[node - n] - n];
    for (int i = 0; i < n; ++i) labels[i] = resolved[point_parent[i] - n];
    free(subtree); free(first_child); free(next_sibling); free(blocked);
    free(stop); free(mark); free(label_map); free(resolved);
}

/* Tree stages from given Prim edges: sort, link, condense, select, label.
 * Optional outputs (NULL to skip) expose every intermediate. */
int F(orc_tree_)(const int *edge_u, const int *edge_v, const REAL *edge_w, int n,
                 long long mcs, REAL epsilon, int depth_override, int *labels,
                 int *order_out, int *left_out, int *right_out, int *sizes_out,
                 REAL *distance_out, int *parent_out, REAL *birth_out,
                 REAL *stability_out, unsigned char *selected_out,
                 int *point_parent_out)
{
    int m = n - 1;
    int *order = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    int *left = (int *)malloc(sizeof(int) * (size_t)n);
    int *right = (int *)malloc(sizeof(int) * (size_t)n);
    int *sizes = (int *)malloc(sizeof(int) * (size_t)n);
    REAL *distance = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    int *parent = (int *)malloc(sizeof(int) * (size_t)n);
    REAL *birth = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    REAL *stability = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    unsigned char *selected = (unsigned char *)malloc((size_t)n);
    int *point_parent = (int *)malloc(sizeof(int) * (size_t)n);
    F(orc_argsort_)(edge_w, m, depth_override, order);
    F(orc_linkage_)(edge_u, edge_v, edge_w, order, n, left, right, sizes, distance);
    int next_label = F(orc_condense_)(left, right, sizes, distance, n, mcs, parent,
                                      birth, stability, selected, point_parent);
    if (order_out) memcpy(order_out, order, sizeof(int) * (size_t)m);
    if (left_out) memcpy(left_out, left, sizeof(int) * (size_t)m);
    if (right_out) memcpy(right_out, right, sizeof(int) * (size_t)m);
    if (sizes_out) memcpy(sizes_out, sizes, sizeof(int) * (size_t)m);
    if (distance_out) memcpy(distance_out, distance, sizeof(REAL) * (size_t)m);
    if (birth_out) memcpy(birth_out, birth, sizeof(REAL) * (size_t)n);
    F(orc_select_)(parent, birth, stability, selected, point_parent, n, next_label,
                   epsilon, labels);
    if (parent_out) memcpy(parent_out, parent, sizeof(int) * (size_t)n);
    if (stability_out) memcpy(stability_out, stability, sizeof(REAL) * (size_t)n);
    if (selected_out) memcpy(selected_out, selected, (size_t)n);
    if (point_parent_out) memcpy(point_parent_out, point_parent, sizeof(int) * (size_t)n);
    free(order); free(left); free(right); free(sizes); free(distance);
    free(parent); free(birth); free(stability); free(selected); free(point_parent);
    return next_label;
}

/* Whole pipeline for B windows of x[B][n][d].  Returns a status word:
 * bit 1 = non-finite input, bit 0 = non-finite pairwise distance.
 * Labels are written only when the status is 0. */
int F(orc_hdbscan_)(const REAL *x, int batch, int n, int d, long long mcs, int ms,
                    REAL epsilon, int *labels)
{
    int status = 0;
    for (size_t i = 0; i < (size_t)batch * n * d; ++i)
        if (!isfinite(x[i])) status |= 2;
    if (status) return status;
    REAL *dist = (REAL *)malloc(sizeof(REAL) * (size_t)n * n);
    REAL *core = (REAL *)calloc((size_t)n, sizeof(REAL));
    int *eu = (int *)malloc(sizeof(int) * (size_t)n);
    int *ev = (int *)malloc(sizeof(int) * (size_t)n);
    REAL *ew = (REAL *)malloc(sizeof(REAL) * (size_t)n);
    for (int b = 0; b < batch && !status; ++b) {
        F(orc_distances_)(x + (size_t)b * n * d, n, d, dist);
        for (size_t i = 0; i < (size_t)n * n; ++i)
            if (!isfinite(dist[i])) { status |= 1; break; }
        if (status) break;
        int unit = ms <= 2;
        if (!unit) F(orc_core_)(dist, n, ms, core);
        F(orc_prim_)(dist, core, n, unit, eu, ev, ew);
        F(orc_tree_)(eu, ev, ew, n, mcs, epsilon, -100000, labels + (size_t)b * n,
                     NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL);
    }
    free(dist); free(core); free(eu); free(ev); free(ew);
    return status;
}

#undef F
#undef ORC_CAT
#undef ORC_CAT2
