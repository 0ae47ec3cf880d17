/* Column ordering oracle: restates src/solver/order.zig. TEST
 * INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   or_order       order.zig:78-147   BTF blocks, sink first; AMD vs natural
 *                                     order per block by Cholesky count
 *   or_blocks      order.zig:154-235  iterative Tarjan SCC
 *   or_chol_count  order.zig:251-311  Liu's elimination tree count
 *   or_diag_fill   order.zig:241-244
 *   amd            order.zig:375-692  quotient-graph AMD with aggressive
 *                                     absorption, hashed supervariables,
 *                                     dense-row deferral
 *   compact        order.zig:321-344
 * The workspace sizes of the reference slab (4 nnz + 4 n + 16 variable
 * words, 8 n + nnz element words) are kept, so compaction triggers where
 * the reference's would; results do not depend on it.
 * Parity pinned by: tests/test_oracle_order.py (the known answers of
 * src/solver/tests.zig:2610-2781: star and arrowhead hubs last, BTF sink
 * block first, upper-triangular order [0,1,2], 1x1; order.zig:721-785
 * permutation and cap properties; the Cholesky count against a brute-force
 * symbolic elimination; AMD's fill never above the natural order's where
 * `order` picks it). Bit-exactness of AMD tie-breaking against the Zig
 * build: parity unpinned beyond those known answers (the reference cannot
 * run here).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NONE OR_NONE
#define SPAN_TAG (1u << 31)

int or_blocks(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *emit, uint32_t *block_ptr)
{
    uint32_t *num = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *low = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint8_t *on = calloc(n ? n : 1, 1);
    uint32_t *scc = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *fv = malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *fpp = malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) num[i] = NONE;
    uint32_t sp = 0, fp = 0, counter = 0, no = 0, nb = 0;
    block_ptr[0] = 0;
    for (uint32_t root = 0; root < n; root++) {
        if (num[root] != NONE) continue;
        num[root] = counter;
        low[root] = counter;
        counter++;
        scc[sp++] = root;
        on[root] = 1;
        fv[fp] = root;
        fpp[fp] = col_ptr[root];
        fp++;
        while (fp > 0) {
            uint32_t v = fv[fp - 1];
            if (fpp[fp - 1] < col_ptr[v + 1]) {
                uint32_t child = row_idx[fpp[fp - 1]];
                fpp[fp - 1]++;
                if (child == v) continue;
                if (num[child] == NONE) {
                    num[child] = counter;
                    low[child] = counter;
                    counter++;
                    scc[sp++] = child;
                    on[child] = 1;
                    fv[fp] = child;
                    fpp[fp] = col_ptr[child];
                    fp++;
                } else if (on[child]) {
                    if (num[child] < low[v]) low[v] = num[child];
                }
                continue;
            }
            if (low[v] == num[v]) {
                for (;;) {
                    sp--;
                    uint32_t m = scc[sp];
                    on[m] = 0;
                    emit[no++] = m;
                    if (m == v) break;
                }
                nb++;
                block_ptr[nb] = no;
            }
            fp--;
            if (fp > 0) {
                uint32_t u = fv[fp - 1];
                if (low[v] < low[u]) low[u] = low[v];
            }
        }
    }
    free(num); free(low); free(on); free(scc); free(fv); free(fpp);
    return (int)nb;
}

uint64_t or_chol_count(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *perm, uint64_t cap)
{
    uint32_t nnz = col_ptr[n];
    uint32_t *pinv = malloc(sizeof(uint32_t) * (n + 1));
    uint32_t *ptr = calloc(n + 1, sizeof(uint32_t));
    uint32_t *lst = malloc(sizeof(uint32_t) * (nnz + 1));
    uint32_t *fill = malloc(sizeof(uint32_t) * (n + 1));
    uint32_t *parent = malloc(sizeof(uint32_t) * (n + 1));
    for (uint32_t k = 0; k < n; k++) pinv[perm[k]] = k;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t a = pinv[row_idx[p]], b = pinv[j];
            if (a != b) ptr[(a > b ? a : b) + 1]++;
        }
    for (uint32_t k = 1; k <= n; k++) ptr[k] += ptr[k - 1];
    memcpy(fill, ptr, sizeof(uint32_t) * n);
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t a = pinv[row_idx[p]], b = pinv[j];
            if (a == b) continue;
            uint32_t hi = a > b ? a : b;
            lst[fill[hi]] = a > b ? b : a;
            fill[hi]++;
        }
    uint32_t *anc = fill;
    for (uint32_t k = 0; k < n; k++) {
        parent[k] = NONE;
        anc[k] = NONE;
        for (uint32_t t = ptr[k]; t < ptr[k + 1]; t++) {
            uint32_t i = lst[t];
            while (i != NONE && i < k) {
                uint32_t next = anc[i];
                anc[i] = k;
                if (next == NONE) parent[i] = k;
                i = next;
            }
        }
    }
    uint32_t *flag = anc;
    uint64_t count = 0;
    for (uint32_t k = 0; k < n; k++) {
        flag[k] = k;
        count++;
        for (uint32_t t = ptr[k]; t < ptr[k + 1]; t++) {
            uint32_t i = lst[t];
            while (flag[i] != k) {
                flag[i] = k;
                count++;
                i = parent[i];
            }
        }
        if (count >= cap) break;
    }
    free(pinv); free(ptr); free(lst); free(fill); free(parent);
    return count;
}

uint64_t or_diag_fill(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *q)
{
    return 2 * (or_chol_count(n, col_ptr, row_idx, q, UINT64_MAX) - n);
}

static size_t compact(uint32_t *adj, size_t top, uint32_t *pe, const uint32_t *len, const uint32_t *live_a, const uint32_t *live_b, uint32_t *first, uint32_t n)
{
    for (uint32_t o = 0; o < n; o++) {
        if (len[o] == 0 || (live_a[o] == 0 && live_b[o] == 0)) continue;
        first[o] = adj[pe[o]];
        adj[pe[o]] = SPAN_TAG | o;
    }
    size_t dst = 0, src = 0;
    while (src < top) {
        if ((adj[src] & SPAN_TAG) == 0) {
            src++;
            continue;
        }
        uint32_t o = adj[src] & ~SPAN_TAG;
        uint32_t l = len[o];
        adj[dst] = first[o];
        memmove(adj + dst + 1, adj + src + 1, sizeof(uint32_t) * (l - 1));
        pe[o] = (uint32_t)dst;
        dst += l;
        src += l;
    }
    return dst;
}

typedef struct { uint32_t *head, *next, *prev; } deglists;

static void dl_insert(deglists *s, uint32_t d, uint32_t i)
{
    s->next[i] = s->head[d];
    s->prev[i] = NONE;
    if (s->head[d] != NONE) s->prev[s->head[d]] = i;
    s->head[d] = i;
}

static void dl_remove(deglists *s, uint32_t d, uint32_t i)
{
    if (s->prev[i] != NONE) s->next[s->prev[i]] = s->next[i];
    else s->head[d] = s->next[i];
    if (s->next[i] != NONE) s->prev[s->next[i]] = s->prev[i];
}

static int same_adj(const uint32_t *va, const uint32_t *va_pe, const uint32_t *va_len, const uint32_t *ea, const uint32_t *ea_pe, const uint32_t *ea_len, uint32_t *mark, uint32_t *era, uint32_t i, uint32_t j)
{
    if (va_len[i] != va_len[j] || ea_len[i] != ea_len[j]) return 0;
    (*era)++;
    uint32_t m = *era;
    for (uint32_t t = 0; t < va_len[i]; t++) mark[va[va_pe[i] + t]] = m;
    for (uint32_t t = 0; t < va_len[j]; t++) if (mark[va[va_pe[j] + t]] != m) return 0;
    (*era)++;
    m = *era;
    for (uint32_t t = 0; t < ea_len[i]; t++) mark[ea[ea_pe[i] + t]] = m;
    for (uint32_t t = 0; t < ea_len[j]; t++) if (mark[ea[ea_pe[j] + t]] != m) return 0;
    return 1;
}

static uint32_t isqrt_u32(uint32_t n)
{
    /* floor(sqrt(n)), std.math.sqrt on an integer */
    uint32_t r = 0;
    while ((uint64_t)(r + 1) * (r + 1) <= n) r++;
    return r;
}

int or_amd(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q)
{
    if (n == 0) return 0;
    uint32_t nnz = col_ptr[n];
    size_t va_cap = 4 * (size_t)nnz + 4 * (size_t)n + 16;
    size_t ea_cap = 8 * (size_t)n + nnz;
    uint32_t *va = malloc(sizeof(uint32_t) * va_cap);
    uint32_t *va_pe = malloc(sizeof(uint32_t) * n);
    uint32_t *va_len = malloc(sizeof(uint32_t) * n);
    uint32_t *ea = malloc(sizeof(uint32_t) * ea_cap);
    uint32_t *ea_pe = malloc(sizeof(uint32_t) * n);
    uint32_t *ea_len = calloc(n, sizeof(uint32_t));
    uint32_t *ea_lim = malloc(sizeof(uint32_t) * n);
    uint32_t *nv = malloc(sizeof(uint32_t) * n);
    uint32_t *elem_alive = calloc(n, sizeof(uint32_t));
    uint32_t *esize = malloc(sizeof(uint32_t) * n);
    uint32_t *deg = malloc(sizeof(uint32_t) * n);
    uint32_t *w_buf = malloc(sizeof(uint32_t) * n);
    uint32_t *mark = calloc(n, sizeof(uint32_t));
    uint32_t *wmark = calloc(n, sizeof(uint32_t));
    uint32_t *hkey = malloc(sizeof(uint32_t) * n);
    uint32_t *hhead = malloc(sizeof(uint32_t) * n);
    uint32_t *hnext = malloc(sizeof(uint32_t) * n);
    uint32_t *hbuckets = malloc(sizeof(uint32_t) * n);
    uint32_t *mlink = malloc(sizeof(uint32_t) * n);
    uint32_t *lp_buf = malloc(sizeof(uint32_t) * n);
    uint32_t *touched = malloc(sizeof(uint32_t) * n);
    uint32_t *first = malloc(sizeof(uint32_t) * n);
    deglists dl;
    dl.head = malloc(sizeof(uint32_t) * (n + 1));
    dl.next = malloc(sizeof(uint32_t) * n);
    dl.prev = malloc(sizeof(uint32_t) * n);
    int rc = 0;
    uint32_t era = 0, erw = 0;
    for (uint32_t i = 0; i < n; i++) {
        nv[i] = 1;
        hhead[i] = NONE;
        mlink[i] = NONE;
    }
    for (uint32_t i = 0; i <= n; i++) dl.head[i] = NONE;

    memset(deg, 0, sizeof(uint32_t) * n);
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t r = row_idx[p];
            if (r == j) continue;
            deg[j]++;
            deg[r]++;
        }
    size_t va_free = 0;
    for (uint32_t i = 0; i < n; i++) {
        va_pe[i] = (uint32_t)va_free;
        va_free += deg[i];
    }
    memset(deg, 0, sizeof(uint32_t) * n);
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t r = row_idx[p];
            if (r == j) continue;
            va[va_pe[j] + deg[j]] = r;
            deg[j]++;
            va[va_pe[r] + deg[r]] = j;
            deg[r]++;
        }
    memcpy(va_len, deg, sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++) {
        era++;
        uint32_t wp = 0, s = va_pe[i];
        for (uint32_t t = 0; t < va_len[i]; t++) {
            uint32_t v = va[s + t];
            if (mark[v] != era) {
                mark[v] = era;
                va[s + wp] = v;
                wp++;
            }
        }
        va_len[i] = wp;
    }

    uint32_t isq = isqrt_u32(n);
    uint32_t dense_thresh = 10 * isq > 16 ? 10 * isq : 16;
    uint32_t ndense = 0;
    for (uint32_t i = 0; i < n; i++) ndense += va_len[i] >= dense_thresh;
    uint32_t dense_start = n - ndense;
    {
        uint32_t di = 0;
        for (uint32_t i = 0; i < n; i++)
            if (va_len[i] >= dense_thresh) {
                nv[i] = 0;
                lp_buf[dense_start + di] = i;
                di++;
            }
    }
    if (ndense > 0) {
        for (uint32_t i = 0; i < n; i++) {
            if (nv[i] == 0) continue;
            uint32_t wp = 0, s = va_pe[i];
            for (uint32_t t = 0; t < va_len[i]; t++) {
                uint32_t v = va[s + t];
                if (nv[v] != 0) {
                    va[s + wp] = v;
                    wp++;
                }
            }
            va_len[i] = wp;
        }
    }
    size_t ea_free = 0;
    for (uint32_t i = 0; i < n; i++) {
        ea_pe[i] = (uint32_t)ea_free;
        ea_lim[i] = 4;
        ea_free += 4;
    }
    for (uint32_t i = 0; i < n; i++) {
        if (nv[i] == 0) continue;
        deg[i] = va_len[i];
        dl_insert(&dl, deg[i], i);
    }

    uint32_t n_amd = n - ndense;
    uint32_t mindeg = 0, k = 0;
    while (k < n_amd) {
        while (dl.head[mindeg] == NONE) mindeg++;
        uint32_t p = dl.head[mindeg];
        dl_remove(&dl, deg[p], p);

        era++;
        mark[p] = era;
        uint32_t nlp = 0;
        for (uint32_t t = 0; t < va_len[p]; t++) {
            uint32_t v = va[va_pe[p] + t];
            if (nv[v] != 0 && mark[v] != era) {
                mark[v] = era;
                lp_buf[nlp++] = v;
            }
        }
        for (uint32_t t = 0; t < ea_len[p]; t++) {
            uint32_t e = ea[ea_pe[p] + t];
            if (elem_alive[e] == 0) continue;
            for (uint32_t u = 0; u < va_len[e]; u++) {
                uint32_t v = va[va_pe[e] + u];
                if (nv[v] != 0 && mark[v] != era) {
                    mark[v] = era;
                    lp_buf[nlp++] = v;
                }
            }
            elem_alive[e] = 0;
        }
        uint32_t nvpiv = nv[p];
        nv[p] = 0;
        elem_alive[p] = 1;

        if (va_free + nlp > va_cap) {
            va_len[p] = 0;
            va_free = compact(va, va_free, va_pe, va_len, nv, elem_alive, first, n);
            if (va_free + nlp > va_cap) { rc = -1; goto done; }
        }
        va_pe[p] = (uint32_t)va_free;
        va_len[p] = nlp;
        memcpy(va + va_free, lp_buf, sizeof(uint32_t) * nlp);
        va_free += nlp;

        uint32_t lpsize = 0;
        for (uint32_t t = 0; t < nlp; t++) lpsize += nv[lp_buf[t]];
        esize[p] = lpsize;
        for (uint32_t t = 0; t < nlp; t++) dl_remove(&dl, deg[lp_buf[t]], lp_buf[t]);

        erw++;
        uint32_t ntouched = 0;
        for (uint32_t t = 0; t < nlp; t++) {
            uint32_t i = lp_buf[t];
            for (uint32_t u = 0; u < ea_len[i]; u++) {
                uint32_t e = ea[ea_pe[i] + u];
                if (elem_alive[e] == 0) continue;
                if (wmark[e] != erw) {
                    wmark[e] = erw;
                    w_buf[e] = esize[e];
                    touched[ntouched++] = e;
                }
                w_buf[e] -= w_buf[e] < nv[i] ? w_buf[e] : nv[i];
            }
        }
        for (uint32_t t = 0; t < ntouched; t++)
            if (w_buf[touched[t]] == 0) elem_alive[touched[t]] = 0;
        for (uint32_t t = 0; t < nlp; t++) {
            uint32_t i = lp_buf[t];
            uint32_t esum = 0, ewp = 0;
            for (uint32_t u = 0; u < ea_len[i]; u++) {
                uint32_t e = ea[ea_pe[i] + u];
                if (elem_alive[e] != 0 && e != p) {
                    ea[ea_pe[i] + ewp] = e;
                    ewp++;
                    esum += w_buf[e];
                }
            }
            if (ewp < ea_lim[i]) {
                ea[ea_pe[i] + ewp] = p;
            } else {
                uint32_t nc = (ewp + 1) * 2;
                if (ea_free + nc > ea_cap) {
                    ea_len[i] = ewp;
                    ea_free = compact(ea, ea_free, ea_pe, ea_len, nv, nv, first, n);
                    for (uint32_t o = 0; o < n; o++) ea_lim[o] = ea_len[o];
                    if (ea_free + nc > ea_cap) { rc = -1; goto done; }
                }
                memmove(ea + ea_free, ea + ea_pe[i], sizeof(uint32_t) * ewp);
                ea[ea_free + ewp] = p;
                ea_pe[i] = (uint32_t)ea_free;
                ea_lim[i] = nc;
                ea_free += nc;
            }
            ea_len[i] = ewp + 1;

            uint32_t asum = 0, vwp = 0;
            for (uint32_t u = 0; u < va_len[i]; u++) {
                uint32_t v = va[va_pe[i] + u];
                if (nv[v] != 0 && mark[v] != era) {
                    va[va_pe[i] + vwp] = v;
                    vwp++;
                    asum += nv[v];
                }
            }
            va_len[i] = vwp;
            uint32_t cap = n_amd - k - nvpiv;
            uint32_t dgi = asum + (lpsize - nv[i]) + esum;
            deg[i] = dgi < cap ? dgi : cap;
        }

        uint32_t nhb = 0;
        for (uint32_t t = 0; t < nlp; t++) {
            uint32_t i = lp_buf[t];
            if (nv[i] == 0) continue;
            uint32_t h = 0;
            for (uint32_t u = 0; u < va_len[i]; u++) h += va[va_pe[i] + u];
            for (uint32_t u = 0; u < ea_len[i]; u++) h += ea[ea_pe[i] + u];
            hkey[i] = h;
            uint32_t b = h % n;
            if (hhead[b] == NONE) hbuckets[nhb++] = b;
            hnext[i] = hhead[b];
            hhead[b] = i;
        }
        for (uint32_t hb = 0; hb < nhb; hb++) {
            uint32_t b = hbuckets[hb];
            for (uint32_t i = hhead[b]; i != NONE; i = hnext[i]) {
                if (nv[i] == 0) continue;
                for (uint32_t j = hnext[i]; j != NONE; j = hnext[j]) {
                    if (nv[j] == 0 || hkey[j] != hkey[i]) continue;
                    if (!same_adj(va, va_pe, va_len, ea, ea_pe, ea_len, mark, &era, i, j)) continue;
                    deg[i] -= deg[i] < nv[j] ? deg[i] : nv[j];
                    nv[i] += nv[j];
                    nv[j] = 0;
                    uint32_t tail = j;
                    while (mlink[tail] != NONE) tail = mlink[tail];
                    mlink[tail] = mlink[i];
                    mlink[i] = j;
                }
            }
            hhead[b] = NONE;
        }
        for (uint32_t t = 0; t < nlp; t++) {
            uint32_t i = lp_buf[t];
            if (nv[i] == 0) continue;
            dl_insert(&dl, deg[i], i);
            if (deg[i] < mindeg) mindeg = deg[i];
        }
        q[k++] = p;
        for (uint32_t mb = mlink[p]; mb != NONE; mb = mlink[mb]) q[k++] = mb;
    }
    memcpy(q + k, lp_buf + dense_start, sizeof(uint32_t) * ndense);
done:
    free(va); free(va_pe); free(va_len); free(ea); free(ea_pe); free(ea_len); free(ea_lim);
    free(nv); free(elem_alive); free(esize); free(deg); free(w_buf); free(mark); free(wmark);
    free(hkey); free(hhead); free(hnext); free(hbuckets); free(mlink); free(lp_buf); free(touched);
    free(first); free(dl.head); free(dl.next); free(dl.prev);
    return rc;
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

int or_order(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q)
{
    if (n == 0) return 0;
    uint32_t *emit = malloc(sizeof(uint32_t) * n);
    uint32_t *block_ptr = malloc(sizeof(uint32_t) * (n + 1));
    uint32_t nb = (uint32_t)or_blocks(n, col_ptr, row_idx, emit, block_ptr);
    uint32_t *blk_of = malloc(sizeof(uint32_t) * n);
    uint32_t *loc = malloc(sizeof(uint32_t) * n);
    int rc = 0;
    for (uint32_t b = 0; b < nb; b++)
        for (uint32_t i = block_ptr[b], li = 0; i < block_ptr[b + 1]; i++, li++) {
            blk_of[emit[i]] = b;
            loc[emit[i]] = li;
        }
    for (uint32_t b = 0; b < nb && rc == 0; b++) {
        uint32_t lo = block_ptr[b], hi = block_ptr[b + 1], bs = hi - lo;
        const uint32_t *verts = emit + lo;
        if (bs == 1) {
            q[lo] = verts[0];
            continue;
        }
        uint32_t *sub_cp = malloc(sizeof(uint32_t) * (bs + 1));
        uint32_t sub_nnz = 0;
        sub_cp[0] = 0;
        for (uint32_t j = 0; j < bs; j++) {
            uint32_t v = verts[j];
            for (uint32_t p = col_ptr[v]; p < col_ptr[v + 1]; p++) sub_nnz += blk_of[row_idx[p]] == b;
            sub_cp[j + 1] = sub_nnz;
        }
        uint32_t *sub_ri = malloc(sizeof(uint32_t) * (sub_nnz + 1));
        uint32_t w = 0;
        for (uint32_t j = 0; j < bs; j++) {
            uint32_t v = verts[j];
            for (uint32_t p = col_ptr[v]; p < col_ptr[v + 1]; p++)
                if (blk_of[row_idx[p]] == b) sub_ri[w++] = loc[row_idx[p]];
        }
        uint32_t *sub_q = malloc(sizeof(uint32_t) * bs);
        uint32_t *nat = malloc(sizeof(uint32_t) * bs);
        if (or_amd(bs, sub_cp, sub_ri, sub_q) != 0) {
            rc = -1;
        } else {
            memcpy(nat, verts, sizeof(uint32_t) * bs);
            qsort(nat, bs, sizeof(uint32_t), cmp_u32);
            for (uint32_t i = 0; i < bs; i++) nat[i] = loc[nat[i]];
            uint64_t amd_count = or_chol_count(bs, sub_cp, sub_ri, sub_q, UINT64_MAX);
            const uint32_t *ord = or_chol_count(bs, sub_cp, sub_ri, nat, amd_count) < amd_count ? nat : sub_q;
            for (uint32_t i = 0; i < bs; i++) q[lo + i] = verts[ord[i]];
        }
        free(sub_cp); free(sub_ri); free(sub_q); free(nat);
    }
    free(emit); free(block_ptr); free(blk_of); free(loc);
    return rc;
}
