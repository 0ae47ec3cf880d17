// Column ordering of a template's pattern, computed once on the host and
// shared by every lane (the reference shares it per template through
// `ColumnOrder`, src/device/Circuit.zig:55-88). Restates
// src/solver/order.zig: iterative Tarjan BTF with the sink block first
// (:154-235); per non-singleton block, AMD on the block's sub-pattern
// (quotient graph, approximate external degree, aggressive absorption,
// hashed supervariables, dense rows last; :375-692) or the natural order
// when its Cholesky count of A + A^T is smaller (:133-146, cholCount
// :251-311). Workspace sizes follow the reference slab (4 nnz + 4 n + 16
// variable words, 8 n + nnz element words), so compaction happens where
// the reference's would.
#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

namespace esp {
namespace {

constexpr uint32_t TAG = 1u << 31;

struct Adj {
    std::vector<uint32_t> buf;
    std::vector<uint32_t> pe, len;
    size_t top = 0;
};

// order.zig:321-344: slides every live span below `top` to the front.
size_t compact(std::vector<uint32_t> &adj, size_t top, std::vector<uint32_t> &pe, const std::vector<uint32_t> &len,
               const std::vector<uint32_t> &live_a, const std::vector<uint32_t> &live_b, std::vector<uint32_t> &first)
{
    const uint32_t n = (uint32_t)pe.size();
    for (uint32_t o = 0; o < n; o++) {
        if (len[o] == 0 || (live_a[o] == 0 && live_b[o] == 0)) continue;
        first[o] = adj[pe[o]];
        adj[pe[o]] = TAG | o;
    }
    size_t dst = 0, src = 0;
    while (src < top) {
        if ((adj[src] & TAG) == 0) {
            src++;
            continue;
        }
        uint32_t o = adj[src] & ~TAG;
        uint32_t l = len[o];
        adj[dst] = first[o];
        std::memmove(&adj[dst + 1], &adj[src + 1], sizeof(uint32_t) * (l - 1));
        pe[o] = (uint32_t)dst;
        dst += l;
        src += l;
    }
    return dst;
}

struct Buckets {
    std::vector<uint32_t> head, next, prev;
    void insert(uint32_t d, uint32_t i)
    {
        next[i] = head[d];
        prev[i] = ESP_NONE;
        if (head[d] != ESP_NONE) prev[head[d]] = i;
        head[d] = i;
    }
    void remove(uint32_t d, uint32_t i)
    {
        if (prev[i] != ESP_NONE) next[prev[i]] = next[i];
        else head[d] = next[i];
        if (next[i] != ESP_NONE) prev[next[i]] = prev[i];
    }
};

bool same_adj(const std::vector<uint32_t> &va, const std::vector<uint32_t> &va_pe, const std::vector<uint32_t> &va_len,
              const std::vector<uint32_t> &ea, const std::vector<uint32_t> &ea_pe, const std::vector<uint32_t> &ea_len,
              std::vector<uint32_t> &mark, uint32_t &era, uint32_t i, uint32_t j)
{
    if (va_len[i] != va_len[j] || ea_len[i] != ea_len[j]) return false;
    uint32_t m = ++era;
    for (uint32_t t = 0; t < va_len[i]; t++) mark[va[va_pe[i] + t]] = m;
    for (uint32_t t = 0; t < va_len[j]; t++)
        if (mark[va[va_pe[j] + t]] != m) return false;
    m = ++era;
    for (uint32_t t = 0; t < ea_len[i]; t++) mark[ea[ea_pe[i] + t]] = m;
    for (uint32_t t = 0; t < ea_len[j]; t++)
        if (mark[ea[ea_pe[j] + t]] != m) return false;
    return true;
}

}  // namespace

uint32_t tarjan_blocks(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *emit, uint32_t *block_ptr)
{
    std::vector<uint32_t> num(n, ESP_NONE), low(n), scc(n), fv(n), fpos(n);
    std::vector<uint8_t> on(n, 0);
    uint32_t sp = 0, fp = 0, counter = 0, no = 0, nb = 0;
    block_ptr[0] = 0;
    for (uint32_t root = 0; root < n; root++) {
        if (num[root] != ESP_NONE) continue;
        num[root] = low[root] = counter++;
        scc[sp++] = root;
        on[root] = 1;
        fv[fp] = root;
        fpos[fp] = col_ptr[root];
        fp++;
        while (fp > 0) {
            uint32_t v = fv[fp - 1];
            if (fpos[fp - 1] < col_ptr[v + 1]) {
                uint32_t child = row_idx[fpos[fp - 1]++];
                if (child == v) continue;
                if (num[child] == ESP_NONE) {
                    num[child] = low[child] = counter++;
                    scc[sp++] = child;
                    on[child] = 1;
                    fv[fp] = child;
                    fpos[fp] = col_ptr[child];
                    fp++;
                } else if (on[child]) {
                    low[v] = std::min(low[v], num[child]);
                }
                continue;
            }
            if (low[v] == num[v]) {
                for (;;) {
                    uint32_t m = scc[--sp];
                    on[m] = 0;
                    emit[no++] = m;
                    if (m == v) break;
                }
                block_ptr[++nb] = no;
            }
            fp--;
            if (fp > 0) low[fv[fp - 1]] = std::min(low[fv[fp - 1]], low[v]);
        }
    }
    return nb;
}

uint64_t chol_count(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const uint32_t *perm, uint64_t cap)
{
    const uint32_t nnz = col_ptr[n];
    std::vector<uint32_t> pinv(n + 1), ptr(n + 1, 0), lst(nnz + 1), fill(n + 1), parent(n + 1);
    for (uint32_t k = 0; k < n; k++) pinv[perm[k]] = k;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t a = pinv[row_idx[p]], b = pinv[j];
            if (a != b) ptr[std::max(a, b) + 1]++;
        }
    for (uint32_t k = 1; k <= n; k++) ptr[k] += ptr[k - 1];
    std::copy(ptr.begin(), ptr.begin() + n, fill.begin());
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t a = pinv[row_idx[p]], b = pinv[j];
            if (a == b) continue;
            uint32_t hi = std::max(a, b);
            lst[fill[hi]++] = std::min(a, b);
        }
    std::vector<uint32_t> &anc = fill;
    for (uint32_t k = 0; k < n; k++) {
        parent[k] = ESP_NONE;
        anc[k] = ESP_NONE;
        for (uint32_t t = ptr[k]; t < ptr[k + 1]; t++) {
            uint32_t i = lst[t];
            while (i != ESP_NONE && i < k) {
                uint32_t next = anc[i];
                anc[i] = k;
                if (next == ESP_NONE) parent[i] = k;
                i = next;
            }
        }
    }
    std::vector<uint32_t> &flag = anc;
    uint64_t count = 0;
    for (uint32_t k = 0; k < n; k++) {
        flag[k] = k;
        count++;
        for (uint32_t t = ptr[k]; t < ptr[k + 1]; t++)
            for (uint32_t i = lst[t]; flag[i] != k; i = parent[i]) {
                flag[i] = k;
                count++;
            }
        if (count >= cap) return count;
    }
    return count;
}

int amd_order(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q)
{
    if (n == 0) return 0;
    const uint32_t nnz = col_ptr[n];
    const size_t va_cap = 4 * (size_t)nnz + 4 * (size_t)n + 16, ea_cap = 8 * (size_t)n + nnz;
    std::vector<uint32_t> va(va_cap), va_pe(n), va_len(n), ea(ea_cap), ea_pe(n), ea_len(n, 0), ea_lim(n);
    std::vector<uint32_t> nv(n, 1), elem_alive(n, 0), esize(n), deg(n, 0), w_buf(n), mark(n, 0), wmark(n, 0);
    std::vector<uint32_t> hkey(n), hhead(n, ESP_NONE), hnext(n), hbuckets(n), mlink(n, ESP_NONE), lp_buf(n), touched(n), first(n);
    Buckets dl{std::vector<uint32_t>(n + 1, ESP_NONE), std::vector<uint32_t>(n), std::vector<uint32_t>(n)};
    uint32_t era = 0, erw = 0;
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            if (row_idx[p] != j) {
                deg[j]++;
                deg[row_idx[p]]++;
            }
    size_t va_free = 0;
    for (uint32_t i = 0; i < n; i++) {
        va_pe[i] = (uint32_t)va_free;
        va_free += deg[i];
    }
    std::fill(deg.begin(), deg.end(), 0);
    for (uint32_t j = 0; j < n; j++)
        for (uint32_t p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            uint32_t r = row_idx[p];
            if (r == j) continue;
            va[va_pe[j] + deg[j]++] = r;
            va[va_pe[r] + deg[r]++] = j;
        }
    va_len = deg;
    for (uint32_t i = 0; i < n; i++) {
        era++;
        uint32_t wp = 0;
        for (uint32_t t = 0; t < va_len[i]; t++) {
            uint32_t v = va[va_pe[i] + t];
            if (mark[v] != era) {
                mark[v] = era;
                va[va_pe[i] + wp++] = v;
            }
        }
        va_len[i] = wp;
    }
    uint32_t isq = 0;
    while ((uint64_t)(isq + 1) * (isq + 1) <= n) isq++;
    const uint32_t dense_thresh = std::max<uint32_t>(16, 10 * isq);
    uint32_t ndense = 0;
    for (uint32_t i = 0; i < n; i++) ndense += va_len[i] >= dense_thresh;
    const uint32_t dense_start = n - ndense;
    for (uint32_t i = 0, di = 0; i < n; i++)
        if (va_len[i] >= dense_thresh) {
            nv[i] = 0;
            lp_buf[dense_start + di++] = i;
        }
    if (ndense > 0)
        for (uint32_t i = 0; i < n; i++) {
            if (nv[i] == 0) continue;
            uint32_t wp = 0;
            for (uint32_t t = 0; t < va_len[i]; t++) {
                uint32_t v = va[va_pe[i] + t];
                if (nv[v] != 0) va[va_pe[i] + wp++] = v;
            }
            va_len[i] = wp;
        }
    size_t ea_free = 0;
    for (uint32_t i = 0; i < n; i++) {
        ea_pe[i] = (uint32_t)ea_free;
        ea_lim[i] = 4;
        ea_free += 4;
    }
    for (uint32_t i = 0; i < n; i++)
        if (nv[i] != 0) {
            deg[i] = va_len[i];
            dl.insert(deg[i], i);
        }
    const uint32_t n_amd = n - ndense;
    uint32_t mindeg = 0, k = 0;
    while (k < n_amd) {
        while (dl.head[mindeg] == ESP_NONE) mindeg++;
        const uint32_t p = dl.head[mindeg];
        dl.remove(deg[p], p);
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
        const uint32_t nvpiv = nv[p];
        nv[p] = 0;
        elem_alive[p] = 1;
        if (va_free + nlp > va_cap) {
            va_len[p] = 0;
            va_free = compact(va, va_free, va_pe, va_len, nv, elem_alive, first);
            if (va_free + nlp > va_cap) return -1;
        }
        va_pe[p] = (uint32_t)va_free;
        va_len[p] = nlp;
        std::copy(lp_buf.begin(), lp_buf.begin() + nlp, va.begin() + va_free);
        va_free += nlp;
        uint32_t lpsize = 0;
        for (uint32_t t = 0; t < nlp; t++) lpsize += nv[lp_buf[t]];
        esize[p] = lpsize;
        for (uint32_t t = 0; t < nlp; t++) dl.remove(deg[lp_buf[t]], lp_buf[t]);
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
                w_buf[e] -= std::min(w_buf[e], nv[i]);
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
                    ea[ea_pe[i] + ewp++] = e;
                    esum += w_buf[e];
                }
            }
            if (ewp < ea_lim[i]) {
                ea[ea_pe[i] + ewp] = p;
            } else {
                uint32_t nc = (ewp + 1) * 2;
                if (ea_free + nc > ea_cap) {
                    ea_len[i] = ewp;
                    ea_free = compact(ea, ea_free, ea_pe, ea_len, nv, nv, first);
                    for (uint32_t o = 0; o < n; o++) ea_lim[o] = ea_len[o];
                    if (ea_free + nc > ea_cap) return -1;
                }
                std::memmove(&ea[ea_free], &ea[ea_pe[i]], sizeof(uint32_t) * ewp);
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
                    va[va_pe[i] + vwp++] = v;
                    asum += nv[v];
                }
            }
            va_len[i] = vwp;
            deg[i] = std::min(asum + (lpsize - nv[i]) + esum, n_amd - k - nvpiv);
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
            if (hhead[b] == ESP_NONE) hbuckets[nhb++] = b;
            hnext[i] = hhead[b];
            hhead[b] = i;
        }
        for (uint32_t hb = 0; hb < nhb; hb++) {
            uint32_t b = hbuckets[hb];
            for (uint32_t i = hhead[b]; i != ESP_NONE; i = hnext[i]) {
                if (nv[i] == 0) continue;
                for (uint32_t j = hnext[i]; j != ESP_NONE; j = hnext[j]) {
                    if (nv[j] == 0 || hkey[j] != hkey[i]) continue;
                    if (!same_adj(va, va_pe, va_len, ea, ea_pe, ea_len, mark, era, i, j)) continue;
                    deg[i] -= std::min(deg[i], nv[j]);
                    nv[i] += nv[j];
                    nv[j] = 0;
                    uint32_t tail = j;
                    while (mlink[tail] != ESP_NONE) tail = mlink[tail];
                    mlink[tail] = mlink[i];
                    mlink[i] = j;
                }
            }
            hhead[b] = ESP_NONE;
        }
        for (uint32_t t = 0; t < nlp; t++) {
            uint32_t i = lp_buf[t];
            if (nv[i] == 0) continue;
            dl.insert(deg[i], i);
            if (deg[i] < mindeg) mindeg = deg[i];
        }
        q[k++] = p;
        for (uint32_t mb = mlink[p]; mb != ESP_NONE; mb = mlink[mb]) q[k++] = mb;
    }
    std::copy(lp_buf.begin() + dense_start, lp_buf.begin() + dense_start + ndense, q + k);
    return 0;
}

int column_order(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, uint32_t *q)
{
    if (n == 0) return 0;
    std::vector<uint32_t> emit(n), block_ptr(n + 1), blk_of(n), loc(n);
    const uint32_t nb = tarjan_blocks(n, col_ptr, row_idx, emit.data(), block_ptr.data());
    for (uint32_t b = 0; b < nb; b++)
        for (uint32_t i = block_ptr[b], li = 0; i < block_ptr[b + 1]; i++, li++) {
            blk_of[emit[i]] = b;
            loc[emit[i]] = li;
        }
    for (uint32_t b = 0; b < nb; b++) {
        const uint32_t lo = block_ptr[b], bs = block_ptr[b + 1] - lo;
        const uint32_t *verts = emit.data() + lo;
        if (bs == 1) {
            q[lo] = verts[0];
            continue;
        }
        std::vector<uint32_t> sub_cp(bs + 1, 0), sub_ri, sub_q(bs), nat(verts, verts + bs);
        for (uint32_t j = 0; j < bs; j++) {
            uint32_t v = verts[j], cnt = 0;
            for (uint32_t p = col_ptr[v]; p < col_ptr[v + 1]; p++) cnt += blk_of[row_idx[p]] == b;
            sub_cp[j + 1] = sub_cp[j] + cnt;
        }
        sub_ri.reserve(sub_cp[bs] + 1);
        for (uint32_t j = 0; j < bs; j++)
            for (uint32_t p = col_ptr[verts[j]]; p < col_ptr[verts[j] + 1]; p++)
                if (blk_of[row_idx[p]] == b) sub_ri.push_back(loc[row_idx[p]]);
        if (amd_order(bs, sub_cp.data(), sub_ri.data(), sub_q.data()) != 0) return -1;
        std::sort(nat.begin(), nat.end());
        for (auto &v : nat) v = loc[v];
        const uint64_t amd_count = chol_count(bs, sub_cp.data(), sub_ri.data(), sub_q.data(), UINT64_MAX);
        const uint32_t *ord = chol_count(bs, sub_cp.data(), sub_ri.data(), nat.data(), amd_count) < amd_count ? nat.data() : sub_q.data();
        for (uint32_t i = 0; i < bs; i++) q[lo + i] = verts[ord[i]];
    }
    return 0;
}

}  // namespace esp
